"""Shared test helpers: golden fixtures as Forest objects, small workloads."""
from __future__ import annotations

from pathlib import Path

import numpy as np

from rfx_amd.forest import Forest, make_dataset

GOLDEN = Path(__file__).resolve().parent / "golden"


def load_golden(name: str) -> dict:
    with np.load(GOLDEN / name, allow_pickle=False) as z:
        g = {k: z[k] for k in z.files}
    n, p = g["x"].shape
    maxnode = 2 * n + 1
    g["forest"] = Forest.from_arrays(n, p, g["nnode"], g["nodestatus"], g["bestvar"], g["xbestsplit"], g["treemap"],
                                     g["nin"], maxnode=maxnode)
    return g


def small_forest(kind: int, n: int, p: int, ncl: int, T: int, seed: int = 0, mtry: int = 0, minndsize: int = 1):
    x, y = make_dataset(kind, n, p, ncl, seed)
    f = Forest.grow(x, y, T, mtry=mtry, minndsize=minndsize, iseed=seed)
    return x, y, f


def keys_from_nodextr(forest: Forest, nodextr: np.ndarray) -> np.ndarray:
    """Host restatement of the key encoding: (global node << 1 | in-bag) when
    the node is terminal, else -1."""
    T, n = nodextr.shape
    keys = np.full((T, n), -1, np.int64)
    for t in range(T):
        b = forest.node_offset[t]
        st = forest.nodestatus[b:forest.node_offset[t + 1]]
        nd = nodextr[t]
        term = st[nd] == -1
        keys[t, term] = ((b + nd[term]) << 1) | (forest.nin[t, term] > 0)
    return keys.astype(np.int32)


def centred(M: np.ndarray) -> np.ndarray:
    D = M.max() - M
    np.fill_diagonal(D, 0.0)
    D2 = D * D
    r = D2.mean(axis=1)
    return -0.5 * (D2 - r[:, None] - r[None, :] + r.mean())


def eig_residuals(B: np.ndarray, X: np.ndarray, lam: np.ndarray) -> np.ndarray:
    """||B v - lambda v|| / (lambda ||v||) per embedding column (v = x / sqrt(lambda))."""
    V = X / np.sqrt(lam[None, :])
    return np.linalg.norm(B @ V - V * lam[None, :], axis=0) / (np.abs(lam) * np.linalg.norm(V, axis=0))
