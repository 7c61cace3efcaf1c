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


def chain_tree(p: int, n: int, length: int, rng, nest: float = 0.01, exit_q: float = 3.2) -> Forest:
    """One tree of random chains shaped like peel chains: each chain node
    splits on a random feature, which side is heavy drawn at random, with a
    threshold that drifts from exit_q towards the centre along the chain
    (|thr| = exit_q - 2.5 * position / length + N(0, 0.05)), so the light
    side takes the extreme of N(0, 1) samples and exits spread along the
    chain, with thresholds out of order here and there; a light child is a
    leaf or, with probability nest, the top of another chain of the same
    length.  Children are allocated in pairs after their parent, as
    rfx_pack_nodes requires."""
    st, bv, xs, tm = [], [], [], []

    def new():
        st.append(-1)
        bv.append(0)
        xs.append(0.0)
        tm.extend([0, 0])
        return len(st) - 1

    pending = [new()]
    budget = 4
    while pending:
        cur = pending.pop(0)
        for j in range(length):
            l = new()
            r = new()
            f = int(rng.integers(p))
            right_heavy = bool(rng.random() < 0.5)
            thr = exit_q - 2.5 * j / length + rng.normal(0.0, 0.05)
            st[cur], bv[cur], xs[cur] = 1, f, (-thr if right_heavy else thr)
            tm[2 * cur], tm[2 * cur + 1] = l, r
            heavy, light = (r, l) if right_heavy else (l, r)
            if budget > 0 and rng.random() < nest:
                budget -= 1
                pending.append(light)
            cur = heavy
    nn = len(st)
    nin = rng.poisson(1.0, n).astype(np.int32)
    return Forest(n, p, np.array([0, nn], np.int64), np.array(st, np.int32), np.array(bv, np.int32),
                  np.array(xs, np.float32), np.array(tm, np.int32), np.zeros(nn, np.int32), nin.reshape(1, n))


def random_chain_forest(n: int, p: int, seed: int, length: int = 1500, trees: int = 3):
    """Samples N(0, 1) with a few NaN-free duplicates of thresholds, random
    chain trees between two grown trees."""
    rng = np.random.default_rng(seed)
    x = rng.standard_normal((n, p)).astype(np.float32)
    y = (x[:, 0] > 0).astype(np.int32)
    g = Forest.grow(x, y, 2, iseed=seed)
    parts = [g.subset(slice(0, 1))] + [chain_tree(p, n, length, rng) for _ in range(trees)] + [g.subset(slice(1, 2))]
    f = Forest.concat(parts)
    # some samples exactly on a threshold of the first chain tree
    b = f.node_offset[1]
    inner = np.nonzero(f.nodestatus[b:f.node_offset[2]] == 1)[0][:50]
    for j, k in enumerate(inner):
        x[j, f.bestvar[b + k]] = f.xbestsplit[b + k]
    return x, f


def chain_index_sections(index: np.ndarray) -> dict:
    """Views of a chain index's sections (writable when index is)."""
    offs = [int(index[16 + 8 * k:24 + 8 * k].view(np.uint64)[0]) for k in range(8)]
    names = ["tree_lab", "tree_path0", "labels", "paths", "lists", "recs", "exits", "total"]
    sec = {nm: index[offs[k]:offs[k + 1]] for k, nm in enumerate(names[:-1])}
    sec["recs"] = sec["recs"].view(np.uint32).reshape(-1, 2)
    sec["exits"] = sec["exits"].view(np.uint32)
    sec["lists"] = sec["lists"].view(np.uint32).reshape(-1, 4)  # rec0 lo, rec0 hi, count, feature | side << 16
    return sec


def chain_jump_walk(forest: Forest, x: np.ndarray, index: np.ndarray, expect: np.ndarray | None = None):
    """Host restatement of the drain kernel's walk with jumps, read from the
    chain index bytes (test side only): nodextr [T, n].  A jump crosses a
    light edge, so a walk takes at most log2(nodes) of them; more than 64
    marks the walk broken (-2).  With expect, returns at the first sample
    whose node differs from expect (then False), or True."""
    u64 = lambda o: int(index[o:o + 8].view(np.uint64)[0])
    magic, ntree, npaths, _ = index[:16].view(np.int32)
    assert int(np.uint32(magic)) == 0x43584652 and ntree == forest.ntree
    offs = [u64(16 + 8 * k) for k in range(8)]
    tree_lab = index[offs[0]:offs[0] + 8 * ntree].view(np.int64)
    tree_path0 = index[offs[1]:offs[1] + 4 * ntree].view(np.int32)
    labels = index[offs[2]:offs[3]].view(np.uint32)
    paths = index[offs[3]:offs[3] + 24 * npaths]
    lists_raw = index[offs[4]:offs[5]]
    recs = index[offs[5]:offs[6]].view(np.uint32).reshape(-1, 2)
    exits = index[offs[6]:offs[7]].view(np.uint32)

    def path(gp):
        r = paths[24 * gp:24 * gp + 24]
        exit0 = int(r[:8].view(np.uint64)[0])
        list0, nlists, length, end = (int(v) for v in r[8:24].view(np.uint32))
        return exit0, list0, nlists, length, end

    def lst(li):
        r = lists_raw[16 * li:16 * li + 16]
        rec0 = int(r[:8].view(np.uint64)[0])
        count = int(r[8:12].view(np.uint32)[0])
        feat, right = (int(v) for v in r[12:16].view(np.uint16))
        return rec0, count, feat, right

    thr_all = recs[:, 0].view(np.float32)
    out = np.empty((forest.ntree, forest.n), np.int32)
    for t in range(forest.ntree):
        b = int(forest.node_offset[t])
        st = forest.nodestatus[b:forest.node_offset[t + 1]]
        bv = forest.bestvar[b:forest.node_offset[t + 1]]
        xs = forest.xbestsplit[b:forest.node_offset[t + 1]]
        tm = forest.treemap[2 * b:2 * int(forest.node_offset[t + 1])]
        lab0 = int(tree_lab[t])
        nn = len(st)
        for i in range(forest.n):
            bag = forest.nin[t, i] > 0
            kt, steps, moves, jumps = 0, 0, 0, 0
            while st[kt] == 1 and (bag or steps < 999):
                moves += 1
                if moves > 2 * nn + 2:  # a broken index can cycle
                    kt = -2
                    break
                lab = int(labels[lab0 + kt]) if lab0 >= 0 else 0xFFFFFFFF
                if lab == 0xFFFFFFFF:
                    kt = tm[2 * kt] if x[i, bv[kt]] <= xs[kt] else tm[2 * kt + 1]
                    steps += 1
                    continue
                jumps += 1
                if jumps > 64:
                    kt = -2
                    break
                exit0, list0, nlists, length, end = path(int(tree_path0[t]) + (lab >> 22))
                p0 = lab & 0x3FFFFF
                e = None
                for li in range(list0, list0 + nlists):
                    rec0, count, feat, right = lst(li)
                    th = thr_all[rec0:rec0 + count]
                    le = x[i, feat] <= th
                    hit = np.nonzero(le if right else ~le)[0]
                    if len(hit):
                        pos = int(recs[rec0 + hit[0], 1])
                        e = pos if e is None else min(e, pos)
                if e is None:
                    kt, steps = end, steps + length - p0
                else:
                    kt, steps = int(exits[exit0 + e]), steps + e - p0 + 1
                if not 0 <= kt < nn:
                    kt = -2
                    break
                if not bag and steps > 999:
                    kt, steps = -1, 999
                    break
            out[t, i] = kt if kt == -2 else (0 if kt < 0 or st[kt] != -1 else kt)
            if expect is not None and out[t, i] != expect[t, i]:
                return False
    return True if expect is not None else out
