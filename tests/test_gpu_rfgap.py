"""GPU parity of the RF-GAP proximity against the oracle's restatement of the
reference's cpu_proximity_rfgap (itself pinned bit for bit to the reference's
own function in test_oracle.py).  The GPU sums the terms of different trees
in no fixed order, so entries are compared at 1e-12 relative; symmetry, the
diagonal rule and the total (a closed form in the out-of-bag counts) are
exact or size-independent checks, also at a size the oracle is not run on."""
from __future__ import annotations

import numpy as np
import pytest
import torch

from oracle import oracle as O
from rfx_amd import _lib
from rfx_amd.forest import Forest, make_dataset
from rfx_amd.proximity import Accumulator, DeviceForest, finishprox_f64, mds_dense_rfgap, outlier_scores

from helpers import centred, eig_residuals, load_golden

pytestmark = pytest.mark.gpu

# (kind, n, p, classes, T, seed): one tile, many tiles with a ragged edge
# n = 50: one 64 x 64 tile pair; n = 640: whole tiles only; the others end in a partial tile
SHAPES = [(_lib.DATA_OVERLAP_MIXTURE, 50, 7, 3, 6, 9), (_lib.DATA_GAUSS_MIXTURE, 640, 12, 4, 11, 2),
          (_lib.DATA_OVERLAP_MIXTURE, 300, 7, 3, 6, 9), (_lib.DATA_GAUSS_MIXTURE, 2500, 12, 4, 21, 3),
          (_lib.DATA_COVTYPE_LIKE, 9000, 54, 7, 8, 5)]


def _device_rfgap(f: Forest, x: np.ndarray, cat=None):
    df = DeviceForest(f, x, cat)
    _, nodextr = df.leaf_assign(want_nodextr=True)
    nin = torch.from_numpy(np.ascontiguousarray(f.nin.reshape(f.ntree, f.n), np.int32)).to(df.device)
    acc = Accumulator(f.n, f.ntree, f.total_nodes)
    P = acc.rfgap(nin, nodextr, df.node_off)
    torch.cuda.synchronize()
    return P, nodextr


def _total(nin: np.ndarray, nd: np.ndarray) -> float:
    # sum_i (share of i's out-of-bag trees whose node holds in-bag cases) + #never out of bag
    T, n = nin.shape
    S = (nin == 0).sum(axis=0)
    share = np.zeros(n)
    for t in range(T):
        has = np.zeros(nd[t].max() + 1, bool)
        has[nd[t][(nin[t] > 0) & (nd[t] >= 0)]] = True
        ok = (nin[t] == 0) & (nd[t] >= 0)
        ok[ok] &= has[nd[t][ok]]
        share += ok
    return float(np.sum(share[S > 0] / S[S > 0]) + np.sum(S == 0))


@pytest.mark.parametrize("shape", SHAPES, ids=lambda s: f"n{s[1]}_T{s[4]}")
def test_rfgap_matches_oracle(gpu, shape):
    kind, n, p, ncl, T, seed = shape
    print(f"rfgap n={n} T={T} seed={seed}", flush=True)
    x, y = make_dataset(kind, n, p, ncl, seed)
    f = Forest.grow(x, y, T, iseed=seed)
    P, nodextr = _device_rfgap(f, x)
    nd = O.leaf_assign(f, x)
    assert np.array_equal(nodextr.cpu().numpy(), nd)
    want = O.prox_rfgap(f.nin.reshape(T, n), nd)
    g = P.cpu().numpy()
    assert np.allclose(g, want, rtol=1e-12, atol=1e-14)
    assert np.array_equal(g == 0, want == 0)
    G = g.reshape(n, n, order="F")
    assert np.array_equal(G, G.T)
    S = (f.nin.reshape(T, n) == 0).sum(axis=0)
    assert np.all(np.diag(G)[S == 0] == 1.0) and np.all(np.diag(G)[S > 0] == 0.0)


def test_rfgap_golden_categorical(gpu):
    # categorical forest whose cases may stop on inner nodes: RF-GAP matches
    # cases by equal reached node, as the reference does
    print("rfgap golden categorical", flush=True)
    g = load_golden("covtype_cat_n240.npz")
    f = g["forest"]
    P, nodextr = _device_rfgap(f, g["x"], g["cat"])
    assert np.array_equal(nodextr.cpu().numpy(), g["nodextr"])
    want = O.prox_rfgap(g["nin"], g["nodextr"])
    assert np.allclose(P.cpu().numpy(), want, rtol=1e-12, atol=1e-14)


def test_rfgap_invariants_at_scale(gpu):
    n, T = 12000, 60
    print(f"rfgap invariants n={n} T={T}", flush=True)
    x, y = make_dataset(_lib.DATA_COVTYPE_LIKE, n, 54, 7, 4)
    f = Forest.grow(x, y, T, iseed=4)
    P, nodextr = _device_rfgap(f, x)
    G = P.view(n, n)
    assert torch.equal(G, G.t())
    nin = f.nin.reshape(T, n)
    S = torch.from_numpy((nin == 0).sum(axis=0)).to(gpu)
    d = torch.diagonal(G)
    assert bool(torch.all(d[S == 0] == 1.0)) and bool(torch.all(d[S > 0] == 0.0))
    assert bool(torch.all(G >= 0)) and bool(torch.all(G <= 1.0))
    total = float(G.sum())
    assert abs(total - _total(nin, nodextr.cpu().numpy())) < 1e-8 * n


@pytest.mark.parametrize("kind,n,p,ncl,T,seed", [(_lib.DATA_COVTYPE_LIKE, 1500, 54, 7, 100, 12),
                                                 (_lib.DATA_OVERLAP_MIXTURE, 1200, 8, 3, 80, 12)])
def test_mds_rfgap_chain(gpu, kind, n, p, ncl, T, seed):
    print(f"rfgap mds n={n} T={T}", flush=True)
    x, y = make_dataset(kind, n, p, ncl, seed)
    f = Forest.grow(x, y, T, iseed=seed)
    P, _ = _device_rfgap(f, x)
    nout = f.nout(2)
    X, eig, it = mds_dense_rfgap(P.float().double(), n, torch.from_numpy(nout).to(gpu))
    Pref = O.prox_rfgap(f.nin.reshape(T, n), O.leaf_assign(f, x)).reshape(n, n, order="F")
    M = O.mds_input_rfgap(Pref, nout)
    Xref, w = O.classical_mds(M, 3)
    Xg = X.cpu().numpy()
    assert np.allclose(eig[:3], w[:3], rtol=1e-9)
    res = eig_residuals(centred(M), Xg, eig[:3])
    print(it, eig[:5], w[:5], res, flush=True)
    assert res.max() < 1e-5
    if w[2] / w[3] > 1.02:
        assert O.embedding_errors(Xg, Xref)["rel_frobenius"] < 1e-5


def test_rfgap_rejects_bad_shapes(gpu):
    x, y = make_dataset(_lib.DATA_GAUSS_MIXTURE, 200, 5, 3, 1)
    f = Forest.grow(x, y, 3, iseed=1)
    df = DeviceForest(f, x)
    _, nodextr = df.leaf_assign(want_nodextr=True)
    acc = Accumulator(200, 3, f.total_nodes)
    nin = torch.from_numpy(np.ascontiguousarray(f.nin.reshape(3, 200), np.int32)).to(df.device)
    with pytest.raises(ValueError):
        acc.rfgap(torch.zeros(10, dtype=torch.int32, device=df.device), nodextr, df.node_off)
    with pytest.raises(ValueError):
        acc.rfgap(nin, nodextr, df.node_off[:2])


@pytest.mark.parametrize("kind,n,p,ncl,T,seed", [(_lib.DATA_GAUSS_MIXTURE, 2500, 12, 4, 21, 3),
                                                 (_lib.DATA_COVTYPE_LIKE, 4100, 54, 7, 30, 5)])
def test_outlier_scores_bitwise(gpu, kind, n, p, ncl, T, seed):
    # RandomForest::compute_outlier_scores on the RF-GAP matrix and on the
    # finishprox output of the OOB counts, bit for bit against the oracle
    print(f"outlier scores n={n} T={T}", flush=True)
    x, y = make_dataset(kind, n, p, ncl, seed)
    f = Forest.grow(x, y, T, iseed=seed)
    P, _ = _device_rfgap(f, x)
    got, med = outlier_scores(P, n)
    want = O.outlier_scores(P.cpu().numpy())
    assert np.array_equal(got.cpu().numpy(), want)
    assert med > 0
    df = DeviceForest(f, x)
    keys, _ = df.leaf_assign()
    acc = Accumulator(n, T, f.total_nodes)
    C = acc.accumulate(keys, _lib.PROX_OOB_ASYM, _lib.LAYOUT_FULL)
    nout = torch.from_numpy(f.nout(2)).to(gpu)
    S = finishprox_f64(C.double(), nout, n, _lib.LAYOUT_FULL)
    got, _ = outlier_scores(S, n)
    assert np.array_equal(got.cpu().numpy(), O.outlier_scores(S.cpu().numpy()))
