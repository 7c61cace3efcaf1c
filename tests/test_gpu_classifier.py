"""The classifier mirror (reference pybind options and proximity methods) on
the GPU, against the reference's golden vectors and the FP64 MDS oracle."""
from __future__ import annotations

import numpy as np
import pytest

from oracle import oracle as O
from rfx_amd.classifier import RandomForestClassifier

from helpers import load_golden

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name", ["gauss_n200.npz", "covtype_cat_n240.npz"])
def test_fit_proximity_matches_reference_golden(gpu, name):
    print(f"classifier golden {name}", flush=True)
    g = load_golden(name)
    x, y = g["x"], g["y"].astype(np.int32)
    clf = RandomForestClassifier(nsample=len(y), ntree=int(g["ntree"]), mdim=x.shape[1], maxcat=int(g["maxcat"]),
                                 mtry=int(g["mtry"]), iseed=int(g["iseed"]), compute_proximity=True, use_gpu=True)
    clf.fit(x, y)
    assert np.array_equal(clf.cat, g["cat"])
    S = clf.get_proximity_matrix()
    assert S.dtype == np.float32 and S.shape == (len(y), len(y))
    exp = g["proxsym"].reshape(len(y), len(y), order="F").astype(np.float32)
    assert np.array_equal(S, exp)
    # the reference chain's MDS: float32 S with nout renormalisation, FP64
    X = clf.compute_mds_3d_cpu()
    M = O.mds_input_reference(g["proxsym"].reshape(len(y), len(y), order="F"), g["nout"])
    Xref, w = O.classical_mds(M, 3)
    print(w[:5], flush=True)
    if w[2] / w[3] > 1.02:
        assert O.embedding_errors(X, Xref)["rel_frobenius"] < 1e-6
    else:
        assert np.allclose(np.sort(np.linalg.norm(X, axis=0) ** 2), np.sort(w[:3]), rtol=1e-8)
    with pytest.raises(RuntimeError):
        clf.get_lowrank_factors()


def test_lowrank_mode(gpu):
    print("classifier low-rank mode", flush=True)
    g = load_golden("gauss_n200.npz")
    x, y = g["x"], g["y"].astype(np.int32)
    clf = RandomForestClassifier(ntree=40, mtry=4, iseed=0, compute_proximity=True, use_gpu=True, use_qlora=True,
                                 quant_mode="int8", rank=16)
    clf.fit(x, y)
    A, B, r = clf.get_lowrank_factors()
    assert r == 16 and A.shape == (200, 16) and A.dtype == np.float64
    with pytest.raises(RuntimeError):
        clf.get_proximity_matrix()
    X = clf.compute_mds_from_factors(k=2)
    assert X.shape == (200, 2) and np.all(np.isfinite(X))
    assert clf.compute_mds_3d_from_factors().shape == (200, 3)


def test_casewise_proximity_and_mds(gpu):
    # use_casewise: the case-wise FP64 accumulation, finishprox and the MDS
    # chain, against the oracle chain on the same forest (the oracle's
    # case-wise V1 is pinned bit for bit to the reference's cpu_proximity)
    print("classifier case-wise", flush=True)
    g = load_golden("gauss_n200.npz")
    x, y = g["x"], g["y"].astype(np.int32)
    clf = RandomForestClassifier(ntree=30, mtry=4, iseed=2, compute_proximity=True, use_gpu=True, use_casewise=True)
    clf.fit(x, y)
    f = clf.forest
    n = len(y)
    pw = O.prox_v1(f, O.leaf_assign(f, x), casewise=True)
    nout = f.nout(2)
    S_ref = O.finishprox(n, nout, pw).reshape(n, n, order="F")
    S = clf.get_proximity_matrix()
    assert S.dtype == np.float32 and S.shape == (n, n)
    assert np.allclose(S, S_ref.astype(np.float32), rtol=1e-6, atol=1e-7)
    X = clf.compute_mds_3d_cpu()
    M = O.mds_input_reference(S_ref, nout)
    Xref, w = O.classical_mds(M, 3)
    if w[2] / w[3] > 1.02:
        assert O.embedding_errors(X, Xref)["rel_frobenius"] < 1e-5
    else:
        assert np.allclose(np.sort(np.linalg.norm(X, axis=0) ** 2), np.sort(w[:3]), rtol=1e-6)


def test_rfgap_proximity_and_mds(gpu):
    # use_rfgap: the RF-GAP matrix as float32 and the use_rfgap MDS chain,
    # against the oracle (pinned bit for bit to the reference's
    # cpu_proximity_rfgap) on the same forest
    print("classifier rf-gap", flush=True)
    g = load_golden("gauss_n200.npz")
    x, y = g["x"], g["y"].astype(np.int32)
    clf = RandomForestClassifier(ntree=30, mtry=4, iseed=2, compute_proximity=True, use_gpu=True, use_rfgap=True,
                                 use_casewise=True)
    clf.fit(x, y)
    f = clf.forest
    n = len(y)
    Pref = O.prox_rfgap(f.nin.reshape(f.ntree, n), O.leaf_assign(f, x)).reshape(n, n, order="F")
    P = clf.get_proximity_matrix()
    assert P.dtype == np.float32 and P.shape == (n, n)
    assert np.allclose(P, Pref.astype(np.float32), rtol=1e-6, atol=1e-7)
    X = clf.compute_mds_3d_cpu()
    Xref, w = O.classical_mds(O.mds_input_rfgap(Pref, f.nout(2)), 3)
    if w[2] / w[3] > 1.02:
        assert O.embedding_errors(X, Xref)["rel_frobenius"] < 1e-5
    else:
        assert np.allclose(np.sort(np.linalg.norm(X, axis=0) ** 2), np.sort(w[:3]), rtol=1e-6)
