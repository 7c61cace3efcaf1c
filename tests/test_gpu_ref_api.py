"""The reference-signature C++ entry points (include/rfx_ref.hpp, through the
pybind11 module rfx_amd._rfxref) on the GPU: per-tree proximity and
finishprox bit-exact against the reference's golden vectors, the per-tree
FP16 upper triangle against the oracle, the matrix MDS against the FP64
classical-MDS oracle, and the low-rank factor object's host helpers."""
from __future__ import annotations

import numpy as np
import pytest

from oracle import oracle as O
from rfx_amd import _lib

from helpers import centred, eig_residuals, load_golden, small_forest

pytestmark = pytest.mark.gpu


def _say(msg: str) -> None:
    print(msg, flush=True)


@pytest.fixture(scope="module")
def R(gpu):
    import rfx_amd._rfxref as mod
    return mod


def _tree(f, t):
    b, e = f.node_offset[t], f.node_offset[t + 1]
    return np.ascontiguousarray(f.nodestatus[b:e], np.int32), int(e - b)


@pytest.mark.parametrize("name", ["gauss_n200.npz", "covtype_cat_n240.npz"])
def test_per_tree_proximity_and_finishprox_match_golden(R, name):
    _say(f"ref api per-tree proximity {name}")
    g = load_golden(name)
    f = g["forest"]
    n = f.n
    prox = np.zeros((n, n), np.float64, order="F")
    for t in range(f.ntree):
        st, _ = _tree(f, t)
        R.gpu_proximity(st, np.ascontiguousarray(g["nodextr"][t], np.int32),
                        np.ascontiguousarray(f.nin[t], np.int32), prox)
    assert np.array_equal(prox.reshape(-1, order="F"), g["prox"])
    S = R.finishprox(np.ascontiguousarray(g["nout"], np.int32), prox)
    assert np.array_equal(S.reshape(-1, order="F").view(np.uint64), g["proxsym"].view(np.uint64))


def test_per_tree_upper_fp16_matches_oracle(R):
    x, y, f = small_forest(_lib.DATA_GAUSS_MIXTURE, 900, 6, 3, 3, seed=21)
    nodextr = O.leaf_assign(f, x)
    for t in range(f.ntree):
        _say(f"ref api upper fp16 tree {t}")
        st, _ = _tree(f, t)
        got = R.upper_triangle_fp16(st, np.ascontiguousarray(nodextr[t], np.int32),
                                    np.ascontiguousarray(f.nin[t], np.int32))
        sub = f.subset(slice(t, t + 1))
        exp = O.counts(sub, nodextr[t:t + 1], _lib.PROX_OOB_UPPER, _lib.LAYOUT_UPPER)
        assert exp.max() <= 1
        assert np.array_equal(got, exp.astype(np.float16).view(np.uint16))


def test_matrix_mds_matches_oracle(R):
    _say("ref api compute_mds_3d")
    x, y, f = small_forest(_lib.DATA_OVERLAP_MIXTURE, 800, 8, 3, 60, seed=22)
    nodextr = O.leaf_assign(f, x)
    n = f.n
    P = O.counts(f, nodextr, _lib.PROX_OOB_ASYM, _lib.LAYOUT_FULL).astype(np.float64)
    nout = f.nout(2)
    S = O.finishprox(n, nout, P).reshape(n, n, order="F")
    Sf = np.asfortranarray(S.astype(np.float32).astype(np.float64))  # the pybind path's float32 round trip
    X = R.compute_mds_3d(Sf, nout, False)
    M = O.mds_input_reference(S, nout)
    Xref, w = O.classical_mds(M, 3)
    lam = np.sum(X * X, axis=0)  # x_c = v_c sqrt(lambda_c)
    _say(f"eigs {w[:5]} got {lam}")
    # eigenvalues and eigen-residuals pin the result whatever the eigengap
    assert np.allclose(lam, w[:3], rtol=1e-9)
    assert eig_residuals(centred(M), X, lam).max() < 1e-5
    if w[2] / w[3] > 1.02:
        assert O.embedding_errors(X, Xref)["rel_frobenius"] < 1e-6
    # without OOB counts the matrix is embedded as it is
    X2 = R.compute_mds_3d(Sf, None, False)
    lam2 = np.sum(X2 * X2, axis=0)
    _, w2 = O.classical_mds(Sf, 3)
    assert np.allclose(lam2, w2[:3], rtol=1e-9)
    assert eig_residuals(centred(Sf), X2, lam2).max() < 1e-5
    # use_rfgap keeps the diagonal of the divided rows instead of a unit one
    X3 = R.compute_mds_3d(Sf, nout, True)
    M3 = M.copy()
    np.fill_diagonal(M3, np.where(nout > 0, np.diag(Sf) / np.where(nout > 0, nout, 1), np.diag(Sf)))
    lam3 = np.sum(X3 * X3, axis=0)
    _, w3 = O.classical_mds(M3, 3)
    assert np.allclose(lam3, w3[:3], rtol=1e-9)
    assert eig_residuals(centred(M3), X3, lam3).max() < 1e-5
    # points on a line: D^2 double-centred has rank 1, fewer than three
    # eigenvalues above 1e-10, and the reference throws
    z = np.linspace(0.0, 1.0, 40)
    D1 = np.abs(z[:, None] - z[None, :])
    with pytest.raises(RuntimeError):
        R.compute_mds_3d(np.asfortranarray(D1.max() - D1), None, False)


@pytest.mark.parametrize("name", ["covtype_cat_n240.npz", "gauss_n200.npz"])
def test_rfgap_reference_signature(R, name):
    # rf::gpu_proximity_rfgap's signature (flattened per-tree nin and reached
    # nodes, no node counts), against the oracle pinned to the reference's
    # cpu_proximity_rfgap; then the use_rfgap MDS renormalisation
    _say(f"rfgap reference signature {name}")
    g = load_golden(name)
    n = g["nin"].shape[1]
    P = R.proximity_rfgap(g["nin"], g["nodextr"])
    want = O.prox_rfgap(g["nin"], g["nodextr"]).reshape(n, n, order="F")
    assert np.allclose(P, want, rtol=1e-12, atol=1e-14)
    # the float32 matrix of the pybind get_proximity_matrix, as the classifier's MDS receives it
    assert np.array_equal(R.outlier_scores(P), O.outlier_scores(np.asfortranarray(P).reshape(-1, order="F")))
    X = R.compute_mds_3d(np.asfortranarray(P.astype(np.float32).astype(np.float64)), g["nout"], True)
    M = O.mds_input_rfgap(want, g["nout"])
    _, w = O.classical_mds(M, 3)
    assert np.allclose(np.sort(np.linalg.norm(X, axis=0) ** 2)[::-1], w[:3], rtol=1e-7)
    assert eig_residuals(centred(M), X, np.linalg.norm(X, axis=0) ** 2).max() < 1e-5


def test_lowrank_object_and_host_helpers(R):
    _say("ref api LowRankProximityMatrix")
    x, y, f = small_forest(_lib.DATA_OVERLAP_MIXTURE, 1500, 8, 3, 40, seed=23)
    nodextr = O.leaf_assign(f, x)
    n, rank = f.n, 16
    m = R.LowRankProximityMatrix(n, rank, _lib.PREC_FP32)
    for t in range(f.ntree):
        st, _ = _tree(f, t)
        m.add_tree(st, np.ascontiguousarray(nodextr[t], np.int32), np.ascontiguousarray(f.nin[t], np.int32))
    m.finalize()
    A, B, r = m.factors()
    assert r == rank and A.shape == (n, rank) and np.array_equal(A, B)
    C = O.counts(f, nodextr, _lib.PROX_ALL, _lib.LAYOUT_FULL).reshape(n, n).astype(np.float64) / f.ntree
    w = np.linalg.eigvalsh(C)[::-1]
    best = np.sqrt(np.sum(w[rank:] ** 2))  # Eckart-Young optimum of a rank-r PSD approximation
    err = np.linalg.norm(C - A @ A.T)
    _say(f"factor error {err:.4f} vs optimum {best:.4f}")
    assert err <= 1.01 * best
    X = m.mds(3)
    assert X.shape == (n, 3) and np.all(np.isfinite(X))


@pytest.mark.parametrize("name", ["gauss_n200.npz", "covtype_cat_n240.npz"])
def test_casewise_switch_of_the_per_tree_proximity(R, name):
    # rf::g_config.use_casewise: cpu_proximity (and gpu_proximity below its
    # n > 5000 branch) weight each in-bag partner by nin(j) / (sum of nin
    # over the leaf); pinned by the reference RandomForest fitted with
    # use_casewise (golden rf_prox_casewise, after cpu_finishprox)
    _say(f"ref api case-wise switch {name}")
    g = load_golden(name)
    f = g["forest"]
    n = f.n
    assert not R.get_casewise()
    try:
        R.set_casewise(True)
        for fn in (R.cpu_proximity, R.gpu_proximity):
            prox = np.zeros((n, n), np.float64, order="F")
            for t in range(f.ntree):
                st, _ = _tree(f, t)
                fn(st, np.ascontiguousarray(g["nodextr"][t], np.int32), np.ascontiguousarray(f.nin[t], np.int32),
                   prox)
            S = R.finishprox(np.ascontiguousarray(g["nout"], np.int32), prox).reshape(-1, order="F")
            assert np.allclose(S, g["rf_prox_casewise"], rtol=1e-12, atol=1e-15)
    finally:
        R.set_casewise(False)


def test_gpu_proximity_dispatch_above_5000_samples(R):
    # n > 5000 with nnode > 10: the rule of the reference's
    # cuda_proximity_kernel (out-of-bag i, any j of its terminal node, i
    # itself included; no case-wise weights) for every out-of-bag row;
    # cpu_proximity keeps the V1 rule at every n
    _say("ref api gpu_proximity dispatch n=5200")
    x, y, f = small_forest(_lib.DATA_OVERLAP_MIXTURE, 5200, 6, 3, 2, seed=25)
    nodextr = O.leaf_assign(f, x)
    n = f.n
    g4 = np.zeros((n, n), np.float64, order="F")
    v1 = np.zeros((n, n), np.float64, order="F")
    for t in range(f.ntree):
        st, _ = _tree(f, t)
        args = (st, np.ascontiguousarray(nodextr[t], np.int32), np.ascontiguousarray(f.nin[t], np.int32))
        R.gpu_proximity(*args, g4)
        R.cpu_proximity(*args, v1)
    want4 = O.counts(f, nodextr, _lib.PROX_OOB_ROW, _lib.LAYOUT_FULL).astype(np.float64)
    want1 = O.counts(f, nodextr, _lib.PROX_OOB_ASYM, _lib.LAYOUT_FULL).astype(np.float64)
    assert np.array_equal(g4.reshape(-1, order="F"), want4)
    assert np.array_equal(v1.reshape(-1, order="F"), want1)
    # below the threshold gpu_proximity is V1 as well
    small = f.samples(np.arange(3000))
    nd = O.leaf_assign(small, np.ascontiguousarray(x[:3000]))
    p = np.zeros((3000, 3000), np.float64, order="F")
    for t in range(f.ntree):
        st, _ = _tree(f, t)
        R.gpu_proximity(st, np.ascontiguousarray(nd[t], np.int32), np.ascontiguousarray(small.nin[t], np.int32), p)
    assert np.array_equal(p.reshape(-1, order="F"), O.counts(small, nd, _lib.PROX_OOB_ASYM, _lib.LAYOUT_FULL))


def test_lowrank_reconstruct_and_fp16_entry(R):
    _say("ref api reconstruct_full_matrix")
    x, y, f = small_forest(_lib.DATA_OVERLAP_MIXTURE, 1200, 8, 3, 20, seed=27)
    nodextr = O.leaf_assign(f, x)
    n, rank = f.n, 24
    m = R.LowRankProximityMatrix(n, rank, _lib.PREC_BF16)
    for t in range(f.ntree):
        st, _ = _tree(f, t)
        m.add_tree(st, np.ascontiguousarray(nodextr[t], np.int32), np.ascontiguousarray(f.nin[t], np.int32))
    m.finalize()
    A, _, r = m.factors()
    # FP32 factors: the exact FP64 sums of their products
    P = m.reconstruct()
    assert np.allclose(P, A @ A.T, rtol=1e-12, atol=1e-13)
    assert np.array_equal(P, P.T)
    # stored BF16 factors: within the storage rounding (2^-8 relative per entry)
    Pb = m.reconstruct_level(_lib.PREC_BF16)
    assert np.abs(Pb - P).max() <= 2 ** -7 * np.abs(A).sum(axis=1).max() ** 2
    # the FP16-triangle entry refuses loudly (the triangle has no leaves)
    with pytest.raises(RuntimeError, match="add_tree_contribution"):
        m.add_fp16_triangle(np.zeros(n * (n + 1) // 2, np.uint16))
