"""CPU: the oracle and the host-side forest contract, pinned against the
reference's golden vectors, published known answers and closed forms."""
from __future__ import annotations

import numpy as np
import pytest

from oracle import oracle as O
from rfx_amd import _lib
from rfx_amd.classifier import detect_categorical
from rfx_amd.forest import Forest, make_dataset

from helpers import GOLDEN as GOLDEN_DIR, load_golden, small_forest

GOLDEN = ["gauss_n200.npz", "covtype_cat_n240.npz"]


def test_mt19937_known_answers():
    L = _lib.lib()
    # mt19937ar.c, init_genrand(5489): first outputs; std::mt19937 10000th output
    first = [L.rfx_mt19937_u32(5489, i) for i in range(5)]
    assert first == [3499211612, 581869302, 3890346734, 3586334585, 545404204]
    assert L.rfx_mt19937_u32(5489, 9999) == 4123659995


def test_bootstrap_closed_form():
    # nin_i = #{draws d : int(float(u_d) * n) == i}, u_d = MT19937 output / (2^32 - 1)
    L = _lib.lib()
    n, seed = 97, 12345 + 3
    nin = np.zeros(n, np.int32)
    assert L.rfx_bootstrap(n, seed, nin.ctypes.data) == 0
    u = np.array([L.rfx_mt19937_u32(seed, d) for d in range(n)], np.float64) / 4294967295.0
    idx = np.minimum((u.astype(np.float32) * np.float32(n)).astype(np.int32), n - 1)
    assert np.array_equal(nin, np.bincount(idx, minlength=n))
    assert nin.sum() == n


@pytest.mark.parametrize("name", GOLDEN)
def test_grower_matches_reference_forest(name):
    g = load_golden(name)
    x, y = g["x"], g["y"]
    f = Forest.grow(x, y, int(g["ntree"]), mtry=int(g["mtry"]), iseed=int(g["iseed"]))
    ref = g["forest"]
    assert np.array_equal(f.node_offset, ref.node_offset)
    assert np.array_equal(f.nodestatus, ref.nodestatus)
    assert np.array_equal(f.bestvar, ref.bestvar)
    assert np.array_equal(f.xbestsplit.view(np.uint32), ref.xbestsplit.view(np.uint32))
    assert np.array_equal(f.treemap, ref.treemap)
    assert np.array_equal(f.nin, ref.nin)
    assert np.array_equal(f.nout(2), g["nout"])


def test_categorical_detection_matches_fixture():
    g = load_golden("covtype_cat_n240.npz")
    cat = detect_categorical(g["x"], int(g["maxcat"]))
    assert np.array_equal(cat, g["cat"])
    assert (cat[:10] == 1).all() and (cat[10:] == 2).all()


@pytest.mark.parametrize("name", GOLDEN)
def test_oracle_leaf_assign_matches_reference(name):
    g = load_golden(name)
    cat = g["cat"] if int(g["maxcat"]) > 1 else None
    nd = O.leaf_assign(g["forest"], g["x"], cat)
    assert np.array_equal(nd, g["nodextr"])


@pytest.mark.parametrize("name", GOLDEN)
def test_oracle_v1_and_counts_match_reference(name):
    g = load_golden(name)
    f = g["forest"]
    prox = O.prox_v1(f, g["nodextr"])
    assert np.array_equal(prox, g["prox"])
    P = O.counts(f, g["nodextr"], O.OOB_ASYM, 0)
    assert np.array_equal(P.astype(np.float64), g["prox"])


@pytest.mark.parametrize("name", GOLDEN)
def test_oracle_finishprox_bitwise(name):
    g = load_golden(name)
    n = g["x"].shape[0]
    S = O.finishprox(n, g["nout"], g["prox"])
    assert np.array_equal(S.view(np.uint64), g["proxsym"].view(np.uint64))


def test_finishprox_closed_form():
    g = load_golden("gauss_n200.npz")
    n = g["x"].shape[0]
    P = g["prox"].reshape(n, n, order="F")
    nout = g["nout"].astype(np.float64)
    r = np.where(nout > 0, 1.0 / np.where(nout > 0, nout, 1.0), 1.0)
    a = (P + P.T) / 2.0
    S = 0.5 * (a * r[:, None] + a * r[None, :])
    np.fill_diagonal(S, 1.0)
    assert np.allclose(S, g["proxsym"].reshape(n, n, order="F"), rtol=0, atol=1e-15)


@pytest.mark.parametrize("variant", [O.ALL, O.OOB_ASYM, O.OOB_UPPER, O.OOB_ROW])
def test_bucket_counts_equal_brute_force(variant):
    x, y, f = small_forest(_lib.DATA_OVERLAP_MIXTURE, 150, 5, 3, 7, seed=2)
    nd = O.leaf_assign(f, x)
    a = O.counts(f, nd, variant, 0)
    b = O.counts_brute(f, nd, variant)
    assert np.array_equal(a, b)
    n = f.n
    if variant not in (O.OOB_ASYM, O.OOB_ROW):
        u = O.counts(f, nd, variant, 1)
        iu = np.triu_indices(n)
        assert np.array_equal(u, b.reshape(n, n, order="F")[iu])


def test_count_invariants():
    x, y, f = small_forest(_lib.DATA_OVERLAP_MIXTURE, 180, 6, 4, 9, seed=4)
    n, T = f.n, f.ntree
    nd = O.leaf_assign(f, x)
    C = O.counts(f, nd, O.ALL, 0).reshape(n, n, order="F").astype(np.int64)
    P = O.counts(f, nd, O.OOB_ASYM, 0).reshape(n, n, order="F").astype(np.int64)
    U = O.counts(f, nd, O.OOB_UPPER, 0).reshape(n, n, order="F").astype(np.int64)
    # numeric features: every sample reaches a leaf in every tree
    assert (np.diag(C) == T).all()
    assert (C == C.T).all()
    sizes = np.zeros(n, np.int64)
    total = 0
    oob_row = np.zeros(n, np.int64)
    for t in range(T):
        _, inv, cnt = np.unique(nd[t], return_inverse=True, return_counts=True)
        sizes += cnt[inv]
        total += int((cnt.astype(np.int64) ** 2).sum())
        inb = f.nin[t] > 0
        inb_per_leaf = np.bincount(inv, weights=inb, minlength=len(cnt))
        oob_row += np.where(~inb, inb_per_leaf[inv], 0).astype(np.int64)
    assert np.array_equal(C.sum(axis=1), sizes)
    assert C.sum() == total
    assert (np.diag(P) == 0).all()
    assert np.array_equal(P.sum(axis=1), oob_row)
    assert np.array_equal(U, P + P.T)


def test_oracle_leaf_assign_against_reference_fresh():
    if not O.ref_available():
        pytest.skip("reference build oracle/_ref not present")
    x, y = make_dataset(_lib.DATA_OVERLAP_MIXTURE, 300, 7, 3, 9)
    r = O.ref_run_forest(x, y, 6, 2, iseed=3, do_prox=True)
    f = Forest.from_arrays(300, 7, r["nnode"], r["nodestatus"], r["bestvar"], r["xbestsplit"], r["treemap"], r["nin"],
                           maxnode=r["maxnode"])
    nd = O.leaf_assign(f, x)
    assert np.array_equal(nd, r["nodextr"])
    assert np.array_equal(O.prox_v1(f, nd), r["prox"])
    g = Forest.grow(x, y, 6, mtry=2, iseed=3)
    assert np.array_equal(g.nodestatus, f.nodestatus) and np.array_equal(g.treemap, f.treemap)


def test_oracle_casewise_against_reference_fresh():
    # use_casewise: the reference's own cpu_proximity (dense branch, n <= 5000)
    # compiled from its sources, against the restatement, bit for bit
    if not O.ref_available():
        pytest.skip("reference build oracle/_ref not present")
    x, y = make_dataset(_lib.DATA_OVERLAP_MIXTURE, 300, 7, 3, 9)
    r = O.ref_run_forest(x, y, 6, 2, iseed=3, do_prox=True, casewise=True)
    f = Forest.from_arrays(300, 7, r["nnode"], r["nodestatus"], r["bestvar"], r["xbestsplit"], r["treemap"], r["nin"],
                           maxnode=r["maxnode"])
    nd = O.leaf_assign(f, x)
    pw = O.prox_v1(f, nd, casewise=True)
    assert np.array_equal(pw, r["prox"])
    assert not np.array_equal(pw, O.prox_v1(f, nd))  # the weights do change the result


def test_casewise_weights_sum_to_one_per_tree():
    # every leaf of a grown tree holds in-bag cases, and an out-of-bag case's
    # weights over its leaf's in-bag cases sum to 1: row i of the case-wise
    # matrix sums to the number of trees in which i is out of bag
    x, _, f = small_forest(_lib.DATA_GAUSS_MIXTURE, 150, 5, 3, 9, seed=6)
    nd = O.leaf_assign(f, x)
    pw = O.prox_v1(f, nd, casewise=True).reshape(f.n, f.n, order="F")
    nin = f.nin.reshape(f.ntree, f.n)
    oob_trees = np.sum(nin == 0, axis=0)
    assert np.allclose(pw.sum(axis=1), oob_trees, rtol=0, atol=1e-12)
    # rows of cases that are never out of bag stay zero
    assert np.all(pw[oob_trees == 0, :] == 0)


@pytest.mark.parametrize("name", GOLDEN)
def test_oracle_rfgap_matches_reference(name):
    # the reference's own cpu_proximity_rfgap compiled from its sources, on the
    # golden forests' nin and reached nodes (the categorical forest has cases
    # that stop on inner nodes), against the restatement, bit for bit
    if not O.ref_available():
        pytest.skip("reference build oracle/_ref not present")
    g = load_golden(name)
    got = O.prox_rfgap(g["nin"], g["nodextr"])
    assert np.array_equal(got, O.ref_proximity_rfgap(g["nin"], g["nodextr"]))


def _rfgap_by_definition(nin: np.ndarray, nodextr: np.ndarray) -> np.ndarray:
    # pGAP(i, j) = 1/|S_i| sum_{t in S_i} c_j(t) [j in J_i(t)] / |M_i(t)|
    # (Rhodes, Cutler and Moon, 2023), evaluated densely per tree, then
    # symmetrised; cases never out of bag keep a unit diagonal
    T, n = nin.shape
    p = np.zeros((n, n))
    for t in range(T):
        same = (nodextr[t][:, None] == nodextr[t][None, :]) & (nodextr[t][:, None] >= 0)
        W = same * (nin[t] > 0)[None, :] * nin[t][None, :].astype(np.float64)
        M = W.sum(axis=1)
        rows = (nin[t] == 0) & (M > 0)
        p[rows] += W[rows] / M[rows, None]
    S = (nin == 0).sum(axis=0)
    p[S > 0] /= S[S > 0, None]
    P = 0.5 * (p + p.T)
    P[np.diag_indices(n)] = np.where(S > 0, np.diag(p), 1.0)
    return P, S


@pytest.mark.parametrize("name", GOLDEN)
def test_rfgap_definition_and_invariants(name):
    g = load_golden(name)
    nin, nd = g["nin"], g["nodextr"]
    n = nin.shape[1]
    got = O.prox_rfgap(nin, nd).reshape(n, n, order="F")
    want, S = _rfgap_by_definition(nin, nd)
    assert np.allclose(got, want, rtol=0, atol=1e-14)
    assert np.array_equal(got, got.T)
    assert np.all(np.diag(got)[S == 0] == 1.0) and np.all(np.diag(got)[S > 0] == 0.0)
    # the symmetric average keeps the total: every out-of-bag case's weights
    # over a node with in-bag cases sum to 1, so the total is sum_i (share of
    # i's out-of-bag trees whose node holds in-bag cases) + #never out of bag
    share = np.zeros(n)
    for t in range(nin.shape[0]):
        has = np.zeros(nd[t].max() + 1, bool)
        has[nd[t][(nin[t] > 0) & (nd[t] >= 0)]] = True
        ok = (nin[t] == 0) & (nd[t] >= 0)
        ok[ok] &= has[nd[t][ok]]
        share += ok
    total = np.sum(share[S > 0] / S[S > 0]) + np.sum(S == 0)
    assert abs(got.sum() - total) < 1e-9


def test_outlier_scores_closed_forms():
    # diagonal proximities of powers of two keep every step exact: raw score
    # 1/d^2, divided by the (n/2)-th smallest (index 2 of 4, not 1)
    P = np.diag([1.0, 2.0, 4.0, 8.0])
    assert np.array_equal(O.outlier_scores(P), np.array([4.0, 1.0, 0.25, 0.0625], np.float32))
    # a row of zeros scores 0, and its 0 takes part in the median: sorted
    # raw scores [0, 1/64, 1/16, 1/4], median 1/16
    P = np.diag([0.0, 2.0, 4.0, 8.0])
    assert np.array_equal(O.outlier_scores(P), np.array([0.0, 4.0, 1.0, 0.25], np.float32))
    # a power-of-two scale of the matrix leaves the scores unchanged bit for bit
    rng = np.random.default_rng(3)
    A = rng.random((301, 301))
    S = 0.5 * (A + A.T)
    assert np.array_equal(O.outlier_scores(S), O.outlier_scores(4.0 * S))
    # against the float64 definition: 1 / sum_j P_ij^2 over the median
    raw = 1.0 / np.sum(S * S, axis=1)
    want = raw / np.sort(raw)[301 // 2]
    assert np.allclose(O.outlier_scores(S), want, rtol=2e-5, atol=0)


def test_mds_oracle_recovers_euclidean_configuration():
    # a proximity built from a Euclidean 3-D configuration: P = max - D  ->
    # classical MDS recovers the centred points up to an orthogonal map
    rng = np.random.default_rng(0)
    Z = rng.normal(size=(60, 3)) * np.array([5.0, 3.0, 1.5])
    Z -= Z.mean(axis=0)
    D = np.sqrt(((Z[:, None, :] - Z[None, :, :]) ** 2).sum(-1))
    P = D.max() - D
    X, w = O.classical_mds(P, 3)
    Dx = np.sqrt(((X[:, None, :] - X[None, :, :]) ** 2).sum(-1))
    assert np.allclose(Dx, D, atol=1e-8)
    G = Z @ Z.T
    assert np.allclose(np.sort(np.linalg.eigvalsh(G))[::-1][:3], w[:3], rtol=1e-10)


def test_embedding_error_metrics():
    rng = np.random.default_rng(1)
    X = rng.normal(size=(50, 3))
    e = O.embedding_errors(X * np.array([1, -1, 1]), X)
    assert e["rel_frobenius"] < 1e-14 and e["max_principal_angle"] < 1e-6
    Y = X.copy()
    Y[:, 2] = rng.normal(size=50)
    assert O.embedding_errors(Y, X)["max_principal_angle"] > 0.1


def test_oob_row_closed_forms():
    """OOB_ROW (cuda/rf_proximity.cu:32-67) against identities that fix it
    independently of the restatement: P4 = P1 + (OOB-OOB same-leaf counts),
    P4_ii = #{t : i out-of-bag on a terminal leaf}, and row sums
    sum_j P4_ij = sum over OOB trees of |leaf_t(i)| (i itself included)."""
    x, y, f = small_forest(_lib.DATA_OVERLAP_MIXTURE, 160, 5, 3, 9, seed=6)
    nd = O.leaf_assign(f, x)
    n, T = f.n, f.ntree
    P4 = O.counts(f, nd, O.OOB_ROW, 0).reshape(n, n, order="F").astype(np.int64)
    P1 = O.counts(f, nd, O.OOB_ASYM, 0).reshape(n, n, order="F").astype(np.int64)
    oo = np.zeros((n, n), np.int64)
    diag = np.zeros(n, np.int64)
    rows = np.zeros(n, np.int64)
    for t in range(T):
        b = f.node_offset[t]
        st = f.nodestatus[b:f.node_offset[t + 1]]
        leaf = nd[t]
        term = st[leaf] == -1
        oob = (f.nin[t] == 0) & term
        same = (leaf[:, None] == leaf[None, :]) & term[:, None] & term[None, :]
        oo += same & oob[:, None] & oob[None, :]
        diag += oob
        size = np.bincount(leaf[term], minlength=len(st))
        rows += np.where(oob, size[leaf], 0)
    assert np.array_equal(P4, P1 + oo)
    assert np.array_equal(np.diag(P4), diag)
    assert np.array_equal(P4.sum(axis=1), rows)


# ---------------------------------------------------------------------------
# Pins by the reference's RandomForest class, compute_outlier_scores and
# compute_mds_3d_cpu themselves (fixtures written by tests/golden/make_golden.py
# from oracle/_ref), and by closed forms of the MDS input.
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("name", GOLDEN)
def test_reference_class_pins_the_oracle_chain(name):
    # RandomForest::fit on the CPU (find_terminal_node for in-bag samples,
    # cpu_testreebag for out-of-bag ones, cpu_proximity per tree, then
    # cpu_finishprox): the oracle's leaf assignment (with the categorical
    # in-bag quirk), V1 counts and finishprox give the same matrix bit for bit
    g = load_golden(name)
    f = g["forest"]
    cat = g["cat"] if int(g["maxcat"]) > 1 else None
    nd = O.leaf_assign(f, g["x"], cat)
    P = O.counts(f, nd, O.OOB_ASYM, 0).astype(np.float64)
    S = O.finishprox(f.n, f.nout(2), P)
    assert np.array_equal(g["rf_nout"], f.nout(2))
    assert np.array_equal(S.view(np.uint64), g["rf_prox"].view(np.uint64))


@pytest.mark.parametrize("name", GOLDEN)
def test_reference_class_pins_casewise(name):
    # the class with use_casewise: weights nin(j) / sum of nin over the leaf,
    # summed tree by tree in FP64, then cpu_finishprox
    g = load_golden(name)
    f = g["forest"]
    cat = g["cat"] if int(g["maxcat"]) > 1 else None
    nd = O.leaf_assign(f, g["x"], cat)
    S = O.finishprox(f.n, f.nout(2), O.prox_v1(f, nd, casewise=True))
    assert np.array_equal(S.view(np.uint64), g["rf_prox_casewise"].view(np.uint64))


def test_reference_outlier_scores_bitwise():
    # RandomForest::compute_outlier_scores of an unsupervised forest's
    # finished proximity, bit for bit
    with np.load(GOLDEN_DIR / "unsup_n160.npz", allow_pickle=False) as z:
        prox, want = z["prox"], z["outlier"]
    n = want.size
    got = O.outlier_scores(prox.reshape(n, n))
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))


@pytest.mark.parametrize("name", GOLDEN)
def test_mds_inputs_pinned_by_reference_mds(name):
    # rf::compute_mds_3d_cpu (dsyevd) on the classifier's inputs against the
    # oracle's renormalisation + FP64 classical MDS: the coordinates agree to
    # rounding, so the renormalisation, the float32 widening, D = max - M,
    # the squaring and centring all match
    g = load_golden(name)
    n = g["x"].shape[0]
    S = g["rf_prox"].reshape(n, n, order="F")
    X, w = O.classical_mds(O.mds_input_reference(S, g["rf_nout"]))
    assert w[2] - w[3] > 1e-3 * w[0]
    assert O.embedding_errors(X, g["mds_ref"])["rel_frobenius"] < 1e-10
    P = g["rfgap"].reshape(n, n, order="F")
    X, w = O.classical_mds(O.mds_input_rfgap(P, g["rf_nout"]))
    assert w[2] - w[3] > 1e-3 * w[0]
    assert O.embedding_errors(X, g["mds_rfgap"])["rel_frobenius"] < 1e-10


def test_mds_input_reference_closed_forms():
    # dyadic entries survive the float32 widening exactly
    rng = np.random.default_rng(5)
    n = 6
    A = rng.integers(0, 64, size=(n, n)) / 64.0
    S = 0.5 * (A + A.T)
    np.fill_diagonal(S, 1.0)
    # uniform nout = k: M = S / k off the diagonal, 1 on it
    M = O.mds_input_reference(S, np.full(n, 4, np.int32))
    want = S / 4.0
    np.fill_diagonal(want, 1.0)
    assert np.array_equal(M, want)
    # unequal nout and a never-out-of-bag row (nout 0: not divided)
    S3 = np.array([[1.0, 0.5, 0.25], [0.5, 1.0, 0.75], [0.25, 0.75, 1.0]])
    M3 = O.mds_input_reference(S3, np.array([2, 4, 0], np.int32))
    assert np.array_equal(M3, np.array([[1.0, 0.5 * (0.25 + 0.125), 0.5 * (0.125 + 0.25)],
                                        [0.5 * (0.25 + 0.125), 1.0, 0.5 * (0.1875 + 0.75)],
                                        [0.5 * (0.125 + 0.25), 0.5 * (0.1875 + 0.75), 1.0]]))
    # a non-dyadic entry is rounded to float32 before the division
    S4 = np.array([[1.0, 0.1], [0.1, 1.0]])
    M4 = O.mds_input_reference(S4, np.array([1, 1], np.int32))
    assert M4[0, 1] == float(np.float32(0.1)) and M4[0, 1] != 0.1


def test_mds_input_rfgap_closed_forms():
    # RF-GAP keeps the divided diagonal (no unit diagonal is imposed) and
    # leaves a zero-nout row undivided
    P = np.array([[2.0, 0.5, 0.25], [0.5, 4.0, 0.75], [0.25, 0.75, 1.0]])
    M = O.mds_input_rfgap(P, np.array([2, 4, 0], np.int32))
    R = np.array([[1.0, 0.25, 0.125], [0.125, 1.0, 0.1875], [0.25, 0.75, 1.0]])
    want = 0.5 * (R + R.T)
    np.fill_diagonal(want, [1.0, 1.0, 1.0])
    assert np.array_equal(M, want)
    M2 = O.mds_input_rfgap(P, np.array([4, 4, 4], np.int32))
    assert np.array_equal(np.diag(M2), np.diag(P) / 4.0)


@pytest.mark.skipif(not O.ref_available(), reason="oracle/_ref not built (needs the reference sources)")
def test_reference_mds_fresh_inputs():
    # compute_mds_3d_cpu on inputs no fixture holds: a random symmetric
    # proximity with random nout, both forms
    rng = np.random.default_rng(11)
    n = 90
    A = rng.random((n, n))
    S = 0.5 * (A + A.T)
    np.fill_diagonal(S, 1.0)
    nout = rng.integers(0, 9, size=n).astype(np.int32)
    for rfgap in (False, True):
        Xr = O.ref_mds_3d(S.astype(np.float32).astype(np.float64).ravel(order="F"), nout, rfgap)
        M = O.mds_input_rfgap(S, nout) if rfgap else O.mds_input_reference(S, nout)
        X, w = O.classical_mds(M)
        assert O.embedding_errors(X, Xr)["rel_frobenius"] < 1e-8 * max(1.0, w[0] / (w[2] - w[3]))


def test_v1_on_a_sample_subset_of_larger_trees():
    # trees grown on 3,000 samples assigned only 150 of them: more leaves
    # than samples; the V1 port still equals the bucket-based counts
    x, y, f = small_forest(_lib.DATA_OVERLAP_MIXTURE, 3000, 6, 3, 4, seed=8)
    idx = np.sort(np.random.default_rng(1).choice(3000, 150, replace=False))
    s = f.samples(idx)
    assert min(int((s.nodestatus[s.node_offset[t]:s.node_offset[t + 1]] == -1).sum()) for t in range(4)) > 150
    nd = O.leaf_assign(s, np.ascontiguousarray(x[idx]))
    assert np.array_equal(O.prox_v1(s, nd), O.counts(s, nd, O.OOB_ASYM, 0).astype(np.float64))


def test_counts_rows_equal_full_matrices():
    # the row-restricted oracle against the full ALL and OOB_UPPER matrices
    # (themselves pinned by the brute force above), rows in any order
    x, y, f = small_forest(_lib.DATA_OVERLAP_MIXTURE, 700, 6, 3, 9, seed=4)
    nd = O.leaf_assign(f, x)
    n = f.n
    C = O.counts(f, nd, O.ALL, 0).reshape(n, n, order="F")
    U = O.counts(f, nd, O.OOB_UPPER, 0).reshape(n, n, order="F")
    U = np.triu(U) + np.triu(U, 1).T
    rows = np.array([699, 0, 350, 3, 350])
    a, o = O.counts_rows(f, nd, rows)
    assert np.array_equal(a, C[rows]) and np.array_equal(o, U[rows])


def _same_forest(a, b) -> bool:
    return (all(np.array_equal(getattr(a, k), getattr(b, k))
                for k in ("node_offset", "nodestatus", "bestvar", "treemap", "nodeclass", "nin"))
            and np.array_equal(a.xbestsplit.view(np.uint32), b.xbestsplit.view(np.uint32)))


@pytest.mark.parametrize("kind,n,p,ncl,T,mtry,minnd", [
    (_lib.DATA_SUBSPACE_MIXTURE, 20000, 32, 10, 3, 6, 5),
    (_lib.DATA_COVTYPE_LIKE, 8000, 54, 7, 3, 7, 1),
])
def test_presorted_grower_matches_reference_at_scale(kind, n, p, ncl, T, mtry, minnd):
    """The presorted grower (lane scans, deferred peel partitions) against the
    compiled reference's cpu_growtree on forests with peel chains and ties."""
    if not O.ref_available():
        pytest.skip("reference build oracle/_ref not present")
    x, y = make_dataset(kind, n, p, ncl, 0)
    r = O.ref_run_forest(x, y, T, mtry, minndsize=minnd, iseed=0, do_prox=False)
    ref = Forest.from_arrays(n, p, r["nnode"], r["nodestatus"], r["bestvar"], r["xbestsplit"], r["treemap"], r["nin"],
                             maxnode=r["maxnode"])
    got = Forest.grow(x, y, T, mtry=mtry, minndsize=minnd, iseed=0)
    ref.nodeclass = got.nodeclass  # the reference driver does not export node classes
    assert _same_forest(got, ref)


@pytest.mark.parametrize("kind,n,p,ncl,T,mtry,minnd", [
    (_lib.DATA_SUBSPACE_MIXTURE, 100000, 32, 10, 4, 6, 5),
    (_lib.DATA_COVTYPE_LIKE, 20000, 54, 7, 8, 7, 1),
    (_lib.DATA_GAUSS_MIXTURE, 3000, 20, 3, 8, 4, 1),
    (_lib.DATA_OVERLAP_MIXTURE, 5000, 10, 4, 8, 3, 2),
])
def test_presorted_grower_equals_per_node_sort_grower(kind, n, p, ncl, T, mtry, minnd):
    """Presorted columns and per-node sorts grow the same trees, node classes included."""
    x, y = make_dataset(kind, n, p, ncl, 1)
    a = Forest.grow(x, y, T, mtry=mtry, minndsize=minnd, iseed=5)
    b = Forest.grow(x, y, T, mtry=mtry, minndsize=minnd, iseed=5, plain=True)
    assert _same_forest(a, b)
