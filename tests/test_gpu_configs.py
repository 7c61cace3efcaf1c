"""GPU parity at the full sizes of the benchmark configurations.

  * n = 200,000, T = 1000 (covtype-like, the metric's configuration): the
    fused ALL + OOB_UPPER upper-triangle counts of the product path checked
    whole by size-independent properties (diagonals, totals against the
    leaf sizes, full row sums) and on seeded rows against the oracle's
    row-restricted counts; leaf assignment of every sample against the
    oracle's.
  * n = 20,000, T = 500: the 3-D MDS with the FP32 operator against an FP64
    host solution of the same centred matrix (eigenvalues, eigen-residuals).
  * the 64-forest batch (instance i from seed i): every instance's OOB upper
    counts against the closed-form total from the oracle's leaf assignment,
    and the smallest instances against the oracle's whole matrices.
"""
from __future__ import annotations

import numpy as np
import pytest
import torch

from oracle import oracle as O
from rfx_amd import _lib
from rfx_amd.forest import Forest, default_mtry, make_dataset
from rfx_amd.proximity import Accumulator, DeviceForest, mds_dense

from helpers import centred, eig_residuals

pytestmark = pytest.mark.gpu


def _say(msg: str) -> None:
    print(msg, flush=True)


def _upper_off(i, n):
    return i * n - i * (i - 1) // 2


def _full_rows(packed: torch.Tensor, n: int, rows) -> np.ndarray:
    """Full rows of the symmetric matrix held as a packed upper triangle."""
    out = np.empty((len(rows), n), np.int32)
    for k, i in enumerate(rows):
        i = int(i)
        right = packed[_upper_off(i, n):_upper_off(i, n) + n - i]
        j = torch.arange(i, device=packed.device, dtype=torch.int64)
        left = packed[j * n - j * (j - 1) // 2 + (i - j)]
        out[k] = torch.cat([left, right]).cpu().numpy()
    return out


@pytest.fixture(scope="module")
def c3(gpu):
    _say("c3: grow n=200000 T=1000")
    x, y = make_dataset(_lib.DATA_COVTYPE_LIKE, 200000, 54, 7, 0)
    f = Forest.grow(x, y, 1000, mtry=7, iseed=0)
    df = DeviceForest(f, x)
    keys, nodextr = df.leaf_assign(want_nodextr=True)
    acc = Accumulator(f.n, f.ntree, f.total_nodes)
    _say("c3: counts")
    a, o = acc.accumulate_pair(keys, _lib.LAYOUT_UPPER, _lib.ALGO_GATHER)
    torch.cuda.synchronize()
    sq = acc.pair_stats(keys)
    del acc
    _say("c3: oracle leaf assignment of every sample")
    nd = O.leaf_assign(f, x)
    d = dict(x=x, f=f, keys=keys, nodextr=nodextr, nd=nd, a=a, o=o, sq=sq)
    yield d
    d.clear()
    del keys, nodextr, a, o
    torch.cuda.empty_cache()


def test_c3_leaf_assignment_every_sample(c3):
    assert np.array_equal(c3["nodextr"].cpu().numpy(), c3["nd"])


def test_c3_counts_whole_matrix_properties(c3):
    f, a, o = c3["f"], c3["a"], c3["o"]
    n, T = f.n, f.ntree
    _say("c3: diagonals and totals")
    ii = torch.arange(n, device=a.device, dtype=torch.int64)
    diag = ii * n - ii * (ii - 1) // 2
    assert bool((a[diag] == T).all())  # all-numeric forest: every sample ends on a leaf
    assert bool((o[diag] == 0).all())
    # the leaf sizes from the oracle's leaf assignment of every sample
    nd = c3["nd"]
    sq = 0
    oob_pairs = 0
    rowsum = np.zeros(n, np.int64)
    for t in range(T):
        leaf = nd[t].astype(np.int64)
        bag = f.nin[t] > 0
        size = np.bincount(leaf)
        sq += int(np.sum(size.astype(np.int64) ** 2))
        oob_pairs += int(np.sum(np.bincount(leaf[bag], minlength=len(size)).astype(np.int64) *
                                np.bincount(leaf[~bag], minlength=len(size))))
        rowsum += size[leaf]
    assert sq == c3["sq"]
    chunk = 1 << 28
    tot_a = sum(int(a[k:k + chunk].sum(dtype=torch.int64)) for k in range(0, a.numel(), chunk))
    tot_o = sum(int(o[k:k + chunk].sum(dtype=torch.int64)) for k in range(0, o.numel(), chunk))
    assert tot_a == (sq + T * n) // 2  # off-diagonal pairs once, the diagonal C_ii = T
    assert tot_o == oob_pairs
    # full row sums of ALL on seeded rows: sum_j C_ij = sum_t |leaf_t(i)|
    rows = np.random.default_rng(5).choice(n, 24, replace=False)
    full = _full_rows(a, n, rows)
    assert np.array_equal(full.sum(axis=1, dtype=np.int64), rowsum[rows])


def test_c3_counts_seeded_rows_against_oracle(c3):
    f, a, o = c3["f"], c3["a"], c3["o"]
    n = f.n
    rng = np.random.default_rng(11)
    rows = np.concatenate([[0, 1, 16383, 16384, 32767, 32768, 100000, n - 2, n - 1],
                           rng.choice(n, 23, replace=False)])
    _say(f"c3: {len(rows)} rows against the oracle")
    want_a, want_o = O.counts_rows(f, c3["nd"], rows)
    assert np.array_equal(_full_rows(a, n, rows), want_a)
    assert np.array_equal(_full_rows(o, n, rows), want_o)


def test_c2_mds_fp32_against_fp64_host(gpu):
    _say("c2 mds: grow n=20000 T=500")
    x, y = make_dataset(_lib.DATA_COVTYPE_LIKE, 20000, 54, 7, 0)
    f = Forest.grow(x, y, 500, mtry=7, iseed=0)
    df = DeviceForest(f, x)
    keys, _ = df.leaf_assign()
    acc = Accumulator(f.n, f.ntree, f.total_nodes)
    C = acc.accumulate(keys, _lib.PROX_ALL, _lib.LAYOUT_UPPER, _lib.ALGO_GATHER)
    X, ev, it = mds_dense(C, _lib.LAYOUT_UPPER, f.n, _lib.MDS_COUNTS, 1.0 / f.ntree, precision=_lib.PREC_FP32)
    X = X.cpu().numpy()
    n = f.n
    Cu = C.cpu().numpy()
    del C
    M = np.zeros((n, n), np.float64)
    M[np.triu_indices(n)] = Cu
    M = M + np.triu(M, 1).T
    M /= f.ntree
    B = centred(M)
    del M
    from scipy.sparse.linalg import eigsh

    _say("c2 mds: FP64 eigsh on the host")
    w, V = eigsh(B, k=4, which="LA", tol=1e-12)
    w = w[::-1]
    lam = ev[:3]
    assert np.allclose(lam, w[:3], rtol=1e-5, atol=0)
    assert np.all(eig_residuals(B, X, lam) < 1e-3)
    if w[2] / w[3] > 1.02:
        Xref = V[:, ::-1][:, :3] * np.sqrt(w[:3])
        assert O.embedding_errors(X, Xref)["rel_frobenius"] < 2e-3


def _c5_instances():
    out = []
    for i in range(64):
        r = np.random.default_rng(i)
        n = int(round(np.exp(r.uniform(np.log(1000), np.log(50000)))))
        p = int(r.integers(8, 65))
        T = int(r.integers(50, 501))
        ncl = int(r.integers(2, 11))
        out.append((i, n, p, T, ncl))
    return out


def test_c5_batch_all_instances(gpu):
    from rfx_amd.batch import ForestBatch

    _say("c5: grow 64 forests")
    forests, xs = [], []
    for i, n, p, T, ncl in _c5_instances():
        x, y = make_dataset(_lib.DATA_GAUSS_MIXTURE, n, p, ncl, i)
        forests.append(Forest.grow(x, y, T, mtry=default_mtry(p), iseed=i))
        xs.append(x)
    dfs = [DeviceForest(f, x) for f, x in zip(forests, xs)]
    b = ForestBatch(dfs, "oob_upper")
    b.autotune(reps=1)
    b.run(use_graph=True)
    b.run(use_graph=True)
    torch.cuda.synchronize()
    _say("c5: check every instance")
    small = set(sorted(range(64), key=lambda k: forests[k].n)[:3])
    for k, (f, x) in enumerate(zip(forests, xs)):
        nd = O.leaf_assign(f, x)
        want = 0
        for t in range(f.ntree):
            leaf = nd[t].astype(np.int64)
            ok = leaf >= 0
            bag = f.nin[t] > 0
            m = int(leaf.max()) + 1
            want += int(np.sum(np.bincount(leaf[ok & bag], minlength=m).astype(np.int64) *
                               np.bincount(leaf[ok & ~bag], minlength=m)))
        out = b.out[k]
        assert int(out.sum(dtype=torch.int64)) == want, k
        if k in small:
            assert np.array_equal(out.cpu().numpy(), O.counts(f, nd, _lib.PROX_OOB_UPPER, _lib.LAYOUT_UPPER)), k


def test_c4_shape_rank128_factor_mds(gpu):
    """The FP8 scale-up workload's data and forest options (16-component
    subspace mixture, p = 32, 10 classes, mtry 6, minndsize 5: deep peeling
    trees) at n = 200,000 and 24 trees: leaf assignment of every sample
    against the oracle, rank-128 factors from the leaf buckets, and the factor
    MDS from FP8-e4m3 / BF16 / FP32 storage on the matrix cores with FP32
    refinement against the FP64 MDS of the same factors."""
    from rfx_amd.proximity import lowrank_factors, mds_factors, mds_factors_mixed, quantize_factors

    _say("c4 shape: grow n=200000 T=24")
    x, y = make_dataset(_lib.DATA_SUBSPACE_MIXTURE, 200000, 32, 10, 0)
    f = Forest.grow(x, y, 24, mtry=6, minndsize=5, iseed=0)
    df = DeviceForest(f, x)
    keys, nodextr = df.leaf_assign(want_nodextr=True)
    _say("c4 shape: oracle leaf assignment")
    assert np.array_equal(nodextr.cpu().numpy(), O.leaf_assign(f, x))
    _say("c4 shape: rank-128 factors")
    U, eig = lowrank_factors(keys, f, rank=128, oversample=16, power_iters=2)
    assert U.shape == (200000, 128) and np.all(np.diff(eig[:128]) <= 1e-9 * eig[0])
    U32, s32 = quantize_factors(U, _lib.PREC_FP32)
    Xref, eref, _ = mds_factors(U32, s32, _lib.PREC_FP32, precision=_lib.PREC_FP64)
    Xref = Xref.cpu().numpy()
    for name, fmt in (("fp8", _lib.PREC_FP8), ("bf16", _lib.PREC_BF16), ("fp32", _lib.PREC_FP32)):
        Uq, sc = quantize_factors(U, fmt)
        X1, _, it1 = mds_factors_mixed(Uq, sc, fmt, U32, refine_max=1)
        X, e, it = mds_factors_mixed(Uq, sc, fmt, U32)
        err1 = O.embedding_errors(X1.cpu().numpy(), Xref)["rel_frobenius"]
        err = O.embedding_errors(X.cpu().numpy(), Xref)["rel_frobenius"]
        _say(f"c4 shape: {name} one sweep {err1:.2e} {it1}, refined {err:.2e} {it}")
        assert np.allclose(e[:3], eref[:3], rtol=1e-4)
        assert err < 1e-3
