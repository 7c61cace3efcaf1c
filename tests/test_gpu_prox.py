"""GPU parity: leaf assignment, co-occurrence counts and finishprox through
the C ABI, bit-exact against the oracle and the reference's golden vectors."""
from __future__ import annotations

import numpy as np
import pytest
import torch

from oracle import oracle as O
from rfx_amd import _lib
from rfx_amd.forest import Forest
from rfx_amd.proximity import Accumulator, DeviceForest, finishprox, upper_size

from helpers import keys_from_nodextr, load_golden, small_forest

pytestmark = pytest.mark.gpu

ALGOS = [_lib.ALGO_DENSE, _lib.ALGO_BUCKET, _lib.ALGO_GATHER]


def _say(msg: str) -> None:
    print(msg, flush=True)


def run_keys(forest: Forest, x: np.ndarray, cat=None):
    df = DeviceForest(forest, x, cat)
    keys, nodextr = df.leaf_assign(want_nodextr=True)
    torch.cuda.synchronize()
    return df, keys, nodextr


def expect(forest, nodextr, variant, layout):
    return O.counts(forest, nodextr, variant, layout)


@pytest.mark.parametrize("name", ["gauss_n200.npz", "covtype_cat_n240.npz"])
def test_leaf_assign_matches_reference_golden(gpu, name):
    _say(f"leaf_assign golden {name}")
    g = load_golden(name)
    cat = g["cat"] if int(g["maxcat"]) > 1 else None
    _, keys, nodextr = run_keys(g["forest"], g["x"], cat)
    assert np.array_equal(nodextr.cpu().numpy(), g["nodextr"])
    assert np.array_equal(keys.cpu().numpy(), keys_from_nodextr(g["forest"], g["nodextr"]))


@pytest.mark.parametrize("name", ["gauss_n200.npz", "covtype_cat_n240.npz"])
@pytest.mark.parametrize("algo", ALGOS)
def test_oob_asym_and_finishprox_match_reference(gpu, name, algo):
    _say(f"oob_asym+finishprox golden {name} algo {algo}")
    g = load_golden(name)
    f = g["forest"]
    n = f.n
    cat = g["cat"] if int(g["maxcat"]) > 1 else None
    _, keys, _ = run_keys(f, g["x"], cat)
    acc = Accumulator(n, f.ntree, f.total_nodes)
    P = acc.accumulate(keys, _lib.PROX_OOB_ASYM, _lib.LAYOUT_FULL, algo)
    assert np.array_equal(P.cpu().numpy().astype(np.float64), g["prox"])
    nout = torch.from_numpy(g["nout"]).to(gpu)
    S = finishprox(P, _lib.LAYOUT_FULL, nout, n, _lib.LAYOUT_FULL)
    assert np.array_equal(S.cpu().numpy().view(np.uint64), g["proxsym"].view(np.uint64))
    U = acc.accumulate(keys, _lib.PROX_OOB_UPPER, _lib.LAYOUT_UPPER, algo)
    S_up = finishprox(U, _lib.LAYOUT_UPPER, nout, n, _lib.LAYOUT_UPPER)
    iu = np.triu_indices(n)
    assert np.array_equal(S_up.cpu().numpy().view(np.uint64),
                          g["proxsym"].reshape(n, n, order="F")[iu].view(np.uint64))
    S_full = finishprox(U, _lib.LAYOUT_UPPER, nout, n, _lib.LAYOUT_FULL)
    assert np.array_equal(S_full.cpu().numpy().view(np.uint64), g["proxsym"].view(np.uint64))


CASES = [
    # (kind, n, p, classes, trees, seed): ragged n (not a multiple of the 128 tile)
    (_lib.DATA_OVERLAP_MIXTURE, 1, 4, 2, 3, 0),
    (_lib.DATA_OVERLAP_MIXTURE, 129, 5, 3, 1, 1),
    (_lib.DATA_OVERLAP_MIXTURE, 700, 9, 4, 37, 2),
    (_lib.DATA_GAUSS_MIXTURE, 2000, 20, 3, 100, 0),
    (_lib.DATA_COVTYPE_LIKE, 3001, 54, 7, 40, 3),
]


@pytest.mark.parametrize("case", CASES, ids=lambda c: f"k{c[0]}_n{c[1]}_T{c[4]}")
def test_counts_all_variants_layouts_algos(gpu, case):
    kind, n, p, ncl, T, seed = case
    _say(f"counts case {case}")
    if n == 1:
        x = np.zeros((1, p), np.float32)
        y = np.zeros(1, np.int32)
        f = Forest.grow(x, y, T, mtry=1)
    else:
        x, y, f = small_forest(kind, n, p, ncl, T, seed)
    df, keys, nodextr = run_keys(f, x)
    nd_ref = O.leaf_assign(f, x)
    assert np.array_equal(nodextr.cpu().numpy(), nd_ref)
    acc = Accumulator(n, T, f.total_nodes)
    for variant, layout in [(_lib.PROX_ALL, _lib.LAYOUT_FULL), (_lib.PROX_ALL, _lib.LAYOUT_UPPER),
                            (_lib.PROX_OOB_ASYM, _lib.LAYOUT_FULL), (_lib.PROX_OOB_UPPER, _lib.LAYOUT_FULL),
                            (_lib.PROX_OOB_UPPER, _lib.LAYOUT_UPPER), (_lib.PROX_OOB_ROW, _lib.LAYOUT_FULL)]:
        want = expect(f, nd_ref, variant, layout)
        for algo in ALGOS:
            got = acc.accumulate(keys, variant, layout, algo).cpu().numpy()
            assert np.array_equal(got, want), f"variant {variant} layout {layout} algo {algo}"
        # accumulate mode adds on top
        base = torch.from_numpy(want.copy()).to(gpu)
        acc.accumulate(keys, variant, layout, _lib.ALGO_DENSE, out=base, accumulate=True)
        assert np.array_equal(base.cpu().numpy(), 2 * want)
    for layout in (_lib.LAYOUT_FULL, _lib.LAYOUT_UPPER):
        wa = expect(f, nd_ref, _lib.PROX_ALL, layout)
        wo = expect(f, nd_ref, _lib.PROX_OOB_UPPER, layout)
        for algo in ALGOS:
            a, o = acc.accumulate_pair(keys, layout, algo)
            assert np.array_equal(a.cpu().numpy(), wa) and np.array_equal(o.cpu().numpy(), wo)
    auto = acc.accumulate(keys, _lib.PROX_ALL, _lib.LAYOUT_UPPER, _lib.ALGO_AUTO).cpu().numpy()
    assert np.array_equal(auto, expect(f, nd_ref, _lib.PROX_ALL, _lib.LAYOUT_UPPER))
    # buckets built once, then several counting calls on them
    acc.build_buckets(keys)
    for algo in (_lib.ALGO_GATHER, _lib.ALGO_BUCKET):
        a, o = acc.accumulate_pair(keys, _lib.LAYOUT_UPPER, algo | _lib.ALGO_PREBUILT)
        assert np.array_equal(a.cpu().numpy(), expect(f, nd_ref, _lib.PROX_ALL, _lib.LAYOUT_UPPER))
        assert np.array_equal(o.cpu().numpy(), expect(f, nd_ref, _lib.PROX_OOB_UPPER, _lib.LAYOUT_UPPER))
        P = acc.accumulate(keys, _lib.PROX_OOB_ASYM, _lib.LAYOUT_FULL, algo | _lib.ALGO_PREBUILT)
        assert np.array_equal(P.cpu().numpy(), expect(f, nd_ref, _lib.PROX_OOB_ASYM, _lib.LAYOUT_FULL))


def test_rows_match_full(gpu):
    _say("rows")
    x, y, f = small_forest(_lib.DATA_COVTYPE_LIKE, 1500, 54, 7, 25, 5)
    _, keys, _ = run_keys(f, x)
    nd = O.leaf_assign(f, x)
    acc = Accumulator(f.n, f.ntree, f.total_nodes)
    for v in (_lib.PROX_ALL, _lib.PROX_OOB_ASYM, _lib.PROX_OOB_UPPER, _lib.PROX_OOB_ROW):
        full = expect(f, nd, v, _lib.LAYOUT_FULL).reshape(f.n, f.n, order="F")
        r = acc.rows(keys, v, 700, 777).cpu().numpy()
        assert np.array_equal(r, full[700:777])


def test_pair_stats(gpu):
    x, y, f = small_forest(_lib.DATA_OVERLAP_MIXTURE, 900, 6, 3, 12, 6)
    _, keys, _ = run_keys(f, x)
    nd = O.leaf_assign(f, x)
    acc = Accumulator(f.n, f.ntree, f.total_nodes)
    want = sum(int((np.unique(nd[t], return_counts=True)[1].astype(np.int64) ** 2).sum()) for t in range(f.ntree))
    assert acc.pair_stats(keys) == want


def test_oob_depth_cap_and_single_leaf_trees(gpu):
    """A hand-built chain tree 1200 splits deep: out-of-bag samples stop after
    999 descents (node 0, no pair), in-bag ones reach the leaf; plus a
    single-leaf tree where every sample shares the root leaf."""
    _say("depth cap")
    depth = 1200
    nn = 2 * depth + 1
    st = np.full(nn, -1, np.int32)
    bv = np.zeros(nn, np.int32)
    xs = np.zeros(nn, np.float32)
    tm = np.zeros(2 * nn, np.int32)
    # node k (k = 0, 2, 4, ...) splits on x <= -1 - k: left is a leaf, right continues
    node = 0
    for d in range(depth):
        st[node] = 1
        xs[node] = -1.0 - d
        tm[2 * node] = node + 1
        tm[2 * node + 1] = node + 2
        node += 2
    n = 6
    x = np.full((n, 1), 5.0, np.float32)  # always goes right: reaches the deepest leaf
    x[5, 0] = -2.5  # leaves at depth 2
    nin = np.array([[1, 0, 1, 0, 0, 1], [1, 0, 0, 1, 0, 1]], np.int32)
    st2 = np.array([-1], np.int32)
    f = Forest.from_arrays(n, 1, np.array([nn, 1]), np.concatenate([st, st2]), np.concatenate([bv, [0]]),
                           np.concatenate([xs, [0]]), np.concatenate([tm, [0, 0]]), nin)
    _, keys, nodextr = run_keys(f, x)
    nd_ref = O.leaf_assign(f, x)
    assert np.array_equal(nodextr.cpu().numpy(), nd_ref)
    assert nd_ref[0, 1] == 0 and nd_ref[0, 0] == 2 * depth  # OOB capped, in-bag reaches the end
    acc = Accumulator(n, 2, f.total_nodes)
    for v, lay in [(_lib.PROX_ALL, _lib.LAYOUT_FULL), (_lib.PROX_OOB_ASYM, _lib.LAYOUT_FULL),
                   (_lib.PROX_OOB_UPPER, _lib.LAYOUT_UPPER), (_lib.PROX_OOB_ROW, _lib.LAYOUT_FULL)]:
        want = expect(f, nd_ref, v, lay)
        for algo in ALGOS:
            assert np.array_equal(acc.accumulate(keys, v, lay, algo).cpu().numpy(), want)


def test_full_size_properties_c2_shape(gpu):
    """n = 20,000, T = 500 (the covtype-like workload): the whole ALL and OOB
    upper matrices of all three algorithms against the oracle's, plus
    invariants (ALL diagonal = T, total = sum over leaves of |leaf|^2) and the
    dense row kernel on a block of rows."""
    _say("c2 full-size properties")
    x, y, f = small_forest(_lib.DATA_COVTYPE_LIKE, 20000, 54, 7, 500, 0, mtry=7)
    df, keys, _ = run_keys(f, x)
    acc = Accumulator(f.n, f.ntree, f.total_nodes)
    a_b, o_b = acc.accumulate_pair(keys, _lib.LAYOUT_UPPER, _lib.ALGO_BUCKET)
    a_d, o_d = acc.accumulate_pair(keys, _lib.LAYOUT_UPPER, _lib.ALGO_DENSE)
    assert torch.equal(a_b, a_d) and torch.equal(o_b, o_d)
    del a_d, o_d
    a_g, o_g = acc.accumulate_pair(keys, _lib.LAYOUT_UPPER, _lib.ALGO_GATHER)
    assert torch.equal(a_b, a_g) and torch.equal(o_b, o_g)
    del a_g, o_g
    n = f.n
    diag = torch.from_numpy(np.array([n * i - i * (i - 1) // 2 for i in range(n)], np.int64)).to(gpu)
    assert bool((a_b[diag] == f.ntree).all())
    sq = acc.pair_stats(keys)
    total = 2 * int(a_b.sum(dtype=torch.int64).item()) - f.ntree * n
    assert total == sq
    nd = O.leaf_assign(f, x)
    assert np.array_equal(o_b.cpu().numpy(), O.counts(f, nd, _lib.PROX_OOB_UPPER, _lib.LAYOUT_UPPER))
    assert np.array_equal(a_b.cpu().numpy(), O.counts(f, nd, _lib.PROX_ALL, _lib.LAYOUT_UPPER))
    rows = acc.rows(keys, _lib.PROX_OOB_UPPER, 12345, 12345 + 64).cpu().numpy()
    for k, i in enumerate(range(12345, 12345 + 64)):
        b = i * n - i * (i - 1) // 2
        assert np.array_equal(o_b[b:b + n - i].cpu().numpy(), rows[k, i:])


def test_long_rows_walk_kernel(gpu):
    """n above the line-kernel limit (32,768): the row-walk gather streams
    each row through LDS chunks; UPPER ALL/OOB and the FULL OOB_ASYM layout
    against the oracle, and the same through the global-atomic form."""
    _say("walk kernel n=33000")
    x, y, f = small_forest(_lib.DATA_COVTYPE_LIKE, 33000, 54, 7, 12, 8, mtry=7)
    df, keys, _ = run_keys(f, x)
    nd = O.leaf_assign(f, x)
    acc = Accumulator(f.n, f.ntree, f.total_nodes)
    a, o = acc.accumulate_pair(keys, _lib.LAYOUT_UPPER, _lib.ALGO_GATHER)
    assert np.array_equal(a.cpu().numpy(), expect(f, nd, _lib.PROX_ALL, _lib.LAYOUT_UPPER))
    want_o = expect(f, nd, _lib.PROX_OOB_UPPER, _lib.LAYOUT_UPPER)
    assert np.array_equal(o.cpu().numpy(), want_o)
    del a
    b = acc.accumulate(keys, _lib.PROX_OOB_UPPER, _lib.LAYOUT_UPPER, _lib.ALGO_BUCKET)
    assert torch.equal(b, o)
    del b, o, want_o
    P = acc.accumulate(keys, _lib.PROX_OOB_ASYM, _lib.LAYOUT_FULL, _lib.ALGO_GATHER)
    assert np.array_equal(P.cpu().numpy(), expect(f, nd, _lib.PROX_OOB_ASYM, _lib.LAYOUT_FULL))


def test_long_rows_many_trees_match_bucket_form(gpu):
    """n = 40,000 and T = 600: the long-row gather with several tree groups
    per lane (cursors of 4 trees) against the independent global-atomic form,
    for ALL + OOB (fused pair) and the FULL OOB_ASYM layout."""
    _say("long rows n=40000 T=600")
    x, y, f = small_forest(_lib.DATA_COVTYPE_LIKE, 40000, 54, 7, 600, 9, mtry=7)
    df, keys, _ = run_keys(f, x)
    acc = Accumulator(f.n, f.ntree, f.total_nodes)
    a, o = acc.accumulate_pair(keys, _lib.LAYOUT_UPPER, _lib.ALGO_GATHER)
    ba = acc.accumulate(keys, _lib.PROX_ALL, _lib.LAYOUT_UPPER, _lib.ALGO_BUCKET)
    assert torch.equal(a, ba)
    del ba
    bo = acc.accumulate(keys, _lib.PROX_OOB_UPPER, _lib.LAYOUT_UPPER, _lib.ALGO_BUCKET)
    assert torch.equal(o, bo)
    # size-independent properties: C_ii = number of trees with a terminal leaf, row sums
    n = f.n
    diag_idx = torch.arange(n, device=a.device, dtype=torch.int64)
    diag_idx = diag_idx * n - diag_idx * (diag_idx - 1) // 2
    assert int(a[diag_idx].max()) <= f.ntree
    del a, o, bo
    P = acc.accumulate(keys, _lib.PROX_OOB_ASYM, _lib.LAYOUT_FULL, _lib.ALGO_GATHER)
    Q = acc.accumulate(keys, _lib.PROX_OOB_ASYM, _lib.LAYOUT_FULL, _lib.ALGO_BUCKET)
    assert torch.equal(P, Q)
    del P, Q
    P = acc.accumulate(keys, _lib.PROX_OOB_ROW, _lib.LAYOUT_FULL, _lib.ALGO_GATHER)
    Q = acc.accumulate(keys, _lib.PROX_OOB_ROW, _lib.LAYOUT_FULL, _lib.ALGO_BUCKET)
    assert torch.equal(P, Q)


@pytest.mark.parametrize("n", [65536, 70001])
def test_leaf_assign_both_tile_configurations(gpu, n):
    """The staged leaf kernel takes 8 x 4 trees per tile up to n = 65,536 and
    4 x 2 trees beyond: both against the oracle, keys and nodextr."""
    _say(f"leaf tiles n={n}")
    x, y, f = small_forest(_lib.DATA_COVTYPE_LIKE, n, 54, 7, 13, 21, mtry=7)
    _, keys, nodextr = run_keys(f, x)
    nd = O.leaf_assign(f, x)
    assert np.array_equal(nodextr.cpu().numpy(), nd)
    assert np.array_equal(keys.cpu().numpy(), keys_from_nodextr(f, nd))
