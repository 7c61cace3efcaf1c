"""GPU parity of the batched path (many forests in one leaf launch, one
segmented leaf-bucket sort and one persistent gather launch): every
problem's UPPER counts bit-exact against the oracle, eager and as replays of
the captured hipGraph, and after autotuning."""
from __future__ import annotations

import numpy as np
import pytest
import torch

from oracle import oracle as O
from rfx_amd import _lib
from rfx_amd.batch import ForestBatch
from rfx_amd.forest import Forest, make_dataset
from rfx_amd.proximity import Accumulator, DeviceForest

pytestmark = pytest.mark.gpu

# (n, p, classes, T, seed): ragged sizes, n not a multiple of 256, T not a
# multiple of the 4-tree leaf groups, a one-tree forest and a tiny problem
SHAPES = [(1000, 8, 3, 50, 1), (37, 5, 2, 3, 2), (2500, 12, 4, 21, 3), (300, 20, 6, 1, 4), (4097, 10, 5, 13, 5)]


def _say(msg: str) -> None:
    print(msg, flush=True)


def _problems(shapes):
    out = []
    for n, p, ncl, T, seed in shapes:
        x, y = make_dataset(_lib.DATA_GAUSS_MIXTURE, n, p, ncl, seed)
        f = Forest.grow(x, y, T, iseed=seed)
        out.append((f, x))
    return out


def _expected(problems, variant):
    exp = []
    for f, x in problems:
        nodextr = O.leaf_assign(f, x)
        exp.append(O.counts(f, nodextr, variant, _lib.LAYOUT_UPPER))
    return exp


@pytest.fixture(scope="module")
def problems(gpu):
    probs = _problems(SHAPES)
    dfs = [DeviceForest(f, x) for f, x in probs]
    return probs, dfs


@pytest.mark.parametrize("variant", ["oob_upper", "all", "pair"])
def test_batch_counts_match_oracle(gpu, problems, variant):
    probs, dfs = problems
    _say(f"batch eager variant {variant} shapes {SHAPES}")
    b = ForestBatch(dfs, variant)
    b.run(use_graph=False)
    torch.cuda.synchronize()
    if variant == "pair":
        ea, eo = _expected(probs, _lib.PROX_ALL), _expected(probs, _lib.PROX_OOB_UPPER)
        for k in range(len(dfs)):
            assert np.array_equal(b.out[k].cpu().numpy(), ea[k]), f"problem {k} ALL"
            assert np.array_equal(b.out2[k].cpu().numpy(), eo[k]), f"problem {k} OOB"
    else:
        v = _lib.PROX_OOB_UPPER if variant == "oob_upper" else _lib.PROX_ALL
        for k, e in enumerate(_expected(probs, v)):
            assert np.array_equal(b.out[k].cpu().numpy(), e), f"problem {k}"
    b.close()


def test_batch_graph_replays_and_tuning(gpu, problems):
    probs, dfs = problems
    exp = _expected(probs, _lib.PROX_OOB_UPPER)
    b = ForestBatch(dfs, "oob_upper")
    for tuning in [(1, 1), (16, 0)]:
        b.set_tuning(*tuning)
        for rep in range(3):
            _say(f"batch graph replay {rep} tuning {tuning}")
            for o in b.out:
                o.fill_(-7)  # eager work between replays; the replay must overwrite all of it
            torch.cuda.synchronize()
            b.run(use_graph=True)
            torch.cuda.synchronize()
            for k, e in enumerate(exp):
                assert np.array_equal(b.out[k].cpu().numpy(), e), f"replay {rep} problem {k}"
    _say("batch autotune")
    ms = b.autotune(reps=1)
    assert ms > 0
    lpi, bpc = b.tuning()
    assert lpi in (1, 4, 16) and bpc >= 1
    for o in b.out:
        o.fill_(-7)
    b.run(use_graph=True)
    torch.cuda.synchronize()
    for k, e in enumerate(exp):
        assert np.array_equal(b.out[k].cpu().numpy(), e), f"after autotune problem {k}"
    b.close()


@pytest.mark.parametrize("variant", ["oob_upper", "pair"])
def test_batch_dense_and_gather_items_mixed(gpu, problems, variant):
    # DENSE tile items and GATHER row items of different forests in one queue
    probs, dfs = problems
    _say(f"batch mixed dense/gather items variant {variant}")
    b = ForestBatch(dfs, variant)
    b.set_algos([_lib.ALGO_DENSE if k % 2 == 0 else _lib.ALGO_GATHER for k in range(len(dfs))])
    for use_graph in (False, True, True):
        for o in b.out:
            o.fill_(-7)
        b.run(use_graph=use_graph)
        torch.cuda.synchronize()
        if variant == "pair":
            ea, eo = _expected(probs, _lib.PROX_ALL), _expected(probs, _lib.PROX_OOB_UPPER)
            for k in range(len(dfs)):
                assert np.array_equal(b.out[k].cpu().numpy(), ea[k]), f"problem {k} ALL"
                assert np.array_equal(b.out2[k].cpu().numpy(), eo[k]), f"problem {k} OOB"
        else:
            for k, e in enumerate(_expected(probs, _lib.PROX_OOB_UPPER)):
                assert np.array_equal(b.out[k].cpu().numpy(), e), f"problem {k}"
    algos, sq = b.stats()
    assert list(algos) == [_lib.ALGO_DENSE if k % 2 == 0 else _lib.ALGO_GATHER for k in range(len(dfs))]
    b.autotune(reps=1)
    algos, sq = b.stats()
    # measured leaf statistics: sum over trees and leaves of |leaf|^2 from the oracle's leaves
    for k, (f, x) in enumerate(probs):
        nd = O.leaf_assign(f, x)
        want = 0
        for t in range(f.ntree):
            b0 = f.node_offset[t]
            st = f.nodestatus[b0:f.node_offset[t + 1]]
            term = nd[t][st[nd[t]] == -1]
            want += int(np.sum(np.bincount(term).astype(np.int64) ** 2))
        assert sq[k] == want, f"problem {k}"
    b.close()


def test_batch_long_rows_match_single_problem_path(gpu):
    # n > the 8192-column LDS chunk: rows span several chunks; checked against
    # the one-problem GATHER path (itself pinned to the oracle)
    shapes = [(20000, 6, 3, 4, 11), (9000, 6, 3, 6, 12)]
    probs = _problems(shapes)
    dfs = [DeviceForest(f, x) for f, x in probs]
    _say(f"batch long rows {shapes}")
    b = ForestBatch(dfs, "pair")
    b.run(use_graph=True)
    torch.cuda.synchronize()
    for k, (df, (f, x)) in enumerate(zip(dfs, probs)):
        keys, _ = df.leaf_assign()
        acc = Accumulator(df.n, df.ntree, df.total_nodes)
        ra, ro = acc.accumulate_pair(keys, _lib.LAYOUT_UPPER, _lib.ALGO_GATHER)
        torch.cuda.synchronize()
        assert torch.equal(b.out[k], ra), f"problem {k} ALL"
        assert torch.equal(b.out2[k], ro), f"problem {k} OOB"
    # and the oracle on the largest problem
    f, x = probs[0]
    exp = O.counts(f, O.leaf_assign(f, x), _lib.PROX_OOB_UPPER, _lib.LAYOUT_UPPER)
    assert np.array_equal(b.out2[0].cpu().numpy(), exp)
    b.close()


@pytest.mark.parametrize("variant", ["oob_upper", "pair"])
def test_batch_own_launches_and_queue_mixed(gpu, variant):
    # after autotune the large forests get their own kernel launches (the
    # DENSE ones on the second stream), the tiny ones stay in the queue;
    # eager runs and graph replays match the one-problem path, which the
    # other tests pin to the oracle, and the oracle itself on one problem
    shapes = [(20000, 6, 3, 20, 21), (6000, 10, 4, 60, 22), (300, 20, 6, 1, 23), (37, 5, 2, 3, 24),
              (1000, 8, 3, 50, 25)]
    probs = _problems(shapes)
    dfs = [DeviceForest(f, x) for f, x in probs]
    _say(f"batch own launches + queue, variant {variant}, shapes {shapes}")
    b = ForestBatch(dfs, variant)
    b.autotune(reps=1)
    own = b.own_launch()
    algos, _ = b.stats()
    _say(f"own launch {list(own)} algos {list(algos)}")
    assert own[0] and not own[3], "the largest forest launches on its own, the 37-sample one is queued"
    ref = []
    for df in dfs:
        keys, _ = df.leaf_assign()
        acc = Accumulator(df.n, df.ntree, df.total_nodes)
        if variant == "pair":
            ref.append(acc.accumulate_pair(keys, _lib.LAYOUT_UPPER, _lib.ALGO_GATHER))
        else:
            ref.append((acc.accumulate(keys, _lib.PROX_OOB_UPPER, _lib.LAYOUT_UPPER, _lib.ALGO_GATHER), None))
    torch.cuda.synchronize()
    for use_graph in (False, True, True):
        for o in b.out:
            o.fill_(-7)
        if variant == "pair":
            for o in b.out2:
                o.fill_(-7)
        torch.cuda.synchronize()
        b.run(use_graph=use_graph)
        torch.cuda.synchronize()
        for k, (ra, ro) in enumerate(ref):
            assert torch.equal(b.out[k], ra), f"graph {use_graph} problem {k}"
            if variant == "pair":
                assert torch.equal(b.out2[k], ro), f"graph {use_graph} problem {k} OOB"
    f, x = probs[1]
    exp = O.counts(f, O.leaf_assign(f, x), _lib.PROX_OOB_UPPER, _lib.LAYOUT_UPPER)
    got = b.out2[1] if variant == "pair" else b.out[1]
    assert np.array_equal(got.cpu().numpy(), exp)
    b.close()


def test_batch_rejects_bad_input(gpu, problems):
    _, dfs = problems
    with pytest.raises(ValueError):
        ForestBatch(dfs, "oob_asym")
    with pytest.raises(RuntimeError):
        ForestBatch([], "all")


@pytest.mark.parametrize("variant", ["oob_upper", "all", "pair"])
def test_batch_own_launch_dense_packed_form(gpu, variant):
    # own-launched DENSE forests whose trees have at most 16,384 nodes run the
    # packed 16-bit compare (two trees per word, odd tree counts padded);
    # against the oracle for ALL and OOB_UPPER, PAIR keeps the 32-bit form
    shapes = [(6000, 10, 4, 61, 31), (4000, 8, 3, 37, 32), (9000, 12, 5, 20, 33)]
    probs = _problems(shapes)
    assert all(max(np.diff(f.node_offset)) <= 16384 for f, _ in probs)
    dfs = [DeviceForest(f, x) for f, x in probs]
    _say(f"batch packed DENSE, variant {variant}")
    b = ForestBatch(dfs, variant)
    b.autotune(reps=1)
    b.set_algos([_lib.ALGO_DENSE] * len(dfs))
    own = b.own_launch()
    assert own[0] and own[2], "the large forests launch on their own"
    b.run(use_graph=False)
    torch.cuda.synchronize()
    v = {"oob_upper": _lib.PROX_OOB_UPPER, "all": _lib.PROX_ALL, "pair": _lib.PROX_ALL}[variant]
    for k, (f, x) in enumerate(probs):
        nd = O.leaf_assign(f, x)
        assert np.array_equal(b.out[k].cpu().numpy(), O.counts(f, nd, v, _lib.LAYOUT_UPPER)), k
        if variant == "pair":
            assert np.array_equal(b.out2[k].cpu().numpy(), O.counts(f, nd, _lib.PROX_OOB_UPPER, _lib.LAYOUT_UPPER)), k
    b.close()


def test_batch_mds_concurrent_matches_sequential(gpu):
    """The per-instance MDS of a batch run concurrently on several streams
    (mds_dense_many) gives each instance the coordinates of its own
    sequential mds_dense call."""
    from rfx_amd.proximity import mds_dense, mds_dense_many

    shapes = [(1500, 8, 3, 40, 11), (900, 12, 4, 30, 12), (2100, 10, 5, 25, 13), (1200, 6, 3, 35, 14),
              (700, 9, 2, 20, 15)]
    probs = _problems(shapes)
    jobs = []
    for f, x in probs:
        keys, _ = DeviceForest(f, x).leaf_assign()
        U = Accumulator(f.n, f.ntree, f.total_nodes).accumulate(keys, _lib.PROX_OOB_UPPER, _lib.LAYOUT_UPPER,
                                                               _lib.ALGO_GATHER)
        jobs.append(dict(counts=U, layout=_lib.LAYOUT_UPPER, n=f.n, mode=_lib.MDS_REFERENCE, scale=1.0,
                         nout=torch.from_numpy(f.nout(2)).to(U.device), precision=_lib.PREC_FP32))
    seq = [mds_dense(**j) for j in jobs]
    _say("sequential MDS done; concurrent on 3 streams")
    conc = mds_dense_many(jobs, streams=3)
    for (Xs, es, _), (Xc, ec, _) in zip(seq, conc):
        assert np.allclose(ec[:3], es[:3], rtol=1e-9, atol=0)
        Xs, Xc = Xs.cpu().numpy(), Xc.cpu().numpy()
        assert np.linalg.norm(Xc - Xs) <= 1e-9 * np.linalg.norm(Xs)


def test_batch_own_launch_override(gpu, problems):
    """Every forest through the persistent queue, or every forest on its own
    launch, gives the autotuned batch's counts (graph replays after each
    change re-capture the sequence)."""
    probs, dfs = problems
    exp = _expected(probs, _lib.PROX_OOB_UPPER)
    b = ForestBatch(dfs, "oob_upper")
    b.autotune(reps=1)
    for flags in (np.zeros(len(dfs), np.int32), np.ones(len(dfs), np.int32), b.own_launch().astype(np.int32)):
        _say(f"batch own-launch flags {flags.tolist()}")
        b.set_own_launch(flags)
        assert np.array_equal(b.own_launch(), flags.astype(bool))
        for _ in range(2):
            b.run(use_graph=True)
            torch.cuda.synchronize()
            for k, e in enumerate(exp):
                assert np.array_equal(b.out[k].cpu().numpy(), e), f"problem {k}"
    b.close()
