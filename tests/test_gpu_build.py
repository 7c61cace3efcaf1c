"""GPU leaf-bucket build: the one-workgroup-per-tree counting sort (n up to
32,767) checked through the counts of the BUCKET and GATHER kernels, which
read the buckets, against the oracle and against the DENSE kernel, which does
not.  Covers the case of trees with more than 65,535 nodes (trees grown on
more samples than are assigned), whose leaf ids need two counting passes,
and both sides of the size limit of that build."""
from __future__ import annotations

import numpy as np
import pytest
import torch

from oracle import oracle as O
from rfx_amd import _lib
from rfx_amd.forest import Forest, make_dataset
from rfx_amd.proximity import Accumulator, DeviceForest

pytestmark = pytest.mark.gpu


def test_wide_leaf_range_two_passes(gpu):
    # 2 trees grown on 200,000 noisy samples (> 65,535 nodes each), assigned
    # to the first 3,000 samples: each tree's reached leaf ids span more than
    # 65,536 values
    print("wide leaf range: grow 200k", flush=True)
    x, y = make_dataset(_lib.DATA_OVERLAP_MIXTURE, 200000, 8, 3, 11)
    g = Forest.grow(x, y, 2, mtry=3, iseed=11)
    n = 3000
    f = Forest(n, 8, g.node_offset, g.nodestatus, g.bestvar, g.xbestsplit, g.treemap, g.nodeclass,
               np.ascontiguousarray(g.nin.reshape(2, -1)[:, :n]))
    xs = np.ascontiguousarray(x[:n])
    nd = O.leaf_assign(f, xs)
    assert all(int(nd[t].max()) - int(nd[t].min()) > 65536 for t in range(2))
    df = DeviceForest(f, xs)
    keys, _ = df.leaf_assign()
    acc = Accumulator(n, 2, f.total_nodes)
    for variant in (_lib.PROX_ALL, _lib.PROX_OOB_ASYM):
        want = O.counts(f, nd, variant, _lib.LAYOUT_FULL)
        for algo in (_lib.ALGO_BUCKET, _lib.ALGO_GATHER):
            got = acc.accumulate(keys, variant, _lib.LAYOUT_FULL, algo)
            assert np.array_equal(got.cpu().numpy(), want), (variant, algo)
    # the UPPER whole-row path builds only pe and the 16-bit members; its
    # two-pass sort uses the leaf ends, still holding the calls above, as
    # scratch without a clear
    print("wide leaf range: UPPER rows", flush=True)
    a, o = acc.accumulate_pair(keys, _lib.LAYOUT_UPPER, _lib.ALGO_GATHER)
    assert np.array_equal(a.cpu().numpy(), O.counts(f, nd, _lib.PROX_ALL, _lib.LAYOUT_UPPER))
    assert np.array_equal(o.cpu().numpy(), O.counts(f, nd, _lib.PROX_OOB_UPPER, _lib.LAYOUT_UPPER))


@pytest.mark.parametrize("n", [32767, 32768])
def test_build_size_limit(gpu, n):
    # the last size of the per-tree counting sort and the first of the radix
    # sort: bucket-based counts equal the DENSE kernel's
    print(f"build size limit n={n}", flush=True)
    x, y = make_dataset(_lib.DATA_COVTYPE_LIKE, n, 54, 7, 2)
    f = Forest.grow(x, y, 3, mtry=7, iseed=2)
    df = DeviceForest(f, x)
    keys, _ = df.leaf_assign()
    acc = Accumulator(n, 3, f.total_nodes)
    a_d, o_d = acc.accumulate_pair(keys, _lib.LAYOUT_UPPER, _lib.ALGO_DENSE)
    a_b, o_b = acc.accumulate_pair(keys, _lib.LAYOUT_UPPER, _lib.ALGO_BUCKET)
    assert torch.equal(a_b, a_d) and torch.equal(o_b, o_d)
    del a_b, o_b
    a_g, o_g = acc.accumulate_pair(keys, _lib.LAYOUT_UPPER, _lib.ALGO_GATHER)
    assert torch.equal(a_g, a_d) and torch.equal(o_g, o_d)


@pytest.mark.parametrize("n", [3000, 34000])
def test_prebuilt_refuses_partial_or_foreign_builds(gpu, n):
    # RFX_ALGO_PREBUILT runs only on a workspace whose last build was a full
    # one of the same shape: after rfx_prox_build_buckets it counts; after a
    # count call that built only part of the buckets (n <= 32,767, UPPER
    # GATHER) or the case-wise call (no long-line window arrays) it refuses
    print(f"prebuilt records n={n}", flush=True)
    x, y = make_dataset(_lib.DATA_COVTYPE_LIKE, n, 54, 7, 4)
    f = Forest.grow(x, y, 6, mtry=7, iseed=4)
    df = DeviceForest(f, x)
    keys, _ = df.leaf_assign()
    acc = Accumulator(n, f.ntree, f.total_nodes)
    nd = O.leaf_assign(f, x)
    want = O.counts(f, nd, _lib.PROX_ALL, _lib.LAYOUT_UPPER)
    acc.build_buckets(keys)
    a, _ = acc.accumulate_pair(keys, _lib.LAYOUT_UPPER, _lib.ALGO_GATHER | _lib.ALGO_PREBUILT)
    assert np.array_equal(a.cpu().numpy(), want)
    # another shape on the same workspace is refused
    with pytest.raises(RuntimeError, match="invalid argument"):
        _lib.check(_lib.lib().rfx_prox_accumulate_pair(
            torch.cuda.current_stream().cuda_stream, keys.data_ptr(), n, f.ntree - 1, f.total_nodes,
            _lib.LAYOUT_UPPER, _lib.ALGO_GATHER | _lib.ALGO_PREBUILT, 0, a.data_ptr(), a.data_ptr(),
            acc.ws.data_ptr(), acc.ws.numel()), "prebuilt")
    if n <= 32767:
        acc.accumulate_pair(keys, _lib.LAYOUT_UPPER, _lib.ALGO_GATHER)  # partial build
        with pytest.raises(RuntimeError, match="invalid argument"):
            acc.accumulate_pair(keys, _lib.LAYOUT_UPPER, _lib.ALGO_BUCKET | _lib.ALGO_PREBUILT)
    else:
        nin = torch.from_numpy(np.ascontiguousarray(f.nin, np.int32)).to(gpu)
        acc.casewise(keys, nin)  # full buckets, no window arrays
        with pytest.raises(RuntimeError, match="invalid argument"):
            acc.accumulate_pair(keys, _lib.LAYOUT_UPPER, _lib.ALGO_GATHER | _lib.ALGO_PREBUILT)
        b = acc.accumulate(keys, _lib.PROX_ALL, _lib.LAYOUT_UPPER, _lib.ALGO_BUCKET | _lib.ALGO_PREBUILT)
        assert np.array_equal(b.cpu().numpy(), want)
    acc.build_buckets(keys)
    a, _ = acc.accumulate_pair(keys, _lib.LAYOUT_UPPER, _lib.ALGO_GATHER | _lib.ALGO_PREBUILT)
    assert np.array_equal(a.cpu().numpy(), want)
