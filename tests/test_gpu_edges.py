"""GPU edge sizes: one sample, a handful of samples, one tree, ragged sample
counts around the 64-lane step of the per-tree bucket build, against the
oracle for every count variant and algorithm and for the case-wise and
RF-GAP matrices."""
from __future__ import annotations

import numpy as np
import pytest
import torch

from oracle import oracle as O
from rfx_amd import _lib
from rfx_amd.forest import Forest, make_dataset
from rfx_amd.proximity import Accumulator, DeviceForest

pytestmark = pytest.mark.gpu

VARIANTS = [(_lib.PROX_ALL, _lib.LAYOUT_FULL), (_lib.PROX_OOB_ASYM, _lib.LAYOUT_FULL),
            (_lib.PROX_OOB_UPPER, _lib.LAYOUT_UPPER), (_lib.PROX_ALL, _lib.LAYOUT_UPPER)]
ALGOS = [_lib.ALGO_DENSE, _lib.ALGO_BUCKET, _lib.ALGO_GATHER]


def _check(f: Forest, x: np.ndarray):
    df = DeviceForest(f, x)
    keys, nodextr = df.leaf_assign(want_nodextr=True)
    nd = O.leaf_assign(f, x)
    assert np.array_equal(nodextr.cpu().numpy(), nd)
    acc = Accumulator(f.n, f.ntree, f.total_nodes)
    for v, lay in VARIANTS:
        want = O.counts(f, nd, v, lay)
        for algo in ALGOS:
            assert np.array_equal(acc.accumulate(keys, v, lay, algo).cpu().numpy(), want), (v, lay, algo)
    nin = torch.from_numpy(np.ascontiguousarray(f.nin.reshape(f.ntree, f.n), np.int32)).to(df.device)
    cw = acc.casewise(keys, nin).cpu().numpy()
    assert np.allclose(cw, O.prox_v1(f, nd, casewise=True), rtol=1e-12, atol=1e-14)
    rg = acc.rfgap(nin, nodextr, df.node_off).cpu().numpy()
    assert np.allclose(rg, O.prox_rfgap(f.nin.reshape(f.ntree, f.n), nd), rtol=1e-12, atol=1e-14)


@pytest.mark.parametrize("n,T", [(2, 1), (3, 2), (63, 1), (64, 3), (65, 9), (129, 2)])
def test_small_and_ragged_sizes(gpu, n, T):
    print(f"edge n={n} T={T}", flush=True)
    x, y = make_dataset(_lib.DATA_GAUSS_MIXTURE, n, 4, 2, 7)
    f = Forest.grow(x, y, T, iseed=7)
    _check(f, x)


def test_one_sample_one_leaf(gpu):
    # a single sample and single-leaf trees, in bag in one tree, out of bag in the other
    print("edge n=1", flush=True)
    f = Forest.from_arrays(1, 1, np.array([1, 1]), np.array([-1, -1], np.int32), np.zeros(2, np.int32),
                           np.zeros(2, np.float32), np.zeros(4, np.int32), np.array([[1], [0]], np.int32))
    _check(f, np.zeros((1, 1), np.float32))
