"""GPU parity of the case-wise (bootstrap-weighted) proximity against the
oracle's restatement of the reference's use_casewise branch (itself pinned
bit for bit to the reference's own cpu_proximity in test_oracle.py).  The GPU
sums the weights of different trees in no fixed order, so entries are
compared at 1e-12 relative, and the per-row sums against their closed form
(the number of trees in which the case is out of bag)."""
from __future__ import annotations

import numpy as np
import pytest
import torch

from oracle import oracle as O
from rfx_amd import _lib
from rfx_amd.forest import Forest, make_dataset
from rfx_amd.proximity import Accumulator, DeviceForest

pytestmark = pytest.mark.gpu

# (kind, n, p, classes, T, seed): one chunk, several 4096-row chunks, ragged
SHAPES = [(_lib.DATA_OVERLAP_MIXTURE, 300, 7, 3, 6, 9), (_lib.DATA_GAUSS_MIXTURE, 2500, 12, 4, 21, 3),
          (_lib.DATA_COVTYPE_LIKE, 9000, 54, 7, 8, 5)]


@pytest.mark.parametrize("shape", SHAPES, ids=lambda s: f"n{s[1]}_T{s[4]}")
def test_casewise_matches_oracle(gpu, shape):
    kind, n, p, ncl, T, seed = shape
    print(f"casewise n={n} T={T} seed={seed}", flush=True)
    x, y = make_dataset(kind, n, p, ncl, seed)
    f = Forest.grow(x, y, T, iseed=seed)
    df = DeviceForest(f, x)
    keys, _ = df.leaf_assign()
    nin = torch.from_numpy(np.ascontiguousarray(f.nin.reshape(T, n), np.int32)).to(df.device)
    acc = Accumulator(n, T, f.total_nodes)
    got = acc.casewise(keys, nin)
    torch.cuda.synchronize()
    want = O.prox_v1(f, O.leaf_assign(f, x), casewise=True)
    g = got.cpu().numpy()
    assert np.allclose(g, want, rtol=1e-12, atol=1e-14)
    assert np.array_equal(g == 0, want == 0)  # the same entries are touched
    rows = g.reshape(n, n, order="F").sum(axis=1)
    assert np.allclose(rows, np.sum(f.nin.reshape(T, n) == 0, axis=0), rtol=0, atol=1e-9)
    # accumulate adds onto what is there
    acc.casewise(keys, nin, out=got, accumulate=True)
    torch.cuda.synchronize()
    assert np.allclose(got.cpu().numpy(), 2 * want, rtol=1e-12, atol=1e-14)


def test_casewise_rejects_bad_nin(gpu):
    x, y = make_dataset(_lib.DATA_GAUSS_MIXTURE, 200, 5, 3, 1)
    f = Forest.grow(x, y, 3, iseed=1)
    df = DeviceForest(f, x)
    keys, _ = df.leaf_assign()
    acc = Accumulator(200, 3, f.total_nodes)
    with pytest.raises(ValueError):
        acc.casewise(keys, torch.zeros(10, dtype=torch.int32, device=df.device))


@pytest.mark.parametrize("n,T", [(25000, 3), (33000, 2)])
def test_casewise_row_chunks(gpu, n, T):
    # two row chunks per column (16-bit members at n = 25,000, 32-bit beyond
    # 32,768), case-wise and RF-GAP against the oracle
    print(f"casewise chunks n={n} T={T}", flush=True)
    x, y = make_dataset(_lib.DATA_COVTYPE_LIKE, n, 54, 7, 8)
    f = Forest.grow(x, y, T, iseed=8)
    df = DeviceForest(f, x)
    keys, nodextr = df.leaf_assign(want_nodextr=True)
    nin = torch.from_numpy(np.ascontiguousarray(f.nin.reshape(T, n), np.int32)).to(df.device)
    acc = Accumulator(n, T, f.total_nodes)
    nd = O.leaf_assign(f, x)
    got = acc.casewise(keys, nin).cpu().numpy()
    assert np.allclose(got, O.prox_v1(f, nd, casewise=True), rtol=1e-12, atol=1e-14)
    del got
    rg = acc.rfgap(nin, nodextr, df.node_off).cpu().numpy()
    assert np.allclose(rg, O.prox_rfgap(f.nin.reshape(T, n), nd), rtol=1e-12, atol=1e-14)
