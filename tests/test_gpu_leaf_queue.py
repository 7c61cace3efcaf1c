"""GPU leaf assignment with the walk queue and chain jumps
(rfx_leaf_assign_ws): trees with long chains of single-sample splits, the
shape trees grown to small nodes on large noisy data take, against the
oracle's walk and against the one-pass kernel.  The hand-built chain trees
reach depth 3,000, so in-bag walks end deep in the drain kernel and
out-of-bag walks hit the 999-descent cap there (src/rf_testreebag.cpp:51-122:
node 0, off-leaf key)."""
from __future__ import annotations

import numpy as np
import pytest
import torch

from oracle import oracle as O
from rfx_amd import _lib
from rfx_amd.forest import Forest, make_dataset
from rfx_amd.proximity import DeviceForest
from helpers import keys_from_nodextr, random_chain_forest

pytestmark = pytest.mark.gpu


def chain_tree(D: int, p: int, n: int, rng) -> Forest:
    """Node 2j (j < D) splits on feature j % p at j + 0.5: its left child
    2j + 1 is a leaf, its right child 2j + 2 goes on; node 2D is a leaf."""
    nn = 2 * D + 1
    st = np.full(nn, -1, np.int32)
    st[0:2 * D:2] = 1
    bv = np.zeros(nn, np.int32)
    bv[0:2 * D:2] = np.arange(D) % p
    xs = np.zeros(nn, np.float32)
    xs[0:2 * D:2] = np.arange(D) + 0.5
    tm = np.zeros(2 * nn, np.int32)
    k = np.arange(0, 2 * D, 2)
    tm[2 * k] = k + 1
    tm[2 * k + 1] = k + 2
    nin = rng.poisson(1.0, n).astype(np.int32)
    return Forest(n, p, np.array([0, nn], np.int64), st, bv, xs, tm, np.zeros(nn, np.int32), nin.reshape(1, n))


def chain_forest(n: int, p: int, seed: int, D: int = 3000):
    """Samples at depth z + 1 of the chains (z = floor(D u^4), u uniform: most
    shallow, some past the cap), 3 chain trees between 4 grown trees."""
    rng = np.random.default_rng(seed)
    z = np.floor((D + 40) * rng.random(n) ** 4)
    x = (z[:, None] + 0.25 * (np.arange(p)[None, :] % 2)).astype(np.float32)
    y = (z % 3).astype(np.int32)
    g = Forest.grow(x, y, 4, iseed=seed)
    parts = [g.subset(slice(0, 2)), chain_tree(D, p, n, rng), chain_tree(D // 2, p, n, rng), g.subset(slice(2, 4)),
             chain_tree(D, p, n, rng)]
    return x, Forest.concat(parts)


@pytest.mark.parametrize("n", [5000, 70000])
def test_chain_trees_match_oracle(gpu, n):
    # n = 5,000 and 70,000 take the two staged configurations (split at 65,536)
    print(f"leaf queue chain trees n={n}", flush=True)
    x, f = chain_forest(n, 6, 3)
    nd = O.leaf_assign(f, x)
    assert nd[2].max() > 2 * 999  # walks past the first-pass cap and the out-of-bag cap
    want = keys_from_nodextr(f, nd)
    df = DeviceForest(f, x)
    assert df.chains is not None
    keys, nodextr = df.leaf_assign(want_nodextr=True)
    assert np.array_equal(nodextr.cpu().numpy(), nd)
    assert np.array_equal(keys.cpu().numpy(), want)
    ws = df._leaf_ws
    offered = int(ws[:8].cpu().numpy().view(np.uint64)[0])
    assert offered > 0  # the drain kernel took part
    k1, n1 = df.leaf_assign(want_nodextr=True, queue=False)
    assert torch.equal(k1, keys) and torch.equal(n1, nodextr)


def test_full_queue_sends_walks_back(gpu):
    # the smallest workspace (1,024 records, one tree per chunk): most offers
    # are refused and those walks end in the first pass; a second call reuses
    # the workspace
    print("leaf queue overflow", flush=True)
    x, f = chain_forest(20000, 5, 4)
    want = keys_from_nodextr(f, O.leaf_assign(f, x))
    df = DeviceForest(f, x)
    small = 256 + 16 * 1024
    for _ in range(2):
        keys, _ = df.leaf_assign(queue_bytes=small)
        assert np.array_equal(keys.cpu().numpy(), want)
        offered = int(df._leaf_ws[:8].cpu().numpy().view(np.uint64)[0])
        assert offered > 1024


def test_queue_on_grown_forest(gpu):
    # an ordinary grown forest, both configurations, its heavy paths of 16 or
    # more nodes indexed: same keys as the one-pass kernel
    for n in (3000, 66000):
        print(f"leaf queue grown n={n}", flush=True)
        x, y = make_dataset(_lib.DATA_COVTYPE_LIKE, n, 54, 7, 5)
        f = Forest.grow(x, y, 12, iseed=5)
        df = DeviceForest(f, x, chain_min_len=16)
        assert df.chains is not None
        k0, _ = df.leaf_assign(queue=False)
        k1, _ = df.leaf_assign()
        assert torch.equal(k0, k1)
        assert np.array_equal(k1.cpu().numpy(), keys_from_nodextr(f, O.leaf_assign(f, x)))


@pytest.mark.parametrize("n,p", [(4000, 7), (70000, 7), (6000, 80)])
def test_jumps_on_drifting_chains(gpu, n, p):
    # chains with random features, both heavy sides, thresholds out of order,
    # nested chains and samples on thresholds (the host restatement of the
    # jumps is pinned against the oracle in test_chain_index.py); p = 80 takes
    # the drain kernel that reads features from global memory (p > 64)
    print(f"leaf queue drifting chains n={n} p={p}", flush=True)
    x, f = random_chain_forest(n, p, 9)
    nd = O.leaf_assign(f, x)
    df = DeviceForest(f, x, chain_min_len=32)
    assert df.chains is not None
    keys, nodextr = df.leaf_assign(want_nodextr=True)
    assert np.array_equal(nodextr.cpu().numpy(), nd)
    assert np.array_equal(keys.cpu().numpy(), keys_from_nodextr(f, nd))
    k1, _ = df.leaf_assign(queue=False)
    assert torch.equal(k1, keys)
