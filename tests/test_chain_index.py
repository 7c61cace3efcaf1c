"""CPU: the chain index (rfx_chain_index) of long heavy paths.  A host
restatement of the drain kernel's jumps, read from the index bytes, must
reach the node that the oracle's plain walk reaches
(src/rf_testreebag.cpp:51-122 and src/rf_random_forest.cpp:2194-2244) for
every sample and tree: chains with random features, both heavy sides,
thresholds out of order, nested chains, samples on thresholds, and walks past
the out-of-bag 999-descent cap."""
from __future__ import annotations

import numpy as np

from helpers import chain_index_sections, chain_jump_walk, random_chain_forest
from oracle import oracle as O


def test_index_layout_and_jumps_match_walk():
    x, f = random_chain_forest(600, 7, 21)
    ci = f.chain_index(min_len=32)
    assert ci is not None and ci.dtype == np.uint8 and len(ci) % 16 == 0
    nd = O.leaf_assign(f, x)
    # some out-of-bag walks pass 999 levels and end at node 0
    deep_oob = (nd[1:4] == 0) & (f.nin[1:4] == 0)
    assert deep_oob.sum() > 0
    got = chain_jump_walk(f, x, ci)
    assert np.array_equal(got, nd)


def test_planted_index_mistakes_are_caught():
    # plausible builder mistakes, planted in copies of a correct index: each
    # must change some sample's node
    x, f = random_chain_forest(1500, 6, 5)
    ci = f.chain_index(min_len=32)
    nd = O.leaf_assign(f, x)
    assert np.array_equal(chain_jump_walk(f, x, ci), nd)

    def planted(edit):
        bad = ci.copy()
        edit(chain_index_sections(bad))
        return not chain_jump_walk(f, x, bad, expect=nd)

    def flip_side(s):  # heavy side of every list inverted
        s["lists"][:, 3] ^= 1 << 16

    def heavy_exit(s):  # the heavy child recorded as the exit
        s["exits"][:] ^= 1

    def late_pos(s):  # record positions one level late
        s["recs"][:, 1] += 1

    def drop_last(s):  # the last record of every list lost
        s["lists"][:, 2] = np.maximum(s["lists"][:, 2], 1) - 1

    for edit in (flip_side, heavy_exit, late_pos, drop_last):
        assert planted(edit), edit.__name__


def test_short_paths_give_no_index():
    # the two grown trees (300 samples) have no heavy path of 400 nodes
    x, f = random_chain_forest(300, 5, 3)
    g = f.subset(slice(0, 1))
    assert g.chain_index(min_len=400) is None
    ci = f.chain_index(min_len=400)
    assert ci is not None
    assert np.array_equal(chain_jump_walk(f, x, ci), O.leaf_assign(f, x))


def test_path_cap_per_tree():
    # min_len 2 on a tree grown on 60,000 samples: more heavy paths than the
    # 1,023 a tree's labels can name; the index keeps the first 1,023 and the
    # rest are walked
    from rfx_amd import _lib
    from rfx_amd.forest import Forest, make_dataset

    x, y = make_dataset(_lib.DATA_COVTYPE_LIKE, 60000, 54, 7, 4)
    f = Forest.grow(x, y, 1, iseed=4)
    idx = np.arange(0, 60000, 150)
    g = f.samples(idx)
    xs = np.ascontiguousarray(x[idx])
    ci = g.chain_index(min_len=2)
    assert int(ci[8:12].view(np.int32)[0]) == 1023
    assert chain_jump_walk(g, xs, ci, expect=O.leaf_assign(g, xs))
