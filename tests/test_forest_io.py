"""CPU: the forest file (Forest.save / Forest.load) round-trips every array
bit for bit, refuses files of another format, and the sample-subset view
keeps each sample's bootstrap counts; a forest grown in chunks of trees is
the forest of one call."""
from __future__ import annotations

import numpy as np
import pytest

from rfx_amd import _lib
from rfx_amd.forest import Forest, make_dataset

from helpers import load_golden


def test_forest_file_round_trip(tmp_path):
    f = load_golden("covtype_cat_n240.npz")["forest"]
    path = tmp_path / "forest.npz"
    f.save(path)
    g = Forest.load(path)
    assert (g.n, g.p) == (f.n, f.p)
    for k in ("node_offset", "nodestatus", "bestvar", "treemap", "nodeclass", "nin"):
        assert np.array_equal(getattr(g, k), getattr(f, k)), k
    assert np.array_equal(g.xbestsplit.view(np.uint32), f.xbestsplit.view(np.uint32))
    assert np.array_equal(g.packed_nodes(), f.packed_nodes())


def test_forest_file_rejects_other_formats(tmp_path):
    path = tmp_path / "other.npz"
    np.savez(path, format=np.array("something else"), n=np.int64(1))
    with pytest.raises(ValueError):
        Forest.load(path)


def test_sample_subset_keeps_bootstrap_counts():
    f = load_golden("gauss_n200.npz")["forest"]
    idx = np.array([0, 7, 31, 199])
    s = f.samples(idx)
    assert s.n == 4 and s.total_nodes == f.total_nodes
    assert np.array_equal(s.nin, f.nin[:, idx])
    assert np.array_equal(s.nout(2), f.nout(2)[idx])


def test_chunked_growth_equals_one_call():
    x, y = make_dataset(_lib.DATA_SUBSPACE_MIXTURE, 3000, 12, 10, 2)
    a = Forest.grow(x, y, 11, mtry=3, minndsize=5, iseed=4)
    seen = []
    b = Forest.grow_chunked(x, y, 11, chunk=4, progress=lambda d, t: seen.append(d), mtry=3, minndsize=5, iseed=4)
    assert seen == [4, 8, 11]
    for k in ("node_offset", "nodestatus", "bestvar", "treemap", "nodeclass", "nin"):
        assert np.array_equal(getattr(a, k), getattr(b, k)), k
    assert np.array_equal(a.xbestsplit.view(np.uint32), b.xbestsplit.view(np.uint32))


def test_progress_counter_growth_equals_plain_call():
    x, y = make_dataset(_lib.DATA_SUBSPACE_MIXTURE, 3000, 12, 10, 3)
    a = Forest.grow(x, y, 9, mtry=3, minndsize=5, iseed=2)
    seen = []
    b = Forest.grow(x, y, 9, mtry=3, minndsize=5, iseed=2, progress=lambda d, t: seen.append((d, t)),
                    progress_every=0.01)
    assert seen and seen[-1] == (9, 9)
    for k in ("node_offset", "nodestatus", "bestvar", "treemap", "nodeclass", "nin"):
        assert np.array_equal(getattr(a, k), getattr(b, k)), k
