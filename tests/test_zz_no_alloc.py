"""The no-allocation entry point rfx_prox_accumulate_pair_noalloc (leaf
buckets in the library's static device array) against the workspace entry
point rfx_prox_accumulate_pair, through the C ABI: bit-identical counts on
two calls in a row, the second on a smaller problem, so the second call runs
on whatever the first left in the shared array.  Also its refusals: a problem
whose buckets exceed the array, and RFX_ALGO_PREBUILT."""
from __future__ import annotations

import numpy as np
import pytest
import torch

from rfx_amd import _lib
from rfx_amd._lib import check, lib
from rfx_amd.forest import Forest, make_dataset
from rfx_amd.proximity import Accumulator, DeviceForest, upper_size

pytestmark = pytest.mark.gpu


def _stream() -> int:
    return torch.cuda.current_stream().cuda_stream


def _problem(n: int, p: int, T: int, seed: int):
    x, y = make_dataset(_lib.DATA_COVTYPE_LIKE, n, p, 7, seed)
    f = Forest.grow(x, y, T, mtry=7, iseed=seed)
    keys, _ = DeviceForest(f, x).leaf_assign()
    return f, keys


def _noalloc(f: Forest, keys: torch.Tensor, layout: int, algo: int) -> tuple[np.ndarray, np.ndarray]:
    elems = upper_size(f.n) if layout == _lib.LAYOUT_UPPER else f.n * f.n
    a = torch.full((elems,), -7, dtype=torch.int32, device=keys.device)
    o = torch.full((elems,), -7, dtype=torch.int32, device=keys.device)
    check(lib().rfx_prox_accumulate_pair_noalloc(_stream(), keys.data_ptr(), f.n, f.ntree, f.total_nodes, layout,
                                                 algo, 0, a.data_ptr(), o.data_ptr()),
          "rfx_prox_accumulate_pair_noalloc")
    return a.cpu().numpy(), o.cpu().numpy()


def _with_workspace(f: Forest, keys: torch.Tensor, layout: int, algo: int) -> tuple[np.ndarray, np.ndarray]:
    a, o = Accumulator(f.n, f.ntree, f.total_nodes).accumulate_pair(keys, layout, algo)
    return a.cpu().numpy(), o.cpu().numpy()


@pytest.mark.parametrize("layout", [_lib.LAYOUT_UPPER, _lib.LAYOUT_FULL])
@pytest.mark.parametrize("algo", [_lib.ALGO_GATHER, _lib.ALGO_BUCKET])
def test_noalloc_matches_workspace_entry(gpu, layout, algo):
    big = _problem(1800, 54, 60, 3)
    small = _problem(700, 54, 24, 4)
    for f, keys in (big, small):  # the second call is smaller and reuses the array
        print(f"noalloc n={f.n} T={f.ntree} layout={layout} algo={algo}", flush=True)
        assert lib().rfx_prox_workspace_bytes(f.n, f.ntree, f.total_nodes, algo) <= _lib.NOALLOC_SCRATCH_BYTES
        got = _noalloc(f, keys, layout, algo)
        want = _with_workspace(f, keys, layout, algo)
        assert np.array_equal(got[0], want[0]) and np.array_equal(got[1], want[1])


def test_noalloc_long_rows(gpu):
    # n > 32,768: radix-sorted buckets and the gallop kernel inside the array
    f, keys = _problem(33000, 54, 12, 5)
    print(f"noalloc long rows n={f.n} T={f.ntree}", flush=True)
    got = _noalloc(f, keys, _lib.LAYOUT_UPPER, _lib.ALGO_GATHER)
    want = _with_workspace(f, keys, _lib.LAYOUT_UPPER, _lib.ALGO_GATHER)
    assert np.array_equal(got[0], want[0]) and np.array_equal(got[1], want[1])


def test_noalloc_refusals(gpu):
    f, keys = _problem(600, 54, 8, 6)
    a = torch.empty(upper_size(f.n), dtype=torch.int32, device=keys.device)
    o = torch.empty_like(a)
    rc = lib().rfx_prox_accumulate_pair_noalloc(_stream(), keys.data_ptr(), f.n, f.ntree, f.total_nodes,
                                                _lib.LAYOUT_UPPER, _lib.ALGO_GATHER | _lib.ALGO_PREBUILT, 0,
                                                a.data_ptr(), o.data_ptr())
    assert rc == -1  # RFX_ERR_INVALID
    # a problem whose bucket workspace exceeds the array is refused on the
    # host before any launch (the keys pointer is never read)
    n_big, T_big, nodes_big = 100000, 400, 2 * 100000 * 400
    assert lib().rfx_prox_workspace_bytes(n_big, T_big, nodes_big, _lib.ALGO_GATHER) > _lib.NOALLOC_SCRATCH_BYTES
    rc = lib().rfx_prox_accumulate_pair_noalloc(_stream(), keys.data_ptr(), n_big, T_big, nodes_big,
                                                _lib.LAYOUT_UPPER, _lib.ALGO_GATHER, 0, a.data_ptr(), o.data_ptr())
    assert rc == -4  # RFX_ERR_WORKSPACE
