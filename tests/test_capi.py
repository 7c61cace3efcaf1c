"""CPU: the C ABI library loads, exports every declared entry point, and
rejects bad arguments with status codes (no GPU work is launched)."""
from __future__ import annotations

import ctypes as C
import re
from pathlib import Path

import numpy as np
import pytest

from rfx_amd import _lib

HEADER = Path(__file__).resolve().parents[1] / "include" / "rfx_prox.h"


def declared_symbols() -> set[str]:
    text = HEADER.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return set(re.findall(r"\b(rfx_[a-z0-9_]+)\s*\(", text))


def test_library_exports_every_declared_symbol():
    L = _lib.lib()
    syms = declared_symbols()
    assert len(syms) >= 15
    for s in sorted(syms):
        assert hasattr(L, s), f"{s} declared in rfx_prox.h but not exported"
    assert syms <= set(_lib.exported_symbols()), "binding table misses a declared symbol"


def test_invalid_arguments_return_status():
    L = _lib.lib()
    assert L.rfx_leaf_assign(None, None, 0, 1, None, None, 1, None, None, None, None) == -1
    # the queued form wants a workspace of at least 256 + 16 * 1024 bytes,
    # 16-byte aligned; it is sized by rfx_leaf_workspace_bytes
    d = 256
    assert L.rfx_leaf_assign_ws(None, d, 10, 2, d, d, 1, d, None, None, d, d, None, 0) == -4
    assert L.rfx_leaf_assign_ws(None, d, 10, 2, d, d, 1, d, None, None, d, d, d, 256 + 16 * 1023) == -4
    assert L.rfx_leaf_assign_ws(None, d, 10, 2, d, d, 1, d, None, None, d, d, d + 8, 1 << 20) == -1
    # the chain index builder: bad arguments, and no index for a one-node tree
    nb = C.c_size_t(0)
    off = np.array([0, 1], np.int64)
    leaf = np.array([0, 0xFFFFFFFF], np.uint32)
    assert L.rfx_chain_index(None, 1, leaf.ctypes.data, 32, None, C.byref(nb)) == -1
    assert L.rfx_chain_index(off.ctypes.data, 1, leaf.ctypes.data, 1, None, C.byref(nb)) == -1
    assert L.rfx_chain_index(off.ctypes.data, 1, leaf.ctypes.data, 32, None, C.byref(nb)) == 0 and nb.value == 0
    assert L.rfx_leaf_workspace_bytes(0, 5) == 0
    assert L.rfx_leaf_workspace_bytes(6, 2) == 256 + 16 * 1024
    assert L.rfx_leaf_workspace_bytes(100, 30) == 256 + 16 * 3000
    assert L.rfx_leaf_workspace_bytes(200000, 1000) == 256 + 16 * (200000 * 1000 // 16)
    assert L.rfx_leaf_workspace_bytes(4000000, 5000) == 256 + 16 * (1 << 26)
    assert L.rfx_prox_accumulate(None, None, 0, 1, 1, 0, 0, 0, 0, None, None, 0) == -1
    assert L.rfx_prox_accumulate(None, None, 10, 1, 1, 9, 0, 1, 0, None, None, 0) == -1
    assert L.rfx_finishprox(None, None, 0, None, 10, None, 0) == -1
    assert L.rfx_prox_rows(None, None, 10, 1, 0, 5, 3, None) == -1
    h = C.c_void_p()
    x = np.zeros((4, 2), np.float32)
    y = np.zeros(4, np.int32)
    assert L.rfx_forest_grow(x.ctypes.data, y.ctypes.data, 4, 2, 1, 1, 1, 1, 0, 0, 0, 0, 1, C.byref(h)) == -1


def test_weighted_proximity_and_outlier_arguments():
    # case-wise, RF-GAP, the RF-GAP MDS and the outlier scores reject bad
    # arguments and short workspaces before any GPU work; dummy non-null
    # pointers are never dereferenced on these paths
    L = _lib.lib()
    d = 256  # a non-null dummy address
    n, T, tn = 1000, 10, 500
    cw = L.rfx_prox_casewise_workspace_bytes(n, T, tn)
    rg = L.rfx_prox_rfgap_workspace_bytes(n, T, tn)
    assert cw >= L.rfx_prox_workspace_bytes(n, T, tn, _lib.ALGO_BUCKET) and rg > cw
    assert L.rfx_prox_casewise_workspace_bytes(0, T, tn) == 0 and L.rfx_prox_rfgap_workspace_bytes(n, 0, tn) == 0
    assert L.rfx_prox_casewise(None, None, d, n, T, tn, 0, d, d, cw) == -1
    assert L.rfx_prox_casewise(None, d, d, n, T, tn, 0, d, d, cw - 1) == -4
    assert L.rfx_prox_rfgap(None, d, None, d, n, T, tn, d, d, rg) == -1
    assert L.rfx_prox_rfgap(None, d, d, d, n, T, tn, d, d, rg - 1) == -4
    assert L.rfx_prox_rfgap(None, d, d, d, n, T, 0, d, d, rg) == -1
    assert L.rfx_outlier_scores(None, None, n, d, None) == -1
    assert L.rfx_outlier_scores(None, d, 0, d, None) == -1
    mb = L.rfx_mds_dense_workspace_bytes(n, _lib.PREC_FP64)
    assert L.rfx_mds_dense_rfgap(None, d, n, None, _lib.PREC_FP64, 10, 1e-8, d, mb, d, None, None) == -1
    assert L.rfx_mds_dense_rfgap(None, d, n, d, _lib.PREC_FP64, 10, 1e-8, d, mb - 1, d, None, None) == -4
    assert L.rfx_mds_dense_rfgap(None, d, 3, d, _lib.PREC_FP64, 10, 1e-8, d, mb, d, None, None) == -1


def test_pack_nodes_rejects_malformed_forest():
    L = _lib.lib()
    off = np.array([0, 3], np.int64)
    st = np.array([1, -1, -1], np.int32)
    bv = np.array([0, 0, 0], np.int32)
    xs = np.zeros(3, np.float32)
    tm = np.array([1, 2, 0, 0, 0, 0], np.int32)
    out = np.empty(6, np.uint32)
    a = lambda v: v.ctypes.data  # noqa: E731
    assert L.rfx_pack_nodes(a(off), 1, a(st), a(bv), a(xs), a(tm), 1, a(out)) == 0
    assert out[1] == (1 << 10) and out[3] == 0xFFFFFFFF
    tm_bad = np.array([1, 1, 0, 0, 0, 0], np.int32)  # right child must be left + 1
    assert L.rfx_pack_nodes(a(off), 1, a(st), a(bv), a(xs), a(tm_bad), 1, a(out)) == -1
    st_bad = np.array([2, -1, -1], np.int32)  # unfinished node
    assert L.rfx_pack_nodes(a(off), 1, a(st_bad), a(bv), a(xs), a(tm), 1, a(out)) == -1


def test_workspace_size_formula():
    L = _lib.lib()
    assert L.rfx_prox_workspace_bytes(1000, 10, 500, _lib.ALGO_DENSE) == 0
    b = L.rfx_prox_workspace_bytes(1000, 10, 500, _lib.ALGO_BUCKET)
    assert b >= 8 * 500 + 8 * 1000 * 10


def test_reference_signature_module_loads():
    import rfx_amd._rfxref as R

    for name in ("gpu_proximity", "upper_triangle_fp16", "finishprox", "compute_mds_3d", "LowRankProximityMatrix",
                 "proximity_rfgap", "outlier_scores"):
        assert hasattr(R, name)


def test_classifier_options_mirror_reference():
    import inspect

    from rfx_amd.classifier import QUANT_FORMATS, RandomForestClassifier

    params = list(inspect.signature(RandomForestClassifier.__init__).parameters)[1:]
    assert params[:12] == ["nsample", "ntree", "mdim", "nclass", "maxcat", "mtry", "maxnode", "minndsize", "ncsplit",
                           "ncmax", "iseed", "compute_proximity"]
    assert params[-6:] == ["rank", "n_threads_cpu", "use_rfgap", "importance_method", "clique_M", "use_casewise"]
    clf = RandomForestClassifier()
    assert clf.rank == 32 and clf.quant_mode == "nf4" and clf.iseed == 12345
    assert set(QUANT_FORMATS) >= {"nf4", "int8", "fp16", "fp32"}
    with pytest.raises(NotImplementedError):
        RandomForestClassifier(use_rfgap=True, use_qlora=True)
    assert RandomForestClassifier(use_rfgap=True).use_rfgap
    with pytest.raises(NotImplementedError):
        RandomForestClassifier(use_casewise=True, use_qlora=True)
    assert RandomForestClassifier(use_casewise=True).use_casewise
    with pytest.raises(RuntimeError):
        clf.get_proximity_matrix()
    with pytest.raises(TypeError):
        clf.fit(np.zeros((10, 2), np.float32), np.zeros(10, np.int64))


def test_prebuilt_on_a_never_built_workspace_is_refused():
    # the build records are host state: the refusal comes before any launch
    L = _lib.lib()
    d = 4096  # a non-null dummy address, never dereferenced on this path
    n, T, tn = 1000, 4, 4000
    wsb = L.rfx_prox_workspace_bytes(n, T, tn, _lib.ALGO_GATHER)
    rc = L.rfx_prox_accumulate(None, d, n, T, tn, _lib.PROX_ALL, _lib.LAYOUT_UPPER,
                               _lib.ALGO_GATHER | _lib.ALGO_PREBUILT, 0, d, d, wsb)
    assert rc == -1
