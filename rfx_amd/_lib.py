"""ctypes binding of librfx.so (the C ABI declared in include/rfx_prox.h).

The library is loaded from the package directory; there is no fallback: a
missing or stale build raises immediately so no silent CPU path can stand in
for the HIP kernels.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

_LIB_PATH = Path(__file__).resolve().parent / "librfx.so"

RFX_OK = 0
STATUS = {
    -1: "invalid argument",
    -2: "allocation failed",
    -3: "HIP error",
    -4: "workspace too small",
    -5: "unsupported combination",
}

DATA_GAUSS_MIXTURE = 0
DATA_COVTYPE_LIKE = 1
DATA_SUBSPACE_MIXTURE = 2
DATA_OVERLAP_MIXTURE = 3

GROW_NO_FMA = 1
GROW_PLAIN = 2

PROX_ALL = 0
PROX_OOB_ASYM = 1
PROX_OOB_UPPER = 2
PROX_PAIR = 3  # batch: ALL into out, OOB_UPPER into out2
PROX_OOB_ROW = 4

LAYOUT_FULL = 0
LAYOUT_UPPER = 1

ALGO_AUTO = 0
ALGO_DENSE = 1
ALGO_BUCKET = 2
ALGO_GATHER = 3
ALGO_PREBUILT = 0x100  # or-ed into BUCKET/GATHER: buckets already in the workspace

MAX_TOTAL_NODES = 1073741823
NOALLOC_SCRATCH_BYTES = 64 << 20  # RFX_NOALLOC_SCRATCH_BYTES

MDS_COUNTS = 0
MDS_REFERENCE = 1
MDS_MATRIX = 2
MDS_RFGAP = 3

PREC_FP32 = 0
PREC_FP64 = 1
PREC_BF16 = 2
PREC_INT8 = 3
PREC_FP8 = 4
PREC_FP16 = 5

_p = C.c_void_p
_i32 = C.c_int32
_i64 = C.c_int64
_u32 = C.c_uint32
_u64 = C.c_uint64
_sz = C.c_size_t

_SIGS = {
    # host side
    "rfx_mt19937_u32": (_u32, [_u32, _i64]),
    "rfx_bootstrap": (_i32, [_i64, _u32, _p]),
    "rfx_make_dataset": (_i32, [_i32, _i64, _i32, _i32, _u64, _p, _p]),
    "rfx_forest_grow": (_i32, [_p, _p, _i64, _i32, _i32, _i32, _i32, _i32, _i32, _i32, _i32, _i32, _i32,
                               C.POINTER(_p)]),
    "rfx_forest_grow_progress": (_i32, [_p, _p, _i64, _i32, _i32, _i32, _i32, _i32, _i32, _i32, _i32, _i32, _i32,
                                        C.POINTER(C.c_int32), C.POINTER(_p)]),
    "rfx_forest_free": (None, [_p]),
    "rfx_forest_ntree": (_i32, [_p]),
    "rfx_forest_nnode": (_i32, [_p, _i32]),
    "rfx_forest_export": (_i32, [_p, _p, _p, _p, _p, _p, _p, _p]),
    "rfx_pack_nodes": (_i32, [_p, _i32, _p, _p, _p, _p, _i32, _p]),
    "rfx_oob_counts": (_i32, [_p, _i64, _i32, _i32, _p]),
    "rfx_pack_inbag_bits": (_i32, [_p, _i64, _i32, _p]),
    # GPU side
    "rfx_leaf_assign": (_i32, [_p, _p, _i64, _i32, _p, _p, _i32, _p, _p, _p, _p]),
    "rfx_leaf_workspace_bytes": (_sz, [_i64, _i32]),
    "rfx_leaf_assign_ws": (_i32, [_p, _p, _i64, _i32, _p, _p, _i32, _p, _p, _p, _p, _p, _p, _sz]),
    "rfx_chain_index": (_i32, [_p, _i32, _p, _i32, _p, _p]),
    "rfx_prox_workspace_bytes": (_sz, [_i64, _i32, _i64, _i32]),
    "rfx_prox_build_buckets": (_i32, [_p, _p, _i64, _i32, _i64, _p, _sz]),
    "rfx_prox_casewise_workspace_bytes": (_sz, [_i64, _i32, _i64]),
    "rfx_prox_casewise": (_i32, [_p, _p, _p, _i64, _i32, _i64, _i32, _p, _p, _sz]),
    "rfx_prox_rfgap_workspace_bytes": (_sz, [_i64, _i32, _i64]),
    "rfx_prox_rfgap": (_i32, [_p, _p, _p, _p, _i64, _i32, _i64, _p, _p, _sz]),
    "rfx_prox_accumulate": (_i32, [_p, _p, _i64, _i32, _i64, _i32, _i32, _i32, _i32, _p, _p, _sz]),
    "rfx_prox_accumulate_pair": (_i32, [_p, _p, _i64, _i32, _i64, _i32, _i32, _i32, _p, _p, _p, _sz]),
    "rfx_prox_accumulate_pair_noalloc": (_i32, [_p, _p, _i64, _i32, _i64, _i32, _i32, _i32, _p, _p]),
    "rfx_prox_rows": (_i32, [_p, _p, _i64, _i32, _i32, _i64, _i64, _p]),
    "rfx_prox_pair_stats": (_i32, [_p, _p, _i64, _i32, _i64, _p, _sz, C.POINTER(C.c_double)]),
    "rfx_finishprox": (_i32, [_p, _p, _i32, _p, _i64, _p, _i32]),
    "rfx_finishprox_f64": (_i32, [_p, _p, _p, _i64, _p, _i32]),
    "rfx_mds_dense_workspace_bytes": (_sz, [_i64, _i32]),
    "rfx_lowrank_workspace_bytes": (_sz, [_i64, _i32, _i64, _i32, _i32]),
    "rfx_lowrank_factors": (_i32, [_p, _p, _i64, _i32, _i64, _p, _i32, _i32, _i32, _u64, _p, _sz, _p, _p]),
    "rfx_quantize_factors": (_i32, [_p, _p, _i64, _i32, _i32, _p, _p]),
    "rfx_reconstruct_factors": (_i32, [_p, _p, _p, _i32, _i64, _i32, _p]),
    "rfx_mds_factors_workspace_bytes": (_sz, [_i64, _i32, _i32]),
    "rfx_mds_factors": (_i32, [_p, _p, _p, _i32, _i64, _i32, _i32, _i32, C.c_double, _p, _sz, _p, _p, _p]),
    "rfx_mds_factors_mixed_workspace_bytes": (_sz, [_i64, _i32]),
    "rfx_mds_factors_mixed": (_i32, [_p, _p, _p, _i32, _p, _i64, _i32, _i32, C.c_double, _i32, C.c_double, _p, _sz,
                                     _p, _p, _p]),
    "rfx_mds_dense_matrix": (_i32, [_p, _p, _i64, _p, _i32, _i32, C.c_double, _p, _sz, _p, _p, _p]),
    "rfx_mds_dense_rfgap": (_i32, [_p, _p, _i64, _p, _i32, _i32, C.c_double, _p, _sz, _p, _p, _p]),
    "rfx_outlier_scores": (_i32, [_p, _p, _i64, _p, _p]),
    "rfx_counts_convert": (_i32, [_p, _p, _i64, _i32, _p]),
    "rfx_batch_create": (_i32, [_p, _i32, _i32, C.POINTER(_p)]),
    "rfx_batch_workspace_bytes": (_sz, [_p]),
    "rfx_batch_run": (_i32, [_p, _p, _p, _sz, _i32]),
    "rfx_batch_autotune": (_i32, [_p, _p, _p, _sz, _i32, C.POINTER(C.c_double)]),
    "rfx_batch_set_tuning": (_i32, [_p, _i32, _i32]),
    "rfx_batch_get_tuning": (_i32, [_p, C.POINTER(_i32), C.POINTER(_i32)]),
    "rfx_batch_set_algos": (_i32, [_p, _p]),
    "rfx_batch_get_stats": (_i32, [_p, _p, _p]),
    "rfx_batch_get_own_launch": (_i32, [_p, _p]),
    "rfx_batch_set_own_launch": (_i32, [_p, _p]),
    "rfx_batch_free": (None, [_p]),
    "rfx_mds_dense": (_i32, [_p, _p, _i32, _i64, _i32, C.c_double, _p, _i32, _i32, C.c_double, _p, _sz, _p, _p,
                             _p]),
}


class BatchProblem(C.Structure):
    """rfx_batch_problem (include/rfx_prox.h)."""
    _fields_ = [("x_col", _p), ("n", _i64), ("p", _i32), ("ntree", _i32), ("nodes", _p), ("node_off", _p),
                ("total_nodes", _i64), ("inbag", _p), ("out", _p), ("out2", _p)]


_lib = None


def lib() -> C.CDLL:
    """The loaded librfx.so with typed entry points (loads on first use)."""
    global _lib
    if _lib is None:
        if not _LIB_PATH.exists():
            raise RuntimeError(f"{_LIB_PATH} is missing: build it with `python -m rfx_amd.build`")
        L = C.CDLL(str(_LIB_PATH))
        for name, (res, args) in _SIGS.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def exported_symbols() -> list[str]:
    return list(_SIGS)


def check(rc: int, what: str) -> None:
    if rc != RFX_OK:
        raise RuntimeError(f"{what} failed: {STATUS.get(rc, rc)} ({rc})")


def ptr(a) -> int | None:
    """Raw address of a numpy array or torch tensor (None passes NULL)."""
    if a is None:
        return None
    if hasattr(a, "data_ptr"):
        return a.data_ptr()
    return a.ctypes.data
