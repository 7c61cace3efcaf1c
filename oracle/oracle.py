"""Python face of the oracle (TEST INFRASTRUCTURE ONLY).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this module, and only as the checker or the timed CPU port; the product
(rfx_amd) never does.

  * ctypes bindings of liboracle.so (oracle.c: the C restatement of the
    reference's leaf assignment, V1 accumulation, integer counts, finishprox)
  * ctypes bindings of _ref/librfxref.so (the reference's own CPU sources)
  * classical MDS in float64 with numpy/scipy (textbook Torgerson scaling),
    the restatement of src/rf_mds_cpu.cpp:88-341 and the error metrics.
"""
from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
ORACLE_SO = HERE / "liboracle.so"
REF_SO = HERE / "_ref" / "librfxref.so"

ALL, OOB_ASYM, OOB_UPPER, OOB_ROW = 0, 1, 2, 4

_p = C.c_void_p
_i32 = C.c_int32
_i64 = C.c_int64


def build(ref: bool = False) -> None:
    cmd = ["make", "-s", "-C", str(HERE), "liboracle.so"]
    if ref:
        cmd.append("ref")
    subprocess.run(cmd, check=True)


_orc = None
_ref = None


def orc() -> C.CDLL:
    global _orc
    if _orc is None:
        if not ORACLE_SO.exists():
            build()
        L = C.CDLL(str(ORACLE_SO))
        L.orc_leaf_assign.argtypes = [_p, _i64, _i32, _i32, _p, _p, _p, _p, _p, _p, _p, _p]
        L.orc_prox_v1.argtypes = [_p, _p, _p, _p, _i64, _i32, _p]
        L.orc_prox_v1_casewise.argtypes = [_p, _p, _p, _p, _i64, _i32, _p]
        L.orc_prox_rfgap.argtypes = [_p, _p, _i64, _i32, _p]
        L.orc_outlier_scores.argtypes = [_p, _i64, _p]
        L.orc_counts_bucket.argtypes = [_p, _p, _p, _p, _i64, _i32, _i32, _i32, _p]
        L.orc_counts_brute.argtypes = [_p, _p, _p, _p, _i64, _i32, _i32, _p]
        L.orc_finishprox.argtypes = [_i64, _p, _p, _p]
        L.orc_counts_rows.argtypes = [_p, _p, _p, _p, _i64, _i32, _p, _i32, _p, _p]
        L.orc_counts_rows.restype = _i32
        for f in (L.orc_leaf_assign, L.orc_prox_v1, L.orc_prox_v1_casewise, L.orc_prox_rfgap, L.orc_outlier_scores,
                  L.orc_counts_bucket,
                  L.orc_counts_brute, L.orc_finishprox):
            f.restype = _i32
        L.orc_num_threads.restype = _i32
        _orc = L
    return _orc


def ref_available() -> bool:
    return REF_SO.exists()


def ref() -> C.CDLL:
    global _ref
    if _ref is None:
        L = C.CDLL(str(REF_SO))
        L.ref_run_forest.argtypes = [_p, _p, _i32, _i32, _i32, _i32, _i32, _i32, _i32, _p, _i32, _i32,
                                     _p, _p, _p, _p, _p, _p, _p, _p, _p, _i32]
        L.ref_run_forest.restype = _i32
        L.ref_finishprox.argtypes = [_i32, _p, _p, _p]
        L.ref_finishprox.restype = _i32
        L.ref_proximity_rfgap.argtypes = [_i32, _i32, _p, _p, _p]
        L.ref_proximity_rfgap.restype = _i32
        L.ref_rf_fit.argtypes = [_p, _p, _i32, _i32, _i32, _i32, _i32, _i32, _i32, _p, _i32, _i32, _i32, _p, _p, _p]
        L.ref_rf_fit.restype = _i32
        L.ref_mds_3d.argtypes = [_p, _i32, _p, _i32, _p]
        L.ref_mds_3d.restype = _i32
        _ref = L
    return _ref


def _a(x):
    return None if x is None else x.ctypes.data


def _ok(rc: int, what: str) -> None:
    if rc != 0:
        raise RuntimeError(f"{what} -> {rc}")


# ---------------------------------------------------------------------------
# restatement
# ---------------------------------------------------------------------------
def leaf_assign(forest, x: np.ndarray, cat: np.ndarray | None = None) -> np.ndarray:
    """nodextr [T, n]: node reached per the reference rules (see oracle.c)."""
    x = np.ascontiguousarray(x, np.float32)
    out = np.empty((forest.ntree, forest.n), np.int32)
    cat = None if cat is None else np.ascontiguousarray(cat, np.int32)
    _ok(orc().orc_leaf_assign(_a(x), forest.n, forest.p, forest.ntree, _a(forest.node_offset), _a(forest.nodestatus),
                              _a(forest.bestvar), _a(forest.xbestsplit), _a(forest.treemap), _a(forest.nin), _a(cat),
                              _a(out)), "orc_leaf_assign")
    return out


def prox_v1(forest, nodextr: np.ndarray, casewise: bool = False) -> np.ndarray:
    """Accumulated FP64 n x n (column-major as a flat array, shape (n, n) F-order);
    casewise: the reference's use_casewise weights nin(kk) / leaf draws."""
    n = forest.n
    prox = np.zeros(n * n, np.float64)
    fn = orc().orc_prox_v1_casewise if casewise else orc().orc_prox_v1
    _ok(fn(_a(forest.nodestatus), _a(forest.node_offset), _a(np.ascontiguousarray(nodextr)), _a(forest.nin), n,
           forest.ntree, _a(prox)), "orc_prox_v1")
    return prox


def prox_rfgap(nin: np.ndarray, nodextr: np.ndarray) -> np.ndarray:
    """RF-GAP proximity, FP64 n x n (symmetric) as a flat array; nin and
    nodextr are [T, n] int32."""
    nin = np.ascontiguousarray(nin, np.int32)
    nodextr = np.ascontiguousarray(nodextr, np.int32)
    T, n = nin.shape
    prox = np.empty(n * n, np.float64)
    _ok(orc().orc_prox_rfgap(_a(nin), _a(nodextr), n, T, _a(prox)), "orc_prox_rfgap")
    return prox


def outlier_scores(prox: np.ndarray) -> np.ndarray:
    """Outlier scores (float32 [n]) of a finished FP64 n x n proximity, given
    as the flat buffer the reference holds."""
    prox = np.ascontiguousarray(prox, np.float64).reshape(-1)
    n = int(round(np.sqrt(prox.size)))
    out = np.empty(n, np.float32)
    _ok(orc().orc_outlier_scores(_a(prox), n, _a(out)), "orc_outlier_scores")
    return out


def counts(forest, nodextr: np.ndarray, variant: int, layout: int) -> np.ndarray:
    n = forest.n
    size = n * n if layout == 0 else n * (n + 1) // 2
    out = np.zeros(size, np.int32)
    _ok(orc().orc_counts_bucket(_a(forest.nodestatus), _a(forest.node_offset), _a(np.ascontiguousarray(nodextr)),
                                _a(forest.nin), n, forest.ntree, variant, layout, _a(out)), "orc_counts_bucket")
    return out


def counts_brute(forest, nodextr: np.ndarray, variant: int) -> np.ndarray:
    n = forest.n
    out = np.zeros(n * n, np.int32)
    _ok(orc().orc_counts_brute(_a(forest.nodestatus), _a(forest.node_offset), _a(np.ascontiguousarray(nodextr)),
                               _a(forest.nin), n, forest.ntree, variant, _a(out)), "orc_counts_brute")
    return out


def counts_rows(forest, nodextr: np.ndarray, rows) -> tuple[np.ndarray, np.ndarray]:
    """Full rows (len(rows) x n) of the ALL and OOB_UPPER count matrices."""
    rows = np.ascontiguousarray(rows, np.int64)
    n = forest.n
    a = np.empty((len(rows), n), np.int32)
    o = np.empty((len(rows), n), np.int32)
    nd = np.ascontiguousarray(nodextr, np.int32)
    _ok(orc().orc_counts_rows(_a(forest.nodestatus), _a(forest.node_offset), _a(nd), _a(forest.nin), n,
                              forest.ntree, _a(rows), len(rows), _a(a), _a(o)), "orc_counts_rows")
    return a, o


def finishprox(n: int, nout: np.ndarray, prox: np.ndarray) -> np.ndarray:
    out = np.empty(n * n, np.float64)
    _ok(orc().orc_finishprox(n, _a(np.ascontiguousarray(nout, np.int32)), _a(np.ascontiguousarray(prox)), _a(out)),
        "orc_finishprox")
    return out


def num_threads() -> int:
    return orc().orc_num_threads()


# ---------------------------------------------------------------------------
# the reference itself (compiled from its sources into _ref/)
# ---------------------------------------------------------------------------
def ref_run_forest(x: np.ndarray, y: np.ndarray, ntree: int, mtry: int, minndsize: int = 1, iseed: int = 12345,
                   cat: np.ndarray | None = None, maxcat: int = 1, do_prox: bool = True,
                   casewise: bool = False) -> dict:
    x = np.ascontiguousarray(x, np.float32)
    y = np.ascontiguousarray(y, np.int32)
    n, p = x.shape
    nclass = max(2, int(y.max()) + 1)
    maxnode = 2 * n + 1
    r = dict(
        nnode=np.zeros(ntree, np.int32), nodestatus=np.zeros((ntree, maxnode), np.int32),
        bestvar=np.zeros((ntree, maxnode), np.int32), xbestsplit=np.zeros((ntree, maxnode), np.float32),
        treemap=np.zeros((ntree, 2 * maxnode), np.int32), nin=np.zeros((ntree, n), np.int32),
        nodextr=np.zeros((ntree, n), np.int32), nout=np.zeros(n, np.int32),
        prox=np.zeros(n * n, np.float64) if do_prox else None,
    )
    cat = None if cat is None else np.ascontiguousarray(cat, np.int32)
    _ok(ref().ref_run_forest(_a(x), _a(y), n, p, nclass, ntree, mtry, minndsize, iseed, _a(cat), maxcat, int(do_prox),
                             _a(r["nnode"]), _a(r["nodestatus"]), _a(r["bestvar"]), _a(r["xbestsplit"]),
                             _a(r["treemap"]), _a(r["nin"]), _a(r["nodextr"]), _a(r["nout"]), _a(r["prox"]),
                             int(casewise)),
        "ref_run_forest")
    r["maxnode"] = maxnode
    return r


def ref_finishprox(n: int, nout: np.ndarray, prox: np.ndarray) -> np.ndarray:
    out = np.empty(n * n, np.float64)
    p = np.array(prox, np.float64, copy=True)
    _ok(ref().ref_finishprox(n, _a(np.ascontiguousarray(nout, np.int32)), _a(p), _a(out)), "ref_finishprox")
    return out


# ---------------------------------------------------------------------------
# classical MDS, float64
# ---------------------------------------------------------------------------
def mds_input_reference(S: np.ndarray, nout: np.ndarray) -> np.ndarray:
    """The matrix compute_mds_3d_cpu embeds when the pybind path hands it the
    float32-rounded finishprox output together with nout
    (python/randomforest_py.cpp:1138-1146, 1225-1243; src/rf_mds_cpu.cpp:154-186):
    M_ij = 0.5 * (f32(S_ij)/nout_i + f32(S_ji)/nout_j), M_ii = 1."""
    Sf = S.astype(np.float32).astype(np.float64)
    d = np.where(nout > 0, nout.astype(np.float64), 1.0)
    R = Sf / d[:, None]
    R = np.where((nout > 0)[:, None], R, Sf)
    M = 0.5 * (R + R.T)
    np.fill_diagonal(M, 1.0)
    return M


def mds_input_rfgap(P: np.ndarray, nout: np.ndarray) -> np.ndarray:
    """The matrix compute_mds_3d_cpu embeds for an RF-GAP proximity
    (src/rf_mds_cpu.cpp:125-153 with use_rfgap, as called at
    python/randomforest_py.cpp:1236-1244 on the float32 matrix of
    get_proximity_matrix): rows divided by nout_i where nout_i > 0, then
    M_ij = 0.5 * (R_ij + R_ji) off the diagonal, M_ii = R_ii."""
    Pf = P.astype(np.float32).astype(np.float64)
    d = np.where(nout > 0, nout.astype(np.float64), 1.0)
    R = Pf / d[:, None]
    R = np.where((nout > 0)[:, None], R, Pf)
    M = 0.5 * (R + R.T)
    np.fill_diagonal(M, np.diag(R))
    return M


def classical_mds(prox: np.ndarray, k: int = 3, eig_floor: float = 1e-10) -> tuple[np.ndarray, np.ndarray]:
    """src/rf_mds_cpu.cpp:196-338 in float64: D = max(prox) - prox with zero
    diagonal, B = -1/2 J D^2 J, top-k eigenpairs with lambda > floor, sorted
    descending, X = V sqrt(lambda).  Returns (X n x k, all top eigenvalues)."""
    from scipy.linalg import eigh

    P = np.asarray(prox, np.float64)
    n = P.shape[0]
    D = P.max() - P
    np.fill_diagonal(D, 0.0)
    D2 = D * D
    r = D2.mean(axis=1)
    c = D2.mean(axis=0)
    g = r.mean()
    B = -0.5 * (D2 - r[:, None] - c[None, :] + g)
    m = min(n, k + 5)
    w, V = eigh(B, subset_by_index=[n - m, n - 1])
    order = np.argsort(w)[::-1]
    w, V = w[order], V[:, order]
    keep = w > eig_floor
    if keep.sum() < k:
        raise RuntimeError("Not enough positive eigenvalues for MDS")
    w, V = w[keep], V[:, keep]
    return V[:, :k] * np.sqrt(w[:k]), w


def embedding_errors(X: np.ndarray, Xref: np.ndarray) -> dict:
    """Sign-aligned relative Frobenius error and largest principal angle."""
    X = np.asarray(X, np.float64)
    Xref = np.asarray(Xref, np.float64)
    s = np.sign(np.sum(X * Xref, axis=0))
    s[s == 0] = 1.0
    rel = float(np.linalg.norm(X * s - Xref) / np.linalg.norm(Xref))
    qa, _ = np.linalg.qr(X)
    qb, _ = np.linalg.qr(Xref)
    sv = np.linalg.svd(qa.T @ qb, compute_uv=False)
    ang = float(np.arccos(np.clip(sv.min(), -1.0, 1.0)))
    return {"rel_frobenius": rel, "max_principal_angle": ang}


def ref_proximity_rfgap(nin: np.ndarray, nodextr: np.ndarray) -> np.ndarray:
    """The reference's own rf::cpu_proximity_rfgap on [T, n] arrays."""
    nin = np.ascontiguousarray(nin, np.int32)
    nodextr = np.ascontiguousarray(nodextr, np.int32)
    T, n = nin.shape
    prox = np.empty(n * n, np.float64)
    _ok(ref().ref_proximity_rfgap(T, n, _a(nin), _a(nodextr), _a(prox)), "ref_proximity_rfgap")
    return prox


def ref_rf_fit(x: np.ndarray, y: np.ndarray | None, ntree: int, mtry: int, minndsize: int = 1, iseed: int = 12345,
               cat: np.ndarray | None = None, maxcat: int = 1, unsupervised: bool = False,
               casewise: bool = False) -> dict:
    """The reference's RandomForest class fitted on the CPU
    (src/rf_random_forest.cpp: fit -> grow_tree_single -> finalize_training):
    its finished proximity (float64, as get_proximity_matrix returns it), nout
    and, for the unsupervised task, compute_outlier_scores."""
    x = np.ascontiguousarray(x, np.float32)
    n, p = x.shape
    if unsupervised:
        yy, nclass = None, 2
    else:
        # RandomForest::setup_classification_task sets nclass to the number of
        # distinct labels but keeps the labels themselves
        # (src/rf_random_forest.cpp:1855-1864), so labels with gaps index past
        # the grower's class arrays (src/rf_growtree.cpp:638, 717): hand it
        # contiguous labels (a monotone relabelling leaves the trees unchanged).
        _, yy = np.unique(np.asarray(y), return_inverse=True)
        yy = np.ascontiguousarray(yy, np.int32)
        nclass = max(2, int(yy.max()) + 1)
    cat = None if cat is None else np.ascontiguousarray(cat, np.int32)
    r = dict(prox=np.empty(n * n, np.float64), nout=np.empty(n, np.int32),
             outlier=np.zeros(n, np.float32) if unsupervised else None)
    _ok(ref().ref_rf_fit(_a(x), _a(yy), n, p, nclass, ntree, mtry, minndsize, iseed, _a(cat), maxcat,
                         2 if unsupervised else 0, int(casewise), _a(r["prox"]), _a(r["nout"]), _a(r["outlier"])),
        "ref_rf_fit")
    return r


def ref_mds_3d(prox: np.ndarray, nout: np.ndarray | None, use_rfgap: bool = False) -> np.ndarray:
    """The reference's rf::compute_mds_3d_cpu (src/rf_mds_cpu.cpp:88-341) on a
    column-major float64 n x n proximity, with dsyevd from this image's LAPACK."""
    P = np.ascontiguousarray(np.asarray(prox, np.float64).ravel())
    n = int(round(np.sqrt(P.size)))
    out = np.empty((n, 3), np.float64)
    nv = None if nout is None else np.ascontiguousarray(nout, np.int32)
    _ok(ref().ref_mds_3d(_a(P), n, _a(nv), int(use_rfgap), _a(out)), "ref_mds_3d")
    return out
