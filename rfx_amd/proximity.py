"""GPU proximity path over the C ABI: leaf assignment, co-occurrence counts,
finishprox.  torch provides device memory and the current HIP stream; all
arithmetic runs in librfx's HIP kernels.
"""
from __future__ import annotations

import numpy as np
import torch

from . import _lib
from ._lib import check, lib
from .forest import Forest


def _stream() -> int:
    return torch.cuda.current_stream().cuda_stream


def _require_gpu() -> None:
    if not torch.cuda.is_available():
        raise RuntimeError("the rfx_amd proximity path runs on an MI355X GPU; no GPU is visible")


def _dptr(t: torch.Tensor | None) -> int | None:
    return None if t is None else t.data_ptr()


def upper_size(n: int) -> int:
    return n * (n + 1) // 2


class DeviceForest:
    """A forest and its sample matrix resident in HBM.

    x is stored column-major (feature m of sample i at m*n + i) so that a
    traversal step of 64 consecutive samples reads one feature column; nodes
    are the packed 8-byte records; the bootstrap is an in-bag bitmask.
    """

    def __init__(self, forest: Forest, x: np.ndarray, cat: np.ndarray | None = None, device: str = "cuda:0",
                 chain_min_len: int = 256):
        _require_gpu()
        if forest.total_nodes > _lib.MAX_TOTAL_NODES:
            raise ValueError("forest too large for 32-bit leaf keys")
        x = np.asarray(x, np.float32)
        if x.shape != (forest.n, forest.p):
            raise ValueError(f"x shape {x.shape} does not match forest ({forest.n}, {forest.p})")
        self.forest = forest
        self.n, self.p, self.ntree = forest.n, forest.p, forest.ntree
        self.total_nodes = forest.total_nodes
        dev = torch.device(device)
        self.device = dev
        self.x_col = torch.from_numpy(np.ascontiguousarray(x.T)).to(dev)
        packed = forest.packed_nodes()
        self.nodes = torch.from_numpy(packed.view(np.int32)).to(dev)
        # heavy paths of at least chain_min_len internal nodes, 0: none.  At
        # 256, covtype-like trees at n = 200,000 (depths below 256) keep the
        # one-pass kernel, which is 1.4 % faster there than the queue.
        ci = forest.chain_index(chain_min_len, packed=packed) if chain_min_len > 0 else None
        self.chains = None if ci is None else torch.from_numpy(ci).to(dev)
        self.node_off = torch.from_numpy(forest.node_offset).to(dev)
        self.inbag = torch.from_numpy(forest.inbag_bits().view(np.int32)).to(dev)
        self.cat = None
        if cat is not None and np.any(np.asarray(cat) != 1):
            self.cat = torch.from_numpy(np.ascontiguousarray(cat, np.int32)).to(dev)

    def leaf_assign(self, keys: torch.Tensor | None = None, nodextr: torch.Tensor | None = None,
                    want_nodextr: bool = False, queue: bool = True,
                    queue_bytes: int | None = None) -> tuple[torch.Tensor, torch.Tensor | None]:
        """keys [T, n] int32 ((global node << 1) | in-bag, -1 off-leaf) and,
        optionally, nodextr [T, n] (the reference's node ids).  When the trees
        hold long chains (a chain index was built) and queue=True, the long
        walks are finished from a queue in a device workspace with jumps along
        the chains (rfx_leaf_assign_ws, queue_bytes of workspace, default
        rfx_leaf_workspace_bytes); otherwise the one-pass kernel.  Both give
        the same outputs."""
        if keys is None:
            keys = torch.empty((self.ntree, self.n), dtype=torch.int32, device=self.device)
        if want_nodextr and nodextr is None:
            nodextr = torch.empty((self.ntree, self.n), dtype=torch.int32, device=self.device)
        if not queue or self.chains is None:
            check(lib().rfx_leaf_assign(_stream(), _dptr(self.x_col), self.n, self.p, _dptr(self.nodes),
                                        _dptr(self.node_off), self.ntree, _dptr(self.inbag), _dptr(self.cat),
                                        _dptr(nodextr), _dptr(keys)), "rfx_leaf_assign")
            return keys, nodextr
        nbytes = queue_bytes or int(lib().rfx_leaf_workspace_bytes(self.n, self.ntree))
        ws = getattr(self, "_leaf_ws", None)
        if ws is None or ws.numel() != nbytes:
            ws = self._leaf_ws = torch.empty(nbytes, dtype=torch.uint8, device=self.device)
        check(lib().rfx_leaf_assign_ws(_stream(), _dptr(self.x_col), self.n, self.p, _dptr(self.nodes),
                                       _dptr(self.node_off), self.ntree, _dptr(self.inbag), _dptr(self.cat),
                                       _dptr(nodextr), _dptr(keys), _dptr(self.chains), _dptr(ws), nbytes),
              "rfx_leaf_assign_ws")
        return keys, nodextr


class Accumulator:
    """Co-occurrence counts from leaf keys, with its workspace held in HBM."""

    def __init__(self, n: int, ntree: int, total_nodes: int, device: str = "cuda:0"):
        _require_gpu()
        self.n, self.ntree, self.total_nodes = n, ntree, total_nodes
        self.device = torch.device(device)
        nb = lib().rfx_prox_workspace_bytes(n, ntree, total_nodes, _lib.ALGO_BUCKET)
        self.ws = torch.empty(max(nb, 256), dtype=torch.uint8, device=self.device)

    def out_elems(self, layout: int) -> int:
        return upper_size(self.n) if layout == _lib.LAYOUT_UPPER else self.n * self.n

    def alloc(self, layout: int) -> torch.Tensor:
        return torch.empty(self.out_elems(layout), dtype=torch.int32, device=self.device)

    def build_buckets(self, keys: torch.Tensor) -> None:
        """Leaf buckets of keys into the workspace (for algo | ALGO_PREBUILT)."""
        check(lib().rfx_prox_build_buckets(_stream(), _dptr(keys), self.n, self.ntree, self.total_nodes,
                                           _dptr(self.ws), self.ws.numel()), "rfx_prox_build_buckets")

    def casewise(self, keys: torch.Tensor, nin: torch.Tensor, out: torch.Tensor | None = None,
                 accumulate: bool = False) -> torch.Tensor:
        """Case-wise weighted proximity (the reference's use_casewise): FP64
        n x n column-major as a flat tensor, row = the out-of-bag case.  nin:
        device int32 [T, n] bootstrap multiplicities."""
        if nin.dtype != torch.int32 or nin.numel() != self.n * self.ntree:
            raise ValueError("nin must be an int32 [T, n] tensor")
        if out is None:
            out = torch.empty(self.n * self.n, dtype=torch.float64, device=self.device)
            accumulate = False
        need = lib().rfx_prox_casewise_workspace_bytes(self.n, self.ntree, self.total_nodes)
        ws = self.ws if self.ws.numel() >= need else torch.empty(need, dtype=torch.uint8, device=self.device)
        check(lib().rfx_prox_casewise(_stream(), _dptr(keys), _dptr(nin), self.n, self.ntree, self.total_nodes,
                                      int(accumulate), _dptr(out), _dptr(ws), ws.numel()), "rfx_prox_casewise")
        return out

    def rfgap(self, nin: torch.Tensor, nodextr: torch.Tensor, node_off: torch.Tensor,
              out: torch.Tensor | None = None) -> torch.Tensor:
        """RF-GAP proximity (the reference's cpu_proximity_rfgap /
        gpu_proximity_rfgap): FP64 n x n, symmetric, as a flat tensor.  nin,
        nodextr: device int32 [T, n]; node_off: device int64 [T + 1]."""
        if nin.dtype != torch.int32 or nin.numel() != self.n * self.ntree:
            raise ValueError("nin must be an int32 [T, n] tensor")
        if nodextr.dtype != torch.int32 or nodextr.numel() != self.n * self.ntree:
            raise ValueError("nodextr must be an int32 [T, n] tensor")
        if node_off.dtype != torch.int64 or node_off.numel() != self.ntree + 1:
            raise ValueError("node_off must be an int64 [T + 1] tensor")
        if out is None:
            out = torch.empty(self.n * self.n, dtype=torch.float64, device=self.device)
        need = lib().rfx_prox_rfgap_workspace_bytes(self.n, self.ntree, self.total_nodes)
        ws = self.ws if self.ws.numel() >= need else torch.empty(need, dtype=torch.uint8, device=self.device)
        check(lib().rfx_prox_rfgap(_stream(), _dptr(nin), _dptr(nodextr), _dptr(node_off), self.n, self.ntree,
                                   self.total_nodes, _dptr(out), _dptr(ws), ws.numel()), "rfx_prox_rfgap")
        return out

    def accumulate(self, keys: torch.Tensor, variant: int, layout: int, algo: int = _lib.ALGO_AUTO,
                   out: torch.Tensor | None = None, accumulate: bool = False) -> torch.Tensor:
        if out is None:
            out = self.alloc(layout)
            accumulate = False
        check(lib().rfx_prox_accumulate(_stream(), _dptr(keys), self.n, self.ntree, self.total_nodes, variant, layout,
                                        algo, int(accumulate), _dptr(out), _dptr(self.ws), self.ws.numel()),
              "rfx_prox_accumulate")
        return out

    def accumulate_pair(self, keys: torch.Tensor, layout: int, algo: int = _lib.ALGO_AUTO,
                        out_all: torch.Tensor | None = None, out_oob: torch.Tensor | None = None,
                        accumulate: bool = False) -> tuple[torch.Tensor, torch.Tensor]:
        if out_all is None or out_oob is None:
            out_all, out_oob = self.alloc(layout), self.alloc(layout)
            accumulate = False
        check(lib().rfx_prox_accumulate_pair(_stream(), _dptr(keys), self.n, self.ntree, self.total_nodes, layout,
                                             algo, int(accumulate), _dptr(out_all), _dptr(out_oob), _dptr(self.ws),
                                             self.ws.numel()), "rfx_prox_accumulate_pair")
        return out_all, out_oob

    def rows(self, keys: torch.Tensor, variant: int, r0: int, r1: int) -> torch.Tensor:
        out = torch.empty((r1 - r0, self.n), dtype=torch.int32, device=self.device)
        check(lib().rfx_prox_rows(_stream(), _dptr(keys), self.n, self.ntree, variant, r0, r1, _dptr(out)),
              "rfx_prox_rows")
        return out

    def pair_stats(self, keys: torch.Tensor) -> float:
        import ctypes as C
        v = C.c_double()
        check(lib().rfx_prox_pair_stats(_stream(), _dptr(keys), self.n, self.ntree, self.total_nodes, _dptr(self.ws),
                                        self.ws.numel(), C.byref(v)), "rfx_prox_pair_stats")
        return v.value


def finishprox(counts: torch.Tensor, in_layout: int, nout: torch.Tensor, n: int,
               out_layout: int = _lib.LAYOUT_FULL) -> torch.Tensor:
    """rf::cpu_finishprox on the GPU (bit-identical float64)."""
    elems = upper_size(n) if out_layout == _lib.LAYOUT_UPPER else n * n
    out = torch.empty(elems, dtype=torch.float64, device=counts.device)
    check(lib().rfx_finishprox(_stream(), _dptr(counts), in_layout, _dptr(nout), n, _dptr(out), out_layout),
          "rfx_finishprox")
    return out


def finishprox_f64(prox: torch.Tensor, nout: torch.Tensor, n: int, out_layout: int = _lib.LAYOUT_FULL) -> torch.Tensor:
    """rf::cpu_finishprox of an FP64 accumulated matrix (FULL column-major,
    e.g. case-wise weights) on the GPU, bit-identical."""
    elems = upper_size(n) if out_layout == _lib.LAYOUT_UPPER else n * n
    out = torch.empty(elems, dtype=torch.float64, device=prox.device)
    check(lib().rfx_finishprox_f64(_stream(), _dptr(prox), _dptr(nout), n, _dptr(out), out_layout),
          "rfx_finishprox_f64")
    return out


def mds_dense_matrix(prox: torch.Tensor, n: int, nout: torch.Tensor | None = None,
                     precision: int = _lib.PREC_FP64, max_iter: int = 1000,
                     tol: float = 1e-10) -> tuple[torch.Tensor, np.ndarray, int]:
    """3-D classical MDS of a given FP64 n x n proximity (FULL column-major),
    with the reference's nout renormalisation when nout is given."""
    import ctypes as C

    nb = lib().rfx_mds_dense_workspace_bytes(n, precision)
    ws = torch.empty(max(nb, 256), dtype=torch.uint8, device=prox.device)
    X = torch.empty((n, 3), dtype=torch.float64, device=prox.device)
    eig = np.zeros(16, np.float64)
    it = C.c_int32()
    check(lib().rfx_mds_dense_matrix(_stream(), _dptr(prox), n, _dptr(nout) if nout is not None else None, precision,
                                     max_iter, tol, _dptr(ws), ws.numel(), _dptr(X), eig.ctypes.data, C.byref(it)),
          "rfx_mds_dense_matrix")
    return X, eig, it.value


def mds_dense_rfgap(prox: torch.Tensor, n: int, nout: torch.Tensor, precision: int = _lib.PREC_FP64,
                    max_iter: int = 1000, tol: float = 1e-10) -> tuple[torch.Tensor, np.ndarray, int]:
    """3-D classical MDS of a given FP64 RF-GAP proximity with the
    reference's use_rfgap renormalisation by nout."""
    import ctypes as C

    nb = lib().rfx_mds_dense_workspace_bytes(n, precision)
    ws = torch.empty(max(nb, 256), dtype=torch.uint8, device=prox.device)
    X = torch.empty((n, 3), dtype=torch.float64, device=prox.device)
    eig = np.zeros(16, np.float64)
    it = C.c_int32()
    check(lib().rfx_mds_dense_rfgap(_stream(), _dptr(prox), n, _dptr(nout), precision, max_iter, tol, _dptr(ws),
                                    ws.numel(), _dptr(X), eig.ctypes.data, C.byref(it)), "rfx_mds_dense_rfgap")
    return X, eig, it.value


def outlier_scores(prox: torch.Tensor, n: int) -> tuple[torch.Tensor, float]:
    """Outlier scores of a finished FP64 n x n proximity (the reference's
    RandomForest::compute_outlier_scores): float32 [n] and the median used."""
    import ctypes as C

    if prox.dtype != torch.float64 or prox.numel() != n * n:
        raise ValueError("prox must be a float64 n x n tensor")
    out = torch.empty(n, dtype=torch.float32, device=prox.device)
    med = C.c_float()
    check(lib().rfx_outlier_scores(_stream(), _dptr(prox), n, _dptr(out), C.byref(med)), "rfx_outlier_scores")
    return out, med.value


def unpack_upper(u: np.ndarray, n: int) -> np.ndarray:
    """Packed row-major upper triangle -> symmetric dense (host helper)."""
    m = np.zeros((n, n), dtype=u.dtype)
    iu = np.triu_indices(n)
    m[iu] = u
    m.T[iu] = u
    return m


def mds_dense(counts: torch.Tensor, layout: int, n: int, mode: int = _lib.MDS_COUNTS, scale: float = 1.0,
              nout: torch.Tensor | None = None, precision: int = _lib.PREC_FP64, max_iter: int = 500,
              tol: float | None = None) -> tuple[torch.Tensor, np.ndarray, int]:
    """Classical 3-D MDS on the GPU from a count matrix; returns (coords n x 3
    float64 on the device, the 16 Ritz values, iterations)."""
    import ctypes as C

    if tol is None:
        tol = 1e-10 if precision == _lib.PREC_FP64 else 1e-5
    nb = lib().rfx_mds_dense_workspace_bytes(n, precision)
    ws = torch.empty(nb, dtype=torch.uint8, device=counts.device)
    coords = torch.empty((n, 3), dtype=torch.float64, device=counts.device)
    eig = np.zeros(16, np.float64)
    it = C.c_int32(0)
    check(lib().rfx_mds_dense(_stream(), _dptr(counts), layout, n, mode, scale, _dptr(nout), precision, max_iter, tol,
                              _dptr(ws), nb, _dptr(coords), eig.ctypes.data, C.addressof(it)), "rfx_mds_dense")
    return coords, eig, it.value


def mds_dense_many(jobs: list[dict], streams: int = 8) -> list[tuple[torch.Tensor, np.ndarray, int]]:
    """The 3-D MDS of many independent count matrices (the per-instance MDS of
    a forest batch), run concurrently: `streams` host threads, each with a
    HIP stream of its own (default priority), take the jobs in order of
    decreasing n, so the small problems' host round trips (Gram matrices and
    Rayleigh-Ritz on the host) and short launches overlap with the large
    problems' kernels.  Each job is the keyword arguments of mds_dense
    (counts, layout, n, and optionally mode, scale, nout, precision,
    max_iter, tol); results come back in job order, as mds_dense returns
    them, with the device synchronised."""
    import concurrent.futures as cf
    import threading

    _require_gpu()
    if not jobs:
        return []
    dev = jobs[0]["counts"].device
    ready = torch.cuda.Event()
    ready.record(torch.cuda.current_stream(dev))
    local = threading.local()

    def run(job):
        s = getattr(local, "stream", None)
        if s is None:
            s = local.stream = torch.cuda.Stream(device=dev)
            s.wait_event(ready)
        with torch.cuda.stream(s):
            r = mds_dense(**job)
        s.synchronize()
        return r

    order = sorted(range(len(jobs)), key=lambda k: -int(jobs[k]["n"]))
    out: list = [None] * len(jobs)
    with cf.ThreadPoolExecutor(max_workers=max(1, streams)) as ex:
        futs = {ex.submit(run, jobs[k]): k for k in order}
        for f in cf.as_completed(futs):
            out[futs[f]] = f.result()
    torch.cuda.synchronize(dev)
    return out


_FMT_DTYPE = {_lib.PREC_FP32: torch.float32, _lib.PREC_BF16: torch.int16, _lib.PREC_INT8: torch.int8,
              _lib.PREC_FP8: torch.uint8}


def lowrank_factors(keys: torch.Tensor, forest: Forest, rank: int = 64, oversample: int = 16, power_iters: int = 2,
                    seed: int = 1) -> tuple[torch.Tensor, np.ndarray]:
    """U (n x rank, float32, device) with C/T ~ U U^T, and the top eigenvalues of C/T."""
    n, T = forest.n, forest.ntree
    nb = lib().rfx_lowrank_workspace_bytes(n, T, forest.total_nodes, rank, oversample)
    ws = torch.empty(nb, dtype=torch.uint8, device=keys.device)
    U = torch.empty((n, rank), dtype=torch.float32, device=keys.device)
    eig = np.zeros(rank, np.float64)
    off = np.ascontiguousarray(forest.node_offset, np.int64)
    check(lib().rfx_lowrank_factors(_stream(), _dptr(keys), n, T, forest.total_nodes, off.ctypes.data, rank,
                                    oversample, power_iters, seed, _dptr(ws), nb, _dptr(U), eig.ctypes.data),
          "rfx_lowrank_factors")
    return U, eig


def quantize_factors(U: torch.Tensor, fmt: int) -> tuple[torch.Tensor, torch.Tensor]:
    """Storage copy of the factors in `fmt` plus per-column float32 scales."""
    n, r = U.shape
    Uq = torch.empty((n, r), dtype=_FMT_DTYPE[fmt], device=U.device)
    sc = torch.empty(r, dtype=torch.float32, device=U.device)
    check(lib().rfx_quantize_factors(_stream(), _dptr(U), n, r, fmt, _dptr(Uq), _dptr(sc)), "rfx_quantize_factors")
    return Uq, sc


def mds_factors(Uq: torch.Tensor, scales: torch.Tensor, fmt: int, precision: int = _lib.PREC_FP32,
                max_iter: int = 500, tol: float | None = None) -> tuple[torch.Tensor, np.ndarray, int]:
    import ctypes as C

    n, r = Uq.shape
    if tol is None:
        tol = 1e-10 if precision == _lib.PREC_FP64 else 1e-5
    nb = lib().rfx_mds_factors_workspace_bytes(n, r, precision)
    ws = torch.empty(nb, dtype=torch.uint8, device=Uq.device)
    coords = torch.empty((n, 3), dtype=torch.float64, device=Uq.device)
    eig = np.zeros(16, np.float64)
    it = C.c_int32(0)
    check(lib().rfx_mds_factors(_stream(), _dptr(Uq), _dptr(scales), fmt, n, r, precision, max_iter, tol, _dptr(ws),
                                nb, _dptr(coords), eig.ctypes.data, C.addressof(it)), "rfx_mds_factors")
    return coords, eig, it.value


def mds_factors_mixed(Uq: torch.Tensor, scales: torch.Tensor, fmt: int, U32: torch.Tensor | None = None,
                      max_iter: int = 500, tol: float = 1e-4, refine_max: int = 60,
                      refine_tol: float = 5e-5) -> tuple[torch.Tensor, np.ndarray, tuple[int, int]]:
    """3-D MDS from stored factors with the bf16 matrix-core operator, then (with
    the FP32 factors U32) refinement sweeps with the FP32-accurate matrix-core
    operator from the converged block.
    Returns (coords n x 3 float64, the 16 Ritz values, (MFMA iterations,
    refinement sweeps))."""
    import ctypes as C

    n, r = Uq.shape
    nb = lib().rfx_mds_factors_mixed_workspace_bytes(n, r)
    ws = torch.empty(nb, dtype=torch.uint8, device=Uq.device)
    coords = torch.empty((n, 3), dtype=torch.float64, device=Uq.device)
    eig = np.zeros(16, np.float64)
    it = (C.c_int32 * 2)()
    check(lib().rfx_mds_factors_mixed(_stream(), _dptr(Uq), _dptr(scales), fmt, _dptr(U32), n, r, max_iter, tol,
                                      refine_max if U32 is not None else 0, refine_tol, _dptr(ws), nb, _dptr(coords),
                                      eig.ctypes.data, C.addressof(it)), "rfx_mds_factors_mixed")
    return coords, eig, (it[0], it[1])
