"""RandomForestClassifier mirror of the reference pybind class, proximity path.

Keeps the reference's constructor options (python/randomforest_py.cpp:617-647,
1352-1385, same names, order and defaults) and its proximity/MDS methods
(1114-1318: get_proximity_matrix, compute_proximity_matrix,
get_lowrank_factors, compute_mds_3d_cpu, compute_mds_from_factors,
compute_mds_3d_from_factors) with the reference's return dtypes and shapes,
backed by the HIP path in librfx:

  fit                  grows the forest on the host exactly as the reference
                       CPU grower (bootstrap seed 12345 + t, grower seed
                       iseed + t), categorical auto-detection as the
                       reference's fit (771-803), then on the GPU: leaf
                       assignment, OOB_UPPER counts (or, with use_qlora, the
                       rank-`rank` factors of C/T) and nout = 2 * #OOB trees.
  get_proximity_matrix float32 n x n of rf::cpu_finishprox's result
                       (1138-1146); raises in low-rank mode as the reference.
  compute_mds_3d_cpu   the reference chain's 3-D MDS (OOB counts ->
                       finishprox -> float32 -> nout renormalisation ->
                       classical MDS), computed on the GPU in FP64.
  compute_mds_from_factors / compute_mds_3d_from_factors
                       MDS from the stored factors on the matrix cores,
                       refined with the FP32 factors (k <= 3, rank <= 128).

With use_casewise the proximity is the reference's case-wise weighted one
(src/rf_proximity.cpp:186-223: an out-of-bag case and an in-bag case of a
leaf add nin(j) / the leaf's bootstrap draws), accumulated in FP64 and fed
through the same finishprox and MDS chain.

With use_rfgap the proximity is the reference's RF-GAP one
(src/rf_random_forest.cpp:1661-1712 -> rf::gpu_proximity_rfgap): FP64,
symmetric, already normalised; get_proximity_matrix returns it as float32 and
compute_mds_3d_cpu applies the reference's use_rfgap renormalisation by nout
(src/rf_mds_cpu.cpp:125-153).  RF-GAP takes precedence over use_casewise, as
in the reference's finalisation.

Options outside the proximity path (importance, prediction, regression
losses) are accepted for signature compatibility; the methods outside the
path raise NotImplementedError.
use_gpu is accepted: this implementation always runs on the MI355X.
"""
from __future__ import annotations

import numpy as np

from . import _lib

# reference quant_mode names -> stored factor format here
QUANT_FORMATS = {
    "fp32": _lib.PREC_FP32,
    "bf16": _lib.PREC_BF16,
    "fp16": _lib.PREC_BF16,  # 16-bit storage: bfloat16 here
    "int8": _lib.PREC_INT8,
    "fp8": _lib.PREC_FP8,
    "nf4": _lib.PREC_FP8,  # the lowest-precision storage here: FP8 e4m3 with per-column scales
}


def detect_categorical(x: np.ndarray, maxcat: int) -> np.ndarray:
    """Categorical auto-detection of the classifier's fit
    (python/randomforest_py.cpp:771-803): a column whose values are all within
    1e-5 of an integer, with 2..maxcat distinct values, min >= 0 and
    max < maxcat, gets cat = number of distinct values; else 1 (numeric)."""
    x = np.asarray(x, np.float32)
    n, p = x.shape
    cat = np.ones(p, np.int32)
    for m in range(p):
        col = x[:, m]
        rounded = np.round(col)
        all_int = bool(np.all(np.abs(col - rounded) <= np.float32(1e-5)))
        nu = len(np.unique(col))
        if all_int and 1 < nu <= maxcat and col.min() >= 0.0 and col.max() < np.float32(maxcat):
            cat[m] = nu
    return cat


class RandomForestClassifier:
    def __init__(self, nsample: int = 1000, ntree: int = 100, mdim: int = 10, nclass: int = 2, maxcat: int = 10,
                 mtry: int = 0, maxnode: int = 0, minndsize: int = 1, ncsplit: int = 25, ncmax: int = 25,
                 iseed: int = 12345, compute_proximity: bool = False, compute_importance: bool = True,
                 compute_local_importance: bool = False, use_gpu: bool = False, use_qlora: bool = False,
                 quant_mode: str = "nf4", use_sparse: bool = False, sparsity_threshold: float = 1e-6,
                 batch_size: int = 0, nodesize: int = 5, cutoff: float = 0.01, show_progress: bool = True,
                 progress_desc: str = "Training Random Forest", gpu_loss_function: str = "gini", rank: int = 32,
                 n_threads_cpu: int = 0, use_rfgap: bool = False, importance_method: str = "local_imp",
                 clique_M: int = -1, use_casewise: bool = False):
        if use_rfgap and use_qlora:
            raise NotImplementedError("RF-GAP applies to the full proximity, not the low-rank factors")
        if use_casewise and use_qlora:
            raise NotImplementedError("case-wise weights apply to the full proximity, not the low-rank factors")
        if quant_mode not in QUANT_FORMATS:
            raise ValueError(f"quant_mode must be one of {sorted(QUANT_FORMATS)}")
        if gpu_loss_function != "gini":
            raise NotImplementedError("only the classification (gini) grower is part of this path")
        self.nsample, self.ntree, self.mdim, self.nclass = nsample, ntree, mdim, nclass
        self.maxcat, self.mtry, self.maxnode, self.minndsize = maxcat, mtry, maxnode, minndsize
        self.iseed = iseed
        self.compute_proximity = compute_proximity
        self.use_gpu, self.use_qlora, self.quant_mode = use_gpu, use_qlora, quant_mode
        self.rank = rank
        self.use_casewise = use_casewise
        self.use_rfgap = use_rfgap
        self.n_threads_cpu = n_threads_cpu
        self.batch_size = batch_size
        # accepted for signature compatibility, outside the proximity path
        self._unused = dict(ncsplit=ncsplit, ncmax=ncmax, compute_importance=compute_importance,
                            compute_local_importance=compute_local_importance, use_sparse=use_sparse,
                            sparsity_threshold=sparsity_threshold, nodesize=nodesize, cutoff=cutoff,
                            show_progress=show_progress, progress_desc=progress_desc,
                            importance_method=importance_method, clique_M=clique_M)
        self.forest = None
        self.cat = None
        self._counts = None  # device OOB_UPPER counts (UPPER int32)
        self._prox_w = None  # device case-wise FP64 n x n (FULL), use_casewise
        self._prox_g = None  # device RF-GAP FP64 n x n (FULL, symmetric), use_rfgap
        self._nout = None
        self._factors = None  # (stored factors, scales, format, U float32)

    # ------------------------------------------------------------------ fit
    def fit(self, X: np.ndarray, y: np.ndarray, sample_weight=None) -> "RandomForestClassifier":
        import torch

        from .forest import Forest
        from .proximity import Accumulator, DeviceForest, lowrank_factors, quantize_factors

        X = np.ascontiguousarray(X, np.float32)
        y = np.asarray(y)
        if y.dtype != np.int32:
            raise TypeError("y must be an int32 array (the reference reinterprets the buffer as int32)")
        n, p = X.shape
        self.nsample, self.mdim = n, p
        self.nclass = max(self.nclass, int(y.max()) + 1)
        self.cat = detect_categorical(X, self.maxcat) if self.maxcat > 1 else np.ones(p, np.int32)
        self.forest = Forest.grow(X, y, self.ntree, mtry=self.mtry, minndsize=self.minndsize, maxnode=self.maxnode,
                                  iseed=self.iseed, nthreads=self.n_threads_cpu)
        self._counts = self._factors = self._prox_w = self._prox_g = None
        if not self.compute_proximity:
            return self
        f = self.forest
        df = DeviceForest(f, X, self.cat)
        keys, nodextr = df.leaf_assign(want_nodextr=self.use_rfgap)
        self._nout = torch.from_numpy(f.nout(2)).to(df.device)
        if self.use_qlora:
            rank = min(self.rank, 128, n)
            U, _ = lowrank_factors(keys, f, rank=rank)
            fmt = QUANT_FORMATS[self.quant_mode]
            Uq, sc = quantize_factors(U, fmt)
            self._factors = (Uq, sc, fmt, U)
        elif self.use_rfgap:
            acc = Accumulator(n, f.ntree, f.total_nodes, device=str(df.device))
            nin = torch.from_numpy(np.ascontiguousarray(f.nin.reshape(f.ntree, n), np.int32)).to(df.device)
            self._prox_g = acc.rfgap(nin, nodextr, df.node_off)
        elif self.use_casewise:
            acc = Accumulator(n, f.ntree, f.total_nodes, device=str(df.device))
            nin = torch.from_numpy(np.ascontiguousarray(f.nin.reshape(f.ntree, n), np.int32)).to(df.device)
            self._prox_w = acc.casewise(keys, nin)
        else:
            acc = Accumulator(n, f.ntree, f.total_nodes, device=str(df.device))
            self._counts = acc.accumulate(keys, _lib.PROX_OOB_UPPER, _lib.LAYOUT_UPPER)
        torch.cuda.synchronize()
        return self

    def _need_fit(self, what: str) -> None:
        if self.forest is None:
            raise RuntimeError(f"Model must be fitted before calling {what}()")

    # ------------------------------------------------------------ proximity
    def get_proximity_matrix(self) -> np.ndarray:
        from .proximity import finishprox, finishprox_f64, unpack_upper

        self._need_fit("get_proximity_matrix")
        if self._factors is not None:
            raise RuntimeError("the proximity is held as low-rank factors (use_qlora=True): use "
                               "get_lowrank_factors() or compute_mds_from_factors()")
        if self._prox_g is not None:
            n = self.nsample
            return self._prox_g.float().cpu().numpy().reshape(n, n, order="F")
        if self._prox_w is not None:
            n = self.nsample
            S = finishprox_f64(self._prox_w, self._nout, n, _lib.LAYOUT_FULL)
            return S.cpu().numpy().reshape(n, n, order="F").astype(np.float32)
        if self._counts is None:
            raise RuntimeError("Proximity matrix not available. Set compute_proximity=True when fitting.")
        n = self.nsample
        S = finishprox(self._counts, _lib.LAYOUT_UPPER, self._nout, n, _lib.LAYOUT_UPPER)
        return unpack_upper(S.cpu().numpy(), n).astype(np.float32)

    compute_proximity_matrix = get_proximity_matrix

    def get_lowrank_factors(self) -> tuple[np.ndarray, np.ndarray, int]:
        self._need_fit("get_lowrank_factors")
        if self._factors is None:
            raise RuntimeError("low-rank factors need compute_proximity=True and use_qlora=True")
        U = self._factors[3].cpu().numpy().astype(np.float64)
        return U, U.copy(), U.shape[1]

    # ------------------------------------------------------------------ MDS
    def compute_mds_3d_cpu(self) -> np.ndarray:
        from .proximity import finishprox_f64, mds_dense, mds_dense_matrix, mds_dense_rfgap

        self._need_fit("compute_mds_3d_cpu")
        if self._prox_g is not None:
            # float32 at the boundary, then the use_rfgap renormalisation and MDS
            X, _, _ = mds_dense_rfgap(self._prox_g.float().double(), self.nsample, self._nout)
            return X.cpu().numpy()
        if self._prox_w is not None:
            # the reference chain: finishprox, float32 at the boundary, nout renormalisation, MDS
            S = finishprox_f64(self._prox_w, self._nout, self.nsample, _lib.LAYOUT_FULL)
            X, _, _ = mds_dense_matrix(S.float().double(), self.nsample, self._nout)
            return X.cpu().numpy()
        if self._counts is None:
            raise RuntimeError("Proximity matrix not available. Set compute_proximity=True when fitting.")
        X, _, _ = mds_dense(self._counts, _lib.LAYOUT_UPPER, self.nsample, _lib.MDS_REFERENCE, nout=self._nout,
                            precision=_lib.PREC_FP64)
        return X.cpu().numpy()

    def compute_mds_from_factors(self, k: int = 3) -> np.ndarray:
        """MDS from the stored factors on the matrix cores, refined with the FP32 factors."""
        from .proximity import mds_factors_mixed

        self._need_fit("compute_mds_from_factors")
        if self._factors is None:
            raise RuntimeError("compute_mds_from_factors needs use_qlora=True and compute_proximity=True")
        if not 1 <= k <= 3:
            raise ValueError("k must be 1, 2 or 3")
        Uq, sc, fmt, U = self._factors
        if U.shape[1] > 128:
            raise ValueError("the factor MDS supports rank <= 128")
        X, _, _ = mds_factors_mixed(Uq, sc, fmt, U)
        return X.cpu().numpy()[:, :k].copy()

    def compute_mds_3d_from_factors(self) -> np.ndarray:
        return self.compute_mds_from_factors(3)

    # --------------------------------------------------------------- shape
    def get_n_samples(self) -> int:
        return self.nsample

    def get_n_features(self) -> int:
        return self.mdim

    def get_n_classes(self) -> int:
        return self.nclass

    def _outside(self, *a, **k):
        raise NotImplementedError("outside the proximity path of this implementation")

    predict = predict_proba = get_oob_error = feature_importances_ = get_local_importance = _outside
    get_oob_predictions = _outside
