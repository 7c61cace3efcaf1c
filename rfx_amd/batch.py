"""A10: co-occurrence counts of many independent forests as one batch.

A ForestBatch plans leaf assignment and UPPER int32 counts of all its
forests as one leaf launch, one leaf-bucket sort and one persistent gather
launch fed by a device work queue (rfx_batch_* in include/rfx_prox.h), and
can replay that sequence as a hipGraph.  The reference has no multi-forest
batching (fit_batch_gpu, src/rf_random_forest.cpp:1364-1570, batches the
trees of one forest).
"""
from __future__ import annotations

import ctypes as C

import torch

from . import _lib
from ._lib import check, lib
from .proximity import DeviceForest, _dptr, _require_gpu, _stream, upper_size

_VARIANTS = {"all": _lib.PROX_ALL, "oob_upper": _lib.PROX_OOB_UPPER, "pair": _lib.PROX_PAIR}


class ForestBatch:
    """Counts of every forest in `forests` (all-numeric DeviceForests).

    variant: "oob_upper" (the reference's OOB fold, V3), "all", or "pair"
    (ALL into out[b], OOB_UPPER into out2[b]).  Outputs are UPPER int32
    tensors of n(n+1)/2, overwritten by every run.
    """

    def __init__(self, forests: list[DeviceForest], variant: str = "oob_upper", device: str = "cuda:0"):
        _require_gpu()
        if variant not in _VARIANTS:
            raise ValueError(f"variant must be one of {sorted(_VARIANTS)}")
        for f in forests:
            if f.cat is not None:
                raise ValueError("ForestBatch handles all-numeric forests only")
        self.forests = list(forests)
        self.variant = variant
        self.device = torch.device(device)
        self.out = [torch.empty(upper_size(f.n), dtype=torch.int32, device=self.device) for f in forests]
        self.out2 = ([torch.empty(upper_size(f.n), dtype=torch.int32, device=self.device) for f in forests]
                     if variant == "pair" else [None] * len(forests))
        probs = (_lib.BatchProblem * len(forests))()
        for k, f in enumerate(forests):
            probs[k] = _lib.BatchProblem(_dptr(f.x_col), f.n, f.p, f.ntree, _dptr(f.nodes), _dptr(f.node_off),
                                         f.total_nodes, _dptr(f.inbag), _dptr(self.out[k]), _dptr(self.out2[k]))
        self._probs = probs
        h = C.c_void_p()
        with torch.cuda.device(self.device):
            check(lib().rfx_batch_create(probs, len(forests), _VARIANTS[variant], C.byref(h)), "rfx_batch_create")
        self._h = h
        nbytes = lib().rfx_batch_workspace_bytes(h)
        self.workspace = torch.empty(max(nbytes, 1), dtype=torch.uint8, device=self.device)

    def run(self, use_graph: bool = False) -> None:
        check(lib().rfx_batch_run(_stream(), self._h, _dptr(self.workspace), self.workspace.numel(),
                                  1 if use_graph else 0), "rfx_batch_run")

    def autotune(self, reps: int = 2) -> float:
        ms = C.c_double()
        check(lib().rfx_batch_autotune(_stream(), self._h, _dptr(self.workspace), self.workspace.numel(), reps,
                                       C.byref(ms)), "rfx_batch_autotune")
        return ms.value

    def set_tuning(self, lines_per_item: int, blocks_per_cu: int = 0) -> None:
        check(lib().rfx_batch_set_tuning(self._h, lines_per_item, blocks_per_cu), "rfx_batch_set_tuning")

    def tuning(self) -> tuple[int, int]:
        a, b = C.c_int32(), C.c_int32()
        check(lib().rfx_batch_get_tuning(self._h, C.byref(a), C.byref(b)), "rfx_batch_get_tuning")
        return a.value, b.value

    def set_algos(self, algos: list[int]) -> None:
        """Per-forest algorithm: _lib.ALGO_GATHER or _lib.ALGO_DENSE."""
        import numpy as np

        a = np.ascontiguousarray(algos, np.int32)
        if a.shape != (len(self.forests),):
            raise ValueError("one algorithm per forest")
        check(lib().rfx_batch_set_algos(self._h, a.ctypes.data), "rfx_batch_set_algos")

    def own_launch(self):
        """Per forest: True when autotune gave it its own kernel launch, False
        when it goes through the persistent queue."""
        import numpy as np

        f = np.zeros(len(self.forests), np.int32)
        check(lib().rfx_batch_get_own_launch(self._h, f.ctypes.data), "rfx_batch_get_own_launch")
        return f.astype(bool)

    def set_own_launch(self, flags) -> None:
        """Override the autotuner's per-forest own-launch flags (all False:
        every forest through the persistent queue)."""
        import numpy as np

        f = np.ascontiguousarray(flags, np.int32)
        if f.shape != (len(self.forests),):
            raise ValueError("one flag per forest")
        check(lib().rfx_batch_set_own_launch(self._h, f.ctypes.data), "rfx_batch_set_own_launch")

    def stats(self):
        """(algorithm per forest, measured sum |leaf|^2 per forest)."""
        import numpy as np

        a = np.zeros(len(self.forests), np.int32)
        q = np.zeros(len(self.forests), np.float64)
        check(lib().rfx_batch_get_stats(self._h, a.ctypes.data, q.ctypes.data), "rfx_batch_get_stats")
        return a, q

    def work(self) -> int:
        """Tree-pair updates of one run: sum over forests of T * n(n+1)/2."""
        return sum(f.ntree * upper_size(f.n) for f in self.forests)

    def close(self) -> None:
        if getattr(self, "_h", None):
            lib().rfx_batch_free(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
