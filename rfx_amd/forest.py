"""Host-side forest: grow (reference CPU semantics), export, pack for the GPU.

The forest is the input of the proximity path.  It is grown by the restated
reference grower in librfx (rfx_forest_grow), bit-identical to
rf::cpu_growtree on the same data and seeds, and exported as the reference's
own per-node arrays (nodestatus, bestvar, xbestsplit, treemap) flattened over
trees by ``node_offset``.  A forest can equally be built from arrays produced
elsewhere (``Forest.from_arrays``), e.g. exported by the reference itself.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _lib
from ._lib import check, lib, ptr


def make_dataset(kind: int, n: int, p: int, nclass: int = 0, seed: int = 0) -> tuple[np.ndarray, np.ndarray]:
    """Deterministic synthetic dataset: x row-major float32 (n, p), y int32."""
    x = np.empty((n, p), dtype=np.float32)
    y = np.empty(n, dtype=np.int32)
    check(lib().rfx_make_dataset(kind, n, p, nclass, seed, ptr(x), ptr(y)), "rfx_make_dataset")
    return x, y


def default_mtry(p: int) -> int:
    """mtry = int(sqrt(float(p))) as the reference classifier sets it
    (python/randomforest_py.cpp:661)."""
    return max(1, int(np.sqrt(np.float32(p))))


@dataclass
class Forest:
    n: int
    p: int
    node_offset: np.ndarray  # int64 [T+1]
    nodestatus: np.ndarray  # int32 [total]
    bestvar: np.ndarray  # int32 [total]
    xbestsplit: np.ndarray  # float32 [total]
    treemap: np.ndarray  # int32 [2*total]
    nodeclass: np.ndarray  # int32 [total]
    nin: np.ndarray  # int32 [T, n]

    @property
    def ntree(self) -> int:
        return len(self.node_offset) - 1

    @property
    def total_nodes(self) -> int:
        return int(self.node_offset[-1])

    def nnode(self, t: int) -> int:
        return int(self.node_offset[t + 1] - self.node_offset[t])

    @classmethod
    def grow(cls, x: np.ndarray, y: np.ndarray, ntree: int, mtry: int = 0, minndsize: int = 1, maxnode: int = 0,
             iseed: int = 12345, boot_seed: int = 12345, fma: bool = True, nthreads: int = 0,
             plain: bool = False, progress=None, progress_every: float = 20.0) -> "Forest":
        """Grow with the reference CPU grower's semantics.  plain=True uses
        per-node sorts instead of presorted columns (same trees, slower).
        progress(trees_done, ntree) is called about every progress_every
        seconds while the trees grow."""
        x = np.ascontiguousarray(x, dtype=np.float32)
        y = np.ascontiguousarray(y, dtype=np.int32)
        n, p = x.shape
        nclass = int(y.max()) + 1
        if nclass < 2:
            nclass = 2
        if mtry <= 0:
            mtry = default_mtry(p)
        h = C.c_void_p()
        flags = (0 if fma else _lib.GROW_NO_FMA) | (_lib.GROW_PLAIN if plain else 0)
        if progress is None:
            check(lib().rfx_forest_grow(ptr(x), ptr(y), n, p, nclass, ntree, mtry, minndsize, maxnode, iseed,
                                        boot_seed, flags, nthreads, C.byref(h)), "rfx_forest_grow")
        else:
            # The call runs on a helper thread (ctypes releases the GIL); this
            # thread reports the finished-tree counter it maintains.
            import threading

            done = C.c_int32(0)
            rc = []
            th = threading.Thread(target=lambda: rc.append(lib().rfx_forest_grow_progress(
                ptr(x), ptr(y), n, p, nclass, ntree, mtry, minndsize, maxnode, iseed, boot_seed, flags, nthreads,
                C.byref(done), C.byref(h))))
            th.start()
            last = -1
            while th.is_alive():
                th.join(progress_every)
                if done.value != last:
                    last = done.value
                    progress(last, ntree)
            check(rc[0], "rfx_forest_grow_progress")
            if done.value != last:
                progress(done.value, ntree)
        try:
            L = lib()
            T = L.rfx_forest_ntree(h)
            total = sum(L.rfx_forest_nnode(h, t) for t in range(T))
            off = np.empty(T + 1, np.int64)
            st = np.empty(total, np.int32)
            bv = np.empty(total, np.int32)
            xs = np.empty(total, np.float32)
            tm = np.empty(2 * total, np.int32)
            nc = np.empty(total, np.int32)
            nin = np.empty((T, n), np.int32)
            check(L.rfx_forest_export(h, ptr(off), ptr(st), ptr(bv), ptr(xs), ptr(tm), ptr(nc), ptr(nin)),
                  "rfx_forest_export")
        finally:
            lib().rfx_forest_free(h)
        return cls(n, p, off, st, bv, xs, tm, nc, nin)

    @classmethod
    def grow_chunked(cls, x: np.ndarray, y: np.ndarray, ntree: int, chunk: int = 64, progress=None,
                     iseed: int = 12345, boot_seed: int = 12345, **kw) -> "Forest":
        """The forest Forest.grow would return, grown chunk trees at a time
        (tree t keeps its seeds iseed + t and boot_seed + t), calling
        progress(trees_done, ntree) after each chunk."""
        parts = []
        for t0 in range(0, ntree, chunk):
            c = min(chunk, ntree - t0)
            parts.append(cls.grow(x, y, c, iseed=iseed + t0, boot_seed=boot_seed + t0, **kw))
            if progress:
                progress(t0 + c, ntree)
        return cls.concat(parts)

    @classmethod
    def concat(cls, parts: list["Forest"]) -> "Forest":
        """The trees of several forests over the same samples, in order."""
        n, p = parts[0].n, parts[0].p
        if any(f.n != n or f.p != p for f in parts):
            raise ValueError("forests over different samples")
        off = [np.zeros(1, np.int64)]
        base = 0
        for f in parts:
            off.append(f.node_offset[1:] + base)
            base += f.total_nodes
        cat = lambda k: np.concatenate([getattr(f, k) for f in parts])
        return cls(n, p, np.concatenate(off), cat("nodestatus"), cat("bestvar"), cat("xbestsplit"), cat("treemap"),
                   cat("nodeclass"), np.concatenate([f.nin for f in parts], axis=0))

    @classmethod
    def from_arrays(cls, n: int, p: int, nnode: np.ndarray, nodestatus: np.ndarray, bestvar: np.ndarray,
                    xbestsplit: np.ndarray, treemap: np.ndarray, nin: np.ndarray, maxnode: int | None = None,
                    nodeclass: np.ndarray | None = None) -> "Forest":
        """Build from per-tree arrays padded to maxnode (the reference's storage,
        src/rf_random_forest.cpp:985-991) or already flattened (maxnode None)."""
        nnode = np.asarray(nnode, np.int64)
        T = len(nnode)
        off = np.zeros(T + 1, np.int64)
        off[1:] = np.cumsum(nnode)
        if maxnode is None:
            st, bv, xs, tm = nodestatus, bestvar, xbestsplit, treemap
        else:
            sel = np.concatenate([t * maxnode + np.arange(nnode[t]) for t in range(T)])
            sel2 = np.concatenate([2 * t * maxnode + np.arange(2 * nnode[t]) for t in range(T)])
            st, bv, xs = nodestatus.reshape(-1)[sel], bestvar.reshape(-1)[sel], xbestsplit.reshape(-1)[sel]
            tm = treemap.reshape(-1)[sel2]
        nc = np.zeros(int(off[-1]), np.int32) if nodeclass is None else np.asarray(nodeclass, np.int32)
        return cls(n, p, off, np.ascontiguousarray(st, np.int32), np.ascontiguousarray(bv, np.int32),
                   np.ascontiguousarray(xs, np.float32), np.ascontiguousarray(tm, np.int32), nc,
                   np.ascontiguousarray(nin, np.int32).reshape(T, n))

    def packed_nodes(self) -> np.ndarray:
        out = np.empty(2 * self.total_nodes, np.uint32)
        check(lib().rfx_pack_nodes(ptr(self.node_offset), self.ntree, ptr(self.nodestatus), ptr(self.bestvar),
                                   ptr(self.xbestsplit), ptr(self.treemap), self.p, ptr(out)), "rfx_pack_nodes")
        return out

    def chain_index(self, min_len: int = 32, packed: np.ndarray | None = None) -> np.ndarray | None:
        """Chain index (rfx_chain_index) of the heavy paths of at least min_len
        internal nodes, as bytes; None when no tree has one."""
        packed = self.packed_nodes() if packed is None else np.ascontiguousarray(packed, np.uint32)
        need = C.c_size_t(0)
        check(lib().rfx_chain_index(ptr(self.node_offset), self.ntree, ptr(packed), min_len, None, C.byref(need)),
              "rfx_chain_index")
        if need.value == 0:
            return None
        out = np.empty(need.value, np.uint8)
        check(lib().rfx_chain_index(ptr(self.node_offset), self.ntree, ptr(packed), min_len, ptr(out), C.byref(need)),
              "rfx_chain_index")
        return out

    def inbag_bits(self) -> np.ndarray:
        W = (self.n + 31) // 32
        out = np.empty(self.ntree * W, np.uint32)
        check(lib().rfx_pack_inbag_bits(ptr(self.nin), self.n, self.ntree, ptr(out)), "rfx_pack_inbag_bits")
        return out

    def nout(self, mult: int = 2) -> np.ndarray:
        """OOB counts; mult=2 reproduces the reference CPU path's nout."""
        out = np.empty(self.n, np.int32)
        check(lib().rfx_oob_counts(ptr(self.nin), self.n, self.ntree, mult, ptr(out)), "rfx_oob_counts")
        return out

    def subset(self, trees: slice) -> "Forest":
        """The forest restricted to a contiguous range of trees."""
        t0, t1, _ = trees.indices(self.ntree)
        b, e = int(self.node_offset[t0]), int(self.node_offset[t1])
        off = self.node_offset[t0:t1 + 1] - b
        return Forest(self.n, self.p, off, self.nodestatus[b:e], self.bestvar[b:e], self.xbestsplit[b:e],
                      self.treemap[2 * b:2 * e], self.nodeclass[b:e], self.nin[t0:t1])

    def samples(self, idx: np.ndarray) -> "Forest":
        """The same trees with only the samples idx (ascending): their bootstrap
        counts are kept, so each keeps its in-bag / out-of-bag status per tree."""
        idx = np.asarray(idx, np.int64)
        return Forest(len(idx), self.p, self.node_offset, self.nodestatus, self.bestvar, self.xbestsplit,
                      self.treemap, self.nodeclass, np.ascontiguousarray(self.nin[:, idx]))

    # Forest file: an uncompressed .npz of the flattened arrays above plus
    # n, p and a format tag; read back with allow_pickle=False, so loading
    # executes nothing from the file.
    FORMAT = "rfx_amd.forest.v1"

    def save(self, path) -> None:
        with open(path, "wb") as fh:
            np.savez(fh, format=np.array(self.FORMAT), n=np.int64(self.n), p=np.int64(self.p),
                     node_offset=self.node_offset, nodestatus=self.nodestatus, bestvar=self.bestvar,
                     xbestsplit=self.xbestsplit, treemap=self.treemap, nodeclass=self.nodeclass, nin=self.nin)

    @classmethod
    def load(cls, path) -> "Forest":
        with np.load(path, allow_pickle=False) as z:
            if str(z["format"]) != cls.FORMAT:
                raise ValueError(f"{path}: not a {cls.FORMAT} forest file")
            n, p = int(z["n"]), int(z["p"])
            off = np.ascontiguousarray(z["node_offset"], np.int64)
            f = cls(n, p, off, np.ascontiguousarray(z["nodestatus"], np.int32),
                    np.ascontiguousarray(z["bestvar"], np.int32), np.ascontiguousarray(z["xbestsplit"], np.float32),
                    np.ascontiguousarray(z["treemap"], np.int32), np.ascontiguousarray(z["nodeclass"], np.int32),
                    np.ascontiguousarray(z["nin"], np.int32).reshape(len(off) - 1, n))
        total = int(off[-1])
        if (len(f.nodestatus), len(f.bestvar), len(f.xbestsplit), len(f.nodeclass), len(f.treemap)) != \
                (total, total, total, total, 2 * total):
            raise ValueError(f"{path}: node arrays do not match node_offset")
        return f
