#!/usr/bin/env python3
"""Benchmark of the MI355X proximity path: one JSON line on stdout.

A step is one pass of the hot path over the resident workload: leaf
assignment of every sample through every tree, then the ALL and OOB upper
int32 count matrices (both exact).  The default workload is the metric's
configuration: the covtype-like problem at n = 200,000 samples, p = 54
features, 7 classes, T = 1000 trees, mtry = 7, the forest grown on the host
by the restated reference grower (growth is setup, not timed).  After the
timed steps it adds the 3-D MDS section: the exact dense MDS of C/T in FP64
and FP32 at full n, rank-64 factors, and the factor MDS in each storage
format, with errors against FP64.

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload c3|c2|c1|c5]
                  [--algo auto|dense|bucket|gather] [--forest-cache PATH]

value = T * n(n+1)/2 tree-pair updates per second (each tree-pair slot counts
once although it feeds both matrices).  The line carries the roofline of the
dominant kernel (HIP events on the launch stream) and a CPU baseline: the
oracle's restatement of the reference's OpenMP path (leaf assignment +
rf::cpu_proximity's accumulation) timed on a bounded sample (at n = 200,000,
on a seeded 20,000-sample subset of the samples through the same forest).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "prox tree·pair updates/s; 3D-MDS s at n=200k, BF16/INT8 vs FP32 @ stated err"
HBM_PEAK_GBS = 8000.0

WORKLOADS = {
    # name: (dataset kind, n, p, classes, trees, mtry, description)
    "c2": (1, 20000, 54, 7, 500, 7, "covtype-like n=20000 p=54 T=500 mtry=7: leaf assignment + ALL and OOB "
                                    "upper-triangle int32 counts"),
    "c1": (0, 2000, 20, 3, 100, 4, "3-class Gaussian mixture n=2000 p=20 T=100 mtry=4: leaf assignment + ALL and "
                                   "OOB upper-triangle int32 counts"),
    "c3": (1, 200000, 54, 7, 1000, 7, "covtype-like n=200000 p=54 T=1000 mtry=7: leaf assignment + ALL and OOB "
                                      "upper-triangle int32 counts; rank-64 factors; factor 3-D MDS per storage "
                                      "format"),
    "c4": (2, 1000000, 32, 10, 500, 6, "16-component Gaussian mixture in a random 8-dim subspace + 24 N(0,1) noise "
                                       "columns, n=1000000 p=32 10 classes T=500 mtry=6 minndsize=5: leaf "
                                       "assignment, rank-128 factors of C/T from the leaf buckets (matrix-free), "
                                       "3-D MDS from FP8-e4m3 / BF16 / FP32 factor storage"),
    "c5": (0, 0, 0, 0, 0, 0, "64 independent forests (instance i: seed i, n log-uniform in [1000, 50000], p "
                             "uniform in [8, 64], T uniform in [50, 500], 2-10 classes, Gaussian mixture, "
                             "mtry = floor(sqrt(p))): leaf assignment + OOB upper-triangle int32 counts of all 64 "
                             "as one batch"),
}


def log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


def measured_traffic(key: str):
    """HBM bytes per launch of the counting kernel for this kernel and shape,
    as measured by the rocprofv3 counter runs summarised under profiles/
    (None when that kernel and shape have not been measured)."""
    path = ROOT / "profiles" / "hbm_traffic.json"
    if not path.exists():
        return None, None
    e = json.loads(path.read_text()).get(key)
    if not e:
        return None, None
    b = (e["fetch_kib"] * e["fetch_scale"] + e["write_kib"] * e["write_scale"]) * 1024
    return int(b), e["summary"]


def _repeat_for(fn, min_seconds: float) -> tuple[int, float]:
    """Run fn until at least min_seconds have elapsed; (passes, seconds)."""
    passes, t0 = 0, time.perf_counter()
    while True:
        fn()
        passes += 1
        dt = time.perf_counter() - t0
        if dt >= min_seconds:
            return passes, dt


def host_cpu() -> dict:
    """Model, sockets and cores of the host CPU (from /proc/cpuinfo)."""
    info = {"model": None, "sockets": None, "physical_cores": None, "logical_cpus": os.cpu_count()}
    try:
        phys, cores, model = set(), {}, None
        cur = None
        for ln in Path("/proc/cpuinfo").read_text().splitlines():
            k, _, v = ln.partition(":")
            k, v = k.strip(), v.strip()
            if k == "model name" and model is None:
                model = v
            elif k == "physical id":
                cur = v
                phys.add(v)
            elif k == "cpu cores" and cur is not None:
                cores[cur] = int(v)
        info.update(model=model, sockets=len(phys) or None, physical_cores=sum(cores.values()) or None)
    except OSError:
        pass
    return info


def cpu_baseline(forest, x, sample_trees: int, min_seconds: float = 10.0) -> dict:
    """The oracle port of the reference CPU path (per tree: leaf assignment,
    then rf::cpu_proximity's accumulating OpenMP branch) on sample_trees trees,
    repeated until about min_seconds of host work have been timed."""
    from oracle import oracle as O

    sub = forest.subset(slice(0, sample_trees))
    n = forest.n

    def one():
        O.prox_v1(sub, O.leaf_assign(sub, x))

    passes, dt = _repeat_for(one, min_seconds)
    P = n * (n + 1) // 2
    return {
        "value": passes * sample_trees * P / dt,
        "unit": "tree·pair updates/s",
        "cores": int(O.num_threads()),
        "kind": "port",
        "sample": f"{passes} passes over the first {sample_trees} of {forest.ntree} trees at n={n}: oracle leaf "
                  f"assignment (src/rf_testreebag.cpp, find_terminal_node) + rf::cpu_proximity dense OpenMP "
                  f"branch (src/rf_proximity.cpp:99-231) into an n x n float64 matrix; {dt:.2f} s",
        "host_cpu": host_cpu(),
        "seconds": dt,
    }


def cpu_baseline_subset(forest, x, m: int = 20000, seed: int = 0, min_seconds: float = 10.0) -> dict:
    """The same port where the reference's n x n float64 matrix cannot be held
    (320 GB at n = 200,000): a seeded subset of m samples, assigned through the
    same forest (each keeps its bootstrap counts), so the pair density, and
    with it the work per tree-pair, is that of the full problem.  The rate is
    reported; the full-size time is extrapolated from it."""
    import numpy as np

    from oracle import oracle as O

    idx = np.sort(np.random.default_rng(seed).choice(forest.n, m, replace=False))
    sub = forest.samples(idx)
    xs = np.ascontiguousarray(x[idx])
    T = forest.ntree

    def one():
        O.prox_v1(sub, O.leaf_assign(sub, xs))

    passes, dt = _repeat_for(one, min_seconds)
    rate = passes * T * (m * (m + 1) // 2) / dt
    full = T * (forest.n * (forest.n + 1) // 2) / rate
    return {
        "value": rate,
        "unit": "tree·pair updates/s",
        "cores": int(O.num_threads()),
        "kind": "port",
        "sample": f"{passes} passes over all {T} trees on a seeded {m}-sample subset (seed {seed}) of the n={forest.n} "
                  f"samples, assigned through the same forest: oracle leaf assignment + rf::cpu_proximity dense "
                  f"OpenMP branch (src/rf_proximity.cpp:99-231) into an m x m float64 matrix; {dt:.2f} s.  "
                  f"Extrapolated to the full problem: {full:.0f} s (the reference itself would need a "
                  f"{8 * forest.n ** 2 / 1e9:.0f} GB float64 matrix and overflows its int32 indices above n = 46,340)",
        "extrapolated_full_s": full,
        "host_cpu": host_cpu(),
    }


def grow_forest(kind: int, n: int, p: int, ncl: int, T: int, mtry: int, cache: str | None):
    """The workload's dataset and forest: the forest is grown by the restated
    reference grower, or read from a forest file written by an earlier run."""
    from rfx_amd.forest import Forest, make_dataset

    t0 = time.perf_counter()
    x, y = make_dataset(kind, n, p, ncl, 0)
    if cache and Path(cache).exists():
        forest = Forest.load(cache)
        if (forest.n, forest.p, forest.ntree) != (n, p, T):
            raise SystemExit(f"{cache} holds another forest ({forest.n}, {forest.p}, {forest.ntree})")
        log(f"forest read from {cache}: {forest.total_nodes} nodes in {time.perf_counter() - t0:.1f} s")
        return x, y, forest
    forest = Forest.grow(x, y, T, mtry=mtry, iseed=0)
    log(f"forest grown: {forest.total_nodes} nodes in {time.perf_counter() - t0:.1f} s")
    if cache:
        forest.save(cache)
    return x, y, forest


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="c3", choices=sorted(WORKLOADS))
    ap.add_argument("--rank", type=int, default=0, help="factor rank (c3: 64, c4: 128)")
    ap.add_argument("--trees", type=int, default=0, help="c4: trees grown (0: the workload's 500)")
    ap.add_argument("--algo", default="auto", choices=["auto", "dense", "bucket", "gather"])
    ap.add_argument("--cpu-trees", type=int, default=0, help="trees in the CPU baseline sample (0: auto)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--mds", action="store_true", help="c5: also time the 3-D MDS of every instance")
    ap.add_argument("--no-mds", action="store_true", help="c3: counts only, no factor / MDS section")
    ap.add_argument("--forest-cache", default=None,
                    help="forest file: read the forest from it if it exists, else grow and write it")
    ap.add_argument("--casewise", action="store_true",
                    help="c1/c2: also time the case-wise weighted FP64 proximity on the same keys")
    ap.add_argument("--rfgap", action="store_true",
                    help="c1/c2: also time the RF-GAP FP64 proximity of the same forest")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="first run one device-to-device copy of 1 GiB (a known-size kernel that scales the HBM "
                         "byte counters of a profiler run)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: the proximity path runs on a single MI355X here "
              "(multi-GPU sharding is not run)", file=sys.stderr)
        return 2

    import numpy as np
    import torch

    from rfx_amd import _lib
    from rfx_amd.forest import Forest, make_dataset
    from rfx_amd.proximity import Accumulator, DeviceForest

    if args.copy_calibration:
        src = torch.ones(1 << 28, dtype=torch.float32, device="cuda:0")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
        log("copy calibration: 1 GiB read + 1 GiB written")
    if args.workload == "c5":
        return bench_c5(args)
    if args.workload == "c4":
        return bench_c4(args)
    NAMES = {_lib.ALGO_DENSE: "dense", _lib.ALGO_BUCKET: "bucket", _lib.ALGO_GATHER: "gather"}
    kind, n, p, ncl, T, mtry, desc = WORKLOADS[args.workload]
    x, y, forest = grow_forest(kind, n, p, ncl, T, mtry, args.forest_cache)

    dev = torch.device("cuda:0")
    df = DeviceForest(forest, x)
    acc = Accumulator(n, T, forest.total_nodes)
    keys = torch.empty((T, n), dtype=torch.int32, device=dev)
    out_all = acc.alloc(_lib.LAYOUT_UPPER)
    out_oob = acc.alloc(_lib.LAYOUT_UPPER)
    df.leaf_assign(keys=keys)
    sq = acc.pair_stats(keys)
    algo = {"auto": _lib.ALGO_AUTO, "dense": _lib.ALGO_DENSE, "bucket": _lib.ALGO_BUCKET,
            "gather": _lib.ALGO_GATHER}[args.algo]
    if algo == _lib.ALGO_AUTO:
        # resolve AUTO once, outside the timed region, by timing each algorithm
        cand = {}
        # row-segment GATHER and the T*n^2 DENSE compare are only candidates
        # at moderate n (see the AUTO rule in accum.hip)
        # (beyond that the global-atomic BUCKET form is 30x slower than
        # GATHER: 2.2 s against 0.07 s per call at n = 200,000, T = 1000)
        small = n <= 65536
        algos = (_lib.ALGO_DENSE, _lib.ALGO_BUCKET, _lib.ALGO_GATHER) if small else (_lib.ALGO_GATHER,)
        reps = 3 if small else 1
        for a in algos:
            acc.accumulate_pair(keys, _lib.LAYOUT_UPPER, a, out_all, out_oob)
            torch.cuda.synchronize()
            s = time.perf_counter()
            for _ in range(reps):
                acc.accumulate_pair(keys, _lib.LAYOUT_UPPER, a, out_all, out_oob)
            torch.cuda.synchronize()
            cand[a] = (time.perf_counter() - s) / reps
        algo = min(cand, key=cand.get)
        log("algo timings (s): " + " ".join(f"{NAMES[a]} {v:.5f}" for a, v in cand.items()))
    algo_name = NAMES[algo]

    stream = torch.cuda.current_stream()
    ev = [tuple(torch.cuda.Event(enable_timing=True) for _ in range(3)) for _ in range(args.steps)]
    sparse = algo in (_lib.ALGO_GATHER, _lib.ALGO_BUCKET)

    # a step: leaf assignment, then one accumulate call (which builds the leaf
    # buckets it needs and runs the counting kernel), each bracketed by events
    # on the launch stream
    def step(e=None):
        if e:
            e[0].record(stream)
        df.leaf_assign(keys=keys)
        if e:
            e[1].record(stream)
        acc.accumulate_pair(keys, _lib.LAYOUT_UPPER, algo, out_all, out_oob)
        if e:
            e[2].record(stream)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    t_start = time.perf_counter()
    for k in range(args.steps):
        step(ev[k])
    torch.cuda.synchronize()
    wall = time.perf_counter() - t_start
    ms_step = wall * 1e3 / args.steps
    leaf_ms = float(np.mean([a.elapsed_time(b) for a, b, _ in ev]))
    acc_ms = float(np.mean([b.elapsed_time(c) for _, b, c in ev]))
    # after the timed steps: the full bucket build as its own call and the
    # counting kernel alone on those buckets (the kernel's duration for the
    # roofline)
    build_ms, kern_ms = 0.0, acc_ms
    if sparse:
        kev = [tuple(torch.cuda.Event(enable_timing=True) for _ in range(3)) for _ in range(5)]
        for a, b, c in kev:
            a.record(stream)
            acc.build_buckets(keys)
            b.record(stream)
            acc.accumulate_pair(keys, _lib.LAYOUT_UPPER, algo | _lib.ALGO_PREBUILT, out_all, out_oob)
            c.record(stream)
        torch.cuda.synchronize()
        build_ms = float(np.mean([a.elapsed_time(b) for a, b, _ in kev]))
        kern_ms = float(np.mean([b.elapsed_time(c) for _, b, c in kev]))
    P = n * (n + 1) // 2
    value = T * P / (ms_step / 1e3)

    # Algorithmic bytes of the counting launch (one pass over all trees):
    # two int32 upper triangles written + the int32 key array read.
    # the GATHER kernel run_gather (accum.hip) picks: whole lines in LDS up to
    # n = 20476 and T = 4096, line chunks up to n = 32768, the chunk-boundary
    # table walk up to 16 boundaries of 16,384 columns (T <= 2048), else the
    # gallop walk
    gather_kernel = ("gather_rows_kernel<PAIR,UPPER>" if n <= 20476 and T <= 4096 else
                     "gather_line_kernel<PAIR,UPPER>" if n <= 32768 else
                     "gather_window_kernel<PAIR,UPPER>" if (n + 16383) // 16384 - 1 <= 16 and T <= 2048 else
                     "gather_gallop_kernel<PAIR,UPPER>")
    KERNEL = {_lib.ALGO_GATHER: gather_kernel, _lib.ALGO_BUCKET: "bucket_pair_kernel<PAIR,UPPER>",
              _lib.ALGO_DENSE: "dense_tile_kernel<PAIR,UPPER>"}
    acc_bytes = 2 * 4 * P + 4 * n * T
    achieved = acc_bytes / (kern_ms / 1e3) / 1e9
    traffic, traffic_src = measured_traffic(f"{KERNEL[algo]} n={n} T={T}")
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                "kernel": KERNEL[algo], "algorithmic_bytes": acc_bytes, "avg_ms": round(kern_ms, 4)}
    if traffic_src:
        roofline["traffic_source"] = traffic_src

    line = {
        "metric": METRIC, "value": value, "unit": "tree·pair updates/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms_step, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "int32", "data": "synthetic",
        "config": {"workload": desc, "n": n, "p": p, "ntree": T, "mtry": mtry, "algo": algo_name,
                   "features": "all numeric (the one-hot columns are 0/1 floats, maxcat 1: no categorical splits)",
                   "layout": "upper", "leaf_assign_ms": round(leaf_ms, 4), "accumulate_ms": round(acc_ms, 4),
                   "bucket_build_call_ms": round(build_ms, 4), "count_kernel_ms": round(kern_ms, 4),
                   "sum_leaf_sq": sq, "pair_density": sq / (T * float(n) * n)},
        "roofline": roofline,
    }
    if not args.no_cpu and n > 50000:
        # the reference CPU path holds an n x n float64 matrix (320 GB at n = 200,000)
        line["cpu_baseline"] = cpu_baseline_subset(forest, x)
    elif not args.no_cpu:
        trees = args.cpu_trees or T
        cb = cpu_baseline(forest, x, trees)
        cb.pop("seconds")
        line["cpu_baseline"] = cb
    if args.casewise and n <= 50000:
        line["casewise"] = casewise_section(acc, keys, forest, x, n, T, not args.no_cpu)
    if args.rfgap and n <= 50000:
        line["rfgap"] = rfgap_section(acc, df, forest, x, n, T, not args.no_cpu)
    if args.workload == "c3" and not args.no_mds:
        del keys
        out_oob = None
        torch.cuda.empty_cache()
        line["mds"] = factor_mds_section(df, forest, x, y, out_all, args.rank or 64, not args.no_cpu)
        out_all = None
    print(json.dumps(line), flush=True)
    return 0


def casewise_section(acc, keys, forest, x, n: int, T: int, cpu: bool) -> dict:
    """The case-wise weighted FP64 proximity (full n x n) on the same keys:
    one call (leaf buckets, leaf draws, column gather) timed with HIP events,
    and the oracle's case-wise V1 (the reference's OpenMP branch) on the host
    cores for a bounded tree sample."""
    import numpy as np
    import torch

    nin = torch.from_numpy(np.ascontiguousarray(forest.nin.reshape(T, n), np.int32)).to(keys.device)
    out = torch.empty(n * n, dtype=torch.float64, device=keys.device)
    acc.casewise(keys, nin, out)
    torch.cuda.synchronize()
    reps = 5
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        acc.casewise(keys, nin, out)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / reps
    res = {"call_ms": round(ms, 4), "output": "FP64 n x n column-major (row = out-of-bag case)",
           "tree_pair_updates_per_s": T * float(n) * n / (ms / 1e3), "written_bytes": 8 * n * n}
    del out
    torch.cuda.empty_cache()
    if cpu:
        from oracle import oracle as O

        passes, dt = _repeat_for(lambda: O.prox_v1(forest, O.leaf_assign(forest, x), casewise=True), 5.0)
        res["cpu_baseline"] = {"value": passes * T * float(n) * n / dt, "unit": "tree·pair updates/s",
                               "cores": int(O.num_threads()), "kind": "port",
                               "sample": f"{passes} passes over all {T} trees: oracle leaf assignment + the "
                                         f"case-wise rf::cpu_proximity branch (src/rf_proximity.cpp:186-223); "
                                         f"{dt:.2f} s"}
    return res


def rfgap_section(acc, df, forest, x, n: int, T: int, cpu: bool) -> dict:
    """The RF-GAP FP64 proximity (full n x n, symmetric) of the same forest:
    one call (keys from the reached nodes, leaf buckets, leaf draws, column
    gather, the mirrored-tile scaling and average) timed with HIP events, and
    the oracle's restatement of the reference's cpu_proximity_rfgap on the
    host cores for a bounded tree sample."""
    import numpy as np
    import torch

    _, nodextr = df.leaf_assign(want_nodextr=True)
    nin = torch.from_numpy(np.ascontiguousarray(forest.nin.reshape(T, n), np.int32)).to(df.device)
    out = torch.empty(n * n, dtype=torch.float64, device=df.device)
    acc.rfgap(nin, nodextr, df.node_off, out)
    torch.cuda.synchronize()
    reps = 5
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        acc.rfgap(nin, nodextr, df.node_off, out)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / reps
    res = {"call_ms": round(ms, 4), "output": "FP64 n x n, symmetric",
           "tree_pair_updates_per_s": T * float(n) * n / (ms / 1e3), "written_bytes": 2 * 8 * n * n}
    del out
    torch.cuda.empty_cache()
    if cpu:
        from oracle import oracle as O

        passes, dt = _repeat_for(lambda: O.prox_rfgap(forest.nin.reshape(T, n), O.leaf_assign(forest, x)), 5.0)
        res["cpu_baseline"] = {"value": passes * T * float(n) * n / dt, "unit": "tree·pair updates/s",
                               "cores": int(O.num_threads()), "kind": "port",
                               "sample": f"{passes} passes over all {T} trees: oracle leaf assignment + "
                                         f"rf::cpu_proximity_rfgap (src/rf_proximity.cpp:238-391); {dt:.2f} s"}
    return res


def factor_mds_section(df, forest, x, y, counts_all, rank: int, cpu: bool) -> dict:
    """The 3-D MDS half of the metric at the workload's n:
      * the exact classical MDS of P = C/T from the ALL count triangle, with
        the operator in FP64 (the FP64 reference on the exact proximity) and
        in FP32 (the plain FP32 GPU form);
      * rank-r factors of P, the FP64 MDS of those factors, and the factor
        MDS from each storage format (FP32, BF16, INT8, FP8-e4m3 with
        per-column scales): the FP32 VALU operator, and the matrix-core
        operator on the stored factors followed by FP32 refinement sweeps;
        errors against the FP64 MDS of the same factors and against the exact
        dense FP64 MDS;
      * the same pipelines rerun on a seeded 20,000-sample subset of the
        samples (its own forest) against its exact dense FP64 MDS;
      * the CPU port of rf::compute_mds_3d_cpu (renormalisation + full dsyevd)
        on the n = 2,000 configuration beside the GPU dense MDS."""
    import numpy as np
    import torch

    from rfx_amd import _lib
    from rfx_amd.forest import Forest, make_dataset
    from rfx_amd.proximity import (Accumulator, DeviceForest, lowrank_factors, mds_dense, mds_factors,
                                   mds_factors_mixed, quantize_factors)

    def errs(X, Xref):
        X = X.cpu().numpy() if hasattr(X, "cpu") else X
        Xref = Xref.cpu().numpy() if hasattr(Xref, "cpu") else Xref
        s = np.sign(np.sum(X * Xref, axis=0))
        s[s == 0] = 1
        rel = float(np.linalg.norm(X * s - Xref) / np.linalg.norm(Xref))
        qa, _ = np.linalg.qr(X)
        qb, _ = np.linalg.qr(Xref)
        ang = float(np.arccos(np.clip(np.linalg.svd(qa.T @ qb, compute_uv=False).min(), -1, 1)))
        return {"rel_err": rel, "max_angle_rad": ang}

    def timed(fn):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        r = fn()
        torch.cuda.synchronize()
        return r, time.perf_counter() - t0

    n, T = forest.n, forest.ntree
    out = {"rank": rank}
    # exact dense MDS of C/T at full n (reads the 80 GB count triangle per block product)
    (Xd, ed, itd), td = timed(lambda: mds_dense(counts_all, _lib.LAYOUT_UPPER, n, _lib.MDS_COUNTS, 1.0 / T,
                                               precision=_lib.PREC_FP64))
    Xexact = Xd.cpu().numpy()
    out["dense_fp64"] = {"mds_s": round(td, 4), "iters": itd, "eigvals": [float(v) for v in ed[:5]],
                         "eigengap_l3_l4": float(ed[2] - ed[3])}
    log(f"dense fp64 mds: {td:.3f} s, {itd} iterations")
    (X32, e32, it32), t32 = timed(lambda: mds_dense(counts_all, _lib.LAYOUT_UPPER, n, _lib.MDS_COUNTS, 1.0 / T,
                                                   precision=_lib.PREC_FP32))
    out["dense_fp32"] = {"mds_s": round(t32, 4), "iters": it32, "vs_dense_fp64": errs(X32, Xexact)}
    log(f"dense fp32 mds: {t32:.3f} s")
    del Xd, X32

    keys, _ = df.leaf_assign()
    (U, eig), t_fac = timed(lambda: lowrank_factors(keys, forest, rank=rank, oversample=16, power_iters=2))
    del keys
    torch.cuda.empty_cache()
    out["factor_build_s"] = round(t_fac, 4)
    out["factor_top_eigs"] = [round(float(v), 4) for v in eig[:6]]
    log(f"factors: {t_fac:.3f} s")
    U32, s32 = quantize_factors(U, _lib.PREC_FP32)
    (X64, e64, it64), t64 = timed(lambda: mds_factors(U32, s32, _lib.PREC_FP32, precision=_lib.PREC_FP64))
    Xref = X64.cpu().numpy()
    out["factor_fp64"] = {"mds_s": round(t64, 4), "iters": it64, "eigvals": [float(v) for v in e64[:5]],
                          "eigengap_l3_l4": float(e64[2] - e64[3]), "vs_dense_fp64": errs(Xref, Xexact)}
    fp32_form_s = None
    for name, fmt in (("fp32", _lib.PREC_FP32), ("bf16", _lib.PREC_BF16), ("int8", _lib.PREC_INT8),
                      ("fp8", _lib.PREC_FP8)):
        Uq, sc = quantize_factors(U, fmt)
        mds_factors(Uq, sc, fmt, precision=_lib.PREC_FP32)  # warm-up
        (X, ev, it), t = timed(lambda: mds_factors(Uq, sc, fmt, precision=_lib.PREC_FP32))
        if name == "fp32":
            fp32_form_s = t
        rec = {"bytes_per_factor_row": rank * {"fp32": 4, "bf16": 2, "int8": 1, "fp8": 1}[name],
               "valu_fp32": {"mds_s": round(t, 4), "iters": it, "vs_factor_fp64": errs(X, Xref),
                             "vs_dense_fp64": errs(X, Xexact)}}
        log(f"factor mds {name} valu: {t:.4f} s")
        mds_factors_mixed(Uq, sc, fmt, U32)  # warm-up
        (X, ev, itm), t = timed(lambda: mds_factors_mixed(Uq, sc, fmt, U32))
        rec["mfma_refined"] = {"mds_s": round(t, 4), "mfma_iters": itm[0], "fp32_refine_sweeps": itm[1],
                               "speedup_vs_fp32_valu_form": round(fp32_form_s / t, 3),
                               "speedup_vs_dense_fp32": round(t32 / t, 3),
                               "vs_factor_fp64": errs(X, Xref), "vs_dense_fp64": errs(X, Xexact)}
        log(f"factor mds {name} mfma+refine: {t:.4f} s iters {itm}")
        out[name] = rec
    del U, U32
    torch.cuda.empty_cache()

    # the seeded 20k-sample subset: both pipelines rerun on it
    m = 20000
    idx = np.sort(np.random.default_rng(7).choice(n, m, replace=False))
    xs, ys = np.ascontiguousarray(x[idx]), np.ascontiguousarray(y[idx])
    fs = Forest.grow(xs, ys, T, mtry=7, iseed=0)
    dfs = DeviceForest(fs, xs)
    ks, _ = dfs.leaf_assign()
    acc = Accumulator(fs.n, fs.ntree, fs.total_nodes)
    Cs = acc.accumulate(ks, _lib.PROX_ALL, _lib.LAYOUT_UPPER, _lib.ALGO_AUTO)
    Xsd, esd, _ = mds_dense(Cs, _lib.LAYOUT_UPPER, m, _lib.MDS_COUNTS, 1.0 / T, precision=_lib.PREC_FP64)
    Xsd = Xsd.cpu().numpy()
    del Cs
    Us, _ = lowrank_factors(ks, fs, rank=rank, oversample=16, power_iters=2)
    Us32, ss32 = quantize_factors(Us, _lib.PREC_FP32)
    sub = {"n": m, "seed": 7, "dense_fp64_eigs": [float(v) for v in esd[:5]],
           "eigengap_l3_l4": float(esd[2] - esd[3])}
    X, _, _ = mds_factors(Us32, ss32, _lib.PREC_FP32, precision=_lib.PREC_FP64)
    sub["factor_fp64"] = errs(X, Xsd)
    for name, fmt in (("fp32", _lib.PREC_FP32), ("bf16", _lib.PREC_BF16), ("int8", _lib.PREC_INT8),
                      ("fp8", _lib.PREC_FP8)):
        Uq, sc = quantize_factors(Us, fmt)
        X, _, _ = mds_factors_mixed(Uq, sc, fmt, Us32)
        sub[name] = errs(X, Xsd)
    out["subset_20k_vs_dense_fp64"] = sub
    del ks, Us, Us32, acc
    torch.cuda.empty_cache()
    if cpu:
        out["cpu_mds_baseline"] = cpu_mds_baseline()
    return out


def cpu_mds_baseline() -> dict:
    """rf::compute_mds_3d_cpu's work on the n = 2,000 configuration (the
    reference chain: OOB counts -> finishprox -> float32 -> nout
    renormalisation -> D, D^2, double centring -> full dsyevd -> top three),
    ported to numpy with LAPACK's dsyevd (scipy's OpenBLAS) on the host cores,
    beside the GPU dense MDS of the same counts (reference chain, FP64 and FP32
    operators)."""
    import numpy as np
    import torch
    from scipy.linalg import eigh

    from oracle import oracle as O
    from rfx_amd import _lib
    from rfx_amd.forest import Forest, make_dataset
    from rfx_amd.proximity import Accumulator, DeviceForest, mds_dense

    kind, n, p, ncl, T, mtry, _ = WORKLOADS["c1"]
    x, y = make_dataset(kind, n, p, ncl, 0)
    f = Forest.grow(x, y, T, mtry=mtry, iseed=0)
    nout = f.nout(2)
    P = O.counts(f, O.leaf_assign(f, x), O.OOB_ASYM, 0).astype(np.float64)
    S = O.finishprox(n, nout, P).reshape(n, n, order="F")

    def cpu_mds():
        M = O.mds_input_reference(S, nout)
        D = M.max() - M
        np.fill_diagonal(D, 0.0)
        D2 = D * D
        r = D2.mean(axis=1)
        B = -0.5 * (D2 - r[:, None] - r[None, :] + r.mean())
        w, V = eigh(B, driver="evd")
        return V[:, ::-1][:, :3] * np.sqrt(w[::-1][:3])

    t0 = time.perf_counter()
    Xc = cpu_mds()
    tc = time.perf_counter() - t0
    df = DeviceForest(f, x)
    keys, _ = df.leaf_assign()
    acc = Accumulator(n, T, f.total_nodes)
    U = acc.accumulate(keys, _lib.PROX_OOB_UPPER, _lib.LAYOUT_UPPER, _lib.ALGO_AUTO)
    nd = torch.from_numpy(nout).to(keys.device)
    res = {"n": n, "ntree": T, "cpu_s": round(tc, 4), "cores": int(O.num_threads()), "kind": "port",
           "lapack": "dsyevd (scipy's bundled OpenBLAS)", "host_cpu": host_cpu()}
    for name, prec in (("gpu_dense_fp64", _lib.PREC_FP64), ("gpu_dense_fp32", _lib.PREC_FP32)):
        mds_dense(U, _lib.LAYOUT_UPPER, n, _lib.MDS_REFERENCE, 1.0, nd, precision=prec)  # warm-up
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        X, _, it = mds_dense(U, _lib.LAYOUT_UPPER, n, _lib.MDS_REFERENCE, 1.0, nd, precision=prec)
        torch.cuda.synchronize()
        tg = time.perf_counter() - t0
        Xg = X.cpu().numpy()
        s = np.sign(np.sum(Xg * Xc, axis=0))
        res[name] = {"mds_s": round(tg, 4), "iters": it, "speedup_vs_cpu": round(tc / tg, 2),
                     "rel_err_vs_cpu": float(np.linalg.norm(Xg * s - Xc) / np.linalg.norm(Xc))}
    return res


def c5_instances(count: int = 64):
    """Instance i of the 64-forest workload, from seed i."""
    import numpy as np

    out = []
    for i in range(count):
        r = np.random.default_rng(i)
        n = int(round(np.exp(r.uniform(np.log(1000), np.log(50000)))))
        p = int(r.integers(8, 65))
        T = int(r.integers(50, 501))
        ncl = int(r.integers(2, 11))
        out.append((i, n, p, T, ncl))
    return out


def bench_c4(args) -> int:
    """The FP8 scale-up workload: n = 1,000,000 samples, T = 500 trees
    (minndsize 5).  The proximity is never materialised (its upper triangle
    would take 2 TB): rank-128 factors of C/T come from the leaf buckets by
    matrix-free randomized subspace iteration, and the 3-D MDS runs on the
    factors stored as FP8-e4m3 and BF16 (per-column scales) on the matrix
    cores with FP32 refinement sweeps, against the FP32 VALU form and the FP64
    MDS of the same factors.  value = seconds of the FP8 factor MDS with one
    FP32 refinement sweep (lower is better); the leaf assignment and factor
    build are reported beside it.  A seeded 20,000-sample subset (its own
    forest) reruns the pipelines against its exact dense FP64 MDS."""
    import numpy as np
    import torch

    from rfx_amd import _lib
    from rfx_amd.forest import Forest, make_dataset
    from rfx_amd.proximity import (Accumulator, DeviceForest, lowrank_factors, mds_dense, mds_factors,
                                   mds_factors_mixed, quantize_factors)

    kind, n, p, ncl, T, mtry, desc = WORKLOADS["c4"]
    T = args.trees or T
    rank = args.rank or 128
    minnd = 5
    t0 = time.perf_counter()
    x, y = make_dataset(kind, n, p, ncl, 0)
    forest = Forest.grow(x, y, T, mtry=mtry, minndsize=minnd, iseed=0,
                         progress=lambda d, tot: log(f"c4 forest: {d}/{tot} trees, {time.perf_counter() - t0:.0f} s"))
    grow_s = time.perf_counter() - t0
    log(f"c4 forest: {forest.total_nodes} nodes in {grow_s:.1f} s")

    def timed(fn):
        torch.cuda.synchronize()
        s = time.perf_counter()
        r = fn()
        torch.cuda.synchronize()
        return r, time.perf_counter() - s

    def errs(X, Xref):
        X = X.cpu().numpy() if hasattr(X, "cpu") else X
        Xref = Xref.cpu().numpy() if hasattr(Xref, "cpu") else Xref
        sg = np.sign(np.sum(X * Xref, axis=0))
        sg[sg == 0] = 1
        rel = float(np.linalg.norm(X * sg - Xref) / np.linalg.norm(Xref))
        qa, _ = np.linalg.qr(X)
        qb, _ = np.linalg.qr(Xref)
        ang = float(np.arccos(np.clip(np.linalg.svd(qa.T @ qb, compute_uv=False).min(), -1, 1)))
        return {"rel_err": rel, "max_angle_rad": ang}

    dev = torch.device("cuda:0")
    df = DeviceForest(forest, x)
    keys = torch.empty((T, n), dtype=torch.int32, device=dev)
    df.leaf_assign(keys=keys)
    stream = torch.cuda.current_stream()
    evs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(5)]
    for a, b in evs:
        a.record(stream)
        df.leaf_assign(keys=keys)
        b.record(stream)
    torch.cuda.synchronize()
    leaf_ms = float(np.mean([a.elapsed_time(b) for a, b in evs]))
    # the one-pass kernel (no chain jumps) once, for comparison: same keys
    keys1 = torch.empty_like(keys)
    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    a.record(stream)
    df.leaf_assign(keys=keys1, queue=False)
    b.record(stream)
    torch.cuda.synchronize()
    onepass_ms = a.elapsed_time(b)
    same_keys = bool(torch.equal(keys1, keys))
    del keys1
    if not same_keys:
        raise SystemExit("c4: the chain-jump leaf assignment differs from the one-pass kernel")
    log(f"c4 leaf assignment one-pass {onepass_ms:.1f} ms (chain index: {df.chains is not None}), keys identical")
    lowrank_factors(keys, forest, rank=rank, oversample=16, power_iters=2)  # warm-up
    (U, eig), fac_s = timed(lambda: lowrank_factors(keys, forest, rank=rank, oversample=16, power_iters=2))
    del keys
    torch.cuda.empty_cache()
    log(f"c4 leaf assignment {leaf_ms:.2f} ms, rank-{rank} factors {fac_s:.3f} s")
    U32, s32 = quantize_factors(U, _lib.PREC_FP32)
    (X64, e64, it64), t64 = timed(lambda: mds_factors(U32, s32, _lib.PREC_FP32, precision=_lib.PREC_FP64))
    Xref = X64.cpu().numpy()
    mds_factors(U32, s32, _lib.PREC_FP32, precision=_lib.PREC_FP32)  # warm-up
    (X32, _, it32), t32 = timed(lambda: mds_factors(U32, s32, _lib.PREC_FP32, precision=_lib.PREC_FP32))
    out = {"factor_fp64": {"mds_s": round(t64, 4), "iters": it64, "eigvals": [float(v) for v in e64[:5]],
                           "eigengap_l3_l4": float(e64[2] - e64[3])},
           "fp32_valu": {"mds_s": round(t32, 4), "iters": it32, "vs_factor_fp64": errs(X32, Xref)}}
    log(f"c4 factor mds fp64 {t64:.3f} s, fp32 valu {t32:.4f} s")
    value = None
    for name, fmt in (("fp8", _lib.PREC_FP8), ("bf16", _lib.PREC_BF16), ("fp32", _lib.PREC_FP32)):
        Uq, sc = quantize_factors(U, fmt)
        rec = {"bytes_per_factor_row": rank * {"fp32": 4, "bf16": 2, "fp8": 1}[name]}
        for label, rmax in (("one_sweep", 1), ("refined", 60)):
            mds_factors_mixed(Uq, sc, fmt, U32, refine_max=rmax)  # warm-up
            (X, _, itm), t = timed(lambda: mds_factors_mixed(Uq, sc, fmt, U32, refine_max=rmax))
            rec[label] = {"mds_s": round(t, 4), "mfma_iters": itm[0], "fp32_refine_sweeps": itm[1],
                          "speedup_vs_fp32_valu": round(t32 / t, 3), "vs_factor_fp64": errs(X, Xref)}
            log(f"c4 factor mds {name} {label}: {t:.4f} s iters {itm}")
            if name == "fp8" and label == "one_sweep":
                value = t
        out[name] = rec
    del U, U32
    torch.cuda.empty_cache()

    # seeded 20k subset: exact dense FP64 MDS of its own forest's C/T against the factor pipelines
    m = 20000
    idx = np.sort(np.random.default_rng(7).choice(n, m, replace=False))
    xs, ys = np.ascontiguousarray(x[idx]), np.ascontiguousarray(y[idx])
    fs = Forest.grow(xs, ys, T, mtry=mtry, minndsize=minnd, iseed=0)
    dfs = DeviceForest(fs, xs)
    ks, _ = dfs.leaf_assign()
    acc = Accumulator(fs.n, fs.ntree, fs.total_nodes)
    Cs = acc.accumulate(ks, _lib.PROX_ALL, _lib.LAYOUT_UPPER, _lib.ALGO_AUTO)
    Xsd, esd, _ = mds_dense(Cs, _lib.LAYOUT_UPPER, m, _lib.MDS_COUNTS, 1.0 / T, precision=_lib.PREC_FP64)
    Xsd = Xsd.cpu().numpy()
    del Cs
    Us, _ = lowrank_factors(ks, fs, rank=rank, oversample=16, power_iters=2)
    Us32, ss32 = quantize_factors(Us, _lib.PREC_FP32)
    sub = {"n": m, "seed": 7, "dense_fp64_eigs": [float(v) for v in esd[:5]],
           "eigengap_l3_l4": float(esd[2] - esd[3])}
    X, _, _ = mds_factors(Us32, ss32, _lib.PREC_FP32, precision=_lib.PREC_FP64)
    sub["factor_fp64"] = errs(X, Xsd)
    for name, fmt in (("fp8", _lib.PREC_FP8), ("bf16", _lib.PREC_BF16), ("fp32", _lib.PREC_FP32)):
        Uq, sc = quantize_factors(Us, fmt)
        X, _, _ = mds_factors_mixed(Uq, sc, fmt, Us32, refine_max=1)
        sub[name + "_one_sweep"] = errs(X, Xsd)
    out["subset_20k_vs_dense_fp64"] = sub

    line = {
        "metric": "3D-MDS s at n=1M from rank-128 FP8-e4m3 factors (one FP32 refinement sweep)",
        "value": value, "unit": "s", "n_gpus": 1, "steps": 1, "warmup": 1, "ms_per_step": value * 1e3,
        "higher_is_better": False, "scaling": "weak", "vs_baseline": None, "dtype": "fp8-e4m3", "data": "synthetic",
        "config": {"workload": desc, "n": n, "p": p, "ntree": T, "mtry": mtry, "minndsize": minnd, "rank": rank,
                   "forest_nodes": forest.total_nodes, "grow_s": round(grow_s, 1), "leaf_assign_ms": round(leaf_ms, 3),
                   "leaf_assign_onepass_ms": round(onepass_ms, 3), "leaf_chain_index": df.chains is not None,
                   "factor_build_s": round(fac_s, 4), "factor_top_eigs": [round(float(v), 4) for v in eig[:6]]},
        "mds": out,
    }
    print(json.dumps(line), flush=True)
    return 0


def bench_c5(args) -> int:
    import numpy as np
    import torch

    from rfx_amd import _lib
    from rfx_amd.batch import ForestBatch
    from rfx_amd.forest import Forest, default_mtry, make_dataset
    from rfx_amd.proximity import Accumulator, DeviceForest

    inst = c5_instances()
    t0 = time.perf_counter()
    forests, xs, dfs = [], [], []
    for i, n, p, T, ncl in inst:
        x, y = make_dataset(_lib.DATA_GAUSS_MIXTURE, n, p, ncl, i)
        f = Forest.grow(x, y, T, mtry=default_mtry(p), iseed=i)
        forests.append(f)
        xs.append(x)
        dfs.append(DeviceForest(f, x))
        if i % 16 == 15:
            log(f"{i + 1} forests grown, {time.perf_counter() - t0:.1f} s")
    log(f"64 forests grown in {time.perf_counter() - t0:.1f} s")
    work = sum(f.ntree * (f.n * (f.n + 1) // 2) for f in forests)

    batch = ForestBatch(dfs, "oob_upper")
    best_ms = batch.autotune(reps=1)
    lpi, bpc = batch.tuning()
    algos, sumsq = batch.stats()
    n_dense = int(np.sum(algos == _lib.ALGO_DENSE))
    log(f"autotune: {lpi} rows per item, {bpc} workgroups per CU, {n_dense} dense / {len(algos) - n_dense} gather "
        f"forests, {best_ms:.2f} ms")

    # 64 sequential launches of the one-problem path with the same per-forest algorithm, same outputs
    seq = []
    for k, (df, f) in enumerate(zip(dfs, forests)):
        keys = torch.empty((f.ntree, f.n), dtype=torch.int32, device=batch.device)
        seq.append((df, Accumulator(f.n, f.ntree, f.total_nodes), keys, batch.out[k], int(algos[k])))

    def run_seq():
        for df, acc, keys, out, algo in seq:
            df.leaf_assign(keys=keys)
            acc.accumulate(keys, _lib.PROX_OOB_UPPER, _lib.LAYOUT_UPPER, algo, out)

    run_seq()
    torch.cuda.synchronize()
    ref = [o.clone() for o in batch.out[:4]]
    s = time.perf_counter()
    run_seq()
    torch.cuda.synchronize()
    seq_s = time.perf_counter() - s
    del seq
    torch.cuda.empty_cache()

    for _ in range(args.warmup):
        batch.run(use_graph=True)
    torch.cuda.synchronize()
    for k in range(4):
        assert torch.equal(ref[k], batch.out[k]), "batched and sequential counts differ"
    del ref
    stream = torch.cuda.current_stream()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t_start = time.perf_counter()
    e0.record(stream)
    for _ in range(args.steps):
        batch.run(use_graph=True)
    e1.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t_start
    ms_step = wall * 1e3 / args.steps
    ev_ms = e0.elapsed_time(e1) / args.steps
    value = work / (ms_step / 1e3)

    # the same step with every forest through the persistent queue (the
    # autotuner's own launches switched off), same outputs
    own = batch.own_launch()
    keep = [o.clone() for o in batch.out[:4]]
    batch.set_own_launch(np.zeros(len(forests), np.int32))
    batch.run(use_graph=True)
    torch.cuda.synchronize()
    for k in range(4):
        assert torch.equal(keep[k], batch.out[k]), "queue-only and autotuned counts differ"
    del keep
    s = time.perf_counter()
    for _ in range(args.steps):
        batch.run(use_graph=True)
    torch.cuda.synchronize()
    queue_s = (time.perf_counter() - s) / args.steps
    batch.set_own_launch(own.astype(np.int32))
    log(f"batched step {ms_step:.2f} ms ({int(own.sum())} own launches), queue only {queue_s * 1e3:.2f} ms")
    # the DENSE forests' packed 16-bit compare: v_xor_b32, v_pk_min_u16 and
    # v_pk_add_u16 per word of two trees and output pair, 1.5 VALU operations per
    # tree-pair, against 256 CUs x 64 lanes x 2.4 GHz
    dense_pairs = sum(f.ntree * (f.n * (f.n + 1) // 2) for f, a in zip(forests, algos) if a == _lib.ALGO_DENSE)
    valu_peak = 256 * 64 * 2.4e9
    valu = {"bound": "valu", "ops_per_tree_pair": 1.5, "dense_tree_pairs": dense_pairs,
            "achieved_ops_per_s": 1.5 * dense_pairs / (ev_ms / 1e3), "peak_ops_per_s": valu_peak,
            "frac": round(1.5 * dense_pairs / (ev_ms / 1e3) / valu_peak, 4),
            "note": "over the whole batched step (the GATHER forests run beside the DENSE ones), so a lower bound "
                    "of the DENSE kernels' own fraction"}
    P_total = sum(f.n * (f.n + 1) // 2 for f in forests)
    nT_total = sum(f.n * f.ntree for f in forests)
    step_bytes = 4 * P_total + 4 * nT_total
    achieved = step_bytes / (ev_ms / 1e3) / 1e9
    line = {
        "metric": METRIC, "value": value, "unit": "tree·pair updates/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms_step, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "int32", "data": "synthetic",
        "config": {"workload": WORKLOADS["c5"][6], "instances": len(forests), "tree_pair_updates": work,
                   "sum_n": int(sum(f.n for f in forests)), "layout": "upper", "graph": True,
                   "rows_per_item": lpi, "workgroups_per_cu": bpc, "dense_forests": n_dense,
                   "gather_forests": len(algos) - n_dense,
                   "pair_density": float(np.sum(sumsq) / sum(f.ntree * float(f.n) ** 2 for f in forests)),
                   "sequential_s": round(seq_s, 5), "batched_s": round(ms_step / 1e3, 5),
                   "speedup_vs_sequential": round(seq_s / (ms_step / 1e3), 3),
                   "sequential_value": work / seq_s, "own_launch_forests": int(own.sum()),
                   "queue_only_s": round(queue_s, 5)},
        "valu": valu,
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                     "kernel": "whole batched step (leaf + segmented sort + persistent dense/gather queue), "
                               "HIP events",
                     "algorithmic_bytes": step_bytes, "avg_ms": round(ev_ms, 4)},
    }
    if not args.no_cpu:
        line["cpu_baseline"] = c5_cpu_baseline(forests, xs)
    if args.mds:
        line["mds"] = c5_mds(batch, forests)
    print(json.dumps(line), flush=True)
    return 0


def c5_cpu_baseline(forests, xs, budget_s: float = 15.0) -> dict:
    """The oracle port of the reference CPU path on whole instances, smallest
    first, until the time budget is spent."""
    from oracle import oracle as O

    order = sorted(range(len(forests)), key=lambda k: forests[k].n)
    done, work, dt = [], 0, 0.0
    for k in order:
        f = forests[k]
        if dt > budget_s:
            break
        t0 = time.perf_counter()
        nd = O.leaf_assign(f, xs[k])
        prox = O.prox_v1(f, nd)
        dt += time.perf_counter() - t0
        del prox, nd
        work += f.ntree * (f.n * (f.n + 1) // 2)
        done.append(k)
    return {"value": work / dt, "unit": "tree·pair updates/s", "cores": int(O.num_threads()), "kind": "port",
            "sample": f"{len(done)} smallest instances whole (n {forests[done[0]].n}..{forests[done[-1]].n}): oracle "
                      f"leaf assignment + rf::cpu_proximity dense OpenMP branch (src/rf_proximity.cpp:99-231), "
                      f"one instance after another; {dt:.2f} s"}


def c5_mds(batch, forests) -> dict:
    """The reference chain's 3-D MDS (OOB counts -> finishprox -> float32 ->
    renormalisation -> classical MDS) of every instance, FP32 operator: the
    64 instances one after another, and the same 64 run concurrently on
    eight streams (mds_dense_many), whose coordinates are compared with the
    sequential ones."""
    import numpy as np
    import torch

    from rfx_amd import _lib
    from rfx_amd.proximity import mds_dense, mds_dense_many

    jobs = [dict(counts=out, layout=_lib.LAYOUT_UPPER, n=f.n, mode=_lib.MDS_REFERENCE, scale=1.0,
                 nout=torch.from_numpy(f.nout(2)).to(out.device), precision=_lib.PREC_FP32)
            for f, out in zip(forests, batch.out)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    seq = [mds_dense(**j) for j in jobs]
    torch.cuda.synchronize()
    t_seq = time.perf_counter() - t0
    mds_dense_many(jobs[:8], streams=8)  # warm-up of the streams' allocations
    t0 = time.perf_counter()
    conc = mds_dense_many(jobs, streams=8)
    t_conc = time.perf_counter() - t0
    diff = max(float(torch.linalg.norm(a[0] - b[0]) / torch.linalg.norm(a[0])) for a, b in zip(seq, conc))
    iters = [r[2] for r in seq]
    return {"sequential_s": round(t_seq, 4), "concurrent_8_streams_s": round(t_conc, 4),
            "speedup": round(t_seq / t_conc, 3), "max_rel_diff_concurrent_vs_sequential": diff,
            "mean_iters": float(np.mean(iters)), "mode": "reference chain, FP32 operator, FP64 Rayleigh-Ritz"}

if __name__ == "__main__":
    sys.exit(main())
