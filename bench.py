#!/usr/bin/env python3
"""Benchmark of the MI355X proximity path: one JSON line on stdout.

A step is one pass of the hot path over the resident workload: leaf
assignment of every sample through every tree, then the ALL and OOB upper
int32 count matrices (both exact).  The default workload is the covtype-like
problem: n = 20,000 samples, p = 54 features, 7 classes, T = 500 trees,
mtry = 7, forest grown on the host by the restated reference grower (growth
is setup, not timed).

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload c2|c1|c3|c5] [--algo auto|dense|bucket|gather]

value = T * n(n+1)/2 tree-pair updates per second (each tree-pair slot counts
once although it feeds both matrices).  The line carries the roofline of the
dominant kernel (HIP events on the launch stream) and a CPU baseline: the
oracle's restatement of the reference's OpenMP path (leaf assignment +
rf::cpu_proximity's accumulation) timed on a bounded sample of trees.
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
        "seconds": dt,
    }


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="c2", choices=sorted(WORKLOADS))
    ap.add_argument("--rank", type=int, default=64, help="factor rank (c3)")
    ap.add_argument("--algo", default="auto", choices=["auto", "dense", "bucket", "gather"])
    ap.add_argument("--cpu-trees", type=int, default=0, help="trees in the CPU baseline sample (0: auto)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--mds", action="store_true", help="c5: also time the 3-D MDS of every instance")
    ap.add_argument("--no-mds", action="store_true", help="c3: counts only, no factor / MDS section")
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
              "(multi-GPU sharding is described in DESIGN.md, not run)", file=sys.stderr)
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
    NAMES = {_lib.ALGO_DENSE: "dense", _lib.ALGO_BUCKET: "bucket", _lib.ALGO_GATHER: "gather"}
    kind, n, p, ncl, T, mtry, desc = WORKLOADS[args.workload]
    t0 = time.perf_counter()
    x, y = make_dataset(kind, n, p, ncl, 0)
    forest = Forest.grow(x, y, T, mtry=mtry, iseed=0)
    log(f"forest grown: {forest.total_nodes} nodes in {time.perf_counter() - t0:.1f} s")

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
        small = n <= 65536
        algos = (_lib.ALGO_DENSE, _lib.ALGO_BUCKET, _lib.ALGO_GATHER) if small else (_lib.ALGO_GATHER, _lib.ALGO_BUCKET)
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
    # n = 20476 and T = 4096, line chunks up to n = 32768, else the gallop walk
    gather_kernel = ("gather_rows_kernel<PAIR,UPPER>" if n <= 20476 and T <= 4096 else
                     "gather_line_kernel<PAIR,UPPER>" if n <= 32768 else "gather_gallop_kernel<PAIR,UPPER>")
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
                   "layout": "upper", "leaf_assign_ms": round(leaf_ms, 4), "accumulate_ms": round(acc_ms, 4),
                   "bucket_build_call_ms": round(build_ms, 4), "count_kernel_ms": round(kern_ms, 4),
                   "sum_leaf_sq": sq, "pair_density": sq / (T * float(n) * n)},
        "roofline": roofline,
    }
    if not args.no_cpu and n > 50000:
        # the reference CPU path holds an n x n float64 matrix (320 GB at n = 200,000)
        line["cpu_baseline"] = None
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
        out_all = out_oob = None
        torch.cuda.empty_cache()
        line["mds"] = factor_mds_section(df, forest, args.rank)
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


def factor_mds_section(df, forest, rank: int) -> dict:
    """Rank-r factors of P = C/T and the 3-D MDS from them in each storage
    format, timed, with the error of each against the FP64 operator on the
    same factors; plus the truncation check on a seeded 20k-sample problem
    (same generator) against the exact dense FP64 MDS."""
    import numpy as np
    import torch

    from rfx_amd import _lib
    from rfx_amd.forest import Forest, make_dataset
    from rfx_amd.proximity import (Accumulator, DeviceForest, lowrank_factors, mds_dense, mds_factors,
                                   mds_factors_mixed, quantize_factors)

    def errs(X, Xref):
        X = X.cpu().numpy()
        s = np.sign(np.sum(X * Xref, axis=0))
        s[s == 0] = 1
        rel = float(np.linalg.norm(X * s - Xref) / np.linalg.norm(Xref))
        qa, _ = np.linalg.qr(X)
        qb, _ = np.linalg.qr(Xref)
        ang = float(np.arccos(np.clip(np.linalg.svd(qa.T @ qb, compute_uv=False).min(), -1, 1)))
        return rel, ang

    def timed(fn):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        r = fn()
        torch.cuda.synchronize()
        return r, time.perf_counter() - t0

    out = {"rank": rank}
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
    out["fp64_s"] = round(t64, 4)
    out["eigvals_fp64"] = [float(v) for v in e64[:5]]
    out["eigengap_l3_l4"] = float(e64[2] - e64[3])
    for name, fmt in (("fp32", _lib.PREC_FP32), ("bf16", _lib.PREC_BF16), ("int8", _lib.PREC_INT8),
                      ("fp8", _lib.PREC_FP8)):
        Uq, sc = quantize_factors(U, fmt)
        mds_factors(Uq, sc, fmt, precision=_lib.PREC_FP32)  # warm-up
        (X, ev, it), t = timed(lambda: mds_factors(Uq, sc, fmt, precision=_lib.PREC_FP32))
        rel, ang = errs(X, Xref)
        out[name] = {"mds_s": round(t, 4), "iters": it, "rel_err_vs_fp64": rel, "max_angle_rad": ang,
                     "bytes_per_factor_row": rank * {"fp32": 4, "bf16": 2, "int8": 1, "fp8": 1}[name]}
        log(f"factor mds {name}: {t:.4f} s rel {rel:.2e}")
        # bf16 matrix-core operator on the stored factors + FP32 refinement sweeps on the FP32 factors
        mds_factors_mixed(Uq, sc, fmt, U32)  # warm-up
        (X, ev, itm), t = timed(lambda: mds_factors_mixed(Uq, sc, fmt, U32))
        rel, ang = errs(X, Xref)
        out[name]["mfma_refined"] = {"mds_s": round(t, 4), "mfma_iters": itm[0], "fp32_refine_sweeps": itm[1],
                                     "rel_err_vs_fp64": rel, "max_angle_rad": ang}
        log(f"factor mds {name} mfma+refine: {t:.4f} s iters {itm} rel {rel:.2e}")
    # truncation: the exact dense FP64 MDS of a seeded 20k problem vs its factor MDS
    xs, ys = make_dataset(_lib.DATA_COVTYPE_LIKE, 20000, 54, 7, 7)
    fs = Forest.grow(xs, ys, forest.ntree, mtry=7, iseed=0)
    dfs = DeviceForest(fs, xs)
    ks, _ = dfs.leaf_assign()
    acc = Accumulator(fs.n, fs.ntree, fs.total_nodes)
    Cs = acc.accumulate(ks, _lib.PROX_ALL, _lib.LAYOUT_UPPER, _lib.ALGO_AUTO)
    Xd, ed, _ = mds_dense(Cs, _lib.LAYOUT_UPPER, fs.n, _lib.MDS_COUNTS, 1.0 / fs.ntree, precision=_lib.PREC_FP64)
    del Cs
    Us, _ = lowrank_factors(ks, fs, rank=rank, oversample=16, power_iters=2)
    Uq, sc = quantize_factors(Us, _lib.PREC_BF16)
    Xs, es, _ = mds_factors(Uq, sc, _lib.PREC_BF16)
    rel, ang = errs(Xs, Xd.cpu().numpy())
    out["subset_20k"] = {"dense_fp64_eigs": [float(v) for v in ed[:4]], "factor_bf16_eigs": [float(v) for v in es[:4]],
                         "rel_err_vs_dense_fp64": rel, "max_angle_rad": ang}
    return out


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
                   "sequential_value": work / seq_s},
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
    renormalisation -> classical MDS) of every instance, FP32 operator."""
    import torch

    from rfx_amd import _lib
    from rfx_amd.proximity import mds_dense

    torch.cuda.synchronize()
    t0 = time.perf_counter()
    iters = []
    for f, out in zip(forests, batch.out):
        nout = torch.from_numpy(f.nout(2)).to(out.device)
        _, _, it = mds_dense(out, _lib.LAYOUT_UPPER, f.n, _lib.MDS_REFERENCE, 1.0, nout, precision=_lib.PREC_FP32)
        iters.append(it)
    torch.cuda.synchronize()
    return {"total_s": round(time.perf_counter() - t0, 4), "mean_iters": float(sum(iters) / len(iters)),
            "mode": "reference chain, FP32 operator, FP64 Rayleigh-Ritz"}


if __name__ == "__main__":
    sys.exit(main())
