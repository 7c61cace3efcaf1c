#!/usr/bin/env python3
"""Benchmark of the MI355X proximity path: one JSON line on stdout.

A step is one pass of the hot path over the resident workload: leaf
assignment of every sample through every tree, then the ALL and OOB upper
int32 count matrices (both exact).  The default workload is the covtype-like
problem: n = 20,000 samples, p = 54 features, 7 classes, T = 500 trees,
mtry = 7, forest grown on the host by the restated reference grower (growth
is setup, not timed).

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload c2|c1] [--algo auto|dense|bucket]

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
}


def log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


def cpu_baseline(forest, x, sample_trees: int) -> dict:
    """The oracle port of the reference CPU path (per tree: leaf assignment,
    then rf::cpu_proximity's accumulating OpenMP branch) on sample_trees trees."""
    import numpy as np

    from oracle import oracle as O

    sub = forest.subset(slice(0, sample_trees))
    n = forest.n
    t0 = time.perf_counter()
    nd = O.leaf_assign(sub, x)
    prox = O.prox_v1(sub, nd)
    dt = time.perf_counter() - t0
    del prox, nd
    P = n * (n + 1) // 2
    return {
        "value": sample_trees * P / dt,
        "unit": "tree·pair updates/s",
        "cores": int(O.num_threads()),
        "kind": "port",
        "sample": f"first {sample_trees} of {forest.ntree} trees at n={n}: oracle leaf assignment "
                  f"(src/rf_testreebag.cpp, find_terminal_node) + rf::cpu_proximity dense OpenMP branch "
                  f"(src/rf_proximity.cpp:99-231) into an n x n float64 matrix; {dt:.2f} s",
        "seconds": dt,
    }


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="c2", choices=sorted(WORKLOADS))
    ap.add_argument("--algo", default="auto", choices=["auto", "dense", "bucket", "gather"])
    ap.add_argument("--cpu-trees", type=int, default=0, help="trees in the CPU baseline sample (0: auto)")
    ap.add_argument("--no-cpu", action="store_true")
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
        for a in (_lib.ALGO_DENSE, _lib.ALGO_BUCKET, _lib.ALGO_GATHER):
            acc.accumulate_pair(keys, _lib.LAYOUT_UPPER, a, out_all, out_oob)
            torch.cuda.synchronize()
            s = time.perf_counter()
            for _ in range(3):
                acc.accumulate_pair(keys, _lib.LAYOUT_UPPER, a, out_all, out_oob)
            torch.cuda.synchronize()
            cand[a] = (time.perf_counter() - s) / 3
        algo = min(cand, key=cand.get)
        log("algo timings (s): " + " ".join(f"{NAMES[a]} {v:.5f}" for a, v in cand.items()))
    algo_name = NAMES[algo]

    stream = torch.cuda.current_stream()
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True),
           torch.cuda.Event(enable_timing=True)) for _ in range(args.steps)]

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
    P = n * (n + 1) // 2
    value = T * P / (ms_step / 1e3)

    # Algorithmic bytes of the accumulation launch (one pass over all trees):
    # two int32 upper triangles written + the int32 key array read.
    acc_bytes = 2 * 4 * P + 4 * n * T
    achieved = acc_bytes / (acc_ms / 1e3) / 1e9
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                "kernel": f"rfx_prox_accumulate_pair ({algo_name})", "algorithmic_bytes": acc_bytes,
                "avg_ms": round(acc_ms, 4)}

    line = {
        "metric": METRIC, "value": value, "unit": "tree·pair updates/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms_step, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "int32", "data": "synthetic",
        "config": {"workload": desc, "n": n, "p": p, "ntree": T, "mtry": mtry, "algo": algo_name,
                   "layout": "upper", "leaf_assign_ms": round(leaf_ms, 4), "accumulate_ms": round(acc_ms, 4),
                   "sum_leaf_sq": sq, "pair_density": sq / (T * float(n) * n)},
        "roofline": roofline,
    }
    if not args.no_cpu:
        trees = args.cpu_trees or T
        cb = cpu_baseline(forest, x, trees)
        cb.pop("seconds")
        line["cpu_baseline"] = cb
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
