"""Write the golden fixtures of the proximity path from the reference itself.

Runs the reference's own CPU sources (compiled by ``make -C oracle ref`` into
oracle/_ref/librfxref.so from the reference tree) on small deterministic
inputs and stores inputs and outputs as compressed .npz data:

  gauss_n200.npz       3-class Gaussian mixture, n=200, p=20, T=12, mtry=4,
                       iseed=0, all features numeric
  covtype_cat_n240.npz covtype-like n=240, p=54, T=8, mtry=7, iseed=0, the
                       classifier's categorical auto-detection with maxcat=10
                       (the 44 binary columns become 2-level categoricals)

Each holds x, y, cat, the forest (nnode, nodestatus, bestvar, xbestsplit,
treemap per tree, padded to maxnode = 2n+1), nin, nodextr, nout (the CPU
path's doubled count), prox (rf::cpu_proximity accumulated over the trees,
float64 column-major) and proxsym (rf::cpu_finishprox).  From the
reference's RandomForest class itself (src/rf_random_forest.cpp, fit on the
CPU) they also hold rf_prox and rf_nout (get_proximity_matrix, get_nout_ptr)
and rf_prox_casewise (the same with use_casewise); from
rf::compute_mds_3d_cpu (src/rf_mds_cpu.cpp, dsyevd from this image's LAPACK)
mds_ref (the classifier's call: float32 proximity, nout, use_rfgap off),
rfgap (rf::cpu_proximity_rfgap) and mds_rfgap (use_rfgap on).

  unsup_n160.npz       unsupervised RandomForest, Gaussian mixture n=160,
                       p=12, T=10, mtry=3, iseed=0: x, prox, nout and the
                       class's compute_outlier_scores (outlier)

Usage: python tests/golden/make_golden.py   (needs /root/reference)
"""
from __future__ import annotations

import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))

from oracle import oracle as O  # noqa: E402
from rfx_amd.forest import make_dataset  # noqa: E402
from rfx_amd.classifier import detect_categorical  # noqa: E402

OUT = Path(__file__).resolve().parent


def one(name: str, kind: int, n: int, p: int, ncl: int, T: int, mtry: int, maxcat: int, seed: int) -> None:
    x, y = make_dataset(kind, n, p, ncl, seed)
    cat = detect_categorical(x, maxcat) if maxcat > 1 else np.ones(p, np.int32)
    r = O.ref_run_forest(x, y, T, mtry, minndsize=1, iseed=0, cat=cat, maxcat=maxcat, do_prox=True)
    proxsym = O.ref_finishprox(n, r["nout"], r["prox"])
    rf = O.ref_rf_fit(x, y, T, mtry, minndsize=1, iseed=0, cat=cat, maxcat=maxcat)
    rfc = O.ref_rf_fit(x, y, T, mtry, minndsize=1, iseed=0, cat=cat, maxcat=maxcat, casewise=True)
    # the classifier's compute_mds_3d_cpu: get_proximity_matrix's float32
    # matrix widened to float64, nout, use_rfgap off
    f32 = lambda a: a.astype(np.float32).astype(np.float64)  # noqa: E731
    mds_ref = O.ref_mds_3d(f32(rf["prox"]), rf["nout"], False)
    rfgap = O.ref_proximity_rfgap(r["nin"].reshape(T, n), r["nodextr"].reshape(T, n))
    mds_rfgap = O.ref_mds_3d(f32(rfgap), rf["nout"], True)
    np.savez_compressed(
        OUT / name, x=x, y=y, cat=cat, ntree=T, mtry=mtry, maxcat=maxcat, iseed=0,
        nnode=r["nnode"], nodestatus=r["nodestatus"], bestvar=r["bestvar"], xbestsplit=r["xbestsplit"],
        treemap=r["treemap"], nin=r["nin"], nodextr=r["nodextr"], nout=r["nout"], prox=r["prox"],
        proxsym=proxsym, rf_prox=rf["prox"], rf_nout=rf["nout"], rf_prox_casewise=rfc["prox"], mds_ref=mds_ref,
        rfgap=rfgap, mds_rfgap=mds_rfgap)
    print(name, "nodes", int(r["nnode"].sum()), "prox sum", float(r["prox"].sum()),
          "class == driver:", bool(np.array_equal(rf["prox"], proxsym)))


def unsupervised(name: str, n: int, p: int, T: int, mtry: int, seed: int) -> None:
    x, _ = make_dataset(0, n, p, 3, seed)
    r = O.ref_rf_fit(x, None, T, mtry, minndsize=1, iseed=0, unsupervised=True)
    np.savez_compressed(OUT / name, x=x, ntree=T, mtry=mtry, iseed=0, prox=r["prox"], nout=r["nout"],
                        outlier=r["outlier"])
    print(name, "prox sum", float(r["prox"].sum()))


if __name__ == "__main__":
    O.build(ref=True)
    one("gauss_n200.npz", 0, 200, 20, 3, 12, 4, 1, 0)
    one("covtype_cat_n240.npz", 1, 240, 54, 7, 8, 7, 10, 0)
    unsupervised("unsup_n160.npz", 160, 12, 10, 3, 7)
