"""Build librfx.so in-tree: HIP sources with hipcc for gfx950, host C++ with g++.

Run as ``python -m rfx_amd.build``.  Every call recompiles everything (a few
seconds per file); objects go to rfx_amd/_build/ and the library to
rfx_amd/librfx.so, which the snapshot carries to the GPU host.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
BUILD = PKG / "_build"
LIB = PKG / "librfx.so"
PYMOD_SRC = CSRC / "ref_py.cpp"  # pybind11 module of the reference-signature entry points

HIP_SOURCES = ["leaf.hip", "accum.hip", "casewise.hip", "finish.hip", "mds.hip", "factor.hip", "batch.hip", "outlier.hip"]
CXX_SOURCES = ["forest.cpp", "shim.cpp", "chains.cpp"]
# Per-file extra flags: finishprox must not contract a*b+c into an FMA.
HIP_EXTRA = {"finish.hip": ["-ffp-contract=off"]}


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found: the MI355X build needs ROCm's hipcc")


def hip_flags() -> list[str]:
    return [
        "-O3", "-std=c++17", "-fPIC", "--offload-arch=gfx950",
        # Packed FP32 VALU ops are kept out of all device code.
        "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
        f"-I{ROOT / 'include'}", f"-I{CSRC}",
        "-Wno-unused-command-line-argument",
    ]


def _run(cmd: list[str]) -> None:
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"build step failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")


def build(verbose: bool = False) -> Path:
    BUILD.mkdir(exist_ok=True)
    jobs = []
    for src in HIP_SOURCES:
        s = CSRC / src
        if not s.exists():
            continue
        o = BUILD / (src + ".o")
        jobs.append([hipcc(), *hip_flags(), *HIP_EXTRA.get(src, []), "-c", str(s), "-o", str(o)])
    for src in CXX_SOURCES:
        o = BUILD / (src + ".o")
        jobs.append(["g++", "-O2", "-std=c++17", "-fPIC", "-pthread", "-ffp-contract=off",
                     "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include",
                     f"-I{ROOT / 'include'}", "-c", str(CSRC / src), "-o", str(o)])
    workers = min(len(jobs), int(os.environ.get("MAX_JOBS", "8")) or 1)
    with cf.ThreadPoolExecutor(max_workers=workers) as ex:
        for cmd, fut in [(c, ex.submit(_run, c)) for c in jobs]:
            fut.result()
            if verbose:
                print(" ".join(cmd[:1] + cmd[-3:]), flush=True)
    objs = [j[-1] for j in jobs]
    tmp = LIB.with_suffix(".so.tmp")
    _run([hipcc(), "-shared", "-fPIC", "--offload-arch=gfx950", "-o", str(tmp), *objs, "-pthread"])
    os.replace(tmp, LIB)
    build_pymod()
    return LIB


def build_pymod() -> Path:
    """rfx_amd/_rfxref*.so, linked to librfx.so beside it."""
    import sysconfig

    import pybind11

    out = PKG / ("_rfxref" + sysconfig.get_config_var("EXT_SUFFIX"))
    tmp = out.with_suffix(".tmp")
    _run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include",
          f"-I{ROOT / 'include'}", f"-I{pybind11.get_include()}", f"-I{sysconfig.get_paths()['include']}",
          str(PYMOD_SRC), f"-L{PKG}", "-lrfx", "-L/opt/rocm/lib", "-lamdhip64", "-Wl,-rpath,$ORIGIN",
          "-Wl,-rpath,/opt/rocm/lib", "-o", str(tmp)])
    os.replace(tmp, out)
    return out


if __name__ == "__main__":
    p = build(verbose="-v" in sys.argv)
    print(p)
