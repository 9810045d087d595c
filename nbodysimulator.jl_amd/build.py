"""Build libnbx.so (the gfx950 HIP engine behind include/nbx.h) in-tree.

Every HIP translation unit is compiled for gfx950 only, with packed-FP32 VALU
instructions disabled (-packed-fp32-ops); the FP64 reference-order kernels are also
built with -ffp-contract=off so that no a*b+c is fused into an FMA.  Object files go
to _build/, the shared library to lib/libnbx.so.

    python nbodysimulator.jl_amd/build.py [-j JOBS]
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIBDIR = os.path.join(HERE, "lib")
LIB = os.path.join(LIBDIR, "libnbx.so")
ARCH = "gfx950"

# (source, extra flags)
SOURCES = [
    ("nbx_ref64.hip", ["-ffp-contract=off"]),
    ("nbx_nonbonded.hip", []),
    ("nbx_sort.hip", []),
    ("nbx_gravity.hip", []),
    ("nbx_batch.hip", []),
    ("nbx_batch.cpp", []),
    ("nbx_api.cpp", ["-ffp-contract=off"]),
    ("nbx_mixed.cpp", []),
    ("nbx_learn.hip", []),
    ("nbx_learn.cpp", []),
]


def hipcc() -> str:
    for cand in (shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found (ROCm is required to build libnbx)")


def compile_one(src: str, extra: list[str]) -> str:
    path = os.path.join(CSRC, src)
    obj = os.path.join(BUILD, src + ".o")
    common = ["-O3", "-fPIC", "-std=c++17", "-Wall", "-I", os.path.join(ROOT, "include"),
              "-I", CSRC]
    if src.endswith(".hip"):
        cmd = [hipcc(), "-x", "hip", f"--offload-arch={ARCH}",
               "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
               "-fno-gpu-rdc", *common, *extra, "-c", path, "-o", obj]
    else:
        cmd = [hipcc(), *common, *extra, "-c", path, "-o", obj]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{res.stdout}\n{res.stderr}")
    return obj


def build(jobs: int = 8, verbose: bool = False) -> str:
    os.makedirs(BUILD, exist_ok=True)
    os.makedirs(LIBDIR, exist_ok=True)
    with cf.ThreadPoolExecutor(max_workers=max(1, jobs)) as ex:
        futs = [ex.submit(compile_one, s, e) for s, e in SOURCES]
        objs = [f.result() for f in futs]
    tmp = LIB + ".tmp"
    cmd = [hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{res.stdout}\n{res.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    return LIB


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("-j", "--jobs", type=int, default=min(8, os.cpu_count() or 1))
    args = ap.parse_args(argv)
    build(args.jobs, verbose=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
