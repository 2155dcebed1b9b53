"""Builds libvlbm.so (HIP, gfx950) in-tree: viscoplastic-lbm_amd/lib/libvlbm.so.

Each translation unit is compiled by hipcc in parallel, then linked into one
shared library exporting the C-ABI of include/vlbm.h.

    python viscoplastic-lbm_amd/build.py [--jobs N]
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf
import os
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
CSRC = PKG / "csrc"
OBJ = PKG / "_build"
LIB = PKG / "lib" / "libvlbm.so"

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("VLBM_OFFLOAD_ARCH", "gfx950")

# Packed FP32 instructions are disabled for all device code on this target.
COMMON = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unknown-pragmas",
]

SOURCES = [
    "kernels_fp64.hip",
    "kernels_fp32.hip",
    "kernels_fp16.hip",
    "kernels_bf16.hip",
    "kernels_common.hip",
    "kernels_ibm.hip",
    "ibm_capi.cpp",
    "kernels_batch.hip",
    "batch_capi.cpp",
    "checkpoint.cpp",
    "capi.cpp",
    "noalloc.hip",
]


def _compile(src: str) -> Path:
    out = OBJ / (Path(src).stem + ".o")
    cmd = [HIPCC, *COMMON, "-c", str(CSRC / src), "-o", str(out)]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{res.stdout}\n{res.stderr}")
    return out


def build(jobs: int | None = None, verbose: bool = True) -> Path:
    OBJ.mkdir(exist_ok=True)
    LIB.parent.mkdir(exist_ok=True)
    jobs = jobs or min(len(SOURCES), os.cpu_count() or 4)
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [HIPCC, "-shared", f"--offload-arch={ARCH}", "-fPIC", "-o", str(tmp), *map(str, objs)]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"link failed:\n{res.stdout}\n{res.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    return LIB


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--jobs", type=int, default=None)
    a = ap.parse_args()
    try:
        build(a.jobs)
    except RuntimeError as e:
        print(e, file=sys.stderr)
        sys.exit(1)
