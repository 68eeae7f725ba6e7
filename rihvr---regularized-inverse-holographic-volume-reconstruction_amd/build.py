"""Builds librihvr.so (HIP, gfx950) and the sparse-inverse-recon executable in-tree.

Every .hip source under csrc/ is compiled with hipcc for gfx950 only, the
host-only .cpp sources (input/output) with the host C++ compiler, and the
objects are linked into one shared library exporting the C ABI of
include/rihvr.h: <package>/_build/librihvr.so.  Packed FP32 instructions are
disabled for all device code.  The executables' main sources (EXECUTABLES) are
linked against that library under <package>/_build/ (runpath $ORIGIN).
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import pathlib
import shutil
import subprocess
import sys

PKG = pathlib.Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
OUT = PKG / "_build"
LIB = OUT / "librihvr.so"
# executables linked against the library: main source -> file name under _build/
EXECUTABLES = {"cli_main.cpp": "sparse-inverse-recon", "make_background_main.cpp": "make-background",
               "segmentation_main.cpp": "segmentation", "ccbg_main.cpp": "ccbg"}
CLI = OUT / EXECUTABLES["cli_main.cpp"]

ARCH = "gfx950"
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
CXXFLAGS = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
    f"-I{ROOT / 'include'}",
]


# Per-source flags.  warmstart.hip: MFMA accumulators in VGPRs instead of AGPRs (its kernels
# otherwise copied their accumulators between AGPR sets at the double-buffered loop's edges):
# training step 371 -> 435 steps/s and the 1024^2 x 128 prediction 23.0 -> 22.6 ms, the same
# results bit for bit, no kernel losing waves per SIMD.
EXTRA_FLAGS = {"warmstart.hip": ["-mllvm", "-amdgpu-mfma-vgpr-form"]}


HOST_CXX = os.environ.get("CXX", "g++")
HOST_CXXFLAGS = ["-O2", "-std=c++17", "-fPIC", "-Wall", f"-I{ROOT / 'include'}"]
ROCM_LIB = pathlib.Path(os.environ.get("ROCM_PATH", "/opt/rocm")) / "lib"


def sources():
    hip = sorted(CSRC.glob("*.hip"))
    host = sorted(p for p in CSRC.glob("*.cpp") if p.name not in EXECUTABLES)
    return hip + host


def _compile(src: pathlib.Path, verbose: bool) -> pathlib.Path:
    obj = OUT / (src.stem + ".o")
    if src.suffix == ".hip":
        cmd = [HIPCC, *CXXFLAGS, *EXTRA_FLAGS.get(src.name, []), "-c", str(src), "-o", str(obj)]
    else:
        cmd = [HOST_CXX, *HOST_CXXFLAGS, "-c", str(src), "-o", str(obj)]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed for {src.name}:\n{r.stdout}\n{r.stderr}")
    if r.stderr.strip() and verbose:
        print(r.stderr, file=sys.stderr)
    return obj


def build(verbose: bool = False, jobs: int | None = None) -> pathlib.Path:
    if shutil.which(HIPCC) is None and not pathlib.Path(HIPCC).exists():
        raise RuntimeError(f"hipcc not found at {HIPCC}")
    OUT.mkdir(parents=True, exist_ok=True)
    srcs = sources()
    jobs = jobs or min(len(srcs), max(1, (os.cpu_count() or 2)))
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(lambda s: _compile(s, verbose), srcs))
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp), *map(str, objs)]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    for main, name in EXECUTABLES.items():
        exe = OUT / name
        tmp = exe.with_name(name + ".tmp")
        cmd = [HOST_CXX, *HOST_CXXFLAGS, str(CSRC / main), "-o", str(tmp), f"-L{OUT}", "-lrihvr",
               f"-L{ROCM_LIB}", f"-Wl,-rpath-link,{ROCM_LIB}", "-Wl,-rpath,$ORIGIN"]
        if verbose:
            print(" ".join(cmd), flush=True)
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"link of {name} failed:\n{r.stdout}\n{r.stderr}")
        os.replace(tmp, exe)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
