"""Builds libtdvmc.so (HIP kernels + C ABI) for gfx950, in-tree.

    python -m tdvmc_amd.build            # compile every source, link

Each translation unit is compiled to an object under tdvmc_amd/_build/ in
parallel and linked into tdvmc_amd/libtdvmc.so.  Packed-FP32 instructions are
disabled for all device code.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libtdvmc.so")

SOURCES = [
    "capi.cpp",
    "config.cpp",
    "writer.cpp",
    "knots.cpp",
    "kernels_eval.hip",
    "kernels_sweep.hip",
    "kernels_moments.hip",
    "kernels_solve.hip",
    "kernels_misc.hip",
    "kernels_obs.hip",
]

HIPCC = os.environ.get("HIPCC", shutil.which("hipcc") or "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
COMMON = ["-O3", "-std=c++17", "-fPIC", "-Wall", "-Wno-unused-function", "-I" + os.path.join(ROOT, "include")]
DEVICE = [f"--offload-arch={ARCH}", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          "-munsafe-fp-atomics"]


def _compile(src: str) -> str:
    obj = os.path.join(BUILD, os.path.basename(src) + ".o")
    path = os.path.join(CSRC, src)
    if src.endswith(".hip"):
        cmd = [HIPCC, "-x", "hip"] + DEVICE + COMMON + ["-c", path, "-o", obj]
    else:
        # host-only C++ that includes the HIP runtime API
        cmd = [HIPCC, "-x", "hip"] + DEVICE + COMMON + ["-c", path, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return obj


def build(verbose: bool = True) -> str:
    os.makedirs(BUILD, exist_ok=True)
    jobs = min(len(SOURCES), max(1, (os.cpu_count() or 4)))
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB)
    build_cpp_check(verbose)
    return LIB


CPP_CHECK_SRC = os.path.join(ROOT, "tests", "cpp", "adapter_check.cpp")
CPP_CHECK_BIN = os.path.join(ROOT, "tests", "cpp", "build", "adapter_check")


def build_cpp_check(verbose: bool = True) -> str:
    """Host-only C++ program over include/tdvmc.hpp (g++, no device code),
    linked against the in-tree libtdvmc.so with a relative run path."""
    os.makedirs(os.path.dirname(CPP_CHECK_BIN), exist_ok=True)
    cmd = ["g++", "-std=c++17", "-O2", "-Wall", "-I" + os.path.join(ROOT, "include"), CPP_CHECK_SRC,
           "-L" + HERE, "-ltdvmc", "-Wl,-rpath,$ORIGIN/../../../tdvmc_amd", "-o", CPP_CHECK_BIN]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"C++ check build failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    if verbose:
        print("built", CPP_CHECK_BIN)
    # the non-uniform knot tables looked up as the kernels do (host only)
    kc = os.path.join(os.path.dirname(CPP_CHECK_BIN), "knots_check")
    cmd = ["g++", "-std=c++17", "-O2", "-Wall", os.path.join(ROOT, "tests", "cpp", "knots_check.cpp"),
           os.path.join(CSRC, "knots.cpp"), "-o", kc]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"knots check build failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    if verbose:
        print("built", kc)
    # the SplinedFunction-shaped view of include/tdvmc_splined_function.hpp
    sc = os.path.join(os.path.dirname(CPP_CHECK_BIN), "splined_check")
    cmd = ["g++", "-std=c++17", "-O2", "-Wall", "-I" + os.path.join(ROOT, "include"),
           os.path.join(ROOT, "tests", "cpp", "splined_check.cpp"), "-L" + HERE, "-ltdvmc",
           "-Wl,-rpath,$ORIGIN/../../../tdvmc_amd", "-o", sc]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"SplinedFunction check build failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    if verbose:
        print("built", sc)
    return CPP_CHECK_BIN


if __name__ == "__main__":
    try:
        build()
    except RuntimeError as e:
        print(e, file=sys.stderr)
        sys.exit(1)
