"""Build libglsx.so (HIP kernels for gfx950 + the C ABI + the gls:: C++ layer).

Usage: python -m glscv_amd.build  (or glscv_amd.build.build()).
Sources are compiled in parallel with hipcc into glscv_amd/_build/ and linked
in-tree as glscv_amd/libglsx.so.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libglsx.so")

ARCH = os.environ.get("GLSX_OFFLOAD_ARCH", "gfx950")

SOURCES = [
    "glsx_api.cpp",
    "fft_kernels.hip",
    "poc.hip",
    "gls_api.cpp",
    "taskgraph.hip",
    "group_kernels.hip",
    "plans.hip",
    "server.hip",
    "split_mfma.hip",
    "spectrum.hip",
    "fft_r2c.hip",
    "noalloc.hip",
    "chain_ops.hip",
]

COMMON = [
    "-O3", "-fPIC", "-std=c++20", "-fvisibility=hidden",
    # no implicit FMA contraction: kernels fuse explicitly (fmaf), so every
    # launch shape produces the same bits for the same transform
    "-ffp-contract=off",
    f"--offload-arch={ARCH}",
    # packed FP32 instructions are kept out of the generated code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I", os.path.join(ROOT, "include"),
    "-Wno-unused-result",
]


def _hipcc() -> str:
    for cand in (shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def _compile(src: str) -> str:
    obj = os.path.join(BUILD, os.path.basename(src) + ".o")
    cmd = [_hipcc(), *COMMON, "-x", "hip", "-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return obj


def build(verbose: bool = False) -> str:
    os.makedirs(BUILD, exist_ok=True)
    srcs = [s for s in SOURCES if os.path.exists(os.path.join(CSRC, s))]
    with ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB + ".tmp"
    cmd = [_hipcc(), "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", tmp, *objs, "-lpthread"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    build_cpp_test(verbose)
    return LIB


CPP_TEST_SRC = os.path.join(ROOT, "tests", "cpp", "test_gls_api.cpp")
CPP_TEST_BIN = os.path.join(ROOT, "tests", "cpp", "test_gls_api")
ORACLE_DIR = os.path.join(ROOT, "oracle")


def build_cpp_test(verbose: bool = False) -> None:
    """The C++ gls:: API test program: a plain host program (g++) linked
    against libglsx.so and the test oracle."""
    if not os.path.exists(CPP_TEST_SRC) or not os.path.exists(os.path.join(ORACLE_DIR, "liboracle.so")):
        return
    cmd = ["g++", "-std=c++17", "-O2", CPP_TEST_SRC, "-I", os.path.join(ROOT, "include"),
           "-L", HERE, "-lglsx", "-L", ORACLE_DIR, "-loracle",
           f"-Wl,-rpath,{HERE}", f"-Wl,-rpath,{ORACLE_DIR}",
           # the repository may be moved or copied elsewhere: also search relative to the binary
           "-Wl,-rpath,$ORIGIN/../../glscv_amd", "-Wl,-rpath,$ORIGIN/../../oracle",
           "-o", CPP_TEST_BIN]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"C++ test build failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    if verbose:
        print(f"built {CPP_TEST_BIN}")


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
