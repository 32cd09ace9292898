"""Build liblfgpu.so (HIP, gfx950) in-tree with hipcc.

    python -m lehrfempp_amd.build

Every translation unit is compiled with -ffp-contract=off (the element
kernels restate the reference's evaluation order and must not be contracted
into FMAs) and with packed FP32 instructions disabled.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "liblfgpu.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("LFGPU_ARCH", "gfx950")

SOURCES = ["lfgpu_core.hip", "lfgpu_assemble.hip", "lfgpu_mfma.hip", "lfgpu_solver.hip",
           "lfgpu_sparse.hip", "lfgpu_post.hip", "lfgpu_edges.hip", "lfgpu_timestep.hip"]
HEADERS = ["lfgpu_internal.h", "lfgpu_fe_device.h", "lfgpu_reduce.h"]

FLAGS = [
    "-O3", "-std=c++17", "-fPIC", "--offload-arch=" + ARCH,
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
    "-I", os.path.join(ROOT, "include"),
]


def _compile(src: str) -> str:
    obj = os.path.join(BUILD, os.path.splitext(src)[0] + ".o")
    cmd = [HIPCC, *FLAGS, "-c", os.path.join(CSRC, src), "-o", obj]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError("hipcc failed for %s:\n%s\n%s" % (src, " ".join(cmd),
                                                           res.stderr[-8000:]))
    return obj


DEMO_SRC = os.path.join(ROOT, "tests", "cpp", "homdir_demo.cc")
DEMO_BIN = os.path.join(ROOT, "tests", "cpp", "homdir_demo")


def build_cpp_demo() -> str:
    """Compile the C++ facade driver (tests/cpp/homdir_demo.cc) with g++
    against liblfgpu.so."""
    cmd = ["g++", "-std=c++20", "-O2", "-Wall", "-I", ROOT, "-I", os.path.join(ROOT, "include"),
           DEMO_SRC, "-L", HERE, "-llfgpu", "-Wl,-rpath,$ORIGIN/../../lehrfempp_amd",
           "-o", DEMO_BIN]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError("g++ failed for the C++ facade driver:\n%s" % res.stderr[-8000:])
    return DEMO_BIN


def build(verbose: bool = False) -> str:
    os.makedirs(BUILD, exist_ok=True)
    with cf.ThreadPoolExecutor(max_workers=len(SOURCES)) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "-shared", "-fPIC", "--offload-arch=" + ARCH, *objs, "-o", tmp]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError("link failed:\n%s\n%s" % (" ".join(cmd), res.stderr[-8000:]))
    os.replace(tmp, LIB)
    build_cpp_demo()
    if verbose:
        print("built", LIB, "and", DEMO_BIN)
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
