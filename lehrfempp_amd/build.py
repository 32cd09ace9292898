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
           "lfgpu_sparse.hip", "lfgpu_post.hip", "lfgpu_edges.hip", "lfgpu_timestep.hip",
           "lfgpu_nonlocal.hip"]
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
BINDING_SRC = os.path.join(ROOT, "tests", "cpp", "binding_demo.cc")
BINDING_BIN = os.path.join(ROOT, "tests", "cpp", "binding_demo")
ORACLE_DIR = os.path.join(ROOT, "oracle")


def binding_demo_cmd(out: str) -> list:
    """g++ line of the LehrFEM++-binding driver (tests/cpp/binding_demo.cc):
    the binding header against the archetype of the reference interfaces,
    linked with liblfgpu.so and, as its checker, the oracle."""
    return ["g++", "-std=c++20", "-O2", "-Wall", "-Wextra", "-I", ROOT, "-I",
            os.path.join(ROOT, "include"), BINDING_SRC, "-L", HERE, "-llfgpu", "-L", ORACLE_DIR,
            "-llfref", "-Wl,-rpath,$ORIGIN/../../lehrfempp_amd:$ORIGIN/../../oracle", "-o", out]


def build_cpp_demo() -> str:
    """Compile the C++ drivers with g++ against liblfgpu.so: the facade
    driver (tests/cpp/homdir_demo.cc) and, when the oracle library is built,
    the LehrFEM++-binding driver (tests/cpp/binding_demo.cc)."""
    cmd = ["g++", "-std=c++20", "-O2", "-Wall", "-I", ROOT, "-I", os.path.join(ROOT, "include"),
           DEMO_SRC, "-L", HERE, "-llfgpu", "-Wl,-rpath,$ORIGIN/../../lehrfempp_amd",
           "-o", DEMO_BIN]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError("g++ failed for the C++ facade driver:\n%s" % res.stderr[-8000:])
    if os.path.exists(os.path.join(ORACLE_DIR, "liblfref.so")):
        res = subprocess.run(binding_demo_cmd(BINDING_BIN), capture_output=True, text=True)
        if res.returncode != 0:
            raise RuntimeError("g++ failed for the binding driver:\n%s" % res.stderr[-8000:])
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
