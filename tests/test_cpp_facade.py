"""The C++ bindings over the C ABI.
* lehrfempp_amd/include/lf_amd/lf_amd.h: the facade compiles with the host
  toolchain, and its driver in the shape of homDir_linfe_demo.cc reproduces
  the fused device path through the reference calling pattern (GPU).
* lehrfempp_amd/include/lf_amd/lehrfem_binding.h: the binding to LehrFEM++'s
  own interfaces compiles against their archetype (tests/cpp/lf_archetype.h),
  and its driver shows the reference call expressions selecting the device
  overloads with triplets, load vectors and the fused system bit-identical to
  the reference's CPU cell loop (GPU)."""
import os
import subprocess

import pytest

from lehrfempp_amd import build as B


def test_facade_driver_compiles(tmp_path):
    out = str(tmp_path / "demo")
    cmd = ["g++", "-std=c++20", "-O0", "-Wall", "-Werror", "-I", B.ROOT, "-I",
           os.path.join(B.ROOT, "include"), B.DEMO_SRC, "-L", B.HERE, "-llfgpu", "-o", out]
    res = subprocess.run(cmd, capture_output=True, text=True)
    assert res.returncode == 0, res.stderr[-4000:]


@pytest.mark.gpu
def test_facade_driver_runs_on_gpu():
    assert os.path.exists(B.DEMO_BIN), "build the driver first (python -m lehrfempp_amd.build)"
    res = subprocess.run([B.DEMO_BIN], capture_output=True, text=True, timeout=120)
    print(res.stdout)
    assert res.returncode == 0, res.stdout + res.stderr
    assert "OK" in res.stdout


def test_binding_driver_compiles(tmp_path):
    out = str(tmp_path / "binding")
    cmd = B.binding_demo_cmd(out)
    cmd.insert(cmd.index("-Wextra") + 1, "-Werror")
    res = subprocess.run(cmd, capture_output=True, text=True)
    assert res.returncode == 0, res.stderr[-4000:]


@pytest.mark.gpu
def test_binding_driver_runs_on_gpu():
    assert os.path.exists(B.BINDING_BIN), "build the driver first (python -m lehrfempp_amd.build)"
    res = subprocess.run([B.BINDING_BIN], capture_output=True, text=True, timeout=120)
    print(res.stdout)
    assert res.returncode == 0, res.stdout + res.stderr
    assert "OK" in res.stdout
