"""C-ABI boundary checks that need no GPU: the library loads, exports every
symbol include/lfgpu.h declares, and its host-side topology helpers agree
with the oracle (and so with the reference's mesh/DOF numbering)."""
import os
import re

import numpy as np
import pytest

from lehrfempp_amd import capi
from oracle import lfref as L

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared_symbols():
    txt = open(os.path.join(ROOT, "include", "lfgpu.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(lfgpu_[a-z0-9_]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    lib = capi.lib()
    names = _declared_symbols()
    assert len(names) >= 30
    for n in names:
        assert hasattr(lib, n), n
    # the ctypes signature table covers the whole header
    assert set(names) == set(capi.SIGNATURES)


def test_ctypes_structs_match_the_c_layout(tmp_path):
    """Every struct the ctypes binding mirrors has the C compiler's size and
    field offsets (a program built by gcc against include/lfgpu.h prints
    them), so option and info blocks cross the ABI unchanged."""
    import ctypes
    import subprocess
    mirrors = {"lfgpu_coeff": capi.Coeff, "lfgpu_pcg_opts": capi.PcgOpts,
               "lfgpu_pcg_info": capi.PcgInfo, "lfgpu_strang_info": capi.StrangInfo}
    lines = ["#include <stdio.h>", "#include <stddef.h>", '#include "lfgpu.h"', "int main(void) {"]
    for cname, py in mirrors.items():
        lines.append('  printf("%s sizeof %%zu\\n", sizeof(%s));' % (cname, cname))
        for fname, _ in py._fields_:
            lines.append('  printf("%s %s %%zu\\n", offsetof(%s, %s));' % (cname, fname, cname, fname))
    lines += ["  return 0;", "}"]
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c11", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)],
                   check=True)
    got = {}
    for ln in subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split("\n"):
        if ln:
            cname, key, val = ln.split()
            got[(cname, key)] = int(val)
    for cname, py in mirrors.items():
        assert got[(cname, "sizeof")] == ctypes.sizeof(py), cname
        for fname, _ in py._fields_:
            assert got[(cname, fname)] == getattr(py, fname).offset, (cname, fname)


def test_error_path_without_gpu_work():
    # invalid arguments are rejected before any device call
    with pytest.raises(capi.LfgpuError) as e:
        capi.tp_triag_mesh(0)
    assert e.value.code == capi.EINVAL
    assert "bad mesh builder" in capi.lib().lfgpu_last_error().decode()


@pytest.mark.parametrize("n", [1, 3, 7])
def test_tp_mesh_matches_oracle(n):
    xy, cells, edges = capi.tp_triag_mesh(n)
    m = L.tp_triag_mesh(n)
    assert np.array_equal(xy, m.xy)
    assert np.array_equal(cells, m.cells)
    assert np.array_equal(edges, m.edges)


def test_rectangular_box_mesh_matches_oracle():
    xy, cells, edges = capi.tp_triag_mesh(5, 3, (-1.0, 0.5, 2.0, 1.75))
    m = L.tp_triag_mesh(5, 3, (-1.0, 0.5, 2.0, 1.75))
    assert np.array_equal(xy, m.xy) and np.array_equal(cells, m.cells)


@pytest.mark.parametrize("with_pre", [True, False])
def test_edges_and_p3_dofs_match_oracle(with_pre):
    n = 4
    xy, cells, edges = capi.tp_triag_mesh(n)
    pre = edges if with_pre else None
    ev, ce, co = capi.mesh_edges(len(xy), cells, pre)
    m = L.TriMesh(xy, cells, pre)
    assert np.array_equal(ev, m.edges)
    assert np.array_equal(ce, m.cell_e) and np.array_equal(co, m.cell_ori)
    cd, nd = capi.dof_table(len(xy), cells, len(ev), ce, co, 3)
    cdo, ndo = m.dofs(3)
    assert nd == ndo and np.array_equal(cd, cdo)


def test_edges_from_cells_only_on_test_mesh0(golden):
    # triangles of test mesh 0 (test_meshes.cc:26-66); created edges are
    # numbered in (min, max) endpoint order (mesh.cc:517-575)
    g = golden["test_mesh0"]
    tri = np.array([c for c in g["cells"] if len(c) == 3], np.int32)
    ev, ce, co = capi.mesh_edges(10, tri)
    keys = [tuple(sorted(e)) for e in ev]
    assert keys == sorted(keys)
    m = L.TriMesh(np.array(g["points"]), tri)
    assert np.array_equal(ev, m.edges) and np.array_equal(co, m.cell_ori)
