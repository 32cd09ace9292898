"""The HIP path (through the C ABI) against the committed golden fixtures
(tests/golden/make_fixtures.py): the 64 x 64 P1 Poisson system of
homDir_linfe_demo.cc:114-175 and the element matrices/vectors of the first
and last 64 cells of the 2048 x 2048 P1 and jittered 512 x 512 P3 meshes.

Bars: CSR pattern, values and load vector bit-exact; element matrices
bit-exact through the exact-order kernel with per-cell alpha; 1e-15 of each
row's largest entry with alpha = exp(theta) evaluated on the device (device
exp vs glibc exp, within one ulp); 1e-14 through the matrix-core contraction
(a different summation order); load vectors 1e-14 of the cell's largest
entry (device sin vs glibc sin).  PCG solutions agree with the direct
solution to kappa * rtol, with kappa ~ 1.7e3 for the Jacobi-preconditioned
64 x 64 Laplacian.
"""
import os

import numpy as np
import pytest

from lehrfempp_amd import capi
from tests.golden import make_fixtures as MF

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


@pytest.fixture(scope="module")
def c1():
    return dict(np.load(os.path.join(GOLD, "c1_poisson_64.npz")))


@pytest.fixture(scope="module")
def cells():
    return dict(np.load(os.path.join(GOLD, "cells_c2_c3.npz")))


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def test_c1_assembly_is_bitexact_and_the_solves_agree(ctx, c1):
    xy, cl = MF.structured_arrays(64)
    mesh, fe = capi.Mesh(ctx, xy, cl), capi.FE(ctx, 1)
    plan = capi.Plan(ctx, mesh, c1["flags"])
    b, x = capi.Vec(ctx, xy.shape[0]), capi.Vec(ctx, xy.shape[0])
    plan.assemble(fe, capi.coeff(capi.COEFF_CONST, 1.0), capi.coeff(capi.COEFF_CONST, 0.0),
                  capi.coeff(capi.COEFF_CONST, 1.0), b)
    rp, col, val = plan.csr.download()
    assert np.array_equal(rp, c1["row_ptr"]) and np.array_equal(col, c1["col"])
    assert np.array_equal(val, c1["val"])
    assert np.array_equal(b.numpy(), c1["rhs"])
    for mode, rtol, tol in ((capi.PCG_FP64, 1e-13, 1e-9), (capi.PCG_MIXED_MATRIX, 1e-12, 5e-9),
                            (capi.PCG_MIXED, 1e-12, 5e-9)):
        x.fill(0.0)
        info = capi.pcg(ctx, plan.csr, b, x, rtol=rtol, mode=mode)
        assert info.converged
        assert rel(x.numpy(), c1["x"]) <= tol, (mode, rel(x.numpy(), c1["x"]))
    for h in (x, b, plan, fe, mesh):
        h.close()


def _row_scaled(a, b):
    scale = np.maximum(np.abs(b).max(axis=-1, keepdims=True), 1e-300)
    return float((np.abs(a - b) / scale).max())


@pytest.mark.parametrize("prefix,degree", [("c2_", 1), ("c3_", 3)])
def test_fixture_cells_on_the_device(ctx, cells, prefix, degree):
    m = MF.cell_mesh(cells[prefix + "verts"])
    if degree == 1:
        g = capi.Mesh(ctx, m.xy, m.cells)
    else:
        cd, nd = m.dofs(3)
        g = capi.Mesh(ctx, m.xy, m.cells, cd, nd)
    fe = capi.FE(ctx, degree)
    va = capi.Vec(ctx, cells[prefix + "alpha"])
    vt = capi.Vec(ctx, cells[prefix + "theta"])
    one = capi.coeff(capi.COEFF_CONST, 1.0)
    K_ref, F_ref = cells[prefix + "mats"], cells[prefix + "vecs"]

    K = capi.element_matrices(ctx, g, fe, capi.coeff(capi.COEFF_CELL, 1.0, va), one)
    assert np.array_equal(K, K_ref), np.abs(K - K_ref).max()
    K = capi.element_matrices(ctx, g, fe, capi.coeff(capi.COEFF_EXPCELL, 1.0, vt), one)
    assert _row_scaled(K, K_ref) <= 1e-15
    K = capi.element_matrices(ctx, g, fe, capi.coeff(capi.COEFF_CELL, 1.0, va), one, mfma=True)
    assert _row_scaled(K, K_ref) <= 1e-14
    F = capi.element_vectors(ctx, g, fe, capi.coeff(capi.COEFF_SINSIN, 1.0))
    assert _row_scaled(F, F_ref) <= 1e-14
    for h in (vt, va, fe, g):
        h.close()
