"""The self-written golden fixtures (tests/golden/make_fixtures.py) against the
oracle, and against facts that do not depend on the oracle's code: closed
forms of the structured mesh and of P1 element matrices, an independent
sparse direct solver, and the known value of the continuous solution.

C1 follows homDir_linfe_demo.cc:114-175 (64 x 64 P1 Poisson, f = 1, zero
Dirichlet data).  The element fixtures are the first and last 64 cells of
the 2048 x 2048 P1 mesh and of the jittered 512 x 512 P3 mesh, evaluated by
ReactionDiffusionElementMatrixProvider (uscalfe/loc_comp_ellbvp.h:266-339)
with gamma = 1 and ScalarLoadElementVectorProvider (:691-746) with
f = sin(pi x) sin(pi y).
"""
import os

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from oracle import lfref as L
from tests.golden import make_fixtures as MF

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden")


@pytest.fixture(scope="module")
def c1():
    return dict(np.load(os.path.join(GOLD, "c1_poisson_64.npz")))


@pytest.fixture(scope="module")
def cells():
    return dict(np.load(os.path.join(GOLD, "cells_c2_c3.npz")))


def csr(g):
    n = len(g["row_ptr"]) - 1
    return sp.csr_matrix((g["val"], g["col"], g["row_ptr"]), shape=(n, n))


# --------------------------------------------------------------------- C1
def test_c1_fixture_is_the_oracle_output(c1):
    now = MF.c1_fixture()
    for k in ("row_ptr", "col", "val", "rhs", "flags"):
        assert np.array_equal(now[k], c1[k]), k
    assert np.allclose(now["x"], c1["x"], rtol=0, atol=1e-15)


def test_c1_counts_match_the_closed_forms(c1):
    # nnz after FixFlaggedSolutionComponents and the boundary DOF count of
    # the 64 x 64 mesh (4n boundary vertices)
    assert len(c1["row_ptr"]) - 1 == 65 * 65
    assert len(c1["val"]) == 27537
    assert int(c1["flags"].sum()) == 256


def test_c1_rows_are_the_five_point_stencil(c1):
    """Interior rows of the P1 Poisson matrix on the structured right
    triangles: 4 on the diagonal, -1 to the four axis neighbours, zero (up to
    rounding) to the two diagonal neighbours; the load of f = 1 is h^2."""
    A = csr(c1).tolil()
    n, h = 64, 1.0 / 64
    for (i, j) in ((1, 1), (17, 40), (32, 32), (63, 63)):
        r = j * (n + 1) + i
        row = dict(zip(A.rows[r], A.data[r]))
        assert abs(row.pop(r) - 4.0) < 1e-14
        for c in (r - 1, r + 1, r - (n + 1), r + (n + 1)):
            if not c1["flags"][c]:
                assert abs(row.pop(c) + 1.0) < 1e-14
            else:
                row.pop(c, None)
        assert all(abs(v) < 1e-14 for v in row.values()), row
        assert abs(c1["rhs"][r] - h * h) < 1e-18
    fixed = np.nonzero(c1["flags"])[0]
    for r in fixed[:16]:
        assert list(A.rows[r]) == [r] and A.data[r] == [1.0] and c1["rhs"][r] == 0.0


def test_c1_solution_against_an_independent_solver_and_the_pde(c1):
    A = csr(c1).tocsc()
    x_lu = spla.spsolve(A, c1["rhs"])
    x = c1["x"]
    assert np.linalg.norm(x - x_lu) <= 1e-12 * np.linalg.norm(x_lu)
    # interior rows cancel terms of size |A| |x| ~ 0.3 down to h^2: the
    # rounding floor of the residual is ~1e-13 relative to ||b||
    assert np.linalg.norm(c1["rhs"] - A @ x) <= 1e-12 * np.linalg.norm(c1["rhs"])
    # -Laplace u = 1 on the unit square, u = 0 on the boundary:
    # u(1/2, 1/2) = 0.0736713532814..., O(h^2) from the P1 value
    assert abs(x[32 * 65 + 32] - 0.07367135328) < 2e-4


def test_all_cores_pcg_baseline_solves_c1(c1):
    """The OpenMP PCG timed as the all-cores CPU baseline solves the same
    system: same iteration count as the serial oracle loop within one or
    two (reductions in another order), solution within kappa * tol of the
    direct one."""
    rp, col, val, b = c1["row_ptr"], c1["col"], c1["val"], c1["rhs"]
    x1, it1, e1 = L.pcg(rp, col, val, b, tol=1e-12)
    x2, it2, e2 = L.pcg_omp(rp, col, val, b, tol=1e-12)
    assert L.omp_threads() >= 1
    assert abs(it1 - it2) <= 2 and e2 < 1e-12
    assert np.linalg.norm(x2 - c1["x"]) <= 1e-8 * np.linalg.norm(c1["x"])


def test_autograd_binding_imports_without_a_gpu():
    """The checking binding is a torch.autograd.Function over the C ABI."""
    import torch
    from lehrfempp_amd import autograd
    assert issubclass(autograd._AdjointJ, torch.autograd.Function)
    assert callable(autograd.AdjointObjective)


# ----------------------------------------------------------- C2/C3 cells
def test_cell_fixture_inputs_are_the_benchmark_meshes(cells):
    n = 2048
    ne = 2 * n * n
    assert np.array_equal(cells["c2_cells"], MF.end_cells(ne))
    # structured vertices: multiples of h, every cell of area h^2 / 2
    v = cells["c2_verts"]
    assert np.array_equal(v * n, np.round(v * n))
    e1, e2 = v[:, 1] - v[:, 0], v[:, 2] - v[:, 0]
    area = 0.5 * np.abs(e1[:, 0] * e2[:, 1] - e1[:, 1] * e2[:, 0])
    assert np.allclose(area, 0.5 / n ** 2, rtol=1e-14, atol=0)
    theta = np.random.default_rng(0).standard_normal(ne)[cells["c2_cells"]]
    assert np.array_equal(cells["c2_theta"], theta)
    assert np.array_equal(cells["c2_alpha"], np.exp(theta))
    # the jittered P3 mesh: boundary vertices stay, interior ones move by
    # at most 0.2 h in each coordinate
    m = 512
    xy, cl = MF.structured_arrays(m)
    base = xy[cl[cells["c3_cells"]]]
    d = np.abs(cells["c3_verts"] - base)
    assert d.max() <= 0.2 / m
    on_bd = (base == 0) | (base == 1)
    assert np.all(d[on_bd.any(axis=2)] == 0)


@pytest.mark.parametrize("prefix,degree", [("c2_", 1), ("c3_", 3)])
def test_cell_fixture_is_the_oracle_output(cells, prefix, degree):
    sub = MF.cell_mesh(cells[prefix + "verts"])
    K = L.element_matrices(sub, degree, cells[prefix + "alpha"], 1.0)
    F = L.element_vectors(sub, degree, L.F_SINSIN)
    assert np.array_equal(K, cells[prefix + "mats"])
    assert np.array_equal(F, cells[prefix + "vecs"])


def test_p1_cell_matrices_match_the_closed_form(cells):
    """K = alpha S + M with the P1 stiffness S_ij = (b_i b_j + c_i c_j) /
    (4 |T|) (b, c: differences of the opposite vertices' coordinates, as in
    uscalfe/lin_fe.cc:60-84) and the P1 mass M = |T| / 12 (1 + delta_ij)."""
    v = cells["c2_verts"]
    K = cells["c2_mats"]
    for k in range(v.shape[0]):
        x, y = v[k, :, 0], v[k, :, 1]
        b = np.array([y[1] - y[2], y[2] - y[0], y[0] - y[1]])
        c = np.array([x[2] - x[1], x[0] - x[2], x[1] - x[0]])
        area = 0.5 * abs(b[0] * c[1] - b[1] * c[0])
        S = (np.outer(b, b) + np.outer(c, c)) / (4 * area)
        M = area / 12 * (np.ones((3, 3)) + np.eye(3))
        ref = cells["c2_alpha"][k] * S + M
        assert np.abs(K[k] - ref).max() <= 1e-14 * np.abs(ref).max(), k


def _triangle_integral(f, verts, order=12):
    """Integral of f over a triangle by a collapsed (Duffy) tensor
    Gauss-Legendre rule, exact for polynomials of degree 2 order - 2."""
    t, w = np.polynomial.legendre.leggauss(order)
    t, w = 0.5 * (t + 1), 0.5 * w
    u, vv = np.meshgrid(t, t, indexing="ij")
    W = np.outer(w, w) * (1 - u)
    s, r = u, vv * (1 - u)
    p = verts[0] + np.multiply.outer(s, verts[1] - verts[0]) + np.multiply.outer(r, verts[2] - verts[0])
    e1, e2 = verts[1] - verts[0], verts[2] - verts[0]
    det = abs(e1[0] * e2[1] - e1[1] * e2[0])
    return det * np.sum(W * f(p[..., 0], p[..., 1]))


@pytest.mark.parametrize("prefix,degree,tol", [("c2_", 1, 1e-5), ("c3_", 3, 1e-12)])
def test_cell_vectors_and_mass_sums(cells, prefix, degree, tol):
    """The partition of unity: sum_i F_i = integral of f over the cell (to
    the quadrature's accuracy), and with gamma = 1 the entries of K sum to
    the cell's area (the stiffness part annihilates constants)."""
    f = lambda x, y: np.sin(np.pi * x) * np.sin(np.pi * y)  # noqa: E731
    for k in range(cells[prefix + "verts"].shape[0]):
        V = cells[prefix + "verts"][k]
        exact = _triangle_integral(f, V)
        got = cells[prefix + "vecs"][k].sum()
        assert abs(got - exact) <= tol * abs(exact) + 1e-300, (k, got, exact)
        e1, e2 = V[1] - V[0], V[2] - V[0]
        area = 0.5 * abs(e1[0] * e2[1] - e1[1] * e2[0])
        K = cells[prefix + "mats"][k]
        assert abs(K.sum() - area) <= 1e-14 * np.abs(K).sum() + 1e-13 * area
        assert np.abs(K - K.T).max() <= 1e-14 * np.abs(K).max()
