"""Pins the CPU oracle (oracle/lfref.cc) against the reference's own golden
vectors, closed forms and invariants.  CPU only.

Each check is chosen so that a plausible slip in the restatement (a dropped
term, a transposed Jacobian, a wrong index or summation order) fails one of
them; the reference test each one mirrors is cited.
"""
import math

import numpy as np
import pytest

from oracle import lfref as L


# ---------------------------------------------------------------- quadrature
@pytest.mark.parametrize("qdeg", [2, 6])
def test_quad_rule_exactness(qdeg):
    # quad/test/make_quad_rule_tests.cc: monomials x^i y^j, i + j <= degree
    pts, w = L.quad_rule(qdeg)
    assert abs(w.sum() - 0.5) < 1e-15
    for i in range(qdeg + 1):
        for j in range(qdeg + 1 - i):
            exact = math.factorial(i) * math.factorial(j) / math.factorial(i + j + 2)
            approx = float(np.sum(w * pts[:, 0] ** i * pts[:, 1] ** j))
            assert abs(approx - exact) < 1e-14, (i, j)
    # one degree higher is not integrated exactly by the degree-2 rule
    if qdeg == 2:
        errs = [abs(np.sum(w * pts[:, 0] ** i * pts[:, 1] ** (3 - i)) -
                    math.factorial(i) * math.factorial(3 - i) / math.factorial(5))
                for i in range(4)]
        assert max(errs) > 1e-4


def test_unsupported_rule():
    with pytest.raises(ValueError):
        L.quad_rule(3)


# ----------------------------------------------------------- shape functions
def test_p1_shape_known_values(golden):
    g = golden["p1_shape_values"]
    N, dN = L.shape_functions(1, np.array(g["points"]))
    # EXPECT_DOUBLE_EQ: equal to within 4 ulp
    assert np.allclose(N.T, np.array(g["values"]), rtol=4 * 2.2e-16, atol=0)
    assert np.array_equal(dN[:, :2], [[-1, -1], [1, 0], [0, 1]])


def _p3_nodes():
    t = 1.0 / 3.0
    return np.array([[0, 0], [1, 0], [0, 1], [t, 0], [2 * t, 0], [2 * t, t], [t, 2 * t],
                     [0, 2 * t], [0, t], [t, t]])


def test_p3_cardinal_basis():
    # lagr_fe_tests.cc:293-305: N_i(node_j) = delta_ij
    N, _ = L.shape_functions(3, _p3_nodes())
    assert np.allclose(N, np.eye(10), atol=1e-14)


@pytest.mark.parametrize("degree", [1, 3])
def test_gradients_match_finite_differences(degree):
    rng = np.random.default_rng(4)
    p = rng.uniform(0.05, 0.45, size=(7, 2))
    _, dN = L.shape_functions(degree, p)
    h = 1e-6
    for d in range(2):
        e = np.zeros(2)
        e[d] = h
        Np, _ = L.shape_functions(degree, p + e)
        Nm, _ = L.shape_functions(degree, p - e)
        fd = (Np - Nm) / (2 * h)
        assert np.allclose(dN[:, d::2], fd, atol=1e-7)
    N, _ = L.shape_functions(degree, p)
    assert np.allclose(N.sum(axis=0), 1.0, atol=1e-14)  # partition of unity


# ------------------------------------------------------------------ geometry
def test_tria_o1_reference_triangle(golden):
    nodes = np.array(golden["tria_o1_nodes"]["nodes"], dtype=float)
    m = L.TriMesh(nodes, np.array([[0, 1, 2]]))
    J, Jit, det = L.geometry(m)
    assert np.array_equal(J[0], [[3.0, 2.0], [1.0, 4.0]])
    assert det[0] == 10.0
    assert np.allclose(Jit[0], np.linalg.inv(J[0]).T, rtol=1e-15)
    # JacobianInverseGramian == J (J^T J)^-1 (check_jacobian_inverse_gramian.cc)
    assert np.allclose(Jit[0], J[0] @ np.linalg.inv(J[0].T @ J[0]), rtol=1e-14)


def test_geometry_random_triangles():
    rng = np.random.default_rng(0)
    xy = rng.uniform(-2, 3, size=(60, 2))
    cells = np.arange(60, dtype=np.int32).reshape(20, 3)
    m = L.TriMesh(xy, cells)
    J, Jit, det = L.geometry(m)
    for c in range(20):
        a, b, d = xy[cells[c]]
        Jn = np.column_stack([b - a, d - a])
        assert np.allclose(J[c], Jn, rtol=0, atol=1e-15)
        assert np.allclose(Jit[c], np.linalg.inv(Jn).T, rtol=1e-13)
        assert abs(det[c] - abs(np.linalg.det(Jn))) < 1e-13 * abs(np.linalg.det(Jn))


def test_degenerate_triangle_rejected():
    m = L.TriMesh(np.array([[0, 0], [1, 1], [2, 2.0]]), np.array([[0, 1, 2]]))
    with pytest.raises(ValueError):
        L.geometry(m)


# ----------------------------------------------------------- mesh / dofs
def test_tp_mesh_counts():
    # mesh/utils/test/tp_triag_mesh_builder_tests.cc: 2x2 -> 8 cells, 16 edges, 9 nodes
    m = L.tp_triag_mesh(2)
    assert (m.ne, m.nedges, m.nv) == (8, 16, 9)
    nf, ef = m.boundary_flags()
    assert nf.sum() == 8 and ef.sum() == 8


def test_tp_mesh_closed_forms():
    n = 3
    m = L.tp_triag_mesh(n)
    V = lambda i, j: i + j * (n + 1)  # noqa: E731
    for i in range(n):
        for j in range(n):
            assert list(m.cells[2 * (i * n + j)]) == [V(i, j), V(i + 1, j + 1), V(i, j + 1)]
            assert list(m.cells[2 * (i * n + j) + 1]) == [V(i, j), V(i + 1, j), V(i + 1, j + 1)]
    # predefined edges keep the builder's order: horizontal, vertical, diagonal
    assert list(m.edges[0]) == [V(0, 0), V(1, 0)]
    assert list(m.edges[n * (n + 1)]) == [V(0, 0), V(0, 1)]
    assert list(m.edges[n * (n + 1) + (n + 1) * n]) == [V(0, 0), V(1, 1)]
    # up cell: e0 = diagonal (+), e1 = horizontal (i, j+1) (-), e2 = vertical (-)
    assert list(m.cell_ori[0]) == [1, -1, -1]
    assert list(m.cell_ori[1]) == [1, 1, -1]


def test_test_mesh0_counts(golden):
    # assembly_tests.cc:154-161 (the triangle-only part: quads are not
    # represented by the triangle oracle, so count the triangles' edges)
    g = golden["test_mesh0"]
    assert g["num_cells"] == len(g["cells"]) and g["num_nodes"] == len(g["points"])
    edges = set()
    for c in g["cells"]:
        for k in range(len(c)):
            a, b = c[k], c[(k + 1) % len(c)]
            edges.add((min(a, b), max(a, b)))
    assert len(edges) == g["num_edges"]


@pytest.mark.parametrize("n", [2, 3, 5])
def test_p3_dofs_and_pattern_counts(n):
    m = L.tp_triag_mesh(n)
    cd, nd = m.dofs(3)
    assert nd == m.nv + 2 * m.nedges + m.ne
    assert set(np.unique(cd)) == set(range(nd))
    rp, col, val, b = L.galerkin_system(m, 3, 1.0, 1.0)
    assert len(col) == 153 * n * n - 234 * n + 109  # nnz after the Dirichlet fix
    # every cell's DOF set is shared consistently: interior edge DOFs are seen
    # in opposite order by the two adjacent cells
    fl = m.dirichlet_flags(3)
    assert fl.sum() == 4 * n * 3


def test_p1_pattern_counts():
    m = L.tp_triag_mesh(64)
    rp, col, val, b = L.galerkin_system(m, 1, 1.0, 0.0)
    assert len(col) == 27537 and m.dirichlet_flags(1).sum() == 256


# ----------------------------------------------------------- element level
def test_p1_element_matrix_closed_form():
    # uscalfe/lin_fe.cc:60-84: area * grad(lambda)^T grad(lambda); mass
    # (area/12)(1 + delta_ij)
    rng = np.random.default_rng(1)
    xy = rng.uniform(0, 1, size=(30, 2))
    cells = np.arange(30, dtype=np.int32).reshape(10, 3)
    m = L.TriMesh(xy, cells)
    alpha = rng.uniform(0.5, 2, size=10)
    gamma = rng.uniform(0.5, 2, size=10)
    K = L.element_matrices(m, 1, alpha, gamma)
    for c in range(10):
        a, b, d = xy[cells[c]]
        area = 0.5 * abs(np.linalg.det(np.column_stack([b - a, d - a])))
        X = np.column_stack([np.ones(3), xy[cells[c]]])
        grads = np.linalg.inv(X)[1:, :]  # 2 x 3
        ref = alpha[c] * area * grads.T @ grads + gamma[c] * area / 12 * (np.ones((3, 3)) + np.eye(3))
        assert np.allclose(K[c], ref, rtol=1e-12, atol=1e-14)


@pytest.mark.parametrize("degree", [1, 3])
def test_element_invariants(degree):
    # bvp_fe_tests.cc:57-60 (diffusion rows sum to 0), loc_comp_test.cc:46-84
    # (1^T M 1 = |K|)
    rng = np.random.default_rng(2)
    xy = rng.uniform(0, 1, size=(24, 2))
    cells = np.arange(24, dtype=np.int32).reshape(8, 3)
    m = L.TriMesh(xy, cells)
    _, _, det = L.geometry(m)
    A = L.element_matrices(m, degree, 1.0, 0.0)
    assert np.abs(A.sum(axis=2)).max() < 1e-10
    M = L.element_matrices(m, degree, 0.0, 1.0)
    assert np.allclose(M.sum(axis=(1, 2)), det / 2, rtol=1e-13)
    assert np.allclose(M, M.transpose(0, 2, 1), rtol=1e-14, atol=1e-16)


def test_p3_stiffness_energy_of_quadratic():
    # the P3 interpolant of v = x^2 is exact; v^T K v = int |grad v|^2 = 4/3
    m = L.tp_triag_mesh(4)
    cd, nd = m.dofs(3)
    rp, col, val, _ = L.set_from_triplets(*L.assemble_coo(cd, L.element_matrices(m, 3, 1.0, 0.0)),
                                          nd) + (None,)
    # DOF locations: evaluate the shape-function nodes through each cell's map
    loc = np.zeros((nd, 2))
    nodes = _p3_nodes()
    for c in range(m.ne):
        a, b, d = m.xy[m.cells[c]]
        for k in range(10):
            loc[cd[c, k]] = a + nodes[k, 0] * (b - a) + nodes[k, 1] * (d - a)
    v = loc[:, 0] ** 2
    Av = L.csr_spmv(rp, col, val, v)
    assert abs(v @ Av - 4.0 / 3.0) < 1e-12


def test_load_vector_equals_mass_times_linear_f():
    # f linear: degree-2 quadrature is exact, so the load vector is M f_I
    rng = np.random.default_rng(3)
    xy = rng.uniform(0, 1, size=(18, 2))
    cells = np.arange(18, dtype=np.int32).reshape(6, 3)
    m = L.TriMesh(xy, cells)
    q = L.qp_coords(m, 1)
    fq = 1 + 2 * q[..., 0] + 3 * q[..., 1]
    v = L.element_vectors(m, 1, L.F_QP, f_qp=fq)
    M = L.element_matrices(m, 1, 0.0, 1.0)
    fv = 1 + 2 * xy[cells][..., 0] + 3 * xy[cells][..., 1]
    assert np.allclose(v, np.einsum("cij,cj->ci", M, fv), rtol=1e-13)
    # constant f: entries sum to f |K|
    _, _, det = L.geometry(m)
    v1 = L.element_vectors(m, 1, L.F_CONST, 2.5)
    assert np.allclose(v1.sum(axis=1), 2.5 * det / 2, rtol=1e-14)
    # sin(pi x) sin(pi y) kind equals the same values passed per QP
    vs = L.element_vectors(m, 1, L.F_SINSIN)
    vq = L.element_vectors(m, 1, L.F_QP, f_qp=np.sin(np.pi * q[..., 0]) * np.sin(np.pi * q[..., 1]))
    assert np.allclose(vs, vq, rtol=1e-15, atol=1e-18)


# ------------------------------------------------- COO / setFromTriplets / fix
def test_golden_assembly_matrix(golden):
    # assembly_tests.cc:219-285 with TestAssembler on test mesh 0
    g = golden["test_mesh0"]
    rows, cols, vals = [], [], []
    for c, cell in enumerate(g["cells"]):
        for i in cell:
            for j in cell:
                rows.append(i)
                cols.append(j)
                vals.append(float(c) if i == j else -float(c))
    rp, col, val = L.set_from_triplets(rows, cols, vals, 10)
    dense = np.zeros((10, 10))
    for r in range(10):
        dense[r, col[rp[r]:rp[r + 1]]] = val[rp[r]:rp[r + 1]]
        assert np.all(np.diff(col[rp[r]:rp[r + 1]]) > 0)
    ref = np.array(golden["test_assembler_matrix"]["matrix"], dtype=float).reshape(10, 10)
    assert np.array_equal(dense, ref)
    # explicit zeros from cell 0 are kept in the pattern
    assert len(col) == len(set(zip(rows, cols)))


def test_set_from_triplets_sums_in_insertion_order():
    # three values whose sum depends on the order: (1e16 + 1) - 1e16 vs ...
    rows = [0, 0, 0]
    cols = [1, 1, 1]
    vals = [1e16, 1.0, -1e16]
    rp, col, val = L.set_from_triplets(rows, cols, vals, 2)
    assert val[0] == (1e16 + 1.0) + (-1e16)
    rp, col, val = L.set_from_triplets(rows, cols, [1.0, 1e16, -1e16], 2)
    assert val[0] == (1.0 + 1e16) + (-1e16)


def test_coo_matvec_golden(golden):
    g = golden["coo_matvec"]
    t = np.array(g["triplets_1based"])
    rows = (t[:, 0] - 1).astype(np.int32)
    cols = (t[:, 1] - 1).astype(np.int32)
    vals = t[:, 2].astype(float)
    x = np.linspace(0.5, 5.0, 10) + np.linspace(9.0, 0.0, 10)
    y = L.coo_matvec(rows, cols, vals, 1.0, x, np.zeros(8))
    assert np.array_equal(y, np.array(g["exact"], dtype=float))


def test_fix_flagged_golden(golden):
    g = golden["fix_dof"]
    n = g["n"]
    rows, cols, vals = [], [], []
    for k in range(n):
        if k > 0:
            rows.append(k); cols.append(k - 1); vals.append(-1.0)
        if k < n - 1:
            rows.append(k); cols.append(k + 1); vals.append(-1.0)
        rows.append(k); cols.append(k); vals.append(2.0)
    b = np.arange(1, n + 1, dtype=float)
    flags = np.zeros(n, np.uint8)
    fv = np.ones(n)
    for k, v in g["fixed"].items():
        flags[int(k)] = 1
        fv[int(k)] = v
    r, c, v, b2 = L.fix_flagged(np.array(rows, np.int32), np.array(cols, np.int32),
                                np.array(vals), flags, fv, b)
    rp, col, val = L.set_from_triplets(r, c, v, n)
    x = L.band_cholesky_solve(rp, col, val, b2)
    assert np.linalg.norm(x - np.array(g["exact"])) < 1e-12


# --------------------------------------------------------- global operator
def test_energies_unit_square(golden):
    # full_gal_tests.cc:49-236 on the unit square
    m = L.tp_triag_mesh(8)
    q = L.qp_coords(m, 1)
    x, y = q[..., 0], q[..., 1]
    coeffs = {"0": 0.0 * x, "1": 1.0 + 0.0 * x, "1+x": 1.0 + x, "1+x^2+y^2": 1 + x * x + y * y}
    cd, nd = m.dofs(1)
    for case in golden["energies_unit_square"]["cases"]:
        A = L.element_matrices(m, 1, coeffs[case["alpha"]], coeffs[case["gamma"]])
        rp, col, val = L.set_from_triplets(*L.assemble_coo(cd, A), nd)
        v = np.ones(nd) if case["v"] == "1" else 2 * m.xy[:, 0] + m.xy[:, 1]
        e = v @ L.csr_spmv(rp, col, val, v)
        assert abs(e - case["energy"]) < 1e-12, case


def test_manufactured_solution_converges():
    # homDir_linfe_demo.cc:45-57: u = sin(pi x) sin(pi y), f = 2 pi^2 u
    errs = []
    for n in (16, 32):
        m = L.tp_triag_mesh(n)
        q = L.qp_coords(m, 1)
        f = 2 * np.pi ** 2 * np.sin(np.pi * q[..., 0]) * np.sin(np.pi * q[..., 1])
        rp, col, val, b = L.galerkin_system(m, 1, 1.0, 0.0, L.F_QP, f_qp=f)
        u = L.band_cholesky_solve(rp, col, val, b)
        ex = np.sin(np.pi * m.xy[:, 0]) * np.sin(np.pi * m.xy[:, 1])
        errs.append(np.abs(u - ex).max())
    assert 3.5 < errs[0] / errs[1] < 4.5  # O(h^2) nodal error


def test_pcg_matches_direct_solve_c1():
    m = L.tp_triag_mesh(64)
    rp, col, val, b = L.galerkin_system(m, 1, 1.0, 0.0)
    x = L.band_cholesky_solve(rp, col, val, b)
    y, it, err = L.pcg(rp, col, val, b, tol=1e-12)
    assert err < 1e-12 and 50 < it < 1000
    assert np.abs(x - y).max() < 1e-12 * np.abs(x).max() * 100
    # residual certificate of the direct solve
    assert np.linalg.norm(L.csr_spmv(rp, col, val, x) - b) < 1e-12 * np.linalg.norm(b)


def test_cpu_path_equals_pipeline():
    m = L.tp_triag_mesh(12)
    cd, nd = m.dofs(1)
    rng = np.random.default_rng(5)
    a = np.exp(rng.standard_normal(m.ne))
    fl = m.dirichlet_flags(1)
    r1 = L.cpu_path_assemble(m, 1, cd, nd, a, 1.0, L.F_SINSIN, 0.0, fl)
    r2 = L.galerkin_system(m, 1, a, 1.0, L.F_SINSIN)
    for u, v in zip(r1, r2):
        assert np.array_equal(u, v)


def test_adjoint_gradient_matches_finite_differences():
    n = 6
    m = L.tp_triag_mesh(n)
    rng = np.random.default_rng(7)
    theta = rng.standard_normal(m.ne)
    fl = m.dirichlet_flags(1)
    Mrp, Mcol, Mval, _ = L.galerkin_system(m, 1, 0.0, 1.0)
    uobs = rng.standard_normal(m.nv) * (1 - fl)

    def solve(th):
        rp, col, val, b = L.galerkin_system(m, 1, np.exp(th), 0.0, L.F_CONST, 1.0)
        return rp, col, val, L.band_cholesky_solve(rp, col, val, b)

    def J(th):
        u = solve(th)[3]
        d = u - uobs
        return 0.5 * d @ L.csr_spmv(Mrp, Mcol, Mval, d)

    rp, col, val, u = solve(theta)
    rhs = L.csr_spmv(Mrp, Mcol, Mval, u - uobs) * (1 - fl)
    lam = L.band_cholesky_solve(rp, col, val, rhs)
    g = L.adjoint_gradient_p1(m, theta, u, lam)
    for _ in range(4):
        d = rng.standard_normal(m.ne)
        eps = 1e-6
        fd = (J(theta + eps * d) - J(theta - eps * d)) / (2 * eps)
        assert abs(fd - g @ d) < 1e-6 * max(1.0, abs(fd))


# ------------------------------------------------------------- error norms
def _interpolant(m, degree, fn):
    """DOF vector of the nodal interpolant of fn (every cell writes the values
    at its evaluation nodes, lagr_fe.h:1115-1133; a consistent DOF map makes
    the writes agree)."""
    cd, nd = m.dofs(degree)
    nodes = np.array([[0, 0], [1, 0], [0, 1]]) if degree == 1 else _p3_nodes()
    u = np.full(nd, np.nan)
    p0, p1, p2 = m.xy[m.cells[:, 0]], m.xy[m.cells[:, 1]], m.xy[m.cells[:, 2]]
    for i, (a, b) in enumerate(nodes):
        x = p0 * (1 - a - b) + p1 * a + p2 * b
        v = fn(x[:, 0], x[:, 1])
        prev = u[cd[:, i]]
        ok = np.isnan(prev) | (np.abs(prev - v) <= 1e-12 * (1 + np.abs(v)))
        assert ok.all(), "DOF map gives one DOF two different nodes"
        u[cd[:, i]] = v
    assert not np.isnan(u).any()
    return u


def _exact_qp(m, degree, fn, grad):
    q = L.qp_coords(m, degree)
    return fn(q[..., 0], q[..., 1]), np.stack(grad(q[..., 0], q[..., 1]), axis=-1)


@pytest.mark.parametrize("degree", [1, 3])
def test_error_norms_vanish_on_the_fe_space(degree):
    # a polynomial of the element's degree is reproduced exactly, so both
    # error integrals vanish (up to rounding); jittered mesh
    m = L.tp_triag_mesh(9)
    rng = np.random.default_rng(2)
    xy = m.xy.copy()
    inner = (xy[:, 0] > 0) & (xy[:, 0] < 1) & (xy[:, 1] > 0) & (xy[:, 1] < 1)
    xy[inner] += rng.uniform(-0.03, 0.03, (inner.sum(), 2))
    m = L.TriMesh(xy, m.cells, m.edges)
    if degree == 1:
        fn = lambda x, y: 2 * x + y - 0.5
        gr = lambda x, y: (2 + 0 * x, 1 + 0 * y)
    else:
        fn = lambda x, y: x ** 3 - 2 * x * y * y + y - 1
        gr = lambda x, y: (3 * x * x - 2 * y * y, -4 * x * y + 1)
    uh = _interpolant(m, degree, fn)
    uq, gq = _exact_qp(m, degree, fn, gr)
    l2, h1 = L.error_norms(m, degree, uh, uq, gq)
    assert l2 < 1e-26 and h1 < 1e-24, (l2, h1)


@pytest.mark.parametrize("degree", [1, 3])
def test_error_norms_of_zero_are_the_norms_of_u(degree):
    # u_h = 0: the integrals are the quadrature of |u|^2 = 1/4 and of
    # |grad u|^2 = pi^2/2 for u = sin(pi x) sin(pi y) on the unit square
    m = L.tp_triag_mesh(48)
    cd, nd = m.dofs(degree)
    l2, h1 = L.error_norms(m, degree, np.zeros(nd))
    tol = 2e-4 if degree == 1 else 1e-10
    assert abs(l2 - 0.25) < tol and abs(h1 - np.pi ** 2 / 2) < 20 * tol, (l2, h1)


def test_error_norms_convergence_rates():
    # homDir_linfe_demo.cc:114-188: -Laplace u = 2 pi^2 sin sin, P1, L2 error
    # O(h^2) and H1-seminorm error O(h)
    errs = []
    for n in (16, 32, 64):
        m = L.tp_triag_mesh(n)
        q = L.qp_coords(m, 1)
        f = 2 * np.pi ** 2 * np.sin(np.pi * q[..., 0]) * np.sin(np.pi * q[..., 1])
        rp, col, val, b = L.galerkin_system(m, 1, 1.0, 0.0, L.F_QP, f_qp=f,
                                            flags=m.dirichlet_flags(1))
        u = L.band_cholesky_solve(rp, col, val, b)
        errs.append(np.sqrt(L.error_norms(m, 1, u)))
    e = np.array(errs)
    r = e[:-1] / e[1:]
    assert np.all((r[:, 0] > 3.7) & (r[:, 0] < 4.3)), r
    assert np.all((r[:, 1] > 1.9) & (r[:, 1] < 2.1)), r


# ------------------------------------------------------ boundary edge terms
SEG_RULES = {  # correctly rounded Gauss-Legendre points / weights on [0, 1]
    2: ([0.2113248654051871, 0.7886751345948129], [0.5, 0.5]),
    4: ([0.06943184420297371, 0.33000947820757187, 0.6699905217924281, 0.9305681557970263],
        [0.17392742256872692, 0.32607257743127305, 0.32607257743127305, 0.17392742256872692]),
}


def test_segment_rule_closed_forms():
    # quad/test/gauss_legendre_tests.cc:14-57 (EXPECT_DOUBLE_EQ = 4 ulps)
    def close(a, b):
        return abs(a - b) <= 4 * np.spacing(max(abs(a), abs(b)))
    p, w = L.segment_rule(1)
    assert p[0] == 0.5 and w[0] == 1.0
    p, w = L.segment_rule(2)
    assert close(p[0], (1 - np.sqrt(1 / 3)) / 2) and close(p[1], (1 + np.sqrt(1 / 3)) / 2)
    assert close(w[0], 0.5) and close(w[1], 0.5)
    p, w = L.segment_rule(3)
    assert close(p[0], (1 - np.sqrt(3 / 5)) / 2) and p[1] == 0.5
    assert close(w[0], 5 / 18) and close(w[1], 4 / 9)
    p, w = L.segment_rule(4)
    assert close(p[0], (1 - np.sqrt(3 / 7 + 2 / 7 * np.sqrt(6 / 5))) / 2)
    assert close(p[1], (1 - np.sqrt(3 / 7 - 2 / 7 * np.sqrt(6 / 5))) / 2)
    assert close(p[2], 1 - p[1]) and close(p[3], 1 - p[0])
    assert close(w[0], (18 - np.sqrt(30)) / 72) and close(w[1], (18 + np.sqrt(30)) / 72)
    # the device tables hold the correctly rounded values: within 1 ulp here
    for n, (pp, ww) in SEG_RULES.items():
        p, w = L.segment_rule(n)
        assert np.all(np.abs(p - pp) <= np.spacing(np.abs(p))), (n, p - pp)
        assert np.all(np.abs(w - ww) <= np.spacing(np.abs(w))), (n, w - ww)
    # exactness: degree 2n - 1
    for n in (1, 2, 3, 4):
        p, w = L.segment_rule(n)
        for k in range(2 * n):
            assert abs(np.dot(w, p ** k) - 1 / (k + 1)) < 1e-15


@pytest.mark.parametrize("degree", [1, 3])
def test_edge_mass_and_load_closed_forms(degree):
    rng = np.random.default_rng(4)
    xy = rng.uniform(0, 1, (6, 2))
    ev = np.array([[0, 1], [2, 3], [4, 5], [1, 0]], np.int32)
    L_e = np.linalg.norm(xy[ev[:, 1]] - xy[ev[:, 0]], axis=1)
    M = L.edge_matrices(xy, ev, degree, L.G_CONST, 1.0)
    v = L.edge_vectors(xy, ev, degree, L.G_CONST, 1.0)
    # integral of N_i: L/2 (P1); Simpson 3/8: L (1/8, 1/8, 3/8, 3/8) (P3)
    wts = np.array([0.5, 0.5]) if degree == 1 else np.array([1, 1, 3, 3]) / 8
    assert np.allclose(v, L_e[:, None] * wts, rtol=1e-14, atol=0)
    assert np.allclose(M.sum(axis=2), L_e[:, None] * wts, rtol=1e-13, atol=0)
    assert np.allclose(M, np.transpose(M, (0, 2, 1)), rtol=0, atol=1e-15)
    if degree == 1:
        ref = np.array([[2, 1], [1, 2]]) / 6
        assert np.allclose(M, L_e[:, None, None] * ref, rtol=1e-14, atol=0)
    # per-edge and per-point coefficients scale linearly
    c = rng.uniform(0.5, 2, 4)
    assert np.allclose(L.edge_matrices(xy, ev, degree, L.G_EDGE, 2.0, c),
                       2.0 * c[:, None, None] * M, rtol=1e-14, atol=0)
    nq = degree + 1
    cq = np.repeat(c[:, None], nq, axis=1)
    assert np.allclose(L.edge_vectors(xy, ev, degree, L.G_QP, 1.0, cq), c[:, None] * v,
                       rtol=1e-14, atol=0)


def _edge_qp_xy(m, sel, degree):
    p, _ = L.segment_rule(degree + 1)
    a, b = m.xy[m.edges[sel, 0]], m.xy[m.edges[sel, 1]]
    return (b[:, None, :] * p[None, :, None] + a[:, None, :] * (1 - p)[None, :, None])


@pytest.mark.parametrize("degree", [1, 3])
def test_robin_problem_reproduces_linear_solution(degree):
    # -Laplace u + u = f in the unit square, du/dn + u = h on the boundary
    # (ellbvp_linfe_demo.cc:272-315 with impedance edges everywhere, no
    # Dirichlet part); u = 1 + x + 2y lies in the FE space, all data are
    # integrated exactly, so the Galerkin solution is its interpolant
    n = 8
    m = L.tp_triag_mesh(n)
    cd, nd = m.dofs(degree)
    q = L.qp_coords(m, degree)
    u = lambda x, y: 1 + x + 2 * y
    f = u(q[..., 0], q[..., 1])
    _, ef = m.boundary_flags()
    sel = np.nonzero(ef)[0]
    pts = _edge_qp_xy(m, sel, degree)
    mid = 0.5 * (m.xy[m.edges[sel, 0]] + m.xy[m.edges[sel, 1]])
    dudn = np.select([mid[:, 0] < 1e-12, mid[:, 0] > 1 - 1e-12, mid[:, 1] < 1e-12],
                     [-1.0, 1.0, -2.0], 2.0)
    h = np.zeros((m.nedges, degree + 1))
    h[sel] = dudn[:, None] + u(pts[..., 0], pts[..., 1])
    rp, col, val, b = L.galerkin_system(
        m, degree, 1.0, 1.0, L.F_QP, f_qp=f, flags=np.zeros(nd, np.uint8),
        edge_terms=(ef, (L.G_CONST, 1.0, None), ef, (L.G_QP, 1.0, h)))
    x = L.band_cholesky_solve(rp, col, val, b)
    xi = _interpolant(m, degree, u)
    assert np.abs(x - xi).max() < 1e-11, np.abs(x - xi).max()


# -- Fisher/KPP Strang splitting ------------------------------------------------
def _fkpp_scipy(n, c=1.2, h=1.0):
    import scipy.sparse as sp
    m = L.tp_triag_mesh(n)
    (rA, cA, vA), (rM, cM, vM) = L.fisher_kpp_matrices(m, c, (L.G_CONST, h, None))
    A = sp.csr_matrix((vA, cA, rA), shape=(m.nv, m.nv))
    M = sp.csr_matrix((vM, cM, rM), shape=(m.nv, m.nv))
    return m, A, M


def test_fisher_kpp_matrices_structure():
    # A: P1 stiffness rows sum to zero, so A 1 is the Robin edge mass
    # applied to ones = h * (edge lengths shared by the node) on boundary
    # nodes; M 1 sums to the area; both symmetric on one pattern
    m, A, M = _fkpp_scipy(6, c=1.2, h=0.5)
    nf, _ = m.boundary_flags()
    a1 = A @ np.ones(m.nv)
    assert np.abs(a1[nf == 0]).max() < 1e-13
    corner = np.array([0, 6, 42, 48])
    expect = np.where(nf == 1, 0.5 * (1 / 6), 0.0)
    expect[corner] = 0.5 * (1 / 6)  # two half-edges of length 1/6 each
    assert np.allclose(a1, expect, atol=1e-14)
    assert abs((M @ np.ones(m.nv)).sum() - 1.0) < 1e-14
    assert abs(A - A.T).max() < 1e-15 and abs(M - M.T).max() < 1e-15
    assert np.array_equal(A.indptr, M.indptr) and np.array_equal(A.indices, M.indices)


def test_strang_diffusion_on_generalised_eigenvectors():
    # A v = mu M v: one diffusion step multiplies v by the SDIRK-2 stability
    # function R(-h mu), h = tau / 2 on the first call and tau afterwards;
    # R is computed from the Butcher tableau, independently of the stepper
    import scipy.linalg as sl
    _, A, M = _fkpp_scipy(5, c=1.2, h=1.0)
    T, nsteps = 0.7, 9
    ss = L.StrangSplit(A, M, T, nsteps, 2.1)
    w, V = sl.eigh(A.toarray(), M.toarray())
    tau = T / nsteps
    for k in (0, 3, len(w) // 2, len(w) - 1):
        v = V[:, k]
        for first, h in ((True, 0.5 * tau), (False, tau)):
            got = ss.diffusion(first, v)
            want = L.sdirk2_stability(-h * w[k]) * v
            assert np.abs(got - want).max() < 1e-11 * np.abs(v).max(), (k, first)


def test_sdirk2_stability_known_values():
    # R(0) = 1; second order: R(z) = 1 + z + z^2/2 + O(z^3); L-stable: R -> 0
    k = L.SDIRK2_KAPPA
    assert abs(k - (1 - 1 / math.sqrt(2))) < 3e-16
    assert L.sdirk2_stability(0.0) == 1.0
    z = 1e-3
    assert abs(L.sdirk2_stability(z) - (1 + z + z * z / 2)) < 1e-9
    assert abs(L.sdirk2_stability(-1e8)) < 1e-7


@pytest.mark.parametrize("nsteps", [1, 2, 3, 10])
def test_strang_reaction_only_is_logistic_flow(nsteps):
    # with A = 0 the diffusion steps are the identity and the splitting
    # composes exact logistic flows: u(t) = K / (1 + (K/u0 - 1) e^{-lambda t})
    # over t = max(m - 2, 0) tau + tau / 2 (the reference's loop runs
    # i = 1 .. m-2)
    import scipy.sparse as sp
    _, A, M = _fkpp_scipy(4)
    Z = sp.csr_matrix((np.zeros_like(A.data), A.indices, A.indptr), shape=A.shape)
    T, lam = 1.3, 2.1
    ss = L.StrangSplit(Z, M, T, nsteps, lam)
    rng = np.random.default_rng(5)
    K = 0.5 + rng.random(A.shape[0])
    u0 = 0.01 + rng.random(A.shape[0])
    got = ss.evolution(K, u0)
    tau = T / nsteps
    t = max(nsteps - 2, 0) * tau + tau / 2
    want = K / (1 + (K / u0 - 1) * math.exp(-lam * t))
    assert np.abs(got - want).max() < 1e-14


def test_strang_conserves_mass_without_reaction_and_robin_term():
    # pure Neumann diffusion (h = 0), lambda = 0: 1^T M u is invariant since
    # 1^T A = 0, and the reaction with lambda = 0 is the identity
    _, A, M = _fkpp_scipy(6, c=1.2, h=0.0)
    ss = L.StrangSplit(A, M, 1.0, 12, 0.0)
    rng = np.random.default_rng(1)
    u0 = rng.random(A.shape[0])
    u = ss.evolution(np.ones(A.shape[0]), u0)
    assert abs((M @ u).sum() - (M @ u0).sum()) < 1e-13
    assert np.abs(u - u0).max() > 1e-3  # it did diffuse


def test_strang_diffusion_is_second_order_in_time():
    # lambda = 0: evolution = exp(-M^{-1} A t) u0 with t = tau/2 + (m-2) tau,
    # up to the SDIRK-2 error O(tau^2)
    import scipy.linalg as sl
    _, A, M = _fkpp_scipy(5, c=1.2, h=1.0)
    Ad, Md = A.toarray(), M.toarray()
    u0 = np.random.default_rng(2).random(A.shape[0])
    errs = []
    for nsteps in (20, 40):
        T = 0.5
        tau = T / nsteps
        t = tau / 2 + (nsteps - 2) * tau
        exact = sl.expm(-np.linalg.solve(Md, Ad) * t) @ u0
        got = L.StrangSplit(A, M, T, nsteps, 0.0).evolution(np.ones(A.shape[0]), u0)
        errs.append(np.abs(got - exact).max())
    assert 3.3 < errs[0] / errs[1] < 4.7, errs


# ----------------------------------------------- Fisher/KPP nonlocal terms
def _nonlocal_case(n=8, jitter=0.0, rmin=0.0, rmax=0.0):
    m = L.tp_triag_mesh(n)
    if jitter:
        rng = np.random.default_rng(7)
        nf, _ = m.boundary_flags()
        xy = m.xy.copy()
        xy[nf == 0] += jitter / n * (rng.random((int((nf == 0).sum()), 2)) - 0.5)
        m = L.TriMesh(xy, m.cells)
    _, ef = m.boundary_flags()
    h = L.nonlocal_gain(m.xy, m.edges, 1, ef, rmin, rmax)
    rows, Lm = L.nonlocal_loss(m.xy, m.edges, m.edge_dofs(1), 1, ef, rmin, rmax)
    return m, ef, h, rows, Lm


def _quadrature_matrices(m, ef, rows):
    """The boundary quadrature as matrices: point coordinates Q, weights W
    (w_k |e| on flagged edges, 0 elsewhere) and the P1 basis values Phi[dof
    row, point], built from the Gauss points (3 -+ sqrt 3) / 6 directly."""
    ev, ed = m.edges, m.edge_dofs(1)
    g = np.array([(3 - math.sqrt(3)) / 6, (3 + math.sqrt(3)) / 6])
    pos = {int(r): i for i, r in enumerate(rows)}
    Q = np.zeros((2 * len(ev), 2))
    W = np.zeros(2 * len(ev))
    Phi = np.zeros((len(rows), 2 * len(ev)))
    for e in range(len(ev)):
        a, b = m.xy[ev[e, 0]], m.xy[ev[e, 1]]
        for k in range(2):
            Q[2 * e + k] = a + (b - a) * g[k]
            if ef[e]:
                W[2 * e + k] = 0.5 * np.linalg.norm(b - a)
                Phi[pos[int(ed[e, 0])], 2 * e + k] = 1 - g[k]
                Phi[pos[int(ed[e, 1])], 2 * e + k] = g[k]
    return Q, W, Phi


@pytest.mark.parametrize("radii", [(0.0, 0.0), (0.3, 0.8)])
def test_nonlocal_loss_is_the_double_boundary_quadrature(radii):
    # L_ij = sum_kl phi_i(x_k) w_k g(x_k, y_l) w_l phi_j(y_l) over the
    # boundary points: the nested AssembleVectorLocally calls of the drivers
    # (modelproblem_island_main.cc:92-126) written as one matrix product
    # from independently built quadrature matrices; a transposed index, a
    # wrong weight or edge length, or a point on the wrong edge fails this
    m, ef, h, rows, Lm = _nonlocal_case(8, jitter=0.5, rmin=radii[0], rmax=radii[1])
    Q, W, Phi = _quadrature_matrices(m, ef, rows)
    D2 = ((Q[:, None, :] - Q[None, :, :]) ** 2).sum(-1)
    G = 1.0 / (1.0 + D2)
    if radii[1] > 0:
        D = np.sqrt(D2)
        G = np.where((D >= radii[0]) & (D <= radii[1]), G, 0.0)
    ref = (Phi * W) @ G @ (Phi * W).T
    assert np.abs(ref - Lm).max() <= 1e-14 * np.abs(Lm).max()
    # the gain at the points: h(x_k) = sum_l w_l g(x_k, y_l)
    hq = G @ W
    flagged = np.repeat(ef != 0, 2)
    assert np.abs(hq[flagged] - h.ravel()[flagged]).max() <= 1e-14 * np.abs(hq).max()
    assert np.all(h[ef == 0] == 0.0)
    # the boundary DOFs are the vertices of the boundary edges
    assert np.array_equal(rows, np.unique(m.edge_dofs(1)[ef != 0]))


def test_nonlocal_loss_rows_sum_to_the_load_of_the_gain():
    # sum_j phi_j = 1 on the boundary, so L 1 is the boundary load of h
    # (ScalarLoadEdgeVectorProvider with f = h), and L is symmetric as g is:
    # the loss and the gain of the model balance
    m, ef, h, rows, Lm = _nonlocal_case(10, jitter=0.6)
    ev_h = L.edge_vectors(m.xy, m.edges, 1, kind=L.G_QP, const=1.0, data=h.ravel())
    bh = np.zeros(m.nv)
    ed = m.edge_dofs(1)
    for e in np.nonzero(ef)[0]:
        bh[ed[e]] += ev_h[e]
    assert np.abs(Lm.sum(1) - bh[rows]).max() <= 1e-14 * np.abs(bh).max()
    assert np.abs(Lm - Lm.T).max() <= 1e-15 * np.abs(Lm).max()
    assert np.all(Lm > 0)  # g > 0 everywhere without a cut


def test_strang_with_nonlocal_terms_conserves_mass():
    # gain h (edge mass term) and loss L balance: 1^T (A - L) = 0 up to
    # rounding, so pure diffusion keeps 1^T M u (lambda = 0)
    import scipy.sparse as sp
    m, ef, h, rows, Lm = _nonlocal_case(8)
    (rA, cA, vA), (rM, cM, vM) = L.fisher_kpp_matrices(m, 1.2, (L.G_QP, 1.0, h))
    A = sp.csr_matrix((vA, cA, rA), shape=(m.nv, m.nv)).toarray()
    A[np.ix_(rows, rows)] -= Lm
    A = sp.csr_matrix(A)
    M = sp.csr_matrix((vM, cM, rM), shape=(m.nv, m.nv))
    assert np.abs(np.ones(m.nv) @ A).max() < 1e-13
    ss = L.StrangSplit(A, M, 1.0, 10, 0.0)
    u0 = np.random.default_rng(3).random(m.nv)
    u = u0
    for k in range(6):
        u = ss.diffusion(k == 0, u)
    assert abs((M @ u).sum() - (M @ u0).sum()) < 1e-13
    assert np.abs(u - u0).max() > 1e-3
