"""Parity of the HIP path (through the C ABI) with the CPU oracle.

Bars: integer/index data (patterns, DOF maps, row pointers) bit-exact;
element-level FP64 values bit-exact where both sides evaluate the same
expression sequence (the kernels are built with -ffp-contract=off), and
otherwise within the tolerance written in each test (1e-12 relative to the
row's infinity norm for assembled matrices, as the element couplings across
right angles are cancellation-prone near-zeros).
"""
import math

import numpy as np
import pytest

from lehrfempp_amd import capi
from oracle import lfref as L

pytestmark = pytest.mark.gpu


def jittered_mesh(n, amount=0.2, seed=11):
    m = L.tp_triag_mesh(n)
    xy = m.xy.copy()
    rng = np.random.default_rng(seed)
    h = 1.0 / n
    inner = (xy[:, 0] > 0) & (xy[:, 0] < 1) & (xy[:, 1] > 0) & (xy[:, 1] < 1)
    xy[inner] += rng.uniform(-amount * h, amount * h, size=(inner.sum(), 2))
    return L.TriMesh(xy, m.cells, m.edges)


def gpu_mesh(ctx, m: L.TriMesh, degree):
    if degree == 1:
        return capi.Mesh(ctx, m.xy, m.cells)
    cd, nd = m.dofs(degree)
    return capi.Mesh(ctx, m.xy, m.cells, cd, nd)


def csr_to_dense_rows(rp, col, val):
    return {(r, int(c)): v for r in range(len(rp) - 1)
            for c, v in zip(col[rp[r]:rp[r + 1]], val[rp[r]:rp[r + 1]])}


def row_scaled_err(rp, val_a, val_b):
    err = 0.0
    for r in range(len(rp) - 1):
        a = val_a[rp[r]:rp[r + 1]]
        b = val_b[rp[r]:rp[r + 1]]
        if len(a):
            scale = max(np.abs(b).max(), 1e-300)
            err = max(err, np.abs(a - b).max() / scale)
    return err


# ------------------------------------------------------------- element level
def test_geometry_bitexact(ctx):
    m = jittered_mesh(16)
    g = gpu_mesh(ctx, m, 1)
    J, Jit, det = capi.geometry(ctx, g)
    Jo, Jito, deto = L.geometry(m)
    assert np.array_equal(J, Jo) and np.array_equal(Jit, Jito) and np.array_equal(det, deto)
    g.close()


@pytest.mark.parametrize("degree", [1, 3])
@pytest.mark.parametrize("kind", ["const", "cell", "qp"])
def test_element_matrices_bitexact(ctx, degree, kind):
    m = jittered_mesh(8 if degree == 3 else 16)
    g = gpu_mesh(ctx, m, degree)
    fe = capi.FE(ctx, degree)
    rng = np.random.default_rng(3)
    q = fe.nqp
    if kind == "const":
        a_h, c_h = 1.7, 0.3
        ca, cg = capi.coeff(capi.COEFF_CONST, a_h), capi.coeff(capi.COEFF_CONST, c_h)
        vecs = []
    elif kind == "cell":
        a_h = np.exp(rng.standard_normal(m.ne))
        c_h = rng.uniform(0, 2, m.ne)
        va, vc = capi.Vec(ctx, a_h), capi.Vec(ctx, c_h)
        ca, cg = capi.coeff(capi.COEFF_CELL, 1.0, va), capi.coeff(capi.COEFF_CELL, 1.0, vc)
        vecs = [va, vc]
    else:
        a_h = rng.uniform(0.5, 2, (m.ne, q))
        c_h = rng.uniform(0, 1, (m.ne, q))
        va, vc = capi.Vec(ctx, a_h), capi.Vec(ctx, c_h)
        ca, cg = capi.coeff(capi.COEFF_QP, 1.0, va), capi.coeff(capi.COEFF_QP, 1.0, vc)
        vecs = [va, vc]
    Kg = capi.element_matrices(ctx, g, fe, ca, cg)
    Ko = L.element_matrices(m, degree, a_h, c_h)
    assert np.array_equal(Kg, Ko), np.abs(Kg - Ko).max()
    for v in vecs:
        v.close()
    fe.close()
    g.close()


@pytest.mark.parametrize("degree", [1, 3])
def test_element_vectors(ctx, degree):
    m = jittered_mesh(8)
    g = gpu_mesh(ctx, m, degree)
    fe = capi.FE(ctx, degree)
    v = capi.element_vectors(ctx, g, fe, capi.coeff(capi.COEFF_CONST, 2.5))
    assert np.array_equal(v, L.element_vectors(m, degree, L.F_CONST, 2.5))
    vs = capi.element_vectors(ctx, g, fe, capi.coeff(capi.COEFF_SINSIN, 1.0))
    vo = L.element_vectors(m, degree, L.F_SINSIN)
    # device sin vs glibc sin may differ in the last bit
    assert np.abs(vs - vo).max() <= 1e-15 * np.abs(vo).max()
    fe.close()
    g.close()


# ---------------------------------------------------------------- assembly
def _assemble_both(ctx, m, degree, alpha_cell, gamma, f_kind, fixed_vals=None):
    cd, nd = m.dofs(degree)
    flags = m.dirichlet_flags(degree)
    g = gpu_mesh(ctx, m, degree)
    fe = capi.FE(ctx, degree)
    plan = capi.Plan(ctx, g, flags)
    va = capi.Vec(ctx, alpha_cell)
    rhs = capi.Vec(ctx, nd)
    fv = capi.Vec(ctx, fixed_vals) if fixed_vals is not None else None
    f = capi.coeff(capi.COEFF_SINSIN if f_kind == L.F_SINSIN else capi.COEFF_CONST, 1.0)
    plan.assemble(fe, capi.coeff(capi.COEFF_CELL, 1.0, va), capi.coeff(capi.COEFF_CONST, gamma),
                  f, rhs, fv)
    rp, col, val = plan.csr.download()
    b = rhs.numpy()
    ro = L.galerkin_system(m, degree, alpha_cell, gamma, f_kind, 1.0, flags=flags,
                           fixed_vals=fixed_vals)
    for h in (va, rhs, fv, plan, fe, g):
        if h is not None:
            h.close()
    return (rp, col, val, b), ro


@pytest.mark.parametrize("degree,n", [(1, 64), (1, 37), (3, 8), (3, 13)])
def test_galerkin_csr_matches_oracle(ctx, degree, n):
    m = jittered_mesh(n) if degree == 3 else L.tp_triag_mesh(n)
    rng = np.random.default_rng(n)
    alpha = np.exp(rng.standard_normal(m.ne))
    (rp, col, val, b), (rpo, colo, valo, bo) = _assemble_both(ctx, m, degree, alpha, 1.0,
                                                              L.F_CONST)
    assert np.array_equal(rp, rpo) and np.array_equal(col, colo)  # pattern bit-exact
    # every entry is summed in triplet order from bit-identical element values
    assert np.array_equal(val, valo), row_scaled_err(rp, val, valo)
    assert np.array_equal(b, bo)


@pytest.mark.parametrize("degree,n", [(1, 1), (1, 2), (1, 3), (3, 1), (3, 2)])
def test_smallest_meshes(ctx, degree, n):
    # the smallest structured meshes: every row fixed (P1, n = 1) or a handful
    # of free rows, fewer rows than one workgroup and than the DIA
    # detection's minimum; nonzero prescribed values make every system
    # nontrivial.  Assembly bit-exact; the PCG solution equals the direct
    # solve and the prescribed values
    m = L.tp_triag_mesh(n)
    rng = np.random.default_rng(40 + n)
    alpha = np.exp(rng.standard_normal(m.ne))
    _, nd = m.dofs(degree)
    fixed_vals = rng.standard_normal(nd)
    (rp, col, val, b), (rpo, colo, valo, bo) = _assemble_both(ctx, m, degree, alpha, 1.0,
                                                              L.F_CONST, fixed_vals=fixed_vals)
    assert np.array_equal(rp, rpo) and np.array_equal(col, colo)
    assert np.array_equal(val, valo) and np.array_equal(b, bo)
    rows = np.repeat(np.arange(nd), np.diff(rp))
    A = capi.coo_to_csr(ctx, rows, col, val, nd, nd)
    vb, x = capi.Vec(ctx, b), capi.Vec(ctx, nd)
    info = capi.pcg(ctx, A, vb, x, rtol=1e-12)
    assert info.converged
    xd = L.band_cholesky_solve(rp, col, val, b)
    assert np.linalg.norm(x.numpy() - xd) <= 1e-10 * np.linalg.norm(xd)
    fl = m.dirichlet_flags(degree) != 0
    assert np.allclose(x.numpy()[fl], fixed_vals[fl], rtol=1e-10, atol=0)
    for h in (x, vb, A):
        h.close()


def test_galerkin_sinsin_rhs_tolerance(ctx):
    m = L.tp_triag_mesh(48)
    alpha = np.exp(np.random.default_rng(0).standard_normal(m.ne))
    (rp, col, val, b), (rpo, colo, valo, bo) = _assemble_both(ctx, m, 1, alpha, 1.0, L.F_SINSIN)
    assert np.array_equal(val, valo)
    assert np.abs(b - bo).max() <= 1e-15 * np.abs(bo).max() * 8


def test_nonzero_dirichlet_values(ctx):
    # FixFlaggedSolutionComponents with prescribed values (fix_dof.h:97-114)
    m = L.tp_triag_mesh(20)
    rng = np.random.default_rng(9)
    alpha = np.exp(rng.standard_normal(m.ne))
    fl = m.dirichlet_flags(1)
    gvals = rng.standard_normal(m.nv) * fl
    (rp, col, val, b), (rpo, colo, valo, bo) = _assemble_both(ctx, m, 1, alpha, 0.5, L.F_CONST,
                                                              fixed_vals=gvals)
    assert np.array_equal(col, colo) and np.array_equal(val, valo)
    assert np.array_equal(b, bo), np.abs(b - bo).max()


def test_assembly_is_deterministic(ctx):
    m = L.tp_triag_mesh(96)
    g = gpu_mesh(ctx, m, 1)
    fe = capi.FE(ctx, 1)
    plan = capi.Plan(ctx, g, m.dirichlet_flags(1))
    va = capi.Vec(ctx, np.exp(np.random.default_rng(1).standard_normal(m.ne)))
    rhs = capi.Vec(ctx, m.nv)
    outs = []
    for _ in range(3):
        plan.assemble(fe, capi.coeff(capi.COEFF_CELL, 1.0, va), capi.coeff(capi.COEFF_CONST, 1.0),
                      capi.coeff(capi.COEFF_SINSIN, 1.0), rhs)
        outs.append((plan.csr.download()[2].copy(), rhs.numpy()))
    for o in outs[1:]:
        assert np.array_equal(o[0], outs[0][0]) and np.array_equal(o[1], outs[0][1])
    for h in (va, rhs, plan, fe, g):
        h.close()


# ------------------------------------------------------------ COO -> CSR
def test_coo_to_csr_golden(ctx, golden):
    g = golden["test_mesh0"]
    rows, cols, vals = [], [], []
    for c, cell in enumerate(g["cells"]):
        for i in cell:
            for j in cell:
                rows.append(i)
                cols.append(j)
                vals.append(float(c) if i == j else -float(c))
    A = capi.coo_to_csr(ctx, rows, cols, vals, 10, 10)
    rp, col, val = A.download()
    dense = np.zeros((10, 10))
    for r in range(10):
        dense[r, col[rp[r]:rp[r + 1]]] = val[rp[r]:rp[r + 1]]
    ref = np.array(golden["test_assembler_matrix"]["matrix"], dtype=float).reshape(10, 10)
    assert np.array_equal(dense, ref)
    rpo, colo, valo = L.set_from_triplets(rows, cols, vals, 10)
    assert np.array_equal(rp, rpo) and np.array_equal(col, colo) and np.array_equal(val, valo)
    A.close()


def test_coo_to_csr_insertion_order_and_random(ctx):
    A = capi.coo_to_csr(ctx, [0, 0, 0], [1, 1, 1], [1e16, 1.0, -1e16], 2, 2)
    assert A.download()[2][0] == (1e16 + 1.0) + (-1e16)
    A.close()
    rng = np.random.default_rng(2)
    nnz = 200000
    r = rng.integers(0, 3000, nnz).astype(np.int32)
    c = rng.integers(0, 2500, nnz).astype(np.int32)
    v = rng.standard_normal(nnz)
    A = capi.coo_to_csr(ctx, r, c, v, 3000, 2500)
    got = A.download()
    exp = L.set_from_triplets(r, c, v, 3000)
    for a, b in zip(got, exp):
        assert np.array_equal(a, b)
    A.close()
    with pytest.raises(capi.LfgpuError) as e:
        capi.coo_to_csr(ctx, [0, 5], [0, 0], [1.0, 1.0], 3, 3)
    assert e.value.code == capi.ERANGE


# ---------------------------------------------------------- SpMV and PCG
def _c1_system(ctx, n=64, alpha=None, gamma=0.0, f_kind=L.F_CONST):
    m = L.tp_triag_mesh(n)
    if alpha is None:
        alpha = np.ones(m.ne)
    g = gpu_mesh(ctx, m, 1)
    fe = capi.FE(ctx, 1)
    plan = capi.Plan(ctx, g, m.dirichlet_flags(1))
    va = capi.Vec(ctx, alpha)
    b = capi.Vec(ctx, m.nv)
    plan.assemble(fe, capi.coeff(capi.COEFF_CELL, 1.0, va), capi.coeff(capi.COEFF_CONST, gamma),
                  capi.coeff(capi.COEFF_SINSIN if f_kind == L.F_SINSIN else capi.COEFF_CONST, 1.0),
                  b)
    return m, g, fe, plan, va, b


def test_spmv_matches_oracle(ctx):
    m, g, fe, plan, va, b = _c1_system(ctx, 80, np.exp(np.random.default_rng(4).standard_normal(2 * 80 * 80)))
    rp, col, val = plan.csr.download()
    x = np.random.default_rng(5).standard_normal(m.nv)
    xv, yv = capi.Vec(ctx, x), capi.Vec(ctx, m.nv)
    capi.spmv(ctx, plan.csr, xv, yv)
    y = yv.numpy()
    yo = L.csr_spmv(rp, col, val, x)
    assert np.abs(y - yo).max() <= 1e-14 * np.abs(yo).max()
    for h in (xv, yv, va, b, plan, fe, g):
        h.close()


def test_pcg_fp64_c1_against_direct_solve(ctx):
    # -Laplace u = 1, 64x64, zero Dirichlet; SparseLU's role is
    # played by the oracle's banded Cholesky
    m, g, fe, plan, va, b = _c1_system(ctx)
    rp, col, val = plan.csr.download()
    bh = b.numpy()
    x_direct = L.band_cholesky_solve(rp, col, val, bh)
    x = capi.Vec(ctx, m.nv)
    info = capi.pcg(ctx, plan.csr, b, x, rtol=1e-10)
    xo, ito, erro = L.pcg(rp, col, val, bh, tol=1e-10)
    assert info.converged and info.relres < 1e-10
    assert abs(info.iters - ito) <= 2  # same algorithm, different summation order
    xg = x.numpy()
    assert np.linalg.norm(xg - x_direct) <= 1e-8 * np.linalg.norm(x_direct)
    # tighter solve reaches the direct solution to ~1e-12
    x.fill(0.0)
    info = capi.pcg(ctx, plan.csr, b, x, rtol=1e-14)
    assert np.linalg.norm(x.numpy() - x_direct) <= 1e-12 * np.linalg.norm(x_direct)
    for h in (x, va, b, plan, fe, g):
        h.close()


@pytest.mark.parametrize("mode", [capi.PCG_MIXED, capi.PCG_MIXED_MATRIX])
def test_pcg_mixed_c1_against_direct_solve(ctx, mode):
    alpha = np.exp(np.random.default_rng(8).standard_normal(2 * 64 * 64))
    m, g, fe, plan, va, b = _c1_system(ctx, 64, alpha, 1.0, L.F_SINSIN)
    rp, col, val = plan.csr.download()
    x_direct = L.band_cholesky_solve(rp, col, val, b.numpy())
    x = capi.Vec(ctx, m.nv)
    info = capi.pcg(ctx, plan.csr, b, x, rtol=1e-10, mode=mode)
    assert info.converged and info.relres < 1e-10 and info.refinements >= 1
    xg = x.numpy()
    r = b.numpy() - L.csr_spmv(rp, col, val, xg)  # FP64 residual certificate
    assert np.linalg.norm(r) < 1e-10 * np.linalg.norm(b.numpy())
    assert np.linalg.norm(xg - x_direct) <= 1e-7 * np.linalg.norm(x_direct)
    for h in (x, va, b, plan, fe, g):
        h.close()


def test_pcg_error_paths(ctx):
    m, g, fe, plan, va, b = _c1_system(ctx, 32)
    x = capi.Vec(ctx, m.nv)
    info = capi.pcg(ctx, plan.csr, b, x, rtol=1e-10, maxit=3, raise_on_noconv=False)
    assert not info.converged and info.iters == 3
    with pytest.raises(capi.LfgpuError) as e:
        capi.pcg(ctx, plan.csr, b, x, rtol=1e-10, maxit=3)
    assert e.value.code == capi.ENOCONV
    # zero rhs gives x = 0 in zero iterations (Eigen semantics)
    z = capi.Vec(ctx, m.nv)
    x.fill(1.0)
    info = capi.pcg(ctx, plan.csr, z, x, rtol=1e-10)
    assert info.iters == 0 and np.all(x.numpy() == 0.0)
    for h in (z, x, va, b, plan, fe, g):
        h.close()


def test_error_codes(ctx):
    with pytest.raises(capi.LfgpuError) as e:
        capi.FE(ctx, 2)
    assert e.value.code == capi.ENORULE
    with pytest.raises(capi.LfgpuError) as e:
        capi.Mesh(ctx, np.array([[0, 0], [1, 1], [2, 2.0]]), np.array([[0, 1, 2]]))
    assert e.value.code == capi.EDEGENERATE
    with pytest.raises(capi.LfgpuError) as e:
        capi.Mesh(ctx, np.array([[0, 0], [1, 0], [0, 1.0]]), np.array([[0, 1, 3]]))
    assert e.value.code == capi.EDEGENERATE


# ------------------------------------------------------------------ adjoint
def test_adjoint_gradient_kernel_matches_oracle(ctx):
    m = jittered_mesh(24)
    rng = np.random.default_rng(12)
    theta = rng.standard_normal(m.ne)
    u = rng.standard_normal(m.nv)
    lam = rng.standard_normal(m.nv)
    g = gpu_mesh(ctx, m, 1)
    fe = capi.FE(ctx, 1)
    vt, vu, vl, vg = (capi.Vec(ctx, theta), capi.Vec(ctx, u), capi.Vec(ctx, lam),
                      capi.Vec(ctx, m.ne))
    capi.adjoint_gradient(ctx, g, fe, vt, vu, vl, vg)
    go = L.adjoint_gradient_p1(m, theta, u, lam)
    # exp on the device may differ from glibc's in the last bit
    assert np.abs(vg.numpy() - go).max() <= 4e-16 * np.abs(go).max() * 4
    for h in (vt, vu, vl, vg, fe, g):
        h.close()


# ------------------------------------------- P3 (sliced-ELL solver storage)
def _p3_system(ctx, n=10):
    m = jittered_mesh(n)
    cd, nd = m.dofs(3)
    flags = m.dirichlet_flags(3)
    g = gpu_mesh(ctx, m, 3)
    fe = capi.FE(ctx, 3)
    plan = capi.Plan(ctx, g, flags)
    alpha = np.exp(np.random.default_rng(n).standard_normal(m.ne))
    va = capi.Vec(ctx, alpha)
    b = capi.Vec(ctx, nd)
    plan.assemble(fe, capi.coeff(capi.COEFF_CELL, 1.0, va), capi.coeff(capi.COEFF_CONST, 1.0),
                  capi.coeff(capi.COEFF_SINSIN, 1.0), b)
    return m, nd, g, fe, plan, va, b


def test_spmv_p3_matches_oracle(ctx):
    m, nd, g, fe, plan, va, b = _p3_system(ctx)
    rp, col, val = plan.csr.download()
    x = np.random.default_rng(6).standard_normal(nd)
    xv, yv = capi.Vec(ctx, x), capi.Vec(ctx, nd)
    capi.spmv(ctx, plan.csr, xv, yv)
    yo = L.csr_spmv(rp, col, val, x)
    assert np.abs(yv.numpy() - yo).max() <= 1e-13 * np.abs(yo).max()
    for h in (xv, yv, va, b, plan, fe, g):
        h.close()


@pytest.mark.parametrize("mode", [capi.PCG_FP64, capi.PCG_MIXED, capi.PCG_MIXED_MATRIX])
def test_pcg_p3_against_direct_solve(ctx, mode):
    m, nd, g, fe, plan, va, b = _p3_system(ctx)
    rp, col, val = plan.csr.download()
    bh = b.numpy()
    xd = L.band_cholesky_solve(rp, col, val, bh)
    x = capi.Vec(ctx, nd)
    info = capi.pcg(ctx, plan.csr, b, x, rtol=1e-10, mode=mode)
    assert info.converged
    xg = x.numpy()
    # FP64 residual certificate of the returned solution
    assert np.linalg.norm(bh - L.csr_spmv(rp, col, val, xg)) < 2e-10 * np.linalg.norm(bh)
    assert np.linalg.norm(xg - xd) <= 1e-7 * np.linalg.norm(xd)
    for h in (x, va, b, plan, fe, g):
        h.close()


def test_pcg_iteration_count_matches_oracle_c1(ctx):
    # Eigen ConjugateGradient iteration count: the merged two-sweep iteration
    # reproduces the textbook recurrences up to rounding
    m, g, fe, plan, va, b = _c1_system(ctx, 48, np.exp(np.random.default_rng(3).standard_normal(2 * 48 * 48)), 1.0)
    rp, col, val = plan.csr.download()
    x = capi.Vec(ctx, m.nv)
    for tol in (1e-6, 1e-10):
        x.fill(0.0)
        info = capi.pcg(ctx, plan.csr, b, x, rtol=tol)
        _, ito, _ = L.pcg(rp, col, val, b.numpy(), tol=tol)
        assert abs(info.iters - ito) <= 2, (tol, info.iters, ito)
    for h in (x, va, b, plan, fe, g):
        h.close()


@pytest.mark.parametrize("degree", [1, 3])
def test_pcg_single_reduction_matches_eigen_iterates(ctx, degree):
    # one fused sweep per iteration (Chronopoulos-Gear recurrences): the same
    # iterates as Eigen's loop in exact arithmetic, so the same iteration
    # count (+-2) and solution; DIA storage at P1, sliced ELL at P3
    if degree == 1:
        m, g, fe, plan, va, b = _c1_system(
            ctx, 48, np.exp(np.random.default_rng(3).standard_normal(2 * 48 * 48)), 1.0)
        nd = m.nv
    else:
        m, nd, g, fe, plan, va, b = _p3_system(ctx, 9)
    rp, col, val = plan.csr.download()
    bh = b.numpy()
    xd = L.band_cholesky_solve(rp, col, val, bh)
    x = capi.Vec(ctx, nd)
    for tol in (1e-6, 1e-10):
        x.fill(0.0)
        info = capi.pcg(ctx, plan.csr, b, x, rtol=tol, single_reduction=True)
        _, ito, _ = L.pcg(rp, col, val, bh, tol=tol)
        assert info.converged and abs(info.iters - ito) <= 2, (tol, info.iters, ito)
        r = bh - L.csr_spmv(rp, col, val, x.numpy())
        assert np.linalg.norm(r) <= 1.5 * tol * np.linalg.norm(bh)
    assert np.linalg.norm(x.numpy() - xd) <= 1e-7 * np.linalg.norm(xd)
    # iteration cap: the maxit-th update is not applied, as in the two-sweep
    # solver, and both report maxit iterations from the same partial solution
    xs = []
    for single in (False, True):
        x.fill(0.0)
        info = capi.pcg(ctx, plan.csr, b, x, rtol=1e-14, maxit=7, single_reduction=single,
                        raise_on_noconv=False)
        assert not info.converged and info.iters == 7
        xs.append(x.numpy())
    assert np.linalg.norm(xs[0] - xs[1]) <= 1e-12 * np.linalg.norm(xs[0])
    # a zero right-hand side returns x = 0 at once (Eigen)
    b.fill(0.0)
    x.fill(3.0)
    info = capi.pcg(ctx, plan.csr, b, x, rtol=1e-10, single_reduction=True)
    assert info.converged and info.iters == 0 and not x.numpy().any()
    for h in (x, va, b, plan, fe, g):
        h.close()


@pytest.mark.parametrize("degree", [1, 3])
def test_pcg_graph_replay_matches_plain_launches(ctx, degree):
    # FP64 iterations replayed from captured chunks (graph=1) give the
    # iterates of plain launches (graph=-1) bit for bit on every replay, with
    # other device work and synchronisations between replays and three
    # solution vectors rotating through the two capture slots; DIA sweeps at
    # P1, sliced-ELL row sweeps at P3
    if degree == 1:
        m, g, fe, plan, va, b = _c1_system(
            ctx, 48, np.exp(np.random.default_rng(5).standard_normal(2 * 48 * 48)), 1.0)
        nd = m.nv
    else:
        m, nd, g, fe, plan, va, b = _p3_system(ctx, 9)
    x0 = capi.Vec(ctx, nd)
    x0.fill(0.0)
    ref = capi.pcg(ctx, plan.csr, b, x0, rtol=1e-10, check_every=8, graph=-1)
    x_ref = x0.numpy()
    assert ref.converged
    xs = [capi.Vec(ctx, nd) for _ in range(3)]
    y = capi.Vec(ctx, nd)
    for rep in range(3):
        # slot hits (0, 1 alternating) and evictions (2)
        for x in (xs[0], xs[1], xs[0], xs[1], xs[2], xs[0]):
            x.fill(0.0)
            info = capi.pcg(ctx, plan.csr, b, x, rtol=1e-10, check_every=8, graph=1)
            assert info.converged and info.iters == ref.iters, (rep, info.iters, ref.iters)
            assert np.array_equal(x.numpy(), x_ref), rep
            capi.spmv(ctx, plan.csr, x, y)  # eager work between replays
            ctx.sync()
        # a plain solve in between leaves the captured chunks valid
        x0.fill(0.0)
        capi.pcg(ctx, plan.csr, b, x0, rtol=1e-10, check_every=8, graph=-1)
        assert np.array_equal(x0.numpy(), x_ref)
    # new values assembled into the same matrix (same storage): the chunk
    # captured for xs[0] is reused and must read them
    va.upload(np.exp(np.random.default_rng(11).standard_normal(m.ne)))
    plan.assemble(fe, capi.coeff(capi.COEFF_CELL, 1.0, va), capi.coeff(capi.COEFF_CONST, 1.0),
                  capi.coeff(capi.COEFF_CONST if degree == 1 else capi.COEFF_SINSIN, 1.0), b)
    x0.fill(0.0)
    ref2 = capi.pcg(ctx, plan.csr, b, x0, rtol=1e-10, check_every=8, graph=-1)
    xs[0].fill(0.0)
    info = capi.pcg(ctx, plan.csr, b, xs[0], rtol=1e-10, check_every=8, graph=1)
    assert info.iters == ref2.iters and np.array_equal(xs[0].numpy(), x0.numpy())
    assert not np.array_equal(x0.numpy(), x_ref)
    for h in xs + [y, x0, va, b, plan, fe, g]:
        h.close()


@pytest.mark.parametrize("degree", [1, 3])
def test_polling_does_not_change_the_iterates(ctx, degree):
    # The host polls the solver state one step behind the device and queues
    # several chunks per poll once the residual history allows an estimate.
    # Chunk length, steps and the empty chunk after the stop must leave the
    # iterates and the iteration count unchanged in every mode (single and
    # batched), and the number of host waits must fall with the estimate.
    if degree == 1:
        m, g, fe, plan, va, b = _c1_system(
            ctx, 40, np.exp(np.random.default_rng(7).standard_normal(2 * 40 * 40)), 1.0)
        nd = m.nv
    else:
        m, nd, g, fe, plan, va, b = _p3_system(ctx, 8)
    x = capi.Vec(ctx, nd)
    for mode in (capi.PCG_FP64, capi.PCG_MIXED, capi.PCG_MIXED_MATRIX):
        runs = []
        for check, graph in ((2, -1), (8, -1), (64, -1), (8, 1)):
            x.fill(0.0)
            info = capi.pcg(ctx, plan.csr, b, x, rtol=1e-10, mode=mode, check_every=check,
                            graph=graph if mode == capi.PCG_FP64 else 0)
            assert info.converged
            runs.append((info.iters, info.syncs, x.numpy()))
        for it, _, xv in runs[1:]:
            assert it == runs[0][0] and np.array_equal(xv, runs[0][2]), mode
        # two iterations per chunk: far fewer waits than chunks once the
        # history gives an estimate (FP64: every poll; mixed: per refinement)
        assert 1 <= runs[0][1] < (runs[0][0] // 2 + 2), (mode, runs[0][:2])
        if mode == capi.PCG_FP64:
            assert runs[0][1] <= runs[0][0] // 6, runs[0][:2]
    # batched: two copies of the problem, the same iterates as the single solve
    x.fill(0.0)
    capi.pcg(ctx, plan.csr, b, x, rtol=1e-10, check_every=64)
    x_single = x.numpy()
    batch = capi.Batch(ctx, plan.csr, np.array([0, nd], np.int32))
    for check in (2, 64):
        x.fill(0.0)
        summ, infos = batch.pcg(b, x, rtol=1e-10, check_every=check)
        assert summ.converged and summ.syncs >= 1
        assert np.linalg.norm(x.numpy() - x_single) <= 1e-12 * np.linalg.norm(x_single)
        if check == 2:
            assert summ.syncs <= summ.iters // 6, (summ.syncs, summ.iters)
    batch.close()
    for h in (x, va, b, plan, fe, g):
        h.close()


# ------------------------------------------------------------ batched solves
def test_batched_pcg_matches_per_problem_solves(ctx):
    sizes = [9, 17, 5, 31, 12, 24, 40, 8]
    xy, cells, flags, voff, coff = capi.batched_tp_mesh(sizes)
    rng = np.random.default_rng(21)
    alpha = np.exp(rng.standard_normal(cells.shape[0]))
    g = capi.Mesh(ctx, xy, cells)
    fe = capi.FE(ctx, 1)
    plan = capi.Plan(ctx, g, flags)
    va = capi.Vec(ctx, alpha)
    b = capi.Vec(ctx, g.ndofs)
    plan.assemble(fe, capi.coeff(capi.COEFF_CELL, 1.0, va), capi.coeff(capi.COEFF_CONST, 0.0),
                  capi.coeff(capi.COEFF_CONST, 1.0), b)
    batch = capi.Batch(ctx, plan.csr, voff)
    x = capi.Vec(ctx, g.ndofs)
    for rep in range(2):  # the second call replays the captured graph
        x.fill(0.0)
        summ, infos = batch.pcg(b, x, rtol=1e-10)
        assert summ.converged
        xg, bh = x.numpy(), b.numpy()
        for k, n in enumerate(sizes):
            m = L.tp_triag_mesh(n)
            a_k = alpha[coff[k]:coff[k + 1]]
            rp, col, val, bo = L.galerkin_system(m, 1, a_k, 0.0, L.F_CONST, 1.0)
            assert np.array_equal(bh[voff[k]:voff[k + 1]], bo)
            xd = L.band_cholesky_solve(rp, col, val, bo)
            xk = xg[voff[k]:voff[k + 1]]
            assert np.linalg.norm(xk - xd) <= 1e-8 * np.linalg.norm(xd), k
            _, ito, _ = L.pcg(rp, col, val, bo, tol=1e-10)
            assert abs(infos[k].iters - ito) <= 2, (k, infos[k].iters, ito)
    for h in (x, batch, b, va, plan, fe, g):
        h.close()


@pytest.mark.parametrize("mode", [capi.PCG_FP64, capi.PCG_MIXED, capi.PCG_MIXED_MATRIX])
def test_batched_with_trivial_problems(ctx, mode):
    # problems of side 1 (every row fixed, zero right-hand side: Eigen's early
    # exit) and 2 (one free row) next to larger ones: the trivial ones stop at
    # iteration 0 with x = 0, every other one matches its direct solve
    sizes = [1, 16, 2, 33, 1]
    xy, cells, flags, voff, coff = capi.batched_tp_mesh(sizes)
    rng = np.random.default_rng(23)
    alpha = np.exp(rng.standard_normal(cells.shape[0]))
    g = capi.Mesh(ctx, xy, cells)
    fe = capi.FE(ctx, 1)
    plan = capi.Plan(ctx, g, flags)
    va = capi.Vec(ctx, alpha)
    b = capi.Vec(ctx, g.ndofs)
    plan.assemble(fe, capi.coeff(capi.COEFF_CELL, 1.0, va), capi.coeff(capi.COEFF_CONST, 0.0),
                  capi.coeff(capi.COEFF_CONST, 1.0), b)
    batch = capi.Batch(ctx, plan.csr, voff)
    x = capi.Vec(ctx, g.ndofs)
    x.fill(1.0)
    summ, infos = batch.pcg(b, x, rtol=1e-10, mode=mode)
    assert summ.converged
    xg, bh = x.numpy(), b.numpy()
    for k, n in enumerate(sizes):
        m = L.tp_triag_mesh(n)
        rp, col, val, bo = L.galerkin_system(m, 1, alpha[coff[k]:coff[k + 1]], 0.0, L.F_CONST,
                                             1.0)
        xk = xg[voff[k]:voff[k + 1]]
        if n == 1:
            assert not np.any(bo) and infos[k].iters == 0 and not np.any(xk), k
            continue
        xd = L.band_cholesky_solve(rp, col, val, bo)
        assert np.linalg.norm(xk - xd) <= 1e-8 * np.linalg.norm(xd), k
    for h in (x, batch, b, va, plan, fe, g):
        h.close()


@pytest.mark.parametrize("mode", [capi.PCG_MIXED, capi.PCG_MIXED_MATRIX])
def test_batched_mixed_pcg_matches_direct_solves(ctx, mode):
    # per-problem reliable updates: each problem stops on its own FP64 true
    # residual; FP64 mode on the same batch in between must still work
    sizes = [9, 17, 5, 31, 12, 24, 40, 8, 3]
    xy, cells, flags, voff, coff = capi.batched_tp_mesh(sizes)
    rng = np.random.default_rng(22)
    alpha = np.exp(rng.standard_normal(cells.shape[0]))
    g = capi.Mesh(ctx, xy, cells)
    fe = capi.FE(ctx, 1)
    plan = capi.Plan(ctx, g, flags)
    va = capi.Vec(ctx, alpha)
    b = capi.Vec(ctx, g.ndofs)
    plan.assemble(fe, capi.coeff(capi.COEFF_CELL, 1.0, va), capi.coeff(capi.COEFF_CONST, 1.0),
                  capi.coeff(capi.COEFF_CONST, 1.0), b)
    bh = b.numpy()
    # a zero right-hand side for one problem: Eigen returns x = 0 at once
    bh[voff[2]:voff[3]] = 0.0
    b.upload(bh)
    batch = capi.Batch(ctx, plan.csr, voff)
    x = capi.Vec(ctx, g.ndofs)
    rp, col, val = plan.csr.download()
    for rep, md in enumerate((mode, capi.PCG_FP64, mode)):
        x.upload(rng.standard_normal(g.ndofs))  # nonzero initial guesses
        summ, infos = batch.pcg(b, x, rtol=1e-10, mode=md)
        assert summ.converged, rep
        xg = x.numpy()
        for k in range(len(sizes)):
            lo, hi = int(voff[k]), int(voff[k + 1])
            sub = slice(int(rp[lo]), int(rp[hi]))
            rpk = np.ascontiguousarray(rp[lo:hi + 1] - rp[lo], np.int32)
            colk = np.ascontiguousarray(col[sub] - lo, np.int32)
            valk = np.ascontiguousarray(val[sub])
            bk = bh[lo:hi]
            if not bk.any():
                assert not xg[lo:hi].any() and infos[k].iters == 0
                continue
            xd = L.band_cholesky_solve(rpk, colk, valk, bk)
            assert np.linalg.norm(xg[lo:hi] - xd) <= 1e-8 * np.linalg.norm(xd), (rep, k)
            # the stopping test is on the true residual in the mixed modes
            r = bk - L.csr_spmv(rpk, colk, valk, xg[lo:hi])
            if md != capi.PCG_FP64:
                assert np.linalg.norm(r) <= 1.05e-10 * np.linalg.norm(bk), (rep, k)
                assert infos[k].refinements >= 1
    for h in (x, batch, b, va, plan, fe, g):
        h.close()


def _bitwise_symmetric(rp, col, val):
    n = len(rp) - 1
    rows = np.repeat(np.arange(n), np.diff(rp)).astype(np.int64)
    key = rows * n + col
    order = np.argsort(key)
    pos = np.searchsorted(key[order], col.astype(np.int64) * n + rows)
    return bool(np.array_equal(val[order][pos].view(np.int64), val.view(np.int64)))


@pytest.mark.parametrize("jitter", [0.0, 0.2])
def test_pcg_fp64_on_symmetric_and_asymmetric_dia_values(ctx, jitter):
    # a structured P1 matrix is symmetric bit for bit, so the FP64 sweep reads
    # only its upper diagonals; a jittered mesh gives the same DIA pattern
    # with values symmetric only up to rounding, which takes the full read.
    # Both must match the direct solve and Eigen's iteration count.
    n = 48
    m = jittered_mesh(n, amount=jitter, seed=5) if jitter else L.tp_triag_mesh(n)
    a = np.exp(np.random.default_rng(6).standard_normal(m.ne))
    g = gpu_mesh(ctx, m, 1)
    fe = capi.FE(ctx, 1)
    plan = capi.Plan(ctx, g, m.dirichlet_flags(1))
    va = capi.Vec(ctx, a)
    b = capi.Vec(ctx, g.ndofs)
    plan.assemble(fe, capi.coeff(capi.COEFF_CELL, 1.0, va), capi.coeff(capi.COEFF_CONST, 1.0),
                  capi.coeff(capi.COEFF_SINSIN, 1.0), b)
    rp, col, val = plan.csr.download()
    assert _bitwise_symmetric(rp, col, val) == (jitter == 0.0)
    x = capi.Vec(ctx, g.ndofs)
    info = capi.pcg(ctx, plan.csr, b, x, rtol=1e-10)
    assert info.converged
    bh = b.numpy()
    xd = L.band_cholesky_solve(rp, col, val, bh)
    assert np.linalg.norm(x.numpy() - xd) <= 1e-8 * np.linalg.norm(xd)
    _, ito, _ = L.pcg(rp, col, val, bh, tol=1e-10)
    assert abs(info.iters - ito) <= 2, (info.iters, ito)
    # y = A x through the same storage equals the CSR product to rounding
    # of the row's terms (A x ~ b is far smaller than the terms a_ij x_j)
    y = capi.Vec(ctx, g.ndofs)
    capi.spmv(ctx, plan.csr, x, y)
    xh = x.numpy()
    yo = L.csr_spmv(rp, col, val, xh)
    scale = L.csr_spmv(rp, col, np.abs(val), np.abs(xh))
    assert np.all(np.abs(y.numpy() - yo) <= 1e-14 * scale + 1e-300)
    for h in (y, x, b, va, plan, fe, g):
        h.close()


def _ragged_system(kind, n, seed):
    """COO of an SPD test matrix: 'diagonal' (no off-diagonal entries at
    all) or 'arrow' (row and column 0 couple to every row, the other rows
    hold only their diagonal and column 0: one row of n entries next to rows
    of two)."""
    rng = np.random.default_rng(seed)
    d = 1.0 + rng.random(n)
    if kind == "diagonal":
        return np.arange(n), np.arange(n), d
    c = 0.5 * rng.random(n - 1) / np.sqrt(n)
    rows = np.concatenate([np.arange(n), np.zeros(n - 1, int), np.arange(1, n)])
    cols = np.concatenate([np.arange(n), np.arange(1, n), np.zeros(n - 1, int)])
    d[0] += 1.0
    return rows, cols, np.concatenate([d, c, c])


@pytest.mark.parametrize("kind", ["diagonal", "arrow"])
@pytest.mark.parametrize("mode", [capi.PCG_FP64, capi.PCG_MIXED, capi.PCG_MIXED_MATRIX])
def test_ragged_and_diagonal_matrices(ctx, kind, mode):
    # sliced-ELL storage at its extremes: slices with no off-diagonal slot at
    # all, and one slice as wide as the matrix beside slices of width one;
    # product and solve against dense numpy
    n = 300
    r, c, v = _ragged_system(kind, n, 33)
    A = capi.coo_to_csr(ctx, r, c, v, n, n)
    Ad = np.zeros((n, n))
    np.add.at(Ad, (r, c), v)
    x0 = np.random.default_rng(34).standard_normal(n)
    vx, vy = capi.Vec(ctx, x0), capi.Vec(ctx, n)
    capi.spmv(ctx, A, vx, vy)
    assert np.all(np.abs(vy.numpy() - Ad @ x0) <= 1e-14 * (np.abs(Ad) @ np.abs(x0)))
    b = capi.Vec(ctx, np.random.default_rng(35).standard_normal(n))
    x = capi.Vec(ctx, n)
    info = capi.pcg(ctx, A, b, x, rtol=1e-12, mode=mode)
    assert info.converged
    st, _, _ = capi.solver_storage(A)
    assert st in (capi.STORAGE_ELL, capi.STORAGE_ELL16), st
    xd = np.linalg.solve(Ad, b.numpy())
    assert np.linalg.norm(x.numpy() - xd) <= 1e-10 * np.linalg.norm(xd)
    if kind == "diagonal" and mode == capi.PCG_FP64:
        # Jacobi is exact: the first residual check stops the loop before
        # it counts the pass, as Eigen's loop does (the oracle reports 0)
        assert info.iters == 0
    for h in (x, b, vy, vx, A):
        h.close()


def _grid_laplacian(m, seed, zero_pair, lower_zero):
    """COO of an SPD 5-point operator on an m x m grid (edge weights
    exp(N(0,1)), a small shift on the diagonal) plus one explicit zero pair
    (i, i+1) / (i+1, i) between the last node of a grid line and the first
    of the next, stored as +0.0 above the diagonal and `lower_zero` below."""
    rng = np.random.default_rng(seed)
    n = m * m
    rows, cols, vals = [], [], []
    diag = np.full(n, 1e-3)
    for di, dj in ((0, 1), (1, 0)):
        ii, jj = np.meshgrid(np.arange(m - di), np.arange(m - dj), indexing="ij")
        a = (ii * m + jj).ravel()
        b = ((ii + di) * m + jj + dj).ravel()
        w = np.exp(rng.standard_normal(a.size))
        rows += [a, b]
        cols += [b, a]
        vals += [-w, -w]
        np.add.at(diag, a, w)
        np.add.at(diag, b, w)
    i = zero_pair
    rows += [np.arange(n), np.array([i, i + 1])]
    cols += [np.arange(n), np.array([i + 1, i])]
    vals += [diag, np.array([0.0, lower_zero])]
    return np.concatenate(rows), np.concatenate(cols), np.concatenate(vals), n


@pytest.mark.parametrize("mode", [capi.PCG_FP64, capi.PCG_MIXED, capi.PCG_MIXED_MATRIX])
def test_dia_upper_diagonal_reads_give_the_full_read_iterates(ctx, mode):
    # on DIA values that are symmetric bit for bit the FP64 product sweep
    # reads only the upper diagonals (the lower ones mirrored); one explicit
    # zero stored as -0.0 below the diagonal and +0.0 above makes the values
    # asymmetric bit for bit, so that copy takes the full read, with the same
    # terms in the same order: the iterates must agree bit for bit, iteration
    # counts included.  (The mixed modes read every diagonal of either copy:
    # the same comparison holds for them.)
    m = 65  # odd: the middle row, from which the DIA offsets are read, is interior
    i = 5 * m - 1
    out = []
    for lower_zero in (0.0, -0.0):
        r, c, v, n = _grid_laplacian(m, 31, i, lower_zero)
        A = capi.coo_to_csr(ctx, r, c, v, n, n)
        b = capi.Vec(ctx, np.random.default_rng(32).standard_normal(n))
        x = capi.Vec(ctx, n)
        info = capi.pcg(ctx, A, b, x, rtol=1e-10, mode=mode)
        st, nd, sym = capi.solver_storage(A)
        assert st == capi.STORAGE_DIA and nd == 4, (st, nd)
        assert sym == (np.signbit(lower_zero) == False), (sym, lower_zero)
        assert info.converged and info.iters > 100, info.iters
        out.append((info.iters, info.refinements, x.numpy().copy()))
        if mode != capi.PCG_FP64:
            rp, col, val = A.download()
            res = b.numpy() - L.csr_spmv(rp, col, val, out[-1][2])
            assert np.linalg.norm(res) <= 1.05e-10 * np.linalg.norm(b.numpy())
        for h in (x, b, A):
            h.close()
    assert out[0][:2] == out[1][:2], (out[0][:2], out[1][:2])
    assert np.array_equal(out[0][2], out[1][2])


def _batch_system(ctx, sizes, extra=None, seed=23):
    """Block-diagonal P1 system of tp meshes of the given sizes, optionally
    followed by one unstructured mesh; returns handles and offsets."""
    xy, cells, flags, voff, coff = capi.batched_tp_mesh(sizes)
    if extra is not None:
        nv0 = xy.shape[0]
        xy = np.vstack([xy, extra.xy])
        cells = np.vstack([cells, extra.cells + nv0]).astype(np.int32)
        flags = np.concatenate([flags, extra.dirichlet_flags(1)])
        voff = np.append(voff, xy.shape[0])
        coff = np.append(coff, cells.shape[0])
    alpha = np.exp(np.random.default_rng(seed).standard_normal(cells.shape[0]))
    g = capi.Mesh(ctx, xy, cells)
    fe = capi.FE(ctx, 1)
    plan = capi.Plan(ctx, g, flags)
    va = capi.Vec(ctx, alpha)
    b = capi.Vec(ctx, g.ndofs)
    plan.assemble(fe, capi.coeff(capi.COEFF_CELL, 1.0, va), capi.coeff(capi.COEFF_CONST, 0.5),
                  capi.coeff(capi.COEFF_CONST, 1.0), b)
    return g, fe, plan, va, b, voff


@pytest.mark.parametrize("mode", [capi.PCG_FP64, capi.PCG_MIXED, capi.PCG_MIXED_MATRIX])
def test_batched_block_dia_storage_gives_the_ell_iterates(ctx, mode):
    # structured problems of different sizes (odd and even n, whose middle
    # rows lie on the boundary or inside) use per-problem DIA storage; one
    # unstructured problem more forces sliced ELL.  Both sum every row in
    # column order, so the structured problems' solutions and iteration
    # counts agree bit for bit.
    sizes = [9, 17, 5, 31, 12, 24, 3, 40, 8]
    A = _batch_system(ctx, sizes)
    B = _batch_system(ctx, sizes, extra=delaunay_mesh(60, 3))
    nshared = int(A[5][-1])
    xs, its = [], []
    for g, fe, plan, va, b, voff in (A, B):
        batch = capi.Batch(ctx, plan.csr, voff)
        assert batch.storage() == -1
        x = capi.Vec(ctx, g.ndofs)
        for _ in range(2):  # the second call replays the captured graph
            x.fill(0.0)
            summ, infos = batch.pcg(b, x, rtol=1e-10, mode=mode)
            assert summ.converged
        xs.append(x.numpy()[:nshared])
        its.append([i.iters for i in infos[:len(sizes)]])
        xs_storage = batch.storage()
        its[-1].append(xs_storage)
        batch.close()
        x.close()
    # (fewer than 32768 rows: the ELL columns are 16-bit offsets)
    assert its[0][-1] == capi.STORAGE_BLOCK_DIA and its[1][-1] == capi.STORAGE_ELL16
    assert its[0][:-1] == its[1][:-1]
    assert np.array_equal(xs[0], xs[1])
    # and the solutions solve the systems
    g, fe, plan, va, b, voff = A
    rp, col, val = plan.csr.download()
    bh = b.numpy()
    for k in range(len(sizes)):
        lo, hi = int(voff[k]), int(voff[k + 1])
        sub = slice(int(rp[lo]), int(rp[hi]))
        rpk = np.ascontiguousarray(rp[lo:hi + 1] - rp[lo], np.int32)
        colk = np.ascontiguousarray(col[sub] - lo, np.int32)
        xd = L.band_cholesky_solve(rpk, colk, np.ascontiguousarray(val[sub]), bh[lo:hi])
        assert np.linalg.norm(xs[0][lo:hi] - xd) <= 1e-8 * np.linalg.norm(xd), k
    for sysh in (A, B):
        for h in sysh[4::-1]:
            h.close()


@pytest.mark.parametrize("extra", [False, True])
def test_batched_assemble_pcg_graph_matches_separate_calls(ctx, extra):
    # lfgpu_batch_assemble_pcg: the first call runs assembly + batched PCG,
    # later calls replay one captured graph (assembly, value refresh,
    # set-up, first chunk) and continue with the chunk graphs.  Every call
    # must equal plan.assemble + batch.pcg bit for bit, in every mode, also
    # when the coefficient data change between replays and with them the
    # bitwise symmetry of the per-problem DIA values (alpha = 1: symmetric;
    # alpha per cell: not), which the replayed product reads on the device.
    # extra: one unstructured problem more, which makes the batch sliced ELL.
    sizes = [9, 17, 5, 24, 12]
    xy, cells, flags, voff, coff = capi.batched_tp_mesh(sizes)
    if extra:
        em = delaunay_mesh(80, 5)
        nv0 = xy.shape[0]
        xy = np.vstack([xy, em.xy])
        cells = np.vstack([cells, em.cells + nv0]).astype(np.int32)
        flags = np.concatenate([flags, em.dirichlet_flags(1)])
        voff = np.append(voff, xy.shape[0])
    g = capi.Mesh(ctx, xy, cells)
    fe = capi.FE(ctx, 1)
    plan = capi.Plan(ctx, g, flags)
    nd = g.ndofs
    theta = capi.Vec(ctx, np.zeros(cells.shape[0]))
    coefs = (capi.coeff(capi.COEFF_EXPCELL, 1.0, theta), capi.coeff(capi.COEFF_CONST, 0.5),
             capi.coeff(capi.COEFF_SINSIN, 1.0))
    b_sep, x_sep = capi.Vec(ctx, nd), capi.Vec(ctx, nd)
    b, x = capi.Vec(ctx, nd), capi.Vec(ctx, nd)
    # one batch for both paths: the same tuned layout, so the same sums
    batch = capi.Batch(ctx, plan.csr, voff)
    rng = np.random.default_rng(3)
    thetas = [np.zeros(cells.shape[0]), rng.standard_normal(cells.shape[0]),
              np.zeros(cells.shape[0]), rng.standard_normal(cells.shape[0])]
    for mode in (capi.PCG_FP64, capi.PCG_MIXED, capi.PCG_MIXED_MATRIX):
        for call, th in enumerate(thetas):
            theta.upload(th)
            plan.assemble(fe, *coefs, b_sep)
            x_sep.fill(0.0)
            s_ref, i_ref = batch.pcg(b_sep, x_sep, rtol=1e-10, mode=mode, check_every=16)
            xr, br = x_sep.numpy(), b_sep.numpy()
            x.fill(0.0)
            s, inf = batch.assemble_pcg(plan, fe, *coefs, b, x, rtol=1e-10, mode=mode,
                                        check_every=16)
            assert s.converged and s.syncs >= 1, (mode, call)
            assert np.array_equal(b.numpy(), br), (mode, call)
            assert [i.iters for i in inf] == [i.iters for i in i_ref], (mode, call)
            assert np.array_equal(x.numpy(), xr), (mode, call)
    for h in (batch, x, b, x_sep, b_sep, theta, plan, fe, g):
        h.close()


def test_batched_p3_problems_against_the_oracle(ctx):
    # a batch of P3 problems, each with its own DOF numbering stacked into
    # one contiguous row range: the device CSR block of every problem equals
    # the oracle's reference-order system on its own mesh (pattern bit for
    # bit, values to 1e-15 of the row), and every mode's solution solves it
    # (oracle banded Cholesky, 1e-8); the one-graph step gives the same bits
    sizes = [5, 9, 3, 12, 7]
    xs, cs, cds, fls, doff, vo = [], [], [], [], [0], 0
    for n in sizes:
        m = L.tp_triag_mesh(n)
        cd, nd = m.dofs(3)
        xs.append(m.xy)
        cs.append(m.cells + vo)
        cds.append(cd + doff[-1])
        fls.append(m.dirichlet_flags(3))
        vo += m.nv
        doff.append(doff[-1] + nd)
    nd = doff[-1]
    cells = np.ascontiguousarray(np.concatenate(cs), np.int32)
    g = capi.Mesh(ctx, np.concatenate(xs), cells, np.concatenate(cds), nd)
    fe = capi.FE(ctx, 3)
    plan = capi.Plan(ctx, g, np.concatenate(fls))
    th = np.random.default_rng(8).standard_normal(cells.shape[0])
    theta = capi.Vec(ctx, th)
    b, x = capi.Vec(ctx, nd), capi.Vec(ctx, nd)
    coefs = (capi.coeff(capi.COEFF_EXPCELL, 1.0, theta), capi.coeff(capi.COEFF_CONST, 0.0),
             capi.coeff(capi.COEFF_CONST, 1.0))
    plan.assemble(fe, *coefs, b)
    rp, col, val = plan.csr.download()
    bh = b.numpy()
    batch = capi.Batch(ctx, plan.csr, np.array(doff, np.int32))
    ref = []
    c0 = 0
    for k, n in enumerate(sizes):
        m = L.tp_triag_mesh(n)
        cdo, ndo = m.dofs(3)
        rpo, colo, valo, bo = L.cpu_path_assemble(m, 3, cdo, ndo, np.exp(th[c0:c0 + m.ne]), 0.0,
                                                  L.F_CONST, 1.0, m.dirichlet_flags(3))
        c0 += m.ne
        lo, hi = doff[k], doff[k + 1]
        assert np.array_equal(rp[lo:hi + 1] - rp[lo], rpo), k
        assert np.array_equal(col[rp[lo]:rp[hi]] - lo, colo), k
        assert row_scaled_err(rpo, val[rp[lo]:rp[hi]], valo) <= 1e-15, k
        assert np.abs(bh[lo:hi] - bo).max() <= 1e-15 * np.abs(bo).max(), k
        ref.append(L.band_cholesky_solve(rpo, colo, valo, bo))
    for mode in (capi.PCG_FP64, capi.PCG_MIXED, capi.PCG_MIXED_MATRIX):
        x.fill(0.0)
        summ, infos = batch.pcg(b, x, rtol=1e-12, mode=mode)
        assert summ.converged, mode
        xv = x.numpy()
        for k in range(len(sizes)):
            xk = xv[doff[k]:doff[k + 1]]
            assert np.linalg.norm(xk - ref[k]) <= 1e-8 * np.linalg.norm(ref[k]), (mode, k)
        for _ in range(2):  # capture, then replay
            x.fill(0.0)
            batch.assemble_pcg(plan, fe, *coefs, b, x, rtol=1e-12, mode=mode)
            assert np.array_equal(x.numpy(), xv), mode
    for h in (batch, x, b, theta, plan, fe, g):
        h.close()


def test_batched_fp64_symmetry_change_between_solves(ctx):
    # one batch, one solution vector, the same chunk length: the values first
    # bitwise symmetric (alpha = 1 on a jittered mesh: G_i . G_j commutes), so
    # the FP64 product reads only the upper diagonals, then asymmetric
    # (alpha per cell: the products round differently) -- the second solve
    # must not replay the first one's captured upper-diagonal product.  Each
    # solve is checked bit for bit against a batch created for that matrix.
    sizes = [12, 20, 7, 16]
    xy, cells, flags, voff, coff = capi.batched_tp_mesh(sizes)
    rng = np.random.default_rng(41)
    xy = xy.copy()
    for k, n in enumerate(sizes):
        lo, hi = int(voff[k]), int(voff[k + 1])
        v = xy[lo:hi]
        inner = (v[:, 0] > 0) & (v[:, 0] < 1) & (v[:, 1] > 0) & (v[:, 1] < 1)
        v[inner] += rng.uniform(-0.2 / n, 0.2 / n, (int(inner.sum()), 2))
    g, fe = capi.Mesh(ctx, xy, cells), capi.FE(ctx, 1)
    plan = capi.Plan(ctx, g, flags)
    va = capi.Vec(ctx, np.exp(rng.standard_normal(cells.shape[0])))
    b, x, xf = capi.Vec(ctx, g.ndofs), capi.Vec(ctx, g.ndofs), capi.Vec(ctx, g.ndofs)
    batch = capi.Batch(ctx, plan.csr, voff)
    for alpha, symmetric in ((capi.coeff(capi.COEFF_CONST, 1.0), True),
                             (capi.coeff(capi.COEFF_CELL, 1.0, va), False)):
        plan.assemble(fe, alpha, capi.coeff(capi.COEFF_CONST, 1.0),
                      capi.coeff(capi.COEFF_CONST, 1.0), b)
        rp, col, val = plan.csr.download()
        assert _bitwise_symmetric(rp, col, val) == symmetric
        for _ in range(2):
            x.fill(0.0)
            summ, infos = batch.pcg(b, x, rtol=1e-10)
            assert summ.converged
        assert batch.storage() == capi.STORAGE_BLOCK_DIA
        fresh = capi.Batch(ctx, plan.csr, voff)
        xf.fill(0.0)
        summ_f, infos_f = fresh.pcg(b, xf, rtol=1e-10)
        fresh.close()
        assert [i.iters for i in infos] == [i.iters for i in infos_f], symmetric
        assert np.array_equal(x.numpy(), xf.numpy()), symmetric
    for h in (batch, xf, x, b, va, plan, fe, g):
        h.close()


def test_batched_rejects_coupled_problems(ctx):
    # a matrix with an entry coupling problem 0 and problem 1 is refused;
    # the same matrix without it is accepted
    n = 6
    rows = np.arange(n, dtype=np.int32)
    for coupled in (True, False):
        r, c, v = list(rows), list(rows), [2.0] * n
        if coupled:
            r += [2, 3]
            c += [3, 2]
            v += [-1.0, -1.0]
        A = capi.coo_to_csr(ctx, np.array(r, np.int32), np.array(c, np.int32), np.array(v),
                            n, n)
        if coupled:
            with pytest.raises(capi.LfgpuError) as e:
                capi.Batch(ctx, A, np.array([0, 3, 6], np.int32))
            assert e.value.code == capi.EINVAL
        else:
            capi.Batch(ctx, A, np.array([0, 3, 6], np.int32)).close()
        A.close()


@pytest.mark.parametrize("degree", [1, 3])
def test_ell_locality_order_single_and_batched_solves(ctx, degree):
    # sliced-ELL storage of a single matrix keeps its rows in the locality
    # order (by smallest column) and solves on permuted copies of x and b;
    # a batch on the same matrix needs the original row order and rebuilds
    # the storage; a single solve afterwards runs on that storage.  All
    # solutions and products must match the oracle.
    m = delaunay_mesh(500, 20 + degree) if degree == 1 else jittered_mesh(12)
    cd, nd = m.dofs(degree)
    g = gpu_mesh(ctx, m, degree)
    fe = capi.FE(ctx, degree)
    plan = capi.Plan(ctx, g, m.dirichlet_flags(degree))
    va = capi.Vec(ctx, np.exp(np.random.default_rng(degree).standard_normal(m.ne)))
    b = capi.Vec(ctx, nd)
    plan.assemble(fe, capi.coeff(capi.COEFF_CELL, 1.0, va), capi.coeff(capi.COEFF_CONST, 1.0),
                  capi.coeff(capi.COEFF_SINSIN, 1.0), b)
    rp, col, val = plan.csr.download()
    bh = b.numpy()
    xd = L.band_cholesky_solve(rp, col, val, bh)
    xv = np.random.default_rng(3).standard_normal(nd)
    x, y = capi.Vec(ctx, xv), capi.Vec(ctx, nd)
    sols = []
    for stage in ("single", "batched", "single again"):
        if stage == "batched":
            batch = capi.Batch(ctx, plan.csr, np.array([0, nd], np.int32))
            x.fill(0.0)
            summ, _ = batch.pcg(b, x, rtol=1e-12)
            # fewer than 32768 rows: the unpermuted storage has 16-bit offsets too
            assert summ.converged and batch.storage() == capi.STORAGE_ELL16
            batch.close()
        else:
            x.fill(0.0)
            info = capi.pcg(ctx, plan.csr, b, x, rtol=1e-12)
            assert info.converged
            assert capi.solver_storage(plan.csr)[0] == capi.STORAGE_ELL16
        sols.append(x.numpy())
        assert np.linalg.norm(sols[-1] - xd) <= 1e-8 * np.linalg.norm(xd), stage
        x.upload(xv)
        capi.spmv(ctx, plan.csr, x, y)
        yo = L.csr_spmv(rp, col, val, xv)
        scale = L.csr_spmv(rp, col, np.abs(val), np.abs(xv))
        assert np.all(np.abs(y.numpy() - yo) <= 1e-14 * scale + 1e-300), stage
    for h in (y, x, b, va, plan, fe, g):
        h.close()


def test_ell_column_offsets_out_of_16_bit_range(ctx):
    # A chain of 70000 unknowns closed into a ring: rows 0 and n-1 couple,
    # so no diagonal set covers the matrix (sliced ELL), and in the locality
    # order row n-1 (smallest column 0) moves next to row 0 while row n-2
    # stays at the end: one stored column lies ~70000 rows from its row and
    # the storage keeps 32-bit columns.  The same chain without the ring
    # coupling but with a coupling inside 16-bit range (rows 0 and 20000)
    # takes 16-bit offsets.  Products and solutions match the oracle's.
    n = 70000
    rng = np.random.default_rng(41)
    for far, expect in ((n - 1, capi.STORAGE_ELL), (20000, capi.STORAGE_ELL16)):
        d = 3.5 + rng.random(n)  # diagonally dominant: SPD
        e = -(0.25 + rng.random(n - 1))
        rows = np.concatenate([np.arange(n), np.arange(n - 1), np.arange(1, n), [0, far]])
        cols = np.concatenate([np.arange(n), np.arange(1, n), np.arange(n - 1), [far, 0]])
        vals = np.concatenate([d, e, e, [-0.25, -0.25]])
        A = capi.coo_to_csr(ctx, rows.astype(np.int32), cols.astype(np.int32), vals, n, n)
        rp, col, val = A.download()
        bh = rng.standard_normal(n)
        b, x, y = capi.Vec(ctx, bh), capi.Vec(ctx, n), capi.Vec(ctx, n)
        xv = rng.standard_normal(n)
        x.upload(xv)
        capi.spmv(ctx, A, x, y)
        assert capi.solver_storage(A)[0] == expect
        yo = L.csr_spmv(rp, col, val, xv)
        scale = L.csr_spmv(rp, col, np.abs(val), np.abs(xv))
        assert np.all(np.abs(y.numpy() - yo) <= 1e-14 * scale)
        xo, _, _ = L.pcg(rp, col, val, bh, tol=1e-13)
        for mode in (capi.PCG_FP64, capi.PCG_MIXED, capi.PCG_MIXED_MATRIX):
            x.fill(0.0)
            info = capi.pcg(ctx, A, b, x, rtol=1e-12, mode=mode)
            assert info.converged, (far, mode)
            assert np.linalg.norm(x.numpy() - xo) <= 1e-10 * np.linalg.norm(xo), (far, mode)
        for h in (y, x, b, A):
            h.close()


# ---------------------------------------------- adjoint gradient end to end
def test_adjoint_gradient_against_finite_differences(ctx):
    # J(theta) = 1/2 (u - u_obs)^T M (u - u_obs), A(exp(theta)) u = f;
    # dJ/dtheta by one adjoint solve + the gradient kernel, against central
    # differences of J along random directions (all solves on the device)
    n = 24
    xy, cells, _ = capi.tp_triag_mesh(n)
    flags = L.tp_triag_mesh(n).dirichlet_flags(1)
    rng = np.random.default_rng(31)
    ne, nv = cells.shape[0], xy.shape[0]
    theta = rng.standard_normal(ne)
    g = capi.Mesh(ctx, xy, cells)
    fe = capi.FE(ctx, 1)
    A = capi.Plan(ctx, g, flags)
    M = capi.Plan(ctx, g, flags)
    M.assemble(fe, capi.coeff(capi.COEFF_CONST, 0.0), capi.coeff(capi.COEFF_CONST, 1.0), None, None)
    vt = capi.Vec(ctx, theta)
    b, u, lam, rhs, grad = (capi.Vec(ctx, nv), capi.Vec(ctx, nv), capi.Vec(ctx, nv),
                            capi.Vec(ctx, nv), capi.Vec(ctx, ne))
    uobs = rng.standard_normal(nv) * (1 - flags) * 0.01
    vobs, d = capi.Vec(ctx, uobs), capi.Vec(ctx, nv)

    def forward(th):
        vt.upload(th)
        A.assemble(fe, capi.coeff(capi.COEFF_EXPCELL, 1.0, vt), capi.coeff(capi.COEFF_CONST, 0.0),
                   capi.coeff(capi.COEFF_CONST, 1.0), b)
        u.fill(0.0)
        capi.pcg(ctx, A.csr, b, u, rtol=1e-14)
        capi.axpby(ctx, 1.0, u, 0.0, d)
        capi.axpby(ctx, -1.0, vobs, 1.0, d)  # d = u - u_obs
        capi.spmv(ctx, M.csr, d, rhs)       # rhs = M d
        return 0.5 * capi.dot(ctx, d, rhs)

    J0 = forward(theta)
    lam.fill(0.0)
    capi.pcg(ctx, A.csr, rhs, lam, rtol=1e-14)   # A symmetric: adjoint = forward operator
    capi.adjoint_gradient(ctx, g, fe, vt, u, lam, grad)
    gr = grad.numpy()
    # the kernel agrees with the oracle's restatement on the same u, lambda
    go = L.adjoint_gradient_p1(L.TriMesh(xy, cells), theta, u.numpy(), lam.numpy())
    assert np.abs(gr - go).max() <= 1e-14 * np.abs(go).max()
    eps = 1e-5
    for _ in range(6):
        dv = rng.standard_normal(ne)
        dv /= np.linalg.norm(dv)
        fd = (forward(theta + eps * dv) - forward(theta - eps * dv)) / (2 * eps)
        assert abs(fd - gr @ dv) <= 1e-6 * max(abs(fd), 1e-3 * abs(J0)), (fd, gr @ dv)
    for h in (vobs, d, b, u, lam, rhs, grad, vt, M, A, fe, g):
        h.close()


# ------------------------------------------------- unstructured / edge cases
def delaunay_mesh(npts, seed):
    from scipy.spatial import Delaunay
    rng = np.random.default_rng(seed)
    g = np.linspace(0.0, 1.0, 9)
    border = np.array([(x, y) for x in g for y in (0.0, 1.0)] + [(x, y) for y in g[1:-1] for x in (0.0, 1.0)])
    pts = np.vstack([border, rng.uniform(0.02, 0.98, size=(npts, 2))])
    tri = Delaunay(pts).simplices.astype(np.int32)
    # counter-clockwise orientation is not required; keep Delaunay's order
    return L.TriMesh(pts, tri)


@pytest.mark.parametrize("degree", [1, 3])
def test_unstructured_mesh_assembly_and_solve(ctx, degree):
    m = delaunay_mesh(400, 5 + degree)
    rng = np.random.default_rng(degree)
    alpha = np.exp(rng.standard_normal(m.ne))
    (rp, col, val, b), (rpo, colo, valo, bo) = _assemble_both(ctx, m, degree, alpha, 1.0,
                                                              L.F_SINSIN)
    assert np.array_equal(rp, rpo) and np.array_equal(col, colo)
    assert np.array_equal(val, valo)
    assert np.abs(b - bo).max() <= 1e-15 * np.abs(bo).max() * 8
    # solve on the device (sliced-ELL storage) vs the direct solve
    A = capi.coo_to_csr(ctx, np.repeat(np.arange(len(rp) - 1), np.diff(rp)), col, val,
                        len(rp) - 1, len(rp) - 1)
    vb, x = capi.Vec(ctx, b), capi.Vec(ctx, len(b))
    info = capi.pcg(ctx, A, vb, x, rtol=1e-12)
    xd = L.band_cholesky_solve(rpo, colo, valo, bo)
    assert info.converged and np.linalg.norm(x.numpy() - xd) <= 1e-9 * np.linalg.norm(xd)
    for h in (vb, x, A):
        h.close()


def test_plan_without_fixed_dofs_is_the_plain_pattern(ctx):
    # no Dirichlet selector: every triplet stays (Neumann matrix, rows sum to 0)
    m = L.tp_triag_mesh(12)
    g = gpu_mesh(ctx, m, 1)
    fe = capi.FE(ctx, 1)
    plan = capi.Plan(ctx, g, None)
    plan.assemble(fe, capi.coeff(capi.COEFF_CONST, 1.0), capi.coeff(capi.COEFF_CONST, 0.0),
                  None, None)
    rp, col, val = plan.csr.download()
    cd, nd = m.dofs(1)
    ro = L.set_from_triplets(*L.assemble_coo(cd, L.element_matrices(m, 1, 1.0, 0.0)), nd)
    for u, v in zip((rp, col, val), ro):
        assert np.array_equal(u, v)
    rs = np.add.reduceat(val, rp[:-1])
    assert np.abs(rs).max() < 1e-12  # bvp_fe_tests.cc:57-60
    for h in (plan, fe, g):
        h.close()


def test_all_dofs_fixed_gives_identity(ctx):
    m = L.tp_triag_mesh(3)
    g = gpu_mesh(ctx, m, 1)
    fe = capi.FE(ctx, 1)
    plan = capi.Plan(ctx, g, np.ones(m.nv, np.uint8))
    vals = capi.Vec(ctx, np.arange(m.nv, dtype=float))
    b = capi.Vec(ctx, m.nv)
    plan.assemble(fe, capi.coeff(capi.COEFF_CONST, 1.0), capi.coeff(capi.COEFF_CONST, 0.0),
                  capi.coeff(capi.COEFF_CONST, 1.0), b, vals)
    rp, col, val = plan.csr.download()
    assert np.array_equal(col, np.arange(m.nv)) and np.all(val == 1.0)
    assert np.array_equal(b.numpy(), np.arange(m.nv, dtype=float))
    for h in (vals, b, plan, fe, g):
        h.close()


def test_invalid_inputs_are_rejected(ctx):
    with pytest.raises(capi.LfgpuError) as e:
        capi.Mesh(ctx, np.zeros((0, 2)), np.zeros((0, 3), np.int32))
    assert e.value.code == capi.EINVAL
    with pytest.raises(capi.LfgpuError) as e:  # DOF index outside [0, ndofs)
        capi.Mesh(ctx, np.array([[0, 0], [1, 0], [0, 1.0]]), np.array([[0, 1, 2]]),
                  np.array([[0, 1, 2, 3, 4, 5, 6, 7, 8, 99]]), 10)
    assert e.value.code == capi.ERANGE
    m = L.tp_triag_mesh(4)
    g = gpu_mesh(ctx, m, 1)
    fe3 = capi.FE(ctx, 3)
    with pytest.raises(capi.LfgpuError) as e:  # P3 element on a P1 DOF layout
        capi.element_matrices(ctx, g, fe3, capi.coeff(), capi.coeff())
    assert e.value.code == capi.EINVAL
    fe1 = capi.FE(ctx, 1)
    short = capi.Vec(ctx, 5)
    with pytest.raises(capi.LfgpuError) as e:  # per-cell data of the wrong length
        capi.element_matrices(ctx, g, fe1, capi.coeff(capi.COEFF_CELL, 1.0, short), capi.coeff())
    assert e.value.code == capi.EINVAL
    for h in (short, fe1, fe3, g):
        h.close()


# ------------------------------------- matrix-core (MFMA) contraction path
def _row_rel(K, Ko):
    """max |K - Ko| / max |Ko row| over the rows of every element matrix."""
    scale = np.maximum(np.abs(Ko).max(axis=-1, keepdims=True), 1e-300)
    return float((np.abs(K - Ko) / scale).max())


@pytest.mark.parametrize("degree", [1, 3])
@pytest.mark.parametrize("kind", ["const", "expcell", "qp", "sinsin"])
def test_mfma_element_matrices_match_oracle(ctx, degree, kind):
    # lfgpu_element_matrices_mfma regroups loc_comp_ellbvp.h:324-337 into
    # coefficient x reference-matrix products; "const"/"expcell" take the
    # folded four-term contraction, "qp"/"sinsin" the per-point one
    m = jittered_mesh(12 if degree == 3 else 24, seed=5)
    g = gpu_mesh(ctx, m, degree)
    fe = capi.FE(ctx, degree)
    rng = np.random.default_rng(21)
    q = fe.nqp
    vecs = []
    if kind == "const":
        a_h, c_h = 2.3, 0.7
        ca, cg = capi.coeff(capi.COEFF_CONST, a_h), capi.coeff(capi.COEFF_CONST, c_h)
    elif kind == "expcell":
        th = rng.standard_normal(m.ne)
        va = capi.Vec(ctx, th)
        vecs.append(va)
        ca, cg = capi.coeff(capi.COEFF_EXPCELL, 1.0, va), capi.coeff(capi.COEFF_CONST, 1.0)
        a_h, c_h = np.exp(th), 1.0
    elif kind == "qp":
        a_h = rng.uniform(0.5, 2, (m.ne, q))
        c_h = rng.uniform(0, 1, (m.ne, q))
        va, vc = capi.Vec(ctx, a_h), capi.Vec(ctx, c_h)
        vecs += [va, vc]
        ca, cg = capi.coeff(capi.COEFF_QP, 1.0, va), capi.coeff(capi.COEFF_QP, 1.0, vc)
    else:
        ca, cg = capi.coeff(capi.COEFF_SINSIN, 1.0), capi.coeff(capi.COEFF_CONST, 0.25)
        P = L.qp_coords(m, degree)
        a_h, c_h = np.sin(np.pi * P[..., 0]) * np.sin(np.pi * P[..., 1]), 0.25
    Km = capi.element_matrices(ctx, g, fe, ca, cg, mfma=True)
    Ke = capi.element_matrices(ctx, g, fe, ca, cg)
    Ko = L.element_matrices(m, degree, a_h, c_h)
    tol = 1e-12  # relative to each element-matrix row's largest entry
    assert _row_rel(Km, Ke) <= tol, _row_rel(Km, Ke)
    if kind in ("const", "qp"):  # device exp/sin vs glibc may differ in the last bit
        assert np.array_equal(Ke, Ko)
    assert _row_rel(Km, Ko) <= tol
    for v in vecs:
        v.close()
    fe.close()
    g.close()


@pytest.mark.parametrize("degree,n", [(1, 50), (3, 11)])
def test_mfma_assembly_matches_exact_assembly(ctx, degree, n):
    m = jittered_mesh(n, seed=17)
    cd, nd = m.dofs(degree)
    flags = m.dirichlet_flags(degree)
    g = gpu_mesh(ctx, m, degree)
    fe = capi.FE(ctx, degree)
    plan = capi.Plan(ctx, g, flags)
    th = capi.Vec(ctx, np.random.default_rng(n).standard_normal(m.ne))
    args = (fe, capi.coeff(capi.COEFF_EXPCELL, 1.0, th), capi.coeff(capi.COEFF_CONST, 1.0),
            capi.coeff(capi.COEFF_SINSIN, 1.0))
    b1, b2 = capi.Vec(ctx, nd), capi.Vec(ctx, nd)
    plan.assemble(*args, b1)
    rp, col, v1 = plan.csr.download()
    plan.assemble(*args, b2, mfma=True)
    rp2, col2, v2 = plan.csr.download()
    assert np.array_equal(rp, rp2) and np.array_equal(col, col2)
    assert row_scaled_err(rp, v2, v1) <= 1e-12
    # the element vectors follow the reference order in both paths
    assert np.array_equal(b1.numpy(), b2.numpy())
    # and the oracle's matrix, pattern and rhs
    alpha = np.exp(th.numpy())
    rpo, colo, valo, bo = L.galerkin_system(m, degree, alpha, 1.0, L.F_SINSIN, 1.0, flags=flags)
    assert np.array_equal(rp2, rpo) and np.array_equal(col2, colo)
    assert row_scaled_err(rp, v2, valo) <= 1e-12
    assert np.abs(b2.numpy() - bo).max() <= 8e-15 * np.abs(bo).max()
    # prescribed boundary values go through the same element matrices
    gv = capi.Vec(ctx, np.random.default_rng(2).standard_normal(nd) * flags)
    plan.assemble(*args, b1, gv)
    plan.assemble(*args, b2, gv, mfma=True)
    assert np.abs(b1.numpy() - b2.numpy()).max() <= 1e-12 * np.abs(b1.numpy()).max()
    # the solve on the contracted matrix agrees with the solve on the exact one
    x = capi.Vec(ctx, nd)
    info = capi.pcg(ctx, plan.csr, b2, x, rtol=1e-12)
    assert info.converged
    xo = L.band_cholesky_solve(rpo, colo, valo, b2.numpy())
    assert np.abs(x.numpy() - xo).max() <= 1e-8 * np.abs(xo).max()
    for h in (gv, x, b1, b2, th, plan, fe, g):
        h.close()


# ------------------------------------------------------------- error norms
@pytest.mark.parametrize("degree", [1, 3])
def test_error_norms_match_oracle(ctx, degree):
    # lfgpu_error_norms vs the oracle's cell-by-cell IntegrateMeshFunction
    # restatement; the sums differ only in summation order (tolerance 1e-12)
    m = jittered_mesh(20 if degree == 1 else 9, seed=8)
    cd, nd = m.dofs(degree)
    g = gpu_mesh(ctx, m, degree)
    fe = capi.FE(ctx, degree)
    rng = np.random.default_rng(5)
    uh = rng.standard_normal(nd) * 0.1
    vu = capi.Vec(ctx, uh)
    l2, h1 = capi.error_norms(ctx, g, fe, vu)
    l2o, h1o = L.error_norms(m, degree, uh)
    assert abs(l2 - l2o) <= 1e-12 * l2o and abs(h1 - h1o) <= 1e-12 * h1o
    q = fe.nqp
    uq = rng.standard_normal((m.ne, q))
    gq = rng.standard_normal((m.ne, q, 2))
    vq, vg = capi.Vec(ctx, uq), capi.Vec(ctx, gq)
    l2, h1 = capi.error_norms(ctx, g, fe, vu, vq, vg)
    l2o, h1o = L.error_norms(m, degree, uh, uq, gq)
    assert abs(l2 - l2o) <= 1e-12 * l2o and abs(h1 - h1o) <= 1e-12 * h1o
    # size errors are rejected
    with pytest.raises(capi.LfgpuError):
        capi.error_norms(ctx, g, fe, vu, vq, vq)
    for h in (vq, vg, vu, fe, g):
        h.close()


def test_manufactured_solution_errors_on_gpu(ctx):
    # the homDir_linfe_demo.cc:114-188 loop end to end on the device: assemble
    # -Laplace u = 2 pi^2 sin sin, solve, integrate the errors; L2 O(h^2),
    # H1-seminorm O(h), and the numbers equal the oracle pipeline's
    errs = []
    for n in (16, 32, 64):
        m = L.tp_triag_mesh(n)
        g = gpu_mesh(ctx, m, 1)
        fe = capi.FE(ctx, 1)
        plan = capi.Plan(ctx, g, m.dirichlet_flags(1))
        b, x = capi.Vec(ctx, m.nv), capi.Vec(ctx, m.nv)
        plan.assemble(fe, capi.coeff(capi.COEFF_CONST, 1.0), capi.coeff(capi.COEFF_CONST, 0.0),
                      capi.coeff(capi.COEFF_SINSIN, 2 * np.pi ** 2), b)
        capi.pcg(ctx, plan.csr, b, x, rtol=1e-13)
        e = np.sqrt(capi.error_norms(ctx, g, fe, x))
        eo = np.sqrt(L.error_norms(m, 1, x.numpy()))
        assert np.allclose(e, eo, rtol=1e-12)
        errs.append(e)
        for h in (b, x, plan, fe, g):
            h.close()
    e = np.array(errs)
    r = e[:-1] / e[1:]
    assert np.all((r[:, 0] > 3.7) & (r[:, 0] < 4.3)) and np.all((r[:, 1] > 1.9) & (r[:, 1] < 2.1)), r


# ------------------------------------------------------ boundary edge terms
def _edge_coeffs(ctx, kind, n_edges, nq, rng):
    """(device coefficient, oracle (kind, const, data), keep-alive vectors)."""
    if kind == "const":
        return capi.coeff(capi.COEFF_CONST, 1.7), (L.G_CONST, 1.7, None), []
    if kind == "edge":
        d = rng.uniform(0.5, 2.0, n_edges)
        v = capi.Vec(ctx, d)
        return capi.coeff(capi.COEFF_CELL, 1.0, v), (L.G_EDGE, 1.0, d), [v]
    if kind == "qp":
        d = rng.uniform(-1.0, 1.0, (n_edges, nq))
        v = capi.Vec(ctx, d)
        return capi.coeff(capi.COEFF_QP, 1.0, v), (L.G_QP, 1.0, d), [v]
    return capi.coeff(capi.COEFF_SINSIN, 2.0), (L.G_SINSIN, 2.0, None), []


@pytest.mark.parametrize("degree", [1, 3])
@pytest.mark.parametrize("kind", ["const", "edge", "qp", "sinsin"])
def test_edge_matrices_and_vectors_match_oracle(ctx, degree, kind):
    # MassEdgeMatrixProvider / ScalarLoadEdgeVectorProvider for every edge
    m = jittered_mesh(7, seed=3)
    g = gpu_mesh(ctx, m, degree)
    ed = m.edge_dofs(degree)
    E = capi.Edges(ctx, g, degree, m.edges, None if degree == 1 else ed)
    rng = np.random.default_rng(6)
    c, co, keep = _edge_coeffs(ctx, kind, m.nedges, degree + 1, rng)
    Mg = capi.edge_matrices(ctx, E, c)
    vg = capi.edge_vectors(ctx, E, c)
    Mo = L.edge_matrices(m.xy, m.edges, degree, *co)
    vo = L.edge_vectors(m.xy, m.edges, degree, *co)
    if kind == "sinsin":  # device sin vs glibc sin may differ in the last bit
        assert np.abs(Mg - Mo).max() <= 1e-15 * np.abs(Mo).max()
        assert np.abs(vg - vo).max() <= 1e-15 * np.abs(vo).max()
    else:
        assert np.array_equal(Mg, Mo) and np.array_equal(vg, vo)
    for h in keep + [E, g]:
        h.close()


def _mixed_bc_setup(m, degree):
    """Dirichlet on x = 0 (with values), impedance (edge mass + load) on
    y = 0 and x = 1, Neumann (load only) on y = 1."""
    _, ef = m.boundary_flags()
    mid = 0.5 * (m.xy[m.edges[:, 0]] + m.xy[m.edges[:, 1]])
    left = (ef == 1) & (mid[:, 0] < 1e-9)
    imp = (ef == 1) & ((mid[:, 1] < 1e-9) | (mid[:, 0] > 1 - 1e-9)) & ~left
    neu = (ef == 1) & ~left & ~imp
    ed = m.edge_dofs(degree)
    cd, nd = m.dofs(degree)
    flags = np.zeros(nd, np.uint8)
    flags[ed[left].ravel()] = 1
    return flags, imp.astype(np.uint8), (imp | neu).astype(np.uint8)


@pytest.mark.parametrize("degree,n", [(1, 24), (3, 7)])
def test_galerkin_with_edge_terms_matches_oracle(ctx, degree, n):
    # ellbvp_linfe_demo.cc:272-330 pattern: cell terms, edge mass on impedance
    # edges, edge loads on Neumann + impedance edges, Dirichlet values on the
    # rest; CSR and rhs bit for bit against the oracle pipeline
    m = jittered_mesh(n, seed=9)
    cd, nd = m.dofs(degree)
    flags, mat_on, vec_on = _mixed_bc_setup(m, degree)
    rng = np.random.default_rng(n)
    alpha = np.exp(rng.standard_normal(m.ne))
    eta_d = rng.uniform(0.5, 2.0, m.nedges)
    h_d = rng.uniform(-1.0, 1.0, (m.nedges, degree + 1))
    gvals = rng.standard_normal(nd) * flags
    g = gpu_mesh(ctx, m, degree)
    E = capi.Edges(ctx, g, degree, m.edges, None if degree == 1 else m.edge_dofs(degree))
    fe = capi.FE(ctx, degree)
    plan = capi.Plan(ctx, g, flags)
    plan.set_edges(E, mat_on, vec_on)
    va, ve, vh, vg = (capi.Vec(ctx, alpha), capi.Vec(ctx, eta_d), capi.Vec(ctx, h_d),
                      capi.Vec(ctx, gvals))
    rhs = capi.Vec(ctx, nd)
    plan.assemble_bd(fe, capi.coeff(capi.COEFF_CELL, 1.0, va), capi.coeff(capi.COEFF_CONST, 0.5),
                     capi.coeff(capi.COEFF_CONST, 1.0), capi.coeff(capi.COEFF_CELL, 1.0, ve),
                     capi.coeff(capi.COEFF_QP, 1.0, vh), rhs, vg)
    rp, col, val = plan.csr.download()
    rpo, colo, valo, bo = L.galerkin_system(
        m, degree, alpha, 0.5, L.F_CONST, 1.0, flags=flags, fixed_vals=gvals,
        edge_terms=(mat_on, (L.G_EDGE, 1.0, eta_d), vec_on, (L.G_QP, 1.0, h_d)))
    assert np.array_equal(rp, rpo) and np.array_equal(col, colo)
    assert np.array_equal(val, valo), row_scaled_err(rp, val, valo)
    assert np.array_equal(rhs.numpy(), bo), np.abs(rhs.numpy() - bo).max()
    # the plain assembly of the same plan leaves the edge terms out
    plan.assemble(fe, capi.coeff(capi.COEFF_CELL, 1.0, va), capi.coeff(capi.COEFF_CONST, 0.5),
                  capi.coeff(capi.COEFF_CONST, 1.0), rhs, vg)
    _, _, val2 = plan.csr.download()
    _, _, valo2, _ = L.galerkin_system(m, degree, alpha, 0.5, L.F_CONST, 1.0, flags=flags,
                                       fixed_vals=gvals)
    assert np.array_equal(val2, valo2)
    # an edge term without its coefficient is an error
    with pytest.raises(capi.LfgpuError):
        plan.assemble_bd(fe, capi.coeff(capi.COEFF_CELL, 1.0, va),
                         capi.coeff(capi.COEFF_CONST, 0.5), None, None, None, rhs)
    for h in (rhs, va, ve, vh, vg, plan, fe, E, g):
        h.close()


@pytest.mark.parametrize("degree", [1, 3])
def test_robin_problem_solution_on_gpu(ctx, degree):
    # -Laplace u + u = f, du/dn + u = h on the whole boundary, u = 1 + x + 2y
    # in the FE space: assembled with the edge terms and solved on the
    # device, the solution is the nodal interpolant of u
    n = 10
    m = L.tp_triag_mesh(n)
    cd, nd = m.dofs(degree)
    q = L.qp_coords(m, degree)
    u = lambda x, y: 1 + x + 2 * y
    _, ef = m.boundary_flags()
    sel = np.nonzero(ef)[0]
    p, _ = L.segment_rule(degree + 1)
    a, b = m.xy[m.edges[sel, 0]], m.xy[m.edges[sel, 1]]
    pts = b[:, None, :] * p[None, :, None] + a[:, None, :] * (1 - p)[None, :, None]
    mid = 0.5 * (a + b)
    dudn = np.select([mid[:, 0] < 1e-12, mid[:, 0] > 1 - 1e-12, mid[:, 1] < 1e-12],
                     [-1.0, 1.0, -2.0], 2.0)
    h = np.zeros((m.nedges, degree + 1))
    h[sel] = dudn[:, None] + u(pts[..., 0], pts[..., 1])
    g = gpu_mesh(ctx, m, degree)
    E = capi.Edges(ctx, g, degree, m.edges, None if degree == 1 else m.edge_dofs(degree))
    fe = capi.FE(ctx, degree)
    plan = capi.Plan(ctx, g, np.zeros(nd, np.uint8))
    plan.set_edges(E, ef, ef)
    vf, vh = capi.Vec(ctx, u(q[..., 0], q[..., 1])), capi.Vec(ctx, h)
    rhs, x = capi.Vec(ctx, nd), capi.Vec(ctx, nd)
    plan.assemble_bd(fe, capi.coeff(capi.COEFF_CONST, 1.0), capi.coeff(capi.COEFF_CONST, 1.0),
                     capi.coeff(capi.COEFF_QP, 1.0, vf), capi.coeff(capi.COEFF_CONST, 1.0),
                     capi.coeff(capi.COEFF_QP, 1.0, vh), rhs)
    info = capi.pcg(ctx, plan.csr, rhs, x, rtol=1e-13)
    assert info.converged
    # nodal interpolant: every cell writes u at its evaluation nodes
    nodes = (np.array([[0, 0], [1, 0], [0, 1]]) if degree == 1 else
             np.array([[0, 0], [1, 0], [0, 1], [1 / 3, 0], [2 / 3, 0], [2 / 3, 1 / 3],
                       [1 / 3, 2 / 3], [0, 2 / 3], [0, 1 / 3], [1 / 3, 1 / 3]]))
    xi = np.zeros(nd)
    p0, p1, p2 = m.xy[m.cells[:, 0]], m.xy[m.cells[:, 1]], m.xy[m.cells[:, 2]]
    for i, (s, t) in enumerate(nodes):
        pt = p0 * (1 - s - t) + p1 * s + p2 * t
        xi[cd[:, i]] = u(pt[:, 0], pt[:, 1])
    assert np.abs(x.numpy() - xi).max() < 1e-10
    for hh in (x, rhs, vf, vh, plan, fe, E, g):
        hh.close()


# --------------------------------------------- Fisher/KPP Strang time stepping
def _fkpp_gpu(ctx, m, c=1.2, h=1.0):
    from lehrfempp_amd import fisher_kpp
    g = capi.Mesh(ctx, m.xy, m.cells)
    E = capi.Edges(ctx, g, 1, m.edges)
    ef = fisher_kpp.boundary_edge_flags(m.nedges, m.cell_e)
    assert np.array_equal(ef, m.boundary_flags()[1])
    A, M = fisher_kpp.assemble_galerkin_matrices(ctx, g, E, ef, capi.coeff(capi.COEFF_CONST, c),
                                                 capi.coeff(capi.COEFF_CONST, h))
    return g, E, A, M


def _scipy(csr_triplet, n):
    import scipy.sparse as sp
    rp, col, val = csr_triplet
    return sp.csr_matrix((val, col, rp), shape=(n, n))


def test_fisher_kpp_matrices_and_csr_combinations(ctx):
    # assembleGalerkinMatrices on the device equals the oracle bit for bit;
    # clone copies, axpby combines value by value on the shared pattern
    m = L.tp_triag_mesh(11)
    g, E, A, M = _fkpp_gpu(ctx, m, c=1.2, h=0.7)
    (ra, ca, va), (rm, cm, vm) = L.fisher_kpp_matrices(m, 1.2, (L.G_CONST, 0.7, None))
    rpA, cA, vA = A.download()
    rpM, cM, vM = M.download()
    assert np.array_equal(rpA, ra) and np.array_equal(cA, ca) and np.array_equal(vA, va)
    assert np.array_equal(rpM, rm) and np.array_equal(cM, cm) and np.array_equal(vM, vm)
    Z = capi.csr_axpby(ctx, 1.0, M, 0.37, A)
    rz, cz, vz = Z.download()
    assert np.array_equal(rz, rpM) and np.array_equal(cz, cM)
    assert np.array_equal(vz, 1.0 * vM + 0.37 * vA)
    Cl = capi.csr_clone(ctx, A)
    assert all(np.array_equal(a, b) for a, b in zip(Cl.download(), (rpA, cA, vA)))
    # same size and nonzero count, different pattern
    n = 5
    X = capi.coo_to_csr(ctx, np.arange(n), np.arange(n), np.ones(n), n, n)
    cols = np.arange(n)
    cols[0] = 1
    Y = capi.coo_to_csr(ctx, np.arange(n), cols, np.ones(n), n, n)
    with pytest.raises(capi.LfgpuError):
        capi.csr_axpby(ctx, 1.0, X, 1.0, Y)
    with pytest.raises(capi.LfgpuError):
        capi.csr_axpby(ctx, 1.0, X, 1.0, A)
    for h in (X, Y, Cl, Z, A, M, E, g):
        h.close()


def test_strang_diffusion_operator_matches_oracle(ctx):
    # diffusionEvolutionOperator: first (tau/2) and later (tau) calls against
    # the oracle's SuperLU stages; PCG to 1e-13, tolerance 1e-11 of max|out|
    from lehrfempp_amd import fisher_kpp
    m = L.tp_triag_mesh(20)
    g, E, A, M = _fkpp_gpu(ctx, m)
    T, steps, lam = 1.0, 10, 2.1
    ss = fisher_kpp.StrangSplit(ctx, A, M, T, steps, lam, rtol=1e-13)
    (ra, ca, va), (rm, cm, vm) = L.fisher_kpp_matrices(m, 1.2, (L.G_CONST, 1.0, None))
    ref = L.StrangSplit(_scipy((ra, ca, va), m.nv), _scipy((rm, cm, vm), m.nv), T, steps, lam)
    mu = np.random.default_rng(3).random(m.nv)
    vmu = capi.Vec(ctx, mu)
    for first in (True, False, True):
        out = ss.diffusionEvolutionOperator(first, vmu)
        info = ss.last_info
        assert info.converged and info.solves == 2 and info.diffusion_steps == 1
        want = ref.diffusion(first, mu)
        got = out.numpy()
        assert np.abs(got - want).max() < 1e-11 * np.abs(want).max(), first
        out.close()
    ss.close()
    for h in (vmu, A, M, E, g):
        h.close()


@pytest.mark.parametrize("steps", [1, 2, 12])
def test_strang_evolution_matches_oracle(ctx, steps):
    # Evolution with the reaction (lambda = 2.1, carrying capacity K varying
    # in space) from a point source as in convergence_meshwidth_main.cc
    from lehrfempp_amd import fisher_kpp
    m = L.tp_triag_mesh(24)
    g, E, A, M = _fkpp_gpu(ctx, m)
    T, lam = 1.0, 2.1
    K = 0.8 + 0.4 * m.xy[:, 0]
    u0 = np.zeros(m.nv)
    u0[13] = 0.3
    u0[300] = 0.5
    ss = fisher_kpp.StrangSplit(ctx, A, M, T, steps, lam, rtol=1e-13)
    vK, vu = capi.Vec(ctx, K), capi.Vec(ctx, u0)
    out = ss.Evolution(vK, vu)
    info = ss.last_info
    nd = 1 + max(steps - 2, 0)
    assert info.converged and info.diffusion_steps == nd and info.solves == 2 * nd
    (ra, ca, va), (rm, cm, vm) = L.fisher_kpp_matrices(m, 1.2, (L.G_CONST, 1.0, None))
    ref = L.StrangSplit(_scipy((ra, ca, va), m.nv), _scipy((rm, cm, vm), m.nv), T, steps, lam)
    want = ref.evolution(K, u0)
    got = out.numpy()
    assert np.abs(got - want).max() < 1e-10 * np.abs(want).max()
    # in-place: out aliases mu
    capi.copy(ctx, vu, out)
    ss._s.evolution(vK, out, out)
    assert np.abs(out.numpy() - want).max() < 1e-10 * np.abs(want).max()
    ss.close()
    for h in (out, vK, vu, A, M, E, g):
        h.close()


def test_strang_reaction_only_is_bitexact(ctx):
    # c = 0, h = 0: A is all (explicit) zeros, the diffusion steps are exact
    # identities and the fused reaction reproduces the oracle bit for bit;
    # both equal the logistic flow over t = (m-2) tau + tau/2
    from lehrfempp_amd import fisher_kpp
    m = L.tp_triag_mesh(16)
    g, E, A, M = _fkpp_gpu(ctx, m, c=0.0, h=0.0)
    T, steps, lam = 1.3, 7, 2.1
    rng = np.random.default_rng(8)
    K = 0.5 + rng.random(m.nv)
    u0 = 0.01 + rng.random(m.nv)
    ss = fisher_kpp.StrangSplit(ctx, A, M, T, steps, lam)
    vK, vu = capi.Vec(ctx, K), capi.Vec(ctx, u0)
    got = ss.Evolution(vK, vu)
    assert ss.last_info.iters == 0
    (ra, ca, va), (rm, cm, vm) = L.fisher_kpp_matrices(m, 0.0, (L.G_CONST, 0.0, None))
    assert not va.any()
    ref = L.StrangSplit(_scipy((ra, ca, va), m.nv), _scipy((rm, cm, vm), m.nv), T, steps, lam)
    want = ref.evolution(K, u0)
    assert np.array_equal(got.numpy(), want)
    tau = T / steps
    flow = K / (1 + (K / u0 - 1) * math.exp(-lam * ((steps - 2) * tau + tau / 2)))
    assert np.abs(want - flow).max() < 1e-14
    ss.close()
    for h in (got, vK, vu, A, M, E, g):
        h.close()


def test_strang_conserves_mass_on_a_large_mesh(ctx):
    # pure Neumann diffusion (h = 0) without reaction (lambda = 0) keeps
    # 1^T M u: the drift per stage is bounded by the solve's residual
    from lehrfempp_amd import fisher_kpp
    n = 256
    m = L.tp_triag_mesh(n)
    g, E, A, M = _fkpp_gpu(ctx, m, c=1.2, h=0.0)
    rng = np.random.default_rng(4)
    u0 = rng.random(m.nv)
    ss = fisher_kpp.StrangSplit(ctx, A, M, 0.5, 8, 0.0, rtol=1e-12)
    vK, vu = capi.Vec(ctx, np.ones(m.nv)), capi.Vec(ctx, u0)
    out = ss.Evolution(vK, vu)
    assert ss.last_info.converged
    Ms = _scipy(M.download(), m.nv)
    u = out.numpy()
    m0 = (Ms @ u0).sum()
    assert abs((Ms @ u).sum() - m0) < 1e-10 * m0
    assert np.abs(u - u0).max() > 1e-2
    ss.close()
    for h in (out, vK, vu, A, M, E, g):
        h.close()


def test_strang_invalid_arguments(ctx):
    m = L.tp_triag_mesh(4)
    g, E, A, M = _fkpp_gpu(ctx, m)
    with pytest.raises(capi.LfgpuError):
        capi.Strang(ctx, A, M, 1.0, 0, 1.0)
    ss = capi.Strang(ctx, A, M, 1.0, 3, 1.0)
    small, ok = capi.Vec(ctx, 3), capi.Vec(ctx, m.nv)
    with pytest.raises(capi.LfgpuError):
        ss.evolution(small, ok, ok)
    with pytest.raises(capi.LfgpuError):
        ss.diffusion(True, ok, small)
    ss.close()
    for h in (small, ok, A, M, E, g):
        h.close()


# ----------------------------------------------- Fisher/KPP nonlocal terms
@pytest.mark.parametrize("degree,radii", [(1, (0.0, 0.0)), (1, (0.25, 0.7)), (3, (0.0, 0.0))])
def test_nonlocal_gain_and_loss_bitexact(ctx, degree, radii):
    # the gain h at the boundary points and the loss block L (the drivers'
    # nested edge loads, modelproblem_island_main.cc:70-126) equal the
    # oracle's bit for bit, with and without the annulus cut
    m = jittered_mesh(9, amount=0.3, seed=4)
    _, ef = m.boundary_flags()
    g = gpu_mesh(ctx, m, degree)
    ed = m.edge_dofs(degree)
    E = capi.Edges(ctx, g, degree, m.edges, None if degree == 1 else ed)
    eta = capi.nonlocal_gain(ctx, E, ef, *radii)
    want = L.nonlocal_gain(m.xy, m.edges, degree, ef, *radii)
    assert np.array_equal(eta.numpy().reshape(want.shape), want)
    rows, Lv = capi.nonlocal_loss(ctx, E, ef, *radii)
    wrows, wL = L.nonlocal_loss(m.xy, m.edges, ed, degree, ef, *radii)
    assert np.array_equal(rows, wrows)
    assert np.array_equal(Lv.numpy().reshape(wL.shape), wL)
    if radii[1] > 0:
        assert np.count_nonzero(wL == 0.0) > 0 and np.count_nonzero(wL) > 0
    # nothing flagged: no rows, a zero gain
    none = np.zeros_like(ef)
    assert capi.nonlocal_loss(ctx, E, none)[0].size == 0
    z = capi.nonlocal_gain(ctx, E, none)
    assert not np.any(z.numpy())
    for h in (z, Lv, eta, E, g):
        h.close()


def _nonlocal_system(ctx, n, c=1.2, shuffle=False):
    """Device A (with the loss block attached) and M of the nonlocal model on
    a jittered n x n mesh (vertices renumbered at random with shuffle), and
    the oracle's dense A - L and M (scipy)."""
    from lehrfempp_amd import fisher_kpp
    import scipy.sparse as sp
    m = jittered_mesh(n, amount=0.3, seed=6)
    if shuffle:
        new = np.random.default_rng(12).permutation(m.nv).astype(np.int32)
        xy = np.empty_like(m.xy)
        xy[new] = m.xy
        m = L.TriMesh(xy, new[m.cells], new[m.edges])
    _, ef = m.boundary_flags()
    g = capi.Mesh(ctx, m.xy, m.cells)
    E = capi.Edges(ctx, g, 1, m.edges)
    hc, eta = fisher_kpp.nonlocal_gain(ctx, E, ef)
    loss = fisher_kpp.nonlocal_loss(ctx, E, ef)
    A, M = fisher_kpp.assemble_galerkin_matrices(ctx, g, E, ef, capi.coeff(capi.COEFF_CONST, c),
                                                 hc, L=loss)
    h = L.nonlocal_gain(m.xy, m.edges, 1, ef)
    rows, Lm = L.nonlocal_loss(m.xy, m.edges, m.edge_dofs(1), 1, ef)
    (ra, ca, va), (rm, cm, vm) = L.fisher_kpp_matrices(m, c, (L.G_QP, 1.0, h))
    Ad = _scipy((ra, ca, va), m.nv).toarray()
    Ad[np.ix_(rows, rows)] -= Lm
    handles = [loss[1], eta, E, g]
    return m, A, M, sp.csr_matrix(Ad), _scipy((rm, cm, vm), m.nv), rows, handles


def test_dense_block_on_sliced_ell(ctx):
    # vertices numbered at random: the stage matrix is stored as sliced ELL
    # in the locality order, so the block rows reach the sweeps through the
    # storage permutation; product and solve against the oracle's A - L
    import scipy.sparse.linalg as sla
    m, A, M, Aref, Mref, rows, handles = _nonlocal_system(ctx, 14, shuffle=True)
    s = 0.21
    S = capi.csr_axpby(ctx, 1.0, M, s, A)
    b = capi.Vec(ctx, np.random.default_rng(10).standard_normal(m.nv))
    u = capi.Vec(ctx, m.nv)
    info = capi.pcg(ctx, S, b, u, rtol=1e-12)
    assert info.converged
    st, _, _ = capi.solver_storage(S)
    assert st in (capi.STORAGE_ELL, capi.STORAGE_ELL16), st
    Sref = (Mref + s * Aref).tocsc()
    uref = sla.spsolve(Sref, b.numpy())
    assert np.linalg.norm(u.numpy() - uref) <= 1e-9 * np.linalg.norm(uref)
    assert (np.linalg.norm(b.numpy() - Sref @ u.numpy())
            <= 1.05e-12 * np.linalg.norm(b.numpy()))
    x = np.random.default_rng(11).standard_normal(m.nv)
    vx, vy = capi.Vec(ctx, x), capi.Vec(ctx, m.nv)
    capi.spmv(ctx, S, vx, vy)
    want = Sref @ x
    assert np.all(np.abs(vy.numpy() - want) <= 1e-14 * (abs(Sref) @ np.abs(x)) + 1e-300)
    for h in (vy, vx, u, b, S, A, M, *handles):
        h.close()


def test_dense_block_product_solve_and_combinations(ctx):
    # A carries the loss block with factor -1: its product, the FP64 PCG on
    # M + s A (axpby carries the block, scaled) and clone agree with the
    # oracle's dense A - L; the modes and paths without block support refuse
    import scipy.sparse.linalg as sla
    m, A, M, Aref, Mref, rows, handles = _nonlocal_system(ctx, 12)
    nb, scale = capi.csr_dense_block(A)
    assert nb == rows.size and scale == -1.0
    x = np.random.default_rng(8).standard_normal(m.nv)
    vx, vy = capi.Vec(ctx, x), capi.Vec(ctx, m.nv)
    capi.spmv(ctx, A, vx, vy)
    want = Aref @ x
    scale_ref = np.abs(Aref).toarray() @ np.abs(x)
    assert np.all(np.abs(vy.numpy() - want) <= 1e-14 * scale_ref + 1e-300)
    s = 0.37
    S = capi.csr_axpby(ctx, 1.0, M, s, A)
    assert capi.csr_dense_block(S) == (nb, -s)
    b = capi.Vec(ctx, np.random.default_rng(9).standard_normal(m.nv))
    u = capi.Vec(ctx, m.nv)
    info = capi.pcg(ctx, S, b, u, rtol=1e-12)
    assert info.converged
    Sref = (Mref + s * Aref).tocsc()
    uref = sla.spsolve(Sref, b.numpy())
    assert np.linalg.norm(u.numpy() - uref) <= 1e-9 * np.linalg.norm(uref)
    # the solver's residual is the block-aware product's
    r = b.numpy() - Sref @ u.numpy()
    assert np.linalg.norm(r) <= 1.05e-12 * np.linalg.norm(b.numpy())
    Cl = capi.csr_clone(ctx, S)
    assert capi.csr_dense_block(Cl) == (nb, -s)
    u2 = capi.Vec(ctx, m.nv)
    capi.pcg(ctx, Cl, b, u2, rtol=1e-12)
    assert np.array_equal(u2.numpy(), u.numpy())
    for mode in (capi.PCG_MIXED, capi.PCG_MIXED_MATRIX):
        with pytest.raises(capi.LfgpuError):
            capi.pcg(ctx, S, b, u, mode=mode)
    with pytest.raises(capi.LfgpuError):
        capi.pcg(ctx, S, b, u, single_reduction=True)
    with pytest.raises(capi.LfgpuError):
        capi.pcg_noalloc(ctx, S, b, u)
    with pytest.raises(capi.LfgpuError):
        capi.csr_axpby(ctx, 1.0, S, 1.0, A)
    # removing the block restores the sparse operator
    capi.csr_set_dense_block(ctx, Cl, [], None)
    assert capi.csr_dense_block(Cl) == (0, 0.0)
    capi.spmv(ctx, Cl, vx, vy)
    Ssp = Mref + s * _scipy(A.download(), m.nv)
    assert np.all(np.abs(vy.numpy() - Ssp @ x) <= 1e-14 * (abs(Ssp) @ np.abs(x)) + 1e-300)
    for h in (u2, Cl, u, b, S, vy, vx, A, M, *handles):
        h.close()


def test_strang_with_nonlocal_terms_matches_oracle(ctx):
    # Evolution with the nonlocal gain and loss (strangsplitting.h:120-201
    # with L != 0) against the oracle's SuperLU stepper on A - L
    from lehrfempp_amd import fisher_kpp
    m, A, M, Aref, Mref, rows, handles = _nonlocal_system(ctx, 16)
    T, steps, lam = 1.0, 8, 2.1
    K = 0.8 + 0.4 * m.xy[:, 0]
    u0 = np.exp(-20 * ((m.xy[:, 0] - 0.3) ** 2 + (m.xy[:, 1] - 0.6) ** 2))
    ss = fisher_kpp.StrangSplit(ctx, A, M, T, steps, lam, rtol=1e-13)
    vK, vu = capi.Vec(ctx, K), capi.Vec(ctx, u0)
    out = ss.Evolution(vK, vu)
    assert ss.last_info.converged
    want = L.StrangSplit(Aref, Mref, T, steps, lam).evolution(K, u0)
    got = out.numpy()
    assert np.abs(got - want).max() < 1e-10 * np.abs(want).max()
    # pure diffusion (lambda = 0) keeps the mass 1^T M u: gain and loss
    # balance at the boundary
    ss0 = fisher_kpp.StrangSplit(ctx, A, M, T, steps, 0.0, rtol=1e-13)
    cur = capi.Vec(ctx, u0)
    for k in range(4):
        nxt = ss0.diffusionEvolutionOperator(k == 0, cur)
        cur.close()
        cur = nxt
    mass0, mass = (Mref @ u0).sum(), (Mref @ cur.numpy()).sum()
    assert abs(mass - mass0) < 1e-11 * mass0
    assert np.abs(cur.numpy() - u0).max() > 1e-3
    # the mixed modes do not take the block: refused when the stepper is made
    for mode in (capi.PCG_MIXED, capi.PCG_MIXED_MATRIX):
        with pytest.raises(capi.LfgpuError):
            fisher_kpp.StrangSplit(ctx, A, M, T, steps, lam, mode=mode)
    for h in (cur, out, vK, vu, A, M, *handles):
        h.close()
    ss.close()
    ss0.close()
