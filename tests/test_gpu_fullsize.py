"""Parity at the benchmark sizes through size-independent properties (the
oracle is too slow there).  The meshes are the 2048 x 2048 P1 mesh with
a = exp(N(0,1)) per cell, gamma = 1, f = sin(pi x) sin(pi y), and the
jittered 512 x 512 P3 mesh with the same coefficient law.  The properties:
* the closed-form counts of the pattern (nnz after
  FixFlaggedSolutionComponents, boundary DOFs);
* fixed rows are identity rows with a zero load;
* rows whose cells touch no fixed DOF sum to gamma * integral(phi_i), because
  the stiffness part annihilates constants and the mass part's row sum is
  the integral of the shape function (exact for both quadrature rules);
* the total load equals the quadrature of f;
* at P1 the solves of every precision mode converge and agree.
Row sums are compared within 1e-12 of the row's absolute sum (cancellation
of O(alpha) stiffness entries against an O(h^2) result).
"""
import numpy as np
import pytest

from lehrfempp_amd import capi

pytestmark = pytest.mark.gpu


def _row_sums(rp, val):
    nr = len(rp) - 1
    s = np.zeros(nr)
    a = np.zeros(nr)
    nz = np.diff(rp) > 0
    s[nz] = np.add.reduceat(val, rp[:-1][nz])
    a[nz] = np.add.reduceat(np.abs(val), rp[:-1][nz])
    return s, a


def _clean_rows(cell_dofs, fixed, ndofs):
    """DOFs none of whose cells holds a fixed DOF (their rows keep every
    column after FixFlaggedSolutionComponents)."""
    bad_cell = fixed[cell_dofs].any(axis=1)
    touched = np.zeros(ndofs, bool)
    touched[cell_dofs[bad_cell].ravel()] = True
    return ~touched & ~fixed


def _cell_areas(xy, cells):
    p = xy[cells]
    e1, e2 = p[:, 1] - p[:, 0], p[:, 2] - p[:, 0]
    return 0.5 * np.abs(e1[:, 0] * e2[:, 1] - e1[:, 1] * e2[:, 0])


def test_c2_full_size_properties(ctx):
    n = 2048
    xy, cells, _ = capi.tp_triag_mesh(n)
    nv, ne = xy.shape[0], cells.shape[0]
    j, i = np.divmod(np.arange(nv), n + 1)
    fixed = (i == 0) | (i == n) | (j == 0) | (j == n)
    mesh, fe = capi.Mesh(ctx, xy, cells), capi.FE(ctx, 1)
    plan = capi.Plan(ctx, mesh, fixed.astype(np.uint8))
    theta = capi.Vec(ctx, np.random.default_rng(0).standard_normal(ne))
    b, x = capi.Vec(ctx, nv), capi.Vec(ctx, nv)
    plan.assemble(fe, capi.coeff(capi.COEFF_EXPCELL, 1.0, theta), capi.coeff(capi.COEFF_CONST, 1.0),
                  capi.coeff(capi.COEFF_SINSIN, 1.0), b)
    rp, col, val = plan.csr.download()
    assert len(val) == 29_323_281 and int(fixed.sum()) == 8192
    bh = b.numpy()
    fr = np.nonzero(fixed)[0]
    assert np.all(rp[fr + 1] - rp[fr] == 1) and np.all(col[rp[fr]] == fr)
    assert np.all(val[rp[fr]] == 1.0) and np.all(bh[fr] == 0.0)
    s, a = _row_sums(rp, val)
    clean = _clean_rows(cells, fixed, nv)
    h2 = 1.0 / n ** 2
    assert np.all(np.abs(s[clean] - h2) <= 1e-12 * a[clean]), np.abs(s[clean] - h2).max()
    # integral of sin(pi x) sin(pi y) over the square is 4 / pi^2; the
    # boundary hat functions miss O(h) of it
    assert abs(bh.sum() - 4 / np.pi ** 2) <= 4.0 / n
    infos = {}
    sols = {}
    for name, md in (("fp64", capi.PCG_FP64), ("mixed", capi.PCG_MIXED),
                     ("mixed_matrix", capi.PCG_MIXED_MATRIX)):
        x.fill(0.0)
        infos[name] = capi.pcg(ctx, plan.csr, b, x, rtol=1e-10, mode=md, check_every=128,
                               maxit=60000)
        assert infos[name].converged
        sols[name] = x.numpy()
    for name in ("mixed", "mixed_matrix"):
        err = np.linalg.norm(sols[name] - sols["fp64"]) / np.linalg.norm(sols["fp64"])
        assert err <= 1e-8, (name, err)
    for h in (x, b, theta, plan, fe, mesh):
        h.close()


def _c3_system(ctx, n=512):
    """The jittered 512 x 512 P3 system of the benchmark (a = exp(N(0,1)) per
    cell, gamma = 1, f = sin sin), assembled on the matrix cores."""
    xy, cells, edges = capi.tp_triag_mesh(n)
    rng = np.random.default_rng(0)
    h = 1.0 / n
    inner = (xy[:, 0] > 0) & (xy[:, 0] < 1) & (xy[:, 1] > 0) & (xy[:, 1] < 1)
    xy = xy.copy()
    xy[inner] += rng.uniform(-0.2 * h, 0.2 * h, size=(int(inner.sum()), 2))
    ev, ce, co = capi.mesh_edges(xy.shape[0], cells, edges)
    cd, nd = capi.dof_table(xy.shape[0], cells, len(ev), ce, co, 3)
    nv = xy.shape[0]
    vb = ~inner
    eb = vb[ev[:, 0]] & vb[ev[:, 1]] & (
        (xy[ev[:, 0], 0] == xy[ev[:, 1], 0]) | (xy[ev[:, 0], 1] == xy[ev[:, 1], 1]))
    fixed = np.zeros(nd, np.uint8)
    fixed[:nv] = vb
    be = np.nonzero(eb)[0]
    fixed[nv + 2 * be] = 1
    fixed[nv + 2 * be + 1] = 1
    mesh = capi.Mesh(ctx, xy, cells, cd, nd)
    fe = capi.FE(ctx, 3)
    plan = capi.Plan(ctx, mesh, fixed)
    theta = capi.Vec(ctx, np.random.default_rng(1).standard_normal(cells.shape[0]))
    b = capi.Vec(ctx, nd)
    plan.assemble(fe, capi.coeff(capi.COEFF_EXPCELL, 1.0, theta), capi.coeff(capi.COEFF_CONST, 1.0),
                  capi.coeff(capi.COEFF_SINSIN, 1.0), b, mfma=True)
    return nd, plan, b, (theta, fe, mesh)


def test_c3_full_size_pcg_modes(ctx):
    """The full-size P3 system (2,362,369 DOFs) in every precision mode:
    each converges (the mixed modes on their FP64 true residual), the
    solutions agree with one another to 1e-8 (rtol 1e-10 times the
    conditioning of the smooth modes), and the FP64 residual of the mixed
    solutions, computed by the device product, is below 2e-10."""
    nd, plan, b, keep = _c3_system(ctx)
    x, y = capi.Vec(ctx, nd), capi.Vec(ctx, nd)
    sols = {}
    for name, md in (("fp64", capi.PCG_FP64), ("mixed", capi.PCG_MIXED),
                     ("mixed_matrix", capi.PCG_MIXED_MATRIX)):
        x.fill(0.0)
        info = capi.pcg(ctx, plan.csr, b, x, rtol=1e-10, mode=md, maxit=60000, check_every=128)
        assert info.converged, name
        sols[name] = x.numpy()
        if md != capi.PCG_FP64:
            capi.spmv(ctx, plan.csr, x, y)
            bh = b.numpy()
            assert np.linalg.norm(bh - y.numpy()) <= 2e-10 * np.linalg.norm(bh), name
    for name in ("mixed", "mixed_matrix"):
        err = np.linalg.norm(sols[name] - sols["fp64"]) / np.linalg.norm(sols["fp64"])
        assert err <= 1e-8, (name, err)
    for h in (y, x, b, plan, *keep):
        h.close()


def test_c3_full_size_row_sums(ctx):
    n = 512
    xy, cells, edges = capi.tp_triag_mesh(n)
    rng = np.random.default_rng(0)
    h = 1.0 / n
    inner = (xy[:, 0] > 0) & (xy[:, 0] < 1) & (xy[:, 1] > 0) & (xy[:, 1] < 1)
    xy = xy.copy()
    xy[inner] += rng.uniform(-0.2 * h, 0.2 * h, size=(int(inner.sum()), 2))
    ev, ce, co = capi.mesh_edges(xy.shape[0], cells, edges)
    cd, nd = capi.dof_table(xy.shape[0], cells, len(ev), ce, co, 3)
    nv = xy.shape[0]
    vb = ~inner
    eb = vb[ev[:, 0]] & vb[ev[:, 1]] & (
        (xy[ev[:, 0], 0] == xy[ev[:, 1], 0]) | (xy[ev[:, 0], 1] == xy[ev[:, 1], 1]))
    fixed = np.zeros(nd, bool)
    fixed[:nv] = vb
    be = np.nonzero(eb)[0]
    fixed[nv + 2 * be] = True
    fixed[nv + 2 * be + 1] = True
    mesh = capi.Mesh(ctx, xy, cells, cd, nd)
    fe = capi.FE(ctx, 3)
    plan = capi.Plan(ctx, mesh, fixed.astype(np.uint8))
    theta = capi.Vec(ctx, np.random.default_rng(1).standard_normal(cells.shape[0]))
    b = capi.Vec(ctx, nd)
    # integral of each P3 Lagrange basis function over a triangle T:
    # vertices |T|/30, edge nodes 3|T|/40, bubble 9|T|/20
    w = np.array([1 / 30] * 3 + [3 / 40] * 6 + [9 / 20])
    area = _cell_areas(xy, cells)
    cdi = np.asarray(cd, np.int64).reshape(cells.shape[0], 10)
    integ = np.bincount(cdi.ravel(), weights=(area[:, None] * w[None, :]).ravel(), minlength=nd)
    clean = _clean_rows(cdi, fixed, nd)
    assert clean.sum() > 0.9 * nd
    for mfma in (False, True):
        plan.assemble(fe, capi.coeff(capi.COEFF_EXPCELL, 1.0, theta),
                      capi.coeff(capi.COEFF_CONST, 1.0), capi.coeff(capi.COEFF_SINSIN, 1.0), b,
                      mfma=mfma)
        rp, col, val = plan.csr.download()
        assert len(val) == 153 * n * n - 234 * n + 109
        s, a = _row_sums(rp, val)
        err = np.abs(s[clean] - integ[clean]) / a[clean]
        assert err.max() <= 1e-12, (mfma, err.max())
        assert abs(b.numpy().sum() - 4 / np.pi ** 2) <= 4.0 / n
    for hh in (b, theta, plan, fe, mesh):
        hh.close()


def test_c4_full_size_batch(ctx):
    """The 256 problems of the batched workload (sides log-uniform in
    [32, 512]): every problem converges in every mode, and each problem's
    solutions agree across modes to 1e-8 (rtol 1e-10 with kappa ~ 1e5)."""
    rng = np.random.default_rng(4)
    sizes = np.rint(np.exp(rng.uniform(np.log(32), np.log(512), 256))).astype(int)
    xy, cells, flags, voff, coff = capi.batched_tp_mesh(sizes)
    assert int(voff[-1]) == int(((sizes + 1) ** 2).sum()) == 14_647_286
    g = np.concatenate([np.random.default_rng(5 + k).standard_normal(int(coff[k + 1] - coff[k]))
                        for k in range(256)])
    mesh, fe = capi.Mesh(ctx, xy, cells), capi.FE(ctx, 1)
    plan = capi.Plan(ctx, mesh, flags)
    theta = capi.Vec(ctx, g)
    b, x = capi.Vec(ctx, xy.shape[0]), capi.Vec(ctx, xy.shape[0])
    plan.assemble(fe, capi.coeff(capi.COEFF_EXPCELL, 1.0, theta), capi.coeff(capi.COEFF_CONST, 0.0),
                  capi.coeff(capi.COEFF_CONST, 1.0), b)
    batch = capi.Batch(ctx, plan.csr, voff)
    sols = {}
    for name, md in (("fp64", capi.PCG_FP64), ("mixed", capi.PCG_MIXED),
                     ("mixed_matrix", capi.PCG_MIXED_MATRIX)):
        x.fill(0.0)
        summ, infos = batch.pcg(b, x, rtol=1e-10, maxit=60000, mode=md)
        assert summ.converged and all(i.converged for i in infos), name
        sols[name] = x.numpy()
    for name in ("mixed", "mixed_matrix"):
        for k in range(256):
            a, r = sols[name][voff[k]:voff[k + 1]], sols["fp64"][voff[k]:voff[k + 1]]
            assert np.linalg.norm(a - r) <= 1e-8 * np.linalg.norm(r), (name, k)
    # the mixed solves are bitwise repeatable: a stop found by one workgroup
    # must not make the problem's other workgroups skip their final fold
    x.fill(0.0)
    batch.pcg(b, x, rtol=1e-10, maxit=60000, mode=capi.PCG_MIXED)
    assert np.array_equal(x.numpy(), sols["mixed"])
    for h in (batch, x, b, theta, plan, fe, mesh):
        h.close()


def test_c5_full_size_adjoint_gradient(ctx):
    """dJ/dtheta on the 1024 x 1024 mesh (theta, theta* ~ N(0,1), u_obs =
    u(theta*) + 1 % noise on interior DOFs) through the autograd objective,
    against central differences along the gradient and along a mix of the
    gradient and a random direction.  (Along a purely random unit direction
    in 2M dimensions the derivative is ~|grad J| / 1400, about 4e-13 here,
    and the solves' noise in J divided by 2 eps limits the agreement to
    ~1e-3.)"""
    from lehrfempp_amd.autograd import AdjointObjective
    n = 1024
    xy, cells, _ = capi.tp_triag_mesh(n)
    nv, ne = xy.shape[0], cells.shape[0]
    j, i = np.divmod(np.arange(nv), n + 1)
    flags = ((i == 0) | (i == n) | (j == 0) | (j == n)).astype(np.uint8)
    theta = np.random.default_rng(1).standard_normal(ne)
    theta_star = np.random.default_rng(2).standard_normal(ne)
    obj = AdjointObjective(ctx, xy, cells, flags, np.zeros(nv), rtol=1e-13)
    _, ustar = obj.value(theta_star)
    obj.close()
    noise = np.random.default_rng(3).standard_normal(nv) * (1 - flags)
    obj = AdjointObjective(ctx, xy, cells, flags, ustar + 0.01 * np.abs(ustar).max() * noise,
                           rtol=1e-13)
    J0, _ = obj.value(theta)
    g = obj.gradient(theta)
    rng = np.random.default_rng(7)
    eps = 1e-5
    gn = g / np.linalg.norm(g)
    r = rng.standard_normal(ne)
    r /= np.linalg.norm(r)
    for dv in (gn, (gn + r) / np.linalg.norm(gn + r)):
        fd = (obj.value(theta + eps * dv)[0] - obj.value(theta - eps * dv)[0]) / (2 * eps)
        assert abs(fd - g @ dv) <= 1e-5 * abs(fd), (fd, g @ dv)
    obj.close()


def test_c5_full_size_fd_64_local_directions(ctx):
    """dJ/dtheta on the 1024 x 1024 mesh against central differences on 64
    random unit directions, each supported on 8 cells drawn from the 5 % of
    cells with the largest |dJ/dtheta_e| (seeded).  A direction spread over all
    2M cells has g.d ~ |g| / 1400, which the solves' noise in J / (2 eps)
    swamps; on a few cells g.d is of the size of single gradient entries.
    eps = 1e-4 (J is smooth in theta: the O(eps^2) term is ~1e-8 relative),
    solves to 1e-13; every direction must agree to 1e-4 relative."""
    from lehrfempp_amd.autograd import AdjointObjective
    n = 1024
    xy, cells, _ = capi.tp_triag_mesh(n)
    nv, ne = xy.shape[0], cells.shape[0]
    j, i = np.divmod(np.arange(nv), n + 1)
    flags = ((i == 0) | (i == n) | (j == 0) | (j == n)).astype(np.uint8)
    theta = np.random.default_rng(1).standard_normal(ne)
    theta_star = np.random.default_rng(2).standard_normal(ne)
    obj = AdjointObjective(ctx, xy, cells, flags, np.zeros(nv), rtol=1e-13)
    _, ustar = obj.value(theta_star)
    obj.close()
    noise = np.random.default_rng(3).standard_normal(nv) * (1 - flags)
    obj = AdjointObjective(ctx, xy, cells, flags, ustar + 0.01 * np.abs(ustar).max() * noise,
                           rtol=1e-13)
    obj.value(theta)
    g = obj.gradient(theta)
    top = np.argsort(-np.abs(g))[: ne // 20]
    rng = np.random.default_rng(11)
    eps = 1e-4
    errs = []
    for _ in range(64):
        cells_d = rng.choice(top, 8, replace=False)
        dv = np.zeros(ne)
        dv[cells_d] = rng.standard_normal(8)
        dv /= np.linalg.norm(dv)
        fd = (obj.value(theta + eps * dv)[0] - obj.value(theta - eps * dv)[0]) / (2 * eps)
        errs.append(abs(fd - g @ dv) / abs(g @ dv))
    obj.close()
    print("fd 64 local directions: max rel err %.2e, median %.2e" % (max(errs),
                                                                     float(np.median(errs))))
    assert max(errs) <= 1e-4, max(errs)
