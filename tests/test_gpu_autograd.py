"""The torch.autograd binding of the adjoint gradient (lehrfempp_amd.autograd),
which exists for checking.  It is checked against the direct C-ABI gradient
(bit for bit), against central differences of J along random directions,
and against the FP64 CPU adjoint built from the oracle's assembly with
direct solves.

J(theta) = 1/2 (u - u_obs)^T M (u - u_obs), with A(exp(theta)) u = b, on a
20 x 20 P1 mesh with zero Dirichlet data.  Tolerances: the device solves
stop at rtol 1e-13.  With kappa ~ 1e3 the gradient is within ~1e-9 of the
direct-solve adjoint, and the test allows 1e-7.  Central differences with
eps = 1e-5 are within 1e-6 of the directional derivative.
"""
import numpy as np
import pytest
import torch

from lehrfempp_amd.autograd import AdjointObjective
from oracle import lfref as L

pytestmark = pytest.mark.gpu


def test_autograd_gradient_matches_fd_and_the_cpu_adjoint(ctx):
    n = 20
    m = L.tp_triag_mesh(n)
    flags = m.dirichlet_flags(1)
    rng = np.random.default_rng(41)
    u_obs = 0.01 * rng.standard_normal(m.nv) * (1 - flags)
    obj = AdjointObjective(ctx, m.xy, m.cells, flags, u_obs, rtol=1e-13)
    theta = rng.standard_normal(m.ne)

    t = torch.tensor(theta, dtype=torch.float64, requires_grad=True)
    J = obj(t)
    J.backward()
    g = t.grad.numpy().copy()
    assert np.array_equal(g, obj.gradient(theta))

    eps = 1e-5
    for _ in range(4):
        dv = rng.standard_normal(m.ne)
        dv /= np.linalg.norm(dv)
        fd = (obj.value(theta + eps * dv)[0] - obj.value(theta - eps * dv)[0]) / (2 * eps)
        assert abs(fd - g @ dv) <= 1e-6 * max(abs(fd), 1e-3 * J.detach().item()), (fd, g @ dv)

    # FP64 CPU adjoint: oracle assembly (A with alpha = exp(theta), the mass
    # matrix with the same fixed DOFs), banded Cholesky for both solves
    rp, col, val, b = L.galerkin_system(m, 1, np.exp(theta), 0.0, L.F_CONST, 1.0, flags=flags)
    u = L.band_cholesky_solve(rp, col, val, b)
    mrp, mcol, mval, _ = L.galerkin_system(m, 1, 0.0, 1.0, L.F_CONST, 0.0, flags=flags)
    d = u - u_obs
    Md = L.csr_spmv(mrp, mcol, mval, d)
    lam = L.band_cholesky_solve(rp, col, val, Md)
    go = L.adjoint_gradient_p1(m, theta, u, lam)
    assert np.linalg.norm(g - go) <= 1e-7 * np.linalg.norm(go)
    assert abs(J.detach().item() - 0.5 * d @ Md) <= 1e-9 * abs(0.5 * d @ Md)

    # the chain rule through torch: d(3 J + J^2)/dtheta = (3 + 2 J) dJ/dtheta
    t2 = torch.tensor(theta, dtype=torch.float64, requires_grad=True)
    J2 = obj(t2)
    (3.0 * J2 + J2 * J2).backward()
    assert np.allclose(t2.grad.numpy(), (3.0 + 2.0 * J2.detach().item()) * g, rtol=1e-13, atol=0)

    # a forward evaluation between forward and backward does not change
    # the gradient (the state is recomputed from the saved theta)
    t3 = torch.tensor(theta, dtype=torch.float64, requires_grad=True)
    J3 = obj(t3)
    obj.value(theta + 0.1)
    J3.backward()
    assert np.array_equal(t3.grad.numpy(), g)
    obj.close()
