"""torch.autograd binding of the adjoint gradient, for checking only.

The objective is J(theta) = 1/2 (u - u_obs)^T M (u - u_obs), where
A(theta) u = b, A(theta) = sum_e exp(theta_e) K_e.  K_e is
ReactionDiffusionElementMatrixProvider with alpha = exp(theta) per cell and
gamma = 0 (uscalfe/loc_comp_ellbvp.h:266-339), and b is the load of a
constant f (:691-746).  M is the P1 mass matrix.  Both matrices go through
FixFlaggedSolutionComponents (assemble/fix_dof.h:86-138) with the same
fixed DOFs and zero prescribed values.

The forward pass assembles A(theta) and b and solves A u = b by Jacobi-PCG
(lfgpu_pcg).  The backward pass solves the adjoint system A lambda =
M (u - u_obs) and evaluates dJ/dtheta_e = -exp(theta_e) lambda_e^T K_e u_e
with lfgpu_adjoint_gradient.  All arithmetic runs in liblfgpu through the C
ABI.  torch only carries theta in and the gradient out, by host copies,
which are sized for checking, not for production use.
"""
from __future__ import annotations

import numpy as np
import torch

from . import capi


class AdjointObjective:
    """J(theta) on a P1 triangle mesh, callable on a float64 torch tensor of
    per-cell log coefficients.  The returned 0-dim tensor back-propagates
    through the adjoint solve."""

    def __init__(self, ctx, xy, cells, fixed_flags, u_obs, f=1.0, rtol=1e-12):
        self.ctx = ctx
        self.rtol = float(rtol)
        self.f = float(f)
        self.mesh = capi.Mesh(ctx, xy, cells)
        self.fe = capi.FE(ctx, 1)
        self.ne, self.nv = int(np.asarray(cells).shape[0]), int(np.asarray(xy).shape[0])
        flags = np.ascontiguousarray(fixed_flags, dtype=np.uint8)
        if flags.shape != (self.nv,):
            raise ValueError("fixed_flags must have one entry per vertex")
        u_obs = np.ascontiguousarray(u_obs, dtype=np.float64)
        if u_obs.shape != (self.nv,):
            raise ValueError("u_obs must have one entry per vertex")
        self.A = capi.Plan(ctx, self.mesh, flags)
        self.M = capi.Plan(ctx, self.mesh, flags)
        self.M.assemble(self.fe, capi.coeff(capi.COEFF_CONST, 0.0),
                        capi.coeff(capi.COEFF_CONST, 1.0), None, None)
        self.theta = capi.Vec(ctx, self.ne)
        self.uobs = capi.Vec(ctx, u_obs)
        self.b, self.u, self.d, self.rhs, self.lam = (capi.Vec(ctx, self.nv) for _ in range(5))
        self.grad = capi.Vec(ctx, self.ne)

    def _state(self, theta: np.ndarray):
        """Assemble A(theta), b and solve A u = b; leaves d = u - u_obs and
        rhs = M d on the device."""
        if theta.shape != (self.ne,):
            raise ValueError("theta must have one entry per cell")
        self.theta.upload(theta)
        self.A.assemble(self.fe, capi.coeff(capi.COEFF_EXPCELL, 1.0, self.theta),
                        capi.coeff(capi.COEFF_CONST, 0.0), capi.coeff(capi.COEFF_CONST, self.f),
                        self.b)
        self.u.fill(0.0)
        capi.pcg(self.ctx, self.A.csr, self.b, self.u, rtol=self.rtol)
        capi.axpby(self.ctx, 1.0, self.u, 0.0, self.d)
        capi.axpby(self.ctx, -1.0, self.uobs, 1.0, self.d)
        capi.spmv(self.ctx, self.M.csr, self.d, self.rhs)

    def value(self, theta: np.ndarray):
        """J(theta) and the state u (host copies)."""
        theta = np.ascontiguousarray(theta, dtype=np.float64)
        self._state(theta)
        return 0.5 * capi.dot(self.ctx, self.d, self.rhs), self.u.numpy()

    def gradient(self, theta: np.ndarray) -> np.ndarray:
        """dJ/dtheta by one adjoint solve and the gradient kernel."""
        theta = np.ascontiguousarray(theta, dtype=np.float64)
        self._state(theta)
        self.lam.fill(0.0)
        capi.pcg(self.ctx, self.A.csr, self.rhs, self.lam, rtol=self.rtol)
        capi.adjoint_gradient(self.ctx, self.mesh, self.fe, self.theta, self.u, self.lam,
                              self.grad)
        return self.grad.numpy()

    def __call__(self, theta: torch.Tensor) -> torch.Tensor:
        return _AdjointJ.apply(theta, self)

    def close(self):
        for h in (self.grad, self.lam, self.rhs, self.d, self.u, self.b, self.uobs, self.theta,
                  self.M, self.A, self.fe, self.mesh):
            h.close()


class _AdjointJ(torch.autograd.Function):
    @staticmethod
    def forward(ctx, theta, obj):
        th = theta.detach().to("cpu", torch.float64).contiguous().numpy()
        J, _ = obj.value(th)
        ctx.obj = obj
        ctx.theta = th.copy()
        ctx.device, ctx.dtype = theta.device, theta.dtype
        return torch.tensor(J, dtype=torch.float64, device=theta.device)

    @staticmethod
    def backward(ctx, grad_out):
        # the state is recomputed from the saved theta, so other forward
        # evaluations between this one and its backward do not interfere
        g = torch.from_numpy(ctx.obj.gradient(ctx.theta))
        g = g.to(ctx.device, ctx.dtype) * grad_out.to(ctx.dtype)
        return g, None
