"""Fisher/KPP time stepping on the device: the host side of
projects/fisher_kpp/strangsplitting.{h,cc} over the C ABI.

``assemble_galerkin_matrices`` mirrors ``FisherKPP::assembleGalerkinMatrices``
(strangsplitting.h:123-174) and ``StrangSplit`` mirrors the class of the same
name (strangsplitting.h:28-76, :176-201; strangsplitting.cc:15-94), with the
method names of the reference.  All arithmetic runs in liblfgpu; this module
only wires the plan, the edge terms and the stepper together.
"""
from __future__ import annotations

import numpy as np

from . import capi


def boundary_edge_flags(nedges, cell_e):
    """flagEntitiesOnBoundary(mesh, 1) (mesh/utils/special_entity_sets.cc:
    40-62): an edge is on the boundary iff exactly one cell has it."""
    cnt = np.bincount(np.asarray(cell_e, np.int64).ravel(), minlength=int(nedges))
    return (cnt == 1).astype(np.uint8)


def assemble_galerkin_matrices(ctx, mesh: capi.Mesh, edges: capi.Edges, bd_edge_flags,
                               c: capi.Coeff, h: capi.Coeff, L=None):
    """(A, M) of assembleGalerkinMatrices (strangsplitting.h:123-174) for the
    P1 space: A = ReactionDiffusionElementMatrixProvider(c, 0) plus
    MassEdgeMatrixProvider(h) on the boundary edges, M =
    ReactionDiffusionElementMatrixProvider(0, 1), both from one plan without
    fixed DOFs, returned as owned capi.Csr copies.  The nonlocal loss matrix
    L (a dense N x N matrix subtracted from A in the reference) must be None
    or all zeros."""
    if L is not None and np.any(np.asarray(L) != 0):
        raise NotImplementedError("a nonzero dense loss matrix L is not supported")
    fe = capi.FE(ctx, 1)
    plan = capi.Plan(ctx, mesh, np.zeros(mesh.ndofs, np.uint8))
    try:
        plan.set_edges(edges, bd_edge_flags, None)
        plan.assemble_bd(fe, c, capi.coeff(capi.COEFF_CONST, 0.0), None, h, None, None)
        A = capi.csr_clone(ctx, plan.csr)
        plan.assemble(fe, capi.coeff(capi.COEFF_CONST, 0.0), capi.coeff(capi.COEFF_CONST, 1.0),
                      None, None)
        M = capi.csr_clone(ctx, plan.csr)
    finally:
        plan.close()
        fe.close()
    return A, M


class StrangSplit:
    """FisherKPP::StrangSplit on the device (lfgpu_strang_*)."""

    def __init__(self, ctx, A: capi.Csr, M: capi.Csr, T, m, lam, rtol=1e-10,
                 mode=capi.PCG_FP64, check_every=0):
        self.ctx = ctx
        self.n = A.info()[0]
        self._s = capi.Strang(ctx, A, M, T, m, lam, rtol=rtol, mode=mode,
                              check_every=check_every)
        self.last_info = None

    def diffusionEvolutionOperator(self, firstcall, mu: capi.Vec) -> capi.Vec:
        """strangsplitting.cc:15-50; returns a new device vector."""
        out = capi.Vec(self.ctx, self.n)
        self.last_info = self._s.diffusion(firstcall, mu, out)
        return out

    def Evolution(self, cap: capi.Vec, mu: capi.Vec) -> capi.Vec:
        """strangsplitting.cc:53-94; returns a new device vector."""
        out = capi.Vec(self.ctx, self.n)
        self.last_info = self._s.evolution(cap, mu, out)
        return out

    def close(self):
        self._s.close()
