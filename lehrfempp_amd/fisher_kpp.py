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


def nonlocal_gain(ctx, edges: capi.Edges, bd_edge_flags, rmin=0.0, rmax=0.0):
    """The gain h of the drivers' nonlocal boundary condition
    (modelproblem_island_main.cc:70-90: localQuadFunction of g(x, .) over the
    boundary edges, strangsplitting.h:78-115) at the edge points, as the
    LFGPU_COEFF_QP coefficient of the edge mass term; g is 1 / (1 + |x-y|^2),
    cut to rmin <= |x - y| <= rmax when rmax > 0.  Returns (coefficient,
    device vector); keep the vector while the coefficient is used."""
    eta = capi.nonlocal_gain(ctx, edges, bd_edge_flags, rmin, rmax)
    return capi.coeff(capi.COEFF_QP, 1.0, eta), eta


def nonlocal_loss(ctx, edges: capi.Edges, bd_edge_flags, rmin=0.0, rmax=0.0):
    """The loss matrix L of the drivers (modelproblem_island_main.cc:92-126)
    on the DOFs of the boundary edges: (rows, device vector of the nb x nb
    block), the form assemble_galerkin_matrices takes for L."""
    return capi.nonlocal_loss(ctx, edges, bd_edge_flags, rmin, rmax)


def _loss_block(ctx, L, n):
    """(rows, device block, owned) of L given as (rows, block) or as the
    reference's dense N x N matrix (owned: the block vector was made here);
    None when L is absent or zero."""
    if L is None:
        return None
    if isinstance(L, tuple):
        rows, B = L
        rows = np.ascontiguousarray(rows, np.int32)
        if isinstance(B, capi.Vec):
            return rows, B, False
        return rows, capi.Vec(ctx, np.ascontiguousarray(B, np.float64).ravel()), True
    Ld = np.asarray(L, np.float64)
    if Ld.shape != (n, n):
        raise ValueError("L must be N x N, or (rows, block)")
    rows = np.nonzero(np.any(Ld != 0, axis=0) | np.any(Ld != 0, axis=1))[0].astype(np.int32)
    if rows.size == 0:
        return None
    return rows, capi.Vec(ctx, np.ascontiguousarray(Ld[np.ix_(rows, rows)]).ravel()), True


def assemble_galerkin_matrices(ctx, mesh: capi.Mesh, edges: capi.Edges, bd_edge_flags,
                               c: capi.Coeff, h: capi.Coeff, L=None):
    """(A, M) of assembleGalerkinMatrices (strangsplitting.h:123-174) for the
    P1 space: A = ReactionDiffusionElementMatrixProvider(c, 0) plus
    MassEdgeMatrixProvider(h) on the boundary edges, minus the nonlocal loss
    L, and M = ReactionDiffusionElementMatrixProvider(0, 1), both from one
    plan without fixed DOFs, returned as owned capi.Csr copies.  L is None,
    the reference's dense N x N matrix, or (rows, block) as nonlocal_loss
    returns it; it is attached to A as a dense block with factor -1 (the
    reference's A_ds - L, :166-170), which the FP64 solver and the time
    stepper honour."""
    blk = _loss_block(ctx, L, mesh.ndofs)
    fe = capi.FE(ctx, 1)
    plan = capi.Plan(ctx, mesh, np.zeros(mesh.ndofs, np.uint8))
    try:
        plan.set_edges(edges, bd_edge_flags, None)
        plan.assemble_bd(fe, c, capi.coeff(capi.COEFF_CONST, 0.0), None, h, None, None)
        A = capi.csr_clone(ctx, plan.csr)
        plan.assemble(fe, capi.coeff(capi.COEFF_CONST, 0.0), capi.coeff(capi.COEFF_CONST, 1.0),
                      None, None)
        M = capi.csr_clone(ctx, plan.csr)
        if blk is not None:
            capi.csr_set_dense_block(ctx, A, blk[0], blk[1], -1.0)
    finally:
        plan.close()
        fe.close()
        if blk is not None and blk[2]:
            blk[1].close()
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
