"""Benchmark of the MI355X-native LehrFEM++ hot path.

Workload (one GPU): reaction-diffusion -div(a grad u) + u = f on the unit
square, 2048 x 2048 structured triangular mesh (8,388,608 cells, 4,198,401
DOFs), P1 Lagrange, FP64 geometry and assembly, a = exp(g) per cell with
g ~ N(0, 1) (numpy default_rng(seed)), f = sin(pi x) sin(pi y), zero
Dirichlet data, Jacobi-PCG to ||r|| / ||b|| <= 1e-10.

One step = assemble the Galerkin matrix and load vector into the cached
pattern (AssembleMatrixLocally + AssembleVectorLocally +
FixFlaggedSolutionComponents + makeSparse) + solve with PCG from x = 0.
The mesh, coefficient field and assembly plan are device resident before the
timed region starts.  value = cells assembled-and-solved per second over the
timed steps (Melem/s).

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--n 2048] [--mode mixed|fp64]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "P1 assembly Melem/s + PCG time to rtol 1e-10 at 4.2M DOF; grad cost / solve"
HBM_PEAK_GBS = 8000.0
MAXIT = 40000  # a few seconds at the largest size; the solves converge far earlier


def log(msg):
    print("[bench] " + msg, file=sys.stderr, flush=True)


def boundary_flags(n):
    j, i = np.divmod(np.arange((n + 1) * (n + 1)), n + 1)
    return ((i == 0) | (i == n) | (j == 0) | (j == n)).astype(np.uint8)


class Problem:
    """Device-resident P1 reaction-diffusion problem on an n x n mesh."""

    def __init__(self, capi, ctx, n, seed, gamma=1.0, f_sinsin=True, log_alpha=None):
        self.capi, self.ctx, self.n = capi, ctx, n
        xy, cells, _ = capi.tp_triag_mesh(n)
        self.ne, self.nv = cells.shape[0], xy.shape[0]
        g = (np.random.default_rng(seed).standard_normal(self.ne)
             if log_alpha is None else log_alpha)
        self.log_alpha = g
        self.flags = boundary_flags(n)
        self.mesh = capi.Mesh(ctx, xy, cells)
        self.fe = capi.FE(ctx, 1)
        t0 = time.perf_counter()
        self.plan = capi.Plan(ctx, self.mesh, self.flags)
        ctx.sync()
        self.plan_s = time.perf_counter() - t0
        self.theta = capi.Vec(ctx, g)
        self.alpha = capi.coeff(capi.COEFF_EXPCELL, 1.0, self.theta)
        self.gamma = capi.coeff(capi.COEFF_CONST, gamma)
        self.f = capi.coeff(capi.COEFF_SINSIN if f_sinsin else capi.COEFF_CONST, 1.0)
        self.b = capi.Vec(ctx, self.nv)
        self.x = capi.Vec(ctx, self.nv)
        self.nnz = self.plan.csr.info()[2]

    def assemble(self):
        return self.plan.assemble(self.fe, self.alpha, self.gamma, self.f, self.b)

    def solve(self, mode, rtol=1e-10, x=None, b=None):
        x = self.x if x is None else x
        x.fill(0.0)
        return self.capi.pcg(self.ctx, self.plan.csr, self.b if b is None else b, x, rtol=rtol,
                             mode=mode, check_every=128, maxit=MAXIT, raise_on_noconv=False)

    def close(self):
        for h in (self.x, self.b, self.theta, self.plan, self.fe, self.mesh):
            h.close()


def spmv_bytes(nnz, n, val_bytes=8, vec_bytes=8):
    return (val_bytes + 4) * nnz + 4 * (n + 1) + 2 * vec_bytes * n


def run_steps(P, mode, steps):
    asm_ms, pcg_ms, iters = [], [], []
    P.ctx.sync()
    t0 = time.perf_counter()
    for _ in range(steps):
        asm_ms.append(P.assemble())
        info = P.solve(mode)
        pcg_ms.append(info.ms)
        iters.append(info.iters)
        if not info.converged:
            raise RuntimeError("PCG did not converge")
    P.ctx.sync()
    wall = time.perf_counter() - t0
    return wall, asm_ms, pcg_ms, iters, info


def cpu_baseline(P, gpu_iters, capi):
    """Reference-faithful CPU path (oracle, one thread) on a bounded sample."""
    from oracle import lfref as L  # the checker; timed here as the CPU baseline
    ns = 512
    m = L.tp_triag_mesh(ns)
    cd, nd = m.dofs(1)
    a = np.exp(np.random.default_rng(0).standard_normal(m.ne))
    fl = m.dirichlet_flags(1)
    t0 = time.perf_counter()
    L.cpu_path_assemble(m, 1, cd, nd, a, 1.0, L.F_SINSIN, 0.0, fl)
    t_cell = (time.perf_counter() - t0) / m.ne
    log("cpu assembly sample: %d cells, %.2f us/cell" % (m.ne, 1e6 * t_cell))
    rp, col, val = P.plan.csr.download()
    b = P.b.numpy()
    ts = []
    for k in (5, 15):
        t0 = time.perf_counter()
        L.pcg(rp, col, val, b, tol=1e-10, maxit=k)
        ts.append(time.perf_counter() - t0)
    t_iter = (ts[1] - ts[0]) / 10
    log("cpu pcg sample: %.1f ms/iteration at N=%d" % (1e3 * t_iter, P.nv))
    step_s = t_cell * P.ne + t_iter * gpu_iters
    return {
        "value": P.ne / step_s / 1e6,
        "unit": "Melem/s",
        "cores": 1,
        "kind": "port",
        "sample": ("assembly timed on a 512x512 mesh (524,288 cells) and scaled per cell to "
                   "%d cells; Jacobi-PCG timed for 10 iterations on the full %d-DOF system "
                   "and scaled to the GPU's %d iterations (extrapolated)"
                   % (P.ne, P.nv, gpu_iters)),
        "assembly_us_per_cell": 1e6 * t_cell,
        "pcg_ms_per_iter": 1e3 * t_iter,
    }


def adjoint_cost(capi, ctx, n=1024, seed=1):
    """cost of dJ/dtheta (adjoint PCG + gradient kernel) / forward solve."""
    P = Problem(capi, ctx, n, seed, gamma=0.0, f_sinsin=False)
    P.assemble()
    uobs = capi.Vec(ctx, P.nv)
    info_f = P.solve(capi.PCG_FP64)
    # adjoint rhs M (u - u_obs) with u_obs = 0 here: the cost does not depend on it
    Mplan_b = capi.Vec(ctx, P.nv)
    lam = capi.Vec(ctx, P.nv)
    capi.axpby(ctx, 1.0, P.x, 0.0, Mplan_b)
    grad = capi.Vec(ctx, P.ne)
    t_solve = []
    for _ in range(2):
        info_f = P.solve(capi.PCG_FP64)
        t_solve.append(info_f.ms)
    t_adj = []
    for _ in range(2):
        lam.fill(0.0)
        ia = capi.pcg(ctx, P.plan.csr, Mplan_b, lam, rtol=1e-10, check_every=128, maxit=MAXIT)
        tg = capi.adjoint_gradient(ctx, P.mesh, P.fe, P.theta, P.x, lam, grad)
        t_adj.append(ia.ms + tg)
    for h in (uobs, Mplan_b, lam, grad):
        h.close()
    P.close()
    return min(t_adj) / min(t_solve)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=2048)
    ap.add_argument("--mode", choices=["mixed", "fp64"], default="fp64")
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--spmv-reps", type=int, default=50)
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path "
              "(see DESIGN.md for how the path would shard); run with --gpus 1",
              file=sys.stderr)
        sys.exit(2)

    from lehrfempp_amd import capi
    ctx = capi.Context(0)
    log("setup n=%d" % args.n)
    P = Problem(capi, ctx, args.n, args.seed)
    log("plan built in %.3f s: N=%d nnz=%d" % (P.plan_s, P.nv, P.nnz))
    mode = capi.PCG_MIXED if args.mode == "mixed" else capi.PCG_FP64

    for _ in range(args.warmup):
        P.assemble()
        P.solve(mode)
    log("warmup done")
    wall, asm_ms, pcg_ms, iters, info = run_steps(P, mode, args.steps)
    log("timed %d steps: %.3f s (assembly %.3f ms, pcg %.1f ms, %d iters, relres %.2e)"
        % (args.steps, wall, np.median(asm_ms), np.median(pcg_ms), iters[-1], info.relres))

    # the other precision, for the mixed-vs-FP64 throughput ratio
    other = capi.PCG_FP64 if mode == capi.PCG_MIXED else capi.PCG_MIXED
    info_o = P.solve(other)
    log("other mode: %.1f ms, %d iters, relres %.2e" % (info_o.ms, info_o.iters, info_o.relres))
    fp64_ms = info_o.ms if mode == capi.PCG_MIXED else float(np.median(pcg_ms))
    mixed_ms = float(np.median(pcg_ms)) if mode == capi.PCG_MIXED else info_o.ms

    # roofline of the dominant kernel: the FP64 sliced-ELL SpMV
    xs, ys = capi.Vec(ctx, np.random.default_rng(1).standard_normal(P.nv)), capi.Vec(ctx, P.nv)
    spmv_ms = capi.spmv_timed(ctx, P.plan.csr, xs, ys, args.spmv_reps)
    sb = spmv_bytes(P.nnz, P.nv)
    achieved = sb / (spmv_ms * 1e-3) / 1e9
    log("spmv: %.1f us, %.0f GB/s" % (1e3 * spmv_ms, achieved))
    xs.close()
    ys.close()

    grad_ratio = None
    if not args.no_adjoint:
        grad_ratio = adjoint_cost(capi, ctx)
        log("adjoint gradient cost / forward solve = %.2f" % grad_ratio)

    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(P, iters[-1], capi)

    asm_bytes = 12 * P.ne + 8 * P.ne + 16 * P.nv + 8 * P.nnz + 8 * P.nv
    out = {
        "metric": METRIC,
        "value": P.ne * args.steps / wall / 1e6,
        "unit": "Melem/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * wall / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if mode == capi.PCG_FP64 else "f64 assembly, f32 inner PCG + f64 refinement",
        "data": "synthetic",
        "config": {"workload": "P1 reaction-diffusion, %dx%d structured triangles, "
                               "a=exp(N(0,1)) per cell, f=sin(pi x)sin(pi y), zero Dirichlet, "
                               "assemble + Jacobi-PCG to rtol 1e-10" % (args.n, args.n),
                   "cells": P.ne, "dofs": P.nv, "nnz": P.nnz, "pcg_mode": args.mode},
        "assembly_melem_s": P.ne / (np.median(asm_ms) * 1e-3) / 1e6,
        "assembly_ms": float(np.median(asm_ms)),
        "assembly_gbs_algorithmic": asm_bytes / (np.median(asm_ms) * 1e-3) / 1e9,
        "pcg_s": float(np.median(pcg_ms)) / 1e3,
        "pcg_iters": int(iters[-1]),
        "pcg_relres": info.relres,
        "pcg_fp64_s": fp64_ms / 1e3,
        "pcg_mixed_s": mixed_ms / 1e3,
        "mixed_speedup_vs_fp64": fp64_ms / mixed_ms,
        "dof_per_s": P.nv / (float(np.median(pcg_ms)) * 1e-3),
        "grad_cost_per_solve": grad_ratio,
        "plan_build_s": P.plan_s,
        "roofline": {"bound": "hbm", "kernel": "k_spmv64 (FP64 sliced-ELL SpMV)",
                     "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "algorithmic_bytes": sb, "avg_us": 1e3 * spmv_ms},
        "cpu_baseline": cpu,
    }
    P.close()
    ctx.close()
    print(json.dumps(out))


if __name__ == "__main__":
    main()
