"""Benchmark of the MI355X-native LehrFEM++ hot path (one GPU).

Default workload (--config p1): reaction-diffusion -div(a grad u) + u = f on
the unit square, 2048 x 2048 structured triangular mesh (8,388,608 cells,
4,198,401 DOFs), P1 Lagrange, FP64 geometry and assembly, a = exp(g) per cell
with g ~ N(0, 1) (numpy default_rng(seed)), f = sin(pi x) sin(pi y), zero
Dirichlet data, Jacobi-PCG to ||r|| / ||b|| <= 1e-10 (Eigen ConjugateGradient
semantics).  One step = assemble the Galerkin matrix and load vector into
the cached pattern + solve from x = 0.  Mesh, coefficient field and assembly
plan are device resident before the timed region.  value = cells assembled
and solved per second over the timed steps (Melem/s).

The default line also carries, as extra keys, the P3 workload (c3) and the
batched workload (c4) below, each precision mode of every solve, the
roofline of the dominant kernels (every sweep launch of 200 PCG iterations
timed by its own hipEvent pair), and, after all device work has finished,
the oracle checks at full size: the oracle (CPU restatement of the
reference pipeline) builds its own mesh, assembles the same problem in the
reference's cell order, compares the pattern bit for bit and the values
entry by entry with the device CSR, solves it by the textbook Jacobi-PCG to
1e-12, and every device solution reports its error against that solution
(err_vs_oracle).  --c2-only / --no-oracle drop those parts.

Other configurations (each prints its own JSON line):
  --config p3       P3 on a 512 x 512 mesh with interior vertices jittered by
                    U(-0.2h, 0.2h), degree-6 quadrature, PCG in FP64 and mixed
  --config batched  256 independent P1 problems, sides log-uniform in
                    [32, 512], a = exp(N(0,1)), f = 1, one assembly launch and
                    one graph-replayed batched PCG
  --config adjoint  1024 x 1024 P1, J = 1/2 (u - u_obs)^T M (u - u_obs), dJ/dtheta
                    by an adjoint solve, checked against central differences
  --config fkpp     Fisher/KPP Strang splitting (SDIRK-2 diffusion + exact
                    logistic reaction) on a 1024 x 1024 mesh, 100 time steps
  --config c1       the 64 x 64 P1 Poisson problem: absolute assembly, SpMV
                    and PCG times (launch-latency bound)
  --config strip    one GPU's share of the default workload split into 8
                    strips (2048 x 256 cells): assembly, SpMV and PCG
                    iteration times next to 1/8 of the full-size ones

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--config p1]
"""
from __future__ import annotations

import argparse
import json
import os
import re
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "P1 assembly Melem/s + PCG time to rtol 1e-10 at 4.2M DOF; grad cost / solve"
HBM_PEAK_GBS = 8000.0
MAXIT = 60000  # a few seconds at the largest size; the solves converge far earlier


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
        self.xy, self.cells = xy, cells
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
    """Algorithmic bytes of one CSR SpMV: values + int32 columns + row
    pointers + x read once + y written once."""
    return (val_bytes + 4) * nnz + 4 * (n + 1) + 2 * vec_bytes * n


def sweep_bytes(mode, n, sym):
    """Bytes one sweep of the DIA Jacobi-PCG must move (each array read or
    written once; the neighbours' p values and, with `sym`, the mirrored
    lower-diagonal values are re-reads that the caches serve).  P1 DIA
    storage: 6 off-diagonals, the row sums and the inverse diagonal.
      FP64 sweep 1 (k_cg_spmv):  p, r, D^-1, row sums and 3 (sym) or 6
                                 diagonals read, q written: 64N or 88N
      FP64 sweep 2 (k_cg_update): p, r, D^-1, x, q read, x, r, p written: 64N
      FP32-Krylov sweep 1 (k_mx_a): p, r, row sums, 6 diagonals read (4 B
                                 each), D^-1 read (bfloat16), q written: 42N
      FP32-Krylov sweep 2 (k_mx_b): p, r, increment, q read, D^-1 read
                                 (bfloat16), increment, r, p written: 30N"""
    if mode == "fp64":
        return (32 + 8 * (3 if sym else 6) + 8) * n, 64 * n
    return 42 * n, 30 * n


def pmc_traffic(kernel, name="pmc_traffic_p1_2048.md"):
    """HBM bytes per dispatch of `kernel` recorded in the committed rocprofv3
    --pmc summary (FETCH_SIZE + WRITE_SIZE of the same dispatches, each scaled
    by the copy-kernel calibration written next to it), or None."""
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles", name)
    try:
        text = open(path).read()
    except OSError:
        return None
    m = re.search(r"^traffic_bytes_per_dispatch %s: (\d+)" % re.escape(kernel), text, re.M)
    return int(m.group(1)) if m else None


def roofline_entry(n, times_ms, sym, mode="fp64"):
    """Roofline of the dominant kernel of the timed step: the first sweep of
    the Jacobi-PCG iteration (q = A p with the three dot products), achieved
    = its algorithmic bytes (sweep_bytes) / its median duration between
    hipEvents on the solver's stream; the second sweep is reported beside
    it.  traffic = HBM bytes per dispatch counted by rocprofv3 (FETCH_SIZE +
    WRITE_SIZE, calibrated on copy kernels) from the committed summary."""
    b1, b2 = sweep_bytes(mode, n, sym)
    k1 = "k_cg_spmv" if mode == "fp64" else "k_mx_a"
    k2 = "k_cg_update" if mode == "fp64" else "k_mx_b"
    a1 = b1 / (times_ms[0] * 1e-3) / 1e9
    a2 = b2 / (times_ms[1] * 1e-3) / 1e9
    it_us = 1e3 * (times_ms[0] + times_ms[1])
    out = {"bound": "hbm", "kernel": "%s (Jacobi-PCG sweep 1: q = A p and 3 dot products, "
                                     "DIA storage%s)" % (k1, ", upper diagonals" if sym else ""),
           "achieved": a1, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": a1 / HBM_PEAK_GBS,
           "traffic": pmc_traffic(k1), "algorithmic_bytes": b1, "avg_us": 1e3 * times_ms[0],
           "timing": "median of per-launch hipEvent pairs on the solver stream",
           "sweep2": {"kernel": k2, "algorithmic_bytes": b2, "avg_us": 1e3 * times_ms[1],
                      "achieved": a2, "frac": a2 / HBM_PEAK_GBS, "traffic": pmc_traffic(k2)},
           "iteration": {"us": it_us, "algorithmic_bytes": b1 + b2,
                         "achieved": (b1 + b2) / (it_us * 1e-6) / 1e9,
                         "frac": (b1 + b2) / (it_us * 1e-6) / 1e9 / HBM_PEAK_GBS}}
    return out


def ell_row_sweep_bytes(n, nnz, col_bytes):
    """Bytes one FP64 row sweep of the sliced-ELL Jacobi-PCG must move (each
    array once; the off-diagonal entries carry a value and a column, padding
    slots are excluded; the gathered p values are re-reads the caches serve).
      k_cgr_spmv:   p, r, D^-1, row sums read, q written (40 B per row) and
                    8 + col_bytes per off-diagonal entry (nnz - N of them:
                    every row stores its diagonal in the CSR)
      k_cgr_update: p, r, D^-1, x, q read, x, r, p written: 64 B per row"""
    return 40 * n + (8 + col_bytes) * (nnz - n), 64 * n


def ell_roofline_entry(n, nnz, times_ms, col_bytes):
    b1, b2 = ell_row_sweep_bytes(n, nnz, col_bytes)
    it_us = 1e3 * (times_ms[0] + times_ms[1] + 2 * times_ms[2])
    a1 = b1 / (times_ms[0] * 1e-3) / 1e9
    a2 = b2 / (times_ms[1] * 1e-3) / 1e9
    return {"bound": "hbm", "kernel": "k_cgr_spmv (sliced-ELL Jacobi-PCG row sweep 1: q = A p "
                                      "and 3 dot products, %d-bit column offsets)"
                                      % (8 * col_bytes),
            "achieved": a1, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": a1 / HBM_PEAK_GBS,
            "traffic": pmc_traffic("k_cgr_spmv") if col_bytes == 2 else None,
            "algorithmic_bytes": b1, "avg_us": 1e3 * times_ms[0],
            "timing": "median of per-launch hipEvent pairs on the solver stream",
            "sweep2": {"kernel": "k_cgr_update", "algorithmic_bytes": b2,
                       "avg_us": 1e3 * times_ms[1], "achieved": a2, "frac": a2 / HBM_PEAK_GBS,
                       "traffic": pmc_traffic("k_cgr_update")},
            "reductions_us": 2e3 * times_ms[2],
            "iteration": {"us": it_us, "algorithmic_bytes": b1 + b2,
                          "achieved": (b1 + b2) / (it_us * 1e-6) / 1e9,
                          "frac": (b1 + b2) / (it_us * 1e-6) / 1e9 / HBM_PEAK_GBS}}


def true_relres(capi, ctx, A, x, b):
    y = capi.Vec(ctx, b.n)
    capi.spmv(ctx, A, x, y)
    r = b.numpy() - y.numpy()
    y.close()
    return float(np.linalg.norm(r) / np.linalg.norm(b.numpy()))


def rel_err(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def row_scaled_err(rp, a, b):
    """max over rows of max |a - b| / max |b| within the row (CSR values)."""
    starts = rp[:-1][np.diff(rp) > 0]
    scale = np.maximum(np.maximum.reduceat(np.abs(b), starts), 1e-300)
    return float((np.maximum.reduceat(np.abs(a - b), starts) / scale).max())


# ---------------------------------------------------------------- configs
def cpu_baseline_p1(ne, nv, rp, col, val, b, gpu_iters, oracle=None):
    """Reference-faithful CPU path (the oracle, one thread): the assembly of
    the full problem when the oracle check ran (else a 512 x 512 sample
    scaled per cell) plus Jacobi-PCG per-iteration times scaled to the GPU's
    iteration count (one thread: 10 iterations timed; all threads: the
    oracle's full solve)."""
    L = oracle_lib()
    if oracle is not None:
        t_cell = oracle["assembly"]["assembly_s"] / ne
        asm_sample = "the full %d-cell assembly (one thread)" % ne
    else:
        m = L.tp_triag_mesh(512)
        cd, nd = m.dofs(1)
        a = np.exp(np.random.default_rng(0).standard_normal(m.ne))
        fl = m.dirichlet_flags(1)
        t0 = time.perf_counter()
        L.cpu_path_assemble(m, 1, cd, nd, a, 1.0, L.F_SINSIN, 0.0, fl)
        t_cell = (time.perf_counter() - t0) / m.ne
        asm_sample = ("a 512x512 sample (524,288 cells) scaled per cell to %d cells"
                      % ne)
    log("cpu assembly: %.2f us/cell" % (1e6 * t_cell))
    ts = []
    for k in (5, 15):
        t0 = time.perf_counter()
        L.pcg(rp, col, val, b, tol=1e-10, maxit=k)
        ts.append(time.perf_counter() - t0)
    t_iter = (ts[1] - ts[0]) / 10
    log("cpu pcg sample: %.1f ms/iteration at N=%d" % (1e3 * t_iter, nv))
    step_s = t_cell * ne + t_iter * gpu_iters
    threads = L.omp_threads()
    if oracle is not None:
        t_iter_mt = oracle["solve"]["s"] / max(1, oracle["solve"]["iters"])
        mt_note = "the oracle's full solve to %g (%d iterations)" % (
            ORACLE_RTOL, oracle["solve"]["iters"])
    else:
        ts = []
        for k in (10, 50):
            t0 = time.perf_counter()
            L.pcg_omp(rp, col, val, b, tol=1e-10, maxit=k)
            ts.append(time.perf_counter() - t0)
        t_iter_mt = (ts[1] - ts[0]) / 40
        mt_note = "40 iterations"
    log("cpu pcg on %d threads: %.2f ms/iteration" % (threads, 1e3 * t_iter_mt))
    step_mt = t_cell * ne + t_iter_mt * gpu_iters
    return {
        "value": ne / step_s / 1e6,
        "unit": "Melem/s",
        "cores": 1,
        "kind": "port",
        "sample": ("assembly (Eval per cell, triplets, Dirichlet fix, setFromTriplets): %s; "
                   "Jacobi-PCG timed for 10 iterations on the full %d-DOF system and scaled "
                   "to the GPU's %d iterations (extrapolated)" % (asm_sample, nv, gpu_iters)),
        "assembly_us_per_cell": 1e6 * t_cell,
        "pcg_ms_per_iter": 1e3 * t_iter,
        "pcg_all_cores": {
            "threads": threads, "pcg_ms_per_iter": 1e3 * t_iter_mt,
            "value": ne / step_mt / 1e6,
            "note": "assembly on one thread as above (the reference's cell loop is serial), "
                    "Jacobi-PCG with OpenMP loops and reductions, per-iteration time from %s, "
                    "scaled to the GPU's iteration count" % mt_note},
    }


def adjoint_run(capi, ctx, n=1024, fd_dirs=0, seed_theta=1, seed_star=2):
    """Adjoint gradient of J on an n x n P1 mesh; returns the cost
    ratio (adjoint solve + gradient kernel) / forward solve and FD errors."""
    rng = np.random.default_rng(seed_theta)
    P = Problem(capi, ctx, n, seed_theta, gamma=0.0, f_sinsin=False)
    theta = P.log_alpha.copy()
    theta_star = np.random.default_rng(seed_star).standard_normal(P.ne)
    M = capi.Plan(ctx, P.mesh, P.flags)
    M.assemble(P.fe, capi.coeff(capi.COEFF_CONST, 0.0), capi.coeff(capi.COEFF_CONST, 1.0),
               None, None)
    # observations: u(theta*) + 1 % Gaussian noise (relative to max|u|) on interior DOFs
    P.theta.upload(theta_star)
    P.assemble()
    P.solve(capi.PCG_FP64, rtol=1e-12)
    ustar = P.x.numpy()
    noise = np.random.default_rng(seed_star + 1).standard_normal(P.nv) * (1 - P.flags)
    uobs = capi.Vec(ctx, ustar + 0.01 * np.abs(ustar).max() * noise)
    d, rhs, lam, grad = (capi.Vec(ctx, P.nv), capi.Vec(ctx, P.nv), capi.Vec(ctx, P.nv),
                         capi.Vec(ctx, P.ne))

    def J(th, rtol=1e-10):
        P.theta.upload(th)
        P.assemble()
        info = P.solve(capi.PCG_FP64, rtol=rtol)
        capi.axpby(ctx, 1.0, P.x, 0.0, d)
        capi.axpby(ctx, -1.0, uobs, 1.0, d)
        capi.spmv(ctx, M.csr, d, rhs)
        return 0.5 * capi.dot(ctx, d, rhs), info

    t_fwd, t_adj = [], []
    for _ in range(3):
        _, info = J(theta)
        t_fwd.append(info.ms)
        lam.fill(0.0)
        ia = capi.pcg(ctx, P.plan.csr, rhs, lam, rtol=1e-10, check_every=128, maxit=MAXIT)
        tg = capi.adjoint_gradient(ctx, P.mesh, P.fe, P.theta, P.x, lam, grad)
        t_adj.append(ia.ms + tg)
    ratio = float(np.median(t_adj) / np.median(t_fwd))
    out = {"grad_cost_per_solve": ratio, "forward_solve_ms": float(np.median(t_fwd)),
           "adjoint_ms": float(np.median(t_adj)), "forward_iters": info.iters,
           "adjoint_iters": ia.iters}
    if fd_dirs:
        # gradient with tight solves, then central differences with tight solves
        J0, _ = J(theta, 1e-13)
        lam.fill(0.0)
        capi.pcg(ctx, P.plan.csr, rhs, lam, rtol=1e-13, check_every=128, maxit=MAXIT)
        capi.adjoint_gradient(ctx, P.mesh, P.fe, P.theta, P.x, lam, grad)
        gr = grad.numpy()
        errs, errs_g = [], []
        eps = 1e-5
        for k in range(fd_dirs):
            dv = rng.standard_normal(P.ne)
            dv /= np.linalg.norm(dv)
            fd = (J(theta + eps * dv, 1e-13)[0] - J(theta - eps * dv, 1e-13)[0]) / (2 * eps)
            errs.append(abs(fd - gr @ dv) / max(abs(fd), 1e-300))
            errs_g.append(abs(fd - gr @ dv) / np.linalg.norm(gr))
            if (k + 1) % 8 == 0:
                log("fd direction %d: max rel err so far %.2e" % (k + 1, max(errs)))
        # along the gradient itself the derivative is |grad J| rather than
        # ~|grad J| / sqrt(Ne), so the solves' noise matters far less there
        gn = gr / np.linalg.norm(gr)
        fdg = (J(theta + eps * gn, 1e-13)[0] - J(theta - eps * gn, 1e-13)[0]) / (2 * eps)
        err_g = abs(fdg - gr @ gn) / abs(fdg)
        log("fd along the gradient: rel err %.2e" % err_g)
        out.update({"J": J0, "fd_directions": fd_dirs, "fd_max_rel_err": float(max(errs)),
                    "fd_median_rel_err": float(np.median(errs)),
                    "fd_max_err_over_grad_norm": float(max(errs_g)),
                    "fd_rel_err_gradient_direction": float(err_g),
                    "grad_norm": float(np.linalg.norm(gr))})
    for h in (uobs, d, rhs, lam, grad, M):
        h.close()
    P.close()
    return out


# ---------------------------------------------------------------- the oracle
# The CPU restatement of the reference pipeline is the checker of the GPU
# results (never timed as the product): it builds its own mesh, assembles the
# same problem with the reference's per-cell loop (one thread) and solves it
# with the textbook Jacobi-PCG (Eigen semantics, plain CSR product) to 1e-12.
ORACLE_RTOL = 1e-12


def oracle_lib():
    # the OpenMP threads of the oracle's PCG sleep between parallel regions
    # (read when its library loads)
    os.environ.setdefault("OMP_WAIT_POLICY", "PASSIVE")
    from oracle import lfref as L  # CPU restatement: the checker
    return L


def oracle_compare_csr(rp_o, col_o, val_o, b_o, rp_g, col_g, val_g, b_g):
    same = bool(np.array_equal(rp_o, rp_g) and np.array_equal(col_o, col_g))
    return {"pattern_bitwise_equal": same,
            "matrix_row_scaled_diff": row_scaled_err(rp_o, val_g, val_o) if same else None,
            "matrix_bitwise_equal": bool(same and np.array_equal(val_o, val_g)),
            "rhs_max_rel_diff": float(np.abs(b_g - b_o).max() / max(np.abs(b_o).max(), 1e-300))}


def oracle_solve(L, rp, col, val, b, threads=True):
    t0 = time.perf_counter()
    if threads:
        x, it, err = L.pcg_omp(rp, col, val, b, tol=ORACLE_RTOL, maxit=400000)
    else:
        x, it, err = L.pcg(rp, col, val, b, tol=ORACLE_RTOL, maxit=400000)
    dt = time.perf_counter() - t0
    r = b - L.csr_spmv(rp, col, val, x)
    return x, {"solver": "Jacobi-PCG, Eigen ConjugateGradient semantics, plain CSR product, "
                         "%s" % ("OpenMP on %d threads" % L.omp_threads() if threads
                                 else "one thread"),
               "rtol": ORACLE_RTOL, "iters": int(it), "relres": float(err),
               "true_relres": float(np.linalg.norm(r) / np.linalg.norm(b)), "s": dt}


def oracle_p1(n, theta, gamma, f_sinsin):
    """Oracle assembly of the n x n P1 problem with alpha = exp(theta)."""
    L = oracle_lib()
    t0 = time.perf_counter()
    m = L.tp_triag_mesh(n)
    cd, nd = m.dofs(1)
    fl = m.dirichlet_flags(1)
    t1 = time.perf_counter()
    rp, col, val, b = L.cpu_path_assemble(m, 1, cd, nd, np.exp(theta), gamma,
                                          L.F_SINSIN if f_sinsin else L.F_CONST, 1.0, fl)
    t2 = time.perf_counter()
    return L, (rp, col, val, b), {"mesh_s": t1 - t0, "assembly_s": t2 - t1}


def run_p1(args, capi, ctx):
    log("setup n=%d" % args.n)
    P = Problem(capi, ctx, args.n, args.seed)
    log("plan built in %.3f s: N=%d nnz=%d" % (P.plan_s, P.nv, P.nnz))
    modes = {"fp64": capi.PCG_FP64, "mixed": capi.PCG_MIXED,
             "mixed_matrix": capi.PCG_MIXED_MATRIX}
    mode = modes[args.mode]
    for _ in range(args.warmup):
        P.assemble()
        P.solve(mode)
    log("warmup done")
    asm_ms, pcg_ms, iters = [], [], []
    ctx.sync()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        asm_ms.append(P.assemble())
        info = P.solve(mode)
        pcg_ms.append(info.ms)
        iters.append(info.iters)
        if not info.converged:
            raise RuntimeError("PCG did not converge")
    ctx.sync()
    wall = time.perf_counter() - t0
    log("timed %d steps: %.3f s (assembly %.3f ms, pcg %.1f ms, %d iters, relres %.2e)"
        % (args.steps, wall, np.median(asm_ms), np.median(pcg_ms), iters[-1], info.relres))

    # every precision mode
    per_mode = {}
    for name, md in modes.items():
        inf = P.solve(md)
        per_mode[name] = {"s": inf.ms / 1e3, "iters": inf.iters, "converged": inf.converged,
                          "relres": inf.relres,
                          "true_relres": true_relres(capi, ctx, P.plan.csr, P.x, P.b),
                          "x": P.x.numpy()}
        log("%s: %.1f ms, %d iters, relres %.2e, true relres %.2e"
            % (name, inf.ms, inf.iters, inf.relres, per_mode[name]["true_relres"]))

    # assembly rate: median over the timed steps and ten more launches
    asm_ms = asm_ms + [P.assemble() for _ in range(10)]

    # the two-phase assembly (matrix-core element matrices, then the gather)
    P.plan.assemble(P.fe, P.alpha, P.gamma, P.f, P.b, mfma=True)
    asm_mfma = [P.plan.assemble(P.fe, P.alpha, P.gamma, P.f, P.b, mfma=True) for _ in range(3)]
    P.assemble()
    log("assembly: row-owner %.3f ms, MFMA two-phase %.3f ms"
        % (np.median(asm_ms), np.median(asm_mfma)))

    # roofline of the dominant kernels: the two sweeps of the FP64 iteration,
    # each launch timed between its own hipEvents (and the FP32-Krylov sweeps)
    _, _, sym = capi.solver_storage(P.plan.csr)
    t64, _ = capi.pcg_sweep_times(ctx, P.plan.csr, P.b, P.x, iters=args.sweep_iters)
    t32, _ = capi.pcg_sweep_times(ctx, P.plan.csr, P.b, P.x, iters=args.sweep_iters,
                                  mode=capi.PCG_MIXED)
    roof = roofline_entry(P.nv, t64, sym)
    roof_mixed = roofline_entry(P.nv, t32, False, mode="mixed")
    log("sweeps fp64: %.1f + %.1f us (frac %.3f / %.3f); fp32-Krylov %.1f + %.1f us"
        % (1e3 * t64[0], 1e3 * t64[1], roof["frac"], roof["sweep2"]["frac"],
           1e3 * t32[0], 1e3 * t32[1]))

    xs = {name: per_mode[name].pop("x") for name in modes}

    adj = None
    if not args.no_adjoint:
        adj = adjoint_run(capi, ctx)
        log("adjoint gradient cost / forward solve = %.2f" % adj["grad_cost_per_solve"])

    asm_bytes = 12 * P.ne + 8 * P.ne + 16 * P.nv + 8 * P.nnz + 8 * P.nv
    fp64_s = per_mode["fp64"]["s"]
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
        "dtype": {"fp64": "f64", "mixed": "f32 Krylov + f64 refinement",
                  "mixed_matrix": "f32 matrix + f64 Krylov"}[args.mode],
        "data": "synthetic",
        "config": {"workload": "P1 reaction-diffusion, %dx%d structured triangles, "
                               "a=exp(N(0,1)) per cell, f=sin(pi x)sin(pi y), zero Dirichlet, "
                               "assemble + Jacobi-PCG to rtol 1e-10" % (args.n, args.n),
                   "cells": P.ne, "dofs": P.nv, "nnz": P.nnz, "pcg_mode": args.mode},
        "assembly_melem_s": P.ne / (np.median(asm_ms) * 1e-3) / 1e6,
        "assembly_ms": float(np.median(asm_ms)),
        "assembly_gbs_algorithmic": asm_bytes / (np.median(asm_ms) * 1e-3) / 1e9,
        "assembly_mfma_two_phase_ms": float(np.median(asm_mfma)),
        "pcg_s": float(np.median(pcg_ms)) / 1e3,
        "pcg_iters": int(iters[-1]),
        "pcg_relres": info.relres,
        "pcg_modes": per_mode,
        "mixed_speedup_vs_fp64": {k: fp64_s / v["s"] for k, v in per_mode.items()},
        "dof_per_s": P.nv / (float(np.median(pcg_ms)) * 1e-3),
        "grad_cost_per_solve": None if adj is None else adj["grad_cost_per_solve"],
        "adjoint": adj,
        "plan_build_s": P.plan_s,
        "roofline": roof,
        "roofline_mixed": roof_mixed,
        "pcg_iteration": {
            "us_in_solve": 1e3 * float(np.median(pcg_ms)) / max(1, iters[-1]),
            "dia_bytes": roof["iteration"]["algorithmic_bytes"],
            "dia_gbs_in_solve": roof["iteration"]["algorithmic_bytes"]
            / (float(np.median(pcg_ms)) * 1e-3 / max(1, iters[-1])) / 1e9,
            "csr_formula_bytes": spmv_bytes(P.nnz, P.nv) + 80 * P.nv,
            "note": "dia_bytes: what the DIA sweeps must move (sweep_bytes); csr_formula_bytes: "
                    "the CSR storage figure 12 nnz + 4 (N+1) + 16 N + 80 N, for reference"},
        "oracle": None,
        "cpu_baseline": None,
    }
    # host copies for the checks that run after every GPU measurement
    rp, col, val = P.plan.csr.download()
    bh = P.b.numpy()
    n, theta, ne, nv = P.n, P.log_alpha, P.ne, P.nv
    P.close()

    def checks():
        # the oracle: its own assembly of the same problem, compared entry by
        # entry, and its own solve, against which every mode's solution is
        # measured; then the CPU baseline
        if not args.no_oracle:
            L, (rpo, colo, valo, bo), ot = oracle_p1(n, theta, 1.0, True)
            cmp = oracle_compare_csr(rpo, colo, valo, bo, rp, col, val, bh)
            log("oracle assembly (%.1f s mesh, %.1f s cells): pattern equal %s, matrix diff "
                "%.2e, rhs diff %.2e" % (ot["mesh_s"], ot["assembly_s"],
                                         cmp["pattern_bitwise_equal"],
                                         cmp["matrix_row_scaled_diff"] or -1,
                                         cmp["rhs_max_rel_diff"]))
            xo, sol = oracle_solve(L, rpo, colo, valo, bo)
            del rpo, colo, valo, bo
            log("oracle solve: %d iters in %.1f s, relres %.2e, true relres %.2e"
                % (sol["iters"], sol["s"], sol["relres"], sol["true_relres"]))
            for name in modes:
                per_mode[name]["err_vs_oracle"] = rel_err(xs[name], xo)
            out["oracle"] = {"assembly": {**cmp, **ot, "cells": ne}, "solve": sol,
                             "note": "the oracle's own mesh, reference-order assembly (alpha = "
                                     "exp and f = sin sin on the host) and PCG; err_vs_oracle = "
                                     "||x_gpu - x_oracle|| / ||x_oracle||"}
        if not args.no_cpu_baseline:
            out["cpu_baseline"] = cpu_baseline_p1(ne, nv, rp, col, val, bh, iters[-1],
                                                  out["oracle"])

    return out, checks


def p3_dirichlet_flags(nv, ev, ce, nd):
    """flagEntitiesOnBoundary for P3 (special_entity_sets.cc:40-86): edges
    of one cell, their endpoints, and the interior DOFs of those edges
    (numbered nv + 2e, nv + 2e + 1)."""
    be = np.nonzero(np.bincount(ce.ravel(), minlength=len(ev)) == 1)[0]
    fl = np.zeros(nd, np.uint8)
    fl[ev[be].ravel()] = 1
    fl[nv + 2 * be] = 1
    fl[nv + 2 * be + 1] = 1
    return fl


def build_p3(capi, ctx, n=512, seed=0):
    """The P3 workload: n x n jittered mesh, zero Dirichlet data on the
    boundary DOFs, a = exp(N(0,1)) per cell, f = sin sin."""
    xy, cells, edges = capi.tp_triag_mesh(n)
    rng = np.random.default_rng(seed)
    h = 1.0 / n
    inner = (xy[:, 0] > 0) & (xy[:, 0] < 1) & (xy[:, 1] > 0) & (xy[:, 1] < 1)
    xy = xy.copy()
    xy[inner] += rng.uniform(-0.2 * h, 0.2 * h, size=(int(inner.sum()), 2))
    ev, ce, co = capi.mesh_edges(xy.shape[0], cells, edges)
    cd, nd = capi.dof_table(xy.shape[0], cells, len(ev), ce, co, 3)
    flags = p3_dirichlet_flags(xy.shape[0], ev, ce, nd)
    mesh = capi.Mesh(ctx, xy, cells, cd, nd)
    fe = capi.FE(ctx, 3)
    plan = capi.Plan(ctx, mesh, flags)
    nnz = plan.csr.info()[2]
    theta = capi.Vec(ctx, np.random.default_rng(seed + 1).standard_normal(cells.shape[0]))
    b, x = capi.Vec(ctx, nd), capi.Vec(ctx, nd)
    args_c = (fe, capi.coeff(capi.COEFF_EXPCELL, 1.0, theta), capi.coeff(capi.COEFF_CONST, 1.0),
              capi.coeff(capi.COEFF_SINSIN, 1.0), b)
    return xy, cells, nd, nnz, mesh, fe, plan, theta, b, x, args_c


def run_p3(args, capi, ctx, n=512, seed=0):
    xy, cells, nd, nnz, mesh, fe, plan, theta, b, x, args_c = build_p3(capi, ctx, n, seed)
    plan.assemble(*args_c)
    asm = [plan.assemble(*args_c) for _ in range(3)]
    log("p3: N=%d nnz=%d assembly (exact element order) %.3f ms" % (nd, nnz, np.median(asm)))
    rp, col_exact, v_exact = plan.csr.download()
    b_exact = b.numpy()
    # element matrices on the matrix cores; the solves below use this matrix
    plan.assemble(*args_c, mfma=True)
    asm_mfma = [plan.assemble(*args_c, mfma=True) for _ in range(3)]
    _, _, v_mfma = plan.csr.download()
    mat_err = row_scaled_err(rp, v_mfma, v_exact)
    rhs_equal = bool(np.array_equal(b.numpy(), b_exact))
    del v_mfma
    log("p3: assembly (MFMA contraction) %.3f ms, matrix row-scaled diff %.2e, rhs equal %s"
        % (np.median(asm_mfma), mat_err, rhs_equal))
    res = {}
    if args.no_solve:
        modes_p3 = ()
    else:
        modes_p3 = (("fp64", capi.PCG_FP64), ("mixed_matrix", capi.PCG_MIXED_MATRIX),
                    ("mixed", capi.PCG_MIXED))
    for name, md in modes_p3:
        x.fill(0.0)
        inf = capi.pcg(ctx, plan.csr, b, x, rtol=1e-10, mode=md, check_every=128, maxit=MAXIT,
                       raise_on_noconv=False)
        res[name] = {"s": inf.ms / 1e3, "iters": inf.iters, "converged": inf.converged,
                     "relres": inf.relres, "x": x.numpy(),
                     "true_relres": true_relres(capi, ctx, plan.csr, x, b)}
        log("p3 %s: %.1f ms, %d iters, true relres %.2e" % (name, inf.ms, inf.iters,
                                                            res[name]["true_relres"]))
    xs = {k: v.pop("x") for k, v in res.items()}
    roof = None
    if not args.no_solve:
        # per-launch sweep durations of 200 FP64 iterations (plain launches)
        t, _ = capi.pcg_sweep_times(ctx, plan.csr, b, x, iters=200, mode=capi.PCG_FP64)
        st = capi.solver_storage(plan.csr)[0]
        roof = ell_roofline_entry(nd, nnz, t, 2 if st == capi.STORAGE_ELL16 else 4)
        log("p3 FP64 sweeps: %.1f + %.1f us (+ 2 x %.1f us reductions), %.0f GB/s per iteration"
            % (1e3 * t[0], 1e3 * t[1], 1e3 * t[2], roof["iteration"]["achieved"]))
    out = {"metric": "P3 assembly Melem/s + PCG time to rtol 1e-10", "config": {
        "workload": "P3 reaction-diffusion, %dx%d jittered (0.2h) triangles, degree-6 "
                    "quadrature, a=exp(N(0,1)), f=sin sin" % (n, n),
        "cells": int(cells.shape[0]), "dofs": int(nd), "nnz": int(nnz)},
        "assembly_ms": float(np.median(asm)),
        "assembly_melem_s": cells.shape[0] / (np.median(asm) * 1e-3) / 1e6,
        "assembly_mfma_ms": float(np.median(asm_mfma)),
        "assembly_mfma_melem_s": cells.shape[0] / (np.median(asm_mfma) * 1e-3) / 1e6,
        "mfma_matrix_row_scaled_diff": mat_err, "mfma_rhs_bitwise_equal": rhs_equal,
        "pcg_matrix": "MFMA-contracted; err_vs_oracle against the oracle's solution of its "
                      "own reference-order system",
        "pcg": res, "roofline": roof, "oracle": None,
        "mixed_speedup_vs_fp64": {k: res["fp64"]["s"] / v["s"] for k, v in res.items()}
        if "fp64" in res else None}
    for hh in (b, x, theta, plan, fe, mesh):
        hh.close()

    def checks():
        # the oracle: its own jittered mesh (same coordinates), reference-order
        # assembly compared with the exact-order device assembly, and its PCG
        # solve, against which the solutions (of the matrix-core system) are
        # measured
        if args.no_oracle:
            return
        L = oracle_lib()
        t0 = time.perf_counter()
        mo = L.tp_triag_mesh(n)
        mo = L.TriMesh(xy, mo.cells, mo.edges)
        cdo, ndo = mo.dofs(3)
        flo = mo.dirichlet_flags(3)
        t1 = time.perf_counter()
        rpo, colo, valo, bo = L.cpu_path_assemble(
            mo, 3, cdo, ndo, np.exp(np.random.default_rng(seed + 1).standard_normal(mo.ne)),
            1.0, L.F_SINSIN, 0.0, flo)
        t2 = time.perf_counter()
        cmp = oracle_compare_csr(rpo, colo, valo, bo, rp, col_exact, v_exact, b_exact)
        log("p3 oracle assembly (%.1f s mesh, %.1f s cells): pattern equal %s, matrix diff "
            "%.2e, rhs diff %.2e" % (t1 - t0, t2 - t1, cmp["pattern_bitwise_equal"],
                                     cmp["matrix_row_scaled_diff"] or -1,
                                     cmp["rhs_max_rel_diff"]))
        sol = None
        if xs:
            xo, sol = oracle_solve(L, rpo, colo, valo, bo)
            log("p3 oracle solve: %d iters in %.1f s, relres %.2e, true relres %.2e"
                % (sol["iters"], sol["s"], sol["relres"], sol["true_relres"]))
            for k, v in res.items():
                v["err_vs_oracle"] = rel_err(xs[k], xo)
        out["oracle"] = {"assembly": {**cmp, "mesh_s": t1 - t0, "assembly_s": t2 - t1},
                         "solve": sol}

    return out, checks


def run_batched(args, capi, ctx, nprob=256, seed=4):
    rng = np.random.default_rng(seed)
    sizes = np.rint(np.exp(rng.uniform(np.log(32), np.log(512), nprob))).astype(int)
    xy, cells, flags, voff, coff = capi.batched_tp_mesh(sizes)
    ne, nv = cells.shape[0], xy.shape[0]
    log("batched: %d problems, %d cells, %d DOFs" % (nprob, ne, nv))
    g = np.concatenate([np.random.default_rng(seed + 1 + k).standard_normal(int(coff[k + 1] - coff[k]))
                        for k in range(nprob)])
    mesh = capi.Mesh(ctx, xy, cells)
    fe = capi.FE(ctx, 1)
    plan = capi.Plan(ctx, mesh, flags)
    theta = capi.Vec(ctx, g)
    b, x = capi.Vec(ctx, nv), capi.Vec(ctx, nv)
    acoef = (fe, capi.coeff(capi.COEFF_EXPCELL, 1.0, theta), capi.coeff(capi.COEFF_CONST, 0.0),
             capi.coeff(capi.COEFF_CONST, 1.0), b)
    plan.assemble(*acoef)
    batch = capi.Batch(ctx, plan.csr, voff)
    rp, col, val = plan.csr.download()
    bh = b.numpy()
    sols = {}
    modes = {"fp64": capi.PCG_FP64, "mixed": capi.PCG_MIXED,
             "mixed_matrix": capi.PCG_MIXED_MATRIX}
    wanted = args.batch_modes.split(",")
    modes = {k: v for k, v in modes.items() if k in wanted}
    per_mode = {}
    for name, md in modes.items():
        x.fill(0.0)
        # tunes / captures the graph
        batch.pcg(b, x, rtol=1e-10, maxit=MAXIT, mode=md, graph=args.batch_graph)
        times = []
        for _ in range(max(1, args.batch_steps)):
            ctx.sync()
            t0 = time.perf_counter()
            asm_ms = plan.assemble(*acoef)
            x.fill(0.0)
            summ, infos = batch.pcg(b, x, rtol=1e-10, maxit=MAXIT, mode=md,
                                    graph=args.batch_graph)
            ctx.sync()
            times.append((time.perf_counter() - t0, asm_ms, summ.ms))
        sols[name] = x.numpy()
        it = [i.iters for i in infos]
        wall = float(np.median([t[0] for t in times]))
        per_mode[name] = {
            "melem_s": ne / wall / 1e6, "ms_per_step": 1e3 * wall,
            "assembly_ms": float(np.median([t[1] for t in times])),
            "pcg_ms": float(np.median([t[2] for t in times])), "iters_max": int(max(it)),
            "iters_min": int(min(it)), "iters_sum": int(sum(it)),
            "refinements_sum": int(sum(i.refinements for i in infos)),
            "converged": bool(summ.converged), "worst_relres": summ.relres,
            "host_syncs": int(summ.syncs), "path": "separate calls"}
        log("batched %s: %.1f ms/step (pcg %.1f ms), iters %d..%d, %d host syncs"
            % (name, 1e3 * wall, per_mode[name]["pcg_ms"], min(it), max(it), summ.syncs))
        if args.batch_graph == 0:
            # the step as one captured graph (assembly, value refresh, set-up,
            # first chunk) plus the captured chunks: lfgpu_batch_assemble_pcg
            x.fill(0.0)
            batch.assemble_pcg(plan, *acoef, x, rtol=1e-10, maxit=MAXIT, mode=md)  # captures
            ftimes = []
            for _ in range(max(1, args.batch_steps)):
                ctx.sync()
                t0 = time.perf_counter()
                x.fill(0.0)
                fs, finfos = batch.assemble_pcg(plan, *acoef, x, rtol=1e-10, maxit=MAXIT, mode=md)
                ctx.sync()
                ftimes.append((time.perf_counter() - t0, fs.ms))
            fwall = float(np.median([t[0] for t in ftimes]))
            same = bool(np.array_equal(x.numpy(), sols[name]) and
                        [i.iters for i in finfos] == it)
            sep = {k: per_mode[name][k] for k in ("melem_s", "ms_per_step", "host_syncs")}
            per_mode[name].update({
                "melem_s": ne / fwall / 1e6, "ms_per_step": 1e3 * fwall,
                "device_ms": float(np.median([t[1] for t in ftimes])),
                "host_syncs": int(fs.syncs), "path": "one captured graph per step "
                "(lfgpu_batch_assemble_pcg) and the captured chunks",
                "bitwise_equal_to_separate_calls": same, "separate_calls": sep})
            log("batched %s one graph: %.1f ms/step (device %.1f ms), %d host syncs, "
                "bitwise equal to the separate calls: %s"
                % (name, 1e3 * fwall, per_mode[name]["device_ms"], fs.syncs, same))
    plain = (None if args.no_plain else
             plain_fp64_per_problem(args, capi, ctx, sizes, g, coff, voff))
    sel = per_mode[args.batch_mode if args.batch_mode in per_mode else next(iter(per_mode))]
    out = {"metric": "batched P1 assemble + PCG, Melem/s", "value": sel["melem_s"],
           "unit": "Melem/s", "pcg_mode": args.batch_mode,
           "config": {"workload": "%d P1 problems, sides log-uniform in [32,512], "
                                  "a=exp(N(0,1)), f=1" % nprob,
                      "cells": int(ne), "dofs": int(nv), "nnz": int(plan.csr.info()[2])},
           "ms_per_step": sel["ms_per_step"], "modes": per_mode,
           "storage": {capi.STORAGE_ELL: "sliced ELL", capi.STORAGE_DIA: "DIA",
                       capi.STORAGE_ELL16: "sliced ELL, 16-bit column offsets",
                       capi.STORAGE_BLOCK_DIA: "per-problem DIA"}.get(batch.storage()),
           "speedup_vs_fp64": ({k: per_mode["fp64"]["ms_per_step"] / v["ms_per_step"]
                                for k, v in per_mode.items()} if "fp64" in per_mode else None),
           "tuned_rows_per_workgroup": summ.refinements, "oracle": None}
    if plain is not None:
        sols["plain_fp64"] = plain.pop("x")
        out["plain_fp64_per_problem"] = plain
        out["speedup_vs_plain_fp64"] = {k: plain["ms_per_step"] / v["ms_per_step"]
                                        for k, v in per_mode.items()}
    for hh in (batch, b, x, theta, plan, fe, mesh):
        hh.close()

    def checks():
        # the oracle: every problem assembled by the reference-order CPU path
        # on its own mesh and solved by the textbook PCG (one thread per
        # problem, problems in parallel), against the blocks of the GPU's
        # matrix and each mode's solution
        if args.no_oracle:
            return
        x_ref, out["oracle"] = oracle_batched(rp, col, val, bh, sizes, g, coff, voff)
        for name, xv in sols.items():
            errs = [rel_err(xv[voff[k]:voff[k + 1]], x_ref[voff[k]:voff[k + 1]])
                    for k in range(nprob)]
            tgt = out["plain_fp64_per_problem"] if name == "plain_fp64" else per_mode[name]
            tgt["worst_err_vs_oracle"] = float(max(errs))
            log("batched %s: worst error vs the oracle %.2e" % (name, max(errs)))

    return out, checks


def run_batched_p3(args, capi, ctx, nprob=256, seed=6):
    """C4 at degree 3: nprob P3 problems on n x n meshes, sides log-uniform
    in [8, 64], each with its own DOF numbering (vertices, edges, cells) and
    zero Dirichlet data on its boundary DOFs, stacked so that every problem
    owns a contiguous row range; a = exp(N(0,1)) per cell, f = 1.  The
    batch storage is sliced ELL (the P3 patterns are not banded), tuned by
    the autotuner for this degree.  Each step: the separate calls
    (assembly, batched PCG) and one captured graph per step
    (lfgpu_batch_assemble_pcg)."""
    rng = np.random.default_rng(seed)
    sizes = np.rint(np.exp(rng.uniform(np.log(8), np.log(64), nprob))).astype(int)
    xs, cs, cds, fls, doff, coff, vo = [], [], [], [], [0], [0], 0
    for n in sizes:
        xy, cells, edges = capi.tp_triag_mesh(int(n))
        ev, ce, co = capi.mesh_edges(xy.shape[0], cells, edges)
        cd, nd = capi.dof_table(xy.shape[0], cells, len(ev), ce, co, 3)
        fl = p3_dirichlet_flags(xy.shape[0], ev, ce, nd)
        xs.append(xy)
        cs.append(cells + vo)
        cds.append(cd + doff[-1])
        fls.append(fl)
        vo += xy.shape[0]
        doff.append(doff[-1] + nd)
        coff.append(coff[-1] + cells.shape[0])
    xy = np.concatenate(xs)
    cells = np.ascontiguousarray(np.concatenate(cs), np.int32)
    cd = np.concatenate(cds)
    nd, ne = int(doff[-1]), int(coff[-1])
    doff = np.array(doff, np.int64)
    log("batched P3: %d problems, %d cells, %d DOFs" % (nprob, ne, nd))
    g = rng.standard_normal(ne)
    mesh = capi.Mesh(ctx, xy, cells, cd, nd)
    fe = capi.FE(ctx, 3)
    plan = capi.Plan(ctx, mesh, np.concatenate(fls))
    theta = capi.Vec(ctx, g)
    b, x = capi.Vec(ctx, nd), capi.Vec(ctx, nd)
    acoef = (fe, capi.coeff(capi.COEFF_EXPCELL, 1.0, theta), capi.coeff(capi.COEFF_CONST, 0.0),
             capi.coeff(capi.COEFF_CONST, 1.0), b)
    plan.assemble(*acoef)
    batch = capi.Batch(ctx, plan.csr, doff)
    rp, col, val = plan.csr.download()
    bh = b.numpy()
    modes = {"fp64": capi.PCG_FP64, "mixed": capi.PCG_MIXED,
             "mixed_matrix": capi.PCG_MIXED_MATRIX}
    per_mode, sols = {}, {}
    for name, md in modes.items():
        x.fill(0.0)
        batch.pcg(b, x, rtol=1e-10, maxit=MAXIT, mode=md)  # tunes, captures the chunk
        times = []
        for _ in range(max(1, args.batch_steps)):
            ctx.sync()
            t0 = time.perf_counter()
            plan.assemble(*acoef)
            x.fill(0.0)
            summ, infos = batch.pcg(b, x, rtol=1e-10, maxit=MAXIT, mode=md)
            ctx.sync()
            times.append(time.perf_counter() - t0)
        sols[name] = x.numpy()
        it = [i.iters for i in infos]
        x.fill(0.0)
        batch.assemble_pcg(plan, *acoef, x, rtol=1e-10, maxit=MAXIT, mode=md)  # captures
        ftimes = []
        for _ in range(max(1, args.batch_steps)):
            ctx.sync()
            t0 = time.perf_counter()
            x.fill(0.0)
            fs, finfos = batch.assemble_pcg(plan, *acoef, x, rtol=1e-10, maxit=MAXIT, mode=md)
            ctx.sync()
            ftimes.append(time.perf_counter() - t0)
        same = bool(np.array_equal(x.numpy(), sols[name]) and [i.iters for i in finfos] == it)
        wall, fwall = float(np.median(times)), float(np.median(ftimes))
        per_mode[name] = {
            "melem_s": ne / fwall / 1e6, "ms_per_step": 1e3 * fwall, "host_syncs": int(fs.syncs),
            "path": "one captured graph per step (lfgpu_batch_assemble_pcg) and the captured "
                    "chunks",
            "bitwise_equal_to_separate_calls": same,
            "separate_calls": {"ms_per_step": 1e3 * wall, "host_syncs": int(summ.syncs)},
            "iters_max": int(max(it)), "iters_min": int(min(it)),
            "refinements_sum": int(sum(i.refinements for i in infos)),
            "converged": bool(fs.converged), "worst_relres": fs.relres}
        log("batched P3 %s: %.1f ms/step one graph (%d syncs), %.1f ms separate (%d syncs), "
            "iters %d..%d, bitwise equal %s" % (name, 1e3 * fwall, fs.syncs, 1e3 * wall,
                                                summ.syncs, min(it), max(it), same))
    out = {"metric": "batched P3 assemble + PCG, Melem/s",
           "config": {"workload": "%d P3 problems, sides log-uniform in [8,64], a=exp(N(0,1)), "
                                  "f=1" % nprob, "cells": ne, "dofs": nd,
                      "nnz": int(plan.csr.info()[2])},
           "storage": {capi.STORAGE_ELL: "sliced ELL", capi.STORAGE_DIA: "DIA",
                       capi.STORAGE_ELL16: "sliced ELL, 16-bit column offsets",
                       capi.STORAGE_BLOCK_DIA: "per-problem DIA"}.get(batch.storage()),
           "tuned_rows_per_workgroup": summ.refinements, "modes": per_mode,
           "speedup_vs_fp64": {k: per_mode["fp64"]["ms_per_step"] / v["ms_per_step"]
                               for k, v in per_mode.items()}, "oracle": None}
    for hh in (batch, b, x, theta, plan, fe, mesh):
        hh.close()

    def checks():
        # each problem assembled by the oracle's reference-order path on its own
        # mesh and solved by its PCG (problems in parallel, one thread each)
        if args.no_oracle:
            return
        import concurrent.futures as cf
        L = oracle_lib()

        def one(k):
            n = int(sizes[k])
            m = L.tp_triag_mesh(n)
            cdo, ndo = m.dofs(3)
            rpo, colo, valo, bo = L.cpu_path_assemble(m, 3, cdo, ndo,
                                                      np.exp(g[coff[k]:coff[k + 1]]), 0.0,
                                                      L.F_CONST, 1.0, m.dirichlet_flags(3))
            lo, hi = int(doff[k]), int(doff[k + 1])
            cmp = oracle_compare_csr(rpo, colo, valo, bo, rp[lo:hi + 1] - rp[lo],
                                     col[rp[lo]:rp[hi]] - lo, val[rp[lo]:rp[hi]], bh[lo:hi])
            xo, _ = oracle_solve(L, rpo, colo, valo, bo, threads=False)
            return xo, cmp

        t0 = time.perf_counter()
        with cf.ThreadPoolExecutor(max_workers=min(16, os.cpu_count() or 1)) as ex:
            res = list(ex.map(one, range(nprob)))
        x_ref = np.concatenate([r[0] for r in res])
        diffs = [r[1]["matrix_row_scaled_diff"] for r in res
                 if r[1]["matrix_row_scaled_diff"] is not None]
        out["oracle"] = {"s": time.perf_counter() - t0,
                         "pattern_bitwise_equal": all(r[1]["pattern_bitwise_equal"] for r in res),
                         "matrix_row_scaled_diff": float(max(diffs)) if diffs else None,
                         "rhs_max_rel_diff": float(max(r[1]["rhs_max_rel_diff"] for r in res))}
        for name, xv in sols.items():
            errs = [rel_err(xv[doff[k]:doff[k + 1]], x_ref[doff[k]:doff[k + 1]])
                    for k in range(nprob)]
            per_mode[name]["worst_err_vs_oracle"] = float(max(errs))
            log("batched P3 %s: worst error vs the oracle %.2e" % (name, max(errs)))

    return out, checks


def oracle_batched(rp, col, val, bg, sizes, g, coff, voff):
    """Oracle systems and solutions of the batched problems (alpha = exp(g),
    gamma = 0, f = 1, zero Dirichlet data), compared block by block with the
    GPU's matrix (rp, col, val) and load vector bg."""
    import concurrent.futures as cf
    L = oracle_lib()

    def one(k):
        n = int(sizes[k])
        m = L.tp_triag_mesh(n)
        cd, nd = m.dofs(1)
        rpo, colo, valo, bo = L.cpu_path_assemble(m, 1, cd, nd, np.exp(g[coff[k]:coff[k + 1]]),
                                                  0.0, L.F_CONST, 1.0, m.dirichlet_flags(1))
        lo, hi = int(voff[k]), int(voff[k + 1])
        rpg = rp[lo:hi + 1] - rp[lo]
        cmp = oracle_compare_csr(rpo, colo, valo, bo, rpg, col[rp[lo]:rp[hi]] - lo,
                                 val[rp[lo]:rp[hi]], bg[lo:hi])
        xo, sol = oracle_solve(L, rpo, colo, valo, bo, threads=False)
        return xo, cmp, sol

    t0 = time.perf_counter()
    workers = min(16, os.cpu_count() or 1)
    with cf.ThreadPoolExecutor(max_workers=workers) as ex:
        res = list(ex.map(one, range(len(sizes))))
    dt = time.perf_counter() - t0
    x_ref = np.concatenate([r[0] for r in res])
    cmps = [r[1] for r in res]
    sols = [r[2] for r in res]
    diffs = [c["matrix_row_scaled_diff"] for c in cmps if c["matrix_row_scaled_diff"] is not None]
    out = {"problems": len(sizes), "s": dt, "threads": workers,
           "pattern_bitwise_equal": all(c["pattern_bitwise_equal"] for c in cmps),
           "matrix_bitwise_equal": all(c["matrix_bitwise_equal"] for c in cmps),
           "matrix_row_scaled_diff": float(max(diffs)) if diffs else None,
           "rhs_max_rel_diff": float(max(c["rhs_max_rel_diff"] for c in cmps)),
           "solve": {"rtol": ORACLE_RTOL, "iters_sum": int(sum(s_["iters"] for s_ in sols)),
                     "worst_true_relres": float(max(s_["true_relres"] for s_ in sols))}}
    log("batched oracle: %d problems in %.1f s (%d threads), pattern equal %s, matrix diff "
        "%.2e" % (len(sizes), dt, workers, out["pattern_bitwise_equal"],
                  out["matrix_row_scaled_diff"] or -1))
    return x_ref, out


def plain_fp64_per_problem(args, capi, ctx, sizes, g, coff, voff):
    """The same problems through the plain FP64 path: one plan, one assembly
    and one lfgpu_pcg call per problem, in sequence."""
    probs = []
    for k, n in enumerate(sizes):
        xy, cells, _ = capi.tp_triag_mesh(int(n))
        j, i = np.divmod(np.arange(xy.shape[0]), n + 1)
        fl = ((i == 0) | (i == n) | (j == 0) | (j == n)).astype(np.uint8)
        mesh = capi.Mesh(ctx, xy, cells)
        plan = capi.Plan(ctx, mesh, fl)
        th = capi.Vec(ctx, g[coff[k]:coff[k + 1]])
        probs.append((mesh, plan, th, capi.Vec(ctx, xy.shape[0]), capi.Vec(ctx, xy.shape[0])))
    fe = capi.FE(ctx, 1)
    one, zero = capi.coeff(capi.COEFF_CONST, 1.0), capi.coeff(capi.COEFF_CONST, 0.0)

    def step():
        its = []
        for mesh, plan, th, b, x in probs:
            plan.assemble(fe, capi.coeff(capi.COEFF_EXPCELL, 1.0, th), zero, one, b)
            x.fill(0.0)
            its.append(capi.pcg(ctx, plan.csr, b, x, rtol=1e-10, maxit=MAXIT).iters)
        return its

    step()  # builds every solver mirror
    times = []
    for _ in range(max(1, args.batch_steps)):
        ctx.sync()
        t0 = time.perf_counter()
        its = step()
        ctx.sync()
        times.append(time.perf_counter() - t0)
    x_all = np.concatenate([pr[4].numpy() for pr in probs])
    wall = float(np.median(times))
    log("plain fp64 per problem: %.1f ms/step, iters %d..%d" % (1e3 * wall, min(its), max(its)))
    for pr in probs:
        for hh in pr[::-1]:
            hh.close()
    fe.close()
    return {"ms_per_step": 1e3 * wall, "calls_per_step": 2 * len(sizes),
            "iters_sum": int(sum(its)), "x": x_all}


def _fkpp_setup(capi, ctx, n, c=1.2, h=1.0, nonlocal_terms=False):
    """Mesh, edges, A, M and u0 of the Fisher/KPP workload; with
    nonlocal_terms the boundary condition of modelproblem_island_main.cc
    (gain h(x) and loss L of g = 1 / (1 + |x - y|^2)) replaces h."""
    from lehrfempp_amd import fisher_kpp
    xy, cells, _ = capi.tp_triag_mesh(n)
    ev, ce, _ = capi.mesh_edges(xy.shape[0], cells)
    mesh = capi.Mesh(ctx, xy, cells)
    edges = capi.Edges(ctx, mesh, 1, ev)
    ef = fisher_kpp.boundary_edge_flags(ev.shape[0], ce)
    if nonlocal_terms:
        ctx.sync()
        t0 = time.perf_counter()
        hc, eta = fisher_kpp.nonlocal_gain(ctx, edges, ef)
        loss = fisher_kpp.nonlocal_loss(ctx, edges, ef)
        ctx.sync()
        log("fkpp nonlocal terms: %d boundary DOFs, gain and loss in %.1f ms"
            % (loss[0].size, 1e3 * (time.perf_counter() - t0)))
        A, M = fisher_kpp.assemble_galerkin_matrices(ctx, mesh, edges, ef,
                                                     capi.coeff(capi.COEFF_CONST, c), hc,
                                                     L=loss)
        eta.close()
        loss[1].close()
    else:
        A, M = fisher_kpp.assemble_galerkin_matrices(ctx, mesh, edges, ef,
                                                     capi.coeff(capi.COEFF_CONST, c),
                                                     capi.coeff(capi.COEFF_CONST, h))
    u0 = np.zeros(xy.shape[0])
    r2 = (xy[:, 0] - 0.3) ** 2 + (xy[:, 1] - 0.4) ** 2
    u0[r2 < 0.01] = 0.3
    return mesh, edges, A, M, u0


def run_fkpp(args, capi, ctx):
    """Fisher/KPP Strang splitting (projects/fisher_kpp): c = 1.2, lambda =
    2.1, T = 1, m steps (convergence_meshwidth_main.cc), Robin coefficient
    h = 1 on the boundary, carrying capacity K = 1, u0 = 0.3 on a disc."""
    from lehrfempp_amd import fisher_kpp
    n, m_steps, T, lam = args.fkpp_n, args.fkpp_steps, 1.0, 2.1
    mesh, edges, A, M, u0 = _fkpp_setup(capi, ctx, n, nonlocal_terms=args.fkpp_nonlocal)
    if args.fkpp_nonlocal:
        args.fkpp_mode, args.fkpp_modes = "fp64", "fp64"  # the block needs FP64
    nv = u0.size
    codes = {"fp64": capi.PCG_FP64, "mixed": capi.PCG_MIXED,
             "mixed_matrix": capi.PCG_MIXED_MATRIX}
    vK, vu = capi.Vec(ctx, np.ones(nv)), capi.Vec(ctx, u0)
    per_mode, results = {}, {}
    wanted = args.fkpp_modes.split(",")
    for name in [args.fkpp_mode] + [k for k in codes if k != args.fkpp_mode and k in wanted]:
        ss = fisher_kpp.StrangSplit(ctx, A, M, T, m_steps, lam, rtol=1e-10, mode=codes[name])
        out = ss.Evolution(vK, vu)  # warm-up: builds the solver storage of A, S1, S2
        out.close()
        runs = []
        for _ in range(max(1, args.steps)):
            ctx.sync()
            t0 = time.perf_counter()
            out = ss.Evolution(vK, vu)
            ctx.sync()
            t_run = time.perf_counter() - t0
            info = ss.last_info
            runs.append((t_run, info.ms, info.iters, info.max_iters, info.solves, info.converged))
            results[name] = out.numpy()
            out.close()
        ss.close()
        w = float(np.median([r[0] for r in runs]))
        per_mode[name] = {"ms_per_step": 1e3 * w / (runs[-1][4] // 2),
                          "iters_per_solve": runs[-1][2] / max(runs[-1][4], 1),
                          "converged": bool(runs[-1][5])}
        if name == args.fkpp_mode:
            wall = w
            _, dev_ms, iters, max_it, solves, conv = runs[-1]
        log("fkpp %s: %.2f ms/step" % (name, per_mode[name]["ms_per_step"]))
    ref = results[args.fkpp_mode]
    for name in per_mode:
        per_mode[name]["max_rel_diff_vs_" + args.fkpp_mode] = float(
            np.abs(results[name] - ref).max() / np.abs(ref).max())
    u = ref
    nd = solves // 2
    res = {"metric": "Fisher/KPP Strang splitting, ms per time step", "unit": "ms",
           "value": 1e3 * wall / nd, "higher_is_better": False,
           "config": {"workload": "Fisher/KPP on %dx%d P1 structured triangles, c=1.2, "
                                  "lambda=2.1, %s, K=1, T=1, %d steps, SDIRK-2 "
                                  "diffusion + exact logistic reaction, stage solves by "
                                  "Jacobi-PCG to 1e-10"
                                  % (n, n, "nonlocal gain h and loss L of "
                                     "g = 1/(1+|x-y|^2)" if args.fkpp_nonlocal
                                     else "Robin h=1", m_steps),
                      "dofs": nv, "nnz": int(A.info()[2])},
           "pcg_mode": args.fkpp_mode,
           "evolution_s": wall, "device_ms": dev_ms, "diffusion_steps": nd,
           "solves": solves, "iters_per_solve": iters / max(solves, 1),
           "max_iters": max_it, "converged": bool(conv),
           "mass_final": float(u.sum() / nv), "modes": per_mode}
    log("fkpp %d^2: %.2f ms/step, %.1f iterations/solve (max %d)"
        % (n, res["value"], res["iters_per_solve"], max_it))
    for hh in (vK, vu, A, M, edges, mesh):
        hh.close()
    if not args.no_cpu_baseline and not args.fkpp_nonlocal:
        res["cpu_baseline"] = fkpp_cpu_sample(capi, ctx, args.fkpp_sample_n, m_steps)
    return res


def fkpp_cpu_sample(capi, ctx, n, m_steps):
    """The oracle's StrangSplit (SciPy SuperLU, one thread) against the device
    stepper on the same n x n problem: time per step and the difference."""
    import scipy.sparse as sp
    from lehrfempp_amd import fisher_kpp
    from oracle import lfref as L  # CPU restatement, timed here as the CPU baseline
    T, lam = 1.0, 2.1
    mesh, edges, A, M, u0 = _fkpp_setup(capi, ctx, n)
    nv = u0.size
    ss = fisher_kpp.StrangSplit(ctx, A, M, T, m_steps, lam, rtol=1e-12)
    vK, vu = capi.Vec(ctx, np.ones(nv)), capi.Vec(ctx, u0)
    out = ss.Evolution(vK, vu)
    out.close()
    ctx.sync()
    t0 = time.perf_counter()
    out = ss.Evolution(vK, vu)
    ctx.sync()
    gpu_s = time.perf_counter() - t0
    nd = ss.last_info.solves // 2
    ug = out.numpy()
    rA, cA, vA = A.download()
    rM, cM, vM = M.download()
    As = sp.csr_matrix((vA, cA, rA), shape=(nv, nv))
    Ms = sp.csr_matrix((vM, cM, rM), shape=(nv, nv))
    t0 = time.perf_counter()
    ref = L.StrangSplit(As, Ms, T, m_steps, lam)
    uc = ref.evolution(np.ones(nv), u0)
    cpu_s = time.perf_counter() - t0
    for hh in (out, vK, vu, ss, A, M, edges, mesh):
        hh.close()
    log("fkpp cpu sample %d^2: cpu %.1f ms/step, gpu %.2f ms/step" %
        (n, 1e3 * cpu_s / nd, 1e3 * gpu_s / nd))
    return {"value": 1e3 * cpu_s / nd, "unit": "ms per time step", "cores": 1, "kind": "port",
            "sample": "oracle StrangSplit with SciPy SuperLU (factorisations included) on the "
                      "%dx%d problem (%d DOFs), %d steps" % (n, n, nv, m_steps),
            "gpu_ms_per_step_same_problem": 1e3 * gpu_s / nd,
            "max_rel_diff_gpu_vs_cpu": float(np.abs(ug - uc).max() / np.abs(uc).max())}


def run_c1(args, capi, ctx, n=64, reps=20):
    """The 64 x 64 P1 Poisson problem (alpha = 1, gamma = 0, f = 1, zero
    Dirichlet data) of homDir_linfe_demo.cc:114-175.  At 4,225 DOFs every
    kernel is launch-latency bound, so only absolute times are reported:
    assembly, one FP64 SpMV and the Jacobi-PCG solve to 1e-10 in each mode
    (medians over `reps`)."""
    P = Problem(capi, ctx, n, args.seed, gamma=0.0, f_sinsin=False,
                log_alpha=np.zeros(2 * n * n))
    asm = [P.assemble() for _ in range(reps)]
    xs, ys = capi.Vec(ctx, np.random.default_rng(1).standard_normal(P.nv)), capi.Vec(ctx, P.nv)
    spmv_us = 1e3 * capi.spmv_timed(ctx, P.plan.csr, xs, ys, 200)
    xs.close()
    ys.close()
    modes = {"fp64": capi.PCG_FP64, "mixed": capi.PCG_MIXED, "mixed_matrix": capi.PCG_MIXED_MATRIX}
    solves = {}

    def solve(md):
        # host checks every 16 iterations: at ~6 us per iteration the empty
        # launches after convergence would otherwise dominate
        P.x.fill(0.0)
        return capi.pcg(ctx, P.plan.csr, P.b, P.x, rtol=1e-10, mode=md, check_every=16,
                        maxit=MAXIT, raise_on_noconv=False)

    for name, md in modes.items():
        runs = [solve(md) for _ in range(5)]
        solves[name] = {"ms": float(np.median([r.ms for r in runs])), "iters": runs[-1].iters,
                        "relres": runs[-1].relres,
                        "true_relres": true_relres(capi, ctx, P.plan.csr, P.x, P.b)}
    log("c1 %dx%d: assembly %.1f us, spmv %.2f us, pcg %s" % (
        n, n, 1e3 * float(np.median(asm)), spmv_us,
        ", ".join("%s %.2f ms (%d it)" % (k, v["ms"], v["iters"]) for k, v in solves.items())))
    out = {"metric": "C1 absolute times (launch-latency bound)",
           "config": {"workload": "P1 Poisson on a %dx%d structured triangle mesh, a=1, f=1, "
                                  "zero Dirichlet" % (n, n),
                      "cells": P.ne, "dofs": P.nv, "nnz": P.nnz},
           "data": "synthetic", "assembly_us": 1e3 * float(np.median(asm)), "spmv_us": spmv_us,
           "pcg": solves, "check_every": 16}
    P.close()
    return out


def run_traffic(args, capi, ctx):
    """A short fixed sequence for HBM byte counters (rocprofv3 --pmc) on the
    2048x2048 system: the FP64 SpMV, 8 FP64 PCG iterations (k_cg_spmv /
    k_cg_update), 8 FP32-Krylov iterations (k_mx_a / k_mx_b), the two
    calibration kernels of known size (a plain copy and a non-temporal copy of
    1 GiB each), then 8 FP64 iterations of the 512 x 512 P3 system (row
    sweeps k_cgr_spmv / k_cgr_update on sliced ELL).  Returns the algorithmic
    bytes per dispatch of each kernel."""
    P = Problem(capi, ctx, args.n, args.seed)
    P.assemble()
    xs, ys = capi.Vec(ctx, np.random.default_rng(1).standard_normal(P.nv)), capi.Vec(ctx, P.nv)
    capi.spmv_timed(ctx, P.plan.csr, xs, ys, 3)
    for md in (capi.PCG_FP64, capi.PCG_MIXED):
        P.x.fill(0.0)
        capi.pcg(ctx, P.plan.csr, P.b, P.x, rtol=1e-10, maxit=8, mode=md, check_every=8,
                 raise_on_noconv=False)
    _, _, sym = capi.solver_storage(P.plan.csr)
    ncopy = 1 << 27  # 1 GiB of doubles
    c0, c1 = capi.Vec(ctx, ncopy), capi.Vec(ctx, ncopy)
    c0.fill(1.0)
    for _ in range(3):
        capi.copy(ctx, c0, c1)
    for _ in range(3):
        capi.copy_nt(ctx, c0, c1)
    ctx.sync()
    b64 = sweep_bytes("fp64", P.nv, sym)
    b32 = sweep_bytes("mixed", P.nv, False)
    out = {"config": "traffic", "n": args.n, "dofs": P.nv, "nnz": P.nnz,
           "spmv_dispatches": 3, "spmv_csr_formula_bytes": spmv_bytes(P.nnz, P.nv),
           "k_cg_spmv_bytes": b64[0], "k_cg_update_bytes": b64[1],
           "k_mx_a_bytes": b32[0], "k_mx_b_bytes": b32[1],
           "copy_dispatches": 3, "copy_nt_dispatches": 3,
           "copy_bytes_read": 8 * ncopy, "copy_bytes_written": 8 * ncopy}
    for h in (c0, c1, xs, ys):
        h.close()
    P.close()
    # the P3 row sweeps (sliced ELL, locality order, 16-bit column offsets):
    # 8 FP64 iterations on the 512 x 512 P3 system
    _, _, nd3, nnz3, mesh, fe, plan, theta, b, x, args_c = build_p3(capi, ctx)
    plan.assemble(*args_c, mfma=True)
    capi.pcg(ctx, plan.csr, b, x, rtol=1e-10, maxit=8, check_every=8, raise_on_noconv=False)
    st = capi.solver_storage(plan.csr)[0]
    e1, e2 = ell_row_sweep_bytes(nd3, nnz3, 2 if st == capi.STORAGE_ELL16 else 4)
    out.update({"p3_dofs": nd3, "p3_nnz": nnz3, "k_cgr_spmv_bytes": e1, "k_cgr_update_bytes": e2})
    for h in (b, x, theta, plan, fe, mesh):
        h.close()
    return out


def run_strip(args, capi, ctx, strips=8, iters=2000):
    """One GPU's share of the default workload when it is split into `strips`
    horizontal strips: an interior 2048 x 256-cell strip of the unit square
    (y in [3/8, 4/8]) with the same coefficient law, Dirichlet rows only
    where the strip meets the domain boundary (x = 0, x = 1).  Times the
    assembly, the FP64 SpMV and a fixed number of PCG iterations (FP64 with
    two sweeps and with a single reduction, FP32-Krylov, FP32-matrix), each
    next to 1/strips of the full-size figure
    measured in the same process.  No halo exchange and no all-reduce are
    included: this is the local compute of one shard."""
    n, ny = args.n, args.n // strips
    k = strips // 2 - 1
    xy, cells, _ = capi.tp_triag_mesh(n, ny, (0.0, k / strips, 1.0, (k + 1) / strips))
    i = np.arange(xy.shape[0]) % (n + 1)
    flags = ((i == 0) | (i == n)).astype(np.uint8)
    g = np.random.default_rng(args.seed).standard_normal(cells.shape[0])
    mesh, fe = capi.Mesh(ctx, xy, cells), capi.FE(ctx, 1)
    plan = capi.Plan(ctx, mesh, flags)
    theta = capi.Vec(ctx, g)
    alpha = capi.coeff(capi.COEFF_EXPCELL, 1.0, theta)
    gamma = capi.coeff(capi.COEFF_CONST, 1.0)
    f = capi.coeff(capi.COEFF_SINSIN, 1.0)
    nv, nnz = xy.shape[0], plan.csr.info()[2]

    def measure(pl, b, x, nrows, nnz_, alpha_):
        pl.assemble(fe, alpha_, gamma, f, b)
        asm = float(np.median([pl.assemble(fe, alpha_, gamma, f, b) for _ in range(10)]))
        xs, ys = capi.Vec(ctx, np.random.default_rng(1).standard_normal(nrows)), capi.Vec(ctx, nrows)
        spmv_us = 1e3 * capi.spmv_timed(ctx, pl.csr, xs, ys, args.spmv_reps)
        xs.close()
        ys.close()
        it = {}
        for name, md, single in (("two_sweep", capi.PCG_FP64, False),
                                 ("single_reduction", capi.PCG_FP64, True),
                                 ("mixed", capi.PCG_MIXED, False),
                                 ("mixed_matrix", capi.PCG_MIXED_MATRIX, False)):
            for _ in range(2):  # the first run also builds the solver's storage
                x.fill(0.0)
                inf = capi.pcg(ctx, pl.csr, b, x, rtol=1e-30, maxit=iters, check_every=128,
                               mode=md, raise_on_noconv=False, single_reduction=single)
            it[name] = 1e3 * inf.ms / max(1, inf.iters)
        return {"assembly_ms": asm, "spmv_us": spmv_us,
                "spmv_gbs_algorithmic": spmv_bytes(nnz_, nrows) / (spmv_us * 1e-6) / 1e9,
                "pcg_iteration_us": it}

    b, x = capi.Vec(ctx, nv), capi.Vec(ctx, nv)
    strip = measure(plan, b, x, nv, nnz, alpha)
    for h in (x, b, plan):
        h.close()
    log("strip %dx%d: assembly %.3f ms, spmv %.2f us, iteration (us) %s"
        % (n, ny, strip["assembly_ms"], strip["spmv_us"],
           ", ".join("%s %.2f" % kv for kv in strip["pcg_iteration_us"].items())))
    P = Problem(capi, ctx, n, args.seed)
    full = measure(P.plan, P.b, P.x, P.nv, P.nnz, P.alpha)
    P.close()
    log("full %dx%d: assembly %.3f ms, spmv %.2f us, iteration (us) %s"
        % (n, n, full["assembly_ms"], full["spmv_us"],
           ", ".join("%s %.2f" % kv for kv in full["pcg_iteration_us"].items())))
    for h in (theta, fe, mesh):
        h.close()
    share = {"assembly_ms": full["assembly_ms"] / strips, "spmv_us": full["spmv_us"] / strips,
             "pcg_iteration_us": {kk: v / strips for kk, v in full["pcg_iteration_us"].items()}}
    return {"metric": "one shard's local compute (assembly, SpMV, PCG iteration)",
            "config": {"workload": "interior %dx%d-cell strip of the %dx%d P1 reaction-diffusion "
                                   "problem (1/%d of it), a=exp(N(0,1)) per cell, "
                                   "f=sin(pi x)sin(pi y)" % (n, ny, n, n, strips),
                       "cells": int(cells.shape[0]), "dofs": int(nv), "nnz": int(nnz)},
            "data": "synthetic", "strip": strip, "full": full, "full_over_strips": share,
            "strip_vs_share": {"assembly": strip["assembly_ms"] / share["assembly_ms"],
                               "spmv": strip["spmv_us"] / share["spmv_us"],
                               "pcg_iteration": {kk: strip["pcg_iteration_us"][kk]
                                                 / share["pcg_iteration_us"][kk]
                                                 for kk in share["pcg_iteration_us"]}}}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=2048)
    ap.add_argument("--config", choices=["p1", "p3", "batched", "batched_p3", "adjoint", "traffic",
                                         "fkpp", "strip", "c1"],
                    default="p1")
    ap.add_argument("--mode", choices=["fp64", "mixed", "mixed_matrix"], default="fp64")
    ap.add_argument("--batch-mode", choices=["fp64", "mixed", "mixed_matrix"], default="mixed")
    ap.add_argument("--batch-modes", default="fp64,mixed,mixed_matrix",
                    help="batched: comma-separated precision modes to run")
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--no-solve", action="store_true", help="p3: assembly only")
    ap.add_argument("--no-plain", action="store_true",
                    help="batched: skip the per-problem plain FP64 comparison")
    ap.add_argument("--no-oracle", action="store_true",
                    help="skip the oracle's assembly and solves (parity at full size)")
    ap.add_argument("--c2-only", action="store_true",
                    help="default configuration: only the 2048^2 P1 problem (no P3 / batched)")
    ap.add_argument("--batch-graph", type=int, default=0, choices=[0, -1],
                    help="batched: 0 replays captured chunks, -1 plain launches")
    ap.add_argument("--batch-steps", type=int, default=3,
                    help="batched: timed repetitions per precision mode")
    ap.add_argument("--sweep-iters", type=int, default=200,
                    help="PCG iterations whose sweeps are timed one by one (roofline)")
    ap.add_argument("--fd-dirs", type=int, default=64)
    ap.add_argument("--spmv-reps", type=int, default=50)
    ap.add_argument("--fkpp-n", type=int, default=1024)
    ap.add_argument("--fkpp-steps", type=int, default=100)
    ap.add_argument("--fkpp-sample-n", type=int, default=256)
    ap.add_argument("--fkpp-mode", choices=["fp64", "mixed", "mixed_matrix"], default="fp64")
    ap.add_argument("--fkpp-nonlocal", action="store_true",
                    help="the nonlocal boundary terms (gain h, loss block L) instead of Robin h=1")
    ap.add_argument("--fkpp-modes", default="fp64,mixed,mixed_matrix",
                    help="fkpp: comma-separated precision modes to run")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: the library has no multi-GPU path; "
              "run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    from lehrfempp_amd import capi
    ctx = capi.Context(0)
    checks = []
    if args.config == "p1":
        out, chk = run_p1(args, capi, ctx)
        checks.append(chk)
        if not args.c2_only:
            # the P3 and batched configurations ride along as extra keys
            out["c3"], chk = run_p3(args, capi, ctx)
            checks.append(chk)
            out["c4"], chk = run_batched(args, capi, ctx)
            checks.append(chk)
            out["c4_p3"], chk = run_batched_p3(args, capi, ctx)
            checks.append(chk)
    elif args.config == "p3":
        out, chk = run_p3(args, capi, ctx)
        checks.append(chk)
    elif args.config == "batched":
        out, chk = run_batched(args, capi, ctx)
        checks.append(chk)
    elif args.config == "batched_p3":
        out, chk = run_batched_p3(args, capi, ctx)
        checks.append(chk)
    elif args.config == "traffic":
        out = run_traffic(args, capi, ctx)
    elif args.config == "fkpp":
        out = run_fkpp(args, capi, ctx)
    elif args.config == "strip":
        out = run_strip(args, capi, ctx)
    elif args.config == "c1":
        out = run_c1(args, capi, ctx)
    else:
        out = adjoint_run(capi, ctx, n=1024, fd_dirs=args.fd_dirs)
        out = {"metric": "adjoint gradient cost / forward solve", "config": {
            "workload": "P1 1024x1024, theta ~ N(0,1) per cell, u_obs = u(theta*) + 1% noise"},
            **out}
    # the oracle checks and the CPU baseline run after every GPU measurement,
    # so that no CPU-side work overlaps the timed device work
    for chk in checks:
        chk()
    ctx.close()
    print(json.dumps(out))


if __name__ == "__main__":
    main()
