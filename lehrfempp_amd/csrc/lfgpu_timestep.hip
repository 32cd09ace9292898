// Time stepping on a fixed pattern: Strang splitting for the Fisher/KPP
// equation (FisherKPP::StrangSplit, projects/fisher_kpp/strangsplitting.h:
// 28-76, :176-201 and strangsplitting.cc:15-94).
//
// Diffusion M u' = -A u advances by the two-stage SDIRK-2 method (kappa =
// 1 - sqrt(2)/2, stage matrix S = M + h kappa A for the step h), the reaction
// u' = lambda u (1 - u/K) by its exact flow.  The two system matrices are
// formed once on the shared pattern; every diffusion step then costs
//   y  = A mu                          (SpMV)
//   S kt1 = y                          (PCG, kt1 = -k1 of the reference)
//   b2 = y - h (1 - kappa) A kt1       (SpMV + fused update)
//   S kt2 = b2                         (PCG, kt2 = -k2)
//   out = (mu - h (1-kappa) kt1) - h kappa kt2   [+ the reaction, fused]
// Working with the negated stages saves a negation pass and changes no bit:
// negation is exact, and CG on (-b, -x0) produces the negated iterates of
// CG on (b, x0).  The update reproduces the reference's
// (mu + s1 k1) + s2 k2 exactly for equal stages; the reaction evaluates
// cap / (1 + (cap / u - 1) e) in the reference's order with e = exp(-lambda
// tau) (or exp(-lambda tau / 2)) taken on the host, so it is bit-identical to
// the reference's formula.
//
// Stage solves are warm-started: stage 1 from the previous step's stage 2,
// stage 2 from this step's stage 1 (the reference factorises S once and
// solves directly; a PCG start vector changes only the iteration count).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>

#include "lfgpu_internal.h"
#include "lfgpu_reduce.h"

using namespace lfgpu;

struct lfgpu_strang {
  lfgpu_ctx *ctx = nullptr;
  lfgpu_csr *A = nullptr, *S1 = nullptr, *S2 = nullptr;  // owned
  int32_t n = 0;
  int m = 0;
  double T = 0.0, tau = 0.0, kappa = 0.0, lambda = 0.0;
  lfgpu_pcg_opts opts{};
  // device work vectors, n each: y, b2, kt1, kt2, cur
  double *work = nullptr;
  bool warm = false;  // kt2 holds a previous stage
  // iterations of the last solve of each stage: the next solve of that stage
  // checks for convergence about 12 times over a similar count (fewer host
  // round trips on long solves, little overshoot on short ones)
  int last_iters[2] = {0, 0};
  // timing events of the stepper's own calls (the solves record the
  // context's pair)
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
};

namespace {

__global__ void k_stage2_rhs(const double *y, const double *z, double s, double *b2,
                             int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) b2[i] = y[i] - s * z[i];
}

// out = (mu - s1 kt1) - s2 kt2, then (react) the logistic flow with factor e
__global__ void k_stage_update(const double *mu, const double *kt1, const double *kt2, double s1,
                               double s2, int react, const double *cap, double e, double *out,
                               int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double v = (mu[i] - s1 * kt1[i]) - s2 * kt2[i];
  if (react) {
    const double K = cap[i];
    v = K / (1.0 + (K / v - 1.0) * e);
  }
  out[i] = v;
}

__global__ void k_copy_vec(const double *x, double *y, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] = x[i];
}

// options of the next solve of stage q: the caller's, with the check
// interval adapted unless the caller fixed one
lfgpu_pcg_opts stage_opts(const lfgpu_strang *st, int q) {
  lfgpu_pcg_opts o = st->opts;
  if (o.check_every <= 0) {
    int c = st->last_iters[q] > 0 ? st->last_iters[q] / 12 : 16;
    c = std::max(8, std::min(128, c));
    o.check_every = c + (c & 1);
  }
  return o;
}

void acc_info(lfgpu_strang_info *info, const lfgpu_pcg_info &pi) {
  info->solves += 1;
  info->iters += pi.iters;
  info->max_iters = std::max(info->max_iters, pi.iters);
  if (!pi.converged) info->converged = 0;
  info->worst_relres = std::max(info->worst_relres, pi.relres);
}

// One diffusion step from mu (device, n) into out; react selects the fused
// reaction with factor e.  out may alias mu.
int diffusion_step(lfgpu_strang *st, bool firstcall, const double *mu, double *out, bool react,
                   const double *cap, double e, lfgpu_strang_info *info) {
  lfgpu_ctx *ctx = st->ctx;
  hipStream_t s = ctx->stream;
  const int64_t n = st->n;
  double *y = st->work, *b2 = st->work + n, *kt1 = st->work + 2 * n, *kt2 = st->work + 3 * n;
  lfgpu_csr *S = firstcall ? st->S1 : st->S2;
  // the reference's scalars: 0.5 * tau_ * (1 - kappa_) and 0.5 * tau_ * kappa_
  // (first call), tau_ * (1 - kappa_) and tau_ * kappa_ (later calls)
  const double h = firstcall ? 0.5 * st->tau : st->tau;
  const double s1 = h * (1.0 - st->kappa), s2 = h * st->kappa;
  const unsigned G = grid_for(n, 256);
  LFG_TRY(spmv_async(ctx, st->A, mu, y));
  if (st->warm)
    hipLaunchKernelGGL(k_copy_vec, dim3(G), dim3(256), 0, s, kt2, kt1, n);
  else
    LFG_HIP(hipMemsetAsync(kt1, 0, (size_t)n * sizeof(double), s));
  LFG_HIP(hipGetLastError());
  lfgpu_vec vb{ctx, y, n}, vx{ctx, kt1, n};
  lfgpu_pcg_info pi{};
  lfgpu_pcg_opts o1 = stage_opts(st, 0);
  int rc = lfgpu_pcg(ctx, S, &vb, &vx, &o1, &pi);
  if (rc != LFGPU_OK && rc != LFGPU_ENOCONV) return rc;
  acc_info(info, pi);
  st->last_iters[0] = pi.iters;
  if (rc == LFGPU_ENOCONV) return rc;
  // b2 = y - s1 A kt1 (b2 holds A kt1 first)
  LFG_TRY(spmv_async(ctx, st->A, kt1, b2));
  hipLaunchKernelGGL(k_stage2_rhs, dim3(G), dim3(256), 0, s, y, b2, s1, b2, n);
  hipLaunchKernelGGL(k_copy_vec, dim3(G), dim3(256), 0, s, kt1, kt2, n);
  LFG_HIP(hipGetLastError());
  lfgpu_vec vb2{ctx, b2, n}, vx2{ctx, kt2, n};
  lfgpu_pcg_opts o2 = stage_opts(st, 1);
  rc = lfgpu_pcg(ctx, S, &vb2, &vx2, &o2, &pi);
  if (rc != LFGPU_OK && rc != LFGPU_ENOCONV) return rc;
  acc_info(info, pi);
  st->last_iters[1] = pi.iters;
  if (rc == LFGPU_ENOCONV) return rc;
  st->warm = true;
  hipLaunchKernelGGL(k_stage_update, dim3(G), dim3(256), 0, s, mu, kt1, kt2, s1, s2,
                     react ? 1 : 0, cap, e, out, n);
  LFG_HIP(hipGetLastError());
  info->diffusion_steps += 1;
  return LFGPU_OK;
}

void info_init(lfgpu_strang_info *info) {
  *info = lfgpu_strang_info{};
  info->converged = 1;
}

int finish(lfgpu_strang *st, int rc, lfgpu_strang_info *info, lfgpu_strang_info *out) {
  lfgpu_ctx *ctx = st->ctx;
  hipError_t e = hipEventRecord(st->ev1, ctx->stream);
  if (e == hipSuccess) e = hipEventSynchronize(st->ev1);
  float ms = 0.f;
  if (e == hipSuccess) e = hipEventElapsedTime(&ms, st->ev0, st->ev1);
  if (e != hipSuccess && rc == LFGPU_OK) rc = hip_fail(e, "time stepping");
  info->ms = ms;
  ctx->last_ms = ms;
  if (out) *out = *info;
  if (rc == LFGPU_ENOCONV) set_error("a stage solve of the time stepper did not converge");
  return rc;
}

}  // namespace

extern "C" int lfgpu_strang_create(lfgpu_ctx *ctx, const lfgpu_csr *A, const lfgpu_csr *M,
                                   double T, int m, double lambda, const lfgpu_pcg_opts *opts,
                                   lfgpu_strang **out) {
  LFG_REQUIRE(ctx && A && M && opts && out, LFGPU_EINVAL, "NULL argument to lfgpu_strang_create");
  *out = nullptr;
  LFG_REQUIRE(A->nrows == A->ncols && M->nrows == A->nrows && M->ncols == A->ncols, LFGPU_EINVAL,
              "A and M must be square matrices of one size");
  LFG_REQUIRE(m >= 1, LFGPU_EINVAL, "the number of time steps must be at least 1");
  LFG_REQUIRE(std::isfinite(T) && std::isfinite(lambda), LFGPU_EINVAL,
              "T and lambda must be finite");
  LFG_REQUIRE(opts->rtol >= 0.0 && opts->maxit >= 0, LFGPU_EINVAL, "bad PCG options");
  LFG_REQUIRE((A->blk_n == 0 && M->blk_n == 0) ||
                  (opts->mode == LFGPU_PCG_FP64 && !opts->single_reduction),
              LFGPU_EINVAL,
              "stage matrices with a dense block are solved in FP64 mode with two sweeps only");
  lfgpu_strang *st = new lfgpu_strang();
  st->ctx = ctx;
  st->n = A->nrows;
  st->m = m;
  st->T = T;
  st->lambda = lambda;
  st->kappa = 1.0 - 0.5 * std::sqrt(2.0);
  st->tau = T / m;
  st->opts = *opts;
  int rc = lfgpu_csr_clone(ctx, A, &st->A);
  // M_ + 0.5 * tau_ * kappa_ * A_  and  M_ + tau_ * kappa_ * A_
  if (rc == LFGPU_OK) rc = lfgpu_csr_axpby(ctx, 1.0, M, 0.5 * st->tau * st->kappa, A, &st->S1);
  if (rc == LFGPU_OK) rc = lfgpu_csr_axpby(ctx, 1.0, M, st->tau * st->kappa, A, &st->S2);
  if (rc == LFGPU_OK)
    rc = dalloc(&st->work, (size_t)5 * std::max<int32_t>(st->n, 1), "time-stepping vectors");
  if (rc == LFGPU_OK) {
    hipError_t e = hipEventCreate(&st->ev0);
    if (e == hipSuccess) e = hipEventCreate(&st->ev1);
    if (e != hipSuccess) rc = hip_fail(e, "lfgpu_strang_create events");
  }
  if (rc != LFGPU_OK) {
    lfgpu_strang_destroy(st);
    return rc;
  }
  *out = st;
  return LFGPU_OK;
}

extern "C" int lfgpu_strang_destroy(lfgpu_strang *st) {
  if (!st) return LFGPU_OK;
  if (st->ctx) (void)hipStreamSynchronize(st->ctx->stream);
  lfgpu_csr_destroy(st->A);
  lfgpu_csr_destroy(st->S1);
  lfgpu_csr_destroy(st->S2);
  dfree(st->work);
  if (st->ev0) (void)hipEventDestroy(st->ev0);
  if (st->ev1) (void)hipEventDestroy(st->ev1);
  delete st;
  return LFGPU_OK;
}

extern "C" int lfgpu_strang_diffusion(lfgpu_ctx *ctx, lfgpu_strang *st, int firstcall,
                                      const lfgpu_vec *mu, lfgpu_vec *out,
                                      lfgpu_strang_info *info) {
  LFG_REQUIRE(ctx && st && mu && out, LFGPU_EINVAL, "NULL argument to lfgpu_strang_diffusion");
  LFG_REQUIRE(ctx == st->ctx, LFGPU_EINVAL, "the stepper belongs to another context");
  LFG_REQUIRE(mu->n == st->n && out->n == st->n, LFGPU_EINVAL, "vector size mismatch");
  lfgpu_strang_info acc;
  info_init(&acc);
  LFG_HIP(hipEventRecord(st->ev0, ctx->stream));
  const int rc = diffusion_step(st, firstcall != 0, mu->d, out->d, false, nullptr, 0.0, &acc);
  return finish(st, rc, &acc, info);
}

extern "C" int lfgpu_strang_evolution(lfgpu_ctx *ctx, lfgpu_strang *st, const lfgpu_vec *cap,
                                      const lfgpu_vec *mu, lfgpu_vec *out,
                                      lfgpu_strang_info *info) {
  LFG_REQUIRE(ctx && st && cap && mu && out, LFGPU_EINVAL,
              "NULL argument to lfgpu_strang_evolution");
  LFG_REQUIRE(ctx == st->ctx, LFGPU_EINVAL, "the stepper belongs to another context");
  LFG_REQUIRE(cap->n == st->n && mu->n == st->n && out->n == st->n, LFGPU_EINVAL,
              "vector size mismatch");
  lfgpu_strang_info acc;
  info_init(&acc);
  hipStream_t s = ctx->stream;
  const int64_t n = st->n;
  double *cur = st->work + 4 * n;
  // exp(-lambda_ * tau_) and exp(-lambda_ * tau_ / 2.) as the reference
  // evaluates them
  const double e_full = std::exp(-st->lambda * st->tau);
  const double e_half = std::exp(-st->lambda * st->tau / 2.);
  // diffusion steps: the first (tau / 2) and m - 2 more (tau); each is
  // followed by a reaction over tau, the last by one over tau / 2
  const int nfull = st->m > 2 ? st->m - 2 : 0;
  LFG_HIP(hipEventRecord(st->ev0, s));
  int rc = diffusion_step(st, true, mu->d, cur, true, cap->d, nfull > 0 ? e_full : e_half, &acc);
  for (int i = 1; i <= nfull && rc == LFGPU_OK; ++i)
    rc = diffusion_step(st, false, cur, cur, true, cap->d, i < nfull ? e_full : e_half, &acc);
  if (rc == LFGPU_OK) {
    hipLaunchKernelGGL(k_copy_vec, dim3(grid_for(n, 256)), dim3(256), 0, s, cur, out->d, n);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) rc = hip_fail(e, "lfgpu_strang_evolution");
  }
  return finish(st, rc, &acc, info);
}
