// Sparse matrices and solvers: device setFromTriplets, a sliced-ELL SpMV,
// deterministic dot products, Jacobi-PCG (FP64 and FP32 with FP64 reliable
// updates) and the adjoint-gradient kernel.
//
// Solver data layout.  The CSR of a matrix is mirrored once into a sliced
// ELL layout with slices of 64 rows (one wavefront): slice s stores
// width_s = max row length of its rows, entries column-major inside the
// slice (element k of row 64 s + l at off_s + 64 k + l), padding entries
// with value 0 and column = row.  Every load of the SpMV is then a fully
// coalesced 64-lane access.  The FP32 mirror shares the column array.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include <rocprim/rocprim.hpp>

#include "lfgpu_fe_device.h"
#include "lfgpu_internal.h"

using namespace lfgpu;

namespace {

constexpr int kSlice = 64;
constexpr int kRedBlocks = 1024;  // fixed grid of the deterministic reductions
constexpr int kMaxGrid = 2048;   // workgroups of the solver sweeps
constexpr int kPartSlots = 6;    // partial-sum arrays of kMaxGrid doubles

inline unsigned grid_for(int64_t n, int block) {
  int64_t g = (n + block - 1) / block;
  return (unsigned)(g < 1 ? 1 : g);
}

// ---------------------------------------------------------------------------
// block reductions (deterministic: fixed shape, fixed order)
// ---------------------------------------------------------------------------
template <int BLOCK>
__device__ __forceinline__ double block_sum(double v, double *sh) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  __syncthreads();
  if (l == 0) sh[w] = v;
  __syncthreads();
  double t = 0.0;
  if (threadIdx.x == 0) {
#pragma unroll
    for (int q = 0; q < BLOCK / 64; ++q) t += sh[q];
  }
  return t;  // valid in thread 0
}

__global__ void k_dot_partial(const double *x, const double *y, int64_t n, double *part) {
  __shared__ double sh[4];
  double v = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256)
    v += x[i] * y[i];
  v = block_sum<256>(v, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = v;
}

__global__ void k_sum_partials(const double *part, int n, double *out) {
  __shared__ double sh[4];
  double v = 0.0;
  for (int i = threadIdx.x; i < n; i += 256) v += part[i];
  v = block_sum<256>(v, sh);
  if (threadIdx.x == 0) *out = v;
}

__global__ void k_axpby(double a, const double *x, double b, double *y, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] = a * x[i] + b * y[i];
}

}  // namespace

// ===========================================================================
// solver cache (sliced ELL mirror + work vectors)
// ===========================================================================
// Solver scalars.  Every kernel recomputes the scalars it needs from the
// previous kernel's per-workgroup partial sums (identical order in every
// workgroup, so all workgroups agree bit for bit); workgroup 0 alone writes
// this state, and values that are read and written within one iteration are
// double buffered by iteration parity.
struct PcgState {
  double rho[2];   // (r, D^-1 r), indexed by iteration parity
  double thr;      // tol^2 ||b||^2 (on entry: tol^2)
  double bb;       // ||b||^2
  double rr;       // last ||r||^2 (recurrence in FP64 mode, true in mixed)
  double rr_ref;   // mixed: ||r||^2 at the last reliable update
  int iter, done, converged, maxit;
  int refine[2];   // mixed: reliable update decided in this iteration (parity)
  int iterp[2];    // mixed: iteration count, indexed by parity
  int refinements;
  int last_refine; // mixed: iteration of the last reliable update
};

struct lfgpu_solver_cache {
  int32_t nslices = 0;
  int64_t ell_len = 0;
  int32_t *slice_off = nullptr;  // [nslices + 1] element offsets
  int32_t *s_col = nullptr;
  double *s_val = nullptr;
  float *s_val32 = nullptr;
  double *dinv = nullptr;
  float *dinv32 = nullptr;
  double *rowsum = nullptr;
  float *rowsum32 = nullptr;
  uint64_t built_version = ~0ull;
  bool have32 = false;
  // work vectors
  double *r = nullptr, *p = nullptr, *q = nullptr;
  float *r32 = nullptr, *p32 = nullptr, *q32 = nullptr, *xa32 = nullptr;
  double *xlo = nullptr;   // mixed: low part of the double-double solution
  double *part = nullptr;  // [kPartSlots x grid] partial sums
  int part_len = 0;
  PcgState *st = nullptr;
};

namespace lfgpu {
void solver_cache_free(lfgpu_csr *A) {
  lfgpu_solver_cache *c = A->cache;
  if (!c) return;
  dfree(c->slice_off);
  dfree(c->s_col);
  dfree(c->s_val);
  dfree(c->s_val32);
  dfree(c->dinv);
  dfree(c->dinv32);
  dfree(c->rowsum);
  dfree(c->rowsum32);
  dfree(c->r);
  dfree(c->p);
  dfree(c->q);
  dfree(c->r32);
  dfree(c->p32);
  dfree(c->q32);
  dfree(c->xa32);
  dfree(c->xlo);
  dfree(c->part);
  dfree(c->st);
  delete c;
  A->cache = nullptr;
}
}  // namespace lfgpu

namespace {

__global__ void k_slice_width(const int32_t *row_ptr, const int32_t *col, int32_t n,
                              int32_t nslices, int32_t *w) {
  const int32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= nslices) return;
  int32_t m = 0;
  const int32_t r0 = s * kSlice, r1 = min(r0 + kSlice, n);
  for (int32_t r = r0; r < r1; ++r) {
    int32_t k = 0;
    for (int32_t p = row_ptr[r]; p < row_ptr[r + 1]; ++p) k += (col[p] != r);
    m = max(m, k);
  }
  w[s] = m * kSlice;
}

// Off-diagonal columns in CSR order; padding points at the row itself.
__global__ void k_fill_ell_cols(const int32_t *row_ptr, const int32_t *col, int32_t n,
                                const int32_t *slice_off, int32_t *s_col) {
  const int32_t r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  const int32_t s = r / kSlice, l = r % kSlice;
  const int32_t off = slice_off[s], w = (slice_off[s + 1] - off) / kSlice;
  int32_t k = 0;
  for (int32_t p = row_ptr[r]; p < row_ptr[r + 1]; ++p)
    if (col[p] != r) s_col[off + (k++) * kSlice + l] = col[p];
  for (; k < w; ++k) s_col[off + k * kSlice + l] = r;
}

// Off-diagonal values, the row sums s_i = sum_j a_ij (Neumaier-compensated,
// so that s_i is accurate to a few ulp of itself even when the row nearly
// cancels) and the Jacobi inverse diagonal (Eigen DiagonalPreconditioner:
// 1/a_ii, or 1 when a_ii is absent or 0).
__global__ void k_fill_ell_vals(const int32_t *row_ptr, const int32_t *col, const double *val,
                                int32_t n, const int32_t *slice_off, double *s_val,
                                float *s_val32, double *rowsum, float *rowsum32, double *dinv,
                                float *dinv32) {
  const int32_t r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  const int32_t s = r / kSlice, l = r % kSlice;
  const int32_t off = slice_off[s], w = (slice_off[s + 1] - off) / kSlice;
  double d = 1.0, sum = 0.0, comp = 0.0;
  int32_t k = 0;
  for (int32_t p = row_ptr[r]; p < row_ptr[r + 1]; ++p) {
    const double v = val[p];
    const double t = sum + v;
    comp += fabs(sum) >= fabs(v) ? (sum - t) + v : (v - t) + sum;
    sum = t;
    if (col[p] == r) {
      if (v != 0.0) d = 1.0 / v;
    } else {
      s_val[off + k * kSlice + l] = v;
      if (s_val32) s_val32[off + k * kSlice + l] = (float)v;
      ++k;
    }
  }
  for (; k < w; ++k) {
    s_val[off + k * kSlice + l] = 0.0;
    if (s_val32) s_val32[off + k * kSlice + l] = 0.0f;
  }
  const double rs = sum + comp;
  rowsum[r] = rs;
  dinv[r] = d;
  if (rowsum32) rowsum32[r] = (float)rs;
  if (dinv32) dinv32[r] = (float)d;
}

int ensure_cache(lfgpu_ctx *ctx, lfgpu_csr *A, bool need32) {
  hipStream_t s = ctx->stream;
  lfgpu_solver_cache *c = A->cache;
  const int32_t n = A->nrows;
  if (!c) {
    c = new lfgpu_solver_cache();
    A->cache = c;
    c->nslices = (n + kSlice - 1) / kSlice;
    int32_t *w = nullptr;
    LFG_TRY(dalloc(&w, (size_t)c->nslices, "slice widths"));
    LFG_TRY(dalloc(&c->slice_off, (size_t)c->nslices + 1, "slice offsets"));
    hipLaunchKernelGGL(k_slice_width, dim3(grid_for(c->nslices, 256)), dim3(256), 0, s,
                       A->row_ptr, A->col, n, c->nslices, w);
    LFG_HIP(hipGetLastError());
    int64_t len = 0;
    int rc = exclusive_scan_i32(s, w, c->slice_off, c->nslices, &len);
    dfree(w);
    LFG_TRY(rc);
    c->ell_len = len;
    LFG_TRY(dalloc(&c->s_col, (size_t)std::max<int64_t>(len, 1), "ELL columns"));
    LFG_TRY(dalloc(&c->s_val, (size_t)std::max<int64_t>(len, 1), "ELL values"));
    LFG_TRY(dalloc(&c->rowsum, (size_t)n, "row sums"));
    LFG_TRY(dalloc(&c->dinv, (size_t)n, "inverse diagonal"));
    hipLaunchKernelGGL(k_fill_ell_cols, dim3(grid_for(n, 256)), dim3(256), 0, s, A->row_ptr,
                       A->col, n, c->slice_off, c->s_col);
    LFG_HIP(hipGetLastError());
    LFG_TRY(dalloc(&c->r, (size_t)n, "PCG r"));
    LFG_TRY(dalloc(&c->p, (size_t)n, "PCG p"));
    LFG_TRY(dalloc(&c->q, (size_t)n, "PCG q"));
    c->part_len = kPartSlots * kMaxGrid;
    LFG_TRY(dalloc(&c->part, (size_t)c->part_len, "reduction partials"));
    LFG_TRY(dalloc(&c->st, 1, "PCG state"));
  }
  if (need32 && !c->have32) {
    LFG_TRY(dalloc(&c->s_val32, (size_t)std::max<int64_t>(c->ell_len, 1), "ELL values FP32"));
    LFG_TRY(dalloc(&c->rowsum32, (size_t)n, "row sums FP32"));
    LFG_TRY(dalloc(&c->dinv32, (size_t)n, "inverse diagonal FP32"));
    LFG_TRY(dalloc(&c->r32, (size_t)n, "PCG r32"));
    LFG_TRY(dalloc(&c->p32, (size_t)n, "PCG p32"));
    LFG_TRY(dalloc(&c->q32, (size_t)n, "PCG q32"));
    LFG_TRY(dalloc(&c->xa32, (size_t)n, "PCG x32"));
    LFG_TRY(dalloc(&c->xlo, (size_t)n, "PCG x low part"));
    c->have32 = true;
    c->built_version = ~0ull;
  }
  if (c->built_version != A->val_version) {
    hipLaunchKernelGGL(k_fill_ell_vals, dim3(grid_for(n, 256)), dim3(256), 0, s, A->row_ptr,
                       A->col, A->val, n, c->slice_off, c->s_val, c->have32 ? c->s_val32 : nullptr,
                       c->rowsum, c->have32 ? c->rowsum32 : nullptr, c->dinv,
                       c->have32 ? c->dinv32 : nullptr);
    LFG_HIP(hipGetLastError());
    c->built_version = A->val_version;
  }
  return LFGPU_OK;
}

// ---------------------------------------------------------------------------
// SpMV on the sliced ELL mirror, in difference form:
//   (A x)_i = sum_{j != i} a_ij (x_j - x_i) + s_i x_i,   s_i = sum_j a_ij.
// Algebraically identical to sum_j a_ij x_j; for stiffness matrices, whose
// rows nearly cancel (s_i = O(h^2) against entries O(1)), it avoids the
// catastrophic cancellation of the plain form on smooth x, which would
// otherwise put the attainable residual of a 2048^2 FP64 solve near 1e-10
// and make an FP32 product useless for the smooth modes.
// ---------------------------------------------------------------------------
template <class V, class X, class S>
__device__ __forceinline__ X ell_row(const int32_t *slice_off, const int32_t *s_col, const V *s_val,
                                     const S *rowsum, const X *x, int32_t r, X xr) {
  const int32_t s = r >> 6, l = r & 63;
  const int32_t off = slice_off[s];
  const int32_t w = (slice_off[s + 1] - off) >> 6;
  const V *vp = s_val + off + l;
  const int32_t *cp = s_col + off + l;
  X acc = 0;
  for (int32_t k = 0; k < w; ++k) acc += (X)vp[64 * k] * (x[cp[64 * k]] - xr);
  return acc + (X)rowsum[r] * xr;
}

// FP32 storage, FP64 arithmetic: products of two FP32 numbers are exact in
// FP64 and the row is summed in FP64, so the only FP32 rounding left in q is
// the final conversion.  Summing in FP32 instead leaves a relative error of
// ~eps32 / h on smooth vectors (7e-4 at h = 1/2048), enough to destroy the
// conjugacy of the CG recurrences.
__device__ __forceinline__ float ell_row32_acc64(const int32_t *slice_off, const int32_t *s_col,
                                                 const float *s_val, const float *rowsum,
                                                 const float *x, int32_t r, float xr) {
  const int32_t s = r >> 6, l = r & 63;
  const int32_t off = slice_off[s];
  const int32_t w = (slice_off[s + 1] - off) >> 6;
  const float *vp = s_val + off + l;
  const int32_t *cp = s_col + off + l;
  const double xd = (double)xr;
  double acc = 0.0;
  for (int32_t k = 0; k < w; ++k) acc += (double)vp[64 * k] * ((double)x[cp[64 * k]] - xd);
  return (float)(acc + (double)rowsum[r] * xd);
}

__global__ __launch_bounds__(256) void k_spmv64(const int32_t *slice_off, const int32_t *s_col,
                                                const double *s_val, const double *rowsum,
                                                const double *x, double *y, int32_t n) {
  const int32_t r = blockIdx.x * 256 + threadIdx.x;
  if (r >= n) return;
  y[r] = ell_row<double, double, double>(slice_off, s_col, s_val, rowsum, x, r, x[r]);
}

// ---------------------------------------------------------------------------
// PCG sweeps.  A fixed grid of G workgroups; workgroup g owns the contiguous
// rows [g R, (g+1) R) (R a multiple of 256, so every wavefront covers one
// 64-row ELL slice) and writes one partial sum per reduced quantity.
// ---------------------------------------------------------------------------
struct Sweep {
  int32_t n;
  int32_t rows_per_group;
  int ngroups;
};

// Sum of ngroups partials, identical in every workgroup, broadcast to all
// threads of the calling workgroup.
__device__ __forceinline__ double all_reduce_partials(const double *part, int ngroups,
                                                      double *sh) {
  double v = 0.0;
  for (int i = threadIdx.x; i < ngroups; i += 256) v += part[i];
  v = block_sum<256>(v, sh);
  __syncthreads();
  if (threadIdx.x == 0) sh[4] = v;
  __syncthreads();
  return sh[4];
}

__device__ __forceinline__ void store_partial(double *part, double v, double *sh) {
  v = block_sum<256>(v, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = v;
}

struct PcgArgs {
  const int32_t *slice_off;
  const int32_t *s_col;
  const double *s_val;
  const double *rowsum;
  const double *dinv;
  double *x, *r, *p, *q;
  double *part;  // slots: 0 pq, 1 rr, 2 rz, 3 bb
  PcgState *st;
  Sweep sw;
};

// setup: r = b - A x, p = D^-1 r; partials of b.b, r.r, r.D^-1 r
__global__ __launch_bounds__(256) void k_pcg_init(PcgArgs a, const double *b) {
  __shared__ double sh[8];
  const int32_t r0 = blockIdx.x * a.sw.rows_per_group;
  const int32_t r1 = min(r0 + a.sw.rows_per_group, a.sw.n);
  double bb = 0.0, rr = 0.0, rz = 0.0;
  for (int32_t i = r0 + threadIdx.x; i < r1; i += 256) {
    const double xi = a.x[i];
    const double ax =
        ell_row<double, double, double>(a.slice_off, a.s_col, a.s_val, a.rowsum, a.x, i, xi);
    const double bi = b[i];
    const double ri = bi - ax;
    a.r[i] = ri;
    const double zi = a.dinv[i] * ri;
    a.p[i] = zi;
    bb += bi * bi;
    rr += ri * ri;
    rz += ri * zi;
  }
  store_partial(a.part + 3 * kMaxGrid, bb, sh);
  store_partial(a.part + 1 * kMaxGrid, rr, sh);
  store_partial(a.part + 2 * kMaxGrid, rz, sh);
}

// one workgroup: finish the setup reductions (Eigen's early exits)
__global__ __launch_bounds__(256) void k_pcg_init_finish(PcgArgs a) {
  __shared__ double sh[8];
  const double bb = all_reduce_partials(a.part + 3 * kMaxGrid, a.sw.ngroups, sh);
  const double rr = all_reduce_partials(a.part + 1 * kMaxGrid, a.sw.ngroups, sh);
  const double rz = all_reduce_partials(a.part + 2 * kMaxGrid, a.sw.ngroups, sh);
  if (threadIdx.x == 0) {
    PcgState *st = a.st;
    st->bb = bb;
    st->rr = rr;
    st->rr_ref = rr;
    st->rho[0] = rz;
    st->iter = 0;
    st->converged = 0;
    st->done = 0;
    st->refinements = 0;
    st->last_refine = 0;
    st->refine[0] = st->refine[1] = 0;
    st->iterp[0] = st->iterp[1] = 0;
    const double thr = st->thr * bb;  // thr holds tol^2 on entry
    st->thr = thr > 2.2250738585072014e-308 ? thr : 2.2250738585072014e-308;
    if (bb == 0.0 || rr < st->thr) {
      st->done = 1;
      st->converged = 1;
    } else if (st->maxit <= 0) {
      st->done = 1;
    }
  }
}

__global__ void k_zero_if_bb0(double *x, int32_t n, const PcgState *st) {
  const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n && st->bb == 0.0) x[i] = 0.0;
}

// q = A p ; partial p.q
__global__ __launch_bounds__(256) void k_pcg_spmv(PcgArgs a) {
  __shared__ double sh[8];
  if (a.st->done) return;
  const int32_t r0 = blockIdx.x * a.sw.rows_per_group;
  const int32_t r1 = min(r0 + a.sw.rows_per_group, a.sw.n);
  double v = 0.0;
  for (int32_t r = r0 + threadIdx.x; r < r1; r += 256) {
    const double pr = a.p[r];
    const double qr =
        ell_row<double, double, double>(a.slice_off, a.s_col, a.s_val, a.rowsum, a.p, r, pr);
    a.q[r] = qr;
    v += pr * qr;
  }
  store_partial(a.part, v, sh);
}

// alpha = rho / (p.q) ; x += alpha p ; r -= alpha q ; partials r.r, r.D^-1 r
__global__ __launch_bounds__(256) void k_pcg_update(PcgArgs a, int par) {
  __shared__ double sh[8];
  if (a.st->done) return;
  const double alpha = a.st->rho[par] / all_reduce_partials(a.part, a.sw.ngroups, sh);
  const int32_t r0 = blockIdx.x * a.sw.rows_per_group;
  const int32_t r1 = min(r0 + a.sw.rows_per_group, a.sw.n);
  double rr = 0.0, rz = 0.0;
  for (int32_t i = r0 + threadIdx.x; i < r1; i += 256) {
    a.x[i] += alpha * a.p[i];
    const double ri = a.r[i] - alpha * a.q[i];
    a.r[i] = ri;
    rr += ri * ri;
    rz += ri * (a.dinv[i] * ri);
  }
  store_partial(a.part + 1 * kMaxGrid, rr, sh);
  store_partial(a.part + 2 * kMaxGrid, rz, sh);
}

// Eigen's convergence test on the recurrence residual, then
// beta = rho' / rho ; p = D^-1 r + beta p
__global__ __launch_bounds__(256) void k_pcg_pupdate(PcgArgs a, int par) {
  __shared__ double sh[8];
  PcgState *st = a.st;
  if (st->done) return;
  const double rr = all_reduce_partials(a.part + 1 * kMaxGrid, a.sw.ngroups, sh);
  const double rz = all_reduce_partials(a.part + 2 * kMaxGrid, a.sw.ngroups, sh);
  if (rr < st->thr) {
    if (blockIdx.x == 0 && threadIdx.x == 0) {
      st->rr = rr;
      st->done = 1;
      st->converged = 1;
    }
    return;
  }
  const double beta = rz / st->rho[par];
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    st->rr = rr;
    st->rho[par ^ 1] = rz;
    st->iter += 1;
    if (st->iter >= st->maxit) st->done = 1;
  }
  const int32_t r0 = blockIdx.x * a.sw.rows_per_group;
  const int32_t r1 = min(r0 + a.sw.rows_per_group, a.sw.n);
  for (int32_t i = r0 + threadIdx.x; i < r1; i += 256) a.p[i] = a.dinv[i] * a.r[i] + beta * a.p[i];
}

// ---------------------------------------------------------------------------
// mixed precision: FP32 matrix, vectors and recurrences (FP64 dot products),
// solution kept as a double-double pair (x, xlo) so that its true residual is
// not limited by rounding x to FP64; reliable updates recompute the true
// residual b - A x - A xlo in FP64 whenever the recurrence residual has
// dropped by delta (or after a caller-set number of iterations).
// ---------------------------------------------------------------------------
struct MixArgs {
  const int32_t *slice_off;
  const int32_t *s_col;
  const float *s_val32;
  const double *s_val;
  const float *rowsum32;
  const double *rowsum;
  const float *dinv32;
  const double *dinv;
  double *x;          // high part of the solution (the caller's vector)
  double *xlo;        // low part
  const double *b;
  double *r64;        // FP64 true residual
  float *r, *p, *q, *xa;
  double *part;       // slots: 0 pq, 1 rr, 2 rz, 4 rr64, 5 rz64
  PcgState *st;
  Sweep sw;
  double delta2;      // reliable update when rr < delta2 * rr_ref
  int max_interval;   // ... or after this many iterations
};

__global__ __launch_bounds__(256) void k_mix_setup(MixArgs a) {
  const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.sw.n) return;
  const float ri = (float)a.r64[i];
  a.r[i] = ri;
  a.p[i] = a.dinv32[i] * ri;
  a.xa[i] = 0.0f;
  a.xlo[i] = 0.0;
}

__global__ __launch_bounds__(256) void k_mix_spmv(MixArgs a) {
  __shared__ double sh[8];
  if (a.st->done) return;
  const int32_t r0 = blockIdx.x * a.sw.rows_per_group;
  const int32_t r1 = min(r0 + a.sw.rows_per_group, a.sw.n);
  double v = 0.0;
  for (int32_t r = r0 + threadIdx.x; r < r1; r += 256) {
    const float pr = a.p[r];
    const float qr = ell_row32_acc64(a.slice_off, a.s_col, a.s_val32, a.rowsum32, a.p, r, pr);
    a.q[r] = qr;
    v += (double)pr * (double)qr;
  }
  store_partial(a.part, v, sh);
}

__global__ __launch_bounds__(256) void k_mix_update(MixArgs a, int par) {
  __shared__ double sh[8];
  if (a.st->done) return;
  const float alpha = (float)(a.st->rho[par] / all_reduce_partials(a.part, a.sw.ngroups, sh));
  const int32_t r0 = blockIdx.x * a.sw.rows_per_group;
  const int32_t r1 = min(r0 + a.sw.rows_per_group, a.sw.n);
  double rr = 0.0, rz = 0.0;
  for (int32_t i = r0 + threadIdx.x; i < r1; i += 256) {
    a.xa[i] += alpha * a.p[i];
    const float ri = a.r[i] - alpha * a.q[i];
    a.r[i] = ri;
    rr += (double)ri * (double)ri;
    rz += (double)ri * (double)(a.dinv32[i] * ri);
  }
  store_partial(a.part + 1 * kMaxGrid, rr, sh);
  store_partial(a.part + 2 * kMaxGrid, rz, sh);
}

// Decide on a reliable update.  Without one: beta, p update.  With one: fold
// the FP32 increment into the double-double solution (TwoSum).
__global__ __launch_bounds__(256) void k_mix_decide(MixArgs a, int par) {
  __shared__ double sh[8];
  PcgState *st = a.st;
  if (st->done) return;
  const double rr = all_reduce_partials(a.part + 1 * kMaxGrid, a.sw.ngroups, sh);
  const double rz = all_reduce_partials(a.part + 2 * kMaxGrid, a.sw.ngroups, sh);
  const int it = st->iterp[par] + 1;
  const bool refine = rr < st->thr || rr < a.delta2 * st->rr_ref ||
                      it - st->last_refine >= a.max_interval || it >= st->maxit;
  const double beta = rz / st->rho[par];
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    st->iterp[par ^ 1] = it;
    st->iter = it;
    st->refine[par] = refine ? 1 : 0;
    if (!refine) st->rho[par ^ 1] = rz;
  }
  const int32_t r0 = blockIdx.x * a.sw.rows_per_group;
  const int32_t r1 = min(r0 + a.sw.rows_per_group, a.sw.n);
  if (!refine) {
    const float fb = (float)beta;
    for (int32_t i = r0 + threadIdx.x; i < r1; i += 256)
      a.p[i] = a.dinv32[i] * a.r[i] + fb * a.p[i];
  } else {
    for (int32_t i = r0 + threadIdx.x; i < r1; i += 256) {
      const double hi = a.x[i], d = (double)a.xa[i];
      const double s = hi + d;
      const double bp = s - hi;
      const double err = (hi - (s - bp)) + (d - bp);
      a.x[i] = s;
      a.xlo[i] += err;
      a.xa[i] = 0.0f;
    }
  }
}

// reliable update, true residual r64 = b - A x - A xlo (FP64, difference form)
__global__ __launch_bounds__(256) void k_mix_resid(MixArgs a, int par) {
  __shared__ double sh[8];
  if (a.st->done || !a.st->refine[par]) return;
  const int32_t r0 = blockIdx.x * a.sw.rows_per_group;
  const int32_t r1 = min(r0 + a.sw.rows_per_group, a.sw.n);
  double rr = 0.0, rz = 0.0;
  for (int32_t i = r0 + threadIdx.x; i < r1; i += 256) {
    const double ah = ell_row<double, double, double>(a.slice_off, a.s_col, a.s_val, a.rowsum,
                                                      a.x, i, a.x[i]);
    const double al = ell_row<double, double, double>(a.slice_off, a.s_col, a.s_val, a.rowsum,
                                                      a.xlo, i, a.xlo[i]);
    const double ri = (a.b[i] - ah) - al;
    a.r64[i] = ri;
    rr += ri * ri;
    rz += ri * (a.dinv[i] * ri);
  }
  store_partial(a.part + 4 * kMaxGrid, rr, sh);
  store_partial(a.part + 5 * kMaxGrid, rz, sh);
}

// reliable update, finish: convergence on the true residual, or restart the
// FP32 recurrence from it (beta rescaled with the new (r, D^-1 r)).
__global__ __launch_bounds__(256) void k_mix_finish(MixArgs a, int par) {
  __shared__ double sh[8];
  PcgState *st = a.st;
  if (st->done || !st->refine[par]) return;
  const double rt = all_reduce_partials(a.part + 4 * kMaxGrid, a.sw.ngroups, sh);
  const double rz = all_reduce_partials(a.part + 5 * kMaxGrid, a.sw.ngroups, sh);
  const bool conv = rt < st->thr;
  const bool stop = conv || st->iterp[par ^ 1] >= st->maxit;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    st->rr = rt;
    st->refinements += 1;
    st->last_refine = st->iterp[par ^ 1];
    if (stop) {
      st->done = 1;
      st->converged = conv ? 1 : 0;
    } else {
      st->rr_ref = rt;
      st->rho[par ^ 1] = rz;
    }
  }
  if (stop) return;
  const float fb = (float)(rz / st->rho[par]);
  const int32_t r0 = blockIdx.x * a.sw.rows_per_group;
  const int32_t r1 = min(r0 + a.sw.rows_per_group, a.sw.n);
  for (int32_t i = r0 + threadIdx.x; i < r1; i += 256) {
    const float ri = (float)a.r64[i];
    a.r[i] = ri;
    a.p[i] = a.dinv32[i] * ri + fb * a.p[i];
  }
}

// x = x + xlo (the solution handed back in FP64)
__global__ __launch_bounds__(256) void k_mix_final(MixArgs a) {
  const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < a.sw.n) a.x[i] += a.xlo[i];
}

Sweep make_sweep(int32_t n) {
  int64_t per = (n + kMaxGrid - 1) / kMaxGrid;
  per = ((per + 255) / 256) * 256;
  if (per < 256) per = 256;
  Sweep s;
  s.n = n;
  s.rows_per_group = (int32_t)per;
  s.ngroups = (int)((n + per - 1) / per);
  if (s.ngroups < 1) s.ngroups = 1;
  return s;
}

}  // namespace

// ===========================================================================
// C ABI: CSR helpers, SpMV, PCG, BLAS-1
// ===========================================================================
extern "C" int lfgpu_csr_info(const lfgpu_csr *A, int32_t *nrows, int32_t *ncols, int64_t *nnz) {
  LFG_REQUIRE(A, LFGPU_EINVAL, "NULL matrix");
  if (nrows) *nrows = A->nrows;
  if (ncols) *ncols = A->ncols;
  if (nnz) *nnz = A->nnz;
  return LFGPU_OK;
}

extern "C" int lfgpu_csr_download(const lfgpu_csr *A, int32_t *row_ptr, int32_t *col,
                                  double *val) {
  LFG_REQUIRE(A, LFGPU_EINVAL, "NULL matrix");
  hipStream_t s = A->ctx->stream;
  if (row_ptr)
    LFG_HIP(hipMemcpyAsync(row_ptr, A->row_ptr, ((size_t)A->nrows + 1) * 4, hipMemcpyDeviceToHost,
                           s));
  if (col && A->nnz)
    LFG_HIP(hipMemcpyAsync(col, A->col, (size_t)A->nnz * 4, hipMemcpyDeviceToHost, s));
  if (val && A->nnz)
    LFG_HIP(hipMemcpyAsync(val, A->val, (size_t)A->nnz * 8, hipMemcpyDeviceToHost, s));
  LFG_HIP(hipStreamSynchronize(s));
  return LFGPU_OK;
}

extern "C" int lfgpu_csr_destroy(lfgpu_csr *A) {
  if (!A) return LFGPU_OK;
  LFG_REQUIRE(!A->owned_by_plan, LFGPU_EINVAL, "this CSR belongs to a plan");
  if (A->ctx) (void)hipStreamSynchronize(A->ctx->stream);
  solver_cache_free(A);
  dfree(A->row_ptr);
  dfree(A->col);
  dfree(A->val);
  delete A;
  return LFGPU_OK;
}

namespace {
__global__ void k_coo_keys(const int32_t *rows, const int32_t *cols, int64_t nnz, int32_t nrows,
                           int32_t ncols, uint64_t *keys, int32_t *idx, int *bad) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nnz) return;
  const int32_t r = rows[t], c = cols[t];
  if (r < 0 || r >= nrows || c < 0 || c >= ncols) {
    atomicAdd(bad, 1);
    keys[t] = 0;
  } else {
    keys[t] = ((uint64_t)(uint32_t)r << 32) | (uint32_t)c;
  }
  idx[t] = (int32_t)t;
}

__global__ void k_coo_heads(const uint64_t *keys, int64_t nnz, int32_t *head) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nnz) return;
  head[t] = (t == 0 || keys[t] != keys[t - 1]) ? 1 : 0;
}

__global__ void k_coo_segsum(const uint64_t *keys, const int32_t *idx, const double *vals,
                             const int32_t *pos, int64_t nnz, int32_t *col, double *val,
                             int32_t *rowcnt) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nnz) return;
  if (t > 0 && keys[t] == keys[t - 1]) return;
  const uint64_t k = keys[t];
  double acc = vals[idx[t]];
  for (int64_t u = t + 1; u < nnz && keys[u] == k; ++u) acc += vals[idx[u]];
  col[pos[t]] = (int32_t)(k & 0xffffffffu);
  val[pos[t]] = acc;
  atomicAdd(&rowcnt[(int32_t)(k >> 32)], 1);
}
}  // namespace

extern "C" int lfgpu_coo_to_csr(lfgpu_ctx *ctx, int64_t nnz, const int32_t *rows,
                                const int32_t *cols, const double *vals, int32_t nrows,
                                int32_t ncols, lfgpu_csr **out) {
  LFG_REQUIRE(ctx && out && nrows > 0 && ncols > 0 && nnz >= 0, LFGPU_EINVAL,
              "bad lfgpu_coo_to_csr arguments (matrix has zero rows or columns)");
  LFG_REQUIRE(nnz == 0 || (rows && cols && vals), LFGPU_EINVAL, "NULL triplet arrays");
  LFG_REQUIRE(nnz < 2147483647LL, LFGPU_ERANGE, "too many triplets");
  *out = nullptr;
  hipStream_t s = ctx->stream;
  lfgpu_csr *A = new lfgpu_csr();
  A->ctx = ctx;
  A->nrows = nrows;
  A->ncols = ncols;
  int32_t *d_rows = nullptr, *d_cols = nullptr, *idx = nullptr, *idx_s = nullptr,
          *head = nullptr, *pos = nullptr, *rowcnt = nullptr;
  uint64_t *keys = nullptr, *keys_s = nullptr;
  double *d_vals = nullptr;
  int *bad = nullptr;
  void *tmp = nullptr;
  auto cleanup = [&]() {
    dfree(d_rows); dfree(d_cols); dfree(idx); dfree(idx_s); dfree(head); dfree(pos);
    dfree(rowcnt); dfree(keys); dfree(keys_s); dfree(d_vals); dfree(bad); dfree(tmp);
  };
  auto fail = [&](int rc) {
    cleanup();
    lfgpu_csr_destroy(A);
    return rc;
  };
#define C_TRY(expr) do { int rc_ = (expr); if (rc_ != LFGPU_OK) return fail(rc_); } while (0)
#define C_HIP(expr) do { hipError_t e_ = (expr); if (e_ != hipSuccess) return fail(hip_fail(e_, #expr)); } while (0)
  const size_t n = (size_t)std::max<int64_t>(nnz, 1);
  C_TRY(dalloc(&A->row_ptr, (size_t)nrows + 1, "CSR row pointers"));
  C_TRY(dalloc(&rowcnt, (size_t)nrows, "row counts"));
  C_HIP(hipMemsetAsync(rowcnt, 0, (size_t)nrows * 4, s));
  if (nnz > 0) {
    C_TRY(dalloc(&d_rows, n, "rows"));
    C_TRY(dalloc(&d_cols, n, "cols"));
    C_TRY(dalloc(&d_vals, n, "vals"));
    C_TRY(dalloc(&idx, n, "idx"));
    C_TRY(dalloc(&idx_s, n, "idx sorted"));
    C_TRY(dalloc(&keys, n, "keys"));
    C_TRY(dalloc(&keys_s, n, "keys sorted"));
    C_TRY(dalloc(&head, n, "heads"));
    C_TRY(dalloc(&pos, n + 1, "positions"));
    C_TRY(dalloc(&bad, 1, "status"));
    C_HIP(hipMemsetAsync(bad, 0, 4, s));
    C_HIP(hipMemcpyAsync(d_rows, rows, nnz * 4, hipMemcpyHostToDevice, s));
    C_HIP(hipMemcpyAsync(d_cols, cols, nnz * 4, hipMemcpyHostToDevice, s));
    C_HIP(hipMemcpyAsync(d_vals, vals, nnz * 8, hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(k_coo_keys, dim3(grid_for(nnz, 256)), dim3(256), 0, s, d_rows, d_cols, nnz,
                       nrows, ncols, keys, idx, bad);
    C_HIP(hipGetLastError());
    int hbad = 0;
    C_HIP(hipMemcpyAsync(&hbad, bad, 4, hipMemcpyDeviceToHost, s));
    C_HIP(hipStreamSynchronize(s));
    if (hbad) {
      set_error("triplet index out of range");
      return fail(LFGPU_ERANGE);
    }
    // stable LSD radix sort: equal keys keep triplet order
    size_t tb = 0;
    C_HIP(rocprim::radix_sort_pairs(nullptr, tb, keys, keys_s, idx, idx_s, (size_t)nnz, 0, 64, s));
    C_TRY(dmalloc(&tmp, tb, "sort scratch"));
    C_HIP(rocprim::radix_sort_pairs(tmp, tb, keys, keys_s, idx, idx_s, (size_t)nnz, 0, 64, s));
    hipLaunchKernelGGL(k_coo_heads, dim3(grid_for(nnz, 256)), dim3(256), 0, s, keys_s, nnz, head);
    C_HIP(hipGetLastError());
    int64_t nu = 0;
    C_TRY(exclusive_scan_i32(s, head, pos, nnz, &nu));
    A->nnz = nu;
    C_TRY(dalloc(&A->col, (size_t)nu, "CSR columns"));
    C_TRY(dalloc(&A->val, (size_t)nu, "CSR values"));
    hipLaunchKernelGGL(k_coo_segsum, dim3(grid_for(nnz, 256)), dim3(256), 0, s, keys_s, idx_s,
                       d_vals, pos, nnz, A->col, A->val, rowcnt);
    C_HIP(hipGetLastError());
  } else {
    C_TRY(dalloc(&A->col, 1, "CSR columns"));
    C_TRY(dalloc(&A->val, 1, "CSR values"));
  }
  int64_t tot = 0;
  C_TRY(exclusive_scan_i32(s, rowcnt, A->row_ptr, nrows, &tot));
  cleanup();
  *out = A;
  return LFGPU_OK;
#undef C_TRY
#undef C_HIP
}

extern "C" int lfgpu_spmv(lfgpu_ctx *ctx, const lfgpu_csr *Ac, const lfgpu_vec *x, lfgpu_vec *y) {
  LFG_REQUIRE(ctx && Ac && x && y, LFGPU_EINVAL, "NULL argument to lfgpu_spmv");
  LFG_REQUIRE(x->n == Ac->ncols && y->n == Ac->nrows, LFGPU_EINVAL, "SpMV size mismatch");
  LFG_REQUIRE(x->d != y->d, LFGPU_EINVAL, "SpMV input and output must differ");
  lfgpu_csr *A = const_cast<lfgpu_csr *>(Ac);
  LFG_REQUIRE(A->nrows == A->ncols, LFGPU_EINVAL, "SpMV needs a square matrix");
  LFG_TRY(ensure_cache(ctx, A, false));
  lfgpu_solver_cache *c = A->cache;
  LFG_HIP(hipEventRecord(ctx->ev0, ctx->stream));
  hipLaunchKernelGGL(k_spmv64, dim3(grid_for(A->nrows, 256)), dim3(256), 0, ctx->stream,
                     c->slice_off, c->s_col, c->s_val, c->rowsum, x->d, y->d, A->nrows);
  LFG_HIP(hipGetLastError());
  LFG_HIP(hipEventRecord(ctx->ev1, ctx->stream));
  LFG_HIP(hipEventSynchronize(ctx->ev1));
  float ms = 0.f;
  LFG_HIP(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
  ctx->last_ms = ms;
  return LFGPU_OK;
}

extern "C" int lfgpu_spmv_timed(lfgpu_ctx *ctx, const lfgpu_csr *Ac, const lfgpu_vec *x,
                                lfgpu_vec *y, int reps, double *avg_ms) {
  LFG_REQUIRE(ctx && Ac && x && y && avg_ms && reps > 0, LFGPU_EINVAL,
              "bad lfgpu_spmv_timed arguments");
  LFG_REQUIRE(x->n == Ac->ncols && y->n == Ac->nrows && x->d != y->d, LFGPU_EINVAL,
              "SpMV size mismatch");
  lfgpu_csr *A = const_cast<lfgpu_csr *>(Ac);
  LFG_REQUIRE(A->nrows == A->ncols, LFGPU_EINVAL, "SpMV needs a square matrix");
  LFG_TRY(ensure_cache(ctx, A, false));
  lfgpu_solver_cache *c = A->cache;
  LFG_HIP(hipEventRecord(ctx->ev0, ctx->stream));
  for (int k = 0; k < reps; ++k)
    hipLaunchKernelGGL(k_spmv64, dim3(grid_for(A->nrows, 256)), dim3(256), 0, ctx->stream,
                       c->slice_off, c->s_col, c->s_val, c->rowsum, x->d, y->d, A->nrows);
  LFG_HIP(hipGetLastError());
  LFG_HIP(hipEventRecord(ctx->ev1, ctx->stream));
  LFG_HIP(hipEventSynchronize(ctx->ev1));
  float ms = 0.f;
  LFG_HIP(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
  *avg_ms = ms / reps;
  ctx->last_ms = ms;
  return LFGPU_OK;
}

extern "C" int lfgpu_pcg(lfgpu_ctx *ctx, const lfgpu_csr *Ac, const lfgpu_vec *b, lfgpu_vec *x,
                         const lfgpu_pcg_opts *opts, lfgpu_pcg_info *info) {
  LFG_REQUIRE(ctx && Ac && b && x && opts, LFGPU_EINVAL, "NULL argument to lfgpu_pcg");
  lfgpu_csr *A = const_cast<lfgpu_csr *>(Ac);
  LFG_REQUIRE(A->nrows == A->ncols && b->n == A->nrows && x->n == A->nrows, LFGPU_EINVAL,
              "PCG size mismatch");
  LFG_REQUIRE(b->d != x->d, LFGPU_EINVAL, "PCG right-hand side and solution must differ");
  LFG_REQUIRE(opts->rtol >= 0.0 && opts->maxit >= 0, LFGPU_EINVAL, "bad PCG options");
  LFG_REQUIRE(opts->mode == LFGPU_PCG_FP64 || opts->mode == LFGPU_PCG_MIXED, LFGPU_EINVAL,
              "unknown PCG mode");
  const bool mixed = opts->mode == LFGPU_PCG_MIXED;
  LFG_TRY(ensure_cache(ctx, A, mixed));
  lfgpu_solver_cache *c = A->cache;
  hipStream_t s = ctx->stream;
  const int32_t n = A->nrows;
  const Sweep sw = make_sweep(n);
  const dim3 G(sw.ngroups), B(256);
  const unsigned gflat = grid_for(n, 256);
  // an even number of iterations per check keeps the parity of the launch
  // index aligned with the iteration count
  int check = opts->check_every > 0 ? opts->check_every : 64;
  check += check & 1;

  PcgState h{};
  h.thr = opts->rtol * opts->rtol;
  h.maxit = opts->maxit;
  LFG_HIP(hipMemcpyAsync(c->st, &h, sizeof(PcgState), hipMemcpyHostToDevice, s));
  LFG_HIP(hipEventRecord(ctx->ev0, s));

  PcgArgs a{c->slice_off, c->s_col, c->s_val, c->rowsum, c->dinv, x->d, c->r, c->p, c->q,
            c->part, c->st, sw};
  hipLaunchKernelGGL(k_pcg_init, G, B, 0, s, a, b->d);
  hipLaunchKernelGGL(k_pcg_init_finish, dim3(1), B, 0, s, a);
  hipLaunchKernelGGL(k_zero_if_bb0, dim3(gflat), B, 0, s, x->d, n, c->st);
  LFG_HIP(hipGetLastError());

  PcgState *hs = reinterpret_cast<PcgState *>(ctx->pinned);
  MixArgs m{c->slice_off, c->s_col, c->s_val32, c->s_val, c->rowsum32, c->rowsum, c->dinv32,
            c->dinv, x->d, c->xlo, b->d, c->r, c->r32, c->p32, c->q32, c->xa32, c->part, c->st,
            sw, 0.09, 2147483647};
  if (opts->refine_delta > 0.0) m.delta2 = opts->refine_delta * opts->refine_delta;
  if (opts->refine_max_interval > 0) m.max_interval = opts->refine_max_interval;
  if (mixed) {
    hipLaunchKernelGGL(k_mix_setup, dim3(gflat), B, 0, s, m);
    LFG_HIP(hipGetLastError());
  }
  for (;;) {
    for (int k = 0; k < check; ++k) {
      const int par = k & 1;
      if (!mixed) {
        hipLaunchKernelGGL(k_pcg_spmv, G, B, 0, s, a);
        hipLaunchKernelGGL(k_pcg_update, G, B, 0, s, a, par);
        hipLaunchKernelGGL(k_pcg_pupdate, G, B, 0, s, a, par);
      } else {
        hipLaunchKernelGGL(k_mix_spmv, G, B, 0, s, m);
        hipLaunchKernelGGL(k_mix_update, G, B, 0, s, m, par);
        hipLaunchKernelGGL(k_mix_decide, G, B, 0, s, m, par);
        hipLaunchKernelGGL(k_mix_resid, G, B, 0, s, m, par);
        hipLaunchKernelGGL(k_mix_finish, G, B, 0, s, m, par);
      }
    }
    LFG_HIP(hipGetLastError());
    LFG_HIP(hipMemcpyAsync(hs, c->st, sizeof(PcgState), hipMemcpyDeviceToHost, s));
    LFG_HIP(hipStreamSynchronize(s));
    if (hs->done) break;
  }
  if (mixed) {
    hipLaunchKernelGGL(k_mix_final, dim3(gflat), B, 0, s, m);
    LFG_HIP(hipGetLastError());
  }
  LFG_HIP(hipEventRecord(ctx->ev1, s));
  LFG_HIP(hipEventSynchronize(ctx->ev1));
  float ms = 0.f;
  LFG_HIP(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
  ctx->last_ms = ms;
  const double relres = hs->bb > 0.0 ? std::sqrt(hs->rr / hs->bb) : 0.0;
  if (info) {
    info->iters = hs->iter;
    info->converged = hs->converged;
    info->refinements = mixed ? hs->refinements : 0;
    info->relres = relres;
    info->ms = ms;
  }
  if (!hs->converged) {
    set_error("PCG did not reach rtol within maxit iterations (relres " +
              std::to_string(relres) + ")");
    return LFGPU_ENOCONV;
  }
  return LFGPU_OK;
}

extern "C" int lfgpu_axpby(lfgpu_ctx *ctx, double a, const lfgpu_vec *x, double b, lfgpu_vec *y) {
  LFG_REQUIRE(ctx && x && y && x->n == y->n, LFGPU_EINVAL, "bad lfgpu_axpby arguments");
  if (y->n == 0) return LFGPU_OK;
  hipLaunchKernelGGL(k_axpby, dim3(grid_for(y->n, 256)), dim3(256), 0, ctx->stream, a, x->d, b,
                     y->d, y->n);
  LFG_HIP(hipGetLastError());
  return LFGPU_OK;
}

extern "C" int lfgpu_dot(lfgpu_ctx *ctx, const lfgpu_vec *x, const lfgpu_vec *y, double *out) {
  LFG_REQUIRE(ctx && x && y && out && x->n == y->n, LFGPU_EINVAL, "bad lfgpu_dot arguments");
  if (!ctx->red) {
    LFG_TRY(dalloc(&ctx->red, kRedBlocks + 1, "reduction scratch"));
    ctx->red_count = kRedBlocks + 1;
  }
  hipLaunchKernelGGL(k_dot_partial, dim3(kRedBlocks), dim3(256), 0, ctx->stream, x->d, y->d, x->n,
                     ctx->red);
  hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(256), 0, ctx->stream, ctx->red, kRedBlocks,
                     ctx->red + kRedBlocks);
  LFG_HIP(hipGetLastError());
  LFG_HIP(hipMemcpyAsync(out, ctx->red + kRedBlocks, 8, hipMemcpyDeviceToHost, ctx->stream));
  LFG_HIP(hipStreamSynchronize(ctx->stream));
  return LFGPU_OK;
}

// ===========================================================================
// adjoint gradient
// ===========================================================================
namespace {
__global__ __launch_bounds__(256) void k_adjoint_p1(const double *xy, const int32_t *cell_v,
                                                    int64_t ncells, const double *tab,
                                                    const double *theta, const double *u,
                                                    const double *lam, double *grad) {
  __shared__ double s_tab[Tab<3, 3>::kLen];
  for (int t = threadIdx.x; t < Tab<3, 3>::kLen; t += blockDim.x) s_tab[t] = tab[t];
  __syncthreads();
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncells) return;
  const Tab<3, 3> T{s_tab};
  const CellXY p = load_cell(xy, cell_v, c);
  Geo g;
  tria_geo(p.x0, p.y0, p.x1, p.y1, p.x2, p.y2, g);
  const double one[3] = {1.0, 1.0, 1.0}, zero[3] = {0.0, 0.0, 0.0};
  const int32_t v[3] = {cell_v[3 * c], cell_v[3 * c + 1], cell_v[3 * c + 2]};
  const double ue[3] = {u[v[0]], u[v[1]], u[v[2]]};
  double s = 0.0;
  for (int i = 0; i < 3; ++i) {
    double row[3];
    element_row<3, 3>(T, g, one, zero, i, row);
    double ku = 0.0;
    for (int j = 0; j < 3; ++j) ku += row[j] * ue[j];
    s += lam[v[i]] * ku;
  }
  grad[c] = -exp(theta[c]) * s;
}
}  // namespace

extern "C" int lfgpu_adjoint_gradient(lfgpu_ctx *ctx, const lfgpu_mesh *m, const lfgpu_fe *fe,
                                      const lfgpu_vec *theta, const lfgpu_vec *u,
                                      const lfgpu_vec *lam, lfgpu_vec *grad) {
  LFG_REQUIRE(ctx && m && fe && theta && u && lam && grad, LFGPU_EINVAL,
              "NULL argument to lfgpu_adjoint_gradient");
  LFG_REQUIRE(m->nloc == 3 && fe->nloc == 3, LFGPU_EINVAL, "adjoint gradient is P1 only");
  LFG_REQUIRE(theta->n == m->ncells && grad->n == m->ncells && u->n == m->ndofs &&
                  lam->n == m->ndofs,
              LFGPU_EINVAL, "adjoint gradient size mismatch");
  LFG_HIP(hipEventRecord(ctx->ev0, ctx->stream));
  hipLaunchKernelGGL(k_adjoint_p1, dim3(grid_for(m->ncells, 256)), dim3(256), 0, ctx->stream,
                     m->xy, m->cell_v, m->ncells, fe->tab, theta->d, u->d, lam->d, grad->d);
  LFG_HIP(hipGetLastError());
  LFG_HIP(hipEventRecord(ctx->ev1, ctx->stream));
  LFG_HIP(hipEventSynchronize(ctx->ev1));
  float ms = 0.f;
  LFG_HIP(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
  ctx->last_ms = ms;
  return LFGPU_OK;
}
