// SpMV and Jacobi-PCG on the device.
//
// Solver mirror of a CSR matrix (built once per pattern, values refreshed
// whenever the CSR values change):
//   * off-diagonal entries only, applied in difference form
//       (A x)_i = sum_{j != i} a_ij (x_j - x_i) + s_i x_i,  s_i = sum_j a_ij,
//     which is algebraically the CSR product.  Stiffness rows nearly cancel
//     (s_i = O(h^2) against entries O(1)); the plain form then loses ~1/h^2
//     ulps on smooth vectors, which at h = 1/2048 puts the attainable FP64
//     residual near 1e-10.  s_i is summed with Neumaier compensation.
//   * storage format chosen per pattern:
//       DIA - when every row's column offsets j - i come from one set of at
//             most 8 offsets (structured meshes): one array per offset, no
//             column indices, every load a coalesced stream;
//       ELL - otherwise: sliced ELL with 64-row slices (one wavefront),
//             entries column-major inside a slice, padding value 0 and
//             column = row.
//   * Jacobi inverse diagonal (Eigen DiagonalPreconditioner semantics).
#include <hip/hip_bf16.h>
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <type_traits>
#include <vector>

#include "lfgpu_internal.h"
#include "lfgpu_reduce.h"

using namespace lfgpu;

namespace {

constexpr int kSlice = 64;
constexpr int kMaxDiag = 8;
constexpr int kMaxGrid = 2048;  // workgroups of the solver sweeps
constexpr int kPartSlots = 8;   // partial-sum arrays of kMaxGrid doubles
constexpr int FMT_ELL = 0, FMT_DIA = 1;
// batches only: DIA storage whose offsets are chosen per problem (a block
// diagonal matrix of structured meshes of different sizes)
constexpr int FMT_BDIA = 2;
// the CSR itself, no mirror (the allocation-free solve): the difference form
// with the row sum accumulated during the product, in the mirror's order
constexpr int FMT_CSR = 3;

}  // namespace

// Solver scalars.  Every kernel recomputes the scalars it needs from the
// previous kernels' per-workgroup partial sums (same order in every
// workgroup, so all workgroups agree bit for bit); only workgroup 0 writes
// this state, and anything read and written within one iteration is double
// buffered by iteration parity.
struct PcgState {
  double rho[2];   // (r, D^-1 r) by parity (mixed mode)
  double thr;      // tol^2 ||b||^2 (on entry: tol^2)
  double bb;       // ||b||^2
  double rr;       // last ||r||^2 (recurrence in FP64 mode, true in mixed)
  double rr_ref;   // mixed: ||r||^2 at the last reliable update
  int iter, done, converged, maxit;
  int refine[2];   // mixed: reliable update decided in this iteration (parity)
  int iterp[2];    // mixed: iteration count by parity
  int refinements;
  int last_refine; // mixed: iteration of the last reliable update
  double alpha[2]; // single-reduction FP64: alpha by parity
  double alpha_s, beta_s;  // row sweeps (sliced ELL): this iteration's alpha, beta
  int stop_pending, conv_pending;  // row-sweep mixed: stop after the fold
};

struct lfgpu_solver_cache {
  int fmt = FMT_ELL;
  int ndiag = 0;
  int off[kMaxDiag] = {0};
  int32_t nslices = 0;
  int64_t mlen = 0;              // stored off-diagonal slots
  int32_t *slice_off = nullptr;  // ELL: [nslices + 1] element offsets
  int32_t *s_col = nullptr;      // ELL columns (int32), or
  int16_t *s_dcol = nullptr;     // ELL column minus row, when every offset fits
  double *val = nullptr;         // ELL / DIA values
  float *val32 = nullptr;
  double *dinv = nullptr;
  float *dinv32 = nullptr;           // FP32-matrix mode: D^-1 in FP32
  __hip_bfloat16 *dinv16 = nullptr;  // FP32-Krylov mode: D^-1 rounded to bfloat16
  double *rowsum = nullptr;
  float *rowsum32 = nullptr;
  uint64_t built_version = ~0ull;
  bool have32 = false;
  // DIA with a mirror-symmetric offset list whose lower-diagonal values
  // equal the mirrored upper-diagonal values bit for bit: the FP64 SpMV
  // sweep of the PCG then reads the upper diagonals only (measured 7 %
  // faster at 2048^2; the FP32 sweeps measured slower that way)
  bool sym64 = false;
  // work vectors
  double *r = nullptr, *p = nullptr, *q = nullptr;
  float *r32 = nullptr, *p32 = nullptr, *q32 = nullptr, *xa32 = nullptr;
  double *xlo = nullptr;   // mixed: low part of the double-double solution
  double *r64 = nullptr;   // mixed with FP64 vectors: true residual
  double *xa64 = nullptr;  // mixed with FP64 vectors: increment accumulator
  double *part = nullptr;  // [kPartSlots x kMaxGrid] partial sums
  double *cg1 = nullptr;   // single-reduction PCG: r, w, s by parity [6 n]
  double *rpart = nullptr; // row-sweep PCG partials [3 x ceil(n / 256)]
  PcgState *st = nullptr;
  // captured chunks of FP64 iterations, two slots (opaque: the slot type
  // holds the kernels' argument blocks, defined with them below)
  void *cg_graphs = nullptr;
  int cg_graph_next = 0;
  // sliced ELL of a single (not batched) matrix: the storage holds the rows
  // in a locality order, perm[new] = old row, iperm[old] = new row, and the
  // solver works on permuted copies of x and b
  int32_t *perm = nullptr, *iperm = nullptr;
  double *xp = nullptr, *bp = nullptr;
  // the CSR's dense block in storage order: blk_srow[c] is the storage row
  // of block row c, blk_smap[r] the block row of storage row r (-1: none),
  // blk_t the block product of the current sweep; blk_n 0: no block
  int32_t blk_n = 0, blk_cap = 0;
  int32_t *blk_srow = nullptr, *blk_smap = nullptr;
  double *blk_t = nullptr;
};

namespace lfgpu {
void cg_graphs_destroy(void *slots);

void solver_cache_free(lfgpu_csr *A) {
  lfgpu_solver_cache *c = A->cache;
  if (!c) return;
  if (c->cg_graphs) cg_graphs_destroy(c->cg_graphs);
  for (void *p : {(void *)c->slice_off, (void *)c->s_col, (void *)c->s_dcol, (void *)c->val, (void *)c->val32,
                  (void *)c->dinv, (void *)c->dinv32, (void *)c->dinv16, (void *)c->rowsum, (void *)c->rowsum32,
                  (void *)c->r, (void *)c->p, (void *)c->q, (void *)c->r32, (void *)c->p32,
                  (void *)c->q32, (void *)c->xa32, (void *)c->xlo, (void *)c->r64, (void *)c->xa64,
                  (void *)c->part, (void *)c->cg1, (void *)c->rpart, (void *)c->st,
                  (void *)c->perm, (void *)c->iperm, (void *)c->xp, (void *)c->bp,
                  (void *)c->blk_srow, (void *)c->blk_smap, (void *)c->blk_t})
    dfree(p);
  delete c;
  A->cache = nullptr;
}
}  // namespace lfgpu

namespace {

// ---------------------------------------------------------------------------
// matrix view handed to kernels
// ---------------------------------------------------------------------------
struct MatView {
  int32_t n;
  int32_t lo;  // first row of the block (a batch problem's rows; else 0)
  int ndiag;
  int off[kMaxDiag];
  const int32_t *__restrict__ slice_off;
  const int32_t *__restrict__ s_col;
  const int16_t *__restrict__ s_dcol;  // non-null: 16-bit column offsets
  const double *__restrict__ val;
  const float *__restrict__ val32;
  const double *__restrict__ rowsum;
  const float *__restrict__ rowsum32;
  const int32_t *__restrict__ row_ptr;  // FMT_CSR
  const int32_t *__restrict__ col;      // FMT_CSR
};

MatView view_of(const lfgpu_solver_cache *c, int32_t n) {
  MatView m;
  m.row_ptr = nullptr;
  m.col = nullptr;
  m.n = n;
  m.lo = 0;
  m.ndiag = c->ndiag;
  for (int d = 0; d < kMaxDiag; ++d) m.off[d] = c->off[d];
  m.slice_off = c->slice_off;
  m.s_col = c->s_col;
  m.s_dcol = c->s_dcol;
  m.val = c->val;
  m.val32 = c->val32;
  m.rowsum = c->rowsum;
  m.rowsum32 = c->rowsum32;
  return m;
}

// One sliced-ELL row: sum over its w slots k of v_k (x_{c_k} - xr), in
// slot order (vp, cp point at the row's first slot; a row's slots are 64
// elements apart).  CT = int32_t: stored columns; CT = int16_t: column minus
// row.  Eight, then four, slots' value and column loads and gathers are
// issued before their terms are added.  NT: the values and columns are
// loaded non-temporally (read once per product, they would otherwise push
// the gathered vector out of the caches); used by the FP64 row sweep of a
// locality-ordered matrix, where it measured faster (2.4M-DOF P3 product
// 104.1 -> 98.4 us), and nowhere else: the FP32 sweeps and the batched
// sweeps (rows in their original order) measured slower with it.
template <bool NT, class T>
__device__ __forceinline__ T ell_load(const T *p) {
  if constexpr (NT)
    return __builtin_nontemporal_load(p);
  else
    return *p;
}

template <bool NT, class CT>
__device__ __forceinline__ int32_t ell_col(const CT *__restrict__ cp, int32_t k, int32_t r) {
  if constexpr (sizeof(CT) == 2)
    return r + (int32_t)ell_load<NT>(&cp[64 * k]);
  else
    return ell_load<NT>(&cp[64 * k]);
}

template <bool NT, class TV, class CT, class XF>
__device__ __forceinline__ double ell_row_sum(const TV *__restrict__ vp, const CT *__restrict__ cp,
                                              int32_t w, int32_t r, double xr, const XF &xf) {
  using XT = decltype(xf(0));
  double acc = 0.0;
  int32_t k = 0;
  for (; k + 8 <= w; k += 8) {
    int32_t c[8];
    TV vv[8];
    XT xx[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      c[u] = ell_col<NT>(cp, k + u, r);
      vv[u] = ell_load<NT>(&vp[64 * (k + u)]);
    }
#pragma unroll
    for (int u = 0; u < 8; ++u) xx[u] = xf(c[u]);
#pragma unroll
    for (int u = 0; u < 8; ++u) acc += (double)vv[u] * ((double)xx[u] - xr);
  }
  for (; k + 4 <= w; k += 4) {
    int32_t c[4];
    TV vv[4];
    XT xx[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      c[u] = ell_col<NT>(cp, k + u, r);
      vv[u] = ell_load<NT>(&vp[64 * (k + u)]);
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) xx[u] = xf(c[u]);
#pragma unroll
    for (int u = 0; u < 4; ++u) acc += (double)vv[u] * ((double)xx[u] - xr);
  }
  for (; k < w; ++k)
    acc += (double)ell_load<NT>(&vp[64 * k]) * ((double)xf(ell_col<NT>(cp, k, r)) - xr);
  return acc;
}

template <bool NT = false, class TV, class XF>
__device__ __forceinline__ double ell_row(const MatView &m, const TV *__restrict__ val, int32_t r,
                                          double xr, const XF &xf) {
  const int32_t s = r >> 6, l = r & 63;
  const int32_t o = m.slice_off[s];
  const int32_t w = (m.slice_off[s + 1] - o) >> 6;
  return m.s_dcol ? ell_row_sum<NT>(val + o + l, m.s_dcol + o + l, w, r, xr, xf)
                  : ell_row_sum<NT>(val + o + l, m.s_col + o + l, w, r, xr, xf);
}

// (A x)_r in FP64 (difference form).  HOIST (DIA): issue every load of the
// row before the first use (one memory round trip instead of one per
// diagonal; the sum keeps the diagonal order).  It costs registers, so it is
// used where it measured faster: the FP64 SpMV and the FP64 PCG SpMV sweep.
// SYM (DIA, hoisted): the lower diagonal d (offset < 0) is read as the
// mirrored upper diagonal ndiag-1-d of row c = r + off[d] (zero when c lies
// before the first row: row 0, or m.lo with BLOCK, a batch problem's first
// row), which holds the same value when the matrix is symmetric bit for bit;
// half the diagonals' bytes, the same sum.  (A runtime bound of 0 instead
// of the constant measured 13% slower in the single-matrix sweep.)
// NT (ELL): non-temporal value and column loads (ell_row_sum).
template <int FMT, bool HOIST = false, bool SYM = false, bool BLOCK = false, bool NT = false>
__device__ __forceinline__ double apply64(const MatView &m, const double *__restrict__ x,
                                          int32_t r, double xr) {
  double acc = 0.0;
  if constexpr (FMT == FMT_CSR) {
    // off-diagonal terms in column order and the Neumaier-compensated row
    // sum of all entries in CSR order: the same sums as the mirror's
    // product (k_fill_vals and the DIA / ELL branches below)
    double sum = 0.0, comp = 0.0;
    for (int32_t k = m.row_ptr[r]; k < m.row_ptr[r + 1]; ++k) {
      const double v = m.val[k];
      const double t = sum + v;
      comp += fabs(sum) >= fabs(v) ? (sum - t) + v : (v - t) + sum;
      sum = t;
      const int32_t c = m.col[k];
      if (c != r) acc += v * (x[c] - xr);
    }
    return acc + (sum + comp) * xr;
  } else if constexpr (FMT == FMT_DIA && HOIST) {
    double v[kMaxDiag], xc[kMaxDiag];
#pragma unroll
    for (int d = 0; d < kMaxDiag; ++d) {
      if (d < m.ndiag) {
        int32_t c = r + m.off[d];
        if (SYM && 2 * d < m.ndiag) {
          const int32_t lo = BLOCK ? m.lo : 0;
          const bool in = c >= lo;
          v[d] = in ? m.val[(int64_t)(m.ndiag - 1 - d) * m.n + c] : 0.0;
          xc[d] = x[in ? c : lo];
        } else {
          c = c < 0 ? 0 : (c >= m.n ? m.n - 1 : c);
          v[d] = m.val[(int64_t)d * m.n + r];
          xc[d] = x[c];
        }
      }
    }
#pragma unroll
    for (int d = 0; d < kMaxDiag; ++d)
      if (d < m.ndiag) acc += v[d] * (xc[d] - xr);
  } else if constexpr (FMT == FMT_DIA) {
#pragma unroll
    for (int d = 0; d < kMaxDiag; ++d) {
      if (d < m.ndiag) {
        int32_t c = r + m.off[d];
        c = c < 0 ? 0 : (c >= m.n ? m.n - 1 : c);
        acc += m.val[(int64_t)d * m.n + r] * (x[c] - xr);
      }
    }
  } else {
    acc = ell_row<NT>(m, m.val, r, xr, [&](int32_t c) { return x[c]; });
  }
  return acc + m.rowsum[r] * xr;
}

// (A x)_r from FP32 storage with FP64 arithmetic: a product of two FP32
// numbers is exact in FP64 and the row is summed in FP64, so the only FP32
// rounding of the result is the final conversion.
template <int FMT, bool HOIST = false>
__device__ __forceinline__ float apply32(const MatView &m, const float *__restrict__ x, int32_t r,
                                         float xr) {
  const double xd = (double)xr;
  double acc = 0.0;
  if constexpr (FMT == FMT_DIA && HOIST) {
    float v[kMaxDiag], xc[kMaxDiag];
#pragma unroll
    for (int d = 0; d < kMaxDiag; ++d) {
      if (d < m.ndiag) {
        int32_t c = r + m.off[d];
        c = c < 0 ? 0 : (c >= m.n ? m.n - 1 : c);
        v[d] = m.val32[(int64_t)d * m.n + r];
        xc[d] = x[c];
      }
    }
#pragma unroll
    for (int d = 0; d < kMaxDiag; ++d)
      if (d < m.ndiag) acc += (double)v[d] * ((double)xc[d] - xd);
  } else if constexpr (FMT == FMT_DIA) {
#pragma unroll
    for (int d = 0; d < kMaxDiag; ++d) {
      if (d < m.ndiag) {
        int32_t c = r + m.off[d];
        c = c < 0 ? 0 : (c >= m.n ? m.n - 1 : c);
        acc += (double)m.val32[(int64_t)d * m.n + r] * ((double)x[c] - xd);
      }
    }
  } else {
    acc = ell_row(m, m.val32, r, xd, [&](int32_t c) { return x[c]; });
  }
  return (float)(acc + (double)m.rowsum32[r] * xd);
}

// ---------------------------------------------------------------------------
// building the mirror
// ---------------------------------------------------------------------------
// Sliced-ELL rows are stored in the order perm (row r of the storage is row
// perm[r] of the CSR; NULL = identity) with columns renamed by iperm; the
// entries of a row keep their CSR order.
__global__ void k_slice_width(const int32_t *row_ptr, const int32_t *col, int32_t n,
                              int32_t nslices, const int32_t *perm, int32_t *w) {
  const int32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= nslices) return;
  int32_t m = 0;
  const int32_t r0 = s * kSlice, r1 = min(r0 + kSlice, n);
  for (int32_t r = r0; r < r1; ++r) {
    const int32_t o = perm ? perm[r] : r;
    int32_t k = 0;
    for (int32_t p = row_ptr[o]; p < row_ptr[o + 1]; ++p) k += (col[p] != o);
    m = max(m, k);
  }
  w[s] = m * kSlice;
}

// flag |= 1 if a stored column lies more than 32767 rows from its stored row
__global__ void k_ell_span(const int32_t *row_ptr, const int32_t *col, int32_t n,
                           const int32_t *perm, const int32_t *iperm, int *flag) {
  const int32_t r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  const int32_t o = perm ? perm[r] : r;
  for (int32_t p = row_ptr[o]; p < row_ptr[o + 1]; ++p) {
    if (col[p] == o) continue;
    const int32_t d = (iperm ? iperm[col[p]] : col[p]) - r;
    if (d > 32767 || d < -32768) {
      atomicOr(flag, 1);
      return;
    }
  }
}

// CT = int32_t: the column; CT = int16_t: column minus row.  Padding slots
// point at the row itself (value 0, and x_r - x_r = 0 in the difference form).
template <class CT>
__global__ void k_fill_ell_cols(const int32_t *row_ptr, const int32_t *col, int32_t n,
                                const int32_t *slice_off, const int32_t *perm,
                                const int32_t *iperm, CT *s_col) {
  const int32_t r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  const int32_t o = perm ? perm[r] : r;
  const int32_t s = r / kSlice, l = r % kSlice;
  const int32_t off = slice_off[s], w = (slice_off[s + 1] - off) / kSlice;
  const int32_t base = sizeof(CT) == 2 ? r : 0;
  int32_t k = 0;
  for (int32_t p = row_ptr[o]; p < row_ptr[o + 1]; ++p)
    if (col[p] != o) s_col[off + (k++) * kSlice + l] = (CT)((iperm ? iperm[col[p]] : col[p]) - base);
  for (; k < w; ++k) s_col[off + k * kSlice + l] = (CT)(r - base);
}

// first locality key of a row: its smallest column (the CSR columns are
// sorted)
__global__ void k_row_key(const int32_t *row_ptr, const int32_t *col, int32_t n, int32_t *key) {
  const int32_t r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  key[r] = row_ptr[r + 1] > row_ptr[r] ? col[row_ptr[r]] : r;
}

// locality key of a row: the smallest first key among its columns (its own
// included).  A row whose own smallest column is far from the mesh position
// of its neighbours (a P3 cell DOF of a corner cell, all of whose vertices
// carry Dirichlet values, couples only to edge DOFs) takes its neighbours'
// key and lands next to them.
__global__ void k_row_key2(const int32_t *row_ptr, const int32_t *col, int32_t n,
                           const int32_t *key1, int32_t *key) {
  const int32_t r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  int32_t k = key1[r];
  for (int32_t p = row_ptr[r]; p < row_ptr[r + 1]; ++p) k = min(k, key1[col[p]]);
  key[r] = k;
}

__global__ void k_gather(const double *__restrict__ src, const int32_t *__restrict__ perm,
                         int32_t n, double *__restrict__ dst) {
  const int32_t r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r < n) dst[r] = src[perm[r]];
}

__global__ void k_scatter(const double *__restrict__ src, const int32_t *__restrict__ perm,
                          int32_t n, double *__restrict__ dst) {
  const int32_t r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r < n) dst[perm[r]] = src[r];
}

struct DiagSet {
  int ndiag;
  int off[kMaxDiag];
};

// flag |= 1 if some off-diagonal entry's offset is not in the set
__global__ void k_dia_check(const int32_t *row_ptr, const int32_t *col, int32_t n, DiagSet ds,
                            int *flag) {
  const int32_t r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  for (int32_t p = row_ptr[r]; p < row_ptr[r + 1]; ++p) {
    const int32_t c = col[p];
    if (c == r) continue;
    bool found = false;
    for (int d = 0; d < ds.ndiag; ++d) found |= (c - r == ds.off[d]);
    if (!found) {
      atomicOr(flag, 1);
      return;
    }
  }
}

// Values, compensated row sums and the inverse diagonal.  DIA: val[d n + r];
// ELL: the slot layout of k_fill_ell_cols.
__global__ void k_fill_vals(const int32_t *row_ptr, const int32_t *col, const double *val,
                            int32_t n, int fmt, DiagSet ds, const int32_t *slice_off,
                            const int32_t *perm, double *mval, float *mval32, double *rowsum,
                            float *rowsum32, double *dinv, float *dinv32,
                            __hip_bfloat16 *dinv16) {
  const int32_t r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  const int32_t o = perm ? perm[r] : r;  // the CSR row stored as row r
  double d = 1.0, sum = 0.0, comp = 0.0;
  if (fmt == FMT_DIA) {
    for (int q = 0; q < ds.ndiag; ++q) {
      mval[(int64_t)q * n + r] = 0.0;
      if (mval32) mval32[(int64_t)q * n + r] = 0.0f;
    }
  }
  int32_t off = 0, w = 0;
  const int32_t s = r / kSlice, l = r % kSlice;
  if (fmt == FMT_ELL) {
    off = slice_off[s];
    w = (slice_off[s + 1] - off) / kSlice;
  }
  int32_t k = 0;
  for (int32_t p = row_ptr[o]; p < row_ptr[o + 1]; ++p) {
    const double v = val[p];
    const double t = sum + v;
    comp += fabs(sum) >= fabs(v) ? (sum - t) + v : (v - t) + sum;
    sum = t;
    const int32_t c = col[p];
    if (c == o) {
      if (v != 0.0) d = 1.0 / v;
      continue;
    }
    int64_t sl;
    if (fmt == FMT_DIA) {
      int q = 0;
      while (q < ds.ndiag - 1 && ds.off[q] != c - r) ++q;
      sl = (int64_t)q * n + r;
    } else {
      sl = off + (int64_t)(k++) * kSlice + l;
    }
    mval[sl] = v;
    if (mval32) mval32[sl] = (float)v;
  }
  if (fmt == FMT_ELL) {
    for (; k < w; ++k) {
      mval[off + (int64_t)k * kSlice + l] = 0.0;
      if (mval32) mval32[off + (int64_t)k * kSlice + l] = 0.0f;
    }
  }
  const double rs = sum + comp;
  rowsum[r] = rs;
  dinv[r] = d;
  if (rowsum32) rowsum32[r] = (float)rs;
  if (dinv32) dinv32[r] = (float)d;
  if (dinv16) dinv16[r] = __float2bfloat16((float)d);
}

// flag = 1 if a lower-diagonal value differs, bit for bit, from the
// mirrored upper-diagonal value (the offsets are ascending and
// mirror-symmetric: off[d] = -off[nd-1-d])
__global__ void k_dia_sym(const double *val, int32_t n, int nd, DiagSet ds, int *flag) {
  const int32_t r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  bool bad = false;
  for (int d = 0; 2 * d < nd; ++d) {
    const int32_t c = r + ds.off[d];
    const double a = val[(int64_t)d * n + r];
    const double b = c >= 0 ? val[(int64_t)(nd - 1 - d) * n + c] : 0.0;
    bad = bad || __double_as_longlong(a) != __double_as_longlong(b);
  }
  if (bad) atomicOr(flag, 1);
}

// Sets the cache's sym64 flag for its current values.
int dia_symmetry(lfgpu_ctx *ctx, lfgpu_solver_cache *c, int32_t n) {
  c->sym64 = false;
  if (c->fmt != FMT_DIA || c->ndiag % 2 != 0) return LFGPU_OK;
  for (int d = 0; d < c->ndiag; ++d)
    if (c->off[d] != -c->off[c->ndiag - 1 - d] || (2 * d < c->ndiag && c->off[d] >= 0))
      return LFGPU_OK;
  DiagSet ds{};
  ds.ndiag = c->ndiag;
  for (int d = 0; d < kMaxDiag; ++d) ds.off[d] = c->off[d];
  hipStream_t s = ctx->stream;
  int *flag = nullptr;
  LFG_TRY(dalloc(&flag, 1, "symmetry flag"));
  hipError_t e = hipMemsetAsync(flag, 0, 4, s);
  if (e == hipSuccess) {
    hipLaunchKernelGGL(k_dia_sym, dim3(grid_for(n, 256)), dim3(256), 0, s, c->val, n, c->ndiag,
                       ds, flag);
    e = hipGetLastError();
  }
  int h = 1;
  if (e == hipSuccess) e = hipMemcpyAsync(&h, flag, 4, hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  dfree(flag);
  if (e != hipSuccess) return hip_fail(e, "DIA symmetry check");
  c->sym64 = h == 0;
  return LFGPU_OK;
}

// DIA when the off-diagonal offsets of a middle row cover every row and
// storing whole diagonals wastes at most 25 %.
int detect_dia(lfgpu_ctx *ctx, const lfgpu_csr *A, DiagSet *ds) {
  hipStream_t s = ctx->stream;
  const int32_t n = A->nrows;
  ds->ndiag = 0;
  if (n < 2 * kSlice) return LFGPU_OK;
  const int32_t mid = n / 2;
  int32_t rp[2];
  LFG_HIP(hipMemcpyAsync(rp, A->row_ptr + mid, 8, hipMemcpyDeviceToHost, s));
  LFG_HIP(hipStreamSynchronize(s));
  const int32_t len = rp[1] - rp[0];
  if (len <= 1 || len > 64) return LFGPU_OK;
  std::vector<int32_t> cols(len);
  LFG_HIP(hipMemcpyAsync(cols.data(), A->col + rp[0], len * 4, hipMemcpyDeviceToHost, s));
  LFG_HIP(hipStreamSynchronize(s));
  DiagSet cand{};
  for (int32_t c : cols)
    if (c != mid) {
      if (cand.ndiag == kMaxDiag) return LFGPU_OK;
      cand.off[cand.ndiag++] = c - mid;
    }
  if (cand.ndiag == 0) return LFGPU_OK;
  const int64_t offdiag_nnz = A->nnz - n;  // FE matrices carry every diagonal
  if ((double)cand.ndiag * n > 1.25 * (double)std::max<int64_t>(offdiag_nnz, 1)) return LFGPU_OK;
  int *flag = nullptr;
  LFG_TRY(dalloc(&flag, 1, "DIA check flag"));
  hipError_t e = hipMemsetAsync(flag, 0, 4, s);
  if (e == hipSuccess) {
    hipLaunchKernelGGL(k_dia_check, dim3(grid_for(n, 256)), dim3(256), 0, s, A->row_ptr, A->col,
                       n, cand, flag);
    e = hipGetLastError();
  }
  int hflag = 1;
  if (e == hipSuccess) e = hipMemcpyAsync(&hflag, flag, 4, hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  dfree(flag);
  if (e != hipSuccess) return hip_fail(e, "DIA detection");
  if (hflag == 0) *ds = cand;
  return LFGPU_OK;
}

// Locality order of the rows of a sliced-ELL matrix.  Rows are ordered by
// their locality key (k_row_key2: the smallest column among the row and its
// neighbours' rows) in windows of 2^kEllWindowBits columns and, inside a
// window, by their length (then by index).  On a finite element matrix whose
// lowest DOFs (the vertices) are numbered along the mesh, the rows of every
// DOF kind (vertex, edge, cell DOFs) near one part of the mesh end up next to
// each other, so a run of consecutive rows gathers from a compact window of
// the (equally permuted) vector instead of from three distant ranges; the
// length order keeps the 64-row slices of rows of one kind, so they carry
// no more padding than in the original order.
constexpr int kEllWindowBits = 9;

int ell_locality_order(lfgpu_ctx *ctx, const lfgpu_csr *A, lfgpu_solver_cache *c) {
  hipStream_t s = ctx->stream;
  const int32_t n = A->nrows;
  int32_t *dkey = nullptr;
  LFG_TRY(dalloc(&dkey, 2 * (size_t)n, "row keys"));
  hipLaunchKernelGGL(k_row_key, dim3(grid_for(n, 256)), dim3(256), 0, s, A->row_ptr, A->col, n,
                     dkey + n);
  hipLaunchKernelGGL(k_row_key2, dim3(grid_for(n, 256)), dim3(256), 0, s, A->row_ptr, A->col, n,
                     dkey + n, dkey);
  std::vector<int32_t> key(n), rp((size_t)n + 1), perm(n), iperm(n);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess)
    e = hipMemcpyAsync(key.data(), dkey, (size_t)n * 4, hipMemcpyDeviceToHost, s);
  if (e == hipSuccess)
    e = hipMemcpyAsync(rp.data(), A->row_ptr, ((size_t)n + 1) * 4, hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  dfree(dkey);
  if (e != hipSuccess) return hip_fail(e, "row keys");
  std::vector<uint64_t> sk(n);
  for (int32_t r = 0; r < n; ++r)
    sk[r] = ((uint64_t)(uint32_t)(key[r] >> kEllWindowBits) << 32) |
            (uint64_t)(uint32_t)(rp[(size_t)r + 1] - rp[r]);
  for (int32_t r = 0; r < n; ++r) perm[r] = r;
  std::stable_sort(perm.begin(), perm.end(), [&](int32_t a, int32_t b) { return sk[a] < sk[b]; });
  for (int32_t r = 0; r < n; ++r) iperm[perm[r]] = r;
  LFG_TRY(dalloc(&c->perm, (size_t)n, "row order"));
  LFG_TRY(dalloc(&c->iperm, (size_t)n, "inverse row order"));
  LFG_TRY(dalloc(&c->xp, (size_t)n, "permuted solution"));
  LFG_TRY(dalloc(&c->bp, (size_t)n, "permuted right-hand side"));
  LFG_HIP(hipMemcpyAsync(c->perm, perm.data(), (size_t)n * 4, hipMemcpyHostToDevice, s));
  LFG_HIP(hipMemcpyAsync(c->iperm, iperm.data(), (size_t)n * 4, hipMemcpyHostToDevice, s));
  LFG_HIP(hipStreamSynchronize(s));
  return LFGPU_OK;
}

// Queues the refresh of the storage's values, row sums and inverse diagonal
// (and their FP32 / bfloat16 copies) from A's current values.
void fill_vals_enqueue(hipStream_t s, const lfgpu_csr *A, lfgpu_solver_cache *c) {
  DiagSet ds{};
  ds.ndiag = c->ndiag;
  for (int d = 0; d < kMaxDiag; ++d) ds.off[d] = c->off[d];
  hipLaunchKernelGGL(k_fill_vals, dim3(grid_for(A->nrows, 256)), dim3(256), 0, s, A->row_ptr,
                     A->col, A->val, A->nrows, c->fmt, ds, c->slice_off, c->perm, c->val,
                     c->have32 ? c->val32 : nullptr, c->rowsum, c->have32 ? c->rowsum32 : nullptr,
                     c->dinv, c->have32 ? c->dinv32 : nullptr, c->have32 ? c->dinv16 : nullptr);
}

// The dense block of a CSR (lfgpu_csr_set_dense_block) as the FP64 sweeps
// see it: the matrix is its sparse part + scale * B on the storage rows
// srow; t = B v of the sweep's vector v, added to row r's product when
// smap[r] >= 0.
struct BlkView {
  const double *B = nullptr;
  const int32_t *srow = nullptr, *smap = nullptr;
  double *t = nullptr;
  double scale = 0.0;
  int32_t n = 0;
};

__global__ void k_fill_i32(int32_t *a, int32_t v, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] = v;
}

// storage rows of the dense block's rows (and the inverse map)
__global__ void k_blk_map(const int32_t *rows, int32_t nb, const int32_t *iperm, int32_t *srow,
                          int32_t *smap) {
  const int32_t c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= nb) return;
  const int32_t r = iperm ? iperm[rows[c]] : rows[c];
  srow[c] = r;
  smap[r] = c;
}

// Jacobi entries of the block rows: 1 / (a_rr + scale B_cc), or 1 for a
// zero diagonal, as k_fill_vals takes 1 / a_rr
__global__ void k_blk_dinv(const int32_t *row_ptr, const int32_t *col, const double *val,
                           const int32_t *rows, int32_t nb, const double *B, double scale,
                           const int32_t *srow, double *dinv) {
  const int32_t c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= nb) return;
  const int32_t r = rows[c];
  double arr = 0.0;
  for (int32_t p = row_ptr[r]; p < row_ptr[r + 1]; ++p)
    if (col[p] == r) arr = val[p];
  const double d = arr + scale * B[(int64_t)c * nb + c];
  dinv[srow[c]] = d != 0.0 ? 1.0 / d : 1.0;
}

// The dense block of A in the storage's row order, and its Jacobi entries
// (after the values' refresh on the same stream).
int block_setup(lfgpu_ctx *ctx, const lfgpu_csr *A, lfgpu_solver_cache *c) {
  hipStream_t s = ctx->stream;
  const int32_t n = A->nrows, nb = A->blk_n;
  c->blk_n = nb;
  if (nb == 0) return LFGPU_OK;
  if (!c->blk_smap) LFG_TRY(dalloc(&c->blk_smap, (size_t)std::max(n, 1), "block row map"));
  if (nb > c->blk_cap) {
    dfree(c->blk_srow);
    dfree(c->blk_t);
    c->blk_srow = nullptr;
    c->blk_t = nullptr;
    c->blk_cap = 0;
    LFG_TRY(dalloc(&c->blk_srow, (size_t)nb, "block storage rows"));
    LFG_TRY(dalloc(&c->blk_t, (size_t)nb, "block product"));
    c->blk_cap = nb;
  }
  hipLaunchKernelGGL(k_fill_i32, dim3(grid_for(n, 256)), dim3(256), 0, s, c->blk_smap, -1,
                     (int64_t)n);
  hipLaunchKernelGGL(k_blk_map, dim3(grid_for(nb, 256)), dim3(256), 0, s, A->blk_rows, nb,
                     c->iperm, c->blk_srow, c->blk_smap);
  hipLaunchKernelGGL(k_blk_dinv, dim3(grid_for(nb, 256)), dim3(256), 0, s, A->row_ptr, A->col,
                     A->val, A->blk_rows, nb, A->blk, A->blk_scale, c->blk_srow, c->dinv);
  LFG_HIP(hipGetLastError());
  return LFGPU_OK;
}

BlkView blk_view(const lfgpu_csr *A, const lfgpu_solver_cache *c) {
  BlkView b;
  if (c->blk_n == 0) return b;
  b.B = A->blk;
  b.srow = c->blk_srow;
  b.smap = c->blk_smap;
  b.t = c->blk_t;
  b.scale = A->blk_scale;
  b.n = c->blk_n;
  return b;
}

// Builds (or refreshes) the solver's storage of A.  allow_perm: a new
// sliced-ELL storage may keep its rows in the locality order; the batched
// solver needs the original row order (its problems own row ranges) and
// rebuilds a permuted storage.
int ensure_cache(lfgpu_ctx *ctx, lfgpu_csr *A, bool need32, bool allow_perm) {
  hipStream_t s = ctx->stream;
  if (A->cache && A->cache->perm && !allow_perm) {
    LFG_HIP(hipStreamSynchronize(s));
    solver_cache_free(A);
  }
  lfgpu_solver_cache *c = A->cache;
  const int32_t n = A->nrows;
  if (!c) {
    c = new lfgpu_solver_cache();
    A->cache = c;
    DiagSet ds{};
    LFG_TRY(detect_dia(ctx, A, &ds));
    if (ds.ndiag > 0) {
      c->fmt = FMT_DIA;
      c->ndiag = ds.ndiag;
      for (int d = 0; d < kMaxDiag; ++d) c->off[d] = d < ds.ndiag ? ds.off[d] : 0;
      c->mlen = (int64_t)ds.ndiag * n;
    } else {
      c->fmt = FMT_ELL;
      if (allow_perm && n > 0) LFG_TRY(ell_locality_order(ctx, A, c));
      c->nslices = (n + kSlice - 1) / kSlice;
      int32_t *w = nullptr;
      LFG_TRY(dalloc(&w, (size_t)c->nslices, "slice widths"));
      LFG_TRY(dalloc(&c->slice_off, (size_t)c->nslices + 1, "slice offsets"));
      hipLaunchKernelGGL(k_slice_width, dim3(grid_for(c->nslices, 256)), dim3(256), 0, s,
                         A->row_ptr, A->col, n, c->nslices, c->perm, w);
      LFG_HIP(hipGetLastError());
      int64_t len = 0;
      int rc = exclusive_scan_i32(s, w, c->slice_off, c->nslices, &len);
      dfree(w);
      LFG_TRY(rc);
      c->mlen = len;
      // 16-bit column offsets when every stored column lies within 32767
      // rows of its row (a locality-ordered FE matrix): 2 of the 12 bytes of
      // an FP64 slot
      int *flag = nullptr;
      LFG_TRY(dalloc(&flag, 1, "ELL span flag"));
      int hflag = 1;
      hipError_t e = hipMemsetAsync(flag, 0, 4, s);
      if (e == hipSuccess) {
        hipLaunchKernelGGL(k_ell_span, dim3(grid_for(n, 256)), dim3(256), 0, s, A->row_ptr,
                           A->col, n, c->perm, c->iperm, flag);
        e = hipGetLastError();
      }
      if (e == hipSuccess) e = hipMemcpyAsync(&hflag, flag, 4, hipMemcpyDeviceToHost, s);
      if (e == hipSuccess) e = hipStreamSynchronize(s);
      dfree(flag);
      if (e != hipSuccess) return hip_fail(e, "ELL column span");
      if (hflag == 0) {
        LFG_TRY(dalloc(&c->s_dcol, (size_t)std::max<int64_t>(len, 1), "ELL column offsets"));
        hipLaunchKernelGGL(k_fill_ell_cols<int16_t>, dim3(grid_for(n, 256)), dim3(256), 0, s,
                           A->row_ptr, A->col, n, c->slice_off, c->perm, c->iperm, c->s_dcol);
      } else {
        LFG_TRY(dalloc(&c->s_col, (size_t)std::max<int64_t>(len, 1), "ELL columns"));
        hipLaunchKernelGGL(k_fill_ell_cols<int32_t>, dim3(grid_for(n, 256)), dim3(256), 0, s,
                           A->row_ptr, A->col, n, c->slice_off, c->perm, c->iperm, c->s_col);
      }
      LFG_HIP(hipGetLastError());
    }
    LFG_TRY(dalloc(&c->val, (size_t)std::max<int64_t>(c->mlen, 1), "solver values"));
    LFG_TRY(dalloc(&c->rowsum, (size_t)n, "row sums"));
    LFG_TRY(dalloc(&c->dinv, (size_t)n, "inverse diagonal"));
    LFG_TRY(dalloc(&c->r, (size_t)n, "PCG r"));
    LFG_TRY(dalloc(&c->p, (size_t)n, "PCG p"));
    LFG_TRY(dalloc(&c->q, (size_t)n, "PCG q"));
    LFG_TRY(dalloc(&c->part, (size_t)kPartSlots * kMaxGrid, "reduction partials"));
    LFG_TRY(dalloc(&c->st, 1, "PCG state"));
  }
  if (need32 && !c->have32) {
    LFG_TRY(dalloc(&c->val32, (size_t)std::max<int64_t>(c->mlen, 1), "solver values FP32"));
    LFG_TRY(dalloc(&c->rowsum32, (size_t)n, "row sums FP32"));
    LFG_TRY(dalloc(&c->dinv32, (size_t)n, "inverse diagonal FP32"));
    LFG_TRY(dalloc(&c->dinv16, (size_t)n, "inverse diagonal BF16"));
    LFG_TRY(dalloc(&c->r32, (size_t)n, "PCG r32"));
    LFG_TRY(dalloc(&c->p32, (size_t)n, "PCG p32"));
    LFG_TRY(dalloc(&c->q32, (size_t)n, "PCG q32"));
    LFG_TRY(dalloc(&c->xa32, (size_t)n, "PCG x32"));
    LFG_TRY(dalloc(&c->xlo, (size_t)n, "PCG x low part"));
    LFG_TRY(dalloc(&c->r64, (size_t)n, "PCG true residual"));
    LFG_TRY(dalloc(&c->xa64, (size_t)n, "PCG increment"));
    c->have32 = true;
    c->built_version = ~0ull;
  }
  if (c->built_version != A->val_version) {
    fill_vals_enqueue(s, A, c);
    LFG_HIP(hipGetLastError());
    LFG_TRY(dia_symmetry(ctx, c, n));
    LFG_TRY(block_setup(ctx, A, c));
    c->built_version = A->val_version;
  }
  return LFGPU_OK;
}

// ---------------------------------------------------------------------------
// SpMV
// ---------------------------------------------------------------------------
template <int FMT>
__global__ __launch_bounds__(256) void k_spmv_f64(MatView m, const double *__restrict__ x,
                                                  double *__restrict__ y) {
  const int32_t r = blockIdx.x * 256 + threadIdx.x;
  if (r >= m.n) return;
  y[r] = apply64<FMT, true>(m, x, r, x[r]);
}

// y = A x with the storage's rows in the locality order: x is gathered into
// that order, and row r of the storage is written to y[perm[r]]
__global__ __launch_bounds__(256) void k_spmv_perm(MatView m, const double *__restrict__ xp,
                                                   const int32_t *__restrict__ perm,
                                                   double *__restrict__ y) {
  const int32_t r = blockIdx.x * 256 + threadIdx.x;
  if (r >= m.n) return;
  y[perm[r]] = apply64<FMT_ELL>(m, xp, r, xp[r]);
}

int launch_spmv(hipStream_t s, const lfgpu_solver_cache *c, int32_t n, const double *x,
                double *y) {
  const MatView m = view_of(c, n);
  if (c->perm) {
    hipLaunchKernelGGL(k_gather, dim3(grid_for(n, 256)), dim3(256), 0, s, x, c->perm, n, c->xp);
    hipLaunchKernelGGL(k_spmv_perm, dim3(grid_for(n, 256)), dim3(256), 0, s, m, c->xp, c->perm,
                       y);
    LFG_HIP(hipGetLastError());
    return LFGPU_OK;
  }
  if (c->fmt == FMT_DIA)
    hipLaunchKernelGGL(k_spmv_f64<FMT_DIA>, dim3(grid_for(n, 256)), dim3(256), 0, s, m, x, y);
  else
    hipLaunchKernelGGL(k_spmv_f64<FMT_ELL>, dim3(grid_for(n, 256)), dim3(256), 0, s, m, x, y);
  LFG_HIP(hipGetLastError());
  return LFGPU_OK;
}

// ---------------------------------------------------------------------------
// PCG sweeps.  A fixed grid of G workgroups; workgroup g owns the contiguous
// rows [g R, (g+1) R) (R a multiple of the workgroup size, so each wavefront covers whole
// 64-row slices) and leaves one partial sum per reduced quantity.
// ---------------------------------------------------------------------------
// Threads per sweep workgroup: 256 for the FP64 sweeps, 512 for the mixed
// sweeps (each measured fastest for its kernels on MI355X at 2048^2).
constexpr int kSB64 = 256;
constexpr int kSBMix = 512;


// Per-kernel timing of the solver sweeps (lfgpu_pcg_sweep_times): while a
// timer is installed on the calling thread, the chunk launchers bracket every
// sweep kernel with a pair of events on the solver's stream.  Kind 0 is the
// first sweep of an iteration (the product), 1 the second (the update), 2 a
// one-workgroup reduction kernel of the row sweeps.
struct SweepTimer {
  std::vector<hipEvent_t> ev;  // start / end pairs in launch order
  std::vector<int> kind;       // one per pair
  size_t used = 0;             // pairs recorded
  hipError_t err = hipSuccess;
};
thread_local SweepTimer *t_timer = nullptr;

inline void tmark_begin(hipStream_t s, int kind) {
  SweepTimer *t = t_timer;
  if (!t || 2 * t->used + 1 >= t->ev.size()) return;
  const hipError_t e = hipEventRecord(t->ev[2 * t->used], s);
  if (e != hipSuccess && t->err == hipSuccess) t->err = e;
  t->kind[t->used] = kind;
}
inline void tmark_end(hipStream_t s) {
  SweepTimer *t = t_timer;
  if (!t || 2 * t->used + 1 >= t->ev.size()) return;
  const hipError_t e = hipEventRecord(t->ev[2 * t->used + 1], s);
  if (e != hipSuccess && t->err == hipSuccess) t->err = e;
  t->used += 1;
}

// Chunks of iterations with the host one step behind the device.  A step
// queues k chunks (launch() queues one), a copy of the solver state into
// pinned host slot j (copy(j)) and an event.  The host queues step j + 1
// before it waits for step j, so the device does not idle at a poll; once
// step j shows that every problem has stopped, step j + 1 runs empty (each
// kernel returns on its done flag) and completes before anything queued
// after it.  remaining(j) reads slot j: < 0 once every problem has stopped,
// else an estimate of the iterations still needed (0: unknown); the next
// step queued takes half of that in chunks, between 1 and
// kMaxChunksPerPoll.  *slot receives the slot of the final state, *syncs
// counts the host waits.
constexpr int kMaxChunksPerPoll = 8;

template <class Launch, class Copy, class Remaining>
int poll_chunks(lfgpu_ctx *ctx, int check, Launch &&launch, Copy &&copy, Remaining &&remaining,
                int *slot, int *syncs) {
  hipStream_t s = ctx->stream;
  auto step = [&](int j, int chunks) -> int {
    for (int q = 0; q < chunks; ++q) LFG_TRY(launch());
    LFG_TRY(copy(j));
    LFG_HIP(hipEventRecord(ctx->evp[j], s));
    return LFGPU_OK;
  };
  int j = 0, k = 1;
  LFG_TRY(step(0, 1));
  for (;;) {
    LFG_TRY(step(j ^ 1, k));
    LFG_HIP(hipEventSynchronize(ctx->evp[j]));
    ++*syncs;
    const double est = remaining(j);
    if (est < 0.0) break;
    // the per-kernel timer keeps one chunk per step: few empty launches
    k = t_timer ? 1
                : (int)std::min<double>(kMaxChunksPerPoll, std::max(1.0, std::floor(0.5 * est / check)));
    j ^= 1;
  }
  *slot = j;
  return LFGPU_OK;
}

// Iterations still needed to bring ||r||^2 from rr down to thr at the rate
// observed since the previous poll (rr0 at iteration it0), at most what the
// cap leaves; 0 when the history gives no estimate.  With thr = 0 (a fixed
// number of iterations) the cap alone.
inline double remaining_iters(double rr0, int it0, double rr, int it, double thr, int maxit) {
  const double cap = std::max(0.0, (double)maxit - it);
  if (!(thr > 0.0)) return cap;
  if (!(rr > thr) || !(rr < rr0) || it <= it0) return 0.0;
  return std::min(cap, (it - it0) * (std::log(thr / rr) / std::log(rr / rr0)));
}

// Host-side history of one problem's polled residual.  rr_now is the
// residual measured at iteration it_rr (every iteration in FP64; at the last
// reliable update in the mixed modes), it_now the current iteration.
struct PollHist {
  double rr = 0.0;
  int it = -1;
  double est = 0.0;  // iterations still needed after iteration `it`
  double next(double rr_now, int it_rr, int it_now, double thr, int maxit) {
    if (it_rr != it) {
      est = it >= 0 ? remaining_iters(rr, it, rr_now, it_rr, thr, maxit) : 0.0;
      rr = rr_now;
      it = it_rr;
    }
    return est > 0.0 ? std::max(1.0, est - (it_now - it)) : 0.0;
  }
};

struct Sweep {
  int32_t n;
  int32_t rows_per_group;
  int ngroups;
};

Sweep make_sweep(int32_t n, int sb, int min_rows_per_thread = 1) {
  int64_t per = (n + kMaxGrid - 1) / kMaxGrid;
  per = ((per + sb - 1) / sb) * sb;
  if (per < (int64_t)min_rows_per_thread * sb) per = (int64_t)min_rows_per_thread * sb;
  Sweep s;
  s.n = n;
  s.rows_per_group = (int32_t)per;
  s.ngroups = (int)((n + per - 1) / per);
  if (s.ngroups < 1) s.ngroups = 1;
  return s;
}

// Row chunk of workgroup b of a G-workgroup sweep.  Workgroups are handed to
// the 8 XCDs round-robin (b -> XCD b % 8), so chunk (b % 8) G/8 + b / 8 gives
// every XCD a contiguous band of rows: the off-diagonal neighbours of a chunk
// (rows +-(n+1) on a structured mesh) then live in chunks that the same XCD,
// and so the same L2, processes at about the same time.  A bijection on
// [0, G); the last G % 8 workgroups keep their own index.
__device__ __forceinline__ int32_t sweep_chunk(int b, int G) {
  const int per = G >> 3;
  return b < 8 * per ? (b & 7) * per + (b >> 3) : b;
}

// one partial per 256-thread workgroup (batched solver)
__device__ __forceinline__ void store_partial(double *part, double v, double *sh) {
  v = block_sum<256>(v, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = v;
}

// Several sums at once over a sweep workgroup of SB threads; every thread
// ends with the totals (summed in the same order in every workgroup).  sh
// needs (SB/64)*K doubles.
template <int SB, int K>
__device__ __forceinline__ void block_sum_k(double (&v)[K], double *sh) {
#pragma unroll
  for (int k = 0; k < K; ++k)
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v[k] += __shfl_xor(v[k], o, 64);
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  __syncthreads();
  if (l == 0)
#pragma unroll
    for (int k = 0; k < K; ++k) sh[(SB / 64) * k + w] = v[k];
  __syncthreads();
#pragma unroll
  for (int k = 0; k < K; ++k) {
    double t = sh[(SB / 64) * k];
#pragma unroll
    for (int q = 1; q < (SB / 64); ++q) t += sh[(SB / 64) * k + q];
    v[k] = t;
  }
}

// Partial sums are stored slot-major (part[slot * kMaxGrid + group]): the
// redundant re-reduction at the start of the next sweep reads, per slot, one
// contiguous run of ngroups doubles in every workgroup.
template <int SB, int K>
__device__ __forceinline__ void store_partials_k(double *part, double (&v)[K], const int (&slots)[K],
                                                 double *sh) {
  block_sum_k<SB, K>(v, sh);
  if (threadIdx.x == 0)
#pragma unroll
    for (int k = 0; k < K; ++k) part[slots[k] * kMaxGrid + blockIdx.x] = v[k];
}

// Every load of a thread's partials is issued before the first addition
// (one L2 round trip at the start of the sweep instead of one per group of
// SB partials); the additions keep the order g = t, t + SB, ...
template <int SB, int K>
__device__ __forceinline__ void reduce_partials_k(const double *part, int ngroups,
                                                  const int (&slots)[K], double (&out)[K],
                                                  double *sh) {
  constexpr int U = (kMaxGrid + SB - 1) / SB;
  double v[K][U];
#pragma unroll
  for (int u = 0; u < U; ++u) {
    const int g = threadIdx.x + u * SB;
#pragma unroll
    for (int k = 0; k < K; ++k) v[k][u] = g < ngroups ? part[slots[k] * kMaxGrid + g] : 0.0;
  }
#pragma unroll
  for (int k = 0; k < K; ++k) {
    out[k] = 0.0;
#pragma unroll
    for (int u = 0; u < U; ++u)
      if (threadIdx.x + u * SB < ngroups) out[k] += v[k][u];
  }
  block_sum_k<SB, K>(out, sh);
}

// partial slots (FP64 mode); P_RR and P_RZ are followed by their parity twin
constexpr int P_PQ = 0, P_QDR = 1, P_QDQ = 2, P_RR = 3, P_RZ = 5, P_BB = 7;


struct CgArgs {
  MatView m;
  const double *__restrict__ dinv;
  double *__restrict__ x;
  double *__restrict__ r;
  double *__restrict__ p;
  double *__restrict__ q;
  double *part;
  PcgState *st;
  Sweep sw;
  BlkView blk;  // n == 0: none
};

// t = B v[srow] for the dense block, one wave per block row, each row summed
// in a fixed order (lane l adds the columns l, l + 64, ... in turn, then a
// butterfly); the loads of 8 columns per lane are issued before their
// products are added.  Returns at once when the solve has stopped (st
// non-null).
__global__ __launch_bounds__(256) void k_blk_gemv(BlkView b, const double *__restrict__ v,
                                                  const PcgState *st) {
  if (st && st->done) return;
  const int32_t row = blockIdx.x * 4 + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (row >= b.n) return;
  const double *__restrict__ Br = b.B + (int64_t)row * b.n;
  double acc = 0.0;
  for (int32_t c0 = lane; c0 < b.n; c0 += 8 * 64) {
    double bv[8], xv[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const int32_t c = c0 + 64 * u;
      const bool in = c < b.n;
      bv[u] = in ? Br[c] : 0.0;
      xv[u] = in ? v[b.srow[c]] : 0.0;
    }
#pragma unroll
    for (int u = 0; u < 8; ++u)
      if (c0 + 64 * u < b.n) acc += bv[u] * xv[u];
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o, 64);
  if (lane == 0) b.t[row] = acc;
}

inline void launch_blk_gemv(hipStream_t s, const BlkView &b, const double *v,
                            const PcgState *st) {
  hipLaunchKernelGGL(k_blk_gemv, dim3(grid_for(b.n, 4)), dim3(256), 0, s, b, v, st);
}

template <bool BLK>
__device__ __forceinline__ double blk_term(const BlkView &b, int32_t r) {
  if constexpr (BLK) {
    const int32_t k = b.smap[r];
    return k >= 0 ? b.scale * b.t[k] : 0.0;
  } else {
    return 0.0;
  }
}

// setup: r = b - A x, p = D^-1 r; partials of b.b, r.r, r.D^-1 r (the parity
// 1 slots, which iteration 0 reads)
// (BLK: the dense block's product B x of the preceding k_blk_gemv is added
// to A x)
template <int FMT, int SB, bool BLK = false>
__global__ __launch_bounds__(SB) void k_cg_init(CgArgs a, const double *__restrict__ b) {
  __shared__ double sh[4 * (SB / 64)];
  const int32_t r0 = sweep_chunk(blockIdx.x, a.sw.ngroups) * a.sw.rows_per_group;
  const int32_t r1 = min(r0 + a.sw.rows_per_group, a.sw.n);
  double bb = 0.0, rr = 0.0, rz = 0.0;
  for (int32_t i = r0 + threadIdx.x; i < r1; i += SB) {
    const double xi = a.x[i];
    const double bi = b[i];
    const double ri = bi - (apply64<FMT>(a.m, a.x, i, xi) + blk_term<BLK>(a.blk, i));
    a.r[i] = ri;
    const double zi = a.dinv[i] * ri;
    a.p[i] = zi;
    bb += bi * bi;
    rr += ri * ri;
    rz += ri * zi;
  }
  double v[3] = {bb, rr, rz};
  const int sl[3] = {P_BB, P_RR + 1, P_RZ + 1};
  store_partials_k<SB, 3>(a.part, v, sl, sh);
}

// one workgroup: Eigen's early exits (zero rhs, converged start, maxit 0)
template <int SB>
__global__ __launch_bounds__(SB) void k_cg_init_finish(CgArgs a) {
  __shared__ double sh[4 * (SB / 64)];
  double t[3];
  const int sl[3] = {P_BB, P_RR + 1, P_RZ + 1};
  reduce_partials_k<SB, 3>(a.part, a.sw.ngroups, sl, t, sh);
  const double bb = t[0], rr = t[1], rz = t[2];
  if (threadIdx.x == 0) {
    PcgState *st = a.st;
    st->bb = bb;
    st->rr = rr;
    st->rr_ref = rr;
    st->rho[0] = rz;
    st->rho[1] = rz;
    st->iter = -1;
    st->iterp[0] = st->iterp[1] = 0;
    st->refine[0] = st->refine[1] = 0;
    st->converged = 0;
    st->done = 0;
    st->refinements = 0;
    st->last_refine = 0;
    const double thr = st->thr * bb;  // thr holds tol^2 on entry
    st->thr = thr > 2.2250738585072014e-308 ? thr : 2.2250738585072014e-308;
    if (bb == 0.0 || rr < st->thr) {
      st->done = 1;
      st->converged = 1;
      st->iter = 0;
    } else if (st->maxit <= 0) {
      st->done = 1;
      st->iter = 0;
    }
  }
}

__global__ void k_zero_if_bb0(double *x, int32_t n, const PcgState *st) {
  const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n && st->bb == 0.0) x[i] = 0.0;
}

// FP64 iteration k, sweep 1 (par = k & 1): Eigen's convergence test on r_k
// (||r_k||^2 left by the previous sweep 2), then q = A p with the partials
// p.q, q.D^-1 r and q.D^-1 q.
// BLK: the dense block's product B p of the preceding k_blk_gemv is added to
// the sparse part's.
template <int FMT, int SB, bool SYM = false, bool BLK = false>
__global__ __launch_bounds__(SB) void k_cg_spmv(CgArgs a, int par) {
  __shared__ double sh[4 * (SB / 64)];
  PcgState *st = a.st;
  if (st->done) return;
  double t[1];
  const int sl_in[1] = {P_RR + (par ^ 1)};
  reduce_partials_k<SB, 1>(a.part, a.sw.ngroups, sl_in, t, sh);
  const double rr = t[0];
  if (rr < st->thr) {  // converged: the iteration count stays as it is
    if (blockIdx.x == 0 && threadIdx.x == 0) {
      st->rr = rr;
      st->done = 1;
      st->converged = 1;
      if (st->iter < 0) st->iter = 0;
    }
    return;
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    st->rr = rr;
    st->iter += 1;
    if (st->iter >= st->maxit) st->done = 1;
  }
  const int32_t r0 = sweep_chunk(blockIdx.x, a.sw.ngroups) * a.sw.rows_per_group;
  const int32_t r1 = min(r0 + a.sw.rows_per_group, a.sw.n);
  double pq = 0.0, qdr = 0.0, qdq = 0.0;
  for (int32_t i = r0 + threadIdx.x; i < r1; i += SB) {
    // loads before the store: the compiler keeps stores and later loads in
    // program order
    const double pi = a.p[i], di = a.dinv[i], ri = a.r[i];
    const double qi = apply64<FMT, true, SYM>(a.m, a.p, i, pi) + blk_term<BLK>(a.blk, i);
    // in the 512-thread sweeps (2^21 rows and more) q goes past the caches:
    // it is read once, by the update sweep (2% faster at 2048^2); the same
    // store measured slower in the mixed and row sweeps
    if constexpr (SB == 512) __builtin_nontemporal_store(qi, &a.q[i]);
    else a.q[i] = qi;
    const double dq = di * qi;
    pq += pi * qi;
    qdr += dq * ri;
    qdq += dq * qi;
  }
  double v[3] = {pq, qdr, qdq};
  const int sl[3] = {P_PQ, P_QDR, P_QDQ};
  store_partials_k<SB, 3>(a.part, v, sl, sh);
}

// FP64 iteration k, sweep 2: alpha = rho_k / (p.q);
//   rho_{k+1} = rho_k - 2 alpha (q, D^-1 r) + alpha^2 (q, D^-1 q), the exact
//   expansion of (r - alpha q, D^-1 (r - alpha q)); beta = rho_{k+1} / rho_k;
//   x += alpha p ; r -= alpha q ; p = D^-1 r + beta p ;
//   partials ||r||^2 and (r, D^-1 r) of the new residual (parity par).
template <int SB>
__global__ __launch_bounds__(SB) void k_cg_update(CgArgs a, int par) {
  __shared__ double sh[4 * (SB / 64)];
  PcgState *st = a.st;
  if (st->done) return;
  double t[4];
  const int sl_in[4] = {P_RZ + (par ^ 1), P_PQ, P_QDR, P_QDQ};
  reduce_partials_k<SB, 4>(a.part, a.sw.ngroups, sl_in, t, sh);
  const double rho = t[0], pq = t[1], qdr = t[2], qdq = t[3];
  const double alpha = rho / pq;
  const double rho_next = rho - 2.0 * alpha * qdr + alpha * alpha * qdq;
  const double beta = rho_next / rho;
  const int32_t r0 = sweep_chunk(blockIdx.x, a.sw.ngroups) * a.sw.rows_per_group;
  const int32_t r1 = min(r0 + a.sw.rows_per_group, a.sw.n);
  double rr = 0.0, rz = 0.0;
  for (int32_t i = r0 + threadIdx.x; i < r1; i += SB) {
    // all loads of the row before its stores; x and q are not read again
    // before the next write, so they bypass the caches and leave room for
    // p, r and D^-1, which the next sweep reads
    const double pi = a.p[i], ri0 = a.r[i], di = a.dinv[i];
    // the 512-thread sweeps run the systems of 2^21 rows and more, whose
    // per-iteration working set exceeds the last-level cache: there x and q
    // bypass it (2048^2: 964 -> 947 ms); at 1024^2 the working set fits and
    // plain accesses measured 3% faster
    double xi, qi;
    if constexpr (SB == 512) {
      xi = __builtin_nontemporal_load(&a.x[i]);
      qi = __builtin_nontemporal_load(&a.q[i]);
    } else {
      xi = a.x[i];
      qi = a.q[i];
    }
    const double ri = ri0 - alpha * qi;
    const double zi = di * ri;
    if constexpr (SB == 512) __builtin_nontemporal_store(xi + alpha * pi, &a.x[i]);
    else a.x[i] = xi + alpha * pi;
    a.r[i] = ri;
    a.p[i] = zi + beta * pi;
    rr += ri * ri;
    rz += ri * zi;
  }
  double v[2] = {rr, rz};
  const int sl[2] = {P_RR + par, P_RZ + par};
  store_partials_k<SB, 2>(a.part, v, sl, sh);
}

// ---------------------------------------------------------------------------
// FP64 PCG on sliced ELL.  The gathers of a row reach several distant parts
// of the vector (vertex, edge and cell DOFs at P3), so chunked sweeps, whose
// active rows are spread over the whole matrix, miss L2 on every column;
// here both sweeps run one row per thread in index order (the active rows
// and columns form a compact moving window) and a one-workgroup kernel
// after each sweep sums the per-workgroup partials in a fixed order and
// takes the scalar decisions.  Same recurrences and stopping test as
// k_cg_spmv / k_cg_update.
// ---------------------------------------------------------------------------
constexpr int kRB = 256;    // rows per workgroup of the row sweeps
constexpr int kRedB = 1024; // threads of the reduction kernels

// This thread's share of the K per-workgroup partial sums (part[k nwg + g],
// g = t, t + kRedB, ...), added in that order; four groups' loads are issued
// before their additions.
template <int K>
__device__ __forceinline__ void sum_partials(const double *part, int nwg, double (&t)[K]) {
#pragma unroll
  for (int k = 0; k < K; ++k) t[k] = 0.0;
  int g = threadIdx.x;
  for (; g + 3 * kRedB < nwg; g += 4 * kRedB) {
    double v[K][4];
#pragma unroll
    for (int u = 0; u < 4; ++u)
#pragma unroll
      for (int k = 0; k < K; ++k) v[k][u] = part[(int64_t)k * nwg + g + u * kRedB];
#pragma unroll
    for (int u = 0; u < 4; ++u)
#pragma unroll
      for (int k = 0; k < K; ++k) t[k] += v[k][u];
  }
  for (; g < nwg; g += kRedB)
#pragma unroll
    for (int k = 0; k < K; ++k) t[k] += part[(int64_t)k * nwg + g];
}

struct RowArgs {
  MatView m;
  const double *__restrict__ dinv;
  double *__restrict__ x;
  double *__restrict__ r;
  double *__restrict__ p;
  double *__restrict__ q;
  double *part;  // slot k of group g at part[k * nwg + g]
  int nwg;
  PcgState *st;
};

template <int K>
__device__ __forceinline__ void row_partials(const RowArgs &a, double (&v)[K], double *sh) {
  block_sum_k<kRB, K>(v, sh);
  if (threadIdx.x == 0)
#pragma unroll
    for (int k = 0; k < K; ++k) a.part[(int64_t)k * a.nwg + blockIdx.x] = v[k];
}

template <int K>
__device__ __forceinline__ void row_totals(const RowArgs &a, double (&t)[K], double *sh) {
  sum_partials<K>(a.part, a.nwg, t);
  block_sum_k<kRedB, K>(t, sh);
}

template <int FMT>
__global__ __launch_bounds__(kRB) void k_cgr_init(RowArgs a, const double *__restrict__ b) {
  __shared__ double sh[4 * (kRB / 64)];
  const int32_t i = blockIdx.x * kRB + threadIdx.x;
  double v[3] = {0.0, 0.0, 0.0};
  if (i < a.m.n) {
    const double bi = b[i];
    const double ri = bi - apply64<FMT>(a.m, a.x, i, a.x[i]);
    a.r[i] = ri;
    const double zi = a.dinv[i] * ri;
    a.p[i] = zi;
    v[0] = bi * bi;
    v[1] = ri * ri;
    v[2] = ri * zi;
  }
  row_partials<3>(a, v, sh);
}

// Eigen's early exits (zero rhs, converged start, maxit 0), else pass 0 starts
__global__ __launch_bounds__(kRedB) void k_cgr_red_init(RowArgs a) {
  __shared__ double sh[4 * (kRedB / 64)];
  double t[3];
  row_totals<3>(a, t, sh);
  if (threadIdx.x == 0) {
    PcgState *st = a.st;
    const double bb = t[0], rr = t[1];
    st->bb = bb;
    st->rr = rr;
    st->rho[0] = t[2];
    st->converged = 0;
    st->done = 0;
    st->refinements = 0;
    const double thr = st->thr * bb;  // thr holds tol^2 on entry
    st->thr = thr > 2.2250738585072014e-308 ? thr : 2.2250738585072014e-308;
    st->iter = 0;
    if (bb == 0.0 || rr < st->thr) {
      st->done = 1;
      st->converged = 1;
    } else if (st->maxit <= 0) {
      st->done = 1;
    }
  }
}

template <int FMT>
__global__ __launch_bounds__(kRB) void k_cgr_spmv(RowArgs a) {
  __shared__ double sh[4 * (kRB / 64)];
  if (a.st->done) return;
  const int32_t i = blockIdx.x * kRB + threadIdx.x;
  double v[3] = {0.0, 0.0, 0.0};
  if (i < a.m.n) {
    const double pi = a.p[i];
    const double qi = apply64<FMT, false, false, false, true>(a.m, a.p, i, pi);
    a.q[i] = qi;
    const double dq = a.dinv[i] * qi;
    v[0] = pi * qi;
    v[1] = dq * a.r[i];
    v[2] = dq * qi;
  }
  row_partials<3>(a, v, sh);
}

// alpha = rho / (p, q); beta from rho' = rho - 2 alpha (q, D^-1 r) +
// alpha^2 (q, D^-1 q), as k_cg_update
__global__ __launch_bounds__(kRedB) void k_cgr_red_s(RowArgs a) {
  __shared__ double sh[4 * (kRedB / 64)];
  if (a.st->done) return;
  double t[3];
  row_totals<3>(a, t, sh);
  if (threadIdx.x == 0) {
    PcgState *st = a.st;
    const double rho = st->rho[0], pq = t[0], qdr = t[1], qdq = t[2];
    const double alpha = rho / pq;
    const double rho_next = rho - 2.0 * alpha * qdr + alpha * alpha * qdq;
    st->alpha_s = alpha;
    st->beta_s = rho_next / rho;
  }
}

__global__ __launch_bounds__(kRB) void k_cgr_update(RowArgs a) {
  __shared__ double sh[4 * (kRB / 64)];
  const PcgState *st = a.st;
  if (st->done) return;
  const double alpha = st->alpha_s, beta = st->beta_s;
  const int32_t i = blockIdx.x * kRB + threadIdx.x;
  double v[2] = {0.0, 0.0};
  if (i < a.m.n) {
    // all loads of the row before its stores (stores and later loads stay
    // in program order)
    const double pi = a.p[i], r0 = a.r[i], di = a.dinv[i];
    const double xi = __builtin_nontemporal_load(&a.x[i]);
    const double qi = __builtin_nontemporal_load(&a.q[i]);
    const double ri = r0 - alpha * qi;
    const double zi = di * ri;
    __builtin_nontemporal_store(xi + alpha * pi, &a.x[i]);
    a.r[i] = ri;
    a.p[i] = zi + beta * pi;
    v[0] = ri * ri;
    v[1] = ri * zi;
  }
  row_partials<2>(a, v, sh);
}

// Eigen's test on the new residual; otherwise the next pass starts (and
// stops before its update once maxit passes are done)
__global__ __launch_bounds__(kRedB) void k_cgr_red_u(RowArgs a) {
  __shared__ double sh[4 * (kRedB / 64)];
  if (a.st->done) return;
  double t[2];
  row_totals<2>(a, t, sh);
  if (threadIdx.x == 0) {
    PcgState *st = a.st;
    st->rr = t[0];
    st->rho[0] = t[1];
    if (t[0] < st->thr) {
      st->done = 1;
      st->converged = 1;
    } else {
      st->iter += 1;
      if (st->iter >= st->maxit) st->done = 1;
    }
  }
}

template <int FMT>
void launch_cgr_chunk(hipStream_t s, const RowArgs &a, int count) {
  const dim3 G(a.nwg), B(kRB), B1(kRedB);
  for (int k = 0; k < count; ++k) {
    tmark_begin(s, 0);
    hipLaunchKernelGGL(k_cgr_spmv<FMT>, G, B, 0, s, a);
    tmark_end(s);
    tmark_begin(s, 2);
    hipLaunchKernelGGL(k_cgr_red_s, dim3(1), B1, 0, s, a);
    tmark_end(s);
    tmark_begin(s, 1);
    hipLaunchKernelGGL(k_cgr_update, G, B, 0, s, a);
    tmark_end(s);
    tmark_begin(s, 2);
    hipLaunchKernelGGL(k_cgr_red_u, dim3(1), B1, 0, s, a);
    tmark_end(s);
  }
}

// ---------------------------------------------------------------------------
// FP64 single-reduction PCG (Chronopoulos-Gear): one sweep per iteration.
//   u = D^-1 r, w = A u;  gamma = (r, u), delta = (w, u)
//   beta = gamma / gamma_prev, alpha = gamma / (delta - beta gamma / alpha_prev)
//   p = u + beta p, s = w + beta s (s = A p), x += alpha p, r -= alpha s,
//   then the next w = A D^-1 r.
// The product needs the new u at the off-diagonal neighbours, which other
// workgroups write in the same sweep; every row therefore re-derives it from
// the previous r, w, s of the neighbour with the same operations its owner
// uses (bit-identical), and r, w, s are double-buffered by iteration parity.
// p and x are only touched by their own row.
// ---------------------------------------------------------------------------
constexpr int C_D = 0, C_RR = 3, C_G = 5;  // partial slots (+parity)

struct Cg1Args {
  MatView m;
  const double *__restrict__ dinv;
  double *__restrict__ x;
  double *__restrict__ p;
  double *r[2], *w[2], *s[2];
  double *part;
  PcgState *st;
  Sweep sw;
};

// The new preconditioned residual of row j, u_j = D^-1 (r_j - alpha (w_j +
// beta s_j)), from the previous iteration's vectors.
__device__ __forceinline__ double cg1_u(const double *__restrict__ dinv,
                                        const double *__restrict__ r,
                                        const double *__restrict__ w,
                                        const double *__restrict__ s, int32_t j, double alpha,
                                        double beta) {
  const double sj = w[j] + beta * s[j];
  const double rj = r[j] - alpha * sj;
  return dinv[j] * rj;
}

// (A u)_i for u given row-wise by the functor U (difference form as apply64)
template <int FMT, class U>
__device__ __forceinline__ double apply64_fn(const MatView &m, int32_t r, double ur, const U &u) {
  double acc = 0.0;
  if constexpr (FMT == FMT_DIA) {
#pragma unroll
    for (int d = 0; d < kMaxDiag; ++d) {
      if (d < m.ndiag) {
        int32_t c = r + m.off[d];
        c = c < 0 ? 0 : (c >= m.n ? m.n - 1 : c);
        acc += m.val[(int64_t)d * m.n + r] * (u(c) - ur);
      }
    }
  } else {
    const int32_t sl = r >> 6, l = r & 63;
    const int32_t o = m.slice_off[sl];
    const int32_t wd = (m.slice_off[sl + 1] - o) >> 6;
    const double *vp = m.val + o + l;
    if (m.s_dcol) {
      const int16_t *cp = m.s_dcol + o + l;
      for (int32_t k = 0; k < wd; ++k) acc += vp[64 * k] * (u(ell_col<false>(cp, k, r)) - ur);
    } else {
      const int32_t *cp = m.s_col + o + l;
      for (int32_t k = 0; k < wd; ++k) acc += vp[64 * k] * (u(cp[64 * k]) - ur);
    }
  }
  return acc + m.rowsum[r] * ur;
}

// setup 1: r = b - A x into r[1]; p = 0, s[1] = 0; partials b.b
template <int FMT, int SB>
__global__ __launch_bounds__(SB) void k_cg1_init_r(Cg1Args a, const double *__restrict__ b) {
  __shared__ double sh[4 * (SB / 64)];
  const int32_t r0 = sweep_chunk(blockIdx.x, a.sw.ngroups) * a.sw.rows_per_group;
  const int32_t r1 = min(r0 + a.sw.rows_per_group, a.sw.n);
  double bb = 0.0;
  for (int32_t i = r0 + threadIdx.x; i < r1; i += SB) {
    const double bi = b[i];
    a.r[1][i] = bi - apply64<FMT>(a.m, a.x, i, a.x[i]);
    a.p[i] = 0.0;
    a.s[1][i] = 0.0;
    bb += bi * bi;
  }
  double v[1] = {bb};
  const int sl[1] = {P_BB};
  store_partials_k<SB, 1>(a.part, v, sl, sh);
}

// setup 2: w = A D^-1 r into w[1]; partials r.r, r.u, w.u (parity 1)
template <int FMT, int SB>
__global__ __launch_bounds__(SB) void k_cg1_init_w(Cg1Args a) {
  __shared__ double sh[4 * (SB / 64)];
  const int32_t r0 = sweep_chunk(blockIdx.x, a.sw.ngroups) * a.sw.rows_per_group;
  const int32_t r1 = min(r0 + a.sw.rows_per_group, a.sw.n);
  const double *__restrict__ rv = a.r[1];
  const double *__restrict__ dinv = a.dinv;
  double rr = 0.0, gam = 0.0, dlt = 0.0;
  for (int32_t i = r0 + threadIdx.x; i < r1; i += SB) {
    const double ri = rv[i];
    const double ui = dinv[i] * ri;
    const double wi =
        apply64_fn<FMT>(a.m, i, ui, [&](int32_t j) { return dinv[j] * rv[j]; });
    a.w[1][i] = wi;
    rr += ri * ri;
    gam += ri * ui;
    dlt += wi * ui;
  }
  double v[3] = {rr, gam, dlt};
  const int sl[3] = {C_RR + 1, C_G + 1, C_D + 1};
  store_partials_k<SB, 3>(a.part, v, sl, sh);
}

// one workgroup: Eigen's early exits
template <int SB>
__global__ __launch_bounds__(SB) void k_cg1_init_finish(Cg1Args a) {
  __shared__ double sh[4 * (SB / 64)];
  double t[2];
  const int sl[2] = {P_BB, C_RR + 1};
  reduce_partials_k<SB, 2>(a.part, a.sw.ngroups, sl, t, sh);
  const double bb = t[0], rr = t[1];
  if (threadIdx.x == 0) {
    PcgState *st = a.st;
    st->bb = bb;
    st->rr = rr;
    st->iter = -1;
    st->iterp[0] = st->iterp[1] = -1;
    st->converged = 0;
    st->done = 0;
    st->refinements = 0;
    const double thr = st->thr * bb;  // thr holds tol^2 on entry
    st->thr = thr > 2.2250738585072014e-308 ? thr : 2.2250738585072014e-308;
    if (bb == 0.0 || rr < st->thr) {
      st->done = 1;
      st->converged = 1;
      st->iter = 0;
    } else if (st->maxit <= 0) {
      st->done = 1;
      st->iter = 0;
    }
  }
}

// iteration (par = iteration & 1): reads the parity par^1 vectors and
// partials, writes parity par.  The stopping test and the iteration count
// follow k_cg_spmv: a converged residual leaves the count as it is, and the
// maxit-th update is not applied.
template <int FMT, int SB>
__global__ __launch_bounds__(SB) void k_cg1(Cg1Args a, int par) {
  __shared__ double sh[4 * (SB / 64)];
  PcgState *st = a.st;
  if (st->done) return;
  double t[3];
  const int sl_in[3] = {C_RR + (par ^ 1), C_G + (par ^ 1), C_D + (par ^ 1)};
  reduce_partials_k<SB, 3>(a.part, a.sw.ngroups, sl_in, t, sh);
  const double rr = t[0], gam = t[1], dlt = t[2];
  const bool writer = blockIdx.x == 0 && threadIdx.x == 0;
  const int itp = st->iterp[par ^ 1];
  if (rr < st->thr) {
    if (writer) {
      st->rr = rr;
      st->done = 1;
      st->converged = 1;
      st->iter = itp < 0 ? 0 : itp;
    }
    return;
  }
  const int it = itp + 1;
  if (it >= st->maxit) {
    if (writer) {
      st->rr = rr;
      st->iter = it;
      st->done = 1;
    }
    return;
  }
  const double beta = it == 0 ? 0.0 : gam / st->rho[par ^ 1];
  const double alpha = it == 0 ? gam / dlt : gam / (dlt - beta * gam / st->alpha[par ^ 1]);
  if (writer) {
    st->rr = rr;
    st->iter = it;
    st->iterp[par] = it;
    st->rho[par] = gam;
    st->alpha[par] = alpha;
  }
  const double *__restrict__ ro = a.r[par ^ 1];
  const double *__restrict__ wo = a.w[par ^ 1];
  const double *__restrict__ so = a.s[par ^ 1];
  double *__restrict__ rn = a.r[par];
  double *__restrict__ wn = a.w[par];
  double *__restrict__ sn = a.s[par];
  const double *__restrict__ dinv = a.dinv;
  const int32_t r0 = sweep_chunk(blockIdx.x, a.sw.ngroups) * a.sw.rows_per_group;
  const int32_t r1 = min(r0 + a.sw.rows_per_group, a.sw.n);
  double rr2 = 0.0, gam2 = 0.0, dlt2 = 0.0;
  for (int32_t i = r0 + threadIdx.x; i < r1; i += SB) {
    const double di = dinv[i];
    const double ri = ro[i];
    const double pi = di * ri + beta * a.p[i];
    const double si = wo[i] + beta * so[i];
    const double rni = ri - alpha * si;
    const double ui = di * rni;
    const double wi = apply64_fn<FMT>(a.m, i, ui, [&](int32_t j) {
      return cg1_u(dinv, ro, wo, so, j, alpha, beta);
    });
    a.x[i] += alpha * pi;
    a.p[i] = pi;
    sn[i] = si;
    rn[i] = rni;
    wn[i] = wi;
    rr2 += rni * rni;
    gam2 += rni * ui;
    dlt2 += wi * ui;
  }
  double v[3] = {rr2, gam2, dlt2};
  const int sl[3] = {C_RR + par, C_G + par, C_D + par};
  store_partials_k<SB, 3>(a.part, v, sl, sh);
}

template <int FMT>
void launch_cg1_chunk(hipStream_t s, const Sweep &sw, const Cg1Args &a, int count) {
  const dim3 G(sw.ngroups), B(kSB64);
  for (int k = 0; k < count; ++k) {
    tmark_begin(s, 0);
    hipLaunchKernelGGL((k_cg1<FMT, kSB64>), G, B, 0, s, a, k & 1);
    tmark_end(s);
  }
}

// ---------------------------------------------------------------------------
// Mixed precision: FP32 matrix, vectors and recurrences (FP64 dot products
// and FP64 row sums in the product), solution kept as a double-double pair
// (x, xlo) so that its true residual is not limited by rounding x to FP64.
// Reliable updates recompute the true residual b - A x - A xlo in FP64 once
// the recurrence residual has dropped by delta (or after a caller-set number
// of iterations) and restart the FP32 recurrence from it.
// ---------------------------------------------------------------------------
// The FP32-Krylov mode preconditions with D^-1 rounded to bfloat16 (8
// significant bits, FP32's exponent range, so no overflow or underflow for
// any diagonal FP32 can hold).  Jacobi-PCG converges with any SPD diagonal;
// a relative perturbation of at most 2^-9 changed the iteration count of the
// 4.2M-DOF P1 system by 0.01% and saves 4 of the two sweeps' bytes per row.
// The FP64 residual of the reliable updates, which decides convergence,
// never sees it.
__device__ __forceinline__ float dinv_bf16(const __hip_bfloat16 *__restrict__ d, int32_t i) {
  return __bfloat162float(d[i]);
}

// D^-1_i in the recurrence precision: bfloat16 for FP32 Krylov vectors
// (VT = float); FP32 for the FP32-matrix mode (VT = double), whose FP64
// recurrence measured 8% more iterations on the 2.4M-DOF P3 system with the
// bfloat16 diagonal
template <class Args>
__device__ __forceinline__ auto mix_dinv(const Args &a, int32_t i) {
  using VT = std::remove_reference_t<decltype(*a.p)>;
  if constexpr (std::is_same_v<VT, float>)
    return dinv_bf16(a.dinv16, i);
  else
    return (double)a.dinv32[i];
}

template <class VT>
struct MixArgs {
  MatView m;
  const float *__restrict__ dinv32;
  const __hip_bfloat16 *__restrict__ dinv16;
  const double *__restrict__ dinv;
  double *__restrict__ x;    // high part of the solution (the caller's vector)
  double *__restrict__ xlo;  // low part
  const double *__restrict__ b;
  double *__restrict__ r64;  // FP64 true residual
  VT *__restrict__ r;
  VT *__restrict__ p;
  VT *__restrict__ q;
  VT *__restrict__ xa;
  double *part;
  PcgState *st;
  Sweep sw;
  double delta2;
  int max_interval;
  double *rpart;  // row sweeps: slot k of group g at rpart[k * nwg + g]
  int nwg;
};

template <class VT>
__global__ __launch_bounds__(256) void k_mix_setup(MixArgs<VT> a) {
  const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.sw.n) return;
  const VT ri = (VT)a.r64[i];
  a.r[i] = ri;
  a.p[i] = mix_dinv(a, i) * ri;
  a.xa[i] = (VT)0;
  a.xlo[i] = 0.0;
}

// (A x)_r from FP32 storage for either vector precision
template <int FMT, bool HOIST = false>
__device__ __forceinline__ float apply_mixed(const MatView &m, const float *x, int32_t r, float xr) {
  return apply32<FMT, HOIST>(m, x, r, xr);
}
template <int FMT, bool HOIST = false>
__device__ __forceinline__ double apply_mixed(const MatView &m, const double *__restrict__ x,
                                              int32_t r, double xr) {
  double acc = 0.0;
  if constexpr (FMT == FMT_DIA && HOIST) {
    float v[kMaxDiag];
    double xc[kMaxDiag];
#pragma unroll
    for (int d = 0; d < kMaxDiag; ++d) {
      if (d < m.ndiag) {
        int32_t c = r + m.off[d];
        c = c < 0 ? 0 : (c >= m.n ? m.n - 1 : c);
        v[d] = m.val32[(int64_t)d * m.n + r];
        xc[d] = x[c];
      }
    }
#pragma unroll
    for (int d = 0; d < kMaxDiag; ++d)
      if (d < m.ndiag) acc += (double)v[d] * (xc[d] - xr);
  } else if constexpr (FMT == FMT_DIA) {
#pragma unroll
    for (int d = 0; d < kMaxDiag; ++d) {
      if (d < m.ndiag) {
        int32_t c = r + m.off[d];
        c = c < 0 ? 0 : (c >= m.n ? m.n - 1 : c);
        acc += (double)m.val32[(int64_t)d * m.n + r] * (x[c] - xr);
      }
    }
  } else {
    acc = ell_row(m, m.val32, r, xr, [&](int32_t c) { return x[c]; });
  }
  return acc + (double)m.rowsum32[r] * xr;
}

// Two sweeps per iteration, as in the FP64 solver.  Sweep A decides, from
// the previous sweep's partials (identically in every workgroup), whether
// this slot is a CG step or a reliable update:
//   step:    q = A32 p, partials p.q, (M q, r), (M q, q)          (M = D^-1 in BF16)
//   update:  r64 = b - A x - A xa - A xlo in FP64 (x, xa, xlo as left by the
//            previous slot), partial ||r64||^2
// Sweep B completes the slot:
//   step:    alpha = rho / p.q, beta from rho' = rho - 2 alpha (Mq, r) +
//            alpha^2 (Mq, q); xa += alpha p, r -= alpha q, p = M r + beta p
//   update:  fold xa into the double-double (x, xlo) with TwoSum; stop on the
//            true residual, else replace the recurrence residual r by r64 and
//            the M r part of the search direction with M r64:
//            p += M (r64 - r), which is the direction the last step would
//            have produced from the true residual
// A reliable update is taken when the recurrence residual has fallen by
// delta since the last one, reaches rtol, after max_interval steps, or at
// maxit.  iterp[par] counts the CG steps up to and including slot par.
constexpr int X_PQ = 0, X_QDR = 1, X_QDQ = 2, X_RT = 0, X_RR = 3, X_RZ = 5;

template <int FMT, class VT, int SB>
__global__ __launch_bounds__(SB) void k_mx_a(MixArgs<VT> a, int par) {
  __shared__ double sh[4 * (SB / 64)];
  PcgState *st = a.st;
  // done: 0 running, 1 stopped, 2 stop decided by the previous sweep; it
  // takes effect here, after every workgroup of that sweep has folded its
  // rows (all workgroups return either way)
  const int done = st->done;
  if (done == 1) return;
  if (done == 2) {
    if (blockIdx.x == 0 && threadIdx.x == 0) st->done = 1;
    return;
  }
  double t[1];
  const int sl_in[1] = {X_RR + (par ^ 1)};
  reduce_partials_k<SB, 1>(a.part, a.sw.ngroups, sl_in, t, sh);
  const double rr = t[0];
  const int itp = st->iterp[par ^ 1];
  const bool refine = rr < st->thr || rr < a.delta2 * st->rr_ref ||
                      itp - st->last_refine >= a.max_interval || itp >= st->maxit;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    st->refine[par] = refine ? 1 : 0;
    st->iterp[par] = refine ? itp : itp + 1;
    if (!refine) st->iter = itp + 1;
  }
  const int32_t r0 = sweep_chunk(blockIdx.x, a.sw.ngroups) * a.sw.rows_per_group;
  const int32_t r1 = min(r0 + a.sw.rows_per_group, a.sw.n);
  if (!refine) {
    double pq = 0.0, qdr = 0.0, qdq = 0.0;
    for (int32_t i = r0 + threadIdx.x; i < r1; i += SB) {
      const VT pi = a.p[i], mi = mix_dinv(a, i), ri = a.r[i];
      const VT qi = apply_mixed<FMT, true>(a.m, a.p, i, pi);
      a.q[i] = qi;
      const double dq = (double)(mi * qi);
      pq += (double)pi * (double)qi;
      qdr += dq * (double)ri;
      qdq += dq * (double)qi;
    }
    double v[3] = {pq, qdr, qdq};
    const int sl[3] = {X_PQ, X_QDR, X_QDQ};
    store_partials_k<SB, 3>(a.part, v, sl, sh);
  } else {
    double rt = 0.0;
    const VT *__restrict__ xa = a.xa;
    for (int32_t i = r0 + threadIdx.x; i < r1; i += SB) {
      const double ah = apply64<FMT>(a.m, a.x, i, a.x[i]);
      const double aa =
          apply64_fn<FMT>(a.m, i, (double)xa[i], [&](int32_t j) { return (double)xa[j]; });
      const double al = apply64<FMT>(a.m, a.xlo, i, a.xlo[i]);
      const double ri = ((a.b[i] - ah) - aa) - al;
      a.r64[i] = ri;
      rt += ri * ri;
    }
    double v[1] = {rt};
    const int sl[1] = {X_RT};
    store_partials_k<SB, 1>(a.part, v, sl, sh);
  }
}

template <class VT, int SB>
__global__ __launch_bounds__(SB) void k_mx_b(MixArgs<VT> a, int par) {
  __shared__ double sh[4 * (SB / 64)];
  PcgState *st = a.st;
  if (st->done == 1) return;  // 2 is set by this sweep's own writer
  const int32_t r0 = sweep_chunk(blockIdx.x, a.sw.ngroups) * a.sw.rows_per_group;
  const int32_t r1 = min(r0 + a.sw.rows_per_group, a.sw.n);
  double rr = 0.0, rz = 0.0;
  if (!st->refine[par]) {
    double t[4];
    const int sl_in[4] = {X_RZ + (par ^ 1), X_PQ, X_QDR, X_QDQ};
    reduce_partials_k<SB, 4>(a.part, a.sw.ngroups, sl_in, t, sh);
    const double rho = t[0], pq = t[1], qdr = t[2], qdq = t[3];
    const double alpha = rho / pq;
    const double rho_next = rho - 2.0 * alpha * qdr + alpha * alpha * qdq;
    const VT fa = (VT)alpha, fb = (VT)(rho_next / rho);
    for (int32_t i = r0 + threadIdx.x; i < r1; i += SB) {
      // plain loads here: the non-temporal form measured 2% slower for
      // this chunked sweep (faster in the batched and row sweeps)
      const VT pi = a.p[i], xai = a.xa[i], r0 = a.r[i], qi = a.q[i];
      const VT mi = mix_dinv(a, i);
      const VT ri = r0 - fa * qi;
      const VT zi = mi * ri;
      a.xa[i] = xai + fa * pi;
      a.r[i] = ri;
      a.p[i] = zi + fb * pi;
      rr += (double)ri * (double)ri;
      rz += (double)ri * (double)zi;
    }
  } else {
    double t[1];
    const int sl_in[1] = {X_RT};
    reduce_partials_k<SB, 1>(a.part, a.sw.ngroups, sl_in, t, sh);
    const double rt = t[0];
    const bool conv = rt < st->thr;
    const int its = st->iterp[par];
    const bool stop = conv || its >= st->maxit;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
      st->rr = rt;
      st->refinements += 1;
      st->last_refine = its;
      if (stop) {
        // 2, not 1: the other workgroups of this sweep return on entry only
        // for 1, so they still fold their rows
        st->done = 2;
        st->converged = conv ? 1 : 0;
      } else {
        st->rr_ref = rt;
      }
    }
    for (int32_t i = r0 + threadIdx.x; i < r1; i += SB) {
      const double hi = a.x[i], d = (double)a.xa[i];
      const double sm = hi + d;
      const double bp = sm - hi;
      const double err = (hi - (sm - bp)) + (d - bp);
      a.x[i] = sm;
      a.xlo[i] += err;
      a.xa[i] = (VT)0;
      if (!stop) {
        const VT ri = (VT)a.r64[i];
        const VT mi = mix_dinv(a, i);
        a.p[i] += mi * (ri - a.r[i]);
        a.r[i] = ri;
        const VT zi = mi * ri;
        rr += (double)ri * (double)ri;
        rz += (double)ri * (double)zi;
      }
    }
    if (stop) return;
  }
  double v[2] = {rr, rz};
  const int sl[2] = {X_RR + par, X_RZ + par};
  store_partials_k<SB, 2>(a.part, v, sl, sh);
}

template <class VT>
__global__ __launch_bounds__(256) void k_mix_final(MixArgs<VT> a) {
  const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < a.sw.n) a.x[i] += a.xlo[i];
}

// ---------------------------------------------------------------------------
// Mixed precision on sliced ELL: the slots of k_mx_a / k_mx_b as one-row-
// per-thread sweeps (see k_cgr_*), each followed by a one-workgroup kernel
// that sums the partials and takes the decisions: after sweep B it decides
// whether the next slot is a CG step or a reliable update; after sweep A it
// forms alpha and beta, or judges the true residual.  A stop found there
// takes effect after sweep B has folded the last increment into x.
// ---------------------------------------------------------------------------
template <class VT, int K>
__device__ __forceinline__ void mxr_partials(const MixArgs<VT> &a, double (&v)[K], double *sh) {
  block_sum_k<kRB, K>(v, sh);
  if (threadIdx.x == 0)
#pragma unroll
    for (int k = 0; k < K; ++k) a.rpart[(int64_t)k * a.nwg + blockIdx.x] = v[k];
}

template <class VT, int K>
__device__ __forceinline__ void mxr_totals(const MixArgs<VT> &a, double (&t)[K], double *sh) {
  sum_partials<K>(a.rpart, a.nwg, t);
  block_sum_k<kRedB, K>(t, sh);
}

// r64 = b - A x, the recurrence started from it; partials b.b, r.r, r.D^-1 r
template <int FMT, class VT>
__global__ __launch_bounds__(kRB) void k_mxr_init(MixArgs<VT> a) {
  __shared__ double sh[4 * (kRB / 64)];
  const int32_t i = blockIdx.x * kRB + threadIdx.x;
  double v[3] = {0.0, 0.0, 0.0};
  if (i < a.m.n) {
    const double bi = a.b[i];
    const double ri = bi - apply64<FMT>(a.m, a.x, i, a.x[i]);
    a.r64[i] = ri;
    const VT rv = (VT)ri;
    a.r[i] = rv;
    a.p[i] = mix_dinv(a, i) * rv;
    a.xa[i] = (VT)0;
    a.xlo[i] = 0.0;
    v[0] = bi * bi;
    v[1] = ri * ri;
    v[2] = ri * (a.dinv[i] * ri);
  }
  mxr_partials<VT, 3>(a, v, sh);
}

template <class VT>
__global__ __launch_bounds__(kRedB) void k_mxr_red_init(MixArgs<VT> a) {
  __shared__ double sh[4 * (kRedB / 64)];
  double t[3];
  mxr_totals<VT, 3>(a, t, sh);
  if (threadIdx.x == 0) {
    PcgState *st = a.st;
    const double bb = t[0], rr = t[1];
    st->bb = bb;
    st->rr = rr;
    st->rr_ref = rr;
    st->rho[0] = t[2];
    st->iter = 0;
    st->iterp[0] = 0;
    st->refine[0] = 0;
    st->refinements = 0;
    st->last_refine = 0;
    st->stop_pending = 0;
    st->conv_pending = 0;
    st->converged = 0;
    st->done = 0;
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

template <int FMT, class VT>
__global__ __launch_bounds__(kRB) void k_mxr_a(MixArgs<VT> a) {
  __shared__ double sh[4 * (kRB / 64)];
  const PcgState *st = a.st;
  if (st->done) return;
  const int32_t i = blockIdx.x * kRB + threadIdx.x;
  if (!st->refine[0]) {
    double v[3] = {0.0, 0.0, 0.0};
    if (i < a.m.n) {
      const VT pi = a.p[i];
      const VT qi = apply_mixed<FMT>(a.m, a.p, i, pi);
      a.q[i] = qi;
      const double dq = (double)(mix_dinv(a, i) * qi);
      v[0] = (double)pi * (double)qi;
      v[1] = dq * (double)a.r[i];
      v[2] = dq * (double)qi;
    }
    mxr_partials<VT, 3>(a, v, sh);
  } else {
    double v[1] = {0.0};
    if (i < a.m.n) {
      const VT *__restrict__ xa = a.xa;
      const double ah = apply64<FMT>(a.m, a.x, i, a.x[i]);
      const double aa =
          apply64_fn<FMT>(a.m, i, (double)xa[i], [&](int32_t j) { return (double)xa[j]; });
      const double al = apply64<FMT>(a.m, a.xlo, i, a.xlo[i]);
      const double ri = ((a.b[i] - ah) - aa) - al;
      a.r64[i] = ri;
      v[0] = ri * ri;
    }
    mxr_partials<VT, 1>(a, v, sh);
  }
}

template <class VT>
__global__ __launch_bounds__(kRedB) void k_mxr_red_a(MixArgs<VT> a) {
  __shared__ double sh[4 * (kRedB / 64)];
  PcgState *st = a.st;
  if (st->done) return;
  if (!st->refine[0]) {
    double t[3];
    mxr_totals<VT, 3>(a, t, sh);
    if (threadIdx.x == 0) {
      const double rho = st->rho[0], pq = t[0], qdr = t[1], qdq = t[2];
      const double alpha = rho / pq;
      const double rho_next = rho - 2.0 * alpha * qdr + alpha * alpha * qdq;
      st->alpha_s = alpha;
      st->beta_s = rho_next / rho;
      st->iterp[0] += 1;
      st->iter = st->iterp[0];
    }
  } else {
    double t[1];
    mxr_totals<VT, 1>(a, t, sh);
    if (threadIdx.x == 0) {
      const double rt = t[0];
      const bool conv = rt < st->thr;
      const int its = st->iterp[0];
      st->rr = rt;
      st->refinements += 1;
      st->last_refine = its;
      if (conv || its >= st->maxit) {
        st->stop_pending = 1;
        st->conv_pending = conv ? 1 : 0;
      } else {
        st->rr_ref = rt;
      }
    }
  }
}

template <class VT>
__global__ __launch_bounds__(kRB) void k_mxr_b(MixArgs<VT> a) {
  __shared__ double sh[4 * (kRB / 64)];
  const PcgState *st = a.st;
  if (st->done) return;
  const int32_t i = blockIdx.x * kRB + threadIdx.x;
  double v[2] = {0.0, 0.0};
  if (!st->refine[0]) {
    const VT fa = (VT)st->alpha_s, fb = (VT)st->beta_s;
    if (i < a.m.n) {
      const VT pi = a.p[i], r0 = a.r[i];
      const VT xai = __builtin_nontemporal_load(&a.xa[i]);
      const VT qi = __builtin_nontemporal_load(&a.q[i]);
      const VT mi = mix_dinv(a, i);
      const VT ri = r0 - fa * qi;
      const VT zi = mi * ri;
      __builtin_nontemporal_store(xai + fa * pi, &a.xa[i]);
      a.r[i] = ri;
      a.p[i] = zi + fb * pi;
      v[0] = (double)ri * (double)ri;
      v[1] = (double)ri * (double)zi;
    }
  } else {
    const bool stop = st->stop_pending != 0;
    if (i < a.m.n) {
      const double hi = a.x[i], d = (double)a.xa[i];
      const double sm = hi + d;
      const double bp = sm - hi;
      const double err = (hi - (sm - bp)) + (d - bp);
      a.x[i] = sm;
      a.xlo[i] += err;
      a.xa[i] = (VT)0;
      if (!stop) {
        const VT ri = (VT)a.r64[i];
        const VT mi = mix_dinv(a, i);
        a.p[i] += mi * (ri - a.r[i]);
        a.r[i] = ri;
        const VT zi = mi * ri;
        v[0] = (double)ri * (double)ri;
        v[1] = (double)ri * (double)zi;
      }
    }
  }
  mxr_partials<VT, 2>(a, v, sh);
}

// ends a slot: a pending stop, or (r, D^-1 r) for the next alpha and the
// step-or-update decision for the next slot
template <class VT>
__global__ __launch_bounds__(kRedB) void k_mxr_red_b(MixArgs<VT> a) {
  __shared__ double sh[4 * (kRedB / 64)];
  PcgState *st = a.st;
  if (st->done) return;
  if (st->stop_pending) {
    if (threadIdx.x == 0) {
      st->done = 1;
      st->converged = st->conv_pending;
    }
    return;
  }
  double t[2];
  mxr_totals<VT, 2>(a, t, sh);
  if (threadIdx.x == 0) {
    const double rr = t[0];
    st->rho[0] = t[1];
    const int itp = st->iterp[0];
    st->refine[0] = (rr < st->thr || rr < a.delta2 * st->rr_ref ||
                     itp - st->last_refine >= a.max_interval || itp >= st->maxit)
                        ? 1
                        : 0;
  }
}

template <int FMT, class VT>
void launch_mxr_chunk(hipStream_t s, const MixArgs<VT> &m, int count) {
  const dim3 G(m.nwg), B(kRB), B1(kRedB);
  for (int k = 0; k < count; ++k) {
    tmark_begin(s, 0);
    hipLaunchKernelGGL((k_mxr_a<FMT, VT>), G, B, 0, s, m);
    tmark_end(s);
    tmark_begin(s, 2);
    hipLaunchKernelGGL(k_mxr_red_a<VT>, dim3(1), B1, 0, s, m);
    tmark_end(s);
    tmark_begin(s, 1);
    hipLaunchKernelGGL(k_mxr_b<VT>, G, B, 0, s, m);
    tmark_end(s);
    tmark_begin(s, 2);
    hipLaunchKernelGGL(k_mxr_red_b<VT>, dim3(1), B1, 0, s, m);
    tmark_end(s);
  }
}

template <int FMT, int SB>
void launch_cg_chunk(hipStream_t s, const Sweep &sw, const CgArgs &a, int count, bool sym) {
  const dim3 G(sw.ngroups), B(SB);
  for (int k = 0; k < count; ++k) {
    const int par = k & 1;
    if (a.blk.n > 0) {
      // the dense block: t = B p before the product sweep adds it
      launch_blk_gemv(s, a.blk, a.p, a.st);
      tmark_begin(s, 0);
      if (FMT == FMT_DIA && sym)
        hipLaunchKernelGGL((k_cg_spmv<FMT, SB, true, true>), G, B, 0, s, a, par);
      else
        hipLaunchKernelGGL((k_cg_spmv<FMT, SB, false, true>), G, B, 0, s, a, par);
      tmark_end(s);
      tmark_begin(s, 1);
      hipLaunchKernelGGL(k_cg_update<SB>, G, B, 0, s, a, par);
      tmark_end(s);
      continue;
    }
    tmark_begin(s, 0);
    if (FMT == FMT_DIA && sym)
      hipLaunchKernelGGL((k_cg_spmv<FMT, SB, true>), G, B, 0, s, a, par);
    else
      hipLaunchKernelGGL((k_cg_spmv<FMT, SB>), G, B, 0, s, a, par);
    tmark_end(s);
    tmark_begin(s, 1);
    hipLaunchKernelGGL(k_cg_update<SB>, G, B, 0, s, a, par);
    tmark_end(s);
  }
}

// ---------------------------------------------------------------------------
// Captured chunks of FP64 iterations.  Below ~1M rows a sweep takes a few
// microseconds and the solve is bound by the host's launch rate (64x64 P1:
// ~15 us per iteration for two sweeps of ~3 us); a chunk of `count`
// iterations is therefore captured into a graph once and replayed.  A slot is
// reused when the captured argument blocks equal the next call's field by
// field (the solution vector is the only argument that changes between
// calls on one matrix); two slots serve callers that alternate between two
// solution vectors (the time stepper's stages).
// ---------------------------------------------------------------------------
enum { CG_CHUNK_DIA = 0, CG_CHUNK_ELL = 1, CG_ROWS_ELL = 2, CG_CHUNK_DIA_WIDE = 3 };

struct CgGraph {
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  int kind = -1, count = 0;
  bool sym = false;
  CgArgs ca{};
  RowArgs ra{};
};

bool same_view(const MatView &x, const MatView &y) {
  if (x.n != y.n || x.lo != y.lo || x.ndiag != y.ndiag || x.slice_off != y.slice_off ||
      x.s_col != y.s_col || x.s_dcol != y.s_dcol || x.val != y.val || x.val32 != y.val32 || x.rowsum != y.rowsum ||
      x.rowsum32 != y.rowsum32 || x.row_ptr != y.row_ptr || x.col != y.col)
    return false;
  for (int d = 0; d < kMaxDiag; ++d)
    if (x.off[d] != y.off[d]) return false;
  return true;
}

bool same_args(const CgArgs &x, const CgArgs &y) {
  return same_view(x.m, y.m) && x.dinv == y.dinv && x.x == y.x && x.r == y.r && x.p == y.p &&
         x.q == y.q && x.part == y.part && x.st == y.st && x.sw.n == y.sw.n &&
         x.sw.rows_per_group == y.sw.rows_per_group && x.sw.ngroups == y.sw.ngroups &&
         x.blk.B == y.blk.B && x.blk.srow == y.blk.srow && x.blk.smap == y.blk.smap &&
         x.blk.t == y.blk.t && x.blk.scale == y.blk.scale && x.blk.n == y.blk.n;
}

bool same_args(const RowArgs &x, const RowArgs &y) {
  return same_view(x.m, y.m) && x.dinv == y.dinv && x.x == y.x && x.r == y.r && x.p == y.p &&
         x.q == y.q && x.part == y.part && x.nwg == y.nwg && x.st == y.st;
}

void launch_fp64_chunk(hipStream_t s, int kind, const CgArgs &ca, const RowArgs &ra, int count,
                       bool sym) {
  if (kind == CG_CHUNK_DIA)
    launch_cg_chunk<FMT_DIA, kSB64>(s, ca.sw, ca, count, sym);
  else if (kind == CG_CHUNK_DIA_WIDE)
    launch_cg_chunk<FMT_DIA, kSBMix>(s, ca.sw, ca, count, sym);
  else if (kind == CG_CHUNK_ELL)
    launch_cg_chunk<FMT_ELL, kSBMix>(s, ca.sw, ca, count, false);
  else
    launch_cgr_chunk<FMT_ELL>(s, ra, count);
}

// One chunk of FP64 iterations, eagerly or from a (re)captured graph.
int run_fp64_chunk(lfgpu_solver_cache *c, hipStream_t s, int kind, const CgArgs &ca,
                   const RowArgs &ra, int count, bool sym, bool use_graph) {
  if (!use_graph || t_timer) {
    launch_fp64_chunk(s, kind, ca, ra, count, sym);
    LFG_HIP(hipGetLastError());
    return LFGPU_OK;
  }
  if (!c->cg_graphs) c->cg_graphs = new CgGraph[2];
  CgGraph *slots = static_cast<CgGraph *>(c->cg_graphs);
  CgGraph *g = nullptr;
  for (int k = 0; k < 2 && !g; ++k) {
    CgGraph &t = slots[k];
    if (t.exec && t.kind == kind && t.count == count && t.sym == sym &&
        (kind == CG_ROWS_ELL ? same_args(t.ra, ra) : same_args(t.ca, ca)))
      g = &t;
  }
  if (!g) {
    g = &slots[c->cg_graph_next];
    c->cg_graph_next ^= 1;
    if (g->exec) (void)hipGraphExecDestroy(g->exec);
    if (g->graph) (void)hipGraphDestroy(g->graph);
    *g = CgGraph{};
    LFG_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    launch_fp64_chunk(s, kind, ca, ra, count, sym);
    hipError_t e = hipStreamEndCapture(s, &g->graph);
    if (e != hipSuccess) {
      g->graph = nullptr;
      return hip_fail(e, "hipStreamEndCapture (PCG chunk)");
    }
    e = hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0);
    if (e != hipSuccess) {
      (void)hipGraphDestroy(g->graph);
      *g = CgGraph{};
      return hip_fail(e, "hipGraphInstantiate (PCG chunk)");
    }
    g->kind = kind;
    g->count = count;
    g->sym = sym;
    g->ca = ca;
    g->ra = ra;
  }
  LFG_HIP(hipGraphLaunch(g->exec, s));
  return LFGPU_OK;
}

template <int FMT, class VT>
void launch_mix_chunk(hipStream_t s, const Sweep &sw, const MixArgs<VT> &m, int count) {
  const dim3 G(sw.ngroups), B(kSBMix);
  for (int k = 0; k < count; ++k) {
    const int par = k & 1;
    tmark_begin(s, 0);
    hipLaunchKernelGGL((k_mx_a<FMT, VT, kSBMix>), G, B, 0, s, m, par);
    tmark_end(s);
    tmark_begin(s, 1);
    hipLaunchKernelGGL((k_mx_b<VT, kSBMix>), G, B, 0, s, m, par);
    tmark_end(s);
  }
}

// poll(launch): poll_chunks with the solve's state copies (done 1, or 2:
// every workgroup has folded its rows, both stop the polling)
template <class VT, class Poll>
int run_mixed(hipStream_t s, const lfgpu_solver_cache *c, const Sweep &sw, const MixArgs<VT> &m,
              int check, Poll &&poll) {
  const unsigned gflat = grid_for(sw.n, 256);
  hipLaunchKernelGGL(k_mix_setup<VT>, dim3(gflat), dim3(256), 0, s, m);
  LFG_HIP(hipGetLastError());
  LFG_TRY(poll([&]() -> int {
    if (c->fmt == FMT_DIA)
      launch_mix_chunk<FMT_DIA, VT>(s, sw, m, check);
    else
      launch_mix_chunk<FMT_ELL, VT>(s, sw, m, check);
    LFG_HIP(hipGetLastError());
    return LFGPU_OK;
  }));
  hipLaunchKernelGGL(k_mix_final<VT>, dim3(gflat), dim3(256), 0, s, m);
  LFG_HIP(hipGetLastError());
  return LFGPU_OK;
}

// mixed precision on sliced ELL (k_mxr_*): set-up, slots until done, x += xlo
template <class VT, class Poll>
int run_mixed_rows(hipStream_t s, const lfgpu_solver_cache *c, const MixArgs<VT> &m, int32_t n,
                   int check, Poll &&poll) {
  const unsigned gflat = grid_for(n, 256);
  hipLaunchKernelGGL((k_mxr_init<FMT_ELL, VT>), dim3(m.nwg), dim3(kRB), 0, s, m);
  hipLaunchKernelGGL(k_mxr_red_init<VT>, dim3(1), dim3(kRedB), 0, s, m);
  hipLaunchKernelGGL(k_zero_if_bb0, dim3(gflat), dim3(256), 0, s, m.x, n, c->st);
  LFG_HIP(hipGetLastError());
  LFG_TRY(poll([&]() -> int {
    launch_mxr_chunk<FMT_ELL, VT>(s, m, check);
    LFG_HIP(hipGetLastError());
    return LFGPU_OK;
  }));
  hipLaunchKernelGGL(k_mix_final<VT>, dim3(gflat), dim3(256), 0, s, m);
  LFG_HIP(hipGetLastError());
  return LFGPU_OK;
}

}  // namespace


// ===========================================================================
// Batched FP64 PCG: many independent systems stored as one block-diagonal
// matrix.  Workgroup g owns rows [r0, r1) of exactly one problem; every
// reduction is a segmented sum over that problem's contiguous range of
// workgroups, so each problem has its own alpha, beta, stopping test and
// iteration count (the same two sweeps per iteration as the single solve).
// ===========================================================================
namespace {

struct BGroup {
  int32_t prob, r0, r1, gb;  // problem, rows, first workgroup of the problem
};
struct BProb {
  int32_t gb, ge;  // workgroups [gb, ge)
};
struct BState {
  double thr, bb, rr;
  int iter, done, converged, maxit;
};

struct BArgs {
  MatView m;
  const double *__restrict__ dinv;
  double *__restrict__ x;
  double *__restrict__ r;
  double *__restrict__ p;
  double *__restrict__ q;
  double *part;  // [kPartSlots x ngroups]
  int ngroups;
  const BGroup *groups;
  const BProb *probs;
  BState *st;
  const int32_t *boff;  // FMT_BDIA: [nprob x kMaxDiag] offsets
  const int32_t *prow;  // FMT_BDIA: [nprob + 1] first rows
  const int *symflag;   // FMT_BDIA: 0 when the values are mirror-symmetric bit for bit
};

// The matrix view of a workgroup of problem prob: FMT_BDIA takes that
// problem's diagonal offsets and is then applied as DIA.
template <int FMT>
__device__ __forceinline__ MatView batch_view(const MatView &m0, const int32_t *boff,
                                              const int32_t *prow, int prob) {
  MatView m = m0;
  if constexpr (FMT == FMT_BDIA) {
#pragma unroll
    for (int d = 0; d < kMaxDiag; ++d) m.off[d] = boff[prob * kMaxDiag + d];
    m.lo = prow[prob];
  }
  return m;
}
template <int FMT>
constexpr int applied_fmt() {
  return FMT == FMT_BDIA ? FMT_DIA : FMT;
}

__device__ __forceinline__ double *bslot(const BArgs &a, int k) { return a.part + (int64_t)k * a.ngroups; }

__device__ __forceinline__ double seg_reduce(const double *part, int gb, int ge, double *sh) {
  double v = 0.0;
  for (int i = gb + threadIdx.x; i < ge; i += 256) v += part[i];
  v = block_sum<256>(v, sh);
  __syncthreads();
  if (threadIdx.x == 0) sh[4] = v;
  __syncthreads();
  return sh[4];
}

// K segmented sums in one pass (one pair of barriers); every thread gets the
// totals.  sh needs 4*K doubles.
template <int K>
__device__ __forceinline__ void seg_reduce_k(const double *const (&slots)[K], int gb, int ge,
                                             double (&out)[K], double *sh) {
#pragma unroll
  for (int k = 0; k < K; ++k) out[k] = 0.0;
  for (int i = gb + threadIdx.x; i < ge; i += 256)
#pragma unroll
    for (int k = 0; k < K; ++k) out[k] += slots[k][i];
  block_sum_k<256, K>(out, sh);
}

// this workgroup's K partials
template <int K>
__device__ __forceinline__ void store_partials_seg(double *const (&slots)[K], double (&v)[K],
                                                   double *sh) {
  block_sum_k<256, K>(v, sh);
  if (threadIdx.x == 0)
#pragma unroll
    for (int k = 0; k < K; ++k) slots[k][blockIdx.x] = v[k];
}

template <int FMT>
__global__ __launch_bounds__(256) void k_bcg_init(BArgs a, const double *__restrict__ b) {
  __shared__ double sh[16];
  const BGroup G = a.groups[blockIdx.x];
  const MatView m = batch_view<FMT>(a.m, a.boff, a.prow, G.prob);
  double bb = 0.0, rr = 0.0, rz = 0.0;
  for (int32_t i = G.r0 + threadIdx.x; i < G.r1; i += 256) {
    const double xi = a.x[i];
    const double bi = b[i];
    const double ri = bi - apply64<applied_fmt<FMT>()>(m, a.x, i, xi);
    a.r[i] = ri;
    const double zi = a.dinv[i] * ri;
    a.p[i] = zi;
    bb += bi * bi;
    rr += ri * ri;
    rz += ri * zi;
  }
  double v[3] = {bb, rr, rz};
  double *const sl[3] = {bslot(a, P_BB), bslot(a, P_RR + 1), bslot(a, P_RZ + 1)};
  store_partials_seg<3>(sl, v, sh);
}

// one workgroup per problem
__global__ __launch_bounds__(256) void k_bcg_init_finish(BArgs a) {
  __shared__ double sh[16];
  const BProb P = a.probs[blockIdx.x];
  double t[2];
  const double *const sl[2] = {bslot(a, P_BB), bslot(a, P_RR + 1)};
  seg_reduce_k<2>(sl, P.gb, P.ge, t, sh);
  const double bb = t[0], rr = t[1];
  if (threadIdx.x == 0) {
    BState *st = a.st + blockIdx.x;
    st->bb = bb;
    st->rr = rr;
    st->iter = -1;
    st->converged = 0;
    st->done = 0;
    const double thr = st->thr * bb;  // thr holds tol^2 on entry
    st->thr = thr > 2.2250738585072014e-308 ? thr : 2.2250738585072014e-308;
    if (bb == 0.0 || rr < st->thr) {
      st->done = 1;
      st->converged = 1;
      st->iter = 0;
    } else if (st->maxit <= 0) {
      st->done = 1;
      st->iter = 0;
    }
  }
}

// every problem's state before a solve (thr holds tol^2 until the set-up
// scales it by ||b||^2); set on the device, so a captured solve needs no copy
__global__ void k_bstate_init(BState *st, int nprob, double tol2, int maxit) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= nprob) return;
  BState v{};
  v.thr = tol2;
  v.maxit = maxit;
  st[p] = v;
}

__global__ void k_bcg_zero_if_bb0(BArgs a) {
  const BGroup G = a.groups[blockIdx.x];
  if (a.st[G.prob].bb != 0.0) return;
  for (int32_t i = G.r0 + threadIdx.x; i < G.r1; i += 256) a.x[i] = 0.0;
}

// SYMMODE 0: every diagonal read; 1: upper diagonals only (values
// mirror-symmetric bit for bit); 2: chosen on the device from *a.symflag,
// for chunks captured before the values they run on are known
template <int FMT, int SYMMODE = 0>
__global__ __launch_bounds__(256) void k_bcg_spmv(BArgs a, int par) {
  __shared__ double sh[16];
  const BGroup G = a.groups[blockIdx.x];
  BState *st = a.st + G.prob;
  if (st->done) return;
  const BProb P = a.probs[G.prob];
  const bool writer = blockIdx.x == (unsigned)G.gb && threadIdx.x == 0;
  const double rr = seg_reduce(bslot(a, P_RR + (par ^ 1)), P.gb, P.ge, sh);
  if (rr < st->thr) {
    if (writer) {
      st->rr = rr;
      st->done = 1;
      st->converged = 1;
      if (st->iter < 0) st->iter = 0;
    }
    return;
  }
  if (writer) {
    st->rr = rr;
    st->iter += 1;
    if (st->iter >= st->maxit) st->done = 1;
  }
  const MatView m = batch_view<FMT>(a.m, a.boff, a.prow, G.prob);
  double pq = 0.0, qdr = 0.0, qdq = 0.0;
  auto rows = [&](auto sym) {
    for (int32_t i = G.r0 + threadIdx.x; i < G.r1; i += 256) {
      const double pi = a.p[i], di = a.dinv[i], ri = a.r[i];
      const double qi =
          apply64<applied_fmt<FMT>(), true, decltype(sym)::value, true>(m, a.p, i, pi);
      __builtin_nontemporal_store(qi, &a.q[i]);  // as in k_cg_spmv
      const double dq = di * qi;
      pq += pi * qi;
      qdr += dq * ri;
      qdq += dq * qi;
    }
  };
  if constexpr (SYMMODE == 2) {
    if (*a.symflag == 0)
      rows(std::true_type{});
    else
      rows(std::false_type{});
  } else {
    rows(std::integral_constant<bool, SYMMODE == 1>{});
  }
  double v[3] = {pq, qdr, qdq};
  double *const sl[3] = {bslot(a, P_PQ), bslot(a, P_QDR), bslot(a, P_QDQ)};
  store_partials_seg<3>(sl, v, sh);
}

__global__ __launch_bounds__(256) void k_bcg_update(BArgs a, int par) {
  __shared__ double sh[16];
  const BGroup G = a.groups[blockIdx.x];
  const BState *st = a.st + G.prob;
  if (st->done) return;
  const BProb P = a.probs[G.prob];
  double t[4];
  const double *const sl[4] = {bslot(a, P_RZ + (par ^ 1)), bslot(a, P_PQ), bslot(a, P_QDR),
                               bslot(a, P_QDQ)};
  seg_reduce_k<4>(sl, P.gb, P.ge, t, sh);
  const double rho = t[0], pq = t[1], qdr = t[2], qdq = t[3];
  const double alpha = rho / pq;
  const double rho_next = rho - 2.0 * alpha * qdr + alpha * alpha * qdq;
  const double beta = rho_next / rho;
  double rr = 0.0, rz = 0.0;
  for (int32_t i = G.r0 + threadIdx.x; i < G.r1; i += 256) {
    const double pi = a.p[i], r0 = a.r[i], di = a.dinv[i];
    const double xi = __builtin_nontemporal_load(&a.x[i]);
    const double qi = __builtin_nontemporal_load(&a.q[i]);
    const double ri = r0 - alpha * qi;
    const double zi = di * ri;
    __builtin_nontemporal_store(xi + alpha * pi, &a.x[i]);
    a.r[i] = ri;
    a.p[i] = zi + beta * pi;
    rr += ri * ri;
    rz += ri * zi;
  }
  double v[2] = {rr, rz};
  double *const so[2] = {bslot(a, P_RR + par), bslot(a, P_RZ + par)};
  store_partials_seg<2>(so, v, sh);
}

template <int FMT>
void launch_bcg_chunk(hipStream_t s, const BArgs &a, int count, int symmode) {
  for (int k = 0; k < count; ++k) {
    const int par = k & 1;
    if constexpr (FMT == FMT_BDIA) {
      if (symmode == 1)
        hipLaunchKernelGGL((k_bcg_spmv<FMT, 1>), dim3(a.ngroups), dim3(256), 0, s, a, par);
      else if (symmode == 2)
        hipLaunchKernelGGL((k_bcg_spmv<FMT, 2>), dim3(a.ngroups), dim3(256), 0, s, a, par);
      else
        hipLaunchKernelGGL(k_bcg_spmv<FMT>, dim3(a.ngroups), dim3(256), 0, s, a, par);
    } else {
      hipLaunchKernelGGL(k_bcg_spmv<FMT>, dim3(a.ngroups), dim3(256), 0, s, a, par);
    }
    hipLaunchKernelGGL(k_bcg_update, dim3(a.ngroups), dim3(256), 0, s, a, par);
  }
}

// ---------------------------------------------------------------------------
// Batched mixed precision: a reliable-update scheme like the single solve's,
// run per problem on the workgroup table.  VT = float: FP32
// matrix, recurrence vectors and increment; VT = double: FP32 matrix only.
// FP64 true residuals, dot products and a double-double solution throughout.
// ---------------------------------------------------------------------------
struct BMState {
  double thr, bb, rr, rr_ref;
  double rho[2];
  int iter, done, converged, maxit;
  int iterp[2], refine[2];
  int last_refine, refinements;
};

// partial slots: the X_* slots of the single solve, plus b.b
constexpr int MB_BB = 7;

template <class VT>
struct BMArgs {
  MatView m;
  const float *__restrict__ dinv32;
  const __hip_bfloat16 *__restrict__ dinv16;
  const double *__restrict__ dinv;
  double *__restrict__ x;
  double *__restrict__ xlo;
  const double *__restrict__ b;
  double *__restrict__ r64;
  VT *__restrict__ r;
  VT *__restrict__ p;
  VT *__restrict__ q;
  VT *__restrict__ xa;
  double *part;  // [kPartSlots x ngroups]
  int ngroups;
  const BGroup *groups;
  const BProb *probs;
  BMState *st;
  double delta2;
  int max_interval;
  const int32_t *boff;  // FMT_BDIA: [nprob x kMaxDiag] offsets
  const int32_t *prow;  // FMT_BDIA: [nprob + 1] first rows
};

template <class VT>
__device__ __forceinline__ double *mslot(const BMArgs<VT> &a, int k) {
  return a.part + (int64_t)k * a.ngroups;
}

// r64 = b - A x in FP64, the FP32 recurrence started from it
template <int FMT, class VT>
__global__ __launch_bounds__(256) void k_bmix_init(BMArgs<VT> a) {
  __shared__ double sh[16];
  const BGroup G = a.groups[blockIdx.x];
  const MatView m = batch_view<FMT>(a.m, a.boff, a.prow, G.prob);
  double bb = 0.0, rr = 0.0, rz = 0.0;
  for (int32_t i = G.r0 + threadIdx.x; i < G.r1; i += 256) {
    const double bi = a.b[i];
    const double ri = bi - apply64<applied_fmt<FMT>()>(m, a.x, i, a.x[i]);
    a.r64[i] = ri;
    const VT rv = (VT)ri;
    a.r[i] = rv;
    a.p[i] = mix_dinv(a, i) * rv;
    a.xa[i] = (VT)0;
    a.xlo[i] = 0.0;
    bb += bi * bi;
    rr += ri * ri;
    rz += ri * (a.dinv[i] * ri);
  }
  double v[3] = {bb, rr, rz};
  double *const sl[3] = {mslot(a, MB_BB), mslot(a, X_RR + 1), mslot(a, X_RZ + 1)};
  store_partials_seg<3>(sl, v, sh);
}

// one workgroup per problem: Eigen's early exits as in k_cg_init_finish
template <class VT>
__global__ __launch_bounds__(256) void k_bmix_init_finish(BMArgs<VT> a) {
  __shared__ double sh[16];
  const BProb P = a.probs[blockIdx.x];
  double t[3];
  const double *const sl[3] = {mslot(a, MB_BB), mslot(a, X_RR + 1), mslot(a, X_RZ + 1)};
  seg_reduce_k<3>(sl, P.gb, P.ge, t, sh);
  const double bb = t[0], rr = t[1], rz = t[2];
  if (threadIdx.x == 0) {
    BMState *st = a.st + blockIdx.x;
    st->bb = bb;
    st->rr = rr;
    st->rr_ref = rr;
    st->rho[0] = st->rho[1] = rz;
    st->iter = -1;
    st->iterp[0] = st->iterp[1] = 0;
    st->refine[0] = st->refine[1] = 0;
    st->converged = 0;
    st->done = 0;
    st->refinements = 0;
    st->last_refine = 0;
    const double thr = st->thr * bb;  // thr holds tol^2 on entry
    st->thr = thr > 2.2250738585072014e-308 ? thr : 2.2250738585072014e-308;
    if (bb == 0.0 || rr < st->thr) {
      st->done = 1;
      st->converged = 1;
      st->iter = 0;
    } else if (st->maxit <= 0) {
      st->done = 1;
      st->iter = 0;
    }
  }
}

__global__ void k_bmstate_init(BMState *st, int nprob, double tol2, int maxit) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= nprob) return;
  BMState v{};
  v.thr = tol2;
  v.maxit = maxit;
  st[p] = v;
}

template <class VT>
__global__ void k_bmix_zero_if_bb0(BMArgs<VT> a) {
  const BGroup G = a.groups[blockIdx.x];
  if (a.st[G.prob].bb != 0.0) return;
  for (int32_t i = G.r0 + threadIdx.x; i < G.r1; i += 256) a.x[i] = 0.0;
}

// One iteration slot per problem in two sweeps, as k_mx_a / k_mx_b: every
// workgroup of a problem takes the same step-or-update decision from that
// problem's segmented partials.
template <int FMT, class VT>
__global__ __launch_bounds__(256) void k_bmx_a(BMArgs<VT> a, int par) {
  __shared__ double sh[16];
  const BGroup G = a.groups[blockIdx.x];
  BMState *st = a.st + G.prob;
  const int done = st->done;  // 0, 1 or 2 as in k_mx_a
  if (done == 1) return;
  const BProb P = a.probs[G.prob];
  const bool writer = blockIdx.x == (unsigned)G.gb && threadIdx.x == 0;
  if (done == 2) {
    if (writer) st->done = 1;
    return;
  }
  const double rr = seg_reduce(mslot(a, X_RR + (par ^ 1)), P.gb, P.ge, sh);
  const int itp = st->iterp[par ^ 1];
  const bool refine = rr < st->thr || rr < a.delta2 * st->rr_ref ||
                      itp - st->last_refine >= a.max_interval || itp >= st->maxit;
  if (writer) {
    st->refine[par] = refine ? 1 : 0;
    st->iterp[par] = refine ? itp : itp + 1;
    if (!refine) st->iter = itp + 1;
  }
  const MatView m = batch_view<FMT>(a.m, a.boff, a.prow, G.prob);
  constexpr int AF = applied_fmt<FMT>();
  if (!refine) {
    double pq = 0.0, qdr = 0.0, qdq = 0.0;
    for (int32_t i = G.r0 + threadIdx.x; i < G.r1; i += 256) {
      const VT pi = a.p[i], mi = mix_dinv(a, i), ri = a.r[i];
      const VT qi = apply_mixed<AF, true>(m, a.p, i, pi);
      a.q[i] = qi;
      const double dq = (double)(mi * qi);
      pq += (double)pi * (double)qi;
      qdr += dq * (double)ri;
      qdq += dq * (double)qi;
    }
    double v[3] = {pq, qdr, qdq};
    double *const sl[3] = {mslot(a, X_PQ), mslot(a, X_QDR), mslot(a, X_QDQ)};
    store_partials_seg<3>(sl, v, sh);
  } else {
    double rt = 0.0;
    const VT *__restrict__ xa = a.xa;
    for (int32_t i = G.r0 + threadIdx.x; i < G.r1; i += 256) {
      const double ah = apply64<AF>(m, a.x, i, a.x[i]);
      const double aa =
          apply64_fn<AF>(m, i, (double)xa[i], [&](int32_t j) { return (double)xa[j]; });
      const double al = apply64<AF>(m, a.xlo, i, a.xlo[i]);
      const double ri = ((a.b[i] - ah) - aa) - al;
      a.r64[i] = ri;
      rt += ri * ri;
    }
    store_partial(mslot(a, X_RT), rt, sh);
  }
}

template <class VT>
__global__ __launch_bounds__(256) void k_bmx_b(BMArgs<VT> a, int par) {
  __shared__ double sh[16];
  const BGroup G = a.groups[blockIdx.x];
  BMState *st = a.st + G.prob;
  if (st->done == 1) return;  // 2 is set by this sweep's own writer
  const BProb P = a.probs[G.prob];
  const bool writer = blockIdx.x == (unsigned)G.gb && threadIdx.x == 0;
  double rr = 0.0, rz = 0.0;
  if (!st->refine[par]) {
    double t[4];
    const double *const sl[4] = {mslot(a, X_RZ + (par ^ 1)), mslot(a, X_PQ), mslot(a, X_QDR),
                                 mslot(a, X_QDQ)};
    seg_reduce_k<4>(sl, P.gb, P.ge, t, sh);
    const double rho = t[0], pq = t[1], qdr = t[2], qdq = t[3];
    const double alpha = rho / pq;
    const double rho_next = rho - 2.0 * alpha * qdr + alpha * alpha * qdq;
    const VT fa = (VT)alpha, fb = (VT)(rho_next / rho);
    for (int32_t i = G.r0 + threadIdx.x; i < G.r1; i += 256) {
      const VT pi = a.p[i], r0 = a.r[i];
      const VT xai = __builtin_nontemporal_load(&a.xa[i]);
      const VT qi = __builtin_nontemporal_load(&a.q[i]);
      const VT mi = mix_dinv(a, i);
      const VT ri = r0 - fa * qi;
      const VT zi = mi * ri;
      __builtin_nontemporal_store(xai + fa * pi, &a.xa[i]);
      a.r[i] = ri;
      a.p[i] = zi + fb * pi;
      rr += (double)ri * (double)ri;
      rz += (double)ri * (double)zi;
    }
  } else {
    const double rt = seg_reduce(mslot(a, X_RT), P.gb, P.ge, sh);
    const bool conv = rt < st->thr;
    const int its = st->iterp[par];
    const bool stop = conv || its >= st->maxit;
    if (writer) {
      st->rr = rt;
      st->refinements += 1;
      st->last_refine = its;
      if (stop) {
        st->done = 2;  // as in k_mx_b
        st->converged = conv ? 1 : 0;
      } else {
        st->rr_ref = rt;
      }
    }
    for (int32_t i = G.r0 + threadIdx.x; i < G.r1; i += 256) {
      const double hi = a.x[i], d = (double)a.xa[i];
      const double sm = hi + d;
      const double bp = sm - hi;
      const double err = (hi - (sm - bp)) + (d - bp);
      a.x[i] = sm;
      a.xlo[i] += err;
      a.xa[i] = (VT)0;
      if (!stop) {
        const VT ri = (VT)a.r64[i];
        const VT mi = mix_dinv(a, i);
        a.p[i] += mi * (ri - a.r[i]);
        a.r[i] = ri;
        const VT zi = mi * ri;
        rr += (double)ri * (double)ri;
        rz += (double)ri * (double)zi;
      }
    }
    if (stop) return;
  }
  double v[2] = {rr, rz};
  double *const so[2] = {mslot(a, X_RR + par), mslot(a, X_RZ + par)};
  store_partials_seg<2>(so, v, sh);
}

template <class VT>
__global__ void k_bmix_final(BMArgs<VT> a) {
  const BGroup G = a.groups[blockIdx.x];
  for (int32_t i = G.r0 + threadIdx.x; i < G.r1; i += 256) a.x[i] += a.xlo[i];
}

template <int FMT, class VT>
void launch_bmix_chunk(hipStream_t s, const BMArgs<VT> &a, int count) {
  const dim3 G(a.ngroups), B(256);
  for (int k = 0; k < count; ++k) {
    const int par = k & 1;
    hipLaunchKernelGGL((k_bmx_a<FMT, VT>), G, B, 0, s, a, par);
    hipLaunchKernelGGL(k_bmx_b<VT>, G, B, 0, s, a, par);
  }
}

}  // namespace

struct lfgpu_batch {
  lfgpu_ctx *ctx = nullptr;
  lfgpu_csr *A = nullptr;
  int nprob = 0;
  std::vector<int32_t> row_off;  // [nprob + 1]
  int rows_per_group = 0;        // 0 until tuned
  int ngroups = 0;
  BGroup *groups = nullptr;
  BProb *probs = nullptr;
  BState *st = nullptr;
  double *part = nullptr;
  BState *h_st = nullptr;  // pinned, two slots of nprob states
  BMState *mst = nullptr;    // mixed-precision states (allocated on first use)
  BMState *h_mst = nullptr;  // pinned, two slots of nprob states
  // one captured chunk of iterations per precision mode and FP64 product
  // variant (batch_slot), so alternating modes replay their own graphs
  // instead of re-capturing; a slot is reused while every captured argument
  // is unchanged
  struct Slot {
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    int count = 0;
    const double *x = nullptr;
    const double *b = nullptr;
    double delta2 = 0.0;
    int max_interval = 0;
    int sym = 0;  // symmode of the captured FP64 product (k_bcg_spmv)
  } slot[5];
  int *h_symflag = nullptr;  // pinned host copy of d_symflag (after h_st)
  // lfgpu_batch_assemble_pcg: assembly, value refill, state set-up and the
  // first chunk captured as one graph per mode, keyed on every argument
  struct FusedKey {
    const lfgpu_plan *plan = nullptr;
    const lfgpu_fe *fe = nullptr;
    lfgpu_coeff alpha{}, gamma{}, f{};
    bool has_f = false;
    const double *rhs = nullptr, *x = nullptr;
    double rtol = 0.0, delta2 = 0.0;
    int maxit = 0, check = 0, max_interval = 0, rows_per_group = 0;
    const float *bval32 = nullptr;
  };
  struct Fused {
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    FusedKey key;
  } fused[3];
  double tune_ms[4] = {0, 0, 0, 0};
  // storage used by the batched sweeps: the matrix's own mirror (DIA or
  // ELL) or per-problem DIA (FMT_BDIA); -1 until the first solve decides
  int fmt = -1;
  int bndiag = 0;
  int32_t *d_row_off = nullptr;  // [nprob + 1]
  int32_t *boff = nullptr;       // [nprob x kMaxDiag]
  double *bval = nullptr;        // [bndiag x n]
  float *bval32 = nullptr;       // [bndiag x n]
  uint64_t bbuilt = ~0ull;       // matrix value version of bval / bval32
  bool bbuilt32 = false;
  // every problem has the same mirror-symmetric offsets and its FP64 values
  // are symmetric bit for bit: the FP64 SpMV sweep reads upper diagonals only
  bool bsymoff = false, bsym = false;
  int *d_symflag = nullptr;  // device copy of the symmetry test (0: symmetric)
};

namespace {

// ---- per-problem DIA storage of a batch ----------------------------------
// problem of row r: the last p with row_off[p] <= r
__device__ __forceinline__ int batch_prob_of(const int32_t *row_off, int nprob, int32_t r) {
  int lo = 0, hi = nprob - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (row_off[mid] <= r) lo = mid; else hi = mid - 1;
  }
  return lo;
}

// For each problem: the ascending union of the off-diagonal offsets of (up
// to) 64 consecutive rows from its middle row on (count -1 if more than
// kMaxDiag; 0 for an empty problem), and its off-diagonal count.  Rows near
// a boundary miss some offsets; the union over a few rows recovers the
// interior stencil.
__global__ void k_bdia_mid(const int32_t *row_ptr, const int32_t *col, const int32_t *row_off,
                           int nprob, int32_t *out) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= nprob) return;
  int32_t *o = out + p * (kMaxDiag + 2);
  const int32_t a = row_off[p], b = row_off[p + 1];
  o[kMaxDiag + 1] = (row_ptr[b] - row_ptr[a]) - (b - a);
  if (b <= a) {
    o[0] = 0;
    return;
  }
  int32_t set[kMaxDiag];
  int k = 0;
  const int32_t rows = b - a, scan = rows < 64 ? rows : 64;
  for (int32_t t = 0; t < scan; ++t) {
    const int32_t r = a + (rows / 2 + t) % rows;
    for (int32_t q = row_ptr[r]; q < row_ptr[r + 1]; ++q) {
      if (col[q] == r) continue;
      const int32_t v = col[q] - r;
      int j = 0;
      while (j < k && set[j] < v) ++j;
      if (j < k && set[j] == v) continue;
      if (k == kMaxDiag) {
        o[0] = -1;
        return;
      }
      for (int i = k; i > j; --i) set[i] = set[i - 1];
      set[j] = v;
      ++k;
    }
  }
  for (int i = 0; i < k; ++i) o[1 + i] = set[i];
  o[0] = k;
}

// flag |= 1 if an off-diagonal entry's offset is missing from its
// problem's set
__global__ void k_bdia_check(const int32_t *row_ptr, const int32_t *col, int32_t n,
                             const int32_t *row_off, int nprob, const int32_t *boff, int nd,
                             int *flag) {
  const int32_t r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  const int p = batch_prob_of(row_off, nprob, r);
  const int32_t *off = boff + p * kMaxDiag;
  for (int32_t q = row_ptr[r]; q < row_ptr[r + 1]; ++q) {
    const int32_t c = col[q];
    if (c == r) continue;
    bool found = false;
    for (int d = 0; d < nd; ++d) found = found || off[d] == c - r;
    if (!found) atomicOr(flag, 1);
  }
}

// values of row r in its problem's diagonal slots (zeros elsewhere)
__global__ void k_bdia_fill(const int32_t *row_ptr, const int32_t *col, const double *val,
                            int32_t n, const int32_t *row_off, int nprob, const int32_t *boff,
                            int nd, double *bval, float *bval32) {
  const int32_t r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  const int p = batch_prob_of(row_off, nprob, r);
  const int32_t *off = boff + p * kMaxDiag;
  for (int d = 0; d < nd; ++d) {
    bval[(int64_t)d * n + r] = 0.0;
    if (bval32) bval32[(int64_t)d * n + r] = 0.0f;
  }
  for (int32_t q = row_ptr[r]; q < row_ptr[r + 1]; ++q) {
    const int32_t c = col[q];
    if (c == r) continue;
    int d = 0;
    while (d < nd - 1 && off[d] != c - r) ++d;
    bval[(int64_t)d * n + r] = val[q];
    if (bval32) bval32[(int64_t)d * n + r] = (float)val[q];
  }
}

// flag = 1 if a lower-diagonal value of row r differs, bit for bit, from the
// mirrored upper-diagonal value of row r + off (offsets of r's problem,
// ascending and mirror-symmetric)
__global__ void k_bdia_sym(const double *bval, int32_t n, const int32_t *row_off, int nprob,
                           const int32_t *boff, int nd, int *flag) {
  const int32_t r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  const int p = batch_prob_of(row_off, nprob, r);
  const int32_t *off = boff + p * kMaxDiag;
  bool bad = false;
  for (int d = 0; 2 * d < nd; ++d) {
    const int32_t c = r + off[d];
    const double a = bval[(int64_t)d * n + r];
    const double b = c >= row_off[p] ? bval[(int64_t)(nd - 1 - d) * n + c] : 0.0;
    bad = bad || __double_as_longlong(a) != __double_as_longlong(b);
  }
  if (bad) atomicOr(flag, 1);
}

// flag = 1 if a row has an entry outside its own problem's rows (the
// matrix is not block diagonal over the row offsets)
__global__ void k_batch_blockcheck(const int32_t *row_ptr, const int32_t *col, int32_t n,
                                   const int32_t *row_off, int nprob, int *flag) {
  const int32_t r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  const int p = batch_prob_of(row_off, nprob, r);
  const int32_t lo = row_off[p], hi = row_off[p + 1];
  bool bad = false;
  for (int32_t k = row_ptr[r]; k < row_ptr[r + 1]; ++k) bad = bad || col[k] < lo || col[k] >= hi;
  if (bad) atomicOr(flag, 1);
}

// Chooses the batch's storage on its first solve: the matrix mirror when it
// is DIA already, per-problem DIA when every problem is a banded matrix
// whose offsets near its middle row cover all of its rows (and whole
// diagonals waste at most half again the batch's off-diagonal count), ELL
// otherwise.
int batch_choose_format(lfgpu_batch *bt) {
  lfgpu_csr *A = bt->A;
  bt->fmt = A->cache->fmt;
  if (bt->fmt == FMT_DIA) return LFGPU_OK;
  hipStream_t s = bt->ctx->stream;
  const int np = bt->nprob;
  const int32_t n = A->nrows;
  int32_t *mid = nullptr;
  LFG_TRY(dalloc(&mid, (size_t)np * (kMaxDiag + 2), "batch middle rows"));
  hipLaunchKernelGGL(k_bdia_mid, dim3(grid_for(np, 256)), dim3(256), 0, s, A->row_ptr, A->col,
                     bt->d_row_off, np, mid);
  std::vector<int32_t> h((size_t)np * (kMaxDiag + 2));
  hipError_t e = hipGetLastError();
  if (e == hipSuccess)
    e = hipMemcpyAsync(h.data(), mid, h.size() * 4, hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  dfree(mid);
  if (e != hipSuccess) return hip_fail(e, "batch storage choice");
  int nd = 0;
  double stored = 0.0, offdiag = 0.0;
  for (int p = 0; p < np; ++p) {
    const int32_t *o = &h[(size_t)p * (kMaxDiag + 2)];
    if (o[0] < 0) return LFGPU_OK;
    stored += (double)o[0] * (bt->row_off[p + 1] - bt->row_off[p]);
    offdiag += o[kMaxDiag + 1];
    nd = std::max(nd, (int)o[0]);
  }
  if (nd == 0 || stored > 1.5 * offdiag) return LFGPU_OK;
  // offsets padded with 0: the diagonal, which holds no off-diagonal value
  std::vector<int32_t> off((size_t)np * kMaxDiag, 0);
  for (int p = 0; p < np; ++p)
    for (int d = 0; d < h[(size_t)p * (kMaxDiag + 2)]; ++d)
      off[(size_t)p * kMaxDiag + d] = h[(size_t)p * (kMaxDiag + 2) + 1 + d];
  LFG_TRY(dalloc(&bt->boff, off.size(), "batch diagonal offsets"));
  LFG_HIP(hipMemcpyAsync(bt->boff, off.data(), off.size() * 4, hipMemcpyHostToDevice, s));
  int *flag = nullptr;
  LFG_TRY(dalloc(&flag, 1, "batch DIA flag"));
  e = hipMemsetAsync(flag, 0, 4, s);
  if (e == hipSuccess) {
    hipLaunchKernelGGL(k_bdia_check, dim3(grid_for(n, 256)), dim3(256), 0, s, A->row_ptr,
                       A->col, n, bt->d_row_off, np, bt->boff, nd, flag);
    e = hipGetLastError();
  }
  int hflag = 1;
  if (e == hipSuccess) e = hipMemcpyAsync(&hflag, flag, 4, hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  dfree(flag);
  if (e != hipSuccess) return hip_fail(e, "batch DIA check");
  if (hflag != 0) return LFGPU_OK;
  LFG_TRY(dalloc(&bt->bval, (size_t)nd * n, "batch DIA values"));
  bt->bndiag = nd;
  bt->fmt = FMT_BDIA;
  bool symoff = nd % 2 == 0;
  for (int p = 0; p < np && symoff; ++p) {
    const int32_t *o = &off[(size_t)p * kMaxDiag];
    if (h[(size_t)p * (kMaxDiag + 2)] != nd) symoff = false;
    for (int d = 0; d < nd && symoff; ++d)
      symoff = o[d] == -o[nd - 1 - d] && (2 * d >= nd || o[d] < 0);
  }
  bt->bsymoff = symoff;
  return LFGPU_OK;
}

__global__ void k_set_flag(int *flag, int v) {
  if (blockIdx.x == 0 && threadIdx.x == 0) *flag = v;
}

// Queues the per-problem DIA value fill and, when the offsets are mirror-
// symmetric, the bitwise symmetry check into bt->d_symflag (0: symmetric;
// 1 otherwise).  Kernels only (capturable).
int batch_fill_enqueue(lfgpu_batch *bt) {
  lfgpu_csr *A = bt->A;
  const int32_t n = A->nrows;
  hipStream_t s = bt->ctx->stream;
  hipLaunchKernelGGL(k_bdia_fill, dim3(grid_for(n, 256)), dim3(256), 0, s, A->row_ptr, A->col,
                     A->val, n, bt->d_row_off, bt->nprob, bt->boff, bt->bndiag, bt->bval,
                     bt->bval32);
  hipLaunchKernelGGL(k_set_flag, dim3(1), dim3(64), 0, s, bt->d_symflag, bt->bsymoff ? 0 : 1);
  if (bt->bsymoff)
    hipLaunchKernelGGL(k_bdia_sym, dim3(grid_for(n, 256)), dim3(256), 0, s, bt->bval, n,
                       bt->d_row_off, bt->nprob, bt->boff, bt->bndiag, bt->d_symflag);
  LFG_HIP(hipGetLastError());
  return LFGPU_OK;
}

// (Re)fills the per-problem DIA values after the matrix values changed.
int batch_fill(lfgpu_batch *bt, bool need32) {
  if (bt->fmt != FMT_BDIA) return LFGPU_OK;
  lfgpu_csr *A = bt->A;
  const int32_t n = A->nrows;
  if (need32 && !bt->bval32) {
    LFG_TRY(dalloc(&bt->bval32, (size_t)bt->bndiag * n, "batch DIA values FP32"));
    bt->bbuilt = ~0ull;
  }
  const bool with32 = bt->bval32 != nullptr;
  if (bt->bbuilt == A->val_version && bt->bbuilt32 == with32) return LFGPU_OK;
  LFG_TRY(batch_fill_enqueue(bt));
  int h = 1;
  hipStream_t s = bt->ctx->stream;
  LFG_HIP(hipMemcpyAsync(&h, bt->d_symflag, 4, hipMemcpyDeviceToHost, s));
  LFG_HIP(hipStreamSynchronize(s));
  bt->bsym = h == 0;
  bt->bbuilt = A->val_version;
  bt->bbuilt32 = with32;
  return LFGPU_OK;
}

// the matrix view of the batched sweeps
MatView batch_matview(const lfgpu_batch *bt) {
  MatView m = view_of(bt->A->cache, bt->A->nrows);
  if (bt->fmt == FMT_BDIA) {
    m.ndiag = bt->bndiag;
    m.val = bt->bval;
    m.val32 = bt->bval32;
  }
  return m;
}

void batch_release_graph(lfgpu_batch *bt) {
  for (auto &g : bt->slot) {
    if (g.exec) (void)hipGraphExecDestroy(g.exec);
    if (g.graph) (void)hipGraphDestroy(g.graph);
    g = lfgpu_batch::Slot{};
  }
  for (auto &g : bt->fused) {
    if (g.exec) (void)hipGraphExecDestroy(g.exec);
    if (g.graph) (void)hipGraphDestroy(g.graph);
    g = lfgpu_batch::Fused{};
  }
}

// slot of a captured chunk: the mode for the mixed modes; for FP64 one slot
// per product variant (symmode 0, 1, 2)
inline int batch_slot(int mode, int sym) {
  return mode != LFGPU_PCG_FP64 ? mode : (sym == 0 ? 0 : sym == 1 ? 3 : 4);
}

// Captures one chunk of `count` iterations into slot `mode` unless the slot
// already holds a chunk with the same arguments.  The stream is drained
// first, so nothing queued before the capture is still pending while it
// runs.
template <class Launch>
int batch_capture(lfgpu_batch *bt, int mode, int count, const double *x, const double *b,
                  double delta2, int max_interval, int sym, Launch &&launch) {
  lfgpu_batch::Slot &g = bt->slot[mode];
  if (g.exec && g.count == count && g.x == x && g.b == b && g.delta2 == delta2 &&
      g.max_interval == max_interval && g.sym == sym)
    return LFGPU_OK;
  if (g.exec) (void)hipGraphExecDestroy(g.exec);
  if (g.graph) (void)hipGraphDestroy(g.graph);
  g = lfgpu_batch::Slot{};
  hipStream_t s = bt->ctx->stream;
  LFG_HIP(hipStreamSynchronize(s));
  LFG_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
  launch();
  hipGraph_t graph = nullptr;
  hipError_t e = hipStreamEndCapture(s, &graph);
  if (e != hipSuccess) return hip_fail(e, "hipStreamEndCapture (batched chunk)");
  hipGraphExec_t exec = nullptr;
  e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
  if (e != hipSuccess) {
    (void)hipGraphDestroy(graph);
    return hip_fail(e, "hipGraphInstantiate (batched chunk)");
  }
  g.graph = graph;
  g.exec = exec;
  g.count = count;
  g.x = x;
  g.b = b;
  g.delta2 = delta2;
  g.max_interval = max_interval;
  g.sym = sym;
  return LFGPU_OK;
}

// Workgroup table for rows_per_group R (a multiple of 256).
int batch_layout(lfgpu_batch *bt, int R) {
  std::vector<BGroup> g;
  std::vector<BProb> pr(bt->nprob);
  for (int p = 0; p < bt->nprob; ++p) {
    const int32_t a = bt->row_off[p], b = bt->row_off[p + 1];
    pr[p].gb = (int32_t)g.size();
    for (int32_t r = a; r < b; r += R) g.push_back(BGroup{p, r, std::min(r + R, b), pr[p].gb});
    if (a == b) g.push_back(BGroup{p, a, a, pr[p].gb});
    pr[p].ge = (int32_t)g.size();
  }
  dfree(bt->groups);
  dfree(bt->probs);
  dfree(bt->part);
  bt->groups = nullptr;
  bt->probs = nullptr;
  bt->part = nullptr;
  bt->ngroups = (int)g.size();
  bt->rows_per_group = R;
  LFG_TRY(dalloc(&bt->groups, g.size(), "batch workgroup table"));
  LFG_TRY(dalloc(&bt->probs, pr.size(), "batch problem table"));
  LFG_TRY(dalloc(&bt->part, (size_t)kPartSlots * g.size(), "batch partials"));
  hipStream_t s = bt->ctx->stream;
  LFG_HIP(hipMemcpyAsync(bt->groups, g.data(), g.size() * sizeof(BGroup), hipMemcpyHostToDevice, s));
  LFG_HIP(hipMemcpyAsync(bt->probs, pr.data(), pr.size() * sizeof(BProb), hipMemcpyHostToDevice, s));
  LFG_HIP(hipStreamSynchronize(s));
  batch_release_graph(bt);
  return LFGPU_OK;
}

BArgs batch_args(lfgpu_batch *bt, double *x) {
  lfgpu_solver_cache *c = bt->A->cache;
  return BArgs{batch_matview(bt), c->dinv, x, c->r, c->p, c->q, bt->part, bt->ngroups,
               bt->groups, bt->probs, bt->st, bt->boff, bt->d_row_off, bt->d_symflag};
}

// states, r = b - A x, p = D^-1 r and Eigen's early exits: kernels only
int batch_init(lfgpu_batch *bt, const double *b, double *x, double rtol, int maxit) {
  hipStream_t s = bt->ctx->stream;
  hipLaunchKernelGGL(k_bstate_init, dim3(grid_for(bt->nprob, 256)), dim3(256), 0, s, bt->st,
                     bt->nprob, rtol * rtol, maxit);
  const BArgs a = batch_args(bt, x);
  if (bt->fmt == FMT_DIA)
    hipLaunchKernelGGL(k_bcg_init<FMT_DIA>, dim3(a.ngroups), dim3(256), 0, s, a, b);
  else if (bt->fmt == FMT_BDIA)
    hipLaunchKernelGGL(k_bcg_init<FMT_BDIA>, dim3(a.ngroups), dim3(256), 0, s, a, b);
  else
    hipLaunchKernelGGL(k_bcg_init<FMT_ELL>, dim3(a.ngroups), dim3(256), 0, s, a, b);
  hipLaunchKernelGGL(k_bcg_init_finish, dim3(bt->nprob), dim3(256), 0, s, a);
  hipLaunchKernelGGL(k_bcg_zero_if_bb0, dim3(a.ngroups), dim3(256), 0, s, a);
  LFG_HIP(hipGetLastError());
  return LFGPU_OK;
}

// symmode of the per-problem DIA product as k_bcg_spmv; -1: from the
// host's symmetry flag of the current values
void batch_chunk(lfgpu_batch *bt, const BArgs &a, int count, int symmode = -1) {
  if (symmode < 0) symmode = bt->bsym ? 1 : 0;
  if (bt->fmt == FMT_DIA)
    launch_bcg_chunk<FMT_DIA>(bt->ctx->stream, a, count, 0);
  else if (bt->fmt == FMT_BDIA)
    launch_bcg_chunk<FMT_BDIA>(bt->ctx->stream, a, count, symmode);
  else
    launch_bcg_chunk<FMT_ELL>(bt->ctx->stream, a, count, 0);
}

// Autotuner: times a few iterations for each candidate workgroup tile (on a
// scratch solution vector, so the caller's initial guess is untouched) and
// keeps the fastest; done once per batch.
int batch_tune(lfgpu_batch *bt, const double *b) {
  static const int cand[4] = {256, 512, 1024, 2048};
  lfgpu_ctx *ctx = bt->ctx;
  double *xs = nullptr;
  LFG_TRY(dalloc(&xs, (size_t)bt->A->nrows, "tuning vector"));
  double best = 1e300;
  int bestR = 256, rc = LFGPU_OK;
  for (int q = 0; q < 4 && rc == LFGPU_OK; ++q) {
    rc = batch_layout(bt, cand[q]);
    if (rc == LFGPU_OK) {
      hipError_t e = hipMemsetAsync(xs, 0, (size_t)bt->A->nrows * sizeof(double), ctx->stream);
      rc = e == hipSuccess ? batch_init(bt, b, xs, 0.0, 1 << 30) : hip_fail(e, "tuning setup");
    }
    if (rc != LFGPU_OK) break;
    const BArgs a = batch_args(bt, xs);
    batch_chunk(bt, a, 2);  // warm
    hipError_t e = hipEventRecord(ctx->ev0, ctx->stream);
    batch_chunk(bt, a, 16);
    if (e == hipSuccess) e = hipGetLastError();
    if (e == hipSuccess) e = hipEventRecord(ctx->ev1, ctx->stream);
    if (e == hipSuccess) e = hipEventSynchronize(ctx->ev1);
    float ms = 0.f;
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1);
    if (e != hipSuccess) {
      rc = hip_fail(e, "tuning run");
      break;
    }
    bt->tune_ms[q] = ms / 16;
    if (ms < best) {
      best = ms;
      bestR = cand[q];
    }
  }
  dfree(xs);
  LFG_TRY(rc);
  return batch_layout(bt, bestR);
}

template <class VT>
BMArgs<VT> batch_mixed_args(lfgpu_batch *bt, const double *b, double *x, double delta2,
                            int max_interval) {
  lfgpu_solver_cache *c = bt->A->cache;
  BMArgs<VT> a;
  a.m = batch_matview(bt);
  a.boff = bt->boff;
  a.prow = bt->d_row_off;
  a.dinv32 = c->dinv32;
  a.dinv16 = c->dinv16;
  a.dinv = c->dinv;
  a.x = x;
  a.xlo = c->xlo;
  a.b = b;
  a.r64 = c->r64;
  if constexpr (std::is_same<VT, float>::value) {
    a.r = c->r32;
    a.p = c->p32;
    a.q = c->q32;
    a.xa = c->xa32;
  } else {
    a.r = c->r;
    a.p = c->p;
    a.q = c->q;
    a.xa = c->xa64;
  }
  a.part = bt->part;
  a.ngroups = bt->ngroups;
  a.groups = bt->groups;
  a.probs = bt->probs;
  a.st = bt->mst;
  a.delta2 = delta2;
  a.max_interval = max_interval;
  return a;
}

template <class VT>
void batch_mixed_chunk(lfgpu_batch *bt, const BMArgs<VT> &a, int count) {
  if (bt->fmt == FMT_DIA)
    launch_bmix_chunk<FMT_DIA, VT>(bt->ctx->stream, a, count);
  else if (bt->fmt == FMT_BDIA)
    launch_bmix_chunk<FMT_BDIA, VT>(bt->ctx->stream, a, count);
  else
    launch_bmix_chunk<FMT_ELL, VT>(bt->ctx->stream, a, count);
}

// Polling callbacks of a batch solve (poll_chunks): copies of the device
// states into the two pinned host slots, and the largest estimate of the
// iterations the problems still running need.  The residual of a mixed
// state was measured at its last reliable update.
inline int rr_iter(const BState &h) { return h.iter; }
inline int rr_iter(const BMState &h) { return h.last_refine; }

template <class ST>
struct BatchPoll {
  lfgpu_batch *bt;
  const ST *dev;
  ST *host;  // two slots of nprob states
  std::vector<PollHist> hist;
  BatchPoll(lfgpu_batch *b, const ST *d, ST *h) : bt(b), dev(d), host(h), hist(b->nprob) {}
  int copy(int j) {
    LFG_HIP(hipMemcpyAsync(host + (size_t)j * bt->nprob, dev, bt->nprob * sizeof(ST),
                           hipMemcpyDeviceToHost, bt->ctx->stream));
    return LFGPU_OK;
  }
  double remaining(int j) {
    const ST *h = host + (size_t)j * bt->nprob;
    double est = -1.0;
    for (int p = 0; p < bt->nprob; ++p)
      if (!h[p].done)
        est = std::max(est, hist[p].next(h[p].rr, rr_iter(h[p]), h[p].iter, h[p].thr, h[p].maxit));
    return est;
  }
};

// states, r64 = b - A x, the FP32 recurrence start and Eigen's early exits:
// kernels only
template <class VT>
int batch_mixed_init(lfgpu_batch *bt, const BMArgs<VT> &a, double rtol, int maxit) {
  hipStream_t s = bt->ctx->stream;
  const dim3 G(a.ngroups), B(256);
  hipLaunchKernelGGL(k_bmstate_init, dim3(grid_for(bt->nprob, 256)), B, 0, s, bt->mst, bt->nprob,
                     rtol * rtol, maxit);
  if (bt->fmt == FMT_DIA)
    hipLaunchKernelGGL((k_bmix_init<FMT_DIA, VT>), G, B, 0, s, a);
  else if (bt->fmt == FMT_BDIA)
    hipLaunchKernelGGL((k_bmix_init<FMT_BDIA, VT>), G, B, 0, s, a);
  else
    hipLaunchKernelGGL((k_bmix_init<FMT_ELL, VT>), G, B, 0, s, a);
  hipLaunchKernelGGL(k_bmix_init_finish<VT>, dim3(bt->nprob), B, 0, s, a);
  hipLaunchKernelGGL(k_bmix_zero_if_bb0<VT>, G, B, 0, s, a);
  LFG_HIP(hipGetLastError());
  return LFGPU_OK;
}

// Mixed-precision batch solve: set-up, graph-replayed chunks until every
// problem has stopped, then x += xlo.
template <class VT>
int batch_run_mixed(lfgpu_batch *bt, const double *b, double *x, const lfgpu_pcg_opts *opts,
                    int check, int mode, int *slot, int *syncs) {
  hipStream_t s = bt->ctx->stream;
  const double delta2 =
      opts->refine_delta > 0.0 ? opts->refine_delta * opts->refine_delta : 0.09;
  const int max_interval = opts->refine_max_interval > 0 ? opts->refine_max_interval : 2147483647;
  const BMArgs<VT> a = batch_mixed_args<VT>(bt, b, x, delta2, max_interval);
  const dim3 G(a.ngroups), B(256);
  LFG_TRY(batch_mixed_init(bt, a, opts->rtol, opts->maxit));
  // the captured arguments depend on the layout (a new layout releases every
  // slot), x, b and the refinement options
  const bool use_graph = opts->graph >= 0;
  if (use_graph)
    LFG_TRY(batch_capture(bt, mode, check, x, b, delta2, max_interval, 0,
                          [&]() { batch_mixed_chunk<VT>(bt, a, check); }));
  BatchPoll<BMState> bp(bt, bt->mst, bt->h_mst);
  LFG_TRY(poll_chunks(
      bt->ctx, check,
      [&]() -> int {
        if (use_graph) {
          LFG_HIP(hipGraphLaunch(bt->slot[mode].exec, s));
        } else {
          batch_mixed_chunk<VT>(bt, a, check);
          LFG_HIP(hipGetLastError());
        }
        return LFGPU_OK;
      },
      [&](int j) { return bp.copy(j); }, [&](int j) { return bp.remaining(j); }, slot, syncs));
  hipLaunchKernelGGL(k_bmix_final<VT>, G, B, 0, s, a);
  LFG_HIP(hipGetLastError());
  return LFGPU_OK;
}

}  // namespace

// ===========================================================================
// C ABI
// ===========================================================================
namespace lfgpu {
int spmv_async(lfgpu_ctx *ctx, lfgpu_csr *A, const double *x, double *y) {
  LFG_TRY(ensure_cache(ctx, A, false, true));
  LFG_TRY(launch_spmv(ctx->stream, A->cache, A->nrows, x, y));
  if (A->blk_n > 0) csr_block_gemv_add(ctx->stream, A, x, y);
  return LFGPU_OK;
}
}  // namespace lfgpu

extern "C" int lfgpu_spmv(lfgpu_ctx *ctx, const lfgpu_csr *Ac, const lfgpu_vec *x, lfgpu_vec *y) {
  LFG_REQUIRE(ctx && Ac && x && y, LFGPU_EINVAL, "NULL argument to lfgpu_spmv");
  LFG_REQUIRE(x->n == Ac->ncols && y->n == Ac->nrows, LFGPU_EINVAL, "SpMV size mismatch");
  LFG_REQUIRE(x->d != y->d, LFGPU_EINVAL, "SpMV input and output must differ");
  lfgpu_csr *A = const_cast<lfgpu_csr *>(Ac);
  LFG_REQUIRE(A->nrows == A->ncols, LFGPU_EINVAL, "SpMV needs a square matrix");
  LFG_TRY(ensure_cache(ctx, A, false, true));
  LFG_HIP(hipEventRecord(ctx->ev0, ctx->stream));
  LFG_TRY(launch_spmv(ctx->stream, A->cache, A->nrows, x->d, y->d));
  if (A->blk_n > 0) csr_block_gemv_add(ctx->stream, A, x->d, y->d);
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
  LFG_TRY(ensure_cache(ctx, A, false, true));
  LFG_HIP(hipEventRecord(ctx->ev0, ctx->stream));
  for (int k = 0; k < reps; ++k) {
    LFG_TRY(launch_spmv(ctx->stream, A->cache, A->nrows, x->d, y->d));
    if (A->blk_n > 0) csr_block_gemv_add(ctx->stream, A, x->d, y->d);
  }
  LFG_HIP(hipEventRecord(ctx->ev1, ctx->stream));
  LFG_HIP(hipEventSynchronize(ctx->ev1));
  float ms = 0.f;
  LFG_HIP(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
  *avg_ms = ms / reps;
  ctx->last_ms = ms;
  return LFGPU_OK;
}

namespace {
// end of a solve: device time, info, status
int pcg_finish(lfgpu_ctx *ctx, const PcgState *hs, bool mixed, lfgpu_pcg_info *info) {
  LFG_HIP(hipEventRecord(ctx->ev1, ctx->stream));
  LFG_HIP(hipEventSynchronize(ctx->ev1));
  float ms = 0.f;
  LFG_HIP(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
  ctx->last_ms = ms;
  const double relres = hs->bb > 0.0 ? std::sqrt(hs->rr / hs->bb) : 0.0;
  if (info) {
    info->iters = hs->iter < 0 ? 0 : hs->iter;
    info->converged = hs->converged;
    info->refinements = mixed ? hs->refinements : 0;
    info->relres = relres;
    info->ms = ms;
    info->syncs = 0;
  }
  if (!hs->converged) {
    char buf[64];
    std::snprintf(buf, sizeof buf, "%.3e", relres);
    set_error(std::string("PCG did not reach rtol within maxit iterations (relres ") + buf + ")");
    return LFGPU_ENOCONV;
  }
  return LFGPU_OK;
}
}  // namespace

extern "C" int lfgpu_pcg(lfgpu_ctx *ctx, const lfgpu_csr *Ac, const lfgpu_vec *b, lfgpu_vec *x,
                         const lfgpu_pcg_opts *opts, lfgpu_pcg_info *info) {
  LFG_REQUIRE(ctx && Ac && b && x && opts, LFGPU_EINVAL, "NULL argument to lfgpu_pcg");
  lfgpu_csr *A = const_cast<lfgpu_csr *>(Ac);
  LFG_REQUIRE(A->nrows == A->ncols && b->n == A->nrows && x->n == A->nrows, LFGPU_EINVAL,
              "PCG size mismatch");
  LFG_REQUIRE(b->d != x->d, LFGPU_EINVAL, "PCG right-hand side and solution must differ");
  LFG_REQUIRE(opts->rtol >= 0.0 && opts->maxit >= 0, LFGPU_EINVAL, "bad PCG options");
  LFG_REQUIRE(opts->mode == LFGPU_PCG_FP64 || opts->mode == LFGPU_PCG_MIXED ||
                  opts->mode == LFGPU_PCG_MIXED_MATRIX,
              LFGPU_EINVAL, "unknown PCG mode");
  LFG_REQUIRE(A->blk_n == 0 || (opts->mode == LFGPU_PCG_FP64 && !opts->single_reduction),
              LFGPU_EINVAL,
              "a matrix with a dense block is solved in FP64 mode with two sweeps only");
  const bool mixed = opts->mode != LFGPU_PCG_FP64;
  LFG_TRY(ensure_cache(ctx, A, mixed, true));
  lfgpu_solver_cache *c = A->cache;
  hipStream_t s = ctx->stream;
  const int32_t n = A->nrows;
  const bool single = opts->mode == LFGPU_PCG_FP64 && opts->single_reduction;
  // the set-up sweep runs on the workgroup shape of the iteration that
  // follows: 512 threads for the mixed sweeps, for FP64 on sliced ELL and for
  // FP64 on DIA from 2^21 rows (2048^2: 929.7 vs 945.8 ms; at 1024^2 and
  // below the 256-thread split is 9-35% faster), else 256 (FP64 on DIA
  // below 2^21 rows, and the single-reduction variant)
  const bool wide = !mixed && !single && c->fmt == FMT_DIA && n >= (1 << 21);
  const int sb = (mixed || wide || (c->fmt == FMT_ELL && !single)) ? kSBMix : kSB64;
  // the chunked mixed sweeps take at least 4 rows per thread: below ~4M
  // rows the natural split (up to kMaxGrid workgroups) spends more on the
  // per-workgroup prologue (partial sums, reliable-update decision) than it
  // gains in parallelism (1024^2: 29.5 -> 23.7 us per FP32-Krylov
  // iteration; the FP64 sweeps are fastest with the natural split).  The
  // FP32-Krylov sweeps take 8 from 2^20 rows and 9 from 2^21 rows (1024^2:
  // 22.0 vs 23.7 us; 2048^2: 719 vs 728 ms, the same iterates), measured
  // values: the trend is not monotonic (8 measured 65 vs 60 us at 2048^2)
  int rpt = 1;
  if (opts->mode == LFGPU_PCG_MIXED)
    rpt = n >= (1 << 21) ? 9 : (n >= (1 << 20) ? 8 : 4);
  else if (mixed)
    rpt = 4;
  const Sweep sw = make_sweep(n, sb, rpt);
  const dim3 G(sw.ngroups), BF(256);
  const unsigned gflat = grid_for(n, 256);
  // an even number of iterations per check keeps the launch parity aligned
  // with the iteration count
  int check = opts->check_every > 0 ? opts->check_every : 64;
  check += check & 1;
  // FP64 two-sweep and row-sweep iterations replay a captured graph of each
  // chunk below 2^20 rows (launch-bound sizes), or as the caller asks
  const bool use_graph = opts->graph > 0 || (opts->graph == 0 && n < (1 << 20));

  if (single && !c->cg1)
    LFG_TRY(dalloc(&c->cg1, (size_t)6 * n, "single-reduction PCG vectors"));
  // sliced ELL: one-row-per-thread sweeps (k_cgr_*, k_mxr_*) in FP64 and
  // the FP32-matrix mode.  The FP32-Krylov mode keeps the chunked sweeps:
  // in the locality order its gathers stay in L2 there too, and it saves
  // the two reduction launches per iteration (2.4M-DOF P3: 82.6 against
  // 90.2 us per iteration; FP64 measured 155.5 against 141.9 us chunked)
  // (a dense block runs the chunked sweeps, which add its product)
  const bool rowsweep =
      !single && c->fmt == FMT_ELL && opts->mode != LFGPU_PCG_MIXED && c->blk_n == 0;
  const int nwg = (int)((n + kRB - 1) / kRB);
  if (rowsweep && !c->rpart)
    LFG_TRY(dalloc(&c->rpart, (size_t)3 * std::max(nwg, 1), "row-sweep partials"));
  // two pinned state slots: the iterations are polled one step behind
  static_assert(2 * sizeof(PcgState) <= 4096, "pinned context buffer holds two states");
  PcgState *hslot = reinterpret_cast<PcgState *>(ctx->pinned);
  int slot = 0, syncs = 0;
  PollHist hist;
  auto poll = [&](auto &&launch) {
    return poll_chunks(
        ctx, check, launch,
        [&](int j) -> int {
          LFG_HIP(hipMemcpyAsync(hslot + j, c->st, sizeof(PcgState), hipMemcpyDeviceToHost, s));
          return LFGPU_OK;
        },
        [&](int j) -> double {
          const PcgState &h = hslot[j];
          if (h.done) return -1.0;
          return mixed ? hist.next(h.rr, h.last_refine, h.iter, h.thr, h.maxit)
                       : hist.next(h.rr, h.iter, h.iter, h.thr, h.maxit);
        },
        &slot, &syncs);
  };
  PcgState h{};
  h.thr = opts->rtol * opts->rtol;
  h.maxit = opts->maxit;
  LFG_HIP(hipMemcpyAsync(c->st, &h, sizeof(PcgState), hipMemcpyHostToDevice, s));
  LFG_HIP(hipEventRecord(ctx->ev0, s));
  // storage rows in the locality order: the solve runs on permuted copies of
  // x and b, and the solution is scattered back before the timing stops
  double *xd = x->d;
  const double *bd = b->d;
  if (c->perm) {
    hipLaunchKernelGGL(k_gather, dim3(grid_for(n, 256)), dim3(256), 0, s, x->d, c->perm, n, c->xp);
    hipLaunchKernelGGL(k_gather, dim3(grid_for(n, 256)), dim3(256), 0, s, b->d, c->perm, n, c->bp);
    LFG_HIP(hipGetLastError());
    xd = c->xp;
    bd = c->bp;
  }
  auto finish = [&](bool mixed_mode) {
    if (c->perm) {
      hipLaunchKernelGGL(k_scatter, dim3(grid_for(n, 256)), dim3(256), 0, s, c->xp, c->perm, n,
                         x->d);
      const hipError_t e = hipGetLastError();
      if (e != hipSuccess) return hip_fail(e, "PCG solution scatter");
    }
    const int rc = pcg_finish(ctx, hslot + slot, mixed_mode, info);
    if (info) info->syncs = syncs;
    return rc;
  };

  const MatView mv = view_of(c, n);
  if (single) {
    Cg1Args a1;
    a1.m = mv;
    a1.dinv = c->dinv;
    a1.x = xd;
    a1.p = c->p;
    for (int q = 0; q < 2; ++q) {
      a1.r[q] = c->cg1 + (size_t)q * n;
      a1.w[q] = c->cg1 + (size_t)(2 + q) * n;
      a1.s[q] = c->cg1 + (size_t)(4 + q) * n;
    }
    a1.part = c->part;
    a1.st = c->st;
    a1.sw = sw;
    const dim3 B(kSB64);
    if (c->fmt == FMT_DIA) {
      hipLaunchKernelGGL((k_cg1_init_r<FMT_DIA, kSB64>), G, B, 0, s, a1, bd);
      hipLaunchKernelGGL((k_cg1_init_w<FMT_DIA, kSB64>), G, B, 0, s, a1);
    } else {
      hipLaunchKernelGGL((k_cg1_init_r<FMT_ELL, kSB64>), G, B, 0, s, a1, bd);
      hipLaunchKernelGGL((k_cg1_init_w<FMT_ELL, kSB64>), G, B, 0, s, a1);
    }
    hipLaunchKernelGGL(k_cg1_init_finish<kSB64>, dim3(1), B, 0, s, a1);
    hipLaunchKernelGGL(k_zero_if_bb0, dim3(gflat), BF, 0, s, xd, n, c->st);
    LFG_HIP(hipGetLastError());
    LFG_TRY(poll([&]() -> int {
      if (c->fmt == FMT_DIA)
        launch_cg1_chunk<FMT_DIA>(s, sw, a1, check);
      else
        launch_cg1_chunk<FMT_ELL>(s, sw, a1, check);
      LFG_HIP(hipGetLastError());
      return LFGPU_OK;
    }));
    return finish(false);
  }
  if (rowsweep && mixed) {
    const double d2 = opts->refine_delta > 0.0 ? opts->refine_delta * opts->refine_delta : 0.09;
    const int mi = opts->refine_max_interval > 0 ? opts->refine_max_interval : 2147483647;
    if (opts->mode == LFGPU_PCG_MIXED) {
      MixArgs<float> m{mv, c->dinv32, c->dinv16, c->dinv, xd, c->xlo, bd, c->r, c->r32, c->p32,
                       c->q32, c->xa32, c->part, c->st, sw, d2, mi, c->rpart, nwg};
      LFG_TRY(run_mixed_rows(s, c, m, n, check, poll));
    } else {
      MixArgs<double> m{mv, c->dinv32, c->dinv16, c->dinv, xd, c->xlo, bd, c->r64, c->r, c->p,
                        c->q, c->xa64, c->part, c->st, sw, d2, mi, c->rpart, nwg};
      LFG_TRY(run_mixed_rows(s, c, m, n, check, poll));
    }
    return finish(true);
  }
  if (rowsweep) {
    const RowArgs ra{mv, c->dinv, xd, c->r, c->p, c->q, c->rpart, nwg, c->st};
    hipLaunchKernelGGL(k_cgr_init<FMT_ELL>, dim3(nwg), dim3(kRB), 0, s, ra, bd);
    hipLaunchKernelGGL(k_cgr_red_init, dim3(1), dim3(kRedB), 0, s, ra);
    hipLaunchKernelGGL(k_zero_if_bb0, dim3(gflat), BF, 0, s, xd, n, c->st);
    LFG_HIP(hipGetLastError());
    LFG_TRY(poll([&]() -> int {
      return run_fp64_chunk(c, s, CG_ROWS_ELL, CgArgs{}, ra, check, false, use_graph);
    }));
    return finish(false);
  }
  // the setup sweep leaves r = b - A x in the buffer the chosen mode reads
  // its true residual from
  double *r_init = opts->mode == LFGPU_PCG_MIXED_MATRIX ? c->r64 : c->r;
  CgArgs a{mv, c->dinv, xd, r_init, c->p, c->q, c->part, c->st, sw};
  a.blk = blk_view(A, c);
  if (a.blk.n > 0) {
    // FP64 only, so the sweeps are 512 wide here unless the matrix is a
    // DIA one below 2^21 rows
    launch_blk_gemv(s, a.blk, xd, nullptr);
    const dim3 B(sb);
    if (sb == kSBMix && c->fmt == FMT_DIA)
      hipLaunchKernelGGL((k_cg_init<FMT_DIA, kSBMix, true>), G, B, 0, s, a, bd);
    else if (sb == kSBMix)
      hipLaunchKernelGGL((k_cg_init<FMT_ELL, kSBMix, true>), G, B, 0, s, a, bd);
    else
      hipLaunchKernelGGL((k_cg_init<FMT_DIA, kSB64, true>), G, B, 0, s, a, bd);
    if (sb == kSBMix)
      hipLaunchKernelGGL(k_cg_init_finish<kSBMix>, dim3(1), B, 0, s, a);
    else
      hipLaunchKernelGGL(k_cg_init_finish<kSB64>, dim3(1), B, 0, s, a);
  } else if (sb == kSBMix) {
    const dim3 B(kSBMix);
    if (c->fmt == FMT_DIA)
      hipLaunchKernelGGL((k_cg_init<FMT_DIA, kSBMix>), G, B, 0, s, a, bd);
    else
      hipLaunchKernelGGL((k_cg_init<FMT_ELL, kSBMix>), G, B, 0, s, a, bd);
    hipLaunchKernelGGL(k_cg_init_finish<kSBMix>, dim3(1), B, 0, s, a);
  } else {
    const dim3 B(kSB64);
    if (c->fmt == FMT_DIA)
      hipLaunchKernelGGL((k_cg_init<FMT_DIA, kSB64>), G, B, 0, s, a, bd);
    else
      hipLaunchKernelGGL((k_cg_init<FMT_ELL, kSB64>), G, B, 0, s, a, bd);
    hipLaunchKernelGGL(k_cg_init_finish<kSB64>, dim3(1), B, 0, s, a);
  }
  hipLaunchKernelGGL(k_zero_if_bb0, dim3(gflat), BF, 0, s, xd, n, c->st);
  LFG_HIP(hipGetLastError());

  const double delta2 =
      opts->refine_delta > 0.0 ? opts->refine_delta * opts->refine_delta : 0.09;
  const int max_interval = opts->refine_max_interval > 0 ? opts->refine_max_interval : 2147483647;
  if (opts->mode == LFGPU_PCG_MIXED) {
    MixArgs<float> m{mv, c->dinv32, c->dinv16, c->dinv, xd, c->xlo, bd, c->r, c->r32, c->p32, c->q32,
                     c->xa32, c->part, c->st, sw, delta2, max_interval};
    LFG_TRY(run_mixed(s, c, sw, m, check, poll));
  } else if (opts->mode == LFGPU_PCG_MIXED_MATRIX) {
    // FP64 vectors: p, q and the increment accumulator reuse the FP64
    // work vectors; r64 needs its own buffer next to the recurrence r.
    MixArgs<double> m{mv, c->dinv32, c->dinv16, c->dinv, xd, c->xlo, bd, c->r64, c->r, c->p, c->q,
                      c->xa64, c->part, c->st, sw, delta2, max_interval};
    LFG_TRY(run_mixed(s, c, sw, m, check, poll));
  } else {
    LFG_TRY(poll([&]() -> int {
      if (c->fmt == FMT_DIA)
        return run_fp64_chunk(c, s, wide ? CG_CHUNK_DIA_WIDE : CG_CHUNK_DIA, a, RowArgs{}, check,
                              c->sym64, use_graph);
      return run_fp64_chunk(c, s, CG_CHUNK_ELL, a, RowArgs{}, check, false, use_graph);
    }));
  }
  return finish(mixed);
}

extern "C" int lfgpu_batch_create(lfgpu_ctx *ctx, lfgpu_csr *A, int nprob,
                                  const int32_t *row_offsets, lfgpu_batch **out) {
  LFG_REQUIRE(ctx && A && row_offsets && out && nprob > 0, LFGPU_EINVAL,
              "bad lfgpu_batch_create arguments");
  *out = nullptr;
  LFG_REQUIRE(A->nrows == A->ncols && row_offsets[0] == 0 && row_offsets[nprob] == A->nrows,
              LFGPU_EINVAL, "row offsets must partition the rows of a square matrix");
  LFG_REQUIRE(A->blk_n == 0, LFGPU_EINVAL, "a batch cannot be made of a matrix with a dense block");
  for (int p = 0; p < nprob; ++p)
    LFG_REQUIRE(row_offsets[p + 1] >= row_offsets[p], LFGPU_EINVAL, "row offsets must increase");
  lfgpu_batch *bt = new lfgpu_batch();
  bt->ctx = ctx;
  bt->A = A;
  bt->nprob = nprob;
  bt->row_off.assign(row_offsets, row_offsets + nprob + 1);
  int rc = dalloc(&bt->st, (size_t)nprob, "batch states");
  if (rc == LFGPU_OK) {
    // two slots: the solve polls one step behind the device
    hipError_t e = hipHostMalloc(&bt->h_st, 2 * nprob * sizeof(BState) + 64, hipHostMallocDefault);
    if (e == hipSuccess) bt->h_symflag = reinterpret_cast<int *>(bt->h_st + 2 * nprob);
    if (e != hipSuccess) rc = hip_fail(e, "pinned batch states");
  }
  if (rc == LFGPU_OK) rc = dalloc(&bt->d_row_off, (size_t)nprob + 1, "batch row offsets");
  if (rc == LFGPU_OK) rc = dalloc(&bt->d_symflag, 1, "batch symmetry flag");
  // every problem is solved on its own rows only: refuse entries that couple
  // two problems
  int *flag = nullptr;
  if (rc == LFGPU_OK) rc = dalloc(&flag, 1, "batch block check flag");
  if (rc == LFGPU_OK) {
    hipStream_t s = ctx->stream;
    int h = 1;
    hipError_t e = hipMemcpyAsync(bt->d_row_off, bt->row_off.data(), ((size_t)nprob + 1) * 4,
                                  hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = hipMemsetAsync(flag, 0, 4, s);
    if (e == hipSuccess && A->nrows > 0) {
      hipLaunchKernelGGL(k_batch_blockcheck, dim3(grid_for(A->nrows, 256)), dim3(256), 0, s,
                         A->row_ptr, A->col, A->nrows, bt->d_row_off, nprob, flag);
      e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpyAsync(&h, flag, 4, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) rc = hip_fail(e, "batch block check");
    else if (h != 0) {
      set_error("lfgpu_batch_create: the matrix couples rows of different problems (it is "
                "not block diagonal over row_offsets)");
      rc = LFGPU_EINVAL;
    }
  }
  dfree(flag);
  if (rc != LFGPU_OK) {
    lfgpu_batch_destroy(bt);
    return rc;
  }
  *out = bt;
  return LFGPU_OK;
}

extern "C" int lfgpu_batch_storage(const lfgpu_batch *bt, int *storage) {
  LFG_REQUIRE(bt && storage, LFGPU_EINVAL, "NULL argument to lfgpu_batch_storage");
  *storage = bt->fmt < 0 ? -1
             : bt->fmt == FMT_BDIA ? LFGPU_STORAGE_BLOCK_DIA
             : bt->fmt == FMT_DIA  ? LFGPU_STORAGE_DIA
             : bt->A->cache && bt->A->cache->s_dcol ? LFGPU_STORAGE_ELL16
                                                    : LFGPU_STORAGE_ELL;
  return LFGPU_OK;
}

extern "C" int lfgpu_batch_destroy(lfgpu_batch *bt) {
  if (!bt) return LFGPU_OK;
  if (bt->ctx) (void)hipStreamSynchronize(bt->ctx->stream);
  batch_release_graph(bt);
  dfree(bt->groups);
  dfree(bt->probs);
  dfree(bt->part);
  dfree(bt->st);
  dfree(bt->mst);
  dfree(bt->d_row_off);
  dfree(bt->d_symflag);
  dfree(bt->boff);
  dfree(bt->bval);
  dfree(bt->bval32);
  if (bt->h_st) (void)hipHostFree(bt->h_st);
  if (bt->h_mst) (void)hipHostFree(bt->h_mst);
  delete bt;
  return LFGPU_OK;
}

namespace {
int batch_report(lfgpu_batch *bt, bool mixed, int slot, int syncs, lfgpu_pcg_info *infos,
                 lfgpu_pcg_info *summary);
}  // namespace

extern "C" int lfgpu_batch_pcg(lfgpu_ctx *ctx, lfgpu_batch *bt, const lfgpu_vec *b, lfgpu_vec *x,
                               const lfgpu_pcg_opts *opts, lfgpu_pcg_info *infos,
                               lfgpu_pcg_info *summary) {
  LFG_REQUIRE(ctx && bt && b && x && opts, LFGPU_EINVAL, "NULL argument to lfgpu_batch_pcg");
  LFG_REQUIRE(b->n == bt->A->nrows && x->n == bt->A->nrows && b->d != x->d, LFGPU_EINVAL,
              "batch PCG size mismatch");
  LFG_REQUIRE(opts->mode == LFGPU_PCG_FP64 || opts->mode == LFGPU_PCG_MIXED ||
                  opts->mode == LFGPU_PCG_MIXED_MATRIX,
              LFGPU_EINVAL, "unknown PCG mode");
  LFG_REQUIRE(opts->rtol >= 0.0 && opts->maxit >= 0, LFGPU_EINVAL, "bad PCG options");
  const bool mixed = opts->mode != LFGPU_PCG_FP64;
  LFG_TRY(ensure_cache(ctx, bt->A, mixed, false));
  hipStream_t s = ctx->stream;
  if (bt->fmt < 0) LFG_TRY(batch_choose_format(bt));
  const float *before32 = bt->bval32;
  LFG_TRY(batch_fill(bt, mixed));
  // a newly allocated FP32 array changes the captured kernel arguments
  if (bt->bval32 != before32) batch_release_graph(bt);
  if (bt->rows_per_group == 0) LFG_TRY(batch_tune(bt, b->d));
  if (mixed && !bt->mst) {
    LFG_TRY(dalloc(&bt->mst, (size_t)bt->nprob, "batch mixed states"));
    LFG_HIP(hipHostMalloc(&bt->h_mst, 2 * bt->nprob * sizeof(BMState), hipHostMallocDefault));
  }
  int check = opts->check_every > 0 ? opts->check_every : 64;
  check += check & 1;
  int slot = 0, syncs = 0;
  LFG_HIP(hipEventRecord(ctx->ev0, s));
  if (opts->mode == LFGPU_PCG_MIXED) {
    LFG_TRY(batch_run_mixed<float>(bt, b->d, x->d, opts, check, opts->mode, &slot, &syncs));
  } else if (opts->mode == LFGPU_PCG_MIXED_MATRIX) {
    LFG_TRY(batch_run_mixed<double>(bt, b->d, x->d, opts, check, opts->mode, &slot, &syncs));
  } else {
    LFG_TRY(batch_init(bt, b->d, x->d, opts->rtol, opts->maxit));
    const BArgs a = batch_args(bt, x->d);
    // One chunk of iterations is captured into a graph once and replayed
    // until every problem has stopped; the captured kernel arguments depend
    // on the batch layout, on x and on the product kernel, which reads the
    // upper diagonals only while the values are bitwise symmetric (bsym is
    // recomputed whenever the values change)
    const bool use_graph = opts->graph >= 0;
    if (use_graph)
      LFG_TRY(batch_capture(bt, batch_slot(LFGPU_PCG_FP64, bt->bsym ? 1 : 0), check, x->d,
                            nullptr, 0.0, 0, bt->bsym ? 1 : 0,
                            [&]() { batch_chunk(bt, a, check); }));
    BatchPoll<BState> bp(bt, bt->st, bt->h_st);
    LFG_TRY(poll_chunks(
        ctx, check,
        [&]() -> int {
          if (use_graph) {
            LFG_HIP(hipGraphLaunch(bt->slot[batch_slot(LFGPU_PCG_FP64, bt->bsym ? 1 : 0)].exec, s));
          } else {
            batch_chunk(bt, a, check);
            LFG_HIP(hipGetLastError());
          }
          return LFGPU_OK;
        },
        [&](int j) { return bp.copy(j); }, [&](int j) { return bp.remaining(j); }, &slot,
        &syncs));
  }
  return batch_report(bt, mixed, slot, syncs, infos, summary);
}

namespace {
// Waits for the solve queued on the context stream since ev0 and fills
// infos / summary from the final state slot.
int batch_report(lfgpu_batch *bt, bool mixed, int slot, int syncs, lfgpu_pcg_info *infos,
                 lfgpu_pcg_info *summary) {
  lfgpu_ctx *ctx = bt->ctx;
  hipStream_t s = ctx->stream;
  LFG_HIP(hipEventRecord(ctx->ev1, s));
  LFG_HIP(hipEventSynchronize(ctx->ev1));
  float ms = 0.f;
  LFG_HIP(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
  ctx->last_ms = ms;
  int conv = 1, maxit_seen = 0;
  double worst = 0.0;
  for (int p = 0; p < bt->nprob; ++p) {
    double bb, rr;
    int iter, converged, refinements = 0;
    if (mixed) {
      const BMState &h = bt->h_mst[(size_t)slot * bt->nprob + p];
      bb = h.bb, rr = h.rr, iter = h.iter, converged = h.converged, refinements = h.refinements;
    } else {
      const BState &h = bt->h_st[(size_t)slot * bt->nprob + p];
      bb = h.bb, rr = h.rr, iter = h.iter, converged = h.converged;
    }
    const double rel = bb > 0.0 ? std::sqrt(rr / bb) : 0.0;
    if (infos) {
      infos[p].iters = iter < 0 ? 0 : iter;
      infos[p].converged = converged;
      infos[p].refinements = refinements;
      infos[p].relres = rel;
      infos[p].ms = ms;
      infos[p].syncs = syncs;
    }
    conv &= converged;
    maxit_seen = std::max(maxit_seen, iter);
    worst = std::max(worst, rel);
  }
  if (summary) {
    summary->iters = maxit_seen;
    summary->converged = conv;
    summary->refinements = bt->rows_per_group;
    summary->relres = worst;
    summary->ms = ms;
    summary->syncs = syncs;
  }
  if (!conv) {
    set_error("batched PCG: some problems did not reach rtol within maxit");
    return LFGPU_ENOCONV;
  }
  return LFGPU_OK;
}
}  // namespace

namespace {
bool same_coeff(const lfgpu_coeff &a, const lfgpu_coeff &b) {
  return a.kind == b.kind && a.value == b.value && a.data == b.data;
}

bool same_key(const lfgpu_batch::FusedKey &a, const lfgpu_batch::FusedKey &b) {
  return a.plan == b.plan && a.fe == b.fe && same_coeff(a.alpha, b.alpha) &&
         same_coeff(a.gamma, b.gamma) && a.has_f == b.has_f &&
         (!a.has_f || same_coeff(a.f, b.f)) && a.rhs == b.rhs && a.x == b.x && a.rtol == b.rtol &&
         a.delta2 == b.delta2 && a.maxit == b.maxit && a.check == b.check &&
         a.max_interval == b.max_interval && a.rows_per_group == b.rows_per_group &&
         a.bval32 == b.bval32;
}

// product variant of the chunks that run on values assembled inside the
// same graph: per-problem DIA with mirror-symmetric offsets reads the
// device's symmetry flag
int fused_symmode(const lfgpu_batch *bt) { return bt->fmt == FMT_BDIA && bt->bsymoff ? 2 : 0; }

// Assembly, the solver storage's value refresh, the solve's set-up and its
// first chunk, as kernels only
int fused_enqueue(lfgpu_ctx *ctx, lfgpu_batch *bt, lfgpu_plan *plan, const lfgpu_fe *fe,
                  const lfgpu_coeff *alpha, const lfgpu_coeff *gamma, const lfgpu_coeff *f,
                  lfgpu_vec *rhs, double *x, const lfgpu_pcg_opts *opts, int check, double delta2,
                  int max_interval) {
  LFG_TRY(assemble_enqueue(ctx, plan, fe, alpha, gamma, f, rhs));
  fill_vals_enqueue(ctx->stream, bt->A, bt->A->cache);
  if (bt->fmt == FMT_BDIA) LFG_TRY(batch_fill_enqueue(bt));
  if (opts->mode == LFGPU_PCG_FP64) {
    LFG_TRY(batch_init(bt, rhs->d, x, opts->rtol, opts->maxit));
    batch_chunk(bt, batch_args(bt, x), check, fused_symmode(bt));
  } else if (opts->mode == LFGPU_PCG_MIXED) {
    const BMArgs<float> a = batch_mixed_args<float>(bt, rhs->d, x, delta2, max_interval);
    LFG_TRY(batch_mixed_init(bt, a, opts->rtol, opts->maxit));
    batch_mixed_chunk<float>(bt, a, check);
  } else {
    const BMArgs<double> a = batch_mixed_args<double>(bt, rhs->d, x, delta2, max_interval);
    LFG_TRY(batch_mixed_init(bt, a, opts->rtol, opts->maxit));
    batch_mixed_chunk<double>(bt, a, check);
  }
  LFG_HIP(hipGetLastError());
  return LFGPU_OK;
}

// The replayed solve: the captured graph, then chunk graphs polled as in
// lfgpu_batch_pcg, then (mixed) x += xlo.
template <class VT>
int fused_replay(lfgpu_batch *bt, const lfgpu_batch::Fused &F, const double *b, double *x,
                 const lfgpu_pcg_opts *opts, int check, double delta2, int max_interval,
                 lfgpu_pcg_info *infos, lfgpu_pcg_info *summary) {
  lfgpu_ctx *ctx = bt->ctx;
  hipStream_t s = ctx->stream;
  const int mode = opts->mode;
  const bool mixed = mode != LFGPU_PCG_FP64;
  BMArgs<VT> am{};
  if (mixed) {
    am = batch_mixed_args<VT>(bt, b, x, delta2, max_interval);
    LFG_TRY(batch_capture(bt, mode, check, x, b, delta2, max_interval, 0,
                          [&]() { batch_mixed_chunk<VT>(bt, am, check); }));
  } else {
    // the chunks read the symmetry flag on the device (symmode 2) until the
    // host has read it with the first state copy; then the variant it names
    const BArgs a = batch_args(bt, x);
    const int symmode = fused_symmode(bt);
    LFG_TRY(batch_capture(bt, batch_slot(mode, symmode), check, x, nullptr, 0.0, 0, symmode,
                          [&]() { batch_chunk(bt, a, check, symmode); }));
    if (symmode == 2)
      for (int v = 0; v < 2; ++v)
        LFG_TRY(batch_capture(bt, batch_slot(mode, v), check, x, nullptr, 0.0, 0, v,
                              [&]() { batch_chunk(bt, a, check, v); }));
  }
  // the replay assembles new values: the host's record of which values the
  // storage and the symmetry flag hold is out of date from here on
  csr_values_changed(bt->A);
  bt->bbuilt = ~0ull;
  int slot = 0, syncs = 0;
  bool first = true;
  // FP64 on per-problem DIA with mirror-symmetric offsets: -1 until the host
  // has read the symmetry flag the graph computed, then 0 (symmetric) or 1
  int hflag = mixed || fused_symmode(bt) != 2 ? 1 : -1;
  const bool devsym = hflag < 0;
  auto launch = [&]() -> int {
    hipGraphExec_t ex = F.exec;
    if (!first)
      ex = bt->slot[mixed ? mode : batch_slot(mode, hflag < 0 ? 2 : (hflag == 0 ? 1 : 0))].exec;
    LFG_HIP(hipGraphLaunch(ex, s));
    first = false;
    return LFGPU_OK;
  };
  LFG_HIP(hipEventRecord(ctx->ev0, s));
  if (mixed) {
    BatchPoll<BMState> bp(bt, bt->mst, bt->h_mst);
    LFG_TRY(poll_chunks(ctx, check, launch, [&](int j) { return bp.copy(j); },
                        [&](int j) { return bp.remaining(j); }, &slot, &syncs));
    hipLaunchKernelGGL(k_bmix_final<VT>, dim3(am.ngroups), dim3(256), 0, s, am);
    LFG_HIP(hipGetLastError());
  } else {
    BatchPoll<BState> bp(bt, bt->st, bt->h_st);
    bool flag_queued = false;
    LFG_TRY(poll_chunks(
        ctx, check, launch,
        [&](int j) -> int {
          LFG_TRY(bp.copy(j));
          if (devsym && !flag_queued) {
            LFG_HIP(hipMemcpyAsync(bt->h_symflag, bt->d_symflag, sizeof(int),
                                   hipMemcpyDeviceToHost, s));
            flag_queued = true;
          }
          return LFGPU_OK;
        },
        [&](int j) {
          // the first step's copies have completed once any step is waited on
          if (devsym && hflag < 0) hflag = *bt->h_symflag != 0 ? 1 : 0;
          return bp.remaining(j);
        },
        &slot, &syncs));
  }
  return batch_report(bt, mixed, slot, syncs, infos, summary);
}
}  // namespace

extern "C" int lfgpu_batch_assemble_pcg(lfgpu_ctx *ctx, lfgpu_batch *bt, lfgpu_plan *plan,
                                        const lfgpu_fe *fe, const lfgpu_coeff *alpha,
                                        const lfgpu_coeff *gamma, const lfgpu_coeff *f,
                                        lfgpu_vec *rhs, lfgpu_vec *x, const lfgpu_pcg_opts *opts,
                                        lfgpu_pcg_info *infos, lfgpu_pcg_info *summary) {
  LFG_REQUIRE(ctx && bt && plan && fe && alpha && gamma && rhs && x && opts, LFGPU_EINVAL,
              "NULL argument to lfgpu_batch_assemble_pcg");
  LFG_REQUIRE(ctx == bt->ctx, LFGPU_EINVAL, "the batch belongs to another context");
  LFG_REQUIRE(&plan->csr == bt->A, LFGPU_EINVAL,
              "the batch was not created on this plan's matrix");
  LFG_REQUIRE(opts->mode == LFGPU_PCG_FP64 || opts->mode == LFGPU_PCG_MIXED ||
                  opts->mode == LFGPU_PCG_MIXED_MATRIX,
              LFGPU_EINVAL, "unknown PCG mode");
  LFG_REQUIRE(rhs->n == bt->A->nrows && x->n == bt->A->nrows && rhs->d != x->d, LFGPU_EINVAL,
              "batch PCG size mismatch");
  int check = opts->check_every > 0 ? opts->check_every : 64;
  check += check & 1;
  const double delta2 =
      opts->refine_delta > 0.0 ? opts->refine_delta * opts->refine_delta : 0.09;
  const int max_interval = opts->refine_max_interval > 0 ? opts->refine_max_interval : 2147483647;
  lfgpu_batch::FusedKey key;
  key.plan = plan;
  key.fe = fe;
  key.alpha = *alpha;
  key.gamma = *gamma;
  key.has_f = f != nullptr;
  if (f) key.f = *f;
  key.rhs = rhs->d;
  key.x = x->d;
  key.rtol = opts->rtol;
  key.delta2 = delta2;
  key.maxit = opts->maxit;
  key.check = check;
  key.max_interval = max_interval;
  key.rows_per_group = bt->rows_per_group;
  key.bval32 = bt->bval32;
  lfgpu_batch::Fused &F = bt->fused[opts->mode];
  if (F.exec && same_key(F.key, key)) {
    if (opts->mode == LFGPU_PCG_MIXED)
      return fused_replay<float>(bt, F, rhs->d, x->d, opts, check, delta2, max_interval, infos,
                                 summary);
    return fused_replay<double>(bt, F, rhs->d, x->d, opts, check, delta2, max_interval, infos,
                                summary);
  }
  // First call with these arguments: the separate calls (which choose the
  // storage, tune the layout and allocate what the mode needs), then the
  // capture of the graph the next calls replay.
  LFG_TRY(lfgpu_assemble_galerkin(ctx, plan, fe, alpha, gamma, f, nullptr, rhs));
  const double asm_ms = ctx->last_ms;
  const int rc = lfgpu_batch_pcg(ctx, bt, rhs, x, opts, infos, summary);
  if (rc != LFGPU_OK && rc != LFGPU_ENOCONV) return rc;
  if (summary) summary->ms += asm_ms;
  key.rows_per_group = bt->rows_per_group;
  key.bval32 = bt->bval32;
  hipStream_t s = ctx->stream;
  if (F.exec) (void)hipGraphExecDestroy(F.exec);
  if (F.graph) (void)hipGraphDestroy(F.graph);
  F = lfgpu_batch::Fused{};
  LFG_HIP(hipStreamSynchronize(s));
  LFG_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
  const int erc = fused_enqueue(ctx, bt, plan, fe, alpha, gamma, f, rhs, x->d, opts, check,
                                delta2, max_interval);
  hipGraph_t graph = nullptr;
  const hipError_t e = hipStreamEndCapture(s, &graph);
  if (erc != LFGPU_OK) {
    if (graph) (void)hipGraphDestroy(graph);
    return erc;
  }
  if (e != hipSuccess) return hip_fail(e, "hipStreamEndCapture (assemble + PCG)");
  hipGraphExec_t exec = nullptr;
  const hipError_t ei = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
  if (ei != hipSuccess) {
    (void)hipGraphDestroy(graph);
    return hip_fail(ei, "hipGraphInstantiate (assemble + PCG)");
  }
  F.graph = graph;
  F.exec = exec;
  F.key = key;
  return rc;
}

extern "C" int lfgpu_solver_storage(const lfgpu_csr *A, int *storage, int *ndiag, int *sym) {
  LFG_REQUIRE(A && storage, LFGPU_EINVAL, "NULL argument to lfgpu_solver_storage");
  const lfgpu_solver_cache *c = A->cache;
  *storage = !c                  ? -1
             : c->fmt == FMT_DIA ? LFGPU_STORAGE_DIA
             : c->s_dcol         ? LFGPU_STORAGE_ELL16
                                 : LFGPU_STORAGE_ELL;
  if (ndiag) *ndiag = c ? c->ndiag : 0;
  if (sym) *sym = c && c->sym64 ? 1 : 0;
  return LFGPU_OK;
}

extern "C" int lfgpu_pcg_sweep_times(lfgpu_ctx *ctx, const lfgpu_csr *A, const lfgpu_vec *b,
                                     lfgpu_vec *x, const lfgpu_pcg_opts *opts, int max_kernels,
                                     double *out_ms, lfgpu_pcg_info *info) {
  LFG_REQUIRE(ctx && A && b && x && opts && out_ms && max_kernels > 0, LFGPU_EINVAL,
              "bad lfgpu_pcg_sweep_times arguments");
  SweepTimer t;
  t.ev.assign(2 * (size_t)max_kernels, nullptr);
  t.kind.assign((size_t)max_kernels, -1);
  hipError_t e = hipSuccess;
  for (size_t k = 0; k < t.ev.size() && e == hipSuccess; ++k) e = hipEventCreate(&t.ev[k]);
  auto destroy = [&]() {
    for (hipEvent_t ev : t.ev)
      if (ev) (void)hipEventDestroy(ev);
  };
  if (e != hipSuccess) {
    destroy();
    return hip_fail(e, "sweep timer events");
  }
  lfgpu_pcg_opts o = *opts;
  o.graph = -1;  // plain launches: every kernel between its own events
  t_timer = &t;
  const int rc = lfgpu_pcg(ctx, A, b, x, &o, info);
  t_timer = nullptr;
  e = hipStreamSynchronize(ctx->stream);
  std::vector<double> d[3];
  for (size_t k = 0; k < t.used && e == hipSuccess && t.err == hipSuccess; ++k) {
    float ms = 0.f;
    e = hipEventElapsedTime(&ms, t.ev[2 * k], t.ev[2 * k + 1]);
    if (t.kind[k] >= 0 && t.kind[k] < 3) d[t.kind[k]].push_back(ms);
  }
  destroy();
  if (e == hipSuccess) e = t.err;
  if (e != hipSuccess) return hip_fail(e, "sweep timing");
  for (int q = 0; q < 3; ++q) {
    out_ms[q] = 0.0;
    if (d[q].empty()) continue;
    std::sort(d[q].begin(), d[q].end());
    const size_t m = d[q].size();
    out_ms[q] = m % 2 ? d[q][m / 2] : 0.5 * (d[q][m / 2 - 1] + d[q][m / 2]);
  }
  return rc;
}

// ===========================================================================
// Allocation-free FP64 Jacobi-PCG (lfgpu_pcg_noalloc).  The work vectors,
// the partial sums and the solver state live in one file-scope device array,
// `scratch`, of fixed size; the product is taken from the CSR itself
// (FMT_CSR), so no solver mirror is built.  The kernels are those of the
// FP64 two-sweep solver with the 256-thread workgroup split lfgpu_pcg uses
// below 2^21 rows, so on a DIA-eligible matrix the iterates equal
// lfgpu_pcg's FP64 iterates.  Each GPU holds its own copy of the array for
// the life of the process; its device address is looked up once per device
// (lfgpu_ctx_create) and used instead of the array's name in host code.
// ===========================================================================
namespace {

constexpr int32_t kNoAllocMaxRows = 1 << 21;
constexpr int64_t kNoAllocVec = kNoAllocMaxRows;  // doubles per work vector
constexpr int64_t kNoAllocPart = (int64_t)kPartSlots * kMaxGrid;
constexpr int64_t kNoAllocState = 32;  // doubles reserved for the PcgState
static_assert(sizeof(PcgState) <= kNoAllocState * sizeof(double), "PcgState too large");
constexpr int64_t kNoAllocDoubles = 4 * kNoAllocVec + kNoAllocPart + kNoAllocState;
constexpr int kNoAllocMaxDevices = 64;

// dinv | r | p | q | partials | state  (64 MiB + 128 KiB + 256 B per GPU)
static __device__ __align__(256) double scratch[kNoAllocDoubles];

double *g_scratch_ptr[kNoAllocMaxDevices] = {nullptr};

// Jacobi inverse diagonal from the CSR, as k_fill_vals: 1 / a_ii, or 1 where
// the diagonal entry is missing or zero (Eigen's DiagonalPreconditioner)
__global__ void k_csr_dinv(const int32_t *row_ptr, const int32_t *col, const double *val,
                           int32_t n, double *dinv) {
  const int32_t r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  double d = 1.0;
  for (int32_t k = row_ptr[r]; k < row_ptr[r + 1]; ++k)
    if (col[k] == r && val[k] != 0.0) d = 1.0 / val[k];
  dinv[r] = d;
}

// The solver state as lfgpu_pcg uploads it (thr = rtol^2, maxit); the array
// keeps the previous call's contents, so every field is reset here.
__global__ void k_noalloc_state(PcgState *st, double thr, int maxit) {
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    *st = PcgState{};
    st->thr = thr;
    st->maxit = maxit;
  }
}

}  // namespace

namespace lfgpu {
// Called by lfgpu_ctx_create with `device` current: records the device
// address of this GPU's copy of `scratch` (once per device).
int noalloc_scratch_init(int device) {
  if (device < 0 || device >= kNoAllocMaxDevices) return LFGPU_OK;  // entry point refuses
  static std::mutex mu;
  std::lock_guard<std::mutex> lock(mu);
  if (g_scratch_ptr[device]) return LFGPU_OK;
  void *ptr = nullptr;
  hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
  if (e != hipSuccess) return hip_fail(e, "hipGetSymbolAddress(scratch)");
  g_scratch_ptr[device] = static_cast<double *>(ptr);
  return LFGPU_OK;
}
}  // namespace lfgpu

extern "C" int lfgpu_pcg_noalloc(lfgpu_ctx *ctx, const lfgpu_csr *A, const lfgpu_vec *b,
                                 lfgpu_vec *x, const lfgpu_pcg_opts *opts,
                                 lfgpu_pcg_info *info) {
  LFG_REQUIRE(ctx && A && b && x && opts, LFGPU_EINVAL, "NULL argument to lfgpu_pcg_noalloc");
  LFG_REQUIRE(A->nrows == A->ncols && b->n == A->nrows && x->n == A->nrows, LFGPU_EINVAL,
              "PCG size mismatch");
  LFG_REQUIRE(b->d != x->d, LFGPU_EINVAL, "PCG right-hand side and solution must differ");
  LFG_REQUIRE(opts->rtol >= 0.0 && opts->maxit >= 0, LFGPU_EINVAL, "bad PCG options");
  LFG_REQUIRE(A->blk_n == 0, LFGPU_EINVAL,
              "lfgpu_pcg_noalloc does not take a matrix with a dense block");
  LFG_REQUIRE(opts->mode == LFGPU_PCG_FP64 && !opts->single_reduction, LFGPU_EINVAL,
              "lfgpu_pcg_noalloc solves in FP64 with two sweeps per iteration only");
  LFG_REQUIRE(A->nrows <= kNoAllocMaxRows, LFGPU_ERANGE,
              "lfgpu_pcg_noalloc: the system has more rows than its fixed work space holds "
              "(2^21 = 2097152)");
  LFG_REQUIRE(ctx->device >= 0 && ctx->device < kNoAllocMaxDevices &&
                  g_scratch_ptr[ctx->device] != nullptr,
              LFGPU_EINVAL, "lfgpu_pcg_noalloc: work space of this device not set up");
  double *base = g_scratch_ptr[ctx->device];
  hipStream_t s = ctx->stream;
  const int32_t n = A->nrows;
  double *dinv = base, *r = base + kNoAllocVec, *p = base + 2 * kNoAllocVec,
         *q = base + 3 * kNoAllocVec, *part = base + 4 * kNoAllocVec;
  PcgState *st = reinterpret_cast<PcgState *>(base + 4 * kNoAllocVec + kNoAllocPart);

  MatView m{};
  m.n = n;
  m.val = A->val;
  m.row_ptr = A->row_ptr;
  m.col = A->col;
  const Sweep sw = make_sweep(n, kSB64);
  const dim3 G(sw.ngroups), B(kSB64), BF(256);
  const unsigned gflat = grid_for(n, 256);
  int check = opts->check_every > 0 ? opts->check_every : 64;
  check += check & 1;

  LFG_HIP(hipEventRecord(ctx->ev0, s));
  hipLaunchKernelGGL(k_noalloc_state, dim3(1), dim3(64), 0, s, st, opts->rtol * opts->rtol,
                     opts->maxit);
  if (n > 0)
    hipLaunchKernelGGL(k_csr_dinv, dim3(gflat), BF, 0, s, A->row_ptr, A->col, A->val, n, dinv);
  const CgArgs a{m, dinv, x->d, r, p, q, part, st, sw};
  hipLaunchKernelGGL((k_cg_init<FMT_CSR, kSB64>), G, B, 0, s, a, b->d);
  hipLaunchKernelGGL(k_cg_init_finish<kSB64>, dim3(1), B, 0, s, a);
  if (n > 0) hipLaunchKernelGGL(k_zero_if_bb0, dim3(gflat), BF, 0, s, x->d, n, st);
  LFG_HIP(hipGetLastError());
  PcgState *hs = reinterpret_cast<PcgState *>(ctx->pinned);
  int syncs = 0;
  for (;;) {
    launch_cg_chunk<FMT_CSR, kSB64>(s, sw, a, check, false);
    LFG_HIP(hipGetLastError());
    LFG_HIP(hipMemcpyAsync(hs, st, sizeof(PcgState), hipMemcpyDeviceToHost, s));
    LFG_HIP(hipStreamSynchronize(s));
    ++syncs;
    if (hs->done) break;
  }
  const int rc = pcg_finish(ctx, hs, false, info);
  if (info) info->syncs = syncs;
  return rc;
}

namespace lfgpu {
void cg_graphs_destroy(void *p) {
  CgGraph *slots = static_cast<CgGraph *>(p);
  for (int k = 0; k < 2; ++k) {
    if (slots[k].exec) (void)hipGraphExecDestroy(slots[k].exec);
    if (slots[k].graph) (void)hipGraphDestroy(slots[k].graph);
  }
  delete[] slots;
}
}  // namespace lfgpu
