// Sparse-matrix utilities of liblfgpu: CSR accessors, device
// setFromTriplets (COOMatrix::makeSparse), deterministic BLAS-1 helpers and
// the adjoint-gradient kernel.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>

#include <rocprim/rocprim.hpp>

#include "lfgpu_fe_device.h"
#include "lfgpu_internal.h"
#include "lfgpu_reduce.h"

using namespace lfgpu;

namespace {

constexpr int kRedBlocks = 1024;  // fixed grid of the deterministic reductions

__global__ void k_dot_partial(const double *x, const double *y, int64_t n, double *part) {
  __shared__ double sh[8];
  double v = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256)
    v += x[i] * y[i];
  v = block_sum<256>(v, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = v;
}

__global__ void k_sum_partials(const double *part, int n, double *out) {
  __shared__ double sh[8];
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
  csr_free_block(A);
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

// ---------------------------------------------------------------------------
// Matrix copies and same-pattern linear combinations
// ---------------------------------------------------------------------------
namespace {

// *bad != 0 if the patterns differ anywhere
__global__ void k_pattern_diff(const int32_t *rp0, const int32_t *c0, const int32_t *rp1,
                               const int32_t *c1, int32_t nrows, int64_t nnz, int *bad) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  bool diff = false;
  if (t <= nrows) diff = rp0[t] != rp1[t];
  if (t < nnz) diff = diff || c0[t] != c1[t];
  if (diff) atomicOr(bad, 1);
}

__global__ void k_vals_axpby(double a, const double *x, double b, const double *y, double *z,
                             int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) z[i] = a * x[i] + b * y[i];
}

// A new matrix with A's pattern; values copied when copy_vals.
int csr_new_like(lfgpu_ctx *ctx, const lfgpu_csr *A, bool copy_vals, lfgpu_csr **out) {
  lfgpu_csr *B = new lfgpu_csr();
  B->ctx = ctx;
  B->nrows = A->nrows;
  B->ncols = A->ncols;
  B->nnz = A->nnz;
  const size_t nz = (size_t)std::max<int64_t>(A->nnz, 1);
  int rc = dalloc(&B->row_ptr, (size_t)A->nrows + 1, "CSR row pointers");
  if (rc == LFGPU_OK) rc = dalloc(&B->col, nz, "CSR columns");
  if (rc == LFGPU_OK) rc = dalloc(&B->val, nz, "CSR values");
  hipError_t e = hipSuccess;
  if (rc == LFGPU_OK)
    e = hipMemcpyAsync(B->row_ptr, A->row_ptr, ((size_t)A->nrows + 1) * sizeof(int32_t),
                       hipMemcpyDeviceToDevice, ctx->stream);
  if (rc == LFGPU_OK && e == hipSuccess && A->nnz)
    e = hipMemcpyAsync(B->col, A->col, (size_t)A->nnz * sizeof(int32_t),
                       hipMemcpyDeviceToDevice, ctx->stream);
  if (rc == LFGPU_OK && e == hipSuccess && copy_vals && A->nnz)
    e = hipMemcpyAsync(B->val, A->val, (size_t)A->nnz * sizeof(double), hipMemcpyDeviceToDevice,
                       ctx->stream);
  if (rc == LFGPU_OK && e != hipSuccess) rc = hip_fail(e, "CSR copy");
  if (rc != LFGPU_OK) {
    lfgpu_csr_destroy(B);
    return rc;
  }
  *out = B;
  return LFGPU_OK;
}

}  // namespace

extern "C" int lfgpu_csr_clone(lfgpu_ctx *ctx, const lfgpu_csr *A, lfgpu_csr **out) {
  LFG_REQUIRE(ctx && A && out, LFGPU_EINVAL, "NULL argument to lfgpu_csr_clone");
  *out = nullptr;
  LFG_TRY(csr_new_like(ctx, A, true, out));
  int rc = A->blk_n > 0 ? csr_copy_block(ctx, A, A->blk_scale, *out) : LFGPU_OK;
  const hipError_t e = hipStreamSynchronize(ctx->stream);
  if (rc == LFGPU_OK && e != hipSuccess) rc = hip_fail(e, "lfgpu_csr_clone");
  if (rc != LFGPU_OK) {
    lfgpu_csr_destroy(*out);
    *out = nullptr;
  }
  return rc;
}

extern "C" int lfgpu_csr_axpby(lfgpu_ctx *ctx, double a, const lfgpu_csr *X, double b,
                               const lfgpu_csr *Y, lfgpu_csr **out) {
  LFG_REQUIRE(ctx && X && Y && out, LFGPU_EINVAL, "NULL argument to lfgpu_csr_axpby");
  *out = nullptr;
  LFG_REQUIRE(X->nrows == Y->nrows && X->ncols == Y->ncols && X->nnz == Y->nnz, LFGPU_EINVAL,
              "lfgpu_csr_axpby needs matrices of equal size and nonzero count");
  LFG_REQUIRE(X->blk_n == 0 || Y->blk_n == 0, LFGPU_EINVAL,
              "lfgpu_csr_axpby: at most one of the matrices may carry a dense block");
  hipStream_t s = ctx->stream;
  if (X->row_ptr != Y->row_ptr || X->col != Y->col) {
    int *bad = nullptr;
    LFG_TRY(dalloc(&bad, 1, "pattern flag"));
    int h = 0;
    hipError_t e = hipMemsetAsync(bad, 0, sizeof(int), s);
    if (e == hipSuccess) {
      const int64_t cnt = std::max<int64_t>((int64_t)X->nrows + 1, X->nnz);
      hipLaunchKernelGGL(k_pattern_diff, dim3(grid_for(cnt, 256)), dim3(256), 0, s, X->row_ptr,
                         X->col, Y->row_ptr, Y->col, X->nrows, X->nnz, bad);
      e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpyAsync(&h, bad, sizeof(int), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    dfree(bad);
    if (e != hipSuccess) return hip_fail(e, "lfgpu_csr_axpby pattern check");
    LFG_REQUIRE(h == 0, LFGPU_EINVAL, "lfgpu_csr_axpby: the two matrices have different patterns");
  }
  lfgpu_csr *Z = nullptr;
  LFG_TRY(csr_new_like(ctx, X, false, &Z));
  if (X->nnz)
    hipLaunchKernelGGL(k_vals_axpby, dim3(grid_for(X->nnz, 256)), dim3(256), 0, s, a, X->val, b,
                       Y->val, Z->val, X->nnz);
  hipError_t e = hipGetLastError();
  // the block of the operand that has one, scaled by its factor
  int rc = LFGPU_OK;
  if (e == hipSuccess && X->blk_n > 0) rc = csr_copy_block(ctx, X, a * X->blk_scale, Z);
  if (e == hipSuccess && Y->blk_n > 0) rc = csr_copy_block(ctx, Y, b * Y->blk_scale, Z);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  if (e != hipSuccess || rc != LFGPU_OK) {
    lfgpu_csr_destroy(Z);
    return rc != LFGPU_OK ? rc : hip_fail(e, "lfgpu_csr_axpby");
  }
  *out = Z;
  return LFGPU_OK;
}
