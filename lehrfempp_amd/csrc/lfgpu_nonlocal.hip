// Nonlocal boundary terms of the Fisher/KPP model and the dense block of a
// CSR that carries them.
//
// projects/fisher_kpp/strangsplitting.h:120-174 assembles
//   A = stiffness(c) + MassEdgeMatrixProvider(h) on the boundary edges - L
// as a dense matrix (A_ds - L, then sparseView), where the gain h and the
// loss L come from one kernel g(x, y) = 1 / (1 + |x - y|^2) over the
// boundary (modelproblem_island_main.cc:70-126; the convergence and
// three-circle drivers cut g to an annulus rmin <= |x - y| <= rmax):
//   h(x)  = localQuadFunction(mesh, {segment rule}, g(x, .), 1, edge_pred)
//           (strangsplitting.h:78-115): the sum over the boundary edges in
//           index order and their points l of (w_l g(x, y_l)) |e|;
//   L_ij  = [AssembleVectorLocally(1, ...) of the edge load of L_j]_i,
//   L_j(x) = [AssembleVectorLocally(1, ...) of the edge load of g(x, .)]_j,
//           both loads ScalarLoadEdgeVectorProvider::Eval
//           (uscalfe/loc_comp_ellbvp.h:905-916).
// L couples the DOFs of the boundary edges only, so it is kept as a dense
// block on those rows and columns next to the sparse part: a CSR then stands
// for (sparse part) + scale * B (lfgpu_csr_set_dense_block).  The products
// of lfgpu_spmv and of the FP64 Jacobi-PCG add the block's term; its Jacobi
// entries include the block's diagonal.
//
// The kernels below evaluate h and L in the reference's operation order,
// one output per thread (sequential sums over the boundary edges), so their
// values are those of the restatement on the CPU bit for bit.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <map>
#include <utility>
#include <vector>

#include "lfgpu_fe_device.h"
#include "lfgpu_internal.h"
#include "lfgpu_reduce.h"

using namespace lfgpu;

namespace {

// g(x, y) as the drivers write it: (x - y).squaredNorm() and, for the
// annulus, (x - y).norm() compared with both radii
__device__ __forceinline__ double nl_kernel(double x0, double x1, double y0, double y1,
                                            double rmin, double rmax) {
  const double dx = x0 - y0, dy = x1 - y1;
  const double s = dx * dx + dy * dy;
  if (rmax > 0.0) {
    const double d = sqrt(s);
    if (!(d >= rmin && d <= rmax)) return 0.0;
  }
  return 1.0 / (1.0 + s);
}

// Points and lengths of the flagged edges: pts[(q NQE + k) 2 + {0, 1}] =
// SegmentO1::Global of point k of flagged edge q, det[q] = its length.
template <int NLE, int NQE>
__global__ void k_nl_points(const double *xy, const int32_t *edge_v, const int32_t *fedges,
                            int32_t nf, const double *tab, double *pts, double *det) {
  const int32_t q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= nf) return;
  const ETab<NLE, NQE> T{tab};
  const int32_t e = fedges[q];
  const double2 a = reinterpret_cast<const double2 *>(xy)[edge_v[2 * e]];
  const double2 b = reinterpret_cast<const double2 *>(xy)[edge_v[2 * e + 1]];
  det[q] = seg_length(a, b);
#pragma unroll
  for (int k = 0; k < NQE; ++k) {
    const double xh = T.px(k);
    pts[(q * NQE + k) * 2] = b.x * xh + a.x * (1 - xh);
    pts[(q * NQE + k) * 2 + 1] = b.y * xh + a.y * (1 - xh);
  }
}

__global__ void k_nl_zero(double *v, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) v[i] = 0.0;
}

// h at point k of flagged edge q: sum += (w_l g(x, y_l)) det over the
// flagged edges in index order
template <int NLE, int NQE>
__global__ void k_nl_gain(const int32_t *fedges, int32_t nf, const double *tab, const double *pts,
                          const double *det, double rmin, double rmax, double *eta) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (int64_t)nf * NQE) return;
  const ETab<NLE, NQE> T{tab};
  const double x0 = pts[2 * t], x1 = pts[2 * t + 1];
  double sum = 0.0;
  for (int32_t q = 0; q < nf; ++q) {
    const double d = det[q];
#pragma unroll
    for (int l = 0; l < NQE; ++l)
      sum += T.w(l) * nl_kernel(x0, x1, pts[(q * NQE + l) * 2], pts[(q * NQE + l) * 2 + 1], rmin,
                                rmax) *
             d;
  }
  const int32_t q = (int32_t)(t / NQE);
  eta[(int64_t)fedges[q] * NQE + (t - (int64_t)q * NQE)] = sum;
}

// Entry j of the assembled edge load of g(x, .): the flagged edges of DOF j
// (list [p0, p1) of (flagged edge, local index)) in index order, each
// element vector entry summed over the points as Eval does.
template <int NLE, int NQE>
__device__ __forceinline__ double nl_column_load(const ETab<NLE, NQE> &T, const int32_t *bq,
                                                 const int8_t *bm, int32_t p0, int32_t p1,
                                                 const double *pts, const double *det, double x0,
                                                 double x1, double rmin, double rmax) {
  double acc = 0.0;
  for (int32_t p = p0; p < p1; ++p) {
    const int32_t q = bq[p];
    const int m = bm[p];
    const double d = det[q];
    double v = 0.0;
#pragma unroll
    for (int l = 0; l < NQE; ++l) {
      const double w =
          (T.w(l) * d) *
          nl_kernel(x0, x1, pts[(q * NQE + l) * 2], pts[(q * NQE + l) * 2 + 1], rmin, rmax);
      v += T.N(m, l) * w;
    }
    acc += v;
  }
  return acc;
}

// L[bi nb + bj]: the edge load of L_j over the flagged edges of DOF i
template <int NLE, int NQE>
__global__ __launch_bounds__(256) void k_nl_loss(int32_t nb, const int32_t *bptr,
                                                 const int32_t *bq, const int8_t *bm,
                                                 const double *tab, const double *pts,
                                                 const double *det, double rmin, double rmax,
                                                 double *L) {
  const int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (t >= (int64_t)nb * nb) return;
  const int32_t bi = (int32_t)(t / nb), bj = (int32_t)(t - (int64_t)bi * nb);
  const ETab<NLE, NQE> T{tab};
  const int32_t j0 = bptr[bj], j1 = bptr[bj + 1];
  double acc = 0.0;
  for (int32_t p = bptr[bi]; p < bptr[bi + 1]; ++p) {
    const int32_t q = bq[p];
    const int m = bm[p];
    const double d = det[q];
    double v = 0.0;
#pragma unroll
    for (int k = 0; k < NQE; ++k) {
      const double x0 = pts[(q * NQE + k) * 2], x1 = pts[(q * NQE + k) * 2 + 1];
      const double w =
          (T.w(k) * d) * nl_column_load<NLE, NQE>(T, bq, bm, j0, j1, pts, det, x0, x1, rmin, rmax);
      v += T.N(m, k) * w;
    }
    acc += v;
  }
  L[t] = acc;
}

// y[rows[b]] += scale (B x[rows])_b, one wave per block row in the order of
// the solver's k_blk_gemv (columns over the lanes, 8 loads in flight per
// lane, then a butterfly)
__global__ __launch_bounds__(256) void k_blk_gemv_add(const double *B, int32_t nb,
                                                      const int32_t *rows, double scale,
                                                      const double *__restrict__ x, double *y) {
  const int32_t row = blockIdx.x * 4 + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (row >= nb) return;
  const double *__restrict__ Br = B + (int64_t)row * nb;
  double acc = 0.0;
  for (int32_t c0 = lane; c0 < nb; c0 += 8 * 64) {
    double bv[8], xv[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const int32_t c = c0 + 64 * u;
      const bool in = c < nb;
      bv[u] = in ? Br[c] : 0.0;
      xv[u] = in ? x[rows[c]] : 0.0;
    }
#pragma unroll
    for (int u = 0; u < 8; ++u)
      if (c0 + 64 * u < nb) acc += bv[u] * xv[u];
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o, 64);
  if (lane == 0) y[rows[row]] += scale * acc;
}

// The flagged edges (ascending) and, for every DOF on them, its (flagged
// edge, local index) pairs in edge order; rows = those DOFs, ascending.
struct FlaggedLists {
  std::vector<int32_t> fedges, rows, bptr, bq;
  std::vector<int8_t> bm;
};

FlaggedLists flagged_lists(const lfgpu_edges *E, const uint8_t *flags) {
  FlaggedLists F;
  std::map<int32_t, std::vector<std::pair<int32_t, int8_t>>> per_dof;
  for (int64_t e = 0; e < E->nedges; ++e) {
    if (!flags[e]) continue;
    const int32_t q = (int32_t)F.fedges.size();
    F.fedges.push_back((int32_t)e);
    for (int m = 0; m < E->nle; ++m)
      per_dof[E->h_dofs[(size_t)e * E->nle + m]].push_back({q, (int8_t)m});
  }
  F.bptr.push_back(0);
  for (const auto &kv : per_dof) {
    F.rows.push_back(kv.first);
    for (const auto &qm : kv.second) {
      F.bq.push_back(qm.first);
      F.bm.push_back(qm.second);
    }
    F.bptr.push_back((int32_t)F.bq.size());
  }
  return F;
}

// Device copies of the flagged edges' points and lengths.
struct NlPoints {
  int32_t *fedges = nullptr;
  double *pts = nullptr, *det = nullptr;
  ~NlPoints() {
    dfree(fedges);
    dfree(pts);
    dfree(det);
  }
};

int nl_points(lfgpu_ctx *ctx, const lfgpu_edges *E, const FlaggedLists &F, NlPoints &P) {
  hipStream_t s = ctx->stream;
  const int32_t nf = (int32_t)F.fedges.size();
  LFG_TRY(dalloc(&P.fedges, (size_t)std::max(nf, 1), "flagged edges"));
  LFG_TRY(dalloc(&P.pts, (size_t)std::max(nf, 1) * E->nqe * 2, "flagged edge points"));
  LFG_TRY(dalloc(&P.det, (size_t)std::max(nf, 1), "flagged edge lengths"));
  if (nf == 0) return LFGPU_OK;
  LFG_HIP(hipMemcpyAsync(P.fedges, F.fedges.data(), nf * sizeof(int32_t), hipMemcpyHostToDevice,
                         s));
  const dim3 g(grid_for(nf, 256));
  if (E->nle == 2)
    hipLaunchKernelGGL((k_nl_points<2, 2>), g, dim3(256), 0, s, E->mesh->xy, E->edge_v, P.fedges,
                       nf, E->tab, P.pts, P.det);
  else
    hipLaunchKernelGGL((k_nl_points<4, 4>), g, dim3(256), 0, s, E->mesh->xy, E->edge_v, P.fedges,
                       nf, E->tab, P.pts, P.det);
  LFG_HIP(hipGetLastError());
  return LFGPU_OK;
}

bool radii_ok(double rmin, double rmax) {
  return std::isfinite(rmin) && std::isfinite(rmax) && (rmax <= 0.0 || rmin <= rmax);
}

}  // namespace

namespace lfgpu {

void csr_free_block(lfgpu_csr *A) {
  dfree(A->blk_rows);
  dfree(A->blk);
  A->blk_rows = nullptr;
  A->blk = nullptr;
  A->blk_n = 0;
  A->blk_scale = 0.0;
}

int csr_copy_block(lfgpu_ctx *ctx, const lfgpu_csr *from, double scale, lfgpu_csr *to) {
  csr_free_block(to);
  const int32_t nb = from->blk_n;
  if (nb == 0) return LFGPU_OK;
  LFG_TRY(dalloc(&to->blk_rows, (size_t)nb, "dense block rows"));
  LFG_TRY(dalloc(&to->blk, (size_t)nb * nb, "dense block"));
  hipStream_t s = ctx->stream;
  LFG_HIP(hipMemcpyAsync(to->blk_rows, from->blk_rows, (size_t)nb * sizeof(int32_t),
                         hipMemcpyDeviceToDevice, s));
  LFG_HIP(hipMemcpyAsync(to->blk, from->blk, (size_t)nb * nb * sizeof(double),
                         hipMemcpyDeviceToDevice, s));
  to->blk_n = nb;
  to->blk_scale = scale;
  csr_values_changed(to);
  return LFGPU_OK;
}

void csr_block_gemv_add(hipStream_t s, const lfgpu_csr *A, const double *x, double *y) {
  hipLaunchKernelGGL(k_blk_gemv_add, dim3(grid_for(A->blk_n, 4)), dim3(256), 0, s, A->blk,
                     A->blk_n, A->blk_rows, A->blk_scale, x, y);
}

}  // namespace lfgpu

extern "C" int lfgpu_csr_set_dense_block(lfgpu_ctx *ctx, lfgpu_csr *A, int32_t nb,
                                         const int32_t *rows, const lfgpu_vec *B, double scale) {
  LFG_REQUIRE(ctx && A, LFGPU_EINVAL, "NULL argument to lfgpu_csr_set_dense_block");
  LFG_REQUIRE(A->nrows == A->ncols, LFGPU_EINVAL, "a dense block needs a square matrix");
  LFG_REQUIRE(nb >= 0 && nb <= A->nrows, LFGPU_EINVAL, "bad dense block size");
  LFG_REQUIRE(std::isfinite(scale), LFGPU_EINVAL, "the dense block's scale must be finite");
  hipStream_t s = ctx->stream;
  LFG_HIP(hipStreamSynchronize(s));
  if (nb == 0) {
    csr_free_block(A);
    csr_values_changed(A);
    return LFGPU_OK;
  }
  LFG_REQUIRE(rows && B, LFGPU_EINVAL, "a dense block needs its rows and values");
  LFG_REQUIRE(B->n >= (int64_t)nb * nb, LFGPU_EINVAL, "the block vector holds fewer than nb^2 values");
  for (int32_t k = 0; k < nb; ++k)
    LFG_REQUIRE(rows[k] >= 0 && rows[k] < A->nrows && (k == 0 || rows[k] > rows[k - 1]),
                LFGPU_EINVAL, "the block rows must be ascending, distinct rows of the matrix");
  csr_free_block(A);
  LFG_TRY(dalloc(&A->blk_rows, (size_t)nb, "dense block rows"));
  LFG_TRY(dalloc(&A->blk, (size_t)nb * nb, "dense block"));
  LFG_HIP(hipMemcpyAsync(A->blk_rows, rows, (size_t)nb * sizeof(int32_t), hipMemcpyHostToDevice,
                         s));
  LFG_HIP(hipMemcpyAsync(A->blk, B->d, (size_t)nb * nb * sizeof(double), hipMemcpyDeviceToDevice,
                         s));
  LFG_HIP(hipStreamSynchronize(s));
  A->blk_n = nb;
  A->blk_scale = scale;
  csr_values_changed(A);
  return LFGPU_OK;
}

extern "C" int lfgpu_csr_dense_block(const lfgpu_csr *A, int32_t *nb, double *scale) {
  LFG_REQUIRE(A, LFGPU_EINVAL, "NULL matrix");
  if (nb) *nb = A->blk_n;
  if (scale) *scale = A->blk_scale;
  return LFGPU_OK;
}

extern "C" int lfgpu_nonlocal_gain(lfgpu_ctx *ctx, const lfgpu_edges *E, const uint8_t *flags,
                                   double rmin, double rmax, lfgpu_vec *eta) {
  LFG_REQUIRE(ctx && E && flags && eta, LFGPU_EINVAL, "NULL argument to lfgpu_nonlocal_gain");
  LFG_REQUIRE(eta->n == E->nedges * E->nqe, LFGPU_EINVAL, "eta must hold nedges*nqe values");
  LFG_REQUIRE(radii_ok(rmin, rmax), LFGPU_EINVAL, "bad kernel radii");
  hipStream_t s = ctx->stream;
  const FlaggedLists F = flagged_lists(E, flags);
  const int32_t nf = (int32_t)F.fedges.size();
  if (eta->n > 0)
    hipLaunchKernelGGL(k_nl_zero, dim3(grid_for(eta->n, 256)), dim3(256), 0, s, eta->d, eta->n);
  NlPoints P;
  LFG_TRY(nl_points(ctx, E, F, P));
  if (nf > 0) {
    const dim3 g(grid_for((int64_t)nf * E->nqe, 64));
    if (E->nle == 2)
      hipLaunchKernelGGL((k_nl_gain<2, 2>), g, dim3(64), 0, s, P.fedges, nf, E->tab, P.pts, P.det,
                         rmin, rmax, eta->d);
    else
      hipLaunchKernelGGL((k_nl_gain<4, 4>), g, dim3(64), 0, s, P.fedges, nf, E->tab, P.pts, P.det,
                         rmin, rmax, eta->d);
  }
  LFG_HIP(hipGetLastError());
  LFG_HIP(hipStreamSynchronize(s));
  return LFGPU_OK;
}

extern "C" int lfgpu_nonlocal_loss(lfgpu_ctx *ctx, const lfgpu_edges *E, const uint8_t *flags,
                                   double rmin, double rmax, int32_t *nb_out, int32_t *rows,
                                   lfgpu_vec *L) {
  LFG_REQUIRE(ctx && E && flags && nb_out, LFGPU_EINVAL, "NULL argument to lfgpu_nonlocal_loss");
  LFG_REQUIRE(radii_ok(rmin, rmax), LFGPU_EINVAL, "bad kernel radii");
  const FlaggedLists F = flagged_lists(E, flags);
  const int32_t nb = (int32_t)F.rows.size();
  *nb_out = nb;
  if (!rows && !L) return LFGPU_OK;  // size query
  LFG_REQUIRE(rows && L, LFGPU_EINVAL, "lfgpu_nonlocal_loss needs both rows and L");
  LFG_REQUIRE(L->n >= (int64_t)nb * nb, LFGPU_EINVAL, "L holds fewer than nb^2 values");
  std::copy(F.rows.begin(), F.rows.end(), rows);
  if (nb == 0) return LFGPU_OK;
  hipStream_t s = ctx->stream;
  NlPoints P;
  LFG_TRY(nl_points(ctx, E, F, P));
  int32_t *bptr = nullptr, *bq = nullptr;
  int8_t *bm = nullptr;
  int rc = dalloc(&bptr, F.bptr.size(), "boundary DOF lists");
  if (rc == LFGPU_OK) rc = dalloc(&bq, F.bq.size(), "boundary DOF edges");
  if (rc == LFGPU_OK) rc = dalloc(&bm, F.bm.size(), "boundary DOF local indices");
  hipError_t e = hipSuccess;
  if (rc == LFGPU_OK)
    e = hipMemcpyAsync(bptr, F.bptr.data(), F.bptr.size() * sizeof(int32_t),
                       hipMemcpyHostToDevice, s);
  if (rc == LFGPU_OK && e == hipSuccess)
    e = hipMemcpyAsync(bq, F.bq.data(), F.bq.size() * sizeof(int32_t), hipMemcpyHostToDevice, s);
  if (rc == LFGPU_OK && e == hipSuccess)
    e = hipMemcpyAsync(bm, F.bm.data(), F.bm.size(), hipMemcpyHostToDevice, s);
  if (rc == LFGPU_OK && e == hipSuccess) {
    const dim3 g(grid_for((int64_t)nb * nb, 256));
    if (E->nle == 2)
      hipLaunchKernelGGL((k_nl_loss<2, 2>), g, dim3(256), 0, s, nb, bptr, bq, bm, E->tab, P.pts,
                         P.det, rmin, rmax, L->d);
    else
      hipLaunchKernelGGL((k_nl_loss<4, 4>), g, dim3(256), 0, s, nb, bptr, bq, bm, E->tab, P.pts,
                         P.det, rmin, rmax, L->d);
    e = hipGetLastError();
  }
  if (rc == LFGPU_OK && e == hipSuccess) e = hipStreamSynchronize(s);
  dfree(bptr);
  dfree(bq);
  dfree(bm);
  LFG_TRY(rc);
  if (e != hipSuccess) return hip_fail(e, "lfgpu_nonlocal_loss");
  return LFGPU_OK;
}
