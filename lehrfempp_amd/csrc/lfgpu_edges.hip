// Boundary (codim-1) terms: edges, the segment element, the edge providers
// and the per-row edge records of an assembly plan.
//
// MassEdgeMatrixProvider (uscalfe/loc_comp_ellbvp.h:367-527) and
// ScalarLoadEdgeVectorProvider (:784-918) are assembled by
// AssembleMatrixLocally(1, ...) / AssembleVectorLocally(1, ...) after the cell
// terms (examples/ellbvp_linfe/ellbvp_linfe_demo.cc:272-315), so their
// triplets follow the cell triplets and setFromTriplets adds them last.  A
// plan with edges therefore keeps, per row, the edge records (edge, local
// row) of the selected edges in edge order with the CSR slot of every local
// column; the row-owner assembly visits them after the cell records.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <string>
#include <type_traits>
#include <vector>

#include "lfgpu_fe_device.h"
#include "lfgpu_internal.h"
#include "lfgpu_reduce.h"

using namespace lfgpu;

namespace {

constexpr uint8_t kDrop = 0xFF;

// make_QuadRule(kSegment, 2p) = GaussLegendre(p + 1) on [0, 1]
// (quad/make_quad_rule.cc:13-18, quad/gauss_quadrature.cc:19-50): the
// correctly rounded points and weights.
constexpr double kGL2Pts[2] = {0.2113248654051871, 0.7886751345948129};
constexpr double kGL2W[2] = {0.5, 0.5};
constexpr double kGL4Pts[4] = {0.06943184420297371, 0.33000947820757187, 0.6699905217924281,
                               0.9305681557970263};
constexpr double kGL4W[4] = {0.17392742256872692, 0.32607257743127305, 0.32607257743127305,
                             0.17392742256872692};

// FeLagrangeO1Segment / FeLagrangeO3Segment values (uscalfe/lagr_fe.h:
// 284-330, :1155-1230), evaluated as the reference's array expressions.
void segment_shape(int degree, double x, double *N) {
  if (degree == 1) {
    N[0] = 1.0 - x;
    N[1] = x;
  } else {
    N[0] = ((4.5 * (1.0 - x)) * (1.0 / 3.0 - x)) * (2.0 / 3.0 - x);
    N[1] = ((4.5 * x) * (x - 1.0 / 3.0)) * (x - 2.0 / 3.0);
    N[2] = ((13.5 * x) * (1 - x)) * (2.0 / 3.0 - x);
    N[3] = ((13.5 * x) * (1 - x)) * (x - 1.0 / 3.0);
  }
}

template <int NLE, int NQE>
__global__ __launch_bounds__(256) void k_edge_matrices(const double *xy, const int32_t *edge_v,
                                                       int64_t nedges, const double *tab,
                                                       CoeffD eta, double *out) {
  const int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (t >= nedges * NLE) return;
  const int64_t e = t / NLE;
  const int i = (int)(t - e * NLE);
  const ETab<NLE, NQE> T{tab};
  const double2 a = reinterpret_cast<const double2 *>(xy)[edge_v[2 * e]];
  const double2 b = reinterpret_cast<const double2 *>(xy)[edge_v[2 * e + 1]];
  double gk[NQE];
#pragma unroll
  for (int k = 0; k < NQE; ++k) gk[k] = edge_coeff<NQE>(eta, e, k, a, b, T.px(k));
  double row[NLE];
  edge_row<NLE, NQE>(T, seg_length(a, b), gk, i, row);
#pragma unroll
  for (int j = 0; j < NLE; ++j) out[t * NLE + j] = row[j];
}

template <int NLE, int NQE>
__global__ __launch_bounds__(256) void k_edge_vectors(const double *xy, const int32_t *edge_v,
                                                      int64_t nedges, const double *tab, CoeffD h,
                                                      double *out) {
  const int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (t >= nedges * NLE) return;
  const int64_t e = t / NLE;
  const int i = (int)(t - e * NLE);
  const ETab<NLE, NQE> T{tab};
  const double2 a = reinterpret_cast<const double2 *>(xy)[edge_v[2 * e]];
  const double2 b = reinterpret_cast<const double2 *>(xy)[edge_v[2 * e + 1]];
  double hk[NQE];
#pragma unroll
  for (int k = 0; k < NQE; ++k) hk[k] = edge_coeff<NQE>(h, e, k, a, b, T.px(k));
  out[t] = edge_load<NLE, NQE>(T, seg_length(a, b), hk, i);
}

void free_edge_records(lfgpu_plan *p) {
  for (void *q : {(void *)p->erec_ptr, (void *)p->erec_key, (void *)p->erec_meta,
                  (void *)p->vrec_ptr, (void *)p->vrec_key})
    dfree(q);
  p->erec_ptr = p->erec_key = p->vrec_ptr = p->vrec_key = nullptr;
  p->erec_meta = nullptr;
  p->edges = nullptr;
  p->edge_mat = p->edge_vec = false;
}

}  // namespace

namespace lfgpu {
void plan_free_edges(lfgpu_plan *p) { free_edge_records(p); }
}  // namespace lfgpu

extern "C" int lfgpu_edges_create(lfgpu_ctx *ctx, const lfgpu_mesh *m, int degree,
                                  int64_t nedges, const int32_t *edge_v,
                                  const int64_t *edge_dofs, lfgpu_edges **out) {
  LFG_REQUIRE(ctx && m && out && (edge_v || nedges == 0), LFGPU_EINVAL,
              "NULL argument to lfgpu_edges_create");
  *out = nullptr;
  LFG_REQUIRE(degree == 1 || degree == 3, LFGPU_ENORULE,
              "No local shape function information for this degree (only 1 and 3)");
  LFG_REQUIRE(m->nloc == (degree == 1 ? 3 : 10), LFGPU_EINVAL,
              "edge degree and the mesh's DOF layout disagree");
  LFG_REQUIRE(nedges >= 0 && nedges < (1LL << 28), LFGPU_ERANGE, "edge count out of range");
  LFG_REQUIRE(degree == 1 || edge_dofs, LFGPU_EINVAL, "degree 3 edges need edge_dofs");
  const int nle = degree + 1, nqe = degree + 1;
  std::vector<int32_t> dofs((size_t)nedges * nle);
  for (int64_t e = 0; e < nedges; ++e) {
    for (int q = 0; q < 2; ++q)
      LFG_REQUIRE(edge_v[2 * e + q] >= 0 && edge_v[2 * e + q] < m->nv, LFGPU_EINVAL,
                  "edge vertex index out of range");
    for (int i = 0; i < nle; ++i) {
      const int64_t d = edge_dofs ? edge_dofs[e * nle + i] : edge_v[2 * e + i];
      LFG_REQUIRE(d >= 0 && d < m->ndofs, LFGPU_EINVAL, "edge DOF index out of range");
      dofs[(size_t)e * nle + i] = (int32_t)d;
    }
  }
  std::vector<double> tab(2 * nqe + nle * nqe);
  const double *P = degree == 1 ? kGL2Pts : kGL4Pts;
  const double *W = degree == 1 ? kGL2W : kGL4W;
  for (int k = 0; k < nqe; ++k) {
    tab[k] = W[k];
    tab[nqe + k] = P[k];
    segment_shape(degree, P[k], tab.data() + 2 * nqe + nle * k);
  }
  lfgpu_edges *E = new lfgpu_edges();
  E->ctx = ctx;
  E->mesh = m;
  E->degree = degree;
  E->nle = nle;
  E->nqe = nqe;
  E->nedges = nedges;
  E->h_dofs = std::move(dofs);
  int rc = dalloc(&E->edge_v, (size_t)std::max<int64_t>(1, 2 * nedges), "edge vertices");
  if (rc == LFGPU_OK)
    rc = dalloc(&E->edge_dofs, (size_t)std::max<int64_t>(1, nle * nedges), "edge DOFs");
  if (rc == LFGPU_OK) rc = dalloc(&E->tab, tab.size(), "segment tables");
  if (rc != LFGPU_OK) {
    lfgpu_edges_destroy(E);
    return rc;
  }
  hipError_t e = hipSuccess;
  if (nedges > 0) {
    e = hipMemcpyAsync(E->edge_v, edge_v, 2 * nedges * sizeof(int32_t), hipMemcpyHostToDevice,
                       ctx->stream);
    if (e == hipSuccess)
      e = hipMemcpyAsync(E->edge_dofs, E->h_dofs.data(), E->h_dofs.size() * sizeof(int32_t),
                         hipMemcpyHostToDevice, ctx->stream);
  }
  if (e == hipSuccess)
    e = hipMemcpyAsync(E->tab, tab.data(), tab.size() * sizeof(double), hipMemcpyHostToDevice,
                       ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  if (e != hipSuccess) {
    lfgpu_edges_destroy(E);
    return hip_fail(e, "lfgpu_edges_create");
  }
  *out = E;
  return LFGPU_OK;
}

extern "C" int lfgpu_edges_destroy(lfgpu_edges *E) {
  if (!E) return LFGPU_OK;
  if (E->ctx) (void)hipStreamSynchronize(E->ctx->stream);
  dfree(E->edge_v);
  dfree(E->edge_dofs);
  dfree(E->tab);
  delete E;
  return LFGPU_OK;
}

extern "C" int lfgpu_edge_matrices(lfgpu_ctx *ctx, const lfgpu_edges *E, const lfgpu_coeff *eta,
                                   lfgpu_vec *out) {
  LFG_REQUIRE(ctx && E && eta && out, LFGPU_EINVAL, "NULL argument to lfgpu_edge_matrices");
  LFG_REQUIRE(out->n == E->nedges * E->nle * E->nle, LFGPU_EINVAL,
              "output must hold nedges*nle*nle values");
  CoeffD c;
  LFG_TRY(resolve_coeff(eta, E->nedges, E->nqe, false, &c, "eta"));
  if (E->nedges == 0) return LFGPU_OK;
  const unsigned g = grid_for(E->nedges * E->nle, 256);
  if (E->nle == 2)
    hipLaunchKernelGGL((k_edge_matrices<2, 2>), dim3(g), dim3(256), 0, ctx->stream, E->mesh->xy,
                       E->edge_v, E->nedges, E->tab, c, out->d);
  else
    hipLaunchKernelGGL((k_edge_matrices<4, 4>), dim3(g), dim3(256), 0, ctx->stream, E->mesh->xy,
                       E->edge_v, E->nedges, E->tab, c, out->d);
  LFG_HIP(hipGetLastError());
  return LFGPU_OK;
}

extern "C" int lfgpu_edge_vectors(lfgpu_ctx *ctx, const lfgpu_edges *E, const lfgpu_coeff *h,
                                  lfgpu_vec *out) {
  LFG_REQUIRE(ctx && E && h && out, LFGPU_EINVAL, "NULL argument to lfgpu_edge_vectors");
  LFG_REQUIRE(out->n == E->nedges * E->nle, LFGPU_EINVAL, "output must hold nedges*nle values");
  CoeffD c;
  LFG_TRY(resolve_coeff(h, E->nedges, E->nqe, false, &c, "h"));
  if (E->nedges == 0) return LFGPU_OK;
  const unsigned g = grid_for(E->nedges * E->nle, 256);
  if (E->nle == 2)
    hipLaunchKernelGGL((k_edge_vectors<2, 2>), dim3(g), dim3(256), 0, ctx->stream, E->mesh->xy,
                       E->edge_v, E->nedges, E->tab, c, out->d);
  else
    hipLaunchKernelGGL((k_edge_vectors<4, 4>), dim3(g), dim3(256), 0, ctx->stream, E->mesh->xy,
                       E->edge_v, E->nedges, E->tab, c, out->d);
  LFG_HIP(hipGetLastError());
  return LFGPU_OK;
}

extern "C" int lfgpu_plan_set_edges(lfgpu_plan *p, const lfgpu_edges *E, const uint8_t *mat_flags,
                                    const uint8_t *vec_flags) {
  LFG_REQUIRE(p && E, LFGPU_EINVAL, "NULL argument to lfgpu_plan_set_edges");
  LFG_REQUIRE(E->mesh == p->mesh, LFGPU_EINVAL, "edges and plan belong to different meshes");
  lfgpu_ctx *ctx = p->ctx;
  hipStream_t s = ctx->stream;
  const int64_t nd = p->mesh->ndofs;
  const int nle = E->nle;
  // host copies of the pattern and the fixed flags
  std::vector<int32_t> rp(nd + 1), col;
  std::vector<uint8_t> fixed(nd);
  LFG_HIP(hipMemcpyAsync(rp.data(), p->csr.row_ptr, (nd + 1) * sizeof(int32_t),
                         hipMemcpyDeviceToHost, s));
  LFG_HIP(hipStreamSynchronize(s));
  col.resize((size_t)rp[nd]);
  if (!col.empty())
    LFG_HIP(hipMemcpyAsync(col.data(), p->csr.col, col.size() * sizeof(int32_t),
                           hipMemcpyDeviceToHost, s));
  if (nd > 0)
    LFG_HIP(hipMemcpyAsync(fixed.data(), p->fixed, nd, hipMemcpyDeviceToHost, s));
  LFG_HIP(hipStreamSynchronize(s));
  const std::vector<int32_t> &ed = E->h_dofs;
  // records per row in edge order (counting sort by row, edges visited in order)
  auto build = [&](const uint8_t *flags, std::vector<int32_t> &ptr, std::vector<int32_t> &key) {
    ptr.assign(nd + 1, 0);
    for (int64_t e = 0; e < E->nedges; ++e)
      if (flags[e])
        for (int i = 0; i < nle; ++i) {
          const int32_t r = ed[(size_t)e * nle + i];
          if (!fixed[r]) ptr[r + 1]++;
        }
    for (int64_t r = 0; r < nd; ++r) ptr[r + 1] += ptr[r];
    key.assign((size_t)ptr[nd], 0);
    std::vector<int32_t> cur(ptr.begin(), ptr.end() - 1);
    for (int64_t e = 0; e < E->nedges; ++e)
      if (flags[e])
        for (int i = 0; i < nle; ++i) {
          const int32_t r = ed[(size_t)e * nle + i];
          if (!fixed[r]) key[cur[r]++] = (int32_t)(e * nle + i);
        }
  };
  std::vector<int32_t> mptr, mkey, vptr, vkey;
  std::vector<uint8_t> meta;
  if (mat_flags) {
    build(mat_flags, mptr, mkey);
    meta.assign(mkey.size() * 4, kDrop);
    for (int64_t r = 0; r < nd; ++r)
      for (int32_t q = mptr[r]; q < mptr[r + 1]; ++q) {
        const int32_t e = mkey[q] / nle;
        for (int j = 0; j < nle; ++j) {
          const int32_t c = ed[(size_t)e * nle + j];
          if (fixed[c]) continue;
          const int32_t *b = col.data() + rp[r], *en = col.data() + rp[r + 1];
          const int32_t *it = std::lower_bound(b, en, c);
          LFG_REQUIRE(it != en && *it == c, LFGPU_EINVAL,
                      "edge couples DOFs that share no cell (not in the cell pattern)");
          meta[(size_t)q * 4 + j] = (uint8_t)(it - b);
        }
      }
  }
  if (vec_flags) build(vec_flags, vptr, vkey);
  free_edge_records(p);
  auto up = [&](auto **dst, const auto &src, const char *what) -> int {
    using T = typename std::remove_reference<decltype(src)>::type::value_type;
    LFG_TRY(dalloc(dst, std::max<size_t>(1, src.size()), what));
    if (!src.empty())
      LFG_HIP(hipMemcpyAsync(*dst, src.data(), src.size() * sizeof(T), hipMemcpyHostToDevice, s));
    return LFGPU_OK;
  };
  int rc = LFGPU_OK;
  if (mat_flags) {
    if (rc == LFGPU_OK) rc = up(&p->erec_ptr, mptr, "edge records");
    if (rc == LFGPU_OK) rc = up(&p->erec_key, mkey, "edge records");
    if (rc == LFGPU_OK) rc = up(&p->erec_meta, meta, "edge records");
  }
  if (vec_flags) {
    if (rc == LFGPU_OK) rc = up(&p->vrec_ptr, vptr, "edge load records");
    if (rc == LFGPU_OK) rc = up(&p->vrec_key, vkey, "edge load records");
  }
  if (rc == LFGPU_OK) {
    hipError_t e = hipStreamSynchronize(s);
    if (e != hipSuccess) rc = hip_fail(e, "lfgpu_plan_set_edges");
  }
  if (rc != LFGPU_OK) {
    free_edge_records(p);
    return rc;
  }
  p->edges = E;
  p->edge_mat = mat_flags != nullptr;
  p->edge_vec = vec_flags != nullptr;
  return LFGPU_OK;
}
