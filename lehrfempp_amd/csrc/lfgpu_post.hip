// Post-processing: error norms of a finite element solution.
//
// The demo drivers finish with (examples/ellbvp_linfe/homDir_linfe_demo.cc:
// 179-188)
//   L2err  = sqrt(IntegrateMeshFunction(mesh, squaredNorm(mf_sol - mf_u), q))
//   H1serr = sqrt(IntegrateMeshFunction(mesh, squaredNorm(mf_grad_sol -
//                                                        mf_grad_u), q))
// with MeshFunctionFE = local_dofs * N (fe/mesh_function_fe.h:54-70),
// MeshFunctionGradFE = J^{-T} (local_dofs * grad N) (fe/mesh_function_grad_fe.h)
// and LocalIntegral = sum_k value_k (w_k |det J|) (fe/fe_tools.h:34-66).  One
// thread per cell evaluates both integrands at the finite element's own
// quadrature points; per-workgroup partials are summed by a single workgroup
// in a fixed order (deterministic; the reference adds cells one by one, so the
// totals agree to rounding).
#include <hip/hip_runtime.h>

#include <algorithm>

#include "lfgpu_fe_device.h"
#include "lfgpu_internal.h"
#include "lfgpu_reduce.h"

using namespace lfgpu;

namespace {

constexpr int kBlocks = 1024;

template <int NLOC, int NQP>
__global__ __launch_bounds__(256) void k_error_norms(const double *xy, const int32_t *cell_v,
                                                     const int32_t *cell_dofs, int64_t ncells,
                                                     const double *tab, const double *uh,
                                                     int exact, const double *u_qp,
                                                     const double *g_qp, double *part) {
  __shared__ double s_tab[Tab<NLOC, NQP>::kLen];
  __shared__ double sh[8];
  for (int t = threadIdx.x; t < Tab<NLOC, NQP>::kLen; t += 256) s_tab[t] = tab[t];
  __syncthreads();
  const Tab<NLOC, NQP> T{s_tab};
  double l2 = 0.0, h1 = 0.0;
  for (int64_t c = (int64_t)blockIdx.x * 256 + threadIdx.x; c < ncells;
       c += (int64_t)gridDim.x * 256) {
    const CellXY p = load_cell(xy, cell_v, c);
    Geo g;
    tria_geo(p.x0, p.y0, p.x1, p.y1, p.x2, p.y2, g);
    double u[NLOC];
#pragma unroll
    for (int i = 0; i < NLOC; ++i) u[i] = uh[cell_dofs[c * NLOC + i]];
    double cl2 = 0.0, ch1 = 0.0;
    for (int k = 0; k < NQP; ++k) {
      double v = 0.0, r0 = 0.0, r1 = 0.0;
#pragma unroll
      for (int i = 0; i < NLOC; ++i) {
        v += u[i] * T.N(i, k);
        r0 += u[i] * T.dx(i, k);
        r1 += u[i] * T.dy(i, k);
      }
      const double g0 = g.jit00 * r0 + g.jit01 * r1;
      const double g1 = g.jit10 * r0 + g.jit11 * r1;
      double ue, ge0, ge1;
      if (exact == LFGPU_EXACT_SINSIN) {
        double x, y;
        tria_global(p.x0, p.y0, p.x1, p.y1, p.x2, p.y2, T.px(k), T.py(k), x, y);
        const double sx = sin(M_PI * x), sy = sin(M_PI * y);
        ue = sx * sy;
        ge0 = M_PI * (cos(M_PI * x) * sy);
        ge1 = M_PI * (sx * cos(M_PI * y));
      } else {
        ue = u_qp[c * NQP + k];
        ge0 = g_qp[(c * NQP + k) * 2];
        ge1 = g_qp[(c * NQP + k) * 2 + 1];
      }
      const double wk = T.w(k) * g.det;
      const double d = v - ue, e0 = g0 - ge0, e1 = g1 - ge1;
      cl2 += (d * d) * wk;
      ch1 += (e0 * e0 + e1 * e1) * wk;
    }
    l2 += cl2;
    h1 += ch1;
  }
  l2 = block_sum<256>(l2, sh);
  __syncthreads();
  h1 = block_sum<256>(h1, sh);
  if (threadIdx.x == 0) {
    part[blockIdx.x] = l2;
    part[kBlocks + blockIdx.x] = h1;
  }
}

__global__ __launch_bounds__(256) void k_error_norms_total(const double *part, int nblocks,
                                                           double *out) {
  __shared__ double sh[8];
  double l2 = 0.0, h1 = 0.0;
  for (int i = threadIdx.x; i < nblocks; i += 256) {
    l2 += part[i];
    h1 += part[kBlocks + i];
  }
  l2 = block_sum<256>(l2, sh);
  __syncthreads();
  h1 = block_sum<256>(h1, sh);
  if (threadIdx.x == 0) {
    out[0] = l2;
    out[1] = h1;
  }
}

}  // namespace

extern "C" int lfgpu_error_norms(lfgpu_ctx *ctx, const lfgpu_mesh *m, const lfgpu_fe *fe,
                                 const lfgpu_vec *uh, int exact, const lfgpu_vec *u_qp,
                                 const lfgpu_vec *grad_qp, double *l2sq, double *h1sq) {
  LFG_REQUIRE(ctx && m && fe && uh && l2sq && h1sq, LFGPU_EINVAL,
              "NULL argument to lfgpu_error_norms");
  LFG_REQUIRE(fe->nloc == m->nloc, LFGPU_EINVAL,
              "finite element and DOF layout disagree on the number of local DOFs");
  LFG_REQUIRE(uh->n == m->ndofs, LFGPU_EINVAL, "uh must have ndofs entries");
  LFG_REQUIRE(exact == LFGPU_EXACT_SINSIN || exact == LFGPU_EXACT_QP, LFGPU_EINVAL,
              "unknown exact-solution kind");
  if (exact == LFGPU_EXACT_QP)
    LFG_REQUIRE(u_qp && grad_qp && u_qp->n == m->ncells * fe->nqp &&
                    grad_qp->n == 2 * m->ncells * fe->nqp,
                LFGPU_EINVAL, "u_qp needs ncells*nqp and grad_qp 2*ncells*nqp values");
  double *buf = nullptr;
  LFG_TRY(dalloc(&buf, (size_t)2 * kBlocks + 2, "error-norm partials"));
  hipStream_t s = ctx->stream;
  const int64_t need = (m->ncells + 255) / 256;
  const int nb = (int)std::max<int64_t>(1, std::min<int64_t>(need, kBlocks));
  const double *uq = exact == LFGPU_EXACT_QP ? u_qp->d : nullptr;
  const double *gq = exact == LFGPU_EXACT_QP ? grad_qp->d : nullptr;
  if (fe->nloc == 3)
    hipLaunchKernelGGL((k_error_norms<3, 3>), dim3(nb), dim3(256), 0, s, m->xy, m->cell_v,
                       m->cell_dofs, m->ncells, fe->tab, uh->d, exact, uq, gq, buf);
  else
    hipLaunchKernelGGL((k_error_norms<10, 12>), dim3(nb), dim3(256), 0, s, m->xy, m->cell_v,
                       m->cell_dofs, m->ncells, fe->tab, uh->d, exact, uq, gq, buf);
  hipLaunchKernelGGL(k_error_norms_total, dim3(1), dim3(256), 0, s, buf, nb, buf + 2 * kBlocks);
  double host[2] = {0.0, 0.0};
  hipError_t e = hipGetLastError();
  if (e == hipSuccess)
    e = hipMemcpyAsync(host, buf + 2 * kBlocks, 2 * sizeof(double), hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  dfree(buf);
  if (e != hipSuccess) return hip_fail(e, "lfgpu_error_norms");
  *l2sq = host[0];
  *h1sq = host[1];
  return LFGPU_OK;
}
