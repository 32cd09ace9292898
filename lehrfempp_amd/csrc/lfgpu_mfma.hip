// Element matrices as a reference-tensor contraction on the matrix cores.
//
// ReactionDiffusionElementMatrixProvider::Eval (uscalfe/loc_comp_ellbvp.h:
// 266-339) sums over the quadrature points k
//   w_k |det J| ( G_i . (alpha_k G_j) + gamma_k N_i N_j ),   G = J^{-T} grad N^.
// J is constant on an affine triangle, so with M = J^{-T} and B = M^T M
//   G_i . G_j = B00 dx_i dx_j + B11 dy_i dy_j + B01 (dx_i dy_j + dy_i dx_j)
// and the element matrix is a short linear combination of fixed reference
// matrices:
//   K_e = sum_{k,t} c_{e,k,t} R_{k,t},
//   c_{e,k,.} = |det| (alpha_k B00, alpha_k B11, alpha_k B01, gamma_k),
//   R_{k,.}(i,j) = w_k (dx_i dx_j, dy_i dy_j, dx_i dy_j + dy_i dx_j, N_i N_j).
// When alpha and gamma are constant on each cell the point sums fold into R
// and four terms remain.  For a batch of 16 cells the matrices are the rows
// of a (16 x 4T) by (4T x nloc^2) product, which a wavefront computes with one
// v_mfma_f64_16x16x4f64 per term quadruple and 16 matrix entries.  Geometry
// and coefficients stay FP64 on the vector units.
//
// The entries agree with the reference's evaluation order up to rounding
// (a few 1e-16 of the row magnitude).  The element vectors follow the
// reference's own order (loc_comp_ellbvp.h:733-740) and are bit-identical to
// the ones the row-owner assembly computes.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "lfgpu_fe_device.h"
#include "lfgpu_internal.h"

using namespace lfgpu;

namespace {

typedef double v4d __attribute__((ext_vector_type(4)));

constexpr int kBlock = 256;          // four wavefronts
constexpr int kCellsPerBlock = 64;   // 16 cells per wavefront
constexpr unsigned kMaxGrid = 65536;

// Operand layout of v_mfma_f64_16x16x4f64 (64 lanes, l = lane):
//   A (16 x 4):  lane l holds A[l % 16][l / 16]
//   B (4 x 16):  lane l holds B[l / 16][l % 16]
//   D (16 x 16): lane l, component v holds D[l / 16 + 4 v][l % 16]
// Here A rows are cells, A columns / B rows are terms and B / D columns are
// element-matrix entries (row-major i * nloc + j).
template <int NLOC, int NQP, int KT>
__global__ __launch_bounds__(kBlock) void k_element_contract(
    const double *xy, const int32_t *cell_v, int64_t ncells, const double *tab,
    const double *rten, CoeffD alpha, CoeffD gamma, CoeffD f, int has_f, double *ke,
    double *fvec, int rowmajor_i) {
  constexpr int NE = NLOC * NLOC;
  constexpr int NEP = (NE + 15) / 16 * 16;
  constexpr int NT = 4 * KT;          // contraction length
  constexpr int NQ = KT > 1 ? 3 : 1;  // staged per-point values: [alpha, gamma,] f
  __shared__ double s_r[NT * NEP];
  __shared__ double s_tab[Tab<NLOC, NQP>::kLen];
  __shared__ double s_q[kBlock / 64][NQ][16][NQP];
  // rows 0..3 of rten hold the folded sums, rows 4.. the per-point terms
  const double *src = rten + (KT == 1 ? 0 : 4 * NEP);
  for (int t = threadIdx.x; t < NT * NEP; t += kBlock) s_r[t] = src[t];
  for (int t = threadIdx.x; t < Tab<NLOC, NQP>::kLen; t += kBlock) s_tab[t] = tab[t];
  __syncthreads();
  const Tab<NLOC, NQP> T{s_tab};
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int ci = lane & 15, q = lane >> 4;
  double(*sq)[16][NQP] = s_q[wave];
  for (int64_t base = (int64_t)blockIdx.x * kCellsPerBlock; base < ncells;
       base += (int64_t)gridDim.x * kCellsPerBlock) {
    const int64_t wbase = base + 16 * wave;
    const int64_t cell = wbase + ci;
    // lanes past the end work on the last cell and store nothing
    const int64_t cc = cell < ncells ? cell : ncells - 1;
    const CellXY p = load_cell(xy, cell_v, cc);
    Geo g;
    tria_geo(p.x0, p.y0, p.x1, p.y1, p.x2, p.y2, g);
    // per-point values, points k = q, q + 4, ... computed by lane group q
    for (int k = q; k < NQP; k += 4) {
      if constexpr (KT > 1) {
        sq[0][ci][k] = coeff_one<NQP>(alpha, cc, p, s_tab, k);
        sq[1][ci][k] = coeff_one<NQP>(gamma, cc, p, s_tab, k);
      }
      if (has_f) sq[NQ - 1][ci][k] = coeff_one<NQP>(f, cc, p, s_tab, k);
    }
    __syncthreads();
    if (has_f && cell < ncells) {
      for (int i = q; i < NLOC; i += 4) {
        double v = 0.0;
        for (int k = 0; k < NQP; ++k) v += (T.w(k) * g.det * sq[NQ - 1][ci][k]) * T.N(i, k);
        fvec[cell * NLOC + i] = v;
      }
    }
    const double bq = q == 0   ? g.jit00 * g.jit00 + g.jit10 * g.jit10
                      : q == 1 ? g.jit01 * g.jit01 + g.jit11 * g.jit11
                               : g.jit00 * g.jit01 + g.jit10 * g.jit11;
    double av[KT];
    if constexpr (KT == 1) {
      av[0] = q == 3 ? g.det * coeff_one<NQP>(gamma, cc, p, s_tab, 0)
                     : g.det * coeff_one<NQP>(alpha, cc, p, s_tab, 0) * bq;
    } else {
#pragma unroll
      for (int k = 0; k < KT; ++k)
        av[k] = q == 3 ? g.det * sq[1][ci][k] : g.det * sq[0][ci][k] * bq;
    }
#pragma unroll
    for (int ct = 0; ct < NEP / 16; ++ct) {
      v4d acc = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
      for (int k = 0; k < KT; ++k)
        acc = __builtin_amdgcn_mfma_f64_16x16x4f64(av[k], s_r[(4 * k + q) * NEP + ct * 16 + ci],
                                                   acc, 0, 0, 0);
      const int e = ct * 16 + ci;
      if (e < NE) {
        // cell-major: ke[cell][i][j]; rowmajor_i: ke[i][cell][j], so the rows
        // i of consecutive cells are adjacent for the row-owner gather
        const int ei = e / NLOC, ej = e - ei * NLOC;
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int64_t rc = wbase + q + 4 * v;
          if (rc < ncells)
            ke[rowmajor_i ? ((int64_t)ei * ncells + rc) * NLOC + ej : rc * NE + e] = acc[v];
        }
      }
    }
    __syncthreads();  // sq is rewritten by the next batch
  }
}

bool cellwise(const CoeffD &c) {
  return c.kind == LFGPU_COEFF_CONST || c.kind == LFGPU_COEFF_CELL ||
         c.kind == LFGPU_COEFF_EXPCELL;
}

template <int NLOC, int NQP>
int launch_contract(hipStream_t s, const lfgpu_mesh *m, const lfgpu_fe *fe, const CoeffD &a,
                    const CoeffD &g, const CoeffD *f, double *ke, double *fvec, int rowmajor_i) {
  const int64_t nb = (m->ncells + kCellsPerBlock - 1) / kCellsPerBlock;
  const unsigned grid = (unsigned)std::max<int64_t>(1, std::min<int64_t>(nb, kMaxGrid));
  const CoeffD fz = f ? *f : CoeffD{LFGPU_COEFF_CONST, 0.0, nullptr};
  if (cellwise(a) && cellwise(g))
    hipLaunchKernelGGL((k_element_contract<NLOC, NQP, 1>), dim3(grid), dim3(kBlock), 0, s, m->xy,
                       m->cell_v, m->ncells, fe->tab, fe->rten, a, g, fz, f ? 1 : 0, ke, fvec,
                       rowmajor_i);
  else
    hipLaunchKernelGGL((k_element_contract<NLOC, NQP, NQP>), dim3(grid), dim3(kBlock), 0, s,
                       m->xy, m->cell_v, m->ncells, fe->tab, fe->rten, a, g, fz, f ? 1 : 0, ke,
                       fvec, rowmajor_i);
  LFG_HIP(hipGetLastError());
  return LFGPU_OK;
}

}  // namespace

namespace lfgpu {

int element_contract(hipStream_t s, const lfgpu_mesh *m, const lfgpu_fe *fe, const CoeffD &alpha,
                     const CoeffD &gamma, const CoeffD *f, double *ke, double *fvec,
                     bool rowmajor_i) {
  if (m->ncells == 0) return LFGPU_OK;
  const int rm = rowmajor_i ? 1 : 0;
  if (fe->nloc == 3) return launch_contract<3, 3>(s, m, fe, alpha, gamma, f, ke, fvec, rm);
  return launch_contract<10, 12>(s, m, fe, alpha, gamma, f, ke, fvec, rm);
}

}  // namespace lfgpu

extern "C" int lfgpu_element_matrices_mfma(lfgpu_ctx *ctx, const lfgpu_mesh *m,
                                           const lfgpu_fe *fe, const lfgpu_coeff *alpha,
                                           const lfgpu_coeff *gamma, lfgpu_vec *out) {
  LFG_REQUIRE(ctx && m && fe && out, LFGPU_EINVAL, "NULL argument to lfgpu_element_matrices_mfma");
  LFG_REQUIRE(fe->nloc == m->nloc, LFGPU_EINVAL,
              "finite element and DOF layout disagree on the number of local DOFs");
  LFG_REQUIRE(out->n == m->ncells * m->nloc * m->nloc, LFGPU_EINVAL,
              "output must hold ncells*nloc*nloc values");
  CoeffD a, g;
  LFG_TRY(resolve_coeff(alpha, m->ncells, fe->nqp, false, &a, "alpha"));
  LFG_TRY(resolve_coeff(gamma, m->ncells, fe->nqp, false, &g, "gamma"));
  return element_contract(ctx->stream, m, fe, a, g, nullptr, out->d, nullptr, false);
}
