// Device-side element computations shared by every kernel that needs
// element matrix entries: the COO-stream kernels, the fused row assembly and
// the adjoint gradient.  All of them go through the same functions, so the
// values a cell contributes are identical whichever kernel computes them.
//
// The translation unit must be compiled with -ffp-contract=off: the
// expressions below restate the reference's evaluation order and must not be
// contracted into FMAs.
#pragma once

#include <hip/hip_runtime.h>

#include "lfgpu_internal.h"

namespace lfgpu {

struct Geo {
  double jit00, jit01, jit10, jit11;  // J^{-T}, row-major
  double det;                         // |det J|
};

// assertNonDegenerateTriangle with tol = 1e-8 (geometry/tria_o1.cc:10-33).
__device__ __forceinline__ bool tria_ok(double x0, double y0, double x1, double y1,
                                        double x2, double y2) {
  const double tol = 1.0e-8;
  const double e0 = (x1 - x0) * (x1 - x0) + (y1 - y0) * (y1 - y0);
  const double e1 = (x2 - x1) * (x2 - x1) + (y2 - y1) * (y2 - y1);
  const double e2 = (x0 - x2) * (x0 - x2) + (y0 - y2) * (y0 - y2);
  const double circum = e0 + e1 + e2;
  if (!(e0 > tol * circum) || !(e1 > tol * circum) || !(e2 > tol * circum)) return false;
  const double area = fabs((x1 - x0) * (y2 - y0) - (y1 - y0) * (x2 - x0));
  return area > tol * circum;
}

// TriaO1 constructor (geometry/tria_o1.cc:57-61).  J^{-T} is Eigen's closed
// 2x2 inverse of J^T; |det J| goes through a one-step partially pivoted LU
// because J has a run-time row count in the reference.
__device__ __forceinline__ void tria_geo(double x0, double y0, double x1, double y1,
                                         double x2, double y2, Geo &g, double *Jout = nullptr) {
  const double a = x1 - x0, c = y1 - y0;  // column 0 of J
  const double b = x2 - x0, d = y2 - y0;  // column 1 of J
  if (Jout) {
    Jout[0] = a;
    Jout[1] = b;
    Jout[2] = c;
    Jout[3] = d;
  }
  const double invdet = 1.0 / (a * d - b * c);
  g.jit00 = d * invdet;
  g.jit10 = -b * invdet;
  g.jit01 = -c * invdet;
  g.jit11 = a * invdet;
  double det;
  if (fabs(c) > fabs(a)) {
    const double l = a / c;
    const double u11 = b - l * d;
    det = -1.0 * (c * u11);
  } else {
    const double l = c / a;
    const double u11 = d - l * b;
    det = 1.0 * (a * u11);
  }
  g.det = fabs(det);
}

// TriaO1::Global for one reference point (geometry/tria_o1.cc:70-74).
__device__ __forceinline__ void tria_global(double x0, double y0, double x1, double y1,
                                            double x2, double y2, double a, double b,
                                            double &x, double &y) {
  const double l0 = 1 - a - b;
  x = x0 * l0 + x1 * a + x2 * b;
  y = y0 * l0 + y1 * a + y2 * b;
}

// MeshFunction value at quadrature point k of a cell.
__device__ __forceinline__ double coeff_at(const CoeffD &c, int64_t cell, int k, int nqp,
                                           double x, double y) {
  switch (c.kind) {
    case LFGPU_COEFF_CONST:
      return c.value;
    case LFGPU_COEFF_CELL:
      return c.value * c.data[cell];
    case LFGPU_COEFF_QP:
      return c.value * c.data[cell * nqp + k];
    case LFGPU_COEFF_SINSIN:
      return c.value * (sin(M_PI * x) * sin(M_PI * y));
    default:  // LFGPU_COEFF_EXPCELL
      return c.value * exp(c.data[cell]);
  }
}

__device__ __forceinline__ bool coeff_needs_xy(const CoeffD &c) {
  return c.kind == LFGPU_COEFF_SINSIN;
}

// Table view: w[NQP] | pts[2 NQP] | N[NLOC NQP] | dN[NLOC 2 NQP], Eigen
// column-major as in PrecomputedScalarReferenceFiniteElement.
template <int NLOC, int NQP>
struct Tab {
  static constexpr int kLen = NQP + 2 * NQP + NLOC * NQP + 2 * NLOC * NQP;
  const double *t;
  __device__ __forceinline__ double w(int k) const { return t[k]; }
  __device__ __forceinline__ double px(int k) const { return t[NQP + 2 * k]; }
  __device__ __forceinline__ double py(int k) const { return t[NQP + 2 * k + 1]; }
  __device__ __forceinline__ double N(int i, int k) const { return t[3 * NQP + i + NLOC * k]; }
  __device__ __forceinline__ double dx(int i, int k) const {
    return t[3 * NQP + NLOC * NQP + i + NLOC * (2 * k)];
  }
  __device__ __forceinline__ double dy(int i, int k) const {
    return t[3 * NQP + NLOC * NQP + i + NLOC * (2 * k + 1)];
  }
};

// Row i of ReactionDiffusionElementMatrixProvider::Eval
// (uscalfe/loc_comp_ellbvp.h:324-337):
//   mat(i,j) = sum_k w_k ( G_i . (alpha_k G_j) + (gamma_k N_i) N_j ),
//   w_k = omega_k |det J|, G = J^{-T} grad N^T,
// with alpha_k = ak(k), gamma_k = gk(k).
// P1: the reference gradients of the linear Lagrange functions are the same
// at every quadrature point (the table holds equal values), so each
// G_j = J^{-T} grad N_j is formed once, with the products of the general
// loop, and reused at every point: the same values, a third of the work.
template <int NLOC, int NQP, class AK, class GK>
__device__ __forceinline__ void element_row_f(const Tab<NLOC, NQP> &T, const Geo &g,
                                              const AK &ak, const GK &gk, int i, double *row) {
#pragma unroll
  for (int j = 0; j < NLOC; ++j) row[j] = 0.0;
  if constexpr (NLOC == 3 && NQP == 3) {
    double G0[3], G1[3];
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      const double gxj = T.dx(j, 0), gyj = T.dy(j, 0);
      G0[j] = g.jit00 * gxj + g.jit01 * gyj;
      G1[j] = g.jit10 * gxj + g.jit11 * gyj;
    }
    double Gi0 = G0[0], Gi1 = G1[0];
#pragma unroll
    for (int j = 1; j < 3; ++j)
      if (i == j) {
        Gi0 = G0[j];
        Gi1 = G1[j];
      }
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      const double w = T.w(k) * g.det;
      const double a = ak(k), c = gk(k);
      const double cNi = c * T.N(i, k);
#pragma unroll
      for (int j = 0; j < 3; ++j) {
        const double s = Gi0 * (a * G0[j]) + Gi1 * (a * G1[j]);
        const double m = cNi * T.N(j, k);
        row[j] += w * (s + m);
      }
    }
  } else {
    for (int k = 0; k < NQP; ++k) {
      const double w = T.w(k) * g.det;
      const double gxi = T.dx(i, k), gyi = T.dy(i, k);
      const double Gi0 = g.jit00 * gxi + g.jit01 * gyi;
      const double Gi1 = g.jit10 * gxi + g.jit11 * gyi;
      const double a = ak(k), c = gk(k);
      const double cNi = c * T.N(i, k);
#pragma unroll
      for (int j = 0; j < NLOC; ++j) {
        const double gxj = T.dx(j, k), gyj = T.dy(j, k);
        const double Gj0 = g.jit00 * gxj + g.jit01 * gyj;
        const double Gj1 = g.jit10 * gxj + g.jit11 * gyj;
        const double s = Gi0 * (a * Gj0) + Gi1 * (a * Gj1);
        const double m = cNi * T.N(j, k);
        row[j] += w * (s + m);
      }
    }
  }
}

template <int NLOC, int NQP>
__device__ __forceinline__ void element_row(const Tab<NLOC, NQP> &T, const Geo &g,
                                            const double *ak, const double *gk, int i,
                                            double *row) {
  element_row_f<NLOC, NQP>(T, g, [&](int k) { return ak[k]; }, [&](int k) { return gk[k]; }, i,
                           row);
}

// Entry i of ScalarLoadElementVectorProvider::Eval (loc_comp_ellbvp.h:733-740),
// f_k = fk(k).
template <int NLOC, int NQP, class FK>
__device__ __forceinline__ double element_load_f(const Tab<NLOC, NQP> &T, const Geo &g,
                                                 const FK &fk, int i) {
  double v = 0.0;
  for (int k = 0; k < NQP; ++k) v += (T.w(k) * g.det * fk(k)) * T.N(i, k);
  return v;
}

template <int NLOC, int NQP>
__device__ __forceinline__ double element_load(const Tab<NLOC, NQP> &T, const Geo &g,
                                               const double *fk, int i) {
  return element_load_f<NLOC, NQP>(T, g, [&](int k) { return fk[k]; }, i);
}

// Value at point k of a coefficient of kind CONST, CELL or QP on one cell
// (the expressions of coeff_at: value, value * data[cell],
// value * data[cell * nqp + k]).
struct PlainCoef {
  double v0;
  const double *q;  // the cell's point values (QP), else null
  __device__ __forceinline__ PlainCoef(const CoeffD &c, int64_t cell, int nqp) {
    v0 = c.value;
    q = nullptr;
    if (c.kind == LFGPU_COEFF_CELL)
      v0 = c.value * c.data[cell];
    else if (c.kind == LFGPU_COEFF_QP)
      q = c.data + cell * nqp;
  }
  __device__ __forceinline__ double operator()(int k) const { return q ? v0 * q[k] : v0; }
};

// Cell data gathered once per (cell) visit.
struct CellXY {
  double x0, y0, x1, y1, x2, y2;
};

__device__ __forceinline__ CellXY load_cell(const double *xy, const int32_t *cell_v,
                                            int64_t cell) {
  const int32_t v0 = cell_v[3 * cell], v1 = cell_v[3 * cell + 1], v2 = cell_v[3 * cell + 2];
  const double2 p0 = reinterpret_cast<const double2 *>(xy)[v0];
  const double2 p1 = reinterpret_cast<const double2 *>(xy)[v1];
  const double2 p2 = reinterpret_cast<const double2 *>(xy)[v2];
  return CellXY{p0.x, p0.y, p1.x, p1.y, p2.x, p2.y};
}

// Coefficient values of one cell at all quadrature points.
template <int NQP>
__device__ __forceinline__ void coeff_qp(const CoeffD &c, int64_t cell, const CellXY &p,
                                         const double *tab, double *out) {
  if (c.kind == LFGPU_COEFF_CONST) {
#pragma unroll
    for (int k = 0; k < NQP; ++k) out[k] = c.value;
  } else if (c.kind == LFGPU_COEFF_CELL) {
    const double v = c.value * c.data[cell];
#pragma unroll
    for (int k = 0; k < NQP; ++k) out[k] = v;
  } else if (c.kind == LFGPU_COEFF_EXPCELL) {
    const double v = c.value * exp(c.data[cell]);
#pragma unroll
    for (int k = 0; k < NQP; ++k) out[k] = v;
  } else {
    for (int k = 0; k < NQP; ++k) {
      double x = 0.0, y = 0.0;
      if (c.kind == LFGPU_COEFF_SINSIN)
        tria_global(p.x0, p.y0, p.x1, p.y1, p.x2, p.y2, tab[NQP + 2 * k], tab[NQP + 2 * k + 1],
                    x, y);
      out[k] = coeff_at(c, cell, k, NQP, x, y);
    }
  }
}

// Segment tables: w[NQE] | pts[NQE] | N[NLE x NQE] with N(i, k) at i + NLE k
// (PrecomputedScalarReferenceFiniteElement of the segment element).
template <int NLE, int NQE>
struct ETab {
  static constexpr int kLen = 2 * NQE + NLE * NQE;
  const double *t;
  __device__ __forceinline__ double w(int k) const { return t[k]; }
  __device__ __forceinline__ double px(int k) const { return t[NQE + k]; }
  __device__ __forceinline__ double N(int i, int k) const { return t[2 * NQE + i + NLE * k]; }
};

// SegmentO1::IntegrationElement (geometry/segment_o1.cc): |v1 - v0|.
__device__ __forceinline__ double seg_length(double2 a, double2 b) {
  const double dx = b.x - a.x, dy = b.y - a.y;
  return sqrt(dx * dx + dy * dy);
}

// Coefficient of an edge term at point k: SegmentO1::Global(x) = v1 x +
// v0 (1 - x) for the analytic kind.
template <int NQE>
__device__ __forceinline__ double edge_coeff(const CoeffD &c, int64_t e, int k, double2 a,
                                             double2 b, double xh) {
  double x = 0.0, y = 0.0;
  if (c.kind == LFGPU_COEFF_SINSIN) {
    x = b.x * xh + a.x * (1 - xh);
    y = b.y * xh + a.y * (1 - xh);
  }
  return coeff_at(c, e, k, NQE, x, y);
}

// Row i of MassEdgeMatrixProvider::Eval (uscalfe/loc_comp_ellbvp.h:510-526):
//   mat += (N_k N_k^T) * w,  w = (omega_k |det|) gamma_k.
template <int NLE, int NQE>
__device__ __forceinline__ void edge_row(const ETab<NLE, NQE> &T, double det, const double *gk,
                                         int i, double *row) {
#pragma unroll
  for (int j = 0; j < NLE; ++j) row[j] = 0.0;
#pragma unroll
  for (int k = 0; k < NQE; ++k) {
    const double w = (T.w(k) * det) * gk[k];
    const double ni = T.N(i, k);
#pragma unroll
    for (int j = 0; j < NLE; ++j) row[j] += (ni * T.N(j, k)) * w;
  }
}

// Entry i of ScalarLoadEdgeVectorProvider::Eval (loc_comp_ellbvp.h:905-916):
//   vec += N_k * w,  w = (omega_k |det|) g_k.
template <int NLE, int NQE>
__device__ __forceinline__ double edge_load(const ETab<NLE, NQE> &T, double det, const double *hk,
                                            int i) {
  double v = 0.0;
#pragma unroll
  for (int k = 0; k < NQE; ++k) v += T.N(i, k) * ((T.w(k) * det) * hk[k]);
  return v;
}

// The value coeff_qp<NQP> stores at point k, for kernels that spread the
// points of one cell over several lanes.
template <int NQP>
__device__ __forceinline__ double coeff_one(const CoeffD &c, int64_t cell, const CellXY &p,
                                            const double *tab, int k) {
  if (c.kind == LFGPU_COEFF_CONST) return c.value;
  if (c.kind == LFGPU_COEFF_CELL) return c.value * c.data[cell];
  if (c.kind == LFGPU_COEFF_EXPCELL) return c.value * exp(c.data[cell]);
  double x = 0.0, y = 0.0;
  if (c.kind == LFGPU_COEFF_SINSIN)
    tria_global(p.x0, p.y0, p.x1, p.y1, p.x2, p.y2, tab[NQP + 2 * k], tab[NQP + 2 * k + 1], x, y);
  return coeff_at(c, cell, k, NQP, x, y);
}

}  // namespace lfgpu
