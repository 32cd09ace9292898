// lfref.cc -- CPU oracle (TEST INFRASTRUCTURE ONLY; see lfref.h).
//
// A from-scratch C++17 restatement of the LehrFEM++ hot path without Eigen.
// Each function names the reference lines it follows.  The expression order
// of the element computations follows the reference's Eigen expressions as
// written; the rounding of Eigen-internal kernels (not present in the
// reference tree) is restated from Eigen 3.4.0's documented semantics.
#include "lfref.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <utility>
#include <vector>

namespace {

// ---------------------------------------------------------------------------
// Quadrature (quad/quad_rules_tria.cc:35-51 and :113-156, literal constants).
// ---------------------------------------------------------------------------
const double kQ2Pts[6] = {0.1666666666666666019,  0.16666666666666679619,
                          0.66666666666666651864, 0.16666666666666665741,
                          0.1666666666666668517,  0.66666666666666674068};
const double kQ2W[3] = {0.16666666666666665741, 0.16666666666666665741,
                        0.16666666666666665741};
const double kQ6Pts[24] = {
    0.21942998254978290928,  0.21942998254978310357,  0.56114003490043395939,
    0.2194299825497829648,   0.21942998254978302031,  0.56114003490043418143,
    0.48013796411221504323,  0.039724071775569962106, 0.48013796411221509874,
    0.48013796411221504323,  0.039724071775569885778, 0.48013796411221526528,
    0.14161901592396813898,  0.019371724361240811457, 0.8390092597147910114,
    0.14161901592396811123,  0.019371724361240960643, 0.83900925971479112242,
    0.83900925971479112242,  0.019371724361240811457, 0.14161901592396813898,
    0.83900925971479123344,  0.019371724361240558188, 0.14161901592396844429};
const double kQ6W[12] = {
    0.085666562076490496147, 0.085666562076490496147, 0.085666562076490496147,
    0.040365544796515481774, 0.040365544796515481774, 0.040365544796515481774,
    0.020317279896830325869, 0.020317279896830325869, 0.020317279896830325869,
    0.020317279896830325869, 0.020317279896830325869, 0.020317279896830325869};

struct Rule {
  int nqp = 0;
  const double *pts = nullptr;
  const double *w = nullptr;
};

// make_QuadRule(kTria, 2*degree) (uscalfe/loc_comp_ellbvp.h:227-228,
// quad/make_quad_rule.cc:38-50).
Rule rule_for_fe_degree(int degree) {
  Rule r;
  if (degree == 1) {
    r.nqp = 3;
    r.pts = kQ2Pts;
    r.w = kQ2W;
  } else if (degree == 3) {
    r.nqp = 12;
    r.pts = kQ6Pts;
    r.w = kQ6W;
  }
  return r;
}

// ---------------------------------------------------------------------------
// Shape functions (uscalfe/lagr_fe.h:104-129 for O1, :1010-1101 for O3).
// N[i + nsf*k], dN[i + nsf*(2k+d)].
// ---------------------------------------------------------------------------
void p1_shape(int npts, const double *pts, double *N, double *dN) {
  for (int k = 0; k < npts; ++k) {
    const double a = pts[2 * k], b = pts[2 * k + 1];
    N[0 + 3 * k] = 1.0 - a - b;  // Ones - row(0) - row(1)
    N[1 + 3 * k] = a;
    N[2 + 3 * k] = b;
    dN[0 + 3 * (2 * k)] = -1.0;
    dN[0 + 3 * (2 * k + 1)] = -1.0;
    dN[1 + 3 * (2 * k)] = 1.0;
    dN[1 + 3 * (2 * k + 1)] = 0.0;
    dN[2 + 3 * (2 * k)] = 0.0;
    dN[2 + 3 * (2 * k + 1)] = 1.0;
  }
}

void p3_shape(int npts, const double *pts, double *N, double *dN) {
  const double t1 = 1 / 3.0, t2 = 2 / 3.0;
  for (int k = 0; k < npts; ++k) {
    const double l1 = pts[2 * k], l2 = pts[2 * k + 1];
    const double l0 = 1 - l1 - l2;
    double *v = N + 10 * k;
    v[0] = 4.5 * l0 * (l0 - t1) * (l0 - t2);
    v[1] = 4.5 * l1 * (l1 - t1) * (l1 - t2);
    v[2] = 4.5 * l2 * (l2 - t1) * (l2 - t2);
    v[3] = 13.5 * l0 * l1 * (l0 - t1);
    v[4] = 13.5 * l0 * l1 * (l1 - t1);
    v[5] = 13.5 * l1 * l2 * (l1 - t1);
    v[6] = 13.5 * l1 * l2 * (l2 - t1);
    v[7] = 13.5 * l2 * l0 * (l2 - t1);
    v[8] = 13.5 * l2 * l0 * (l0 - t1);
    v[9] = 27.0 * l0 * l1 * l2;
    // temp.row(r) for r < 10 is d/dx1 of shape function r, temp.row(10+r)
    // is d/dx2 (the 20 x npts reshape of lagr_fe.h:1063-1099).
    double *gx = dN + 10 * (2 * k);
    double *gy = dN + 10 * (2 * k + 1);
    gx[0] = -4.5 * ((l0 - t1) * (l0 - t2) + l0 * (l0 - t2) + l0 * (l0 - t1));
    gy[0] = -4.5 * ((l0 - t1) * (l0 - t2) + l0 * (l0 - t2) + l0 * (l0 - t1));
    gx[1] = 4.5 * ((l1 - t1) * (l1 - t2) + l1 * (l1 - t2) + l1 * (l1 - t1));
    gy[1] = 0.0;
    gx[2] = 0.0;
    gy[2] = 4.5 * ((l2 - t1) * (l2 - t2) + l2 * (l2 - t2) + l2 * (l2 - t1));
    gx[3] = 13.5 * (-l1 * (l0 - t1) + l0 * (l0 - t1) - l0 * l1);
    gy[3] = -13.5 * (l1 * (l0 - t1) + l0 * l1);
    gx[4] = 13.5 * (-l1 * (l1 - t1) + l0 * (l1 - t1) + l0 * l1);
    gy[4] = -13.5 * l1 * (l1 - t1);
    gx[5] = 13.5 * (l2 * (l1 - t1) + l1 * l2);
    gy[5] = 13.5 * (l1 * (l1 - t1));
    gx[6] = 13.5 * (l2 * (l2 - t1));
    gy[6] = 13.5 * (l1 * (l2 - t1) + l1 * l2);
    gx[7] = -13.5 * l2 * (l2 - t1);
    gy[7] = 13.5 * (l0 * (l2 - t1) - l2 * (l2 - t1) + l0 * l2);
    gx[8] = -13.5 * (l2 * (l0 - t1) + l2 * l0);
    gy[8] = 13.5 * (l0 * (l0 - t1) - l2 * (l0 - t1) - l2 * l0);
    gx[9] = 27.0 * (-l1 * l2 + l0 * l2);
    gy[9] = 27.0 * (-l1 * l2 + l0 * l1);
  }
}

struct FeTables {
  int nsf = 0, nqp = 0;
  Rule rule;
  std::vector<double> N, dN;
};

bool make_tables(int degree, FeTables &t) {
  t.rule = rule_for_fe_degree(degree);
  if (t.rule.nqp == 0) return false;
  t.nqp = t.rule.nqp;
  t.nsf = degree == 1 ? 3 : 10;
  t.N.assign(t.nsf * t.nqp, 0.0);
  t.dN.assign(t.nsf * 2 * t.nqp, 0.0);
  if (degree == 1)
    p1_shape(t.nqp, t.rule.pts, t.N.data(), t.dN.data());
  else
    p3_shape(t.nqp, t.rule.pts, t.N.data(), t.dN.data());
  return true;
}

// ---------------------------------------------------------------------------
// Segments (codim-1 entities): Gauss-Legendre rule, Lagrange shape functions,
// SegmentO1 geometry.
// ---------------------------------------------------------------------------
// GaussLegendre(n) on [0, 1] (quad/gauss_quadrature.cc:19-50): Newton on the
// Legendre recurrence from z = cos(pi (i + 3/4) / (n + 1/2)), then
// points 0.5 (1 -+ z), weights 1 / ((1 - z^2) p'^2).  The reference iterates in
// 57-bit binary floating point; this restatement uses long double (64-bit
// mantissa on x86), so the doubles agree to rounding (the reference's own test
// pins them to 4 ulps, quad/test/gauss_legendre_tests.cc:14-57).
void gauss_legendre01(int n, double *pts, double *w) {
  const long double pi = 3.141592653589793238462643383279502884L;
  const int m = (n + 1) / 2;
  for (int i = 0; i < m; ++i) {
    long double z = std::cos(pi * (i + 0.75L) / (n + 0.5L)), z1, pp;
    int guard = 0;
    do {
      long double p1 = 1.0L, p2 = 0.0L, p3;
      for (int j = 0; j < n; ++j) {
        p3 = p2;
        p2 = p1;
        p1 = ((2.0L * j + 1.0L) * z * p2 - j * p3) / (j + 1.0L);
      }
      pp = n * (z * p1 - p2) / (z * z - 1.0L);
      z1 = z;
      z = z1 - p1 / pp;
    } while (std::fabs(z - z1) > 1e-19L && ++guard < 100);
    pts[i] = (double)(0.5L * (1.0L - z));
    pts[n - 1 - i] = (double)(0.5L * (1.0L + z));
    w[i] = (double)(1.0L / ((1.0L - z * z) * pp * pp));
    w[n - 1 - i] = w[i];
  }
}

// FeLagrangeO1Segment / FeLagrangeO3Segment shape function values
// (uscalfe/lagr_fe.h:284-330, :1155-1230) at reference point x.
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

struct SegTables {
  int nsf = 0, nqp = 0;
  double pts[8], w[8], N[8 * 4];
};

// make_QuadRule(kSegment, 2 * degree) (quad/make_quad_rule.cc:13-18): degree + 1
// Gauss points, as the edge providers' constructors choose
// (uscalfe/loc_comp_ellbvp.h:383-396, :806-820).
void make_seg_tables(int degree, SegTables &t) {
  t.nsf = degree + 1;
  t.nqp = (2 * degree) / 2 + 1;
  gauss_legendre01(t.nqp, t.pts, t.w);
  for (int k = 0; k < t.nqp; ++k) segment_shape(degree, t.pts[k], t.N + t.nsf * k);
}

// SegmentO1 (geometry/segment_o1.cc): |v1 - v0| and Global(x) = v1 x + v0 (1 - x).
double seg_length(const double *a, const double *b) {
  const double dx = b[0] - a[0], dy = b[1] - a[1];
  return std::sqrt(dx * dx + dy * dy);
}

// ---------------------------------------------------------------------------
// TriaO1 (geometry/tria_o1.cc:10-68).
// ---------------------------------------------------------------------------
struct Geo {
  double J[2][2];
  double JinvT[2][2];
  double det;
};

bool non_degenerate(const double *p0, const double *p1, const double *p2) {
  // assertNonDegenerateTriangle, tol = 1e-8 (tria_o1.h:16-18, tria_o1.cc:10-33)
  const double tol = 1.0e-8;
  auto sq = [](double a, double b) { return a * a + b * b; };
  const double e0 = sq(p1[0] - p0[0], p1[1] - p0[1]);
  const double e1 = sq(p2[0] - p1[0], p2[1] - p1[1]);
  const double e2 = sq(p0[0] - p2[0], p0[1] - p2[1]);
  const double circum = e0 + e1 + e2;
  if (!(e0 > tol * circum) || !(e1 > tol * circum) || !(e2 > tol * circum))
    return false;
  const double area = std::fabs((p1[0] - p0[0]) * (p2[1] - p0[1]) -
                                (p1[1] - p0[1]) * (p2[0] - p0[0]));
  return area > tol * circum;
}

void tria_o1(const double *p0, const double *p1, const double *p2, Geo &g) {
  // jacobian_ << v1 - v0, v2 - v0  (tria_o1.cc:57)
  g.J[0][0] = p1[0] - p0[0];
  g.J[1][0] = p1[1] - p0[1];
  g.J[0][1] = p2[0] - p0[0];
  g.J[1][1] = p2[1] - p0[1];
  // jacobian_.transpose().inverse() (tria_o1.cc:60): Eigen's fixed 2x2
  // inverse of m = J^T: invdet = 1/(m00 m11 - m10 m01).
  const double m00 = g.J[0][0], m01 = g.J[1][0], m10 = g.J[0][1],
               m11 = g.J[1][1];
  const double invdet = 1.0 / (m00 * m11 - m10 * m01);
  g.JinvT[0][0] = m11 * invdet;
  g.JinvT[1][0] = -m10 * invdet;
  g.JinvT[0][1] = -m01 * invdet;
  g.JinvT[1][1] = m00 * invdet;
  // std::abs(jacobian_.determinant()) (tria_o1.cc:61): jacobian_ has a
  // dynamic row count, so Eigen evaluates the determinant through a partial
  // pivoting LU (first maximal |pivot| wins, one elimination step).
  const double a = g.J[0][0], b = g.J[0][1], c = g.J[1][0], d = g.J[1][1];
  double det;
  if (std::fabs(c) > std::fabs(a)) {
    const double l = a / c;
    const double u11 = b - l * d;
    det = -1.0 * (c * u11);
  } else {
    const double l = c / a;
    const double u11 = d - l * b;
    det = 1.0 * (a * u11);
  }
  g.det = std::fabs(det);
}

// ReactionDiffusionElementMatrixProvider::Eval QP loop
// (uscalfe/loc_comp_ellbvp.h:324-337).  mat is row-major nsf x nsf.
void rd_eval(const FeTables &t, const Geo &g, const double *alpha,
             const double *gamma, double *mat) {
  const int n = t.nsf;
  for (int i = 0; i < n * n; ++i) mat[i] = 0.0;
  std::vector<double> G(2 * n);
  for (int k = 0; k < t.nqp; ++k) {
    const double w = t.rule.w[k] * g.det;
    const double *gx = &t.dN[n * (2 * k)];
    const double *gy = &t.dN[n * (2 * k + 1)];
    // trf_grad = JinvT * grad^T (2 x n)
    for (int j = 0; j < n; ++j) {
      G[j] = g.JinvT[0][0] * gx[j] + g.JinvT[0][1] * gy[j];
      G[n + j] = g.JinvT[1][0] * gx[j] + g.JinvT[1][1] * gy[j];
    }
    const double ak = alpha[k], ck = gamma[k];
    const double *Nk = &t.N[n * k];
    for (int i = 0; i < n; ++i) {
      for (int j = 0; j < n; ++j) {
        const double s = G[i] * (ak * G[j]) + G[n + i] * (ak * G[n + j]);
        const double m = (ck * Nk[i]) * Nk[j];
        mat[i * n + j] += w * (s + m);
      }
    }
  }
}

// ScalarLoadElementVectorProvider::Eval (uscalfe/loc_comp_ellbvp.h:733-740).
void load_eval(const FeTables &t, const Geo &g, const double *f, double *v) {
  const int n = t.nsf;
  for (int i = 0; i < n; ++i) v[i] = 0.0;
  for (int k = 0; k < t.nqp; ++k) {
    const double s = t.rule.w[k] * g.det * f[k];
    const double *Nk = &t.N[n * k];
    for (int i = 0; i < n; ++i) v[i] += s * Nk[i];
  }
}

// TriaO1::Global (geometry/tria_o1.cc:70-74).
void tria_global(const double *p0, const double *p1, const double *p2,
                 double a, double b, double *x) {
  const double l0 = 1 - a - b;
  for (int d = 0; d < 2; ++d) x[d] = p0[d] * l0 + p1[d] * a + p2[d] * b;
}

double f_eval(int kind, double c, double x, double y) {
  if (kind == 0) return c;
  return std::sin(M_PI * x) * std::sin(M_PI * y);
}

inline const double *P(const double *xy, int v) { return xy + 2 * v; }

}  // namespace

extern "C" {

int lfref_quad_rule(int degree, double *pts, double *w) {
  Rule r;
  if (degree == 2) r = rule_for_fe_degree(1);
  else if (degree == 6) r = rule_for_fe_degree(3);
  else return -1;
  std::memcpy(pts, r.pts, sizeof(double) * 2 * r.nqp);
  std::memcpy(w, r.w, sizeof(double) * r.nqp);
  return r.nqp;
}

int lfref_shape_functions(int degree, int npts, const double *pts, double *N,
                          double *dN) {
  if (degree == 1) {
    p1_shape(npts, pts, N, dN);
    return 3;
  }
  if (degree == 3) {
    p3_shape(npts, pts, N, dN);
    return 10;
  }
  return -1;
}

void lfref_tp_triag_mesh(int nx, int ny, double x0, double y0, double x1,
                         double y1, double *xy, int32_t *cell_v,
                         int32_t *edge_v) {
  const double hx = (x1 - x0) / nx, hy = (y1 - y0) / ny;
  auto V = [nx](int i, int j) { return i + j * (nx + 1); };
  int cnt = 0;
  for (int j = 0; j <= ny; ++j)
    for (int i = 0; i <= nx; ++i, ++cnt) {
      xy[2 * cnt] = x0 + (unsigned)i * hx;
      xy[2 * cnt + 1] = y0 + (unsigned)j * hy;
    }
  int e = 0;
  for (int i = 0; i < nx; ++i)
    for (int j = 0; j <= ny; ++j, ++e) {
      edge_v[2 * e] = V(i, j);
      edge_v[2 * e + 1] = V(i + 1, j);
    }
  for (int i = 0; i <= nx; ++i)
    for (int j = 0; j < ny; ++j, ++e) {
      edge_v[2 * e] = V(i, j);
      edge_v[2 * e + 1] = V(i, j + 1);
    }
  for (int i = 0; i < nx; ++i)
    for (int j = 0; j < ny; ++j, ++e) {
      edge_v[2 * e] = V(i, j);
      edge_v[2 * e + 1] = V(i + 1, j + 1);
    }
  int c = 0;
  for (int i = 0; i < nx; ++i)
    for (int j = 0; j < ny; ++j, c += 2) {
      int32_t *u = cell_v + 3 * c;
      u[0] = V(i, j);
      u[1] = V(i + 1, j + 1);
      u[2] = V(i, j + 1);
      int32_t *l = cell_v + 3 * (c + 1);
      l[0] = V(i, j);
      l[1] = V(i + 1, j);
      l[2] = V(i + 1, j + 1);
    }
}

int lfref_mesh_edges(int nv, int ncells, const int32_t *cell_v, int npre,
                     const int32_t *pre_edge_v, int max_edges,
                     int32_t *edge_v_out, int32_t *cell_e, int8_t *cell_ori) {
  (void)nv;
  struct EdgeData {
    int a, b;  // endpoints as first registered
    int index;
  };
  std::map<std::pair<int, int>, EdgeData> emap;
  for (int e = 0; e < npre; ++e) {
    const int a = pre_edge_v[2 * e], b = pre_edge_v[2 * e + 1];
    emap[{std::min(a, b), std::max(a, b)}] = EdgeData{a, b, e};
  }
  // cells register the edges they miss, in cell order (mesh.cc:311-430)
  for (int c = 0; c < ncells; ++c)
    for (int k = 0; k < 3; ++k) {
      const int a = cell_v[3 * c + k], b = cell_v[3 * c + (k + 1) % 3];
      auto key = std::make_pair(std::min(a, b), std::max(a, b));
      if (emap.find(key) == emap.end()) emap[key] = EdgeData{a, b, -1};
    }
  int next = npre;
  for (auto &kv : emap)
    if (kv.second.index < 0) kv.second.index = next++;
  if (next > max_edges) return -1;
  for (auto &kv : emap) {
    edge_v_out[2 * kv.second.index] = kv.second.a;
    edge_v_out[2 * kv.second.index + 1] = kv.second.b;
  }
  for (int c = 0; c < ncells; ++c)
    for (int k = 0; k < 3; ++k) {
      const int a = cell_v[3 * c + k], b = cell_v[3 * c + (k + 1) % 3];
      const EdgeData &ed = emap[{std::min(a, b), std::max(a, b)}];
      cell_e[3 * c + k] = ed.index;
      // triangle.cc:71-77: positive iff edge node 0 is cell vertex k
      cell_ori[3 * c + k] = (ed.a == a) ? 1 : -1;
    }
  return next;
}

int64_t lfref_dof_table(int nv, int ncells, int nedges, const int32_t *cell_v,
                        const int32_t *cell_e, const int8_t *cell_ori,
                        int ndof_edge, int ndof_cell, int64_t *cell_dofs) {
  const int nloc = 3 + 3 * ndof_edge + ndof_cell;
  const int64_t edge_base = nv;
  const int64_t cell_base = edge_base + (int64_t)nedges * ndof_edge;
  for (int c = 0; c < ncells; ++c) {
    int64_t *d = cell_dofs + (int64_t)c * nloc;
    int p = 0;
    for (int k = 0; k < 3; ++k) d[p++] = cell_v[3 * c + k];
    for (int k = 0; k < 3; ++k) {
      const int64_t off = edge_base + (int64_t)cell_e[3 * c + k] * ndof_edge;
      if (cell_ori[3 * c + k] > 0)
        for (int j = 0; j < ndof_edge; ++j) d[p++] = off + j;
      else
        for (int j = ndof_edge - 1; j >= 0; --j) d[p++] = off + j;
    }
    for (int j = 0; j < ndof_cell; ++j)
      d[p++] = cell_base + (int64_t)c * ndof_cell + j;
  }
  return cell_base + (int64_t)ncells * ndof_cell;
}

void lfref_boundary_flags(int nv, int ncells, int nedges,
                          const int32_t *cell_e, const int32_t *edge_v,
                          uint8_t *node_flag, uint8_t *edge_flag) {
  std::vector<int> cnt(nedges, 0);
  for (int c = 0; c < ncells; ++c)
    for (int k = 0; k < 3; ++k) cnt[cell_e[3 * c + k]]++;
  std::memset(node_flag, 0, nv);
  std::memset(edge_flag, 0, nedges);
  for (int e = 0; e < nedges; ++e)
    if (cnt[e] == 1) {
      edge_flag[e] = 1;
      node_flag[edge_v[2 * e]] = 1;
      node_flag[edge_v[2 * e + 1]] = 1;
    }
}

int lfref_geometry(int ncells, const double *xy, const int32_t *cell_v,
                   double *J, double *JinvT, double *det) {
  for (int c = 0; c < ncells; ++c) {
    const double *p0 = P(xy, cell_v[3 * c]), *p1 = P(xy, cell_v[3 * c + 1]),
                 *p2 = P(xy, cell_v[3 * c + 2]);
    if (!non_degenerate(p0, p1, p2)) return -(1 + c);
    Geo g;
    tria_o1(p0, p1, p2, g);
    std::memcpy(J + 4 * c, g.J, sizeof g.J);
    std::memcpy(JinvT + 4 * c, g.JinvT, sizeof g.JinvT);
    det[c] = g.det;
  }
  return 0;
}

int lfref_element_matrices(int degree, int ncells, const double *xy,
                           const int32_t *cell_v, const double *alpha,
                           const double *gamma, double *mats) {
  FeTables t;
  if (!make_tables(degree, t)) return -1000;
  const int n = t.nsf, q = t.nqp;
  for (int c = 0; c < ncells; ++c) {
    const double *p0 = P(xy, cell_v[3 * c]), *p1 = P(xy, cell_v[3 * c + 1]),
                 *p2 = P(xy, cell_v[3 * c + 2]);
    if (!non_degenerate(p0, p1, p2)) return -(1 + c);
    Geo g;
    tria_o1(p0, p1, p2, g);
    rd_eval(t, g, alpha + (int64_t)c * q, gamma + (int64_t)c * q,
            mats + (int64_t)c * n * n);
  }
  return 0;
}

int lfref_qp_coords(int degree, int ncells, const double *xy,
                    const int32_t *cell_v, double *out) {
  FeTables t;
  if (!make_tables(degree, t)) return -1000;
  for (int c = 0; c < ncells; ++c) {
    const double *p0 = P(xy, cell_v[3 * c]), *p1 = P(xy, cell_v[3 * c + 1]),
                 *p2 = P(xy, cell_v[3 * c + 2]);
    for (int k = 0; k < t.nqp; ++k)
      tria_global(p0, p1, p2, t.rule.pts[2 * k], t.rule.pts[2 * k + 1],
                  out + ((int64_t)c * t.nqp + k) * 2);
  }
  return 0;
}

int lfref_element_vectors(int degree, int ncells, const double *xy,
                          const int32_t *cell_v, int f_kind, double f_const,
                          const double *f_qp, double *vecs) {
  FeTables t;
  if (!make_tables(degree, t)) return -1000;
  const int n = t.nsf, q = t.nqp;
  std::vector<double> f(q);
  for (int c = 0; c < ncells; ++c) {
    const double *p0 = P(xy, cell_v[3 * c]), *p1 = P(xy, cell_v[3 * c + 1]),
                 *p2 = P(xy, cell_v[3 * c + 2]);
    if (!non_degenerate(p0, p1, p2)) return -(1 + c);
    Geo g;
    tria_o1(p0, p1, p2, g);
    for (int k = 0; k < q; ++k) {
      if (f_kind == 2) {
        f[k] = f_qp[(int64_t)c * q + k];
      } else {
        double x[2];
        tria_global(p0, p1, p2, t.rule.pts[2 * k], t.rule.pts[2 * k + 1], x);
        f[k] = f_eval(f_kind, f_const, x[0], x[1]);
      }
    }
    load_eval(t, g, f.data(), vecs + (int64_t)c * n);
  }
  return 0;
}

void lfref_assemble_coo(int ncells, int nloc, const int64_t *cell_dofs,
                        const double *mats, int32_t *rows, int32_t *cols,
                        double *vals) {
  int64_t t = 0;
  for (int c = 0; c < ncells; ++c) {
    const int64_t *d = cell_dofs + (int64_t)c * nloc;
    const double *m = mats + (int64_t)c * nloc * nloc;
    for (int i = 0; i < nloc; ++i)
      for (int j = 0; j < nloc; ++j, ++t) {
        rows[t] = (int32_t)d[i];
        cols[t] = (int32_t)d[j];
        vals[t] = m[i * nloc + j];
      }
  }
}

void lfref_assemble_vector(int ncells, int nloc, const int64_t *cell_dofs,
                           const double *vecs, int64_t ndofs, double *b) {
  for (int64_t i = 0; i < ndofs; ++i) b[i] = 0.0;
  for (int c = 0; c < ncells; ++c)
    for (int i = 0; i < nloc; ++i)
      b[cell_dofs[(int64_t)c * nloc + i]] += vecs[(int64_t)c * nloc + i];
}

void lfref_coo_matvec(int64_t nnz, const int32_t *rows, const int32_t *cols,
                      const double *vals, double alpha, const double *x,
                      double *y) {
  for (int64_t t = 0; t < nnz; ++t) y[rows[t]] += vals[t] * (alpha * x[cols[t]]);
}

int64_t lfref_fix_flagged(int64_t nnz, int32_t *rows, int32_t *cols,
                          double *vals, int64_t ndofs, const uint8_t *flags,
                          const double *fixed_vals, double *b) {
  // b -= A * x_fixed (fix_dof.h:97-106)
  std::vector<double> tmp(ndofs, 0.0);
  for (int64_t k = 0; k < ndofs; ++k)
    if (flags[k]) tmp[k] = fixed_vals[k];
  lfref_coo_matvec(nnz, rows, cols, vals, -1.0, tmp.data(), b);
  for (int64_t k = 0; k < ndofs; ++k)
    if (flags[k]) b[k] = fixed_vals[k];
  // stable remove_if of triplets touching a fixed row or column
  // (fix_dof.h:117-119, coomatrix.h:108-115)
  int64_t w = 0;
  for (int64_t t = 0; t < nnz; ++t) {
    if (flags[rows[t]] || flags[cols[t]]) continue;
    rows[w] = rows[t];
    cols[w] = cols[t];
    vals[w] = vals[t];
    ++w;
  }
  // unit diagonal for fixed DOFs, in DOF order (fix_dof.h:132-137)
  for (int64_t k = 0; k < ndofs; ++k)
    if (flags[k]) {
      rows[w] = (int32_t)k;
      cols[w] = (int32_t)k;
      vals[w] = 1.0;
      ++w;
    }
  return w;
}

int64_t lfref_set_from_triplets(int64_t nnz, const int32_t *rows,
                                const int32_t *cols, const double *vals,
                                int32_t nrows, int32_t *row_ptr, int32_t *col,
                                double *val) {
  // Stable bucket by row, then by column within the row; duplicates are
  // summed in triplet order (the insertion order Eigen's setFromTriplets
  // preserves).
  std::vector<int64_t> cnt(nrows + 1, 0);
  for (int64_t t = 0; t < nnz; ++t) cnt[rows[t] + 1]++;
  for (int32_t r = 0; r < nrows; ++r) cnt[r + 1] += cnt[r];
  std::vector<int64_t> order(nnz);
  {
    std::vector<int64_t> pos(cnt.begin(), cnt.end() - 1);
    for (int64_t t = 0; t < nnz; ++t) order[pos[rows[t]]++] = t;
  }
  int64_t out = 0;
  row_ptr[0] = 0;
  std::vector<int64_t> seg;
  for (int32_t r = 0; r < nrows; ++r) {
    seg.assign(order.begin() + cnt[r], order.begin() + cnt[r + 1]);
    std::stable_sort(seg.begin(), seg.end(), [&](int64_t a, int64_t b) {
      return cols[a] < cols[b];
    });
    for (size_t s = 0; s < seg.size();) {
      const int32_t c = cols[seg[s]];
      double acc = vals[seg[s]];
      size_t e = s + 1;
      while (e < seg.size() && cols[seg[e]] == c) acc += vals[seg[e++]];
      col[out] = c;
      val[out] = acc;
      ++out;
      s = e;
    }
    row_ptr[r + 1] = (int32_t)out;
  }
  return out;
}

int lfref_band_cholesky_solve(int32_t n, const int32_t *row_ptr,
                              const int32_t *col, const double *val,
                              const double *b, double *x) {
  int64_t bw = 0;
  for (int32_t i = 0; i < n; ++i)
    for (int32_t p = row_ptr[i]; p < row_ptr[i + 1]; ++p)
      bw = std::max<int64_t>(bw, (int64_t)i - col[p]);
  if ((double)n * (double)(bw + 1) > 4.0e8) return -2;
  const int64_t W = bw + 1;
  // L stored by rows: L[i][j] at band[i*W + (j - i + bw)], j in [i-bw, i]
  std::vector<double> band((size_t)n * W, 0.0);
  for (int32_t i = 0; i < n; ++i)
    for (int32_t p = row_ptr[i]; p < row_ptr[i + 1]; ++p)
      if (col[p] <= i && i - col[p] <= bw)
        band[(size_t)i * W + (col[p] - i + bw)] += val[p];
  for (int32_t i = 0; i < n; ++i) {
    const int64_t j0 = std::max<int64_t>(0, i - bw);
    for (int64_t j = j0; j <= i; ++j) {
      double s = band[(size_t)i * W + (j - i + bw)];
      const int64_t k0 = std::max<int64_t>(j0, j - bw);
      for (int64_t k = k0; k < j; ++k)
        s -= band[(size_t)i * W + (k - i + bw)] * band[(size_t)j * W + (k - j + bw)];
      if (j == i) {
        if (!(s > 0.0)) return -1;
        band[(size_t)i * W + bw] = std::sqrt(s);
      } else {
        band[(size_t)i * W + (j - i + bw)] = s / band[(size_t)j * W + bw];
      }
    }
  }
  std::vector<double> y(n);
  for (int32_t i = 0; i < n; ++i) {
    double s = b[i];
    for (int64_t k = std::max<int64_t>(0, i - bw); k < i; ++k)
      s -= band[(size_t)i * W + (k - i + bw)] * y[k];
    y[i] = s / band[(size_t)i * W + bw];
  }
  for (int32_t i = n - 1; i >= 0; --i) {
    double s = y[i];
    for (int64_t k = i + 1; k <= std::min<int64_t>(n - 1, i + bw); ++k)
      s -= band[(size_t)k * W + (i - k + bw)] * x[k];
    x[i] = s / band[(size_t)i * W + bw];
  }
  return 0;
}

void lfref_csr_spmv(int32_t n, const int32_t *row_ptr, const int32_t *col,
                    const double *val, const double *x, double *y) {
  for (int32_t i = 0; i < n; ++i) {
    double s = 0.0;
    for (int32_t p = row_ptr[i]; p < row_ptr[i + 1]; ++p) s += val[p] * x[col[p]];
    y[i] = s;
  }
}

int lfref_pcg(int32_t n, const int32_t *row_ptr, const int32_t *col,
              const double *val, const double *b, double *x, double tol,
              int maxit, double *err) {
  std::vector<double> dinv(n, 1.0), r(n), p(n), z(n), q(n);
  for (int32_t i = 0; i < n; ++i)
    for (int32_t k = row_ptr[i]; k < row_ptr[i + 1]; ++k)
      if (col[k] == i && val[k] != 0.0) dinv[i] = 1.0 / val[k];
  auto dot = [n](const double *u, const double *v) {
    double s = 0.0;
    for (int32_t i = 0; i < n; ++i) s += u[i] * v[i];
    return s;
  };
  lfref_csr_spmv(n, row_ptr, col, val, x, q.data());
  for (int32_t i = 0; i < n; ++i) r[i] = b[i] - q[i];
  const double bb = dot(b, b);
  if (bb == 0.0) {
    for (int32_t i = 0; i < n; ++i) x[i] = 0.0;
    *err = 0.0;
    return 0;
  }
  const double thr = std::max(tol * tol * bb, 2.2250738585072014e-308);
  double rr = dot(r.data(), r.data());
  if (rr < thr) {
    *err = std::sqrt(rr / bb);
    return 0;
  }
  for (int32_t i = 0; i < n; ++i) p[i] = dinv[i] * r[i];
  double absNew = dot(r.data(), p.data());
  int it = 0;
  while (it < maxit) {
    lfref_csr_spmv(n, row_ptr, col, val, p.data(), q.data());
    const double alpha = absNew / dot(p.data(), q.data());
    for (int32_t i = 0; i < n; ++i) {
      x[i] += alpha * p[i];
      r[i] -= alpha * q[i];
    }
    rr = dot(r.data(), r.data());
    if (rr < thr) break;
    for (int32_t i = 0; i < n; ++i) z[i] = dinv[i] * r[i];
    const double absOld = absNew;
    absNew = dot(r.data(), z.data());
    const double beta = absNew / absOld;
    for (int32_t i = 0; i < n; ++i) p[i] = z[i] + beta * p[i];
    ++it;
  }
  *err = std::sqrt(rr / bb);
  return it;
}

int64_t lfref_cpu_path_assemble(int degree, int ncells, const double *xy,
                                const int32_t *cell_v, const int64_t *cell_dofs,
                                int64_t ndofs, const double *alpha_cell,
                                double gamma_const, int f_kind, double f_const,
                                const uint8_t *fixed, int32_t *row_ptr,
                                int32_t *col, double *val, double *b) {
  FeTables t;
  if (!make_tables(degree, t)) return -1000;
  const int n = t.nsf, q = t.nqp;
  struct Trip {
    int32_t r, c;
    double v;
  };
  std::vector<Trip> trips;  // grows by push_back like COOMatrix::AddToEntry
  for (int64_t i = 0; i < ndofs; ++i) b[i] = 0.0;
  for (int c = 0; c < ncells; ++c) {
    const double *p0 = P(xy, cell_v[3 * c]), *p1 = P(xy, cell_v[3 * c + 1]),
                 *p2 = P(xy, cell_v[3 * c + 2]);
    if (!non_degenerate(p0, p1, p2)) return -(1 + c);
    Geo g;
    tria_o1(p0, p1, p2, g);
    // per-cell heap objects, as the provider's Eval returns them
    std::vector<double> a(q, alpha_cell[c]), gm(q, gamma_const);
    std::vector<double> mat(n * n);
    rd_eval(t, g, a.data(), gm.data(), mat.data());
    const int64_t *d = cell_dofs + (int64_t)c * n;
    for (int i = 0; i < n; ++i)
      for (int j = 0; j < n; ++j)
        trips.push_back(Trip{(int32_t)d[i], (int32_t)d[j], mat[i * n + j]});
  }
  for (int c = 0; c < ncells; ++c) {
    const double *p0 = P(xy, cell_v[3 * c]), *p1 = P(xy, cell_v[3 * c + 1]),
                 *p2 = P(xy, cell_v[3 * c + 2]);
    Geo g;
    tria_o1(p0, p1, p2, g);
    std::vector<double> f(q), v(n);
    for (int k = 0; k < q; ++k) {
      double x[2];
      tria_global(p0, p1, p2, t.rule.pts[2 * k], t.rule.pts[2 * k + 1], x);
      f[k] = f_eval(f_kind, f_const, x[0], x[1]);
    }
    load_eval(t, g, f.data(), v.data());
    const int64_t *d = cell_dofs + (int64_t)c * n;
    for (int i = 0; i < n; ++i) b[d[i]] += v[i];
  }
  // FixFlaggedSolutionComponents with zero prescribed values
  for (auto &tr : trips) b[tr.r] += tr.v * (-1.0 * 0.0);
  for (int64_t k = 0; k < ndofs; ++k)
    if (fixed[k]) b[k] = 0.0;
  trips.erase(std::remove_if(trips.begin(), trips.end(),
                             [&](const Trip &tr) {
                               return fixed[tr.r] || fixed[tr.c];
                             }),
              trips.end());
  for (int64_t k = 0; k < ndofs; ++k)
    if (fixed[k]) trips.push_back(Trip{(int32_t)k, (int32_t)k, 1.0});
  std::vector<int32_t> R(trips.size()), C(trips.size());
  std::vector<double> Vv(trips.size());
  for (size_t i = 0; i < trips.size(); ++i) {
    R[i] = trips[i].r;
    C[i] = trips[i].c;
    Vv[i] = trips[i].v;
  }
  return lfref_set_from_triplets((int64_t)trips.size(), R.data(), C.data(),
                                 Vv.data(), (int32_t)ndofs, row_ptr, col, val);
}

int lfref_segment_rule(int n, double *pts, double *w) {
  if (n < 1 || n > 8) return -1;
  gauss_legendre01(n, pts, w);
  return 0;
}

// Coefficient on edge e at point k: 0 constant c, 1 c * data[e],
// 2 c * data[e * nq + k], 3 c * sin(pi x) sin(pi y) at the mapped point.
static double edge_coeff(int kind, double c, const double *data, int e, int k, int nq,
                         const double *a, const double *b, double xhat) {
  if (kind == 0) return c;
  if (kind == 1) return c * data[e];
  if (kind == 2) return c * data[(int64_t)e * nq + k];
  const double x = b[0] * xhat + a[0] * (1 - xhat), y = b[1] * xhat + a[1] * (1 - xhat);
  return c * (std::sin(M_PI * x) * std::sin(M_PI * y));
}

int lfref_edge_matrices(int degree, int nedges, const double *xy, const int32_t *edge_v,
                        int g_kind, double g_const, const double *g_data, double *out) {
  if (degree != 1 && degree != 3) return -1000;
  SegTables t;
  make_seg_tables(degree, t);
  const int n = t.nsf;
  for (int e = 0; e < nedges; ++e) {
    const double *a = P(xy, edge_v[2 * e]), *b = P(xy, edge_v[2 * e + 1]);
    const double det = seg_length(a, b);
    double *mat = out + (int64_t)e * n * n;
    for (int q = 0; q < n * n; ++q) mat[q] = 0.0;
    for (int k = 0; k < t.nqp; ++k) {
      // mat += (N_k N_k^T) * w, w = (omega_k det_k) gamma_k (loc_comp_ellbvp.h:510-526)
      const double w =
          (t.w[k] * det) * edge_coeff(g_kind, g_const, g_data, e, k, t.nqp, a, b, t.pts[k]);
      const double *Nk = t.N + n * k;
      for (int i = 0; i < n; ++i)
        for (int jj = 0; jj < n; ++jj) mat[i * n + jj] += (Nk[i] * Nk[jj]) * w;
    }
  }
  return 0;
}

int lfref_edge_vectors(int degree, int nedges, const double *xy, const int32_t *edge_v,
                       int g_kind, double g_const, const double *g_data, double *out) {
  if (degree != 1 && degree != 3) return -1000;
  SegTables t;
  make_seg_tables(degree, t);
  const int n = t.nsf;
  for (int e = 0; e < nedges; ++e) {
    const double *a = P(xy, edge_v[2 * e]), *b = P(xy, edge_v[2 * e + 1]);
    const double det = seg_length(a, b);
    double *vec = out + (int64_t)e * n;
    for (int i = 0; i < n; ++i) vec[i] = 0.0;
    for (int k = 0; k < t.nqp; ++k) {
      // vec += N_k * w, w = (omega_k det_k) g_k (loc_comp_ellbvp.h:905-916)
      const double w =
          (t.w[k] * det) * edge_coeff(g_kind, g_const, g_data, e, k, t.nqp, a, b, t.pts[k]);
      const double *Nk = t.N + n * k;
      for (int i = 0; i < n; ++i) vec[i] += Nk[i] * w;
    }
  }
  return 0;
}

// FisherKPP's nonlocal boundary kernel (projects/fisher_kpp/
// modelproblem_island_main.cc:76-80, strangsplitting_main.cc:160-165,
// convergence_tau_main.cc:108-113): g(x, y) = 1 / (1 + |x - y|^2), and with
// rmax > 0 zero unless rmin <= |x - y| <= rmax.
static double nl_kernel(const double *x, const double *y, double rmin, double rmax) {
  const double dx = x[0] - y[0], dy = x[1] - y[1];
  const double s = dx * dx + dy * dy;  // (x - y).squaredNorm()
  if (rmax > 0.0) {
    const double d = std::sqrt(s);  // (x - y).norm()
    if (!(d >= rmin && d <= rmax)) return 0.0;
  }
  return 1.0 / (1.0 + s);
}

// The gain h(x) = localQuadFunction(mesh, {segment rule}, g(x, .), 1,
// edge_pred) (strangsplitting.h:78-115): sum over the flagged edges in index
// order and their points l of (w_l g(x, y_l)) det.
static double nl_gain_at(const SegTables &t, int nedges, const double *xy, const int32_t *edge_v,
                         const uint8_t *flags, double rmin, double rmax, const double *x) {
  double sum = 0.0;
  for (int e = 0; e < nedges; ++e) {
    if (!flags[e]) continue;
    const double *a = P(xy, edge_v[2 * e]), *b = P(xy, edge_v[2 * e + 1]);
    const double det = seg_length(a, b);
    for (int l = 0; l < t.nqp; ++l) {
      const double y[2] = {b[0] * t.pts[l] + a[0] * (1 - t.pts[l]),
                           b[1] * t.pts[l] + a[1] * (1 - t.pts[l])};
      sum += t.w[l] * nl_kernel(x, y, rmin, rmax) * det;
    }
  }
  return sum;
}

int lfref_nonlocal_gain(int degree, int nedges, const double *xy, const int32_t *edge_v,
                        const uint8_t *flags, double rmin, double rmax, double *out) {
  if (degree != 1 && degree != 3) return -1000;
  SegTables t;
  make_seg_tables(degree, t);
  for (int e = 0; e < nedges; ++e) {
    const double *a = P(xy, edge_v[2 * e]), *b = P(xy, edge_v[2 * e + 1]);
    for (int k = 0; k < t.nqp; ++k) {
      // MassEdgeMatrixProvider evaluates h at SegmentO1::Global of its points
      const double x[2] = {b[0] * t.pts[k] + a[0] * (1 - t.pts[k]),
                           b[1] * t.pts[k] + a[1] * (1 - t.pts[k])};
      out[(int64_t)e * t.nqp + k] =
          flags[e] ? nl_gain_at(t, nedges, xy, edge_v, flags, rmin, rmax, x) : 0.0;
    }
  }
  return 0;
}

// The loss matrix L of modelproblem_island_main.cc:92-126: column j (a DOF
// of a flagged edge) is AssembleVectorLocally(1, ...) of the edge load with
// f = L_j, L_j(x) = [AssembleVectorLocally(1, ...) of the edge load with
// f = g(x, .)]_j; both loads are ScalarLoadEdgeVectorProvider::Eval
// (uscalfe/loc_comp_ellbvp.h:905-916) over the flagged edges in index order.
// rows_out receives the DOFs of the flagged edges in ascending order; L_out
// the nb x nb block on them (row-major); returns nb, or -1 if it exceeds
// max_nb.
int lfref_nonlocal_loss(int degree, int nedges, const double *xy, const int32_t *edge_v,
                        const int64_t *edge_dofs, const uint8_t *flags, double rmin, double rmax,
                        int max_nb, int32_t *rows_out, double *L_out) {
  if (degree != 1 && degree != 3) return -1000;
  SegTables t;
  make_seg_tables(degree, t);
  const int nle = t.nsf, nq = t.nqp;
  // the flagged edges of every DOF, in index order, with its local index
  std::map<int64_t, std::vector<std::pair<int, int>>> dof_edges;
  for (int e = 0; e < nedges; ++e)
    if (flags[e])
      for (int m = 0; m < nle; ++m) dof_edges[edge_dofs[(int64_t)e * nle + m]].push_back({e, m});
  const int nb = (int)dof_edges.size();
  if (nb > max_nb) return -1;
  std::vector<int64_t> rows;
  for (const auto &kv : dof_edges) rows.push_back(kv.first);
  for (int b = 0; b < nb; ++b) rows_out[b] = (int32_t)rows[b];
  auto qp = [&](int e, int k, double *x) {
    const double *a = P(xy, edge_v[2 * e]), *b = P(xy, edge_v[2 * e + 1]);
    x[0] = b[0] * t.pts[k] + a[0] * (1 - t.pts[k]);
    x[1] = b[1] * t.pts[k] + a[1] * (1 - t.pts[k]);
  };
  auto det_of = [&](int e) { return seg_length(P(xy, edge_v[2 * e]), P(xy, edge_v[2 * e + 1])); };
  // L_j(x): entry j of the assembled edge load of g(x, .)
  auto Lj = [&](int bj, const double *x) {
    double acc = 0.0;
    for (const auto &em : dof_edges[rows[bj]]) {
      const int e = em.first, m = em.second;
      const double det = det_of(e);
      double v = 0.0;
      for (int l = 0; l < nq; ++l) {
        double y[2];
        qp(e, l, y);
        const double w = (t.w[l] * det) * nl_kernel(x, y, rmin, rmax);
        v += t.N[nle * l + m] * w;
      }
      acc += v;
    }
    return acc;
  };
  for (int bi = 0; bi < nb; ++bi)
    for (int bj = 0; bj < nb; ++bj) {
      double acc = 0.0;
      for (const auto &em : dof_edges[rows[bi]]) {
        const int e = em.first, m = em.second;
        const double det = det_of(e);
        double v = 0.0;
        for (int k = 0; k < nq; ++k) {
          double x[2];
          qp(e, k, x);
          const double w = (t.w[k] * det) * Lj(bj, x);
          v += t.N[nle * k + m] * w;
        }
        acc += v;
      }
      L_out[(int64_t)bi * nb + bj] = acc;
    }
  return nb;
}

int lfref_error_norms(int degree, int ncells, const double *xy,
                      const int32_t *cell_v, const int64_t *cell_dofs,
                      const double *uh, int exact_kind, const double *u_qp,
                      const double *g_qp, double *out) {
  FeTables t;
  if (!make_tables(degree, t)) return -1000;
  const int n = t.nsf, q = t.nqp;
  double l2 = 0.0, h1 = 0.0;
  std::vector<double> u(n);
  for (int c = 0; c < ncells; ++c) {
    const double *p0 = P(xy, cell_v[3 * c]), *p1 = P(xy, cell_v[3 * c + 1]),
                 *p2 = P(xy, cell_v[3 * c + 2]);
    if (!non_degenerate(p0, p1, p2)) return -(1 + c);
    Geo g;
    tria_o1(p0, p1, p2, g);
    for (int i = 0; i < n; ++i) u[i] = uh[cell_dofs[(int64_t)c * n + i]];
    double cl2 = 0.0, ch1 = 0.0;
    for (int k = 0; k < q; ++k) {
      // MeshFunctionFE: local_dofs * N;  MeshFunctionGradFE: J^{-T} (local_dofs * grad N)
      double v = 0.0, r0 = 0.0, r1 = 0.0;
      for (int i = 0; i < n; ++i) {
        v += u[i] * t.N[n * k + i];
        r0 += u[i] * t.dN[n * (2 * k) + i];
        r1 += u[i] * t.dN[n * (2 * k + 1) + i];
      }
      const double g0 = g.JinvT[0][0] * r0 + g.JinvT[0][1] * r1;
      const double g1 = g.JinvT[1][0] * r0 + g.JinvT[1][1] * r1;
      double ue, ge0, ge1;
      if (exact_kind == 0) {
        double x[2];
        tria_global(p0, p1, p2, t.rule.pts[2 * k], t.rule.pts[2 * k + 1], x);
        const double sx = std::sin(M_PI * x[0]), sy = std::sin(M_PI * x[1]);
        ue = sx * sy;
        ge0 = M_PI * (std::cos(M_PI * x[0]) * sy);
        ge1 = M_PI * (sx * std::cos(M_PI * x[1]));
      } else {
        ue = u_qp[(int64_t)c * q + k];
        ge0 = g_qp[((int64_t)c * q + k) * 2];
        ge1 = g_qp[((int64_t)c * q + k) * 2 + 1];
      }
      const double wk = t.rule.w[k] * g.det;
      const double d = v - ue, e0 = g0 - ge0, e1 = g1 - ge1;
      cl2 += (d * d) * wk;
      ch1 += (e0 * e0 + e1 * e1) * wk;
    }
    l2 += cl2;
    h1 += ch1;
  }
  out[0] = l2;
  out[1] = h1;
  return 0;
}

int lfref_adjoint_gradient_p1(int ncells, const double *xy,
                              const int32_t *cell_v, const double *theta,
                              const double *u, const double *lam,
                              double *grad) {
  FeTables t;
  make_tables(1, t);
  const double one[3] = {1.0, 1.0, 1.0}, zero[3] = {0.0, 0.0, 0.0};
  double K[9];
  for (int c = 0; c < ncells; ++c) {
    const int32_t *v = cell_v + 3 * c;
    const double *p0 = P(xy, v[0]), *p1 = P(xy, v[1]), *p2 = P(xy, v[2]);
    if (!non_degenerate(p0, p1, p2)) return -(1 + c);
    Geo g;
    tria_o1(p0, p1, p2, g);
    rd_eval(t, g, one, zero, K);
    double s = 0.0;
    for (int i = 0; i < 3; ++i) {
      double ku = 0.0;
      for (int j = 0; j < 3; ++j) ku += K[i * 3 + j] * u[v[j]];
      s += lam[v[i]] * ku;
    }
    grad[c] = -std::exp(theta[c]) * s;
  }
  return 0;
}

}  // extern "C"
