// lfref_omp.cc -- the all-cores CPU timing baseline of the Jacobi-PCG.
//
// TEST INFRASTRUCTURE ONLY: loaded by the CPU-baseline leg of the benchmark
// and by the CPU tests, never by the product library.  Kept out of
// liblfref.so so that the OpenMP runtime is loaded only where this is timed.
//
// The recurrences are those of lfref_pcg (Eigen 3.4.0 ConjugateGradient with
// DiagonalPreconditioner, as called after the assembly of
// examples/ellbvp_linfe/homDir_linfe_demo.cc:171-174 in place of SparseLU).
// The loops and dot products are split over OpenMP threads, so the sums are
// taken in a different order and the iterates differ from lfref_pcg in the
// last bits.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <vector>

#include <omp.h>

namespace {

void spmv(int32_t n, const int32_t *rp, const int32_t *col, const double *val, const double *x,
          double *y) {
#pragma omp parallel for schedule(static)
  for (int32_t i = 0; i < n; ++i) {
    double s = 0.0;
    for (int32_t k = rp[i]; k < rp[i + 1]; ++k) s += val[k] * x[col[k]];
    y[i] = s;
  }
}

double dot(int32_t n, const double *u, const double *v) {
  double s = 0.0;
#pragma omp parallel for schedule(static) reduction(+ : s)
  for (int32_t i = 0; i < n; ++i) s += u[i] * v[i];
  return s;
}

}  // namespace

extern "C" int lfref_omp_threads(void) { return omp_get_max_threads(); }

// Same contract as lfref_pcg: x0 in x, stop when ||r||^2 < tol^2 ||b||^2;
// returns the iteration count, *err = ||r|| / ||b||.
extern "C" int lfref_pcg_omp(int32_t n, const int32_t *rp, const int32_t *col, const double *val,
                             const double *b, double *x, double tol, int maxit, double *err) {
  std::vector<double> dinv(n, 1.0), r(n), p(n), z(n), q(n);
#pragma omp parallel for schedule(static)
  for (int32_t i = 0; i < n; ++i)
    for (int32_t k = rp[i]; k < rp[i + 1]; ++k)
      if (col[k] == i && val[k] != 0.0) dinv[i] = 1.0 / val[k];
  spmv(n, rp, col, val, x, q.data());
#pragma omp parallel for schedule(static)
  for (int32_t i = 0; i < n; ++i) r[i] = b[i] - q[i];
  const double bb = dot(n, b, b);
  if (bb == 0.0) {
    std::fill(x, x + n, 0.0);
    *err = 0.0;
    return 0;
  }
  const double thr = std::max(tol * tol * bb, 2.2250738585072014e-308);
  double rr = dot(n, r.data(), r.data());
  if (rr < thr) {
    *err = std::sqrt(rr / bb);
    return 0;
  }
#pragma omp parallel for schedule(static)
  for (int32_t i = 0; i < n; ++i) p[i] = dinv[i] * r[i];
  double absNew = dot(n, r.data(), p.data());
  int it = 0;
  while (it < maxit) {
    spmv(n, rp, col, val, p.data(), q.data());
    const double alpha = absNew / dot(n, p.data(), q.data());
#pragma omp parallel for schedule(static)
    for (int32_t i = 0; i < n; ++i) {
      x[i] += alpha * p[i];
      r[i] -= alpha * q[i];
    }
    rr = dot(n, r.data(), r.data());
    if (rr < thr) break;
#pragma omp parallel for schedule(static)
    for (int32_t i = 0; i < n; ++i) z[i] = dinv[i] * r[i];
    const double absOld = absNew;
    absNew = dot(n, r.data(), z.data());
    const double beta = absNew / absOld;
#pragma omp parallel for schedule(static)
    for (int32_t i = 0; i < n; ++i) p[i] = z[i] + beta * p[i];
    ++it;
  }
  *err = std::sqrt(rr / bb);
  return it;
}
