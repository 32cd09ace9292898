// Driver in the shape of examples/ellbvp_linfe/homDir_linfe_demo.cc:114-175,
// written against the lf::amd facade: the reference calling pattern
// (COOMatrix + AssembleMatrixLocally + AssembleVectorLocally +
// FixFlaggedSolutionComponents + makeSparse) and the fused device path must
// give the same CSR matrix and right-hand side bit for bit; the PCG solve
// of -Laplace u = 1 on the unit square must match the known peak value.
// Exit status 0 on success.
#include <cmath>
#include <cstdio>
#include <vector>

#include "lehrfempp_amd/include/lf_amd/lf_amd.h"

namespace amd = lf::amd;

static int fails = 0;
#define EXPECT(cond, ...)                 \
  do {                                    \
    if (!(cond)) {                        \
      std::printf("FAIL: " __VA_ARGS__);  \
      std::printf("\n");                  \
      ++fails;                            \
    }                                     \
  } while (0)

static void run(const amd::Context &ctx, int degree, unsigned n) {
  auto mesh_p = amd::mesh::TPTriagMeshBuilder()
                    .setBottomLeftCorner(0.0, 0.0)
                    .setTopRightCorner(1.0, 1.0)
                    .setNumXCells(n)
                    .setNumYCells(n)
                    .Build();
  std::shared_ptr<const amd::uscalfe::FeSpaceLagrange> fe_space;
  if (degree == 1)
    fe_space = std::make_shared<amd::uscalfe::FeSpaceLagrangeO1>(ctx, mesh_p);
  else
    fe_space = std::make_shared<amd::uscalfe::FeSpaceLagrangeO3>(ctx, mesh_p);
  const auto &dofh = fe_space->LocGlobMap();
  const amd::size_type N_dofs = dofh.NumDofs();

  // reference pattern
  amd::assemble::COOMatrix<double> A(N_dofs, N_dofs);
  amd::uscalfe::ReactionDiffusionElementMatrixProvider elmat_builder(
      fe_space, amd::uscalfe::MeshFunction::Constant(1.0),
      amd::uscalfe::MeshFunction::Constant(0.0));
  amd::assemble::AssembleMatrixLocally(0, dofh, dofh, elmat_builder, A);
  std::vector<double> phi(N_dofs, 0.0);
  amd::uscalfe::ScalarLoadElementVectorProvider elvec_builder(
      fe_space, amd::uscalfe::MeshFunction::Constant(1.0));
  amd::assemble::AssembleVectorLocally(0, dofh, elvec_builder, phi);
  const std::vector<uint8_t> bd = fe_space->BoundaryDofFlags();
  amd::assemble::FixFlaggedSolutionComponents<double>(
      [&bd](amd::gdof_idx_t k) { return std::make_pair(bd[k] != 0, 0.0); }, A, phi);
  const amd::assemble::CsrMatrix A_crs = A.makeSparse(ctx);

  // fused device path
  amd::assemble::GalerkinAssembler gal(fe_space, bd);
  amd::DeviceVector rhs(ctx, N_dofs);
  const auto f = amd::uscalfe::MeshFunction::Constant(1.0);
  gal.Assemble(amd::uscalfe::MeshFunction::Constant(1.0),
               amd::uscalfe::MeshFunction::Constant(0.0), &f, &rhs);

  std::vector<int32_t> rp1, c1, rp2, c2;
  std::vector<double> v1, v2;
  A_crs.download(rp1, c1, v1);
  gal.matrix().download(rp2, c2, v2);
  EXPECT(rp1 == rp2 && c1 == c2, "P%d n=%u: CSR patterns differ", degree, n);
  EXPECT(v1 == v2, "P%d n=%u: CSR values differ", degree, n);
  EXPECT(phi == rhs.host(), "P%d n=%u: right-hand sides differ", degree, n);

  amd::DeviceVector u(ctx, N_dofs);
  const auto info = amd::solve::PCG(ctx, gal.matrix(), rhs, u, 1e-12);
  const auto uh = u.host();
  double umax = 0.0;
  for (double v : uh) umax = std::max(umax, v);
  // max of the solution of -Laplace u = 1 with u = 0 on the unit square
  // boundary is 0.0736713...; FE values at the mesh nodes approach it at
  // O(h^2)
  const double tol = degree == 1 ? 2e-4 * (64.0 / n) * (64.0 / n) : 2e-5;
  EXPECT(info.converged && std::fabs(umax - 0.07367135) < tol, "P%d n=%u: max u = %.8f", degree,
         n, umax);
  std::printf("P%d n=%u: N=%u nnz=%lld pcg %d iters, max u = %.8f\n", degree, n, N_dofs,
              (long long)gal.matrix().nnz(), info.iters, umax);
}

int main() {
  try {
    amd::Context ctx(0);
    run(ctx, 1, 64);
    run(ctx, 3, 12);
    // error handling: an unsupported degree is an LfException
    bool threw = false;
    try {
      auto m = amd::mesh::TPTriagMeshBuilder().setNumXCells(2).setNumYCells(2).Build();
      amd::uscalfe::FeSpaceLagrange bad(ctx, m, 2);
    } catch (const amd::LfException &e) {
      threw = e.code() == LFGPU_ENORULE;
    }
    EXPECT(threw, "degree 2 must raise LfException(LFGPU_ENORULE)");
  } catch (const std::exception &e) {
    std::printf("FAIL: exception %s\n", e.what());
    return 2;
  }
  std::printf(fails ? "FAILED\n" : "OK\n");
  return fails ? 1 : 0;
}
