// The LehrFEM++ call sites of examples/ellbvp_linfe/homDir_linfe_demo.cc:
// 130-150 bound to the device through lf_amd/lehrfem_binding.h, against the
// archetype of the reference interfaces (tests/cpp/lf_archetype.h).
//
// For P1 and P3 on a jittered mesh, the same call expressions
//     lf::assemble::AssembleMatrixLocally(0, dofh, dofh, elmat_builder, A);
//     lf::assemble::AssembleVectorLocally(0, dofh, elvec_builder, phi);
// run once with the reference's CPU providers (the generic cell loop) and
// once with the GPU providers (the constrained device overloads).  The COO
// triplets must agree in order, indices and values bit for bit, and so must
// the load vectors; the returning overload and the single-entity Eval give
// the same matrices; the fused device system (DeviceGalerkinSystem) equals
// setFromTriplets of the fixed triplets bit for bit, and its PCG solution
// solves the system.  Exit status 0 on success.
#include <cmath>
#include <cstdio>
#include <random>
#include <vector>

#include "tests/cpp/lf_archetype.h"
// the binding comes after the reference headers (here: their archetype)
#include "lehrfempp_amd/include/lf_amd/lehrfem_binding.h"

static int fails = 0;
#define EXPECT(cond, ...)                \
  do {                                   \
    if (!(cond)) {                       \
      std::printf("FAIL: " __VA_ARGS__); \
      std::printf("\n");                 \
      ++fails;                           \
    }                                    \
  } while (0)

// the reference's dispatch: a GPU provider selects the constrained overload
static_assert(lf::assemble::EntityMatrixProvider<lf::amd::binding::GpuReactionDiffusionElementMatrixProvider<
                  const lf::uscalfe::FeSpaceLagrangeO1<double>, lf::amd::binding::DeviceCoefficient,
                  lf::amd::binding::DeviceCoefficient>>);
static_assert(lf::amd::binding::BatchedEntityMatrixProvider<
              lf::amd::binding::GpuReactionDiffusionElementMatrixProvider<
                  const lf::uscalfe::FeSpaceLagrangeO1<double>, lf::amd::binding::DeviceCoefficient,
                  lf::amd::binding::DeviceCoefficient>>);

template <class TRIP>
static bool same_triplets(const TRIP &a, const TRIP &b) {
  if (a.size() != b.size()) return false;
  for (size_t k = 0; k < a.size(); ++k)
    if (a[k].r != b[k].r || a[k].c != b[k].c || a[k].v != b[k].v) return false;
  return true;
}

static void run(int degree, int n) {
  std::mt19937_64 rng(7 + degree);
  std::uniform_real_distribution<double> jit(-0.2 / n, 0.2 / n);
  auto mesh_p = lf::archetype::tp_mesh(n, [&](int) { return std::make_pair(jit(rng), jit(rng)); });
  std::shared_ptr<const lf::uscalfe::UniformScalarFESpace<double>> fe_space;
  if (degree == 1)
    fe_space = std::make_shared<const lf::uscalfe::FeSpaceLagrangeO1<double>>(mesh_p);
  else
    fe_space = std::make_shared<const lf::uscalfe::FeSpaceLagrangeO3<double>>(mesh_p);
  const lf::assemble::DofHandler &dofh{fe_space->LocGlobMap()};
  const lf::base::size_type N_dofs(dofh.NumDofs());

  // MeshFunctions in the reference's form (MeshFunctionGlobal of a lambda)
  auto mf_alpha = [](const lf::mesh::Entity &e, const lf::archetype::Dense &local) {
    const auto X = e.Geometry()->Global(local);
    std::vector<double> v(local.cols());
    for (int k = 0; k < local.cols(); ++k) v[k] = 1.0 + X(0, k) * X(0, k) + 0.5 * X(1, k);
    return v;
  };
  auto mf_gamma = [](const lf::mesh::Entity &, const lf::archetype::Dense &local) {
    return std::vector<double>(local.cols(), 0.5);
  };
  auto mf_f = [](const lf::mesh::Entity &e, const lf::archetype::Dense &local) {
    const auto X = e.Geometry()->Global(local);
    std::vector<double> v(local.cols());
    for (int k = 0; k < local.cols(); ++k) v[k] = std::cos(X(0, k)) + X(1, k);
    return v;
  };

  // the reference providers (CPU)
  lf::uscalfe::ReactionDiffusionElementMatrixProvider elmat_cpu(fe_space, mf_alpha, mf_gamma);
  lf::uscalfe::ScalarLoadElementVectorProvider elvec_cpu(fe_space, mf_f);
  // the device providers, same constructor arguments
  lf::amd::binding::GpuReactionDiffusionElementMatrixProvider elmat_gpu(fe_space, mf_alpha,
                                                                       mf_gamma);
  lf::amd::binding::GpuScalarLoadElementVectorProvider elvec_gpu(fe_space, mf_f);

  lf::assemble::COOMatrix<double> A_cpu(N_dofs, N_dofs), A_gpu(N_dofs, N_dofs);
  lf::assemble::AssembleMatrixLocally(0, dofh, dofh, elmat_cpu, A_cpu);
  lf::assemble::AssembleMatrixLocally(0, dofh, dofh, elmat_gpu, A_gpu);
  EXPECT(same_triplets(A_cpu.triplets(), A_gpu.triplets()),
         "P%d: device triplets differ from the reference loop's", degree);
  EXPECT(A_gpu.triplets().size() == (size_t)mesh_p->NumEntities(0) * (degree == 1 ? 9 : 100),
         "P%d: triplet count", degree);
  auto A_ret = lf::assemble::AssembleMatrixLocally<lf::assemble::COOMatrix<double>>(
      0, dofh, dofh, elmat_gpu);
  EXPECT(same_triplets(A_ret.triplets(), A_cpu.triplets()), "P%d: returning overload", degree);
  // single-entity fallback
  const lf::mesh::Entity &c7 = *mesh_p->Entities(0)[7];
  const auto M7 = elmat_gpu.Eval(c7), M7c = elmat_cpu.Eval(c7);
  bool eq7 = M7.rows() == M7c.rows() && M7.cols() == M7c.cols();
  for (int i = 0; eq7 && i < M7.rows(); ++i)
    for (int j = 0; j < M7.cols(); ++j) eq7 = eq7 && M7(i, j) == M7c(i, j);
  EXPECT(eq7, "P%d: Eval(cell) differs", degree);

  std::vector<double> phi_cpu(N_dofs, 0.0), phi_gpu(N_dofs, 0.0);
  lf::assemble::AssembleVectorLocally(0, dofh, elvec_cpu, phi_cpu);
  lf::assemble::AssembleVectorLocally(0, dofh, elvec_gpu, phi_gpu);
  EXPECT(phi_cpu == phi_gpu, "P%d: device load vector differs", degree);

  // device coefficients: gamma = 0.5 evaluated on the device gives the same
  // triplets as the MeshFunction evaluated on the host
  lf::amd::binding::GpuReactionDiffusionElementMatrixProvider elmat_dev(
      fe_space, mf_alpha, lf::amd::binding::DeviceCoefficient::Constant(0.5));
  lf::assemble::COOMatrix<double> A_dev(N_dofs, N_dofs);
  lf::assemble::AssembleMatrixLocally(0, dofh, dofh, elmat_dev, A_dev);
  EXPECT(same_triplets(A_dev.triplets(), A_cpu.triplets()), "P%d: device coefficient", degree);

  // boundary DOFs: flagEntitiesOnBoundary through the oracle (nodes and,
  // at P3, the interior DOFs of boundary edges)
  const auto &ah = static_cast<const lf::archetype::UniformDofHandler &>(dofh);
  const int nv = (int)mesh_p->NumEntities(2), nc = (int)mesh_p->NumEntities(0);
  std::vector<uint8_t> nf(nv), ef(ah.nedges());
  lfref_boundary_flags(nv, nc, ah.nedges(), ah.cell_e().data(), ah.edge_v().data(), nf.data(),
                       ef.data());
  std::vector<uint8_t> fixed(N_dofs, 0);
  for (int v = 0; v < nv; ++v) fixed[v] = nf[v];
  if (degree == 3)
    for (int e = 0; e < ah.nedges(); ++e)
      if (ef[e]) fixed[nv + 2 * e] = fixed[nv + 2 * e + 1] = 1;
  auto selector = [&](lf::assemble::gdof_idx_t k) { return std::make_pair(fixed[k] != 0, 0.0); };

  // FixFlaggedSolutionComponents (fix_dof.h:86-138) with zero prescribed
  // values on the triplets, then setFromTriplets on the device
  std::vector<int32_t> rr, cc;
  std::vector<double> vv;
  for (const auto &t : A_cpu.triplets())
    if (!fixed[t.r] && !fixed[t.c]) {
      rr.push_back(t.r);
      cc.push_back(t.c);
      vv.push_back(t.v);
    }
  for (lf::base::size_type k = 0; k < N_dofs; ++k)
    if (fixed[k]) {
      rr.push_back((int32_t)k);
      cc.push_back((int32_t)k);
      vv.push_back(1.0);
    }
  auto &ctx = lf::amd::binding::DefaultContext();
  lfgpu_csr *Aref = nullptr;
  lf::amd::check(lfgpu_coo_to_csr(ctx.get(), (int64_t)rr.size(), rr.data(), cc.data(), vv.data(),
                                  (int32_t)N_dofs, (int32_t)N_dofs, &Aref));
  std::vector<double> b_ref = phi_cpu;
  for (lf::base::size_type k = 0; k < N_dofs; ++k)
    if (fixed[k]) b_ref[k] = 0.0;

  lf::amd::binding::DeviceGalerkinSystem sys(elmat_gpu, elvec_gpu, selector);
  sys.Assemble();
  int64_t nnz_a = 0, nnz_b = 0;
  lf::amd::check(lfgpu_csr_info(Aref, nullptr, nullptr, &nnz_a));
  lf::amd::check(lfgpu_csr_info(sys.matrix(), nullptr, nullptr, &nnz_b));
  std::vector<int32_t> rpa(N_dofs + 1), rpb(N_dofs + 1), ca(nnz_a), cb(nnz_b);
  std::vector<double> va(nnz_a), vb(nnz_b);
  lf::amd::check(lfgpu_csr_download(Aref, rpa.data(), ca.data(), va.data()));
  lf::amd::check(lfgpu_csr_download(sys.matrix(), rpb.data(), cb.data(), vb.data()));
  EXPECT(rpa == rpb && ca == cb, "P%d: fused pattern differs from setFromTriplets", degree);
  EXPECT(va == vb, "P%d: fused values differ from setFromTriplets", degree);
  EXPECT(sys.rhs().host() == b_ref, "P%d: fused load vector differs", degree);
  lfgpu_csr_destroy(Aref);

  // the device solve in place of SparseLU: residual certificate on the host
  lfgpu_pcg_info info{};
  const auto x = sys.Solve(1e-10, LFGPU_PCG_FP64, &info);
  double rn = 0.0, bn = 0.0;
  for (lf::base::size_type r = 0; r < N_dofs; ++r) {
    double s = 0.0;
    for (int32_t k = rpb[r]; k < rpb[r + 1]; ++k) s += vb[k] * x[cb[k]];
    rn += (b_ref[r] - s) * (b_ref[r] - s);
    bn += b_ref[r] * b_ref[r];
  }
  EXPECT(info.converged && std::sqrt(rn) <= 1e-8 * std::sqrt(bn),
         "P%d: PCG residual %.3e (iters %d)", degree, std::sqrt(rn / bn), info.iters);
  std::printf("P%d n=%d: %u DOFs, %zu triplets, PCG %d iterations, relres %.2e\n", degree, n,
              N_dofs, A_gpu.triplets().size(), info.iters, std::sqrt(rn / bn));
}

int main() {
  try {
    run(1, 24);
    run(3, 12);
  } catch (const std::exception &e) {
    std::printf("FAIL: exception %s\n", e.what());
    return 1;
  }
  if (fails) {
    std::printf("%d check(s) failed\n", fails);
    return 1;
  }
  std::printf("OK\n");
  return 0;
}
