// lehrfem_binding.h -- liblfgpu behind LehrFEM++'s own interfaces.
//
// Include this header after the LehrFEM++ headers of a call site
// (lf/assemble, lf/uscalfe, lf/geometry, lf/quad).  It adds, without
// touching any reference type:
//
//  * lf::amd::binding::GpuReactionDiffusionElementMatrixProvider and
//    lf::amd::binding::GpuScalarLoadElementVectorProvider.  Their
//    constructors take the arguments of
//    lf::uscalfe::ReactionDiffusionElementMatrixProvider (fe_space, alpha,
//    gamma; uscalfe/loc_comp_ellbvp.h:125-147) and
//    lf::uscalfe::ScalarLoadElementVectorProvider (fe_space, f; :592-605),
//    with the same class template argument deduction.  The coefficients are
//    the reference's MeshFunctions (any callable (const Entity&, local
//    points) -> std::vector<double>, mesh/utils/mesh_function_traits.h:148-154),
//    evaluated on the host at the quadrature points of every cell, or
//    lf::amd::binding::DeviceCoefficient values that the device evaluates
//    itself.  Both providers model the reference concepts
//    EntityMatrixProvider / EntityVectorProvider (isActive, Eval;
//    assemble/assemble_concepts.h:95-99, :210-214): Eval(cell) is the
//    single-entity fallback.
//
//  * More-constrained overloads of lf::assemble::AssembleMatrixLocally and
//    lf::assemble::AssembleVectorLocally (assemble/assembler.h:114-186,
//    :298-327), selected by the concepts BatchedEntityMatrixProvider /
//    BatchedEntityVectorProvider, which refine the reference concepts.  An
//    existing call such as
//        lf::assemble::AssembleMatrixLocally(0, dofh, dofh, elmat_builder, A);
//    (examples/ellbvp_linfe/homDir_linfe_demo.cc:130-150) picks them up when
//    elmat_builder is a GPU provider, and nothing else changes: the
//    overloads read the mesh and the DofHandler through the reference
//    interfaces (Mesh::Entities / Index, Entity::SubEntities / Geometry,
//    DofHandler::GlobalDofIndices / NumDofs), compute every element matrix
//    (vector) in one device launch and add the triplets (entries) to the
//    caller's TMPMATRIX (VECTOR) in the reference's order: entities in
//    Entities(codim) order, then i, then j.  The reference's returning
//    overloads (assembler.h:209-249) forward to them as well.
//
//  * lf::amd::binding::DeviceGalerkinSystem: the same objects assembled into
//    a device-resident CSR matrix and load vector by the fused kernel
//    (AssembleMatrixLocally + AssembleVectorLocally +
//    FixFlaggedSolutionComponents + makeSparse, no host triplets), for the
//    device PCG solve that replaces SparseLU (homDir_linfe_demo.cc:171-174).
//
// Cells must be triangles (TriaO1 geometry) and the FE space Lagrangian of
// degree 1 or 3, as the device kernels support.  All device work runs on a
// process-wide context on device 0 (DefaultContext()).
#pragma once

#include <concepts>
#include <cstdint>
#include <map>
#include <memory>
#include <span>
#include <type_traits>
#include <utility>
#include <vector>

#include "lf_amd.h"

namespace lf::amd::binding {

// the context every binding object runs on (created on first use)
inline lf::amd::Context &DefaultContext() {
  static lf::amd::Context ctx(0);
  return ctx;
}

// A coefficient the device evaluates itself (no host evaluation per cell):
// the lfgpu_coeff kinds other than per-point data.
class DeviceCoefficient {
 public:
  static DeviceCoefficient Constant(double v) { return {LFGPU_COEFF_CONST, v, nullptr}; }
  static DeviceCoefficient SinSin(double scale = 1.0) { return {LFGPU_COEFF_SINSIN, scale, nullptr}; }
  // value * data[cell index] / value * exp(data[cell index])
  static DeviceCoefficient CellData(std::shared_ptr<const lf::amd::DeviceVector> d, double v = 1.0) {
    return {LFGPU_COEFF_CELL, v, std::move(d)};
  }
  static DeviceCoefficient ExpCellData(std::shared_ptr<const lf::amd::DeviceVector> d,
                                       double v = 1.0) {
    return {LFGPU_COEFF_EXPCELL, v, std::move(d)};
  }
  lfgpu_coeff c() const { return lfgpu_coeff{kind_, value_, data_ ? data_->get() : nullptr}; }

 private:
  DeviceCoefficient(int k, double v, std::shared_ptr<const lf::amd::DeviceVector> d)
      : kind_(k), value_(v), data_(std::move(d)) {}
  int kind_;
  double value_;
  std::shared_ptr<const lf::amd::DeviceVector> data_;
};

namespace detail {

// element matrix / vector types of the reference (Eigen::MatrixXd /
// Eigen::VectorXd), named through reference functions that return them
using ElemMat = std::remove_cvref_t<decltype(lf::geometry::Corners(
    std::declval<const lf::geometry::Geometry &>()))>;
using ElemVec = std::remove_cvref_t<decltype(std::declval<const lf::quad::QuadRule &>().Weights())>;

// The cell loop's view of a LehrFEM++ FE space, flattened once: vertex
// coordinates, cell -> vertex indices and GlobalDofIndices of every cell (by
// cell index), read through the Mesh / Entity / DofHandler interfaces.
struct FlatSpace {
  int degree = 0, nloc = 0;
  int64_t ncells = 0, nverts = 0, ndofs = 0;
  std::vector<double> xy;
  std::vector<int32_t> cell_v;
  std::vector<int64_t> cell_dofs;
  std::shared_ptr<const void> mesh_key;  // the mesh this was built from
};

template <class DOFH>
std::shared_ptr<FlatSpace> flatten(const DOFH &dofh, int degree) {
  auto mesh = dofh.Mesh();
  auto fs = std::make_shared<FlatSpace>();
  fs->degree = degree;
  fs->nloc = degree == 1 ? 3 : 10;
  fs->ncells = mesh->NumEntities(0);
  fs->nverts = mesh->NumEntities(2);
  fs->ndofs = dofh.NumDofs();
  fs->mesh_key = mesh;
  fs->xy.assign((size_t)fs->nverts * 2, 0.0);
  fs->cell_v.assign((size_t)fs->ncells * 3, 0);
  fs->cell_dofs.assign((size_t)fs->ncells * fs->nloc, 0);
  for (const lf::mesh::Entity *cell : mesh->Entities(0)) {
    if (cell->RefEl() != lf::base::RefEl::kTria())
      throw LfException(LFGPU_EINVAL, "lf::amd::binding: the device kernels take triangles only");
    const auto c = mesh->Index(*cell);
    const auto verts = cell->SubEntities(2);
    const auto X = lf::geometry::Corners(*cell->Geometry());
    for (int k = 0; k < 3; ++k) {
      const auto v = mesh->Index(*verts[k]);
      fs->cell_v[(size_t)c * 3 + k] = (int32_t)v;
      fs->xy[(size_t)v * 2] = X(0, k);
      fs->xy[(size_t)v * 2 + 1] = X(1, k);
    }
    const auto d = dofh.GlobalDofIndices(*cell);
    if ((int)d.size() != fs->nloc || (int)dofh.NumLocalDofs(*cell) != fs->nloc)
      throw LfException(LFGPU_EINVAL, "lf::amd::binding: unexpected number of local DOFs");
    for (int i = 0; i < fs->nloc; ++i) fs->cell_dofs[(size_t)c * fs->nloc + i] = d[i];
  }
  return fs;
}

// Device mesh + reference element of a flattened space, with the quadrature
// rule the reference providers select (make_QuadRule(kTria, 2 degree),
// uscalfe/loc_comp_ellbvp.h:227-228) handed to the device explicitly.
class DeviceSpace {
 public:
  explicit DeviceSpace(std::shared_ptr<FlatSpace> fs) : fs_(std::move(fs)) {
    auto &ctx = DefaultContext();
    // P1: the DOFs of a cell must be its vertex indices (lfgpu_mesh_create)
    const bool vertex_dofs = fs_->degree == 1 && [&] {
      for (size_t k = 0; k < fs_->cell_v.size(); ++k)
        if (fs_->cell_dofs[k] != fs_->cell_v[k]) return false;
      return fs_->ndofs == fs_->nverts;
    }();
    if (fs_->degree == 1 && !vertex_dofs)
      throw LfException(LFGPU_EINVAL, "lf::amd::binding: P1 DOFs must be the vertex indices");
    check(lfgpu_mesh_create(ctx.get(), fs_->nverts, fs_->xy.data(), fs_->ncells,
                            fs_->cell_v.data(), fs_->nloc,
                            fs_->degree == 1 ? nullptr : fs_->cell_dofs.data(), fs_->ndofs,
                            &mesh_));
    const lf::quad::QuadRule qr = lf::quad::make_QuadRule(lf::base::RefEl::kTria(),
                                                          2 * fs_->degree);
    nqp_ = (int)qr.NumPoints();
    std::vector<double> pts(2 * (size_t)nqp_), w(nqp_);
    for (int k = 0; k < nqp_; ++k) {
      pts[2 * k] = qr.Points()(0, k);
      pts[2 * k + 1] = qr.Points()(1, k);
      w[k] = qr.Weights()[k];
    }
    check(lfgpu_fe_create(ctx.get(), fs_->degree, nqp_, pts.data(), w.data(), &fe_));
  }
  ~DeviceSpace() {
    lfgpu_fe_destroy(fe_);
    lfgpu_mesh_destroy(mesh_);
  }
  DeviceSpace(const DeviceSpace &) = delete;
  DeviceSpace &operator=(const DeviceSpace &) = delete;
  const FlatSpace &flat() const { return *fs_; }
  lfgpu_mesh *mesh() const { return mesh_; }
  lfgpu_fe *fe() const { return fe_; }
  int nqp() const { return nqp_; }

 private:
  std::shared_ptr<FlatSpace> fs_;
  lfgpu_mesh *mesh_ = nullptr;
  lfgpu_fe *fe_ = nullptr;
  int nqp_ = 0;
};

template <class T>
inline constexpr bool is_device_coefficient_v = std::is_same_v<std::remove_cvref_t<T>,
                                                                DeviceCoefficient>;

// A provider coefficient as the device sees it: device kinds pass through;
// a MeshFunction is evaluated on the host at the quadrature points of every
// cell (in cell index order) and handed over as per-point data.
template <class COEFF, class MESH>
lfgpu_coeff device_coeff(const COEFF &coeff, const MESH &mesh, int degree, int nqp,
                         std::unique_ptr<lf::amd::DeviceVector> &storage) {
  if constexpr (is_device_coefficient_v<COEFF>) {
    return coeff.c();
  } else {
    const lf::quad::QuadRule qr = lf::quad::make_QuadRule(lf::base::RefEl::kTria(), 2 * degree);
    std::vector<double> vals((size_t)mesh.NumEntities(0) * nqp);
    for (const lf::mesh::Entity *cell : mesh.Entities(0)) {
      const auto c = mesh.Index(*cell);
      const auto v = coeff(*cell, qr.Points());
      if ((int)v.size() != nqp)
        throw LfException(LFGPU_EINVAL, "lf::amd::binding: a MeshFunction returned the wrong "
                                        "number of values");
      for (int k = 0; k < nqp; ++k) vals[(size_t)c * nqp + k] = static_cast<double>(v[k]);
    }
    storage = std::make_unique<lf::amd::DeviceVector>(DefaultContext(), vals);
    return lfgpu_coeff{LFGPU_COEFF_QP, 1.0, storage->get()};
  }
}

// degree of the FE space's triangle layout (UniformScalarFESpace::
// ShapeFunctionLayout(RefEl), uniform_scalar_fe_space.h:133)
template <class FES>
int space_degree(const FES &fes) {
  const auto *layout = fes.ShapeFunctionLayout(lf::base::RefEl::kTria());
  if (layout == nullptr)
    throw LfException(LFGPU_ENORULE, "lf::amd::binding: the FE space has no triangle layout");
  const int d = (int)layout->Degree();
  if (d != 1 && d != 3)
    throw LfException(LFGPU_ENORULE, "lf::amd::binding: only Lagrange degree 1 and 3 run on "
                                     "the device");
  return d;
}

}  // namespace detail

// ReactionDiffusionElementMatrixProvider (uscalfe/loc_comp_ellbvp.h:85-339)
// with the element matrices of all cells computed on the device.
template <class FESPACE, class DIFF_COEFF, class REACTION_COEFF>
class GpuReactionDiffusionElementMatrixProvider {
 public:
  using ElemMat = detail::ElemMat;
  // marks the provider for the device overload of AssembleMatrixLocally
  using lf_amd_batched = void;

  GpuReactionDiffusionElementMatrixProvider(std::shared_ptr<const FESPACE> fe_space,
                                            DIFF_COEFF alpha, REACTION_COEFF gamma)
      : fes_(std::move(fe_space)), alpha_(std::move(alpha)), gamma_(std::move(gamma)),
        degree_(detail::space_degree(*fes_)) {}
  GpuReactionDiffusionElementMatrixProvider(const GpuReactionDiffusionElementMatrixProvider &) =
      delete;
  GpuReactionDiffusionElementMatrixProvider(GpuReactionDiffusionElementMatrixProvider &&) noexcept =
      default;

  bool isActive(const lf::mesh::Entity & /*cell*/) { return true; }

  // single-entity fallback: the cell's matrix out of the batched result
  ElemMat Eval(const lf::mesh::Entity &cell) {
    const auto &m = ElementMatrices(fes_->LocGlobMap());
    const int n = nloc();
    const auto c = fes_->Mesh()->Index(cell);
    ElemMat out(n, n);
    for (int i = 0; i < n; ++i)
      for (int j = 0; j < n; ++j) out(i, j) = m[((size_t)c * n + i) * n + j];
    return out;
  }

  // every cell's element matrix, row-major, by cell index (computed once)
  const std::vector<double> &ElementMatrices(const lf::assemble::DofHandler &dofh) {
    if (mats_.empty()) {
      ensure_space(dofh);
      std::unique_ptr<lf::amd::DeviceVector> sa, sg;
      const auto &mesh = *fes_->Mesh();
      const lfgpu_coeff a = detail::device_coeff(alpha_, mesh, degree_, dev_->nqp(), sa);
      const lfgpu_coeff g = detail::device_coeff(gamma_, mesh, degree_, dev_->nqp(), sg);
      const int n = nloc();
      lf::amd::DeviceVector out(DefaultContext(), dev_->flat().ncells * n * n);
      check(lfgpu_element_matrices(DefaultContext().get(), dev_->mesh(), dev_->fe(), &a, &g,
                                   out.get()));
      mats_ = out.host();
    }
    return mats_;
  }

  int nloc() const { return degree_ == 1 ? 3 : 10; }
  int degree() const { return degree_; }
  const FESPACE &space() const { return *fes_; }
  const DIFF_COEFF &alpha() const { return alpha_; }
  const REACTION_COEFF &gamma() const { return gamma_; }

 private:
  void ensure_space(const lf::assemble::DofHandler &dofh) {
    if (!dev_) dev_ = std::make_shared<detail::DeviceSpace>(detail::flatten(dofh, degree_));
  }
  std::shared_ptr<const FESPACE> fes_;
  DIFF_COEFF alpha_;
  REACTION_COEFF gamma_;
  int degree_;
  std::shared_ptr<detail::DeviceSpace> dev_;
  std::vector<double> mats_;
};

template <class PTR, class DIFF_COEFF, class REACTION_COEFF>
GpuReactionDiffusionElementMatrixProvider(PTR fe_space, DIFF_COEFF alpha, REACTION_COEFF gamma)
    -> GpuReactionDiffusionElementMatrixProvider<typename PTR::element_type, DIFF_COEFF,
                                                 REACTION_COEFF>;

// ScalarLoadElementVectorProvider (uscalfe/loc_comp_ellbvp.h:562-746) with
// the element vectors of all cells computed on the device.
template <class FESPACE, class FUNCTOR>
class GpuScalarLoadElementVectorProvider {
 public:
  using ElemVec = detail::ElemVec;
  using lf_amd_batched = void;

  GpuScalarLoadElementVectorProvider(std::shared_ptr<const FESPACE> fe_space, FUNCTOR f)
      : fes_(std::move(fe_space)), f_(std::move(f)), degree_(detail::space_degree(*fes_)) {}
  GpuScalarLoadElementVectorProvider(const GpuScalarLoadElementVectorProvider &) = delete;
  GpuScalarLoadElementVectorProvider(GpuScalarLoadElementVectorProvider &&) noexcept = default;

  bool isActive(const lf::mesh::Entity & /*cell*/) { return true; }

  ElemVec Eval(const lf::mesh::Entity &cell) {
    const auto &v = ElementVectors(fes_->LocGlobMap());
    const int n = nloc();
    const auto c = fes_->Mesh()->Index(cell);
    ElemVec out(n);
    for (int i = 0; i < n; ++i) out[i] = v[(size_t)c * n + i];
    return out;
  }

  const std::vector<double> &ElementVectors(const lf::assemble::DofHandler &dofh) {
    if (vecs_.empty()) {
      if (!dev_) dev_ = std::make_shared<detail::DeviceSpace>(detail::flatten(dofh, degree_));
      std::unique_ptr<lf::amd::DeviceVector> sf;
      const lfgpu_coeff f = detail::device_coeff(f_, *fes_->Mesh(), degree_, dev_->nqp(), sf);
      lf::amd::DeviceVector out(DefaultContext(), dev_->flat().ncells * nloc());
      check(lfgpu_element_vectors(DefaultContext().get(), dev_->mesh(), dev_->fe(), &f,
                                  out.get()));
      vecs_ = out.host();
    }
    return vecs_;
  }

  int nloc() const { return degree_ == 1 ? 3 : 10; }
  const FUNCTOR &f() const { return f_; }

 private:
  std::shared_ptr<const FESPACE> fes_;
  FUNCTOR f_;
  int degree_;
  std::shared_ptr<detail::DeviceSpace> dev_;
  std::vector<double> vecs_;
};

template <class PTR, class FUNCTOR>
GpuScalarLoadElementVectorProvider(PTR fe_space, FUNCTOR f)
    -> GpuScalarLoadElementVectorProvider<typename PTR::element_type, FUNCTOR>;

// The reference concepts refined by the batched interface: subsumption makes
// the overloads below more constrained than the reference's.
template <class EMP>
concept BatchedEntityMatrixProvider =
    lf::assemble::EntityMatrixProvider<EMP> &&
    requires(EMP &emp, const lf::assemble::DofHandler &dofh) {
      typename EMP::lf_amd_batched;
      { emp.ElementMatrices(dofh) } -> std::same_as<const std::vector<double> &>;
      { emp.nloc() } -> std::convertible_to<int>;
    };

template <class EVP>
concept BatchedEntityVectorProvider =
    lf::assemble::EntityVectorProvider<EVP> &&
    requires(EVP &evp, const lf::assemble::DofHandler &dofh) {
      typename EVP::lf_amd_batched;
      { evp.ElementVectors(dofh) } -> std::same_as<const std::vector<double> &>;
      { evp.nloc() } -> std::convertible_to<int>;
    };

// The fused device path for the same reference objects: pattern + scatter
// plan built once from the DofHandler and the selector of
// FixFlaggedSolutionComponents (fix_dof.h:86-138; fixed values are zero),
// then every Assemble() writes the CSR values and the load vector on the
// device in the reference's summation order.
class DeviceGalerkinSystem {
 public:
  template <class FES, class DIFF_COEFF, class REACTION_COEFF, class FUNCTOR, class SELECTOR>
  DeviceGalerkinSystem(GpuReactionDiffusionElementMatrixProvider<FES, DIFF_COEFF, REACTION_COEFF> &emp,
                       GpuScalarLoadElementVectorProvider<FES, FUNCTOR> &evp,
                       SELECTOR &&selector)
      : dofh_(&emp.space().LocGlobMap()) {
    auto &ctx = DefaultContext();
    const int degree = emp.degree();
    space_ = std::make_shared<detail::DeviceSpace>(detail::flatten(*dofh_, degree));
    const int64_t nd = space_->flat().ndofs;
    std::vector<uint8_t> flags((size_t)nd, 0);
    for (int64_t k = 0; k < nd; ++k) flags[k] = selector((lf::assemble::gdof_idx_t)k).first ? 1 : 0;
    check(lfgpu_plan_create(ctx.get(), space_->mesh(), flags.data(), &plan_));
    const auto &mesh = *emp.space().Mesh();
    const int nqp = space_->nqp();
    alpha_ = detail::device_coeff(emp.alpha(), mesh, degree, nqp, sa_);
    gamma_ = detail::device_coeff(emp.gamma(), mesh, degree, nqp, sg_);
    f_ = detail::device_coeff(evp_coeff(evp), mesh, degree, nqp, sf_);
    rhs_ = std::make_unique<lf::amd::DeviceVector>(ctx, nd);
  }
  ~DeviceGalerkinSystem() { lfgpu_plan_destroy(plan_); }
  DeviceGalerkinSystem(const DeviceGalerkinSystem &) = delete;
  DeviceGalerkinSystem &operator=(const DeviceGalerkinSystem &) = delete;

  // fused assembly; returns the device time in ms
  double Assemble() {
    auto &ctx = DefaultContext();
    check(lfgpu_assemble_galerkin(ctx.get(), plan_, space_->fe(), &alpha_, &gamma_, &f_, nullptr,
                                  rhs_->get()));
    return ctx.last_ms();
  }
  lfgpu_csr *matrix() const {
    lfgpu_csr *c = nullptr;
    check(lfgpu_plan_csr(plan_, &c));
    return c;
  }
  const lf::amd::DeviceVector &rhs() const { return *rhs_; }
  // Jacobi-PCG on the device in place of SparseLU::solve (x starts at zero)
  std::vector<double> Solve(double rtol = 1e-10, int mode = LFGPU_PCG_FP64,
                            lfgpu_pcg_info *info_out = nullptr) const {
    auto &ctx = DefaultContext();
    lf::amd::DeviceVector x(ctx, rhs_->size());
    check(lfgpu_vec_fill(x.get(), 0.0));
    lfgpu_pcg_opts o{rtol, (int)(2 * rhs_->size() + 100), mode, 64, 0.0, 0, 0, 0};
    lfgpu_pcg_info info{};
    check(lfgpu_pcg(ctx.get(), matrix(), rhs_->get(), x.get(), &o, &info));
    if (info_out) *info_out = info;
    return x.host();
  }

 private:
  template <class FES, class FUNCTOR>
  static const FUNCTOR &evp_coeff(GpuScalarLoadElementVectorProvider<FES, FUNCTOR> &evp) {
    return evp.f();
  }
  const lf::assemble::DofHandler *dofh_;
  std::shared_ptr<detail::DeviceSpace> space_;
  lfgpu_plan *plan_ = nullptr;
  lfgpu_coeff alpha_{}, gamma_{}, f_{};
  std::unique_ptr<lf::amd::DeviceVector> sa_, sg_, sf_, rhs_;
};

}  // namespace lf::amd::binding

namespace lf::assemble {

// AssembleMatrixLocally (assembler.h:114-186) for a device provider: one
// launch computes every element matrix, then the triplets are added in the
// reference's order.  Codim != 0, or different trial and test handlers, take
// the reference loop over the provider's single-entity Eval.
template <typename TMPMATRIX, lf::amd::binding::BatchedEntityMatrixProvider ENTITY_MATRIX_PROVIDER>
void AssembleMatrixLocally(dim_t codim, const DofHandler &dof_handler_trial,
                           const DofHandler &dof_handler_test,
                           ENTITY_MATRIX_PROVIDER &entity_matrix_provider, TMPMATRIX &matrix) {
  auto mesh = dof_handler_trial.Mesh();
  if (codim != 0 || &dof_handler_trial != &dof_handler_test) {
    for (const lf::mesh::Entity *entity : mesh->Entities(codim)) {
      if (!entity_matrix_provider.isActive(*entity)) continue;
      const auto nr = dof_handler_test.NumLocalDofs(*entity);
      const auto nc = dof_handler_trial.NumLocalDofs(*entity);
      const auto row_idx = dof_handler_test.GlobalDofIndices(*entity);
      const auto col_idx = dof_handler_trial.GlobalDofIndices(*entity);
      const auto elem_mat = entity_matrix_provider.Eval(*entity);
      for (int i = 0; i < (int)nr; ++i)
        for (int j = 0; j < (int)nc; ++j) matrix.AddToEntry(row_idx[i], col_idx[j], elem_mat(i, j));
    }
    return;
  }
  const std::vector<double> &mats = entity_matrix_provider.ElementMatrices(dof_handler_trial);
  const int n = entity_matrix_provider.nloc();
  for (const lf::mesh::Entity *cell : mesh->Entities(0)) {
    if (!entity_matrix_provider.isActive(*cell)) continue;
    const auto row_idx = dof_handler_test.GlobalDofIndices(*cell);
    const auto col_idx = dof_handler_trial.GlobalDofIndices(*cell);
    const double *m = mats.data() + (size_t)mesh->Index(*cell) * n * n;
    for (int i = 0; i < n; ++i)
      for (int j = 0; j < n; ++j) matrix.AddToEntry(row_idx[i], col_idx[j], m[i * n + j]);
  }
}

// AssembleVectorLocally (assembler.h:298-327) for a device provider.
template <typename VECTOR, lf::amd::binding::BatchedEntityVectorProvider ENTITY_VECTOR_PROVIDER>
void AssembleVectorLocally(dim_t codim, const DofHandler &dof_handler,
                           ENTITY_VECTOR_PROVIDER &entity_vector_provider, VECTOR &resultvector) {
  auto mesh = dof_handler.Mesh();
  if (codim != 0) {
    for (const lf::mesh::Entity *entity : mesh->Entities(codim)) {
      if (!entity_vector_provider.isActive(*entity)) continue;
      const auto len = dof_handler.NumLocalDofs(*entity);
      const auto dof_idx = dof_handler.GlobalDofIndices(*entity);
      const auto elem_vec = entity_vector_provider.Eval(*entity);
      for (int i = 0; i < (int)len; ++i) resultvector[dof_idx[i]] += elem_vec[i];
    }
    return;
  }
  const std::vector<double> &vecs = entity_vector_provider.ElementVectors(dof_handler);
  const int n = entity_vector_provider.nloc();
  for (const lf::mesh::Entity *cell : mesh->Entities(0)) {
    if (!entity_vector_provider.isActive(*cell)) continue;
    const auto dof_idx = dof_handler.GlobalDofIndices(*cell);
    const double *v = vecs.data() + (size_t)mesh->Index(*cell) * n;
    for (int i = 0; i < n; ++i) resultvector[dof_idx[i]] += v[i];
  }
}

}  // namespace lf::assemble
