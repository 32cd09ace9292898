// lf_amd.h -- header-only C++ facade over the liblfgpu C ABI that keeps
// LehrFEM++'s calling pattern for the Galerkin pipeline
// (examples/ellbvp_linfe/homDir_linfe_demo.cc:114-175):
//
//   auto fe_space = std::make_shared<lf::amd::uscalfe::FeSpaceLagrangeO1>(ctx, mesh);
//   const auto &dofh = fe_space->LocGlobMap();
//   lf::amd::assemble::COOMatrix<double> A(dofh.NumDofs(), dofh.NumDofs());
//   lf::amd::uscalfe::ReactionDiffusionElementMatrixProvider elmat(fe_space, alpha, gamma);
//   lf::amd::assemble::AssembleMatrixLocally(0, dofh, dofh, elmat, A);
//   std::vector<double> phi(dofh.NumDofs(), 0.0);
//   lf::amd::uscalfe::ScalarLoadElementVectorProvider elvec(fe_space, f);
//   lf::amd::assemble::AssembleVectorLocally(0, dofh, elvec, phi);
//   lf::amd::assemble::FixFlaggedSolutionComponents<double>(selector, A, phi);
//   auto A_crs = A.makeSparse();              // device CSR (setFromTriplets)
//
// The element matrices and vectors of the GPU providers are computed by the
// device kernels for all cells at once (the batched overloads below); the
// COO container, DOF handler and Dirichlet fix keep the reference's
// semantics on the host.  For the fast device-resident path (no triplets on
// the host at all) use assemble::GalerkinAssembler.
// Errors are reported as lf::amd::LfException (LfException in the
// reference, base/lf_exception.h:21-35).
#pragma once

#include <algorithm>
#include <cstdint>
#include <functional>
#include <memory>
#include <span>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "../../../include/lfgpu.h"

namespace lf::amd {

using size_type = unsigned int;  // base/types.h:20
using glb_idx_t = unsigned int;  // base/types.h:24
using gdof_idx_t = int64_t;      // assemble/assembly_types.h:22
using dim_t = unsigned int;

class LfException : public std::runtime_error {
 public:
  LfException(int code, const std::string &what) : std::runtime_error(what), code_(code) {}
  int code() const { return code_; }

 private:
  int code_;
};

inline void check(int rc) {
  if (rc != LFGPU_OK) throw LfException(rc, lfgpu_last_error());
}

// ---------------------------------------------------------------- runtime
class Context {
 public:
  explicit Context(int device = 0) { check(lfgpu_ctx_create(device, &h_)); }
  ~Context() { lfgpu_ctx_destroy(h_); }
  Context(const Context &) = delete;
  Context &operator=(const Context &) = delete;
  lfgpu_ctx *get() const { return h_; }
  void sync() const { check(lfgpu_ctx_sync(h_)); }
  double last_ms() const {
    double ms = 0;
    check(lfgpu_ctx_last_ms(h_, &ms));
    return ms;
  }

 private:
  lfgpu_ctx *h_ = nullptr;
};

class DeviceVector {
 public:
  DeviceVector(const Context &ctx, int64_t n) : n_(n) {
    check(lfgpu_vec_create(ctx.get(), n, nullptr, &h_));
  }
  DeviceVector(const Context &ctx, const std::vector<double> &v) : n_((int64_t)v.size()) {
    check(lfgpu_vec_create(ctx.get(), n_, v.data(), &h_));
  }
  ~DeviceVector() { lfgpu_vec_destroy(h_); }
  DeviceVector(const DeviceVector &) = delete;
  DeviceVector &operator=(const DeviceVector &) = delete;
  lfgpu_vec *get() const { return h_; }
  int64_t size() const { return n_; }
  std::vector<double> host() const {
    std::vector<double> v(n_);
    check(lfgpu_vec_download(h_, v.data(), n_));
    return v;
  }
  void upload(const std::vector<double> &v) { check(lfgpu_vec_upload(h_, v.data(), (int64_t)v.size())); }

 private:
  int64_t n_ = 0;
  lfgpu_vec *h_ = nullptr;
};

// ------------------------------------------------------------------- mesh
namespace mesh {

// Triangle mesh in hybrid2d::Mesh numbering: vertex coordinates, cell ->
// vertex table, edges (predefined first, then (min,max) order) and the
// relative edge orientations.
struct TriaMesh {
  std::vector<double> xy;       // [nv x 2]
  std::vector<int32_t> cell_v;  // [ncells x 3]
  std::vector<int32_t> edge_v;  // [nedges x 2]
  std::vector<int32_t> cell_e;  // [ncells x 3]
  std::vector<int8_t> cell_ori; // [ncells x 3]
  size_type NumNodes() const { return (size_type)(xy.size() / 2); }
  size_type NumCells() const { return (size_type)(cell_v.size() / 3); }
  size_type NumEdges() const { return (size_type)(edge_v.size() / 2); }

  // builds edges + orientations from the cells (and predefined edges)
  void finalize(const std::vector<int32_t> &pre_edges = {}) {
    const int64_t nc = NumCells();
    const int64_t npre = (int64_t)pre_edges.size() / 2;
    std::vector<int32_t> ev(2 * (npre + 3 * nc));
    cell_e.assign(3 * nc, 0);
    cell_ori.assign(3 * nc, 0);
    int64_t ne = 0;
    check(lfgpu_mesh_edges(NumNodes(), nc, cell_v.data(), npre,
                           npre ? pre_edges.data() : nullptr, npre + 3 * nc, ev.data(),
                           cell_e.data(), cell_ori.data(), &ne));
    ev.resize(2 * ne);
    edge_v.swap(ev);
  }
};

// TPTriagMeshBuilder (mesh/utils/tp_triag_mesh_builder.cc:18-178)
class TPTriagMeshBuilder {
 public:
  TPTriagMeshBuilder &setBottomLeftCorner(double x, double y) {
    x0_ = x;
    y0_ = y;
    return *this;
  }
  TPTriagMeshBuilder &setTopRightCorner(double x, double y) {
    x1_ = x;
    y1_ = y;
    return *this;
  }
  TPTriagMeshBuilder &setNumXCells(size_type n) {
    nx_ = n;
    return *this;
  }
  TPTriagMeshBuilder &setNumYCells(size_type n) {
    ny_ = n;
    return *this;
  }
  std::shared_ptr<TriaMesh> Build() const {
    auto m = std::make_shared<TriaMesh>();
    const int64_t nv = (int64_t)(nx_ + 1) * (ny_ + 1), nc = 2 * (int64_t)nx_ * ny_;
    const int64_t ne = (int64_t)nx_ * ny_ + (int64_t)(nx_ + 1) * ny_ + (int64_t)nx_ * (ny_ + 1);
    m->xy.resize(2 * nv);
    m->cell_v.resize(3 * nc);
    std::vector<int32_t> pre(2 * ne);
    check(lfgpu_tp_triag_mesh((int32_t)nx_, (int32_t)ny_, x0_, y0_, x1_, y1_, m->xy.data(),
                              m->cell_v.data(), pre.data()));
    m->finalize(pre);
    return m;
  }

 private:
  double x0_ = 0, y0_ = 0, x1_ = 1, y1_ = 1;
  size_type nx_ = 1, ny_ = 1;
};

// flagEntitiesOnBoundary (mesh/utils/special_entity_sets.cc:40-62) for
// codim 2 (nodes) and codim 1 (edges).
inline std::vector<bool> flagEntitiesOnBoundary(const TriaMesh &m, dim_t codim) {
  std::vector<int> cnt(m.NumEdges(), 0);
  for (int32_t e : m.cell_e) cnt[e]++;
  std::vector<bool> f(codim == 2 ? m.NumNodes() : m.NumEdges(), false);
  for (size_type e = 0; e < m.NumEdges(); ++e)
    if (cnt[e] == 1) {
      if (codim == 1) {
        f[e] = true;
      } else {
        f[m.edge_v[2 * e]] = true;
        f[m.edge_v[2 * e + 1]] = true;
      }
    }
  return f;
}

}  // namespace mesh

// --------------------------------------------------------------- assemble
namespace assemble {

// UniformFEDofHandler (assemble/dofhandler.cc:86-305) for {node: 1,
// edge: ndof_edge, cell: ndof_cell}; cells are addressed by index.
class UniformFEDofHandler {
 public:
  UniformFEDofHandler(std::shared_ptr<const mesh::TriaMesh> m, int ndof_edge, int ndof_cell)
      : mesh_(std::move(m)), nloc_(3 + 3 * ndof_edge + ndof_cell) {
    dofs_.resize((size_t)mesh_->NumCells() * nloc_);
    check(lfgpu_dof_table(mesh_->NumNodes(), mesh_->NumCells(), mesh_->NumEdges(),
                          mesh_->cell_v.data(), mesh_->cell_e.data(), mesh_->cell_ori.data(),
                          ndof_edge, ndof_cell, dofs_.data(), &ndofs_));
    edge_dofs_ = ndof_edge;
  }
  size_type NumDofs() const { return (size_type)ndofs_; }
  size_type NumLocalDofs(glb_idx_t /*cell*/) const { return (size_type)nloc_; }
  std::span<const gdof_idx_t> GlobalDofIndices(glb_idx_t cell) const {
    return {dofs_.data() + (size_t)cell * nloc_, (size_t)nloc_};
  }
  const std::vector<gdof_idx_t> &table() const { return dofs_; }
  std::shared_ptr<const mesh::TriaMesh> Mesh() const { return mesh_; }
  int edge_dofs() const { return edge_dofs_; }

 private:
  std::shared_ptr<const mesh::TriaMesh> mesh_;
  int nloc_;
  int edge_dofs_ = 0;
  int64_t ndofs_ = 0;
  std::vector<gdof_idx_t> dofs_;
};

class CsrMatrix;

// COOMatrix (assemble/coomatrix.h:51-260): triplets in insertion order.
template <typename SCALAR>
class COOMatrix {
 public:
  struct Triplet {
    int32_t r, c;
    SCALAR v;
    int32_t row() const { return r; }
    int32_t col() const { return c; }
    SCALAR value() const { return v; }
  };
  using TripletVec = std::vector<Triplet>;
  COOMatrix(size_type rows, size_type cols) : rows_(rows), cols_(cols) {}
  size_type rows() const { return rows_; }
  size_type cols() const { return cols_; }
  void AddToEntry(gdof_idx_t i, gdof_idx_t j, SCALAR inc) {
    rows_ = std::max<size_type>(rows_, (size_type)i + 1);
    cols_ = std::max<size_type>(cols_, (size_type)j + 1);
    trip_.push_back(Triplet{(int32_t)i, (int32_t)j, inc});
  }
  void setZero() { trip_.clear(); }
  template <typename PREDICATE>
  void setZero(PREDICATE &&pred) {
    trip_.erase(std::remove_if(trip_.begin(), trip_.end(),
                               [&](const Triplet &t) { return pred(t.r, t.c); }),
                trip_.end());
  }
  TripletVec &triplets() { return trip_; }
  const TripletVec &triplets() const { return trip_; }
  // resvec += M (alpha vec), in triplet order (coomatrix.h:248-260)
  template <typename V, typename R>
  void MatVecMult(SCALAR alpha, const V &vec, R &resvec) const {
    for (const Triplet &t : trip_) resvec[t.r] += t.v * (alpha * vec[t.c]);
  }
  // makeSparse (coomatrix.h:172-180): setFromTriplets on the device
  CsrMatrix makeSparse(const Context &ctx) const;

 private:
  size_type rows_, cols_;
  TripletVec trip_;
};

class CsrMatrix {
 public:
  CsrMatrix(lfgpu_csr *h, bool owned) : h_(h), owned_(owned) {}
  CsrMatrix(CsrMatrix &&o) noexcept : h_(o.h_), owned_(o.owned_) { o.h_ = nullptr; }
  ~CsrMatrix() {
    if (owned_ && h_) lfgpu_csr_destroy(h_);
  }
  lfgpu_csr *get() const { return h_; }
  int64_t nnz() const {
    int64_t n = 0;
    check(lfgpu_csr_info(h_, nullptr, nullptr, &n));
    return n;
  }
  int32_t rows() const {
    int32_t r = 0;
    check(lfgpu_csr_info(h_, &r, nullptr, nullptr));
    return r;
  }
  void download(std::vector<int32_t> &rp, std::vector<int32_t> &col, std::vector<double> &val) const {
    rp.resize(rows() + 1);
    col.resize(nnz());
    val.resize(nnz());
    check(lfgpu_csr_download(h_, rp.data(), col.data(), val.data()));
  }

 private:
  lfgpu_csr *h_;
  bool owned_;
};

template <typename SCALAR>
CsrMatrix COOMatrix<SCALAR>::makeSparse(const Context &ctx) const {
  if (!(rows_ > 0 && cols_ > 0))
    throw LfException(LFGPU_EINVAL, "matrix has zero rows or columns, this is probably an error.");
  std::vector<int32_t> r(trip_.size()), c(trip_.size());
  std::vector<double> v(trip_.size());
  for (size_t k = 0; k < trip_.size(); ++k) {
    r[k] = trip_[k].r;
    c[k] = trip_[k].c;
    v[k] = trip_[k].v;
  }
  lfgpu_csr *h = nullptr;
  check(lfgpu_coo_to_csr(ctx.get(), (int64_t)trip_.size(), r.data(), c.data(), v.data(),
                         (int32_t)rows_, (int32_t)cols_, &h));
  return CsrMatrix(h, true);
}

// FixFlaggedSolutionComponents (assemble/fix_dof.h:86-138)
template <typename SCALAR, typename SELECTOR, typename RHSVECTOR>
void FixFlaggedSolutionComponents(SELECTOR &&selectvals, COOMatrix<SCALAR> &A, RHSVECTOR &b) {
  const size_type N = A.cols();
  if (A.rows() != N || (size_t)N != (size_t)b.size())
    throw LfException(LFGPU_EINVAL, "FixFlaggedSolutionComponents: size mismatch");
  std::vector<SCALAR> tmp(N, SCALAR());
  for (gdof_idx_t k = 0; k < (gdof_idx_t)N; ++k) {
    const auto s = selectvals(k);
    if (s.first) tmp[k] = s.second;
  }
  A.MatVecMult(-1.0, tmp, b);
  for (gdof_idx_t k = 0; k < (gdof_idx_t)N; ++k) {
    const auto s = selectvals(k);
    if (s.first) b[k] = s.second;
  }
  A.setZero([&](gdof_idx_t i, gdof_idx_t j) { return selectvals(i).first || selectvals(j).first; });
  for (gdof_idx_t k = 0; k < (gdof_idx_t)N; ++k)
    if (selectvals(k).first) A.AddToEntry(k, k, 1.0);
}

}  // namespace assemble

// ---------------------------------------------------------------- uscalfe
namespace uscalfe {

// Lagrange FE space of degree 1 or 3 on a triangle mesh, with its device
// mesh/DOF table and reference element tables.
class FeSpaceLagrange {
 public:
  FeSpaceLagrange(const Context &ctx, std::shared_ptr<const mesh::TriaMesh> m, int degree)
      : ctx_(&ctx), degree_(degree),
        dofh_(m, degree == 3 ? 2 : 0, degree == 3 ? 1 : 0) {
    if (degree != 1 && degree != 3)
      throw LfException(LFGPU_ENORULE, "only Lagrange degree 1 and 3 are available");
    const auto &mm = *m;
    check(lfgpu_mesh_create(ctx.get(), mm.NumNodes(), mm.xy.data(), mm.NumCells(),
                            mm.cell_v.data(), degree == 1 ? 3 : 10,
                            degree == 1 ? nullptr : dofh_.table().data(), dofh_.NumDofs(),
                            &dmesh_));
    check(lfgpu_fe_create(ctx.get(), degree, 0, nullptr, nullptr, &fe_));
  }
  ~FeSpaceLagrange() {
    lfgpu_fe_destroy(fe_);
    lfgpu_mesh_destroy(dmesh_);
  }
  FeSpaceLagrange(const FeSpaceLagrange &) = delete;
  FeSpaceLagrange &operator=(const FeSpaceLagrange &) = delete;
  const assemble::UniformFEDofHandler &LocGlobMap() const { return dofh_; }
  std::shared_ptr<const mesh::TriaMesh> Mesh() const { return dofh_.Mesh(); }
  int Degree() const { return degree_; }
  const Context &ctx() const { return *ctx_; }
  lfgpu_mesh *device_mesh() const { return dmesh_; }
  lfgpu_fe *device_fe() const { return fe_; }
  // flags of the DOFs on the boundary (boundary nodes + boundary edge DOFs)
  std::vector<uint8_t> BoundaryDofFlags() const {
    const auto &m = *Mesh();
    auto nb = mesh::flagEntitiesOnBoundary(m, 2);
    auto eb = mesh::flagEntitiesOnBoundary(m, 1);
    std::vector<uint8_t> f(dofh_.NumDofs(), 0);
    for (size_type v = 0; v < m.NumNodes(); ++v) f[v] = nb[v];
    for (size_type e = 0; e < m.NumEdges() && dofh_.edge_dofs(); ++e)
      if (eb[e])
        for (int j = 0; j < dofh_.edge_dofs(); ++j) f[m.NumNodes() + 2 * e + j] = 1;
    return f;
  }

 private:
  const Context *ctx_;
  int degree_;
  assemble::UniformFEDofHandler dofh_;
  lfgpu_mesh *dmesh_ = nullptr;
  lfgpu_fe *fe_ = nullptr;
};

struct FeSpaceLagrangeO1 : FeSpaceLagrange {
  FeSpaceLagrangeO1(const Context &ctx, std::shared_ptr<const mesh::TriaMesh> m)
      : FeSpaceLagrange(ctx, std::move(m), 1) {}
};
struct FeSpaceLagrangeO3 : FeSpaceLagrange {
  FeSpaceLagrangeO3(const Context &ctx, std::shared_ptr<const mesh::TriaMesh> m)
      : FeSpaceLagrange(ctx, std::move(m), 3) {}
};

// Mesh functions the GPU providers accept (MeshFunctionConstant, a per-cell
// data set, sin(pi x) sin(pi y), exp of per-cell data).
class MeshFunction {
 public:
  static MeshFunction Constant(double v) { return MeshFunction(LFGPU_COEFF_CONST, v, nullptr); }
  static MeshFunction SinSin(double scale = 1.0) {
    return MeshFunction(LFGPU_COEFF_SINSIN, scale, nullptr);
  }
  static MeshFunction CellData(const DeviceVector &v, double scale = 1.0) {
    return MeshFunction(LFGPU_COEFF_CELL, scale, &v);
  }
  static MeshFunction ExpCellData(const DeviceVector &v, double scale = 1.0) {
    return MeshFunction(LFGPU_COEFF_EXPCELL, scale, &v);
  }
  lfgpu_coeff c() const { return lfgpu_coeff{kind_, value_, data_ ? data_->get() : nullptr}; }

 private:
  MeshFunction(int k, double v, const DeviceVector *d) : kind_(k), value_(v), data_(d) {}
  int kind_;
  double value_;
  const DeviceVector *data_;
};

// ReactionDiffusionElementMatrixProvider (uscalfe/loc_comp_ellbvp.h:85-339).
// Eval(cell) returns the cell's element matrix; all element matrices are
// computed on the device on first use (and by the batched assembler).
class ReactionDiffusionElementMatrixProvider {
 public:
  ReactionDiffusionElementMatrixProvider(std::shared_ptr<const FeSpaceLagrange> fes,
                                         MeshFunction alpha, MeshFunction gamma)
      : fes_(std::move(fes)), alpha_(alpha), gamma_(gamma) {}
  bool isActive(glb_idx_t) const { return true; }
  std::vector<double> Eval(glb_idx_t cell) {
    ensure();
    const int n = nloc();
    return std::vector<double>(mats_.begin() + (size_t)cell * n * n,
                               mats_.begin() + (size_t)(cell + 1) * n * n);
  }
  const std::vector<double> &all() {
    ensure();
    return mats_;
  }
  const FeSpaceLagrange &space() const { return *fes_; }
  const MeshFunction &alpha() const { return alpha_; }
  const MeshFunction &gamma() const { return gamma_; }
  int nloc() const { return fes_->Degree() == 1 ? 3 : 10; }

 private:
  void ensure() {
    if (!mats_.empty()) return;
    const int n = nloc();
    DeviceVector out(fes_->ctx(), (int64_t)fes_->Mesh()->NumCells() * n * n);
    const lfgpu_coeff a = alpha_.c(), g = gamma_.c();
    check(lfgpu_element_matrices(fes_->ctx().get(), fes_->device_mesh(), fes_->device_fe(), &a,
                                 &g, out.get()));
    mats_ = out.host();
  }
  std::shared_ptr<const FeSpaceLagrange> fes_;
  MeshFunction alpha_, gamma_;
  std::vector<double> mats_;
};

// ScalarLoadElementVectorProvider (uscalfe/loc_comp_ellbvp.h:562-746)
class ScalarLoadElementVectorProvider {
 public:
  ScalarLoadElementVectorProvider(std::shared_ptr<const FeSpaceLagrange> fes, MeshFunction f)
      : fes_(std::move(fes)), f_(f) {}
  bool isActive(glb_idx_t) const { return true; }
  std::vector<double> Eval(glb_idx_t cell) {
    ensure();
    const int n = nloc();
    return std::vector<double>(vecs_.begin() + (size_t)cell * n, vecs_.begin() + (size_t)(cell + 1) * n);
  }
  const std::vector<double> &all() {
    ensure();
    return vecs_;
  }
  const MeshFunction &f() const { return f_; }
  int nloc() const { return fes_->Degree() == 1 ? 3 : 10; }

 private:
  void ensure() {
    if (!vecs_.empty()) return;
    DeviceVector out(fes_->ctx(), (int64_t)fes_->Mesh()->NumCells() * nloc());
    const lfgpu_coeff f = f_.c();
    check(lfgpu_element_vectors(fes_->ctx().get(), fes_->device_mesh(), fes_->device_fe(), &f,
                                out.get()));
    vecs_ = out.host();
  }
  std::shared_ptr<const FeSpaceLagrange> fes_;
  MeshFunction f_;
  std::vector<double> vecs_;
};

}  // namespace uscalfe

namespace assemble {

// AssembleMatrixLocally (assemble/assembler.h:114-186) for the GPU provider:
// the element matrices come from one device launch; the triplets are added in
// the reference's order (cell, then i, then j).
template <typename TMPMATRIX>
void AssembleMatrixLocally(dim_t codim, const UniformFEDofHandler &trial,
                           const UniformFEDofHandler &test,
                           uscalfe::ReactionDiffusionElementMatrixProvider &emp,
                           TMPMATRIX &matrix) {
  if (codim != 0) throw LfException(LFGPU_EINVAL, "only cell-based assembly (codim 0)");
  const auto &mats = emp.all();
  const size_type nc = trial.Mesh()->NumCells();
  const int n = emp.nloc();
  for (glb_idx_t c = 0; c < nc; ++c) {
    if (!emp.isActive(c)) continue;
    auto rows = test.GlobalDofIndices(c);
    auto cols = trial.GlobalDofIndices(c);
    const double *m = mats.data() + (size_t)c * n * n;
    for (int i = 0; i < n; ++i)
      for (int j = 0; j < n; ++j) matrix.AddToEntry(rows[i], cols[j], m[i * n + j]);
  }
}

// AssembleVectorLocally (assemble/assembler.h:298-327)
template <typename VECTOR>
void AssembleVectorLocally(dim_t codim, const UniformFEDofHandler &dofh,
                           uscalfe::ScalarLoadElementVectorProvider &evp, VECTOR &resultvector) {
  if (codim != 0) throw LfException(LFGPU_EINVAL, "only cell-based assembly (codim 0)");
  const auto &vecs = evp.all();
  const size_type nc = dofh.Mesh()->NumCells();
  const int n = evp.nloc();
  for (glb_idx_t c = 0; c < nc; ++c) {
    auto d = dofh.GlobalDofIndices(c);
    for (int i = 0; i < n; ++i) resultvector[d[i]] += vecs[(size_t)c * n + i];
  }
}

// Device-resident fast path: AssembleMatrixLocally + AssembleVectorLocally +
// FixFlaggedSolutionComponents + makeSparse in one fused launch into a
// pattern built once (lfgpu_plan_* / lfgpu_assemble_galerkin).
class GalerkinAssembler {
 public:
  GalerkinAssembler(std::shared_ptr<const uscalfe::FeSpaceLagrange> fes,
                    const std::vector<uint8_t> &fixed_flags)
      : fes_(std::move(fes)) {
    check(lfgpu_plan_create(fes_->ctx().get(), fes_->device_mesh(),
                            fixed_flags.empty() ? nullptr : fixed_flags.data(), &plan_));
    lfgpu_csr *c = nullptr;
    check(lfgpu_plan_csr(plan_, &c));
    csr_ = std::make_unique<CsrMatrix>(c, false);
  }
  ~GalerkinAssembler() { lfgpu_plan_destroy(plan_); }
  GalerkinAssembler(const GalerkinAssembler &) = delete;
  GalerkinAssembler &operator=(const GalerkinAssembler &) = delete;
  // returns the device time in ms
  double Assemble(const uscalfe::MeshFunction &alpha, const uscalfe::MeshFunction &gamma,
                  const uscalfe::MeshFunction *f, DeviceVector *rhs,
                  const DeviceVector *fixed_vals = nullptr) {
    const lfgpu_coeff a = alpha.c(), g = gamma.c();
    lfgpu_coeff fc{};
    if (f) fc = f->c();
    check(lfgpu_assemble_galerkin(fes_->ctx().get(), plan_, fes_->device_fe(), &a, &g,
                                  f ? &fc : nullptr, fixed_vals ? fixed_vals->get() : nullptr,
                                  rhs ? rhs->get() : nullptr));
    return fes_->ctx().last_ms();
  }
  const CsrMatrix &matrix() const { return *csr_; }

 private:
  std::shared_ptr<const uscalfe::FeSpaceLagrange> fes_;
  lfgpu_plan *plan_ = nullptr;
  std::unique_ptr<CsrMatrix> csr_;
};

}  // namespace assemble

// ------------------------------------------------------------------ solve
namespace solve {

// Jacobi-PCG with Eigen ConjugateGradient semantics (the SparseLU of the
// demo drivers is replaced by this iterative solve on the device).
inline lfgpu_pcg_info PCG(const Context &ctx, const assemble::CsrMatrix &A, const DeviceVector &b,
                          DeviceVector &x, double rtol = 1e-10, int maxit = 0,
                          int mode = LFGPU_PCG_FP64) {
  lfgpu_pcg_opts o{rtol, maxit > 0 ? maxit : 2 * A.rows(), mode, 64, 0.0, 0, 0, 0};
  lfgpu_pcg_info info{};
  check(lfgpu_pcg(ctx.get(), A.get(), b.get(), x.get(), &o, &info));
  return info;
}

}  // namespace solve
}  // namespace lf::amd
