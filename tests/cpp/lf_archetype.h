// lf_archetype.h -- an archetype of the LehrFEM++ interfaces that
// lehrfempp_amd/include/lf_amd/lehrfem_binding.h binds to (test
// infrastructure only).
//
// LehrFEM++ itself cannot be compiled here (Eigen, Boost and spdlog come
// from a network fetch), so the binding is exercised against classes with
// the reference's names, signatures and semantics for exactly the members it
// uses, in the manner of the reference's own archetypes
// (assemble/assemble_concepts.h:95-108):
//   lf::base::RefEl                       base/ref_el.h:31-38
//   lf::mesh::Entity / Mesh               mesh/entity.h:42-150, mesh_interface.h:73-209
//   lf::geometry::Geometry, Corners       geometry/geometry_interface.h:21-227, :261
//   lf::quad::QuadRule, make_QuadRule     quad/quad_rule.h:58-199, make_quad_rule.h:37
//   lf::assemble::DofHandler              assemble/dofhandler.h:112-228
//   lf::assemble::EntityMatrixProvider / EntityVectorProvider concepts
//                                         assemble/assemble_concepts.h:95-99, :210-214
//   lf::assemble::AssembleMatrixLocally / AssembleVectorLocally, the generic
//                                         cell loops of assemble/assembler.h:114-186, :298-327
//   lf::assemble::COOMatrix               assemble/coomatrix.h:51-225
//   lf::uscalfe::FeSpaceLagrangeO1/O3, ReactionDiffusionElementMatrixProvider,
//   ScalarLoadElementVectorProvider       uscalfe/*.h (the CPU providers evaluate
//                                         one cell through the oracle, oracle/lfref.h)
// Dense matrices stand in for Eigen's (a column-major dynamic matrix with
// operator()(i, j), operator[] for vectors, rows() / cols() / size()).
#pragma once

#include <algorithm>
#include <array>
#include <concepts>
#include <cstdint>
#include <memory>
#include <span>
#include <stdexcept>
#include <utility>
#include <vector>

#include "oracle/lfref.h"

namespace lf::base {
using size_type = unsigned int;
using glb_idx_t = unsigned int;
using dim_t = unsigned int;

class RefEl {
 public:
  static constexpr RefEl kPoint() { return RefEl(0); }
  static constexpr RefEl kSegment() { return RefEl(1); }
  static constexpr RefEl kTria() { return RefEl(2); }
  static constexpr RefEl kQuad() { return RefEl(3); }
  constexpr unsigned Id() const { return id_; }
  constexpr bool operator==(const RefEl &o) const { return id_ == o.id_; }
  constexpr bool operator!=(const RefEl &o) const { return id_ != o.id_; }

 private:
  constexpr explicit RefEl(unsigned id) : id_(id) {}
  unsigned id_;
};
}  // namespace lf::base

namespace lf::archetype {
// column-major dynamic matrix (the role of Eigen::MatrixXd / VectorXd)
class Dense {
 public:
  Dense() = default;
  Dense(int r, int c) : r_(r), c_(c), a_((size_t)r * c, 0.0) {}
  explicit Dense(int n) : Dense(n, 1) {}
  double &operator()(int i, int j) { return a_[(size_t)j * r_ + i]; }
  double operator()(int i, int j) const { return a_[(size_t)j * r_ + i]; }
  double &operator[](int i) { return a_[i]; }
  double operator[](int i) const { return a_[i]; }
  int rows() const { return r_; }
  int cols() const { return c_; }
  int size() const { return r_ * c_; }
  const double *data() const { return a_.data(); }

 private:
  int r_ = 0, c_ = 0;
  std::vector<double> a_;
};
}  // namespace lf::archetype

namespace lf::geometry {
// TriaO1: the three corners of a triangle (geometry/tria_o1.h:23-66)
class Geometry {
 public:
  explicit Geometry(std::array<double, 6> corners) : c_(corners) {}
  // x = v0 (1 - x1 - x2) + v1 x1 + v2 x2 for the local points (tria_o1.cc:70-74)
  lf::archetype::Dense Global(const lf::archetype::Dense &local) const {
    lf::archetype::Dense g(2, local.cols());
    for (int k = 0; k < local.cols(); ++k) {
      const double s = local(0, k), t = local(1, k), w = 1.0 - s - t;
      for (int d = 0; d < 2; ++d) g(d, k) = c_[d] * w + c_[2 + d] * s + c_[4 + d] * t;
    }
    return g;
  }
  const std::array<double, 6> &corners() const { return c_; }

 private:
  std::array<double, 6> c_;  // (x0, y0, x1, y1, x2, y2)
};
inline lf::archetype::Dense Corners(const Geometry &geo) {
  lf::archetype::Dense X(2, 3);
  for (int k = 0; k < 3; ++k)
    for (int d = 0; d < 2; ++d) X(d, k) = geo.corners()[2 * k + d];
  return X;
}
}  // namespace lf::geometry

namespace lf::mesh {
class Entity {
 public:
  virtual ~Entity() = default;
  virtual unsigned Codim() const = 0;
  virtual std::span<const Entity *const> SubEntities(unsigned rel_codim) const = 0;
  virtual const geometry::Geometry *Geometry() const = 0;
  virtual base::RefEl RefEl() const = 0;
};

class Mesh {
 public:
  virtual ~Mesh() = default;
  virtual std::span<const Entity *const> Entities(unsigned codim) const = 0;
  virtual base::size_type NumEntities(unsigned codim) const = 0;
  virtual base::size_type Index(const Entity &e) const = 0;
};
}  // namespace lf::mesh

namespace lf::quad {
class QuadRule {
 public:
  QuadRule(lf::archetype::Dense pts, lf::archetype::Dense w)
      : pts_(std::move(pts)), w_(std::move(w)) {}
  const lf::archetype::Dense &Points() const { return pts_; }
  const lf::archetype::Dense &Weights() const { return w_; }
  base::size_type NumPoints() const { return (base::size_type)pts_.cols(); }

 private:
  lf::archetype::Dense pts_, w_;
};
// the triangle rules of quad/make_quad_rule.cc:38-50 (through the oracle)
inline QuadRule make_QuadRule(base::RefEl ref_el, unsigned degree) {
  if (ref_el != base::RefEl::kTria()) throw std::invalid_argument("archetype: triangles only");
  double p[24], w[12];
  const int n = lfref_quad_rule((int)degree, p, w);
  if (n < 0) throw std::invalid_argument("archetype: no rule of that degree");
  lf::archetype::Dense P(2, n), W(n);
  for (int k = 0; k < n; ++k) {
    P(0, k) = p[2 * k];
    P(1, k) = p[2 * k + 1];
    W[k] = w[k];
  }
  return QuadRule(P, W);
}
}  // namespace lf::quad

namespace lf::assemble {
using gdof_idx_t = int64_t;
using size_type = base::size_type;
using glb_idx_t = base::glb_idx_t;
using dim_t = base::dim_t;

class DofHandler {
 public:
  virtual ~DofHandler() = default;
  virtual size_type NumDofs() const = 0;
  virtual size_type NumLocalDofs(const mesh::Entity &entity) const = 0;
  virtual std::span<const gdof_idx_t> GlobalDofIndices(const mesh::Entity &entity) const = 0;
  virtual std::shared_ptr<const mesh::Mesh> Mesh() const = 0;
};

template <class EMP>
concept EntityMatrixProvider = requires(EMP &emp, const mesh::Entity &e) {
  { emp.isActive(e) } -> std::same_as<bool>;
  { emp.Eval(e) } -> std::same_as<lf::archetype::Dense>;
};
template <class EVP>
concept EntityVectorProvider = requires(EVP &evp, const mesh::Entity &e) {
  { evp.isActive(e) } -> std::same_as<bool>;
  { evp.Eval(e) } -> std::same_as<lf::archetype::Dense>;
};

// the reference's cell loop (assembler.h:114-186)
template <typename TMPMATRIX, EntityMatrixProvider ENTITY_MATRIX_PROVIDER>
void AssembleMatrixLocally(dim_t codim, const DofHandler &dof_handler_trial,
                           const DofHandler &dof_handler_test,
                           ENTITY_MATRIX_PROVIDER &entity_matrix_provider, TMPMATRIX &matrix) {
  auto mesh = dof_handler_trial.Mesh();
  for (const lf::mesh::Entity *entity : mesh->Entities(codim)) {
    if (entity_matrix_provider.isActive(*entity)) {
      const size_type nrows_loc = dof_handler_test.NumLocalDofs(*entity);
      const size_type ncols_loc = dof_handler_trial.NumLocalDofs(*entity);
      std::span<const gdof_idx_t> row_idx(dof_handler_test.GlobalDofIndices(*entity));
      std::span<const gdof_idx_t> col_idx(dof_handler_trial.GlobalDofIndices(*entity));
      const auto elem_mat{entity_matrix_provider.Eval(*entity)};
      for (int i = 0; i < (int)nrows_loc; i++)
        for (int j = 0; j < (int)ncols_loc; j++)
          matrix.AddToEntry(row_idx[i], col_idx[j], elem_mat(i, j));
    }
  }
}
// the returning overload (assembler.h:209-218)
template <typename TMPMATRIX, EntityMatrixProvider ENTITY_MATRIX_PROVIDER>
TMPMATRIX AssembleMatrixLocally(dim_t codim, const DofHandler &dof_handler_trial,
                                const DofHandler &dof_handler_test,
                                ENTITY_MATRIX_PROVIDER &entity_matrix_provider) {
  TMPMATRIX matrix{dof_handler_test.NumDofs(), dof_handler_trial.NumDofs()};
  matrix.setZero();
  AssembleMatrixLocally<TMPMATRIX, ENTITY_MATRIX_PROVIDER>(codim, dof_handler_trial,
                                                           dof_handler_test,
                                                           entity_matrix_provider, matrix);
  return matrix;
}

// assembler.h:298-327
template <typename VECTOR, EntityVectorProvider ENTITY_VECTOR_PROVIDER>
void AssembleVectorLocally(dim_t codim, const DofHandler &dof_handler,
                           ENTITY_VECTOR_PROVIDER &entity_vector_provider, VECTOR &resultvector) {
  auto mesh = dof_handler.Mesh();
  for (const lf::mesh::Entity *entity : mesh->Entities(codim)) {
    if (entity_vector_provider.isActive(*entity)) {
      const size_type veclen = dof_handler.NumLocalDofs(*entity);
      const std::span<const gdof_idx_t> dof_idx(dof_handler.GlobalDofIndices(*entity));
      const auto elem_vec{entity_vector_provider.Eval(*entity)};
      for (int i = 0; i < (int)veclen; i++) resultvector[dof_idx[i]] += elem_vec[i];
    }
  }
}

// coomatrix.h:51-225 (triplets in insertion order)
template <typename SCALAR>
class COOMatrix {
 public:
  struct Triplet {
    int32_t r, c;
    SCALAR v;
  };
  COOMatrix(size_type num_rows, size_type num_cols) : rows_(num_rows), cols_(num_cols) {}
  void AddToEntry(gdof_idx_t i, gdof_idx_t j, SCALAR increment) {
    trip_.push_back(Triplet{(int32_t)i, (int32_t)j, increment});
  }
  void setZero() { trip_.clear(); }
  const std::vector<Triplet> &triplets() const { return trip_; }
  size_type rows() const { return rows_; }
  size_type cols() const { return cols_; }

 private:
  size_type rows_, cols_;
  std::vector<Triplet> trip_;
};
}  // namespace lf::assemble

namespace lf::archetype {

// A triangle mesh: vertices (codim 2) and cells (codim 0) as entity objects,
// Entities() in index order, numbering of the reference builder (through the
// oracle's TPTriagMeshBuilder restatement) with optionally moved vertices.
class TriaMesh : public lf::mesh::Mesh {
 public:
  class Vertex : public lf::mesh::Entity {
   public:
    unsigned Codim() const override { return 2; }
    std::span<const Entity *const> SubEntities(unsigned) const override { return {}; }
    const lf::geometry::Geometry *Geometry() const override { return nullptr; }
    lf::base::RefEl RefEl() const override { return lf::base::RefEl::kPoint(); }
  };
  class Cell : public lf::mesh::Entity {
   public:
    Cell(std::array<const Entity *, 3> v, lf::geometry::Geometry g) : v_(v), g_(std::move(g)) {}
    unsigned Codim() const override { return 0; }
    std::span<const Entity *const> SubEntities(unsigned rel_codim) const override {
      if (rel_codim != 2) throw std::invalid_argument("archetype cells carry vertices only");
      return {v_.data(), 3};
    }
    const lf::geometry::Geometry *Geometry() const override { return &g_; }
    lf::base::RefEl RefEl() const override { return lf::base::RefEl::kTria(); }

   private:
    std::array<const Entity *, 3> v_;
    lf::geometry::Geometry g_;
  };

  TriaMesh(std::vector<double> xy, std::vector<int32_t> cell_v, std::vector<int32_t> pre_edges)
      : xy_(std::move(xy)), cell_v_(std::move(cell_v)), pre_(std::move(pre_edges)) {
    const size_t nv = xy_.size() / 2, nc = cell_v_.size() / 3;
    verts_.resize(nv);
    for (size_t v = 0; v < nv; ++v) vptr_.push_back(&verts_[v]);
    cells_.reserve(nc);
    for (size_t c = 0; c < nc; ++c) {
      std::array<const lf::mesh::Entity *, 3> vv;
      std::array<double, 6> g;
      for (int k = 0; k < 3; ++k) {
        const int32_t v = cell_v_[3 * c + k];
        vv[k] = &verts_[v];
        g[2 * k] = xy_[2 * v];
        g[2 * k + 1] = xy_[2 * v + 1];
      }
      cells_.emplace_back(vv, lf::geometry::Geometry(g));
    }
    for (auto &c : cells_) cptr_.push_back(&c);
  }
  std::span<const lf::mesh::Entity *const> Entities(unsigned codim) const override {
    if (codim == 0) return {cptr_.data(), cptr_.size()};
    if (codim == 2) return {vptr_.data(), vptr_.size()};
    throw std::invalid_argument("archetype mesh: no edge entities");
  }
  lf::base::size_type NumEntities(unsigned codim) const override {
    return (lf::base::size_type)Entities(codim).size();
  }
  lf::base::size_type Index(const lf::mesh::Entity &e) const override {
    if (e.Codim() == 0) return (lf::base::size_type)(static_cast<const Cell *>(&e) - cells_.data());
    return (lf::base::size_type)(static_cast<const Vertex *>(&e) - verts_.data());
  }
  const std::vector<double> &xy() const { return xy_; }
  const std::vector<int32_t> &cell_v() const { return cell_v_; }
  const std::vector<int32_t> &pre_edges() const { return pre_; }

 private:
  std::vector<double> xy_;
  std::vector<int32_t> cell_v_, pre_;
  std::vector<Vertex> verts_;
  std::vector<Cell> cells_;
  std::vector<const lf::mesh::Entity *> vptr_, cptr_;
};

// n x n structured mesh of the unit square (TPTriagMeshBuilder), interior
// vertices moved by (dx, dy) = jitter(v)
template <class JITTER>
std::shared_ptr<const TriaMesh> tp_mesh(int n, JITTER &&jitter) {
  const int nv = (n + 1) * (n + 1), nc = 2 * n * n, ne = n * n + 2 * (n + 1) * n;
  std::vector<double> xy(2 * nv);
  std::vector<int32_t> cv(3 * nc), ev(2 * ne);
  lfref_tp_triag_mesh(n, n, 0.0, 0.0, 1.0, 1.0, xy.data(), cv.data(), ev.data());
  for (int v = 0; v < nv; ++v) {
    const int i = v % (n + 1), j = v / (n + 1);
    if (i > 0 && i < n && j > 0 && j < n) {
      const auto d = jitter(v);
      xy[2 * v] += d.first;
      xy[2 * v + 1] += d.second;
    }
  }
  return std::make_shared<const TriaMesh>(std::move(xy), std::move(cv), std::move(ev));
}

// UniformFEDofHandler of a Lagrangian space (dofhandler.cc:86-305) through
// the oracle's restatement of the numbering
class UniformDofHandler : public lf::assemble::DofHandler {
 public:
  UniformDofHandler(std::shared_ptr<const TriaMesh> m, int degree)
      : mesh_(std::move(m)), nloc_(degree == 1 ? 3 : 10) {
    const int nv = (int)(mesh_->xy().size() / 2), nc = (int)(mesh_->cell_v().size() / 3);
    const int npre = (int)(mesh_->pre_edges().size() / 2);
    std::vector<int32_t> ev(2 * (npre + 3 * nc)), ce(3 * nc);
    std::vector<int8_t> co(3 * nc);
    const int ne = lfref_mesh_edges(nv, nc, mesh_->cell_v().data(), npre,
                                    mesh_->pre_edges().data(), npre + 3 * nc, ev.data(),
                                    ce.data(), co.data());
    dofs_.resize((size_t)nc * nloc_);
    ndofs_ = lfref_dof_table(nv, nc, ne, mesh_->cell_v().data(), ce.data(), co.data(),
                             degree == 1 ? 0 : 2, degree == 1 ? 0 : 1, dofs_.data());
    edge_v_.assign(ev.begin(), ev.begin() + 2 * ne);
    cell_e_ = std::move(ce);
    nedges_ = ne;
  }
  lf::assemble::size_type NumDofs() const override { return (lf::assemble::size_type)ndofs_; }
  lf::assemble::size_type NumLocalDofs(const lf::mesh::Entity &) const override {
    return (lf::assemble::size_type)nloc_;
  }
  std::span<const lf::assemble::gdof_idx_t> GlobalDofIndices(
      const lf::mesh::Entity &cell) const override {
    return {dofs_.data() + (size_t)mesh_->Index(cell) * nloc_, (size_t)nloc_};
  }
  std::shared_ptr<const lf::mesh::Mesh> Mesh() const override { return mesh_; }
  const TriaMesh &tria() const { return *mesh_; }
  int nedges() const { return nedges_; }
  const std::vector<int32_t> &edge_v() const { return edge_v_; }
  const std::vector<int32_t> &cell_e() const { return cell_e_; }

 private:
  std::shared_ptr<const TriaMesh> mesh_;
  int nloc_;
  int64_t ndofs_ = 0;
  int nedges_ = 0;
  std::vector<lf::assemble::gdof_idx_t> dofs_;
  std::vector<int32_t> edge_v_, cell_e_;
};
}  // namespace lf::archetype

namespace lf::uscalfe {
// the triangle layout of a Lagrangian FE space (Degree())
struct TriaLayout {
  unsigned degree;
  unsigned Degree() const { return degree; }
};

// UniformScalarFESpace<SCALAR> of degree 1 / 3 (uniform_scalar_fe_space.h:49-181)
template <typename SCALAR>
class UniformScalarFESpace {
 public:
  using Scalar = SCALAR;
  UniformScalarFESpace(std::shared_ptr<const lf::archetype::TriaMesh> m, unsigned degree)
      : mesh_(m), layout_{degree}, dofh_(m, (int)degree) {}
  std::shared_ptr<const lf::mesh::Mesh> Mesh() const { return mesh_; }
  const lf::assemble::DofHandler &LocGlobMap() const { return dofh_; }
  const TriaLayout *ShapeFunctionLayout(lf::base::RefEl ref_el) const {
    return ref_el == lf::base::RefEl::kTria() ? &layout_ : nullptr;
  }

 private:
  std::shared_ptr<const lf::archetype::TriaMesh> mesh_;
  TriaLayout layout_;
  lf::archetype::UniformDofHandler dofh_;
};
template <typename SCALAR>
struct FeSpaceLagrangeO1 : UniformScalarFESpace<SCALAR> {
  explicit FeSpaceLagrangeO1(std::shared_ptr<const lf::archetype::TriaMesh> m)
      : UniformScalarFESpace<SCALAR>(std::move(m), 1) {}
};
template <typename SCALAR>
struct FeSpaceLagrangeO3 : UniformScalarFESpace<SCALAR> {
  explicit FeSpaceLagrangeO3(std::shared_ptr<const lf::archetype::TriaMesh> m)
      : UniformScalarFESpace<SCALAR>(std::move(m), 3) {}
};

namespace detail {
// a MeshFunction at the rule's points of one cell, as the reference's
// providers call them (loc_comp_ellbvp.h:316-317)
template <class MF>
std::vector<double> at_points(const MF &mf, const lf::mesh::Entity &cell,
                              const lf::quad::QuadRule &qr) {
  auto v = mf(cell, qr.Points());
  return std::vector<double>(v.begin(), v.end());
}
inline std::vector<double> cell_xy(const lf::mesh::Entity &cell) {
  const auto &c = cell.Geometry()->corners();
  return std::vector<double>(c.begin(), c.end());
}
}  // namespace detail

// the reference provider on the CPU: one cell through the oracle's
// restatement of Eval (loc_comp_ellbvp.h:266-339)
template <typename SCALAR, class DIFF_COEFF, class REACTION_COEFF>
class ReactionDiffusionElementMatrixProvider {
 public:
  ReactionDiffusionElementMatrixProvider(std::shared_ptr<const UniformScalarFESpace<SCALAR>> fes,
                                         DIFF_COEFF alpha, REACTION_COEFF gamma)
      : fes_(std::move(fes)), alpha_(std::move(alpha)), gamma_(std::move(gamma)),
        degree_((int)fes_->ShapeFunctionLayout(lf::base::RefEl::kTria())->Degree()),
        qr_(lf::quad::make_QuadRule(lf::base::RefEl::kTria(), 2 * degree_)) {}
  bool isActive(const lf::mesh::Entity &) { return true; }
  lf::archetype::Dense Eval(const lf::mesh::Entity &cell) {
    const int n = degree_ == 1 ? 3 : 10;
    const auto xy = detail::cell_xy(cell);
    const int32_t cv[3] = {0, 1, 2};
    const auto a = detail::at_points(alpha_, cell, qr_), g = detail::at_points(gamma_, cell, qr_);
    std::vector<double> m((size_t)n * n);
    if (lfref_element_matrices(degree_, 1, xy.data(), cv, a.data(), g.data(), m.data()) != 0)
      throw std::runtime_error("degenerate cell");
    lf::archetype::Dense M(n, n);
    for (int i = 0; i < n; ++i)
      for (int j = 0; j < n; ++j) M(i, j) = m[(size_t)i * n + j];
    return M;
  }

 private:
  std::shared_ptr<const UniformScalarFESpace<SCALAR>> fes_;
  DIFF_COEFF alpha_;
  REACTION_COEFF gamma_;
  int degree_;
  lf::quad::QuadRule qr_;
};
template <class PTR, class DIFF_COEFF, class REACTION_COEFF>
ReactionDiffusionElementMatrixProvider(PTR fe_space, DIFF_COEFF alpha, REACTION_COEFF gamma)
    -> ReactionDiffusionElementMatrixProvider<typename PTR::element_type::Scalar, DIFF_COEFF,
                                              REACTION_COEFF>;

// loc_comp_ellbvp.h:691-746 on the CPU, one cell through the oracle
template <typename SCALAR, class FUNCTOR>
class ScalarLoadElementVectorProvider {
 public:
  ScalarLoadElementVectorProvider(std::shared_ptr<const UniformScalarFESpace<SCALAR>> fes,
                                  FUNCTOR f)
      : fes_(std::move(fes)), f_(std::move(f)),
        degree_((int)fes_->ShapeFunctionLayout(lf::base::RefEl::kTria())->Degree()),
        qr_(lf::quad::make_QuadRule(lf::base::RefEl::kTria(), 2 * degree_)) {}
  bool isActive(const lf::mesh::Entity &) { return true; }
  lf::archetype::Dense Eval(const lf::mesh::Entity &cell) {
    const int n = degree_ == 1 ? 3 : 10;
    const auto xy = detail::cell_xy(cell);
    const int32_t cv[3] = {0, 1, 2};
    const auto f = detail::at_points(f_, cell, qr_);
    std::vector<double> v(n);
    if (lfref_element_vectors(degree_, 1, xy.data(), cv, 2, 0.0, f.data(), v.data()) != 0)
      throw std::runtime_error("degenerate cell");
    lf::archetype::Dense V(n);
    for (int i = 0; i < n; ++i) V[i] = v[i];
    return V;
  }

 private:
  std::shared_ptr<const UniformScalarFESpace<SCALAR>> fes_;
  FUNCTOR f_;
  int degree_;
  lf::quad::QuadRule qr_;
};
template <class PTR, class FUNCTOR>
ScalarLoadElementVectorProvider(PTR fe_space, FUNCTOR f)
    -> ScalarLoadElementVectorProvider<typename PTR::element_type::Scalar, FUNCTOR>;
}  // namespace lf::uscalfe
