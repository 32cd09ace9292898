// Internal declarations of liblfgpu (not part of the C ABI).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/lfgpu.h"

namespace lfgpu {

void set_error(const std::string &msg);
void clear_error();

// Returns a non-OK status with the HIP error text recorded.
int hip_fail(hipError_t e, const char *what);

#define LFG_HIP(expr)                                     \
  do {                                                    \
    hipError_t lfg_e_ = (expr);                           \
    if (lfg_e_ != hipSuccess) return lfgpu::hip_fail(lfg_e_, #expr); \
  } while (0)

#define LFG_REQUIRE(cond, code, msg) \
  do {                               \
    if (!(cond)) {                   \
      lfgpu::set_error(msg);         \
      return (code);                 \
    }                                \
  } while (0)

#define LFG_TRY(expr)               \
  do {                              \
    int lfg_rc_ = (expr);           \
    if (lfg_rc_ != LFGPU_OK) return lfg_rc_; \
  } while (0)

// Device allocation that records an LFGPU_EOOM on failure.
int dmalloc(void **p, size_t bytes, const char *what);
template <class T>
int dalloc(T **p, size_t count, const char *what) {
  return dmalloc(reinterpret_cast<void **>(p), count * sizeof(T), what);
}
void dfree(void *p);

// Coefficient as seen by kernels (lfgpu_coeff resolved to a device pointer).
struct CoeffD {
  int kind;
  double value;
  const double *data;
};
int resolve_coeff(const lfgpu_coeff *c, int64_t ncells, int nqp, bool allow_null,
                  CoeffD *out, const char *name);

// Device-wide exclusive scan of int32 counts into int32 offsets (n+1 outputs).
int exclusive_scan_i32(hipStream_t s, const int32_t *in, int32_t *out, int64_t n,
                       int64_t *total);

}  // namespace lfgpu

struct lfgpu_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  hipEvent_t evp[2] = {nullptr, nullptr};  // solver polls (no timing)
  double last_ms = 0.0;
  // pinned host mirror of small solver status words
  void *pinned = nullptr;
  // reusable device scratch for reductions (partials + scalars)
  double *red = nullptr;
  size_t red_count = 0;
};

struct lfgpu_vec {
  lfgpu_ctx *ctx = nullptr;
  double *d = nullptr;
  int64_t n = 0;
};

struct lfgpu_mesh {
  lfgpu_ctx *ctx = nullptr;
  int64_t nv = 0, ncells = 0, ndofs = 0;
  int nloc = 0;
  double *xy = nullptr;         // [nv x 2]
  int32_t *cell_v = nullptr;    // [ncells x 3]
  int32_t *cell_dofs = nullptr; // [ncells x nloc]; aliases cell_v for P1
  bool dofs_alias = false;
};

struct lfgpu_fe {
  lfgpu_ctx *ctx = nullptr;
  int degree = 0, nloc = 0, nqp = 0;
  // device table: w[nqp] | pts[2 nqp] | N[nloc nqp] | dN[nloc 2 nqp]
  double *tab = nullptr;
  int tab_len = 0;
  // reference tensors of the element matrix, [4 + 4 nqp][nep] row-major:
  // rows 0..3 are the quadrature sums of the four terms (dx dx, dy dy,
  // dx dy + dy dx, N N), rows 4 + 4k + t the weighted term t at point k;
  // nep = nloc^2 rounded up to a multiple of 16
  double *rten = nullptr;
  int nep = 0;
};

struct lfgpu_solver_cache;  // lfgpu_solver.hip

struct lfgpu_csr {
  lfgpu_ctx *ctx = nullptr;
  int32_t nrows = 0, ncols = 0;
  int64_t nnz = 0;
  int32_t *row_ptr = nullptr;
  int32_t *col = nullptr;
  double *val = nullptr;
  bool owned_by_plan = false;
  lfgpu_solver_cache *cache = nullptr;  // built lazily by the solver
  uint64_t val_version = 0;             // bumped whenever val changes
  // Optional dense block (lfgpu_csr_set_dense_block): the matrix stands for
  // its sparse part + blk_scale * B, B [blk_n x blk_n] row-major on the rows
  // and columns blk_rows (ascending), both device arrays owned here.
  int32_t blk_n = 0;
  int32_t *blk_rows = nullptr;
  double *blk = nullptr;
  double blk_scale = 0.0;
};

struct lfgpu_plan {
  lfgpu_ctx *ctx = nullptr;
  const lfgpu_mesh *mesh = nullptr;
  lfgpu_csr csr;
  int32_t *rec_ptr = nullptr;  // [ndofs + 1]
  int32_t *rec_key = nullptr;  // [R] cell*nloc + i, cell order within a row
  uint8_t *rec_meta = nullptr; // [R x meta_stride]: i, slot of each local col
  int meta_stride = 0;
  int64_t nrec = 0;
  uint8_t *fixed = nullptr;    // [ndofs]
  bool has_fixed = false;
  int block = 256;             // rows per workgroup of the assembly kernel
  int max_block_nnz = 0;       // LDS doubles needed by one workgroup
  double *ke = nullptr;        // [ncells x nloc^2] element matrices (contraction path)
  // coefficient values evaluated once per cell before the row-owner
  // assembly (exp per cell, sin sin per point): three slots of ncells x nqp
  double *coef = nullptr;
  int64_t coef_len = 0;
  double *fe = nullptr;        // [ncells x nloc] element vectors (contraction path)
  // boundary edge terms (lfgpu_plan_set_edges)
  const lfgpu_edges *edges = nullptr;
  int32_t *erec_ptr = nullptr;  // [ndofs + 1] edge-matrix records per row
  int32_t *erec_key = nullptr;  // [Rm] edge * nle + i, in edge order within a row
  uint8_t *erec_meta = nullptr; // [Rm x 4] CSR slot of each local column (0xFF: fixed)
  int32_t *vrec_ptr = nullptr;  // [ndofs + 1] edge-load records per row
  int32_t *vrec_key = nullptr;  // [Rv] edge * nle + i
  bool edge_mat = false, edge_vec = false;
};

// Codim-1 entities for the boundary terms: edge e runs from edge_v[2e] to
// edge_v[2e+1] (its orientation); its DOFs are the two endpoint DOFs then
// its interior DOFs (P3).
struct lfgpu_edges {
  lfgpu_ctx *ctx = nullptr;
  const lfgpu_mesh *mesh = nullptr;
  int degree = 0, nle = 0, nqe = 0;
  int64_t nedges = 0;
  int32_t *edge_v = nullptr;     // [E x 2]
  int32_t *edge_dofs = nullptr;  // [E x nle]
  double *tab = nullptr;         // w[nqe] | pts[nqe] | N[nle x nqe], N(i,k) at i + nle k
  std::vector<int32_t> h_dofs;   // host copy of edge_dofs
};

namespace lfgpu {
void solver_cache_free(lfgpu_csr *A);
// Element matrices (and, if has_f, element vectors) of every cell through the
// reference-tensor contraction on the matrix cores (lfgpu_mfma.hip): ke is
// [cell][i][j], or with rowmajor_i [i][cell][j].
int element_contract(hipStream_t s, const lfgpu_mesh *m, const lfgpu_fe *fe, const CoeffD &alpha,
                     const CoeffD &gamma, const CoeffD *f, double *ke, double *fvec,
                     bool rowmajor_i);
// Release the edge records of a plan (lfgpu_edges.hip).
void plan_free_edges(lfgpu_plan *p);
inline void csr_values_changed(lfgpu_csr *A) { A->val_version++; }
// Dense block of a CSR (lfgpu_nonlocal.hip): copy of from's block scaled by
// `scale` into to (which must have none); release.
int csr_copy_block(lfgpu_ctx *ctx, const lfgpu_csr *from, double scale, lfgpu_csr *to);
void csr_free_block(lfgpu_csr *A);
// y[rows[b]] += scale (B x[rows])_b for the block of A, queued on stream s
// (CSR order x and y).
void csr_block_gemv_add(hipStream_t s, const lfgpu_csr *A, const double *x, double *y);
// y = A x (FP64) on the context stream, without synchronisation; builds the
// solver's storage of A on first use (lfgpu_solver.hip).
int spmv_async(lfgpu_ctx *ctx, lfgpu_csr *A, const double *x, double *y);
// Queues the fused row-owner assembly of lfgpu_assemble_galerkin on the
// context stream: no events, no synchronisation, no host bookkeeping of the
// CSR's value version (so it can be captured into a graph;
// lfgpu_assemble.hip).
int assemble_enqueue(lfgpu_ctx *ctx, lfgpu_plan *p, const lfgpu_fe *fe,
                     const lfgpu_coeff *alpha, const lfgpu_coeff *gamma, const lfgpu_coeff *f,
                     lfgpu_vec *rhs);
// Looks up this device's address of the allocation-free solver's work space
// (lfgpu_solver.hip); called by lfgpu_ctx_create with the device current.
int noalloc_scratch_init(int device);
}  // namespace lfgpu
