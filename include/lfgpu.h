/*
 * lfgpu.h -- C ABI of the MI355X-native LehrFEM++ assembly + PCG path.
 *
 * The library (lehrfempp_amd/liblfgpu.so) replaces the data-parallel part of
 * LehrFEM++'s Galerkin pipeline
 *
 *   ReactionDiffusionElementMatrixProvider / ScalarLoadElementVectorProvider
 *     -> AssembleMatrixLocally / AssembleVectorLocally into a COOMatrix
 *     -> FixFlaggedSolutionComponents -> COOMatrix::makeSparse
 *     -> sparse solve
 *
 * with device-resident HIP kernels.  All entry points return an int status
 * (LFGPU_OK or one of the error codes below) and never throw; the text of the
 * last error of the calling thread is available from lfgpu_last_error().
 * Host buffers passed in are owned by the caller and only read/written during
 * the call; device buffers are owned by the library's handles.
 * All work of one context is ordered on that context's HIP stream; a context
 * must be used by one host thread at a time.
 *
 * Index conventions: DOF indices are gdof_idx_t = int64 at the boundary
 * (reference assemble/assembly_types.h:22) and int32 on the device
 * (every supported problem has fewer than 2^31 DOFs and nonzeros).
 */
#ifndef LFGPU_H
#define LFGPU_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------ */
#define LFGPU_OK 0
#define LFGPU_EINVAL 1      /* bad argument / size mismatch (LF_ASSERT_MSG)   */
#define LFGPU_EDEGENERATE 2 /* degenerate triangle (tria_o1.cc:10-48)         */
#define LFGPU_ENOCONV 3     /* iterative solver hit maxit before rtol        */
#define LFGPU_EOOM 4        /* device allocation failed                      */
#define LFGPU_EHIP 5        /* HIP runtime error                             */
#define LFGPU_ENORULE 6     /* no quadrature rule / FE (LfException,
                               uscalfe/loc_comp_ellbvp.h:278-287)             */
#define LFGPU_ERANGE 7      /* index out of range or size limit exceeded     */

/* Text of the last error on this thread ("" if none). */
const char *lfgpu_last_error(void);

/* ---- opaque handles ---------------------------------------------------- */
typedef struct lfgpu_ctx lfgpu_ctx;   /* device + stream + event timer      */
typedef struct lfgpu_vec lfgpu_vec;   /* device FP64 vector                 */
typedef struct lfgpu_mesh lfgpu_mesh; /* vertex coords + cell->vertex/DOF   */
typedef struct lfgpu_fe lfgpu_fe;     /* quadrature rule + shape tables     */
typedef struct lfgpu_plan lfgpu_plan; /* CSR pattern + scatter records      */
typedef struct lfgpu_csr lfgpu_csr;   /* device CSR matrix                  */
typedef struct lfgpu_edges lfgpu_edges; /* codim-1 entities + segment FE    */

/* ---- context ----------------------------------------------------------- */
int lfgpu_ctx_create(int device, lfgpu_ctx **out);
int lfgpu_ctx_destroy(lfgpu_ctx *ctx);
/* Block until all work queued on the context's stream has finished. */
int lfgpu_ctx_sync(lfgpu_ctx *ctx);
/* Device time (hipEvent pair on the context stream) of the most recent
 * assembly / SpMV / PCG call, in milliseconds. */
int lfgpu_ctx_last_ms(lfgpu_ctx *ctx, double *ms);

/* ---- device vectors ---------------------------------------------------- */
int lfgpu_vec_create(lfgpu_ctx *ctx, int64_t n, const double *host_init,
                     lfgpu_vec **out);
int lfgpu_vec_destroy(lfgpu_vec *v);
int lfgpu_vec_size(const lfgpu_vec *v, int64_t *n);
int lfgpu_vec_upload(lfgpu_vec *v, const double *host, int64_t n);
int lfgpu_vec_download(const lfgpu_vec *v, double *host, int64_t n);
int lfgpu_vec_fill(lfgpu_vec *v, double value);
/* y = x on the context stream (both of the same size, different vectors). */
int lfgpu_vec_copy(lfgpu_ctx *ctx, const lfgpu_vec *x, lfgpu_vec *y);
/* y = x with non-temporal (cache-bypassing) loads and stores: calibration of
 * HBM byte counters for such streams (measurement support, no reference
 * counterpart). */
int lfgpu_vec_copy_nt(lfgpu_ctx *ctx, const lfgpu_vec *x, lfgpu_vec *y);

/* ---- mesh + DOF handler ------------------------------------------------
 * Replaces the cell loop's view of the mesh and of UniformFEDofHandler
 * (assemble/dofhandler.cc:86-305): xy [nv x 2] vertex coordinates,
 * cell_v [ncells x 3] vertex indices in the cell's local order, cell_dofs
 * [ncells x nloc] the GlobalDofIndices of each cell (NULL for P1, where the
 * DOFs are the vertex indices).  Degenerate cells (tria_o1.cc:10-48,
 * tolerance 1e-8) are rejected with LFGPU_EDEGENERATE. */
int lfgpu_mesh_create(lfgpu_ctx *ctx, int64_t nv, const double *xy,
                      int64_t ncells, const int32_t *cell_v, int nloc,
                      const int64_t *cell_dofs, int64_t ndofs,
                      lfgpu_mesh **out);
int lfgpu_mesh_destroy(lfgpu_mesh *m);

/* ---- host-side topology helpers (no GPU work) ---------------------------
 * TPTriagMeshBuilder::Build (mesh/utils/tp_triag_mesh_builder.cc:18-178):
 * xy [(nx+1)(ny+1) x 2], cell_v [2 nx ny x 3], edge_v [E x 2] with
 * E = nx ny + (nx+1) ny + nx (ny+1), in the builder's insertion order. */
int lfgpu_tp_triag_mesh(int32_t nx, int32_t ny, double x0, double y0,
                        double x1, double y1, double *xy, int32_t *cell_v,
                        int32_t *edge_v);
/* hybrid2d::Mesh edge numbering (mesh/hybrid2d/mesh.cc:236-575) and
 * relative orientations (mesh/hybrid2d/triangle.cc:71-77) for a triangle
 * mesh with npre predefined edges; *nedges receives E (edge_v_out holds
 * capacity max_edges).  Then UniformFEDofHandler::initIndexArrays
 * (assemble/dofhandler.cc:142-284) for {node:1, edge:ndof_edge,
 * cell:ndof_cell}: cell_dofs [ncells x (3+3 ndof_edge+ndof_cell)]. */
int lfgpu_mesh_edges(int64_t nv, int64_t ncells, const int32_t *cell_v,
                     int64_t npre, const int32_t *pre_edge_v,
                     int64_t max_edges, int32_t *edge_v_out, int32_t *cell_e,
                     int8_t *cell_ori, int64_t *nedges);
int lfgpu_dof_table(int64_t nv, int64_t ncells, int64_t nedges,
                    const int32_t *cell_v, const int32_t *cell_e,
                    const int8_t *cell_ori, int ndof_edge, int ndof_cell,
                    int64_t *cell_dofs, int64_t *ndofs);

/* ---- reference finite element + quadrature ------------------------------
 * PrecomputedScalarReferenceFiniteElement
 * (uscalfe/precomputed_scalar_reference_finite_element.h:72-78) for
 * FeLagrangeO1Tria (degree 1) or FeLagrangeO3Tria (degree 3)
 * (uscalfe/lagr_fe.h:57-150, :945-1140).  pts/w == NULL selects
 * make_QuadRule(kTria, 2*degree) (quad/make_quad_rule.cc:38-50), which is
 * what the providers' first constructor does (loc_comp_ellbvp.h:210-231);
 * otherwise pts is [2 x nqp] column-major (a user quad_rule_collection_t). */
int lfgpu_fe_create(lfgpu_ctx *ctx, int degree, int nqp, const double *pts,
                    const double *w, lfgpu_fe **out);
int lfgpu_fe_destroy(lfgpu_fe *fe);
int lfgpu_fe_info(const lfgpu_fe *fe, int *nloc, int *nqp);

/* ---- coefficients (MeshFunction arguments of the providers) ------------- */
#define LFGPU_COEFF_CONST 0  /* MeshFunctionConstant(value)                  */
#define LFGPU_COEFF_CELL 1   /* value * data[cell]  (CodimMeshDataSet)       */
#define LFGPU_COEFF_QP 2     /* value * data[cell*nqp + k]                   */
#define LFGPU_COEFF_SINSIN 3 /* value * sin(pi x) sin(pi y) at TriaO1::Global */
#define LFGPU_COEFF_EXPCELL 4 /* value * exp(data[cell])                      */
typedef struct lfgpu_coeff {
  int kind;
  double value;
  const lfgpu_vec *data;
} lfgpu_coeff;

/* ---- element-level kernels (COO value streams, reference order) ---------
 * Geometry of every cell as TriaO1 precomputes it (geometry/tria_o1.cc:50-68):
 * J and J^{-T} row-major 2x2 per cell, |det J| per cell. */
int lfgpu_geometry(lfgpu_ctx *ctx, const lfgpu_mesh *m, lfgpu_vec *J,
                   lfgpu_vec *JinvT, lfgpu_vec *det);
/* ReactionDiffusionElementMatrixProvider::Eval for all cells
 * (uscalfe/loc_comp_ellbvp.h:266-339): out [ncells x nloc x nloc], row-major
 * per cell; that is the triplet value stream of AssembleMatrixLocally
 * (assemble/assembler.h:166-179) in its order. */
int lfgpu_element_matrices(lfgpu_ctx *ctx, const lfgpu_mesh *m,
                           const lfgpu_fe *fe, const lfgpu_coeff *alpha,
                           const lfgpu_coeff *gamma, lfgpu_vec *out);
/* The same element matrices as lfgpu_element_matrices, computed as a
 * contraction of per-cell coefficients with fixed reference matrices on the
 * FP64 matrix cores (the B^T D B products of loc_comp_ellbvp.h:324-337
 * regrouped).  Entries agree with lfgpu_element_matrices up to rounding,
 * within 1e-12 of the row's largest entry, not bit for bit. */
int lfgpu_element_matrices_mfma(lfgpu_ctx *ctx, const lfgpu_mesh *m,
                                const lfgpu_fe *fe, const lfgpu_coeff *alpha,
                                const lfgpu_coeff *gamma, lfgpu_vec *out);
/* ScalarLoadElementVectorProvider::Eval for all cells
 * (uscalfe/loc_comp_ellbvp.h:691-746): out [ncells x nloc]. */
int lfgpu_element_vectors(lfgpu_ctx *ctx, const lfgpu_mesh *m,
                          const lfgpu_fe *fe, const lfgpu_coeff *f,
                          lfgpu_vec *out);

/* ---- assembly plan: sparsity pattern + deterministic scatter ------------
 * Builds, on the device, the compressed pattern that
 * FixFlaggedSolutionComponents (assemble/fix_dof.h:86-138) followed by
 * COOMatrix::makeSparse (coomatrix.h:172-180) yields for the triplets of
 * AssembleMatrixLocally: rows/cols of fixed DOFs are removed except for a
 * unit diagonal.  fixed_flags [ndofs] (host, NULL = no fixed DOFs) is the
 * selector of FixFlaggedSolutionComponents.  The plan also holds, per row,
 * its contributing (cell, local row) pairs in cell order with the CSR slot
 * of every local column, so re-assembly sums each entry in triplet order
 * without atomics. */
int lfgpu_plan_create(lfgpu_ctx *ctx, const lfgpu_mesh *m,
                      const uint8_t *fixed_flags, lfgpu_plan **out);
int lfgpu_plan_destroy(lfgpu_plan *p);
/* Borrow the plan's CSR (values are written by lfgpu_assemble_galerkin). */
int lfgpu_plan_csr(lfgpu_plan *p, lfgpu_csr **csr);

/* Fused AssembleMatrixLocally + AssembleVectorLocally +
 * FixFlaggedSolutionComponents + makeSparse: writes the plan's CSR values and
 * rhs [ndofs].  fixed_vals [ndofs] are the prescribed values of the fixed
 * DOFs (NULL = zero), applied as b -= A x_fixed, b[fixed] = value
 * (fix_dof.h:97-114).  f may be NULL (rhs then only carries the Dirichlet
 * terms), rhs may be NULL (matrix only). */
int lfgpu_assemble_galerkin(lfgpu_ctx *ctx, lfgpu_plan *p, const lfgpu_fe *fe,
                            const lfgpu_coeff *alpha, const lfgpu_coeff *gamma,
                            const lfgpu_coeff *f, const lfgpu_vec *fixed_vals,
                            lfgpu_vec *rhs);

/* lfgpu_assemble_galerkin with the element matrices taken from
 * lfgpu_element_matrices_mfma (kept in a plan-owned buffer of
 * ncells x nloc^2 doubles, allocated on the first call) and gathered in the
 * same triplet order.  Pattern and rhs (without prescribed nonzero values)
 * are identical to lfgpu_assemble_galerkin; matrix values agree up to the
 * element matrices' rounding. */
int lfgpu_assemble_galerkin_mfma(lfgpu_ctx *ctx, lfgpu_plan *p,
                                 const lfgpu_fe *fe, const lfgpu_coeff *alpha,
                                 const lfgpu_coeff *gamma, const lfgpu_coeff *f,
                                 const lfgpu_vec *fixed_vals, lfgpu_vec *rhs);

/* ---- boundary (codim-1) terms -------------------------------------------
 * MassEdgeMatrixProvider (uscalfe/loc_comp_ellbvp.h:367-527) and
 * ScalarLoadEdgeVectorProvider (:784-918), assembled with
 * AssembleMatrixLocally(1, ...) / AssembleVectorLocally(1, ...) after the
 * cell terms, as the impedance/Neumann part of
 * examples/ellbvp_linfe/ellbvp_linfe_demo.cc:272-315.
 * lfgpu_edges_create: edge_v [nedges x 2] the endpoints of every edge in its
 * orientation (lfgpu_mesh_edges returns them), edge_dofs [nedges x (degree+1)]
 * the edge's GlobalDofIndices (dofhandler.cc:160-181: endpoint DOFs, then
 * interior DOFs; NULL for degree 1, where they are the endpoints).  The
 * segment element is FeLagrangeO1Segment / FeLagrangeO3Segment
 * (lagr_fe.h:284-330, :1155-1230) with make_QuadRule(kSegment, 2 degree)
 * (Gauss-Legendre, degree + 1 points; quad/make_quad_rule.cc:13-18). */
int lfgpu_edges_create(lfgpu_ctx *ctx, const lfgpu_mesh *m, int degree,
                       int64_t nedges, const int32_t *edge_v,
                       const int64_t *edge_dofs, lfgpu_edges **out);
int lfgpu_edges_destroy(lfgpu_edges *e);
/* MassEdgeMatrixProvider::Eval / ScalarLoadEdgeVectorProvider::Eval for every
 * edge: out [nedges x nle x nle] row-major, [nedges x nle].  Coefficient
 * kinds as for cells with "cell" read as "edge": CONST, CELL (per edge), QP
 * (per edge and point), SINSIN (at SegmentO1::Global of the point), EXPCELL. */
int lfgpu_edge_matrices(lfgpu_ctx *ctx, const lfgpu_edges *e,
                        const lfgpu_coeff *eta, lfgpu_vec *out);
int lfgpu_edge_vectors(lfgpu_ctx *ctx, const lfgpu_edges *e,
                       const lfgpu_coeff *h, lfgpu_vec *out);
/* Attach boundary terms to a plan: mat_flags [nedges] selects the edges of
 * the edge mass term, vec_flags [nedges] those of the edge load term (host
 * arrays, NULL = no such term) -- the edge selectors of the two providers.
 * The edge DOFs of a cell mesh couple only DOFs of one cell, so the pattern
 * is unchanged; the plan gains per-row edge records in edge order. */
int lfgpu_plan_set_edges(lfgpu_plan *p, const lfgpu_edges *e,
                         const uint8_t *mat_flags, const uint8_t *vec_flags);
/* lfgpu_assemble_galerkin plus the plan's edge terms: eta the coefficient
 * of the edge mass term, h the edge source (either may be NULL when the plan
 * has no such term).  Triplet order as the reference: cell triplets, then
 * edge triplets; rhs = cell loads, then edge loads, then the Dirichlet
 * update over all removed triplets in that order. */
int lfgpu_assemble_galerkin_bd(lfgpu_ctx *ctx, lfgpu_plan *p, const lfgpu_fe *fe,
                               const lfgpu_coeff *alpha, const lfgpu_coeff *gamma,
                               const lfgpu_coeff *f, const lfgpu_coeff *eta,
                               const lfgpu_coeff *h, const lfgpu_vec *fixed_vals,
                               lfgpu_vec *rhs);

/* ---- CSR matrices ------------------------------------------------------ */
int lfgpu_csr_info(const lfgpu_csr *A, int32_t *nrows, int32_t *ncols,
                   int64_t *nnz);
int lfgpu_csr_download(const lfgpu_csr *A, int32_t *row_ptr, int32_t *col,
                       double *val);
/* COOMatrix::makeSparse -> setFromTriplets (coomatrix.h:172-180) on the
 * device for host triplets: stable sort by (row, col), duplicates summed in
 * triplet order, explicit zeros kept. */
int lfgpu_coo_to_csr(lfgpu_ctx *ctx, int64_t nnz, const int32_t *rows,
                     const int32_t *cols, const double *vals, int32_t nrows,
                     int32_t ncols, lfgpu_csr **out);
int lfgpu_csr_destroy(lfgpu_csr *A);
/* y = A x (FP64). */
int lfgpu_spmv(lfgpu_ctx *ctx, const lfgpu_csr *A, const lfgpu_vec *x,
               lfgpu_vec *y);

/* y = A x launched reps times back to back between one hipEvent pair on the
 * context stream; *avg_ms is the mean duration of one SpMV (measurement
 * support for roofline figures; no reference counterpart). */
int lfgpu_spmv_timed(lfgpu_ctx *ctx, const lfgpu_csr *A, const lfgpu_vec *x,
                     lfgpu_vec *y, int reps, double *avg_ms);

/* ---- Jacobi-preconditioned CG ------------------------------------------
 * Replaces Eigen::SparseLU in the demo drivers
 * (examples/ellbvp_linfe/homDir_linfe_demo.cc:171-174) for SPD systems.
 * Stopping rule and iteration count follow Eigen 3.4
 * ConjugateGradient<.., DiagonalPreconditioner>: stop when
 * ||b - A x||_2 < rtol ||b||_2.
 * mode LFGPU_PCG_FP64: everything in FP64.
 * mode LFGPU_PCG_MIXED: matrix, vectors and CG recurrences in FP32 (dot
 *   products and SpMV row sums accumulated in FP64) with reliable updates:
 *   the true residual b - A x is recomputed in FP64 whenever the recurrence
 *   residual has fallen by refine_delta, the solution is accumulated as a
 *   double-double pair, and convergence is judged on the FP64 true residual.
 * mode LFGPU_PCG_MIXED_MATRIX: as LFGPU_PCG_MIXED, but only the matrix values
 *   (and the Jacobi diagonal) are stored in FP32; the Krylov vectors stay
 *   FP64. */
#define LFGPU_PCG_FP64 0
#define LFGPU_PCG_MIXED 1
#define LFGPU_PCG_MIXED_MATRIX 2
typedef struct lfgpu_pcg_opts {
  double rtol;     /* relative residual target                    */
  int maxit;       /* iteration cap (Eigen: 2*n by default)       */
  int mode;        /* LFGPU_PCG_FP64 or LFGPU_PCG_MIXED            */
  int check_every; /* iterations per chunk; the host polls the state
                    * one step behind the device, every 1 to 8 chunks
                    * (sized from the residual history); chunks after
                    * the stop run empty                             */
  /* mixed mode only (0 selects the default):                        */
  double refine_delta;     /* reliable update once ||r|| fell by this  */
                           /* factor since the last one (default 0.3)  */
  int refine_max_interval; /* ... or after this many iterations        */
                           /* (default: never forced)                  */
  /* FP64 mode only: 1 = one fused sweep per iteration with the
   * single-reduction (Chronopoulos-Gear) recurrences, which give Eigen's
   * iterates in exact arithmetic and read the off-diagonal neighbours'
   * preconditioned residual on the fly; 0 = two sweeps with Eigen's own
   * recurrences (default). */
  int single_reduction;
  /* FP64 two-sweep and row-sweep iterations: 0 = replay each chunk of
   * check_every iterations from a captured hipGraph below 2^20 rows, where
   * the solve is bound by the launch rate (default); 1 = always; -1 = never
   * (plain launches).  The iterates are the same either way.  Up to two
   * captured chunks are kept per matrix, keyed by the solution vector. */
  int graph;
} lfgpu_pcg_opts;
typedef struct lfgpu_pcg_info {
  int iters;       /* Eigen's iteration count                     */
  int converged;   /* 1 if rtol was reached                       */
  int refinements; /* FP64 true-residual replacements (mixed)     */
  double relres;   /* final ||b - A x|| / ||b|| (FP64)            */
  double ms;       /* device time of the solve                    */
  int syncs;       /* host waits on the solver state (polls)      */
} lfgpu_pcg_info;
int lfgpu_pcg(lfgpu_ctx *ctx, const lfgpu_csr *A, const lfgpu_vec *b,
              lfgpu_vec *x, const lfgpu_pcg_opts *opts, lfgpu_pcg_info *info);

/* Storage of the solver's copy of A, built by the first SpMV / PCG call
 * (storage -1 before): LFGPU_STORAGE_DIA, LFGPU_STORAGE_ELL or
 * LFGPU_STORAGE_ELL16 (defined with the batched solves below), *ndiag the
 * stored off-diagonals of DIA, *sym 1
 * when the DIA values are mirror-symmetric bit for bit, so that the FP64
 * product sweep reads the upper diagonals only (ndiag and sym may be NULL).
 * Introspection for measurement; no reference counterpart. */
int lfgpu_solver_storage(const lfgpu_csr *A, int *storage, int *ndiag, int *sym);

/* lfgpu_pcg with every sweep kernel of the iterations bracketed by its own
 * hipEvent pair on the context stream (plain launches, no graph replay), for
 * measuring kernel durations; at most max_kernels kernels are timed.
 * out_ms[0], out_ms[1], out_ms[2]: median duration in milliseconds of the
 * first sweep of an iteration (the product), the second sweep (the vector
 * update) and the one-workgroup reduction kernels of the sliced-ELL row
 * sweeps (0 where a kind did not run).  Measurement support; no reference
 * counterpart.  Returns what lfgpu_pcg returns. */
int lfgpu_pcg_sweep_times(lfgpu_ctx *ctx, const lfgpu_csr *A,
                          const lfgpu_vec *b, lfgpu_vec *x,
                          const lfgpu_pcg_opts *opts, int max_kernels,
                          double *out_ms, lfgpu_pcg_info *info);

/* lfgpu_pcg in mode LFGPU_PCG_FP64 (two sweeps per iteration) for callers
 * that cannot allocate memory: no reference counterpart; written for code
 * that must not call hipMalloc (serving loops, harnesses that forbid
 * allocation).  It allocates nothing: its work vectors, partial sums and
 * solver state live in a fixed device array of the library (64 MiB per GPU,
 * held until the process ends), the product is taken from A's CSR directly
 * (no solver storage is built) and no graph is captured.  Systems of up to
 * 2^21 rows (LFGPU_ERANGE above); other modes and single_reduction are
 * rejected (LFGPU_EINVAL); opts->graph is ignored.  On a matrix that
 * lfgpu_pcg stores as DIA the iterates are lfgpu_pcg's.  All calls on one
 * GPU share the array: calls that overlap on one GPU (several contexts or
 * threads) race and must be serialised by the caller. */
int lfgpu_pcg_noalloc(lfgpu_ctx *ctx, const lfgpu_csr *A, const lfgpu_vec *b,
                      lfgpu_vec *x, const lfgpu_pcg_opts *opts,
                      lfgpu_pcg_info *info);

/* ---- batched solves --------------------------------------------------
 * Many independent SPD systems stored as one block-diagonal matrix (e.g. the
 * plan of a mesh that is the disjoint union of the problems' meshes, so one
 * assembly launch assembles all of them).  row_offsets [nprob+1] (host)
 * partitions the rows; problem p owns rows [row_offsets[p], row_offsets[p+1])
 * and gets its own Jacobi-PCG (Eigen semantics per problem).  All problems
 * advance together, one pair of sweeps per iteration; chunks of iterations
 * are captured into a hipGraph once and replayed, and the workgroup tile is
 * chosen by timing candidates on the first solve.  No reference
 * counterpart: a batch replaces nprob separate solve calls.
 * lfgpu_batch_create checks on the device that every entry of A lies inside
 * its row's problem block and returns LFGPU_EINVAL otherwise.
 * opts->graph: >= 0 replays the captured chunk (default); -1 launches each
 * chunk's kernels directly (the same iterates, no graph is captured). */
typedef struct lfgpu_batch lfgpu_batch;
int lfgpu_batch_create(lfgpu_ctx *ctx, lfgpu_csr *A, int nprob,
                       const int32_t *row_offsets, lfgpu_batch **out);
int lfgpu_batch_destroy(lfgpu_batch *bt);
/* infos: nprob entries (or NULL); summary: iters = max over problems,
 * converged = all converged, relres = worst, refinements = the tuned rows
 * per workgroup, ms = device time.  All three modes of lfgpu_pcg: in the
 * mixed modes every problem runs its own reliable updates and stops on its
 * FP64 true residual (infos[p].refinements counts them). */
int lfgpu_batch_pcg(lfgpu_ctx *ctx, lfgpu_batch *bt, const lfgpu_vec *b,
                    lfgpu_vec *x, const lfgpu_pcg_opts *opts,
                    lfgpu_pcg_info *infos, lfgpu_pcg_info *summary);
/* Matrix storage the batched sweeps use, chosen on the first solve (-1
 * before it): whole-matrix DIA, per-problem DIA (every problem banded, e.g.
 * structured meshes of different sizes: diagonals with per-problem offsets,
 * no column indices) or sliced ELL.  All give the same iterates. */
#define LFGPU_STORAGE_ELL 0
#define LFGPU_STORAGE_DIA 1
#define LFGPU_STORAGE_BLOCK_DIA 2
/* sliced ELL whose columns are stored as 16-bit offsets from the row
 * (chosen when every stored column lies within 32767 rows of its row) */
#define LFGPU_STORAGE_ELL16 3
int lfgpu_batch_storage(const lfgpu_batch *bt, int *storage);
/* lfgpu_assemble_galerkin of plan p (without prescribed Dirichlet values)
 * followed by lfgpu_batch_pcg on the batch created over p's matrix, with the
 * same results bit for bit.  The first call with a set of arguments runs the
 * two calls and captures a graph holding the assembly, the refresh of the
 * solver's copy of the values, the solve's set-up and its first chunk of
 * iterations; later calls with the same arguments (handles, coefficient
 * descriptors, vectors and options; the coefficient data may change) replay
 * it as one launch and continue with the captured chunks, polled as in
 * lfgpu_batch_pcg (summary->syncs).  summary->ms covers assembly and solve.
 * The batched form of the north-star pipeline (assemble, then solve, per
 * problem; no single reference counterpart: AssembleMatrixLocally +
 * AssembleVectorLocally + the solve of homDir_linfe_demo.cc:130-174). */
int lfgpu_batch_assemble_pcg(lfgpu_ctx *ctx, lfgpu_batch *bt, lfgpu_plan *p,
                             const lfgpu_fe *fe, const lfgpu_coeff *alpha,
                             const lfgpu_coeff *gamma, const lfgpu_coeff *f,
                             lfgpu_vec *rhs, lfgpu_vec *x, const lfgpu_pcg_opts *opts,
                             lfgpu_pcg_info *infos, lfgpu_pcg_info *summary);

/* ---- error norms ---------------------------------------------------------
 * Squared L2 and H1-seminorm errors of the finite element function with DOF
 * vector uh [ndofs] against an exact solution, integrated with the finite
 * element's quadrature rule: the post-processing of the demo drivers
 * (examples/ellbvp_linfe/homDir_linfe_demo.cc:179-188), i.e.
 * IntegrateMeshFunction (fe/fe_tools.h:34-66, :107-125) of
 * squaredNorm(MeshFunctionFE - u) (fe/mesh_function_fe.h:54-70) and of
 * squaredNorm(MeshFunctionGradFE - grad u) (fe/mesh_function_grad_fe.h).
 * exact LFGPU_EXACT_SINSIN: u = sin(pi x) sin(pi y) (homDir_linfe_demo.cc:
 * 44-52); LFGPU_EXACT_QP: u_qp [ncells x nqp] and grad_qp [ncells x nqp x 2]
 * hold the exact values at the rule's points.  Cells are summed by a fixed
 * reduction tree, so totals agree with the reference's cell-by-cell sum to
 * rounding. */
#define LFGPU_EXACT_SINSIN 0
#define LFGPU_EXACT_QP 1
int lfgpu_error_norms(lfgpu_ctx *ctx, const lfgpu_mesh *m, const lfgpu_fe *fe,
                      const lfgpu_vec *uh, int exact, const lfgpu_vec *u_qp,
                      const lfgpu_vec *grad_qp, double *l2sq, double *h1sq);

/* ---- adjoint gradient ----------------------------------------------------
 * For the P1 problem A(theta) u = f with A(theta) = sum_e exp(theta_e) K_e
 * (+ reaction) assembled by the plan, J(u) = 1/2 (u-u_obs)^T M (u-u_obs) and
 * the adjoint solution lambda of A lambda = M (u - u_obs):
 *   grad[e] = -exp(theta_e) lambda_e^T K_e u_e,
 * K_e the unit-coefficient element stiffness of
 * ReactionDiffusionElementMatrixProvider::Eval (alpha = 1, gamma = 0). */
int lfgpu_adjoint_gradient(lfgpu_ctx *ctx, const lfgpu_mesh *m,
                           const lfgpu_fe *fe, const lfgpu_vec *theta,
                           const lfgpu_vec *u, const lfgpu_vec *lam,
                           lfgpu_vec *grad);

/* ---- time stepping on a fixed pattern ------------------------------------
 * Strang splitting of the Fisher/KPP equation
 *   M u' = -A u  (diffusion)  |  u' = lambda u (1 - u / K)  (reaction)
 * as FisherKPP::StrangSplit (projects/fisher_kpp/strangsplitting.h:28-76,
 * :176-201; strangsplitting.cc:15-94).  A and M share one sparsity pattern
 * (both assembled by one plan).  The stepper keeps its own copy of A and
 * forms S1 = M + (tau kappa / 2) A and S2 = M + tau kappa A once, with
 * tau = T / m and kappa = 1 - sqrt(2)/2 (SDIRK-2).  The two SparseLU
 * factorisations of the reference become Jacobi-PCG solves to opts->rtol on
 * S1 / S2, warm-started from the previous stage.  The nonlocal loss matrix L
 * of the reference (a dense N x N matrix subtracted from A) is not
 * supported: A is used as given (L = 0). */
typedef struct lfgpu_strang lfgpu_strang;
typedef struct lfgpu_strang_info {
  int diffusion_steps; /* calls of the diffusion evolution operator */
  int solves;          /* PCG solves (two per diffusion step)     */
  int64_t iters;       /* PCG iterations summed over the solves   */
  int max_iters;       /* largest iteration count of one solve    */
  int converged;       /* 1 if every solve reached rtol           */
  double worst_relres; /* largest final relative residual         */
  double ms;           /* device time of the call                 */
} lfgpu_strang_info;
/* StrangSplit ctor: T the final time, m the number of time steps (>= 1),
 * lambda the growth factor; opts the PCG settings of every solve
 * (check_every <= 0: adapted per stage to the previous solve's count). */
int lfgpu_strang_create(lfgpu_ctx *ctx, const lfgpu_csr *A, const lfgpu_csr *M,
                        double T, int m, double lambda,
                        const lfgpu_pcg_opts *opts, lfgpu_strang **out);
int lfgpu_strang_destroy(lfgpu_strang *s);
/* diffusionEvolutionOperator (strangsplitting.cc:15-50): one SDIRK-2 step of
 * M u' = -A u over tau / 2 (firstcall != 0) or tau from mu; out may alias
 * mu. */
int lfgpu_strang_diffusion(lfgpu_ctx *ctx, lfgpu_strang *s, int firstcall,
                           const lfgpu_vec *mu, lfgpu_vec *out,
                           lfgpu_strang_info *info);
/* Evolution (strangsplitting.cc:53-94): half a diffusion step, then
 * max(m - 2, 0) times { reaction over tau, diffusion over tau } (the
 * reference's loop bounds), then a reaction over tau / 2, where the
 * reaction is the exact logistic flow
 *   u <- cap / (1 + (cap / u - 1) exp(-lambda t))
 * with carrying capacity cap [n].  out may alias mu.  A stage solve that
 * misses rtol within opts->maxit returns LFGPU_ENOCONV at once (info holds
 * the counts so far, out is not written). */
int lfgpu_strang_evolution(lfgpu_ctx *ctx, lfgpu_strang *s,
                           const lfgpu_vec *cap, const lfgpu_vec *mu,
                           lfgpu_vec *out, lfgpu_strang_info *info);

/* ---- CSR helpers for drivers that combine matrices ----------------------- */
/* A copy of A's pattern and values (Eigen::SparseMatrix copy, as
 * `A_ = galerkinpair.first` in strangsplitting.h:183-184), e.g. to keep an
 * assembled matrix while the plan assembles another operator. */
int lfgpu_csr_clone(lfgpu_ctx *ctx, const lfgpu_csr *A, lfgpu_csr **out);
/* Z = a X + b Y for X and Y with identical patterns (checked), Z a new
 * matrix on that pattern; the sum `M_ + 0.5 * tau_ * kappa_ * A_` of
 * strangsplitting.h:196-199 with a = 1. */
int lfgpu_csr_axpby(lfgpu_ctx *ctx, double a, const lfgpu_csr *X, double b,
                    const lfgpu_csr *Y, lfgpu_csr **out);

/* ---- nonlocal boundary terms of the Fisher/KPP model ------------------- */
/* The kernel g(x, y) = 1 / (1 + |x - y|^2) of the model's nonlocal boundary
 * condition (projects/fisher_kpp/modelproblem_island_main.cc:76-80); with
 * rmax > 0 it is 0 unless rmin <= |x - y| <= rmax (the annulus of
 * strangsplitting_main.cc:160-165 and convergence_tau_main.cc:108-113).
 * flags [nedges] (host) select the boundary edges (the drivers' edge_pred).
 *
 * Gain: eta [nedges x nqe] receives h(x) = localQuadFunction(mesh, {segment
 * rule}, g(x, .), 1, edge_pred) (strangsplitting.h:78-115) at the segment
 * rule's points of every flagged edge (0 on the others), the coefficient of
 * the drivers' MassEdgeMatrixProvider(fe_space, mf_h, edge_pred): pass it as
 * an LFGPU_COEFF_QP coefficient to lfgpu_assemble_galerkin_bd. */
int lfgpu_nonlocal_gain(lfgpu_ctx *ctx, const lfgpu_edges *e,
                        const uint8_t *flags, double rmin, double rmax,
                        lfgpu_vec *eta);
/* Loss: the matrix L the drivers build column by column
 * (modelproblem_island_main.cc:92-126: column j = AssembleVectorLocally(1,
 * ...) of the edge load of L_j(x) = [AssembleVectorLocally(1, ...) of the
 * edge load of g(x, .)]_j), restricted to the nb DOFs of the flagged edges
 * (every other entry is 0): *nb receives nb, rows (host, capacity nb) those
 * DOFs in ascending order and L (device, >= nb^2 values) the block,
 * row-major.  rows = L = NULL queries nb only. */
int lfgpu_nonlocal_loss(lfgpu_ctx *ctx, const lfgpu_edges *e,
                        const uint8_t *flags, double rmin, double rmax,
                        int32_t *nb, int32_t *rows, lfgpu_vec *L);
/* Attaches a dense block to A: A then stands for (its sparse part) +
 * scale * B, B [nb x nb] row-major on the rows and columns rows[0..nb)
 * (host, ascending), copied.  This is the dense `A_ds - L` of
 * strangsplitting.h:166-170 (scale -1, B = L) without densifying the sparse
 * part.  lfgpu_spmv, lfgpu_pcg in mode LFGPU_PCG_FP64 (two sweeps; its
 * Jacobi entries include scale * B's diagonal), lfgpu_csr_clone and
 * lfgpu_csr_axpby (the block of the one operand that has one, scaled by its
 * factor) honour it; lfgpu_csr_download returns the sparse part; the other
 * modes, the batch and lfgpu_pcg_noalloc reject such a matrix.  nb = 0
 * removes the block. */
int lfgpu_csr_set_dense_block(lfgpu_ctx *ctx, lfgpu_csr *A, int32_t nb,
                              const int32_t *rows, const lfgpu_vec *B,
                              double scale);
/* The attached block's size and scale (0 and 0.0 without one). */
int lfgpu_csr_dense_block(const lfgpu_csr *A, int32_t *nb, double *scale);

/* ---- small BLAS-1 helpers used by drivers ------------------------------- */
/* y = a*x + b*y */
int lfgpu_axpby(lfgpu_ctx *ctx, double a, const lfgpu_vec *x, double b,
                lfgpu_vec *y);
/* *out = x . y (FP64, deterministic order) */
int lfgpu_dot(lfgpu_ctx *ctx, const lfgpu_vec *x, const lfgpu_vec *y,
              double *out);

#ifdef __cplusplus
}
#endif
#endif /* LFGPU_H */
