// Assembly plan (pattern + deterministic scatter records) and the fused
// row-owner assembly kernel.
//
// Reference semantics being reproduced (assemble/assembler.h:114-186,
// fix_dof.h:86-138, coomatrix.h:172-180): the triplets of cell c come in
// order c, i, j; FixFlaggedSolutionComponents removes every triplet touching
// a fixed row or column and appends (d, d, 1); setFromTriplets sums equal
// (row, col) triplets in triplet order.  Hence an entry (r, s) of the final
// matrix is the sum, in increasing cell order, of M_c(i, j) over the cells c
// with dof(c, i) = r and dof(c, j) = s.
//
// Plan (once per mesh, on the device): a counting sort of the (cell, i)
// pairs by row r = dof(c, i), made stable by sorting each row's short list by
// (cell, i); per row the sorted unique column set; and for every (row, cell)
// record the CSR slot of each local column.  Assembly then runs one thread
// per row that visits its cells in increasing order, computes the cell's
// element-matrix row on the fly and adds it into workgroup-local LDS
// accumulators laid out exactly like the row's CSR segment.  No atomics, and
// each entry is summed in the reference's triplet order.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <string>
#include <vector>

#include "lfgpu_fe_device.h"
#include "lfgpu_internal.h"

using namespace lfgpu;

namespace {

constexpr uint8_t kDrop = 0xFF;  // local column removed by the Dirichlet fix
constexpr int kMaxRowLen = 254;  // slots are stored as bytes

inline unsigned grid_for(int64_t n, int block) {
  int64_t g = (n + block - 1) / block;
  return (unsigned)(g < 1 ? 1 : g);
}

__global__ void k_count_records(const int32_t *cell_dofs, int64_t nt, const uint8_t *fixed,
                                int32_t *cnt) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nt) return;
  const int32_t d = cell_dofs[t];
  if (!fixed[d]) atomicAdd(&cnt[d], 1);
}

__global__ void k_fill_records(const int32_t *cell_dofs, int64_t nt, const uint8_t *fixed,
                               const int32_t *rec_ptr, int32_t *cursor, int32_t *rec_key) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nt) return;
  const int32_t d = cell_dofs[t];
  if (fixed[d]) return;
  const int32_t pos = rec_ptr[d] + atomicAdd(&cursor[d], 1);
  rec_key[pos] = (int32_t)t;  // t = cell * nloc + i
}

// Per row: order the records by (cell, i) and collect the sorted unique
// columns into tmp[rec_ptr[r] * nloc ...].
__global__ void k_row_pattern(int64_t ndofs, int nloc, const int32_t *cell_dofs,
                              const uint8_t *fixed, const int32_t *rec_ptr, int32_t *rec_key,
                              int32_t *tmp, int32_t *row_len) {
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= ndofs) return;
  if (fixed[r]) {
    row_len[r] = 1;
    return;
  }
  const int32_t b = rec_ptr[r], e = rec_ptr[r + 1];
  for (int32_t p = b + 1; p < e; ++p) {  // insertion sort of a short list
    const int32_t k = rec_key[p];
    int32_t q = p - 1;
    while (q >= b && rec_key[q] > k) {
      rec_key[q + 1] = rec_key[q];
      --q;
    }
    rec_key[q + 1] = k;
  }
  int32_t *c = tmp + (int64_t)b * nloc;
  int32_t m = 0;
  for (int32_t p = b; p < e; ++p) {
    const int64_t cell = rec_key[p] / nloc;
    for (int j = 0; j < nloc; ++j) {
      const int32_t col = cell_dofs[cell * nloc + j];
      if (fixed[col]) continue;
      int32_t q = m - 1;  // sorted insert, skipping duplicates
      bool dup = false;
      while (q >= 0 && c[q] >= col) {
        if (c[q] == col) {
          dup = true;
          break;
        }
        --q;
      }
      if (dup) continue;
      for (int32_t s = m; s > q + 1; --s) c[s] = c[s - 1];
      c[q + 1] = col;
      ++m;
    }
  }
  row_len[r] = m;
}

__global__ void k_row_finalize(int64_t ndofs, int nloc, int meta_stride,
                               const int32_t *cell_dofs, const uint8_t *fixed,
                               const int32_t *rec_ptr, const int32_t *rec_key, const int32_t *tmp,
                               const int32_t *row_ptr, int32_t *col, uint8_t *meta,
                               int *too_long) {
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= ndofs) return;
  const int32_t o = row_ptr[r], len = row_ptr[r + 1] - o;
  if (fixed[r]) {
    col[o] = (int32_t)r;
    return;
  }
  if (len > kMaxRowLen) atomicMax(too_long, len);
  const int32_t b = rec_ptr[r], e = rec_ptr[r + 1];
  const int32_t *c = tmp + (int64_t)b * nloc;
  for (int32_t s = 0; s < len; ++s) col[o + s] = c[s];
  for (int32_t p = b; p < e; ++p) {
    const int32_t key = rec_key[p];
    const int64_t cell = key / nloc;
    uint8_t *mp = meta + (int64_t)p * meta_stride;
    mp[0] = (uint8_t)(key - cell * nloc);
    for (int j = 0; j < nloc; ++j) {
      const int32_t cj = cell_dofs[cell * nloc + j];
      uint8_t slot = kDrop;
      if (!fixed[cj]) {
        int32_t lo = 0, hi = len - 1;
        while (lo < hi) {
          const int32_t mid = (lo + hi) >> 1;
          if (c[mid] < cj) lo = mid + 1; else hi = mid;
        }
        slot = (uint8_t)lo;
      }
      mp[1 + j] = slot;
    }
  }
}

__global__ void k_block_nnz_max(const int32_t *row_ptr, int64_t n, int block, int *out) {
  const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t r0 = b * block;
  if (r0 >= n) return;
  const int64_t r1 = r0 + block < n ? r0 + block : n;
  atomicMax(out, row_ptr[r1] - row_ptr[r0]);
}

}  // namespace

extern "C" int lfgpu_plan_destroy(lfgpu_plan *p) {
  if (!p) return LFGPU_OK;
  if (p->ctx) (void)hipStreamSynchronize(p->ctx->stream);
  solver_cache_free(&p->csr);
  dfree(p->csr.row_ptr);
  dfree(p->csr.col);
  dfree(p->csr.val);
  dfree(p->rec_ptr);
  dfree(p->rec_key);
  dfree(p->rec_meta);
  dfree(p->fixed);
  dfree(p->ke);
  dfree(p->fe);
  dfree(p->coef);
  plan_free_edges(p);
  delete p;
  return LFGPU_OK;
}

extern "C" int lfgpu_plan_create(lfgpu_ctx *ctx, const lfgpu_mesh *m, const uint8_t *fixed_flags,
                                 lfgpu_plan **out) {
  LFG_REQUIRE(ctx && m && out, LFGPU_EINVAL, "NULL argument to lfgpu_plan_create");
  *out = nullptr;
  hipStream_t s = ctx->stream;
  const int64_t nd = m->ndofs, nt = m->ncells * m->nloc;
  const int nloc = m->nloc;
  lfgpu_plan *p = new lfgpu_plan();
  p->ctx = ctx;
  p->mesh = m;
  p->csr.ctx = ctx;
  p->csr.nrows = (int32_t)nd;
  p->csr.ncols = (int32_t)nd;
  p->csr.owned_by_plan = true;
  p->meta_stride = nloc == 3 ? 4 : 12;
  int32_t *cnt = nullptr, *cursor = nullptr, *tmp = nullptr, *row_len = nullptr;
  int *dflag = nullptr;
  auto fail = [&](int rc) {
    dfree(cnt);
    dfree(cursor);
    dfree(tmp);
    dfree(row_len);
    dfree(dflag);
    lfgpu_plan_destroy(p);
    return rc;
  };
#define PLAN_TRY(expr)                 \
  do {                                 \
    int rc_ = (expr);                  \
    if (rc_ != LFGPU_OK) return fail(rc_); \
  } while (0)
#define PLAN_HIP(expr)                                        \
  do {                                                        \
    hipError_t e_ = (expr);                                   \
    if (e_ != hipSuccess) return fail(hip_fail(e_, #expr));   \
  } while (0)

  PLAN_TRY(dalloc(&p->fixed, (size_t)nd, "fixed flags"));
  if (fixed_flags) {
    std::vector<uint8_t> fl(fixed_flags, fixed_flags + nd);
    for (auto &v : fl) v = v ? 1 : 0;
    p->has_fixed = std::any_of(fl.begin(), fl.end(), [](uint8_t v) { return v != 0; });
    PLAN_HIP(hipMemcpyAsync(p->fixed, fl.data(), nd, hipMemcpyHostToDevice, s));
    PLAN_HIP(hipStreamSynchronize(s));
  } else {
    PLAN_HIP(hipMemsetAsync(p->fixed, 0, nd, s));
  }
  // 1) records per row
  PLAN_TRY(dalloc(&cnt, (size_t)nd, "record counts"));
  PLAN_TRY(dalloc(&cursor, (size_t)nd, "record cursors"));
  PLAN_TRY(dalloc(&p->rec_ptr, (size_t)nd + 1, "record offsets"));
  PLAN_HIP(hipMemsetAsync(cnt, 0, nd * 4, s));
  PLAN_HIP(hipMemsetAsync(cursor, 0, nd * 4, s));
  hipLaunchKernelGGL(k_count_records, dim3(grid_for(nt, 256)), dim3(256), 0, s, m->cell_dofs, nt,
                     p->fixed, cnt);
  PLAN_HIP(hipGetLastError());
  int64_t nrec = 0;
  PLAN_TRY(exclusive_scan_i32(s, cnt, p->rec_ptr, nd, &nrec));
  p->nrec = nrec;
  PLAN_TRY(dalloc(&p->rec_key, (size_t)nrec, "record keys"));
  PLAN_TRY(dalloc(&p->rec_meta, (size_t)nrec * p->meta_stride, "record slots"));
  hipLaunchKernelGGL(k_fill_records, dim3(grid_for(nt, 256)), dim3(256), 0, s, m->cell_dofs, nt,
                     p->fixed, p->rec_ptr, cursor, p->rec_key);
  PLAN_HIP(hipGetLastError());
  // 2) per-row ordering and column sets
  PLAN_TRY(dalloc(&tmp, (size_t)std::max<int64_t>(nrec, 1) * nloc, "column scratch"));
  PLAN_TRY(dalloc(&row_len, (size_t)nd, "row lengths"));
  hipLaunchKernelGGL(k_row_pattern, dim3(grid_for(nd, 128)), dim3(128), 0, s, nd, nloc,
                     m->cell_dofs, p->fixed, p->rec_ptr, p->rec_key, tmp, row_len);
  PLAN_HIP(hipGetLastError());
  PLAN_TRY(dalloc(&p->csr.row_ptr, (size_t)nd + 1, "CSR row pointers"));
  int64_t nnz = 0;
  PLAN_TRY(exclusive_scan_i32(s, row_len, p->csr.row_ptr, nd, &nnz));
  p->csr.nnz = nnz;
  PLAN_TRY(dalloc(&p->csr.col, (size_t)nnz, "CSR columns"));
  PLAN_TRY(dalloc(&p->csr.val, (size_t)nnz, "CSR values"));
  PLAN_TRY(dalloc(&dflag, 4, "status words"));
  PLAN_HIP(hipMemsetAsync(dflag, 0, 16, s));
  hipLaunchKernelGGL(k_row_finalize, dim3(grid_for(nd, 128)), dim3(128), 0, s, nd, nloc,
                     p->meta_stride, m->cell_dofs, p->fixed, p->rec_ptr, p->rec_key, tmp,
                     p->csr.row_ptr, p->csr.col, p->rec_meta, dflag);
  PLAN_HIP(hipGetLastError());
  // 3) launch shape of the assembly kernel: rows per workgroup such that the
  // workgroup's CSR segment fits the LDS accumulator budget (64 KiB).
  int blocks[3] = {256, 128, 64};
  int maxnnz[3] = {0, 0, 0};
  for (int q = 0; q < 3; ++q) {
    const int64_t nb = (nd + blocks[q] - 1) / blocks[q];
    hipLaunchKernelGGL(k_block_nnz_max, dim3(grid_for(nb, 256)), dim3(256), 0, s, p->csr.row_ptr,
                       nd, blocks[q], dflag + 1 + q);
    PLAN_HIP(hipGetLastError());
  }
  int hflag[4] = {0, 0, 0, 0};
  PLAN_HIP(hipMemcpyAsync(hflag, dflag, 16, hipMemcpyDeviceToHost, s));
  PLAN_HIP(hipStreamSynchronize(s));
  if (hflag[0] > 0) {
    set_error("row with " + std::to_string(hflag[0]) + " entries exceeds the plan limit of " +
              std::to_string(kMaxRowLen));
    return fail(LFGPU_ERANGE);
  }
  p->block = 0;
  for (int q = 0; q < 3; ++q) {
    maxnnz[q] = hflag[1 + q];
    if (!p->block && maxnnz[q] <= 8192) {
      p->block = blocks[q];
      p->max_block_nnz = maxnnz[q];
    }
  }
  if (!p->block) {
    set_error("assembly workgroup segment exceeds the LDS budget");
    return fail(LFGPU_ERANGE);
  }
  dfree(cnt);
  dfree(cursor);
  dfree(tmp);
  dfree(row_len);
  dfree(dflag);
  *out = p;
  return LFGPU_OK;
#undef PLAN_TRY
#undef PLAN_HIP
}

extern "C" int lfgpu_plan_csr(lfgpu_plan *p, lfgpu_csr **csr) {
  LFG_REQUIRE(p && csr, LFGPU_EINVAL, "NULL argument");
  *csr = &p->csr;
  return LFGPU_OK;
}

// ===========================================================================
// fused row-owner assembly
// ===========================================================================
namespace {

struct AsmArgs {
  const double *xy;
  const int32_t *cell_v;
  const int32_t *cell_dofs;
  const int32_t *rec_ptr;
  const int32_t *rec_key;
  const uint8_t *rec_meta;
  const uint8_t *fixed;
  const int32_t *row_ptr;
  double *val;
  double *rhs;             // may be null
  const double *fixed_vals;  // may be null
  const double *tab;
  CoeffD alpha, gamma, f;
  int has_f;
  int64_t nrows;
  // boundary edge terms (null pointers: none)
  const int32_t *erec_ptr, *erec_key;
  const uint8_t *erec_meta;
  const int32_t *vrec_ptr, *vrec_key;
  const int32_t *edge_v, *edge_dofs;
  const double *etab;
  CoeffD eta, h;
};

template <int NLOC, int MS>
__device__ __forceinline__ void load_meta(const uint8_t *meta, int32_t p, uint8_t *out) {
  if constexpr (MS == 4) {
    const uint32_t w = reinterpret_cast<const uint32_t *>(meta)[p];
#pragma unroll
    for (int q = 0; q < 4; ++q) out[q] = (uint8_t)(w >> (8 * q));
  } else {
    const uint32_t *wp = reinterpret_cast<const uint32_t *>(meta + (int64_t)p * MS);
#pragma unroll
    for (int r = 0; r < MS / 4; ++r) {
      const uint32_t w = wp[r];
#pragma unroll
      for (int q = 0; q < 4; ++q) out[4 * r + q] = (uint8_t)(w >> (8 * q));
    }
  }
}

// Row i of edge e's mass matrix and its CSR slots (edge record q).
template <int NLE, int NQE>
__device__ __forceinline__ void edge_record_row(const AsmArgs &a, const ETab<NLE, NQE> &ET,
                                                int32_t key, int64_t &e, double *row) {
  e = key / NLE;
  const int i = key - (int)e * NLE;
  const double2 pa = reinterpret_cast<const double2 *>(a.xy)[a.edge_v[2 * e]];
  const double2 pb = reinterpret_cast<const double2 *>(a.xy)[a.edge_v[2 * e + 1]];
  double gk[NQE];
#pragma unroll
  for (int k = 0; k < NQE; ++k) gk[k] = edge_coeff<NQE>(a.eta, e, k, pa, pb, ET.px(k));
  edge_row<NLE, NQE>(ET, seg_length(pa, pb), gk, i, row);
}

// PRE: the cell coefficients went through the per-cell pass.  The P1
// kernel with the costly coefficient kinds is compute-bound (exp and sines
// recomputed on every visit of a cell): asking for 3 waves per SIMD caps it
// at 168 registers (from 169 it would use, with 2 spilled) and measured 13%
// faster on the 2048x2048 mesh than 2 waves, with the same bits.
template <int NLOC, int NQP, int MS, int BLOCK, bool EDGES, bool PRE>
__global__ __launch_bounds__(BLOCK, (NLOC == 3 && !PRE && BLOCK == 256 ? 3 : 1)) void
k_assemble_rows(AsmArgs a) {
  constexpr int NLE = NLOC == 3 ? 2 : 4, NQE = NLE;
  extern __shared__ double s_acc[];
  __shared__ double s_tab[Tab<NLOC, NQP>::kLen];
  __shared__ double s_etab[ETab<NLE, NQE>::kLen];
  for (int t = threadIdx.x; t < Tab<NLOC, NQP>::kLen; t += BLOCK) s_tab[t] = a.tab[t];
  if (EDGES)
    for (int t = threadIdx.x; t < ETab<NLE, NQE>::kLen; t += BLOCK) s_etab[t] = a.etab[t];
  const int64_t r0 = (int64_t)blockIdx.x * BLOCK;
  const int64_t r1 = r0 + BLOCK < a.nrows ? r0 + BLOCK : a.nrows;
  const int32_t base = a.row_ptr[r0];
  const int32_t cnt = a.row_ptr[r1] - base;
  for (int t = threadIdx.x; t < cnt; t += BLOCK) s_acc[t] = 0.0;
  __syncthreads();
  const Tab<NLOC, NQP> T{s_tab};
  const int64_t r = r0 + threadIdx.x;
  if (r < r1) {
    const int32_t off = a.row_ptr[r] - base;
    if (a.fixed[r]) {
      s_acc[off] = 1.0;
      if (a.rhs) a.rhs[r] = a.fixed_vals ? a.fixed_vals[r] : 0.0;
    } else {
      double b = 0.0;
      const int32_t pb = a.rec_ptr[r], pe = a.rec_ptr[r + 1];
      for (int32_t p = pb; p < pe; ++p) {
        const int32_t key = a.rec_key[p];
        const int64_t cell = key / NLOC;
        uint8_t mt[MS];
        load_meta<NLOC, MS>(a.rec_meta, p, mt);
        const int i = mt[0];
        const CellXY c = load_cell(a.xy, a.cell_v, cell);
        Geo g;
        tria_geo(c.x0, c.y0, c.x1, c.y1, c.x2, c.y2, g);
        double row[NLOC];
        if constexpr (PRE) {
          element_row_f<NLOC, NQP>(T, g, PlainCoef(a.alpha, cell, NQP),
                                   PlainCoef(a.gamma, cell, NQP), i, row);
        } else {
          double ak[NQP], gk[NQP];
          coeff_qp<NQP>(a.alpha, cell, c, s_tab, ak);
          coeff_qp<NQP>(a.gamma, cell, c, s_tab, gk);
          element_row<NLOC, NQP>(T, g, ak, gk, i, row);
        }
#pragma unroll
        for (int j = 0; j < NLOC; ++j) {
          const uint8_t s = mt[1 + j];
          if (s != kDrop) s_acc[off + s] += row[j];
        }
        if (a.has_f) {
          if constexpr (PRE) {
            b += element_load_f<NLOC, NQP>(T, g, PlainCoef(a.f, cell, NQP), i);
          } else {
            double fk[NQP];
            coeff_qp<NQP>(a.f, cell, c, s_tab, fk);
            b += element_load<NLOC, NQP>(T, g, fk, i);
          }
        }
      }
      if constexpr (EDGES) {
        // AssembleMatrixLocally(1, ...) / AssembleVectorLocally(1, ...) come
        // after the cell terms: edge triplets and edge loads in edge order
        const ETab<NLE, NQE> ET{s_etab};
        if (a.erec_ptr) {
          for (int32_t q = a.erec_ptr[r]; q < a.erec_ptr[r + 1]; ++q) {
            int64_t e;
            double row[NLE];
            edge_record_row<NLE, NQE>(a, ET, a.erec_key[q], e, row);
            const uint32_t w = reinterpret_cast<const uint32_t *>(a.erec_meta)[q];
#pragma unroll
            for (int j = 0; j < NLE; ++j) {
              const uint8_t sl = (uint8_t)(w >> (8 * j));
              if (sl != kDrop) s_acc[off + sl] += row[j];
            }
          }
        }
        if (a.vrec_ptr) {
          for (int32_t q = a.vrec_ptr[r]; q < a.vrec_ptr[r + 1]; ++q) {
            const int32_t key = a.vrec_key[q];
            const int64_t e = key / NLE;
            const int i = key - (int)e * NLE;
            const double2 pa = reinterpret_cast<const double2 *>(a.xy)[a.edge_v[2 * e]];
            const double2 pb = reinterpret_cast<const double2 *>(a.xy)[a.edge_v[2 * e + 1]];
            double hk[NQE];
#pragma unroll
            for (int k = 0; k < NQE; ++k) hk[k] = edge_coeff<NQE>(a.h, e, k, pa, pb, ET.px(k));
            b += edge_load<NLE, NQE>(ET, seg_length(pa, pb), hk, i);
          }
        }
      }
      if (a.fixed_vals) {
        // b += A(r, col) * (-1 * x_fixed[col]) over the removed triplets, in
        // triplet order (fix_dof.h:97-106 via COOMatrix::MatVecMult).
        for (int32_t p = pb; p < pe; ++p) {
          const int32_t key = a.rec_key[p];
          const int64_t cell = key / NLOC;
          uint8_t mt[MS];
          load_meta<NLOC, MS>(a.rec_meta, p, mt);
          bool any = false;
#pragma unroll
          for (int j = 0; j < NLOC; ++j) any |= (mt[1 + j] == kDrop);
          if (!any) continue;
          const CellXY c = load_cell(a.xy, a.cell_v, cell);
          Geo g;
          tria_geo(c.x0, c.y0, c.x1, c.y1, c.x2, c.y2, g);
          double row[NLOC];
          if constexpr (PRE) {
            element_row_f<NLOC, NQP>(T, g, PlainCoef(a.alpha, cell, NQP),
                                     PlainCoef(a.gamma, cell, NQP), mt[0], row);
          } else {
            double ak[NQP], gk[NQP];
            coeff_qp<NQP>(a.alpha, cell, c, s_tab, ak);
            coeff_qp<NQP>(a.gamma, cell, c, s_tab, gk);
            element_row<NLOC, NQP>(T, g, ak, gk, mt[0], row);
          }
#pragma unroll
          for (int j = 0; j < NLOC; ++j)
            if (mt[1 + j] == kDrop)
              b += row[j] * (-1.0 * a.fixed_vals[a.cell_dofs[cell * NLOC + j]]);
        }
        if constexpr (EDGES) {
          if (a.erec_ptr) {
            const ETab<NLE, NQE> ET{s_etab};
            for (int32_t q = a.erec_ptr[r]; q < a.erec_ptr[r + 1]; ++q) {
              const uint32_t w = reinterpret_cast<const uint32_t *>(a.erec_meta)[q];
              bool any = false;
#pragma unroll
              for (int j = 0; j < NLE; ++j) any |= ((uint8_t)(w >> (8 * j)) == kDrop);
              if (!any) continue;
              int64_t e;
              double row[NLE];
              edge_record_row<NLE, NQE>(a, ET, a.erec_key[q], e, row);
#pragma unroll
              for (int j = 0; j < NLE; ++j)
                if ((uint8_t)(w >> (8 * j)) == kDrop)
                  b += row[j] * (-1.0 * a.fixed_vals[a.edge_dofs[e * NLE + j]]);
            }
          }
        }
      }
      if (a.rhs) a.rhs[r] = b;
    }
  }
  __syncthreads();
  for (int t = threadIdx.x; t < cnt; t += BLOCK) a.val[base + t] = s_acc[t];
}

// Coefficients whose value costs transcendental functions (EXPCELL: exp per
// cell; SINSIN: two sines per point) are evaluated once per cell, with the
// same expressions as coeff_qp, into out (EXPCELL: one value per cell;
// SINSIN: NQP per cell).  The row-owner kernel then reads them as a CELL /
// QP coefficient of value 1, and 1 * v == v, so every value it uses is the
// one it would have computed itself, for each of the cell's local rows.
template <int NQP>
__global__ void k_coeff_cells(CoeffD c, const double *xy, const int32_t *cell_v, int64_t ncells,
                              const double *tab, double *out) {
  const int64_t cell = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (cell >= ncells) return;
  if (c.kind == LFGPU_COEFF_EXPCELL) {
    out[cell] = c.value * exp(c.data[cell]);
    return;
  }
  const CellXY p = load_cell(xy, cell_v, cell);
  double v[NQP];
  coeff_qp<NQP>(c, cell, p, tab, v);
#pragma unroll
  for (int k = 0; k < NQP; ++k) out[cell * NQP + k] = v[k];
}

bool costly(const CoeffD &c) {
  return c.kind == LFGPU_COEFF_EXPCELL || c.kind == LFGPU_COEFF_SINSIN;
}

// Queues the per-cell evaluation of the costly coefficients of a into the
// plan's buffer and points a at it.
template <int NQP>
int precompute_coeffs(hipStream_t s, lfgpu_plan *p, AsmArgs &a) {
  CoeffD *cs[3] = {&a.alpha, &a.gamma, &a.f};
  const int64_t nc = p->mesh->ncells;
  const int64_t len = 3 * std::max<int64_t>(nc * NQP, 1);
  if (!costly(a.alpha) && !costly(a.gamma) && !(a.has_f && costly(a.f))) return LFGPU_OK;
  if (p->coef_len < len) {
    dfree(p->coef);
    p->coef = nullptr;
    p->coef_len = 0;
    LFG_TRY(dalloc(&p->coef, (size_t)len, "per-cell coefficient values"));
    p->coef_len = len;
  }
  for (int q = 0; q < 3; ++q) {
    CoeffD &c = *cs[q];
    if (!costly(c) || (q == 2 && !a.has_f) || nc == 0) continue;
    double *out = p->coef + q * (len / 3);
    hipLaunchKernelGGL(k_coeff_cells<NQP>, dim3(grid_for(nc, 256)), dim3(256), 0, s, c, a.xy,
                       a.cell_v, nc, a.tab, out);
    c = CoeffD{c.kind == LFGPU_COEFF_EXPCELL ? LFGPU_COEFF_CELL : LFGPU_COEFF_QP, 1.0, out};
  }
  LFG_HIP(hipGetLastError());
  return LFGPU_OK;
}

template <int NLOC, int NQP, int MS, bool EDGES, bool PRE>
int launch_assemble(hipStream_t s, const lfgpu_plan *p, const AsmArgs &a) {
  const size_t lds = (size_t)std::max(p->max_block_nnz, 1) * sizeof(double);
  const unsigned grid = grid_for(a.nrows, p->block);
  switch (p->block) {
    case 256:
      hipLaunchKernelGGL((k_assemble_rows<NLOC, NQP, MS, 256, EDGES, PRE>), dim3(grid), dim3(256), lds,
                         s, a);
      break;
    case 128:
      hipLaunchKernelGGL((k_assemble_rows<NLOC, NQP, MS, 128, EDGES, PRE>), dim3(grid), dim3(128), lds,
                         s, a);
      break;
    default:
      hipLaunchKernelGGL((k_assemble_rows<NLOC, NQP, MS, 64, EDGES, PRE>), dim3(grid), dim3(64), lds, s,
                         a);
      break;
  }
  LFG_HIP(hipGetLastError());
  return LFGPU_OK;
}

// Row-owner gather of precomputed element matrices and vectors: the same
// record walk and summation order as k_assemble_rows, reading row i of cell
// c from ke[(c nloc + i) nloc ...] instead of evaluating it.
struct KeArgs {
  const int32_t *cell_dofs;
  const int32_t *rec_ptr;
  const int32_t *rec_key;
  const uint8_t *rec_meta;
  const uint8_t *fixed;
  const int32_t *row_ptr;
  const double *ke;          // [i][cell][j] at P1, [cell][i][j] at P3 (ke_row)
  const double *fvec;        // may be null
  double *val;
  double *rhs;               // may be null
  const double *fixed_vals;  // may be null
  int64_t nrows;
  int64_t ncells;
};

// Row i of cell c's element matrix (key = c NLOC + i).  At P1 the matrices
// are stored [i][cell][j]: a workgroup's rows then read the rows i of
// neighbouring cells from adjacent addresses (2048^2 two-phase assembly:
// gather 1320 -> 876 us, contraction 565 -> 590 us).  At P3 the same layout
// scatters the contraction's stores for no gain in the gather (137 + 429 us
// against 166 + 408 us), so the cells stay [cell][i][j].
template <int NLOC>
constexpr bool ke_rowmajor_i() {
  return NLOC == 3;
}

template <int NLOC>
__device__ __forceinline__ const double *ke_row(const KeArgs &a, int32_t key) {
  if constexpr (ke_rowmajor_i<NLOC>()) {
    const int64_t c = key / NLOC;
    return a.ke + ((int64_t)(key - c * NLOC) * a.ncells + c) * NLOC;
  } else {
    return a.ke + (int64_t)key * NLOC;
  }
}

template <int NLOC, int MS, int BLOCK>
__global__ __launch_bounds__(BLOCK) void k_assemble_rows_ke(KeArgs a) {
  extern __shared__ double s_acc[];
  const int64_t r0 = (int64_t)blockIdx.x * BLOCK;
  const int64_t r1 = r0 + BLOCK < a.nrows ? r0 + BLOCK : a.nrows;
  const int32_t base = a.row_ptr[r0];
  const int32_t cnt = a.row_ptr[r1] - base;
  for (int t = threadIdx.x; t < cnt; t += BLOCK) s_acc[t] = 0.0;
  __syncthreads();
  const int64_t r = r0 + threadIdx.x;
  if (r < r1) {
    const int32_t off = a.row_ptr[r] - base;
    if (a.fixed[r]) {
      s_acc[off] = 1.0;
      if (a.rhs) a.rhs[r] = a.fixed_vals ? a.fixed_vals[r] : 0.0;
    } else {
      double b = 0.0;
      const int32_t pb = a.rec_ptr[r], pe = a.rec_ptr[r + 1];
      for (int32_t p = pb; p < pe; ++p) {
        const int32_t key = a.rec_key[p];
        uint8_t mt[MS];
        load_meta<NLOC, MS>(a.rec_meta, p, mt);
        const double *row = ke_row<NLOC>(a, key);
#pragma unroll
        for (int j = 0; j < NLOC; ++j) {
          const uint8_t sl = mt[1 + j];
          if (sl != kDrop) s_acc[off + sl] += row[j];
        }
        if (a.fvec) b += a.fvec[key];
      }
      if (a.fixed_vals) {
        for (int32_t p = pb; p < pe; ++p) {
          const int32_t key = a.rec_key[p];
          const int64_t cell = key / NLOC;
          uint8_t mt[MS];
          load_meta<NLOC, MS>(a.rec_meta, p, mt);
          const double *row = ke_row<NLOC>(a, key);
#pragma unroll
          for (int j = 0; j < NLOC; ++j)
            if (mt[1 + j] == kDrop)
              b += row[j] * (-1.0 * a.fixed_vals[a.cell_dofs[cell * NLOC + j]]);
        }
      }
      if (a.rhs) a.rhs[r] = b;
    }
  }
  __syncthreads();
  for (int t = threadIdx.x; t < cnt; t += BLOCK) a.val[base + t] = s_acc[t];
}

template <int NLOC, int MS>
int launch_assemble_ke(hipStream_t s, const lfgpu_plan *p, const KeArgs &a) {
  const size_t lds = (size_t)std::max(p->max_block_nnz, 1) * sizeof(double);
  const unsigned grid = grid_for(a.nrows, p->block);
  switch (p->block) {
    case 256:
      hipLaunchKernelGGL((k_assemble_rows_ke<NLOC, MS, 256>), dim3(grid), dim3(256), lds, s, a);
      break;
    case 128:
      hipLaunchKernelGGL((k_assemble_rows_ke<NLOC, MS, 128>), dim3(grid), dim3(128), lds, s, a);
      break;
    default:
      hipLaunchKernelGGL((k_assemble_rows_ke<NLOC, MS, 64>), dim3(grid), dim3(64), lds, s, a);
      break;
  }
  LFG_HIP(hipGetLastError());
  return LFGPU_OK;
}

}  // namespace

extern "C" int lfgpu_assemble_galerkin_mfma(lfgpu_ctx *ctx, lfgpu_plan *p, const lfgpu_fe *fe,
                                            const lfgpu_coeff *alpha, const lfgpu_coeff *gamma,
                                            const lfgpu_coeff *f, const lfgpu_vec *fixed_vals,
                                            lfgpu_vec *rhs) {
  LFG_REQUIRE(ctx && p && fe, LFGPU_EINVAL, "NULL argument to lfgpu_assemble_galerkin_mfma");
  const lfgpu_mesh *m = p->mesh;
  LFG_REQUIRE(fe->nloc == m->nloc, LFGPU_EINVAL,
              "finite element and DOF layout disagree on the number of local DOFs");
  LFG_REQUIRE(!rhs || rhs->n == m->ndofs, LFGPU_EINVAL, "rhs must have ndofs entries");
  LFG_REQUIRE(!fixed_vals || fixed_vals->n == m->ndofs, LFGPU_EINVAL,
              "fixed_vals must have ndofs entries");
  CoeffD ca, cg, cf;
  LFG_TRY(resolve_coeff(alpha, m->ncells, fe->nqp, false, &ca, "alpha"));
  LFG_TRY(resolve_coeff(gamma, m->ncells, fe->nqp, false, &cg, "gamma"));
  LFG_TRY(resolve_coeff(f, m->ncells, fe->nqp, true, &cf, "f"));
  const int nloc = m->nloc;
  if (!p->ke) LFG_TRY(dalloc(&p->ke, (size_t)std::max<int64_t>(1, m->ncells * nloc * nloc),
                             "element matrices"));
  if (f && !p->fe) LFG_TRY(dalloc(&p->fe, (size_t)std::max<int64_t>(1, m->ncells * nloc),
                                  "element vectors"));
  KeArgs a;
  a.cell_dofs = m->cell_dofs;
  a.rec_ptr = p->rec_ptr;
  a.rec_key = p->rec_key;
  a.rec_meta = p->rec_meta;
  a.fixed = p->fixed;
  a.row_ptr = p->csr.row_ptr;
  a.ke = p->ke;
  a.fvec = f ? p->fe : nullptr;
  a.val = p->csr.val;
  a.rhs = rhs ? rhs->d : nullptr;
  a.fixed_vals = (fixed_vals && p->has_fixed) ? fixed_vals->d : nullptr;
  a.nrows = m->ndofs;
  a.ncells = m->ncells;
  LFG_HIP(hipEventRecord(ctx->ev0, ctx->stream));
  LFG_TRY(element_contract(ctx->stream, m, fe, ca, cg, f ? &cf : nullptr, p->ke, p->fe,
                           nloc == 3 ? ke_rowmajor_i<3>() : ke_rowmajor_i<10>()));
  if (nloc == 3)
    LFG_TRY((launch_assemble_ke<3, 4>(ctx->stream, p, a)));
  else
    LFG_TRY((launch_assemble_ke<10, 12>(ctx->stream, p, a)));
  LFG_HIP(hipEventRecord(ctx->ev1, ctx->stream));
  LFG_HIP(hipEventSynchronize(ctx->ev1));
  float ms = 0.f;
  LFG_HIP(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
  ctx->last_ms = ms;
  csr_values_changed(&p->csr);
  return LFGPU_OK;
}

namespace {

// Arguments of the row-owner assembly for a plan; edge pointers null.
int asm_args(lfgpu_plan *p, const lfgpu_fe *fe, const lfgpu_coeff *alpha, const lfgpu_coeff *gamma,
             const lfgpu_coeff *f, const lfgpu_vec *fixed_vals, lfgpu_vec *rhs, AsmArgs &a) {
  const lfgpu_mesh *m = p->mesh;
  LFG_REQUIRE(fe->nloc == m->nloc, LFGPU_EINVAL,
              "finite element and DOF layout disagree on the number of local DOFs");
  LFG_REQUIRE(!rhs || rhs->n == m->ndofs, LFGPU_EINVAL, "rhs must have ndofs entries");
  LFG_REQUIRE(!fixed_vals || fixed_vals->n == m->ndofs, LFGPU_EINVAL,
              "fixed_vals must have ndofs entries");
  a = AsmArgs{};
  a.xy = m->xy;
  a.cell_v = m->cell_v;
  a.cell_dofs = m->cell_dofs;
  a.rec_ptr = p->rec_ptr;
  a.rec_key = p->rec_key;
  a.rec_meta = p->rec_meta;
  a.fixed = p->fixed;
  a.row_ptr = p->csr.row_ptr;
  a.val = p->csr.val;
  a.rhs = rhs ? rhs->d : nullptr;
  a.fixed_vals = (fixed_vals && p->has_fixed) ? fixed_vals->d : nullptr;
  a.tab = fe->tab;
  LFG_TRY(resolve_coeff(alpha, m->ncells, fe->nqp, false, &a.alpha, "alpha"));
  LFG_TRY(resolve_coeff(gamma, m->ncells, fe->nqp, false, &a.gamma, "gamma"));
  LFG_TRY(resolve_coeff(f, m->ncells, fe->nqp, true, &a.f, "f"));
  a.has_f = f != nullptr;
  a.nrows = m->ndofs;
  return LFGPU_OK;
}

// P3: the per-cell coefficient pass, then the row-owner kernel built for
// the plain coefficient kinds it leaves.  A P3 cell is visited by up to 10
// rows; and without the costly kinds' code the kernel needs 135 registers
// instead of 256 plus 1.5 KB of spills per lane (512^2 P3: 1.32 ms without
// the pass, 1.08 ms with the pass and the generic kernel, 0.83 ms now).
// P1: the generic kernel.  A P1 cell is visited by 3 rows, and the pass
// with the plain kernel (102 instead of 169 registers, twice the resident
// waves) measured slower at 2048^2 (0.89 vs 0.79 ms): the pass's own
// traffic and the extra gathers outweigh the work saved.
template <bool EDGES>
int queue_assembly(hipStream_t s, lfgpu_plan *p, AsmArgs a) {
  if (p->mesh->nloc == 3) return launch_assemble<3, 3, 4, EDGES, false>(s, p, a);
  LFG_TRY(precompute_coeffs<12>(s, p, a));
  return launch_assemble<10, 12, 12, EDGES, true>(s, p, a);
}

template <bool EDGES>
int run_assembly(lfgpu_ctx *ctx, lfgpu_plan *p, const AsmArgs &a) {
  LFG_HIP(hipEventRecord(ctx->ev0, ctx->stream));
  LFG_TRY(queue_assembly<EDGES>(ctx->stream, p, a));
  LFG_HIP(hipEventRecord(ctx->ev1, ctx->stream));
  LFG_HIP(hipEventSynchronize(ctx->ev1));
  float ms = 0.f;
  LFG_HIP(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
  ctx->last_ms = ms;
  csr_values_changed(&p->csr);
  return LFGPU_OK;
}

}  // namespace

namespace lfgpu {
int assemble_enqueue(lfgpu_ctx *ctx, lfgpu_plan *p, const lfgpu_fe *fe,
                     const lfgpu_coeff *alpha, const lfgpu_coeff *gamma, const lfgpu_coeff *f,
                     lfgpu_vec *rhs) {
  AsmArgs a;
  LFG_TRY(asm_args(p, fe, alpha, gamma, f, nullptr, rhs, a));
  return queue_assembly<false>(ctx->stream, p, a);
}
}  // namespace lfgpu

extern "C" int lfgpu_assemble_galerkin(lfgpu_ctx *ctx, lfgpu_plan *p, const lfgpu_fe *fe,
                                       const lfgpu_coeff *alpha, const lfgpu_coeff *gamma,
                                       const lfgpu_coeff *f, const lfgpu_vec *fixed_vals,
                                       lfgpu_vec *rhs) {
  LFG_REQUIRE(ctx && p && fe, LFGPU_EINVAL, "NULL argument to lfgpu_assemble_galerkin");
  AsmArgs a;
  LFG_TRY(asm_args(p, fe, alpha, gamma, f, fixed_vals, rhs, a));
  return run_assembly<false>(ctx, p, a);
}

extern "C" int lfgpu_assemble_galerkin_bd(lfgpu_ctx *ctx, lfgpu_plan *p, const lfgpu_fe *fe,
                                          const lfgpu_coeff *alpha, const lfgpu_coeff *gamma,
                                          const lfgpu_coeff *f, const lfgpu_coeff *eta,
                                          const lfgpu_coeff *h, const lfgpu_vec *fixed_vals,
                                          lfgpu_vec *rhs) {
  LFG_REQUIRE(ctx && p && fe, LFGPU_EINVAL, "NULL argument to lfgpu_assemble_galerkin_bd");
  LFG_REQUIRE(p->edges, LFGPU_EINVAL, "the plan has no edge terms (lfgpu_plan_set_edges)");
  LFG_REQUIRE(!p->edge_mat || eta, LFGPU_EINVAL, "the plan's edge mass term needs eta");
  LFG_REQUIRE(!p->edge_vec || h, LFGPU_EINVAL, "the plan's edge load term needs h");
  AsmArgs a;
  LFG_TRY(asm_args(p, fe, alpha, gamma, f, fixed_vals, rhs, a));
  const lfgpu_edges *E = p->edges;
  a.edge_v = E->edge_v;
  a.edge_dofs = E->edge_dofs;
  a.etab = E->tab;
  if (p->edge_mat) {
    LFG_TRY(resolve_coeff(eta, E->nedges, E->nqe, false, &a.eta, "eta"));
    a.erec_ptr = p->erec_ptr;
    a.erec_key = p->erec_key;
    a.erec_meta = p->erec_meta;
  }
  if (p->edge_vec) {
    LFG_TRY(resolve_coeff(h, E->nedges, E->nqe, false, &a.h, "h"));
    a.vrec_ptr = p->vrec_ptr;
    a.vrec_key = p->vrec_key;
  }
  return run_assembly<true>(ctx, p, a);
}
