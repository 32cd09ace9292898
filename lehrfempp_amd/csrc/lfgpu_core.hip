// liblfgpu core: errors, context, vectors, mesh, reference FE tables, the
// element-level kernels, the device-built assembly plan and the fused
// row-owner assembly.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include <rocprim/rocprim.hpp>

#include "lfgpu_fe_device.h"
#include "lfgpu_internal.h"

// ===========================================================================
// errors and small utilities
// ===========================================================================
namespace lfgpu {

static thread_local std::string g_last_error;

void set_error(const std::string &msg) { g_last_error = msg; }
void clear_error() { g_last_error.clear(); }

int hip_fail(hipError_t e, const char *what) {
  set_error(std::string(what) + ": " + hipGetErrorString(e));
  return e == hipErrorOutOfMemory ? LFGPU_EOOM : LFGPU_EHIP;
}

int dmalloc(void **p, size_t bytes, const char *what) {
  *p = nullptr;
  if (bytes == 0) bytes = 16;
  hipError_t e = hipMalloc(p, bytes);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    *p = nullptr;
    set_error(std::string("device allocation of ") + std::to_string(bytes) + " bytes for " +
              what + " failed: " + hipGetErrorString(e));
    return LFGPU_EOOM;
  }
  return LFGPU_OK;
}

void dfree(void *p) {
  if (p) (void)hipFree(p);
}

int resolve_coeff(const lfgpu_coeff *c, int64_t ncells, int nqp, bool allow_null, CoeffD *out,
                  const char *name) {
  if (!c) {
    LFG_REQUIRE(allow_null, LFGPU_EINVAL, std::string("coefficient ") + name + " is NULL");
    *out = CoeffD{LFGPU_COEFF_CONST, 0.0, nullptr};
    return LFGPU_OK;
  }
  out->kind = c->kind;
  out->value = c->value;
  out->data = nullptr;
  switch (c->kind) {
    case LFGPU_COEFF_CONST:
    case LFGPU_COEFF_SINSIN:
      return LFGPU_OK;
    case LFGPU_COEFF_CELL:
    case LFGPU_COEFF_EXPCELL:
      LFG_REQUIRE(c->data && c->data->n == ncells, LFGPU_EINVAL,
                  std::string("coefficient ") + name + ": per-cell data must have ncells entries");
      out->data = c->data->d;
      return LFGPU_OK;
    case LFGPU_COEFF_QP:
      LFG_REQUIRE(c->data && c->data->n == ncells * nqp, LFGPU_EINVAL,
                  std::string("coefficient ") + name +
                      ": per-QP data must have ncells*nqp entries");
      out->data = c->data->d;
      return LFGPU_OK;
    default:
      set_error(std::string("coefficient ") + name + ": unknown kind");
      return LFGPU_EINVAL;
  }
}

int exclusive_scan_i32(hipStream_t s, const int32_t *in, int32_t *out, int64_t n,
                       int64_t *total) {
  // out[0..n] = exclusive prefix sums, out[n] = total
  size_t tmp_bytes = 0;
  LFG_HIP(rocprim::exclusive_scan(nullptr, tmp_bytes, in, out, 0, (size_t)n,
                                  rocprim::plus<int32_t>(), s));
  void *tmp = nullptr;
  LFG_TRY(dmalloc(&tmp, tmp_bytes, "scan scratch"));
  hipError_t e = rocprim::exclusive_scan(tmp, tmp_bytes, in, out, 0, (size_t)n,
                                         rocprim::plus<int32_t>(), s);
  if (e != hipSuccess) {
    dfree(tmp);
    return hip_fail(e, "rocprim::exclusive_scan");
  }
  int32_t last_in = 0, last_out = 0;
  if (n > 0) {
    e = hipMemcpyAsync(&last_in, in + n - 1, 4, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipMemcpyAsync(&last_out, out + n - 1, 4, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
  }
  dfree(tmp);
  if (e != hipSuccess) return hip_fail(e, "scan total");
  const int64_t tot = (int64_t)last_in + (int64_t)last_out;
  LFG_REQUIRE(tot < 2147483647LL, LFGPU_ERANGE, "prefix sum exceeds int32 range");
  int32_t t32 = (int32_t)tot;
  LFG_HIP(hipMemcpyAsync(out + n, &t32, 4, hipMemcpyHostToDevice, s));
  LFG_HIP(hipStreamSynchronize(s));
  if (total) *total = tot;
  return LFGPU_OK;
}

static inline unsigned grid_for(int64_t n, int block) {
  int64_t g = (n + block - 1) / block;
  if (g < 1) g = 1;
  return (unsigned)g;
}

// RAII-free timer helpers on the context stream.
static int timer_start(lfgpu_ctx *ctx) {
  LFG_HIP(hipEventRecord(ctx->ev0, ctx->stream));
  return LFGPU_OK;
}
static int timer_stop(lfgpu_ctx *ctx) {
  LFG_HIP(hipEventRecord(ctx->ev1, ctx->stream));
  LFG_HIP(hipEventSynchronize(ctx->ev1));
  float ms = 0.f;
  LFG_HIP(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
  ctx->last_ms = ms;
  return LFGPU_OK;
}

}  // namespace lfgpu

using namespace lfgpu;

// ===========================================================================
// context and vectors
// ===========================================================================
extern "C" const char *lfgpu_last_error(void) { return g_last_error.c_str(); }

extern "C" int lfgpu_ctx_create(int device, lfgpu_ctx **out) {
  LFG_REQUIRE(out, LFGPU_EINVAL, "out is NULL");
  *out = nullptr;
  int ndev = 0;
  LFG_HIP(hipGetDeviceCount(&ndev));
  LFG_REQUIRE(device >= 0 && device < ndev, LFGPU_EINVAL, "device index out of range");
  LFG_HIP(hipSetDevice(device));
  LFG_TRY(lfgpu::noalloc_scratch_init(device));
  lfgpu_ctx *c = new lfgpu_ctx();
  c->device = device;
  hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipEventCreate(&c->ev0);
  if (e == hipSuccess) e = hipEventCreate(&c->ev1);
  for (int k = 0; k < 2 && e == hipSuccess; ++k)
    e = hipEventCreateWithFlags(&c->evp[k], hipEventDisableTiming);
  if (e == hipSuccess) e = hipHostMalloc(&c->pinned, 4096, hipHostMallocDefault);
  if (e != hipSuccess) {
    lfgpu_ctx_destroy(c);
    return hip_fail(e, "lfgpu_ctx_create");
  }
  *out = c;
  return LFGPU_OK;
}

extern "C" int lfgpu_ctx_destroy(lfgpu_ctx *c) {
  if (!c) return LFGPU_OK;
  (void)hipSetDevice(c->device);
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  dfree(c->red);
  if (c->pinned) (void)hipHostFree(c->pinned);
  if (c->ev0) (void)hipEventDestroy(c->ev0);
  if (c->ev1) (void)hipEventDestroy(c->ev1);
  for (hipEvent_t ev : c->evp)
    if (ev) (void)hipEventDestroy(ev);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
  return LFGPU_OK;
}

extern "C" int lfgpu_ctx_sync(lfgpu_ctx *c) {
  LFG_REQUIRE(c, LFGPU_EINVAL, "ctx is NULL");
  LFG_HIP(hipStreamSynchronize(c->stream));
  return LFGPU_OK;
}

extern "C" int lfgpu_ctx_last_ms(lfgpu_ctx *c, double *ms) {
  LFG_REQUIRE(c && ms, LFGPU_EINVAL, "NULL argument");
  *ms = c->last_ms;
  return LFGPU_OK;
}

extern "C" int lfgpu_vec_create(lfgpu_ctx *ctx, int64_t n, const double *host_init,
                                lfgpu_vec **out) {
  LFG_REQUIRE(ctx && out && n >= 0, LFGPU_EINVAL, "bad lfgpu_vec_create arguments");
  *out = nullptr;
  lfgpu_vec *v = new lfgpu_vec();
  v->ctx = ctx;
  v->n = n;
  int rc = dalloc(&v->d, (size_t)n, "vector");
  if (rc != LFGPU_OK) {
    delete v;
    return rc;
  }
  hipError_t e;
  if (host_init)
    e = hipMemcpyAsync(v->d, host_init, n * sizeof(double), hipMemcpyHostToDevice, ctx->stream);
  else
    e = hipMemsetAsync(v->d, 0, n * sizeof(double), ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  if (e != hipSuccess) {
    lfgpu_vec_destroy(v);
    return hip_fail(e, "lfgpu_vec_create");
  }
  *out = v;
  return LFGPU_OK;
}

extern "C" int lfgpu_vec_destroy(lfgpu_vec *v) {
  if (!v) return LFGPU_OK;
  if (v->ctx) (void)hipStreamSynchronize(v->ctx->stream);
  dfree(v->d);
  delete v;
  return LFGPU_OK;
}

extern "C" int lfgpu_vec_size(const lfgpu_vec *v, int64_t *n) {
  LFG_REQUIRE(v && n, LFGPU_EINVAL, "NULL argument");
  *n = v->n;
  return LFGPU_OK;
}

extern "C" int lfgpu_vec_upload(lfgpu_vec *v, const double *host, int64_t n) {
  LFG_REQUIRE(v && (host || n == 0), LFGPU_EINVAL, "NULL argument");
  LFG_REQUIRE(n == v->n, LFGPU_EINVAL, "size mismatch in lfgpu_vec_upload");
  LFG_HIP(hipMemcpyAsync(v->d, host, n * sizeof(double), hipMemcpyHostToDevice, v->ctx->stream));
  LFG_HIP(hipStreamSynchronize(v->ctx->stream));
  return LFGPU_OK;
}

extern "C" int lfgpu_vec_download(const lfgpu_vec *v, double *host, int64_t n) {
  LFG_REQUIRE(v && (host || n == 0), LFGPU_EINVAL, "NULL argument");
  LFG_REQUIRE(n == v->n, LFGPU_EINVAL, "size mismatch in lfgpu_vec_download");
  LFG_HIP(hipMemcpyAsync(host, v->d, n * sizeof(double), hipMemcpyDeviceToHost, v->ctx->stream));
  LFG_HIP(hipStreamSynchronize(v->ctx->stream));
  return LFGPU_OK;
}

__global__ void k_fill(double *x, int64_t n, double v) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) x[i] = v;
}

extern "C" int lfgpu_vec_fill(lfgpu_vec *v, double value) {
  LFG_REQUIRE(v, LFGPU_EINVAL, "NULL argument");
  if (v->n == 0) return LFGPU_OK;
  hipLaunchKernelGGL(k_fill, dim3(grid_for(v->n, 256)), dim3(256), 0, v->ctx->stream, v->d, v->n,
                     value);
  LFG_HIP(hipGetLastError());
  return LFGPU_OK;
}

// y = x with 16-byte accesses: the reference point for HBM byte counters
// (n doubles read, n written).
__global__ void k_copy(const double2 *__restrict__ x, double2 *__restrict__ y, int64_t n2) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n2) y[i] = x[i];
}

// y = x with non-temporal loads and stores: the calibration kernel of the HBM
// byte counters for streams that bypass the caches
__global__ void k_copy_nt(const double *__restrict__ x, double *__restrict__ y, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) __builtin_nontemporal_store(__builtin_nontemporal_load(&x[i]), &y[i]);
}

extern "C" int lfgpu_vec_copy_nt(lfgpu_ctx *ctx, const lfgpu_vec *x, lfgpu_vec *y) {
  LFG_REQUIRE(ctx && x && y && x->n == y->n, LFGPU_EINVAL, "bad lfgpu_vec_copy_nt arguments");
  LFG_REQUIRE(x->d != y->d, LFGPU_EINVAL, "lfgpu_vec_copy_nt needs two different vectors");
  if (x->n == 0) return LFGPU_OK;
  hipLaunchKernelGGL(k_copy_nt, dim3(grid_for(x->n, 256)), dim3(256), 0, ctx->stream, x->d, y->d,
                     x->n);
  LFG_HIP(hipGetLastError());
  return LFGPU_OK;
}

extern "C" int lfgpu_vec_copy(lfgpu_ctx *ctx, const lfgpu_vec *x, lfgpu_vec *y) {
  LFG_REQUIRE(ctx && x && y && x->n == y->n, LFGPU_EINVAL, "bad lfgpu_vec_copy arguments");
  LFG_REQUIRE(x->d != y->d, LFGPU_EINVAL, "lfgpu_vec_copy needs two different vectors");
  if (x->n == 0) return LFGPU_OK;
  const int64_t n2 = x->n / 2;
  if (n2 > 0)
    hipLaunchKernelGGL(k_copy, dim3(grid_for(n2, 256)), dim3(256), 0, ctx->stream,
                       reinterpret_cast<const double2 *>(x->d), reinterpret_cast<double2 *>(y->d),
                       n2);
  if (x->n & 1)
    LFG_HIP(hipMemcpyAsync(y->d + x->n - 1, x->d + x->n - 1, sizeof(double),
                           hipMemcpyDeviceToDevice, ctx->stream));
  LFG_HIP(hipGetLastError());
  return LFGPU_OK;
}

// ===========================================================================
// host topology helpers
// ===========================================================================
extern "C" int lfgpu_tp_triag_mesh(int32_t nx, int32_t ny, double x0, double y0, double x1,
                                   double y1, double *xy, int32_t *cell_v, int32_t *edge_v) {
  LFG_REQUIRE(nx > 0 && ny > 0 && xy && cell_v, LFGPU_EINVAL, "bad mesh builder arguments");
  LFG_REQUIRE(x1 > x0 && y1 > y0, LFGPU_EINVAL, "empty rectangle");
  LFG_REQUIRE((int64_t)(nx + 1) * (ny + 1) < 2147483647LL / 2, LFGPU_ERANGE, "mesh too large");
  // tp_triag_mesh_builder.cc:40-61: h = size / n, node (i,j) at corner + i h
  const double hx = (x1 - x0) / nx, hy = (y1 - y0) / ny;
  const int64_t stride = (int64_t)nx + 1;
  for (int64_t j = 0; j <= ny; ++j)
    for (int64_t i = 0; i <= nx; ++i) {
      const int64_t v = i + j * stride;
      xy[2 * v] = x0 + (double)(uint32_t)i * hx;
      xy[2 * v + 1] = y0 + (double)(uint32_t)j * hy;
    }
  // cells in builder order (:140-176): per (i, j) the "up" then the "low" one
  int64_t c = 0;
  for (int64_t i = 0; i < nx; ++i)
    for (int64_t j = 0; j < ny; ++j, c += 2) {
      const int32_t a = (int32_t)(i + j * stride), b = (int32_t)(i + 1 + j * stride);
      const int32_t d = (int32_t)(i + (j + 1) * stride), e = (int32_t)(i + 1 + (j + 1) * stride);
      int32_t *u = cell_v + 3 * c;
      u[0] = a; u[1] = e; u[2] = d;
      int32_t *l = u + 3;
      l[0] = a; l[1] = b; l[2] = e;
    }
  if (edge_v) {
    // horizontal (:68-88), vertical (:90-110), diagonal (:112-133)
    int64_t k = 0;
    for (int64_t i = 0; i < nx; ++i)
      for (int64_t j = 0; j <= ny; ++j, ++k) {
        edge_v[2 * k] = (int32_t)(i + j * stride);
        edge_v[2 * k + 1] = (int32_t)(i + 1 + j * stride);
      }
    for (int64_t i = 0; i <= nx; ++i)
      for (int64_t j = 0; j < ny; ++j, ++k) {
        edge_v[2 * k] = (int32_t)(i + j * stride);
        edge_v[2 * k + 1] = (int32_t)(i + (j + 1) * stride);
      }
    for (int64_t i = 0; i < nx; ++i)
      for (int64_t j = 0; j < ny; ++j, ++k) {
        edge_v[2 * k] = (int32_t)(i + j * stride);
        edge_v[2 * k + 1] = (int32_t)(i + 1 + (j + 1) * stride);
      }
  }
  return LFGPU_OK;
}

extern "C" int lfgpu_mesh_edges(int64_t nv, int64_t ncells, const int32_t *cell_v, int64_t npre,
                                const int32_t *pre_edge_v, int64_t max_edges,
                                int32_t *edge_v_out, int32_t *cell_e, int8_t *cell_ori,
                                int64_t *nedges) {
  LFG_REQUIRE(cell_v && edge_v_out && cell_e && cell_ori && nedges, LFGPU_EINVAL,
              "NULL argument");
  LFG_REQUIRE(npre == 0 || pre_edge_v, LFGPU_EINVAL, "predefined edges missing");
  // Sorted key (min, max) -> slot.  Predefined edges keep their position
  // (mesh.cc:236-274); edges first met in a cell take the orientation of
  // that cell's local edge and are numbered after all predefined edges in
  // key order (mesh.cc:517-575).
  struct Info {
    int32_t a, b;
    int64_t index;
  };
  std::vector<std::pair<uint64_t, Info>> items;
  items.reserve((size_t)(npre + 3 * ncells));
  auto key = [](int32_t a, int32_t b) {
    const uint32_t lo = (uint32_t)std::min(a, b), hi = (uint32_t)std::max(a, b);
    return ((uint64_t)lo << 32) | hi;
  };
  for (int64_t e = 0; e < npre; ++e) {
    const int32_t a = pre_edge_v[2 * e], b = pre_edge_v[2 * e + 1];
    LFG_REQUIRE(a >= 0 && b >= 0 && a < nv && b < nv && a != b, LFGPU_ERANGE,
                "invalid predefined edge");
    items.push_back({key(a, b), Info{a, b, e}});
  }
  for (int64_t c = 0; c < ncells; ++c)
    for (int k = 0; k < 3; ++k) {
      const int32_t a = cell_v[3 * c + k], b = cell_v[3 * c + (k + 1) % 3];
      LFG_REQUIRE(a >= 0 && b >= 0 && a < nv && b < nv && a != b, LFGPU_ERANGE,
                  "invalid cell vertex index");
      // order tag: predefined edges first, then cells in order -> the
      // stable sort keeps the first registration at the front of each key
      items.push_back({key(a, b), Info{a, b, -1}});
    }
  std::stable_sort(items.begin(), items.end(),
                   [](const std::pair<uint64_t, Info> &x, const std::pair<uint64_t, Info> &y) {
                     return x.first < y.first;
                   });
  std::vector<std::pair<uint64_t, Info>> uniq;
  uniq.reserve(items.size() / 2 + 1);
  for (size_t s = 0; s < items.size(); ++s)
    if (uniq.empty() || uniq.back().first != items[s].first) uniq.push_back(items[s]);
  int64_t next = npre;
  for (auto &u : uniq)
    if (u.second.index < 0) u.second.index = next++;
  LFG_REQUIRE(next <= max_edges, LFGPU_ERANGE, "edge capacity too small");
  for (auto &u : uniq) {
    edge_v_out[2 * u.second.index] = u.second.a;
    edge_v_out[2 * u.second.index + 1] = u.second.b;
  }
  for (int64_t c = 0; c < ncells; ++c)
    for (int k = 0; k < 3; ++k) {
      const int32_t a = cell_v[3 * c + k], b = cell_v[3 * c + (k + 1) % 3];
      const uint64_t kk = key(a, b);
      auto it = std::lower_bound(
          uniq.begin(), uniq.end(), kk,
          [](const std::pair<uint64_t, Info> &x, uint64_t v) { return x.first < v; });
      cell_e[3 * c + k] = (int32_t)it->second.index;
      cell_ori[3 * c + k] = (it->second.a == a) ? 1 : -1;
    }
  *nedges = next;
  return LFGPU_OK;
}

extern "C" int lfgpu_dof_table(int64_t nv, int64_t ncells, int64_t nedges, const int32_t *cell_v,
                               const int32_t *cell_e, const int8_t *cell_ori, int ndof_edge,
                               int ndof_cell, int64_t *cell_dofs, int64_t *ndofs) {
  LFG_REQUIRE(cell_v && cell_dofs && ndofs && ndof_edge >= 0 && ndof_cell >= 0, LFGPU_EINVAL,
              "bad lfgpu_dof_table arguments");
  LFG_REQUIRE(ndof_edge == 0 || (cell_e && cell_ori), LFGPU_EINVAL, "edge tables missing");
  const int nloc = 3 + 3 * ndof_edge + ndof_cell;
  // dofhandler.cc:142-284: node DOFs first, then edge interiors in edge
  // order, then cell interiors; per cell: vertices, edges (reversed when the
  // relative orientation is negative), interior.
  const int64_t eb = nv, cb = nv + nedges * ndof_edge;
  for (int64_t c = 0; c < ncells; ++c) {
    int64_t *d = cell_dofs + c * nloc;
    int p = 0;
    for (int k = 0; k < 3; ++k) d[p++] = cell_v[3 * c + k];
    for (int k = 0; k < 3 && ndof_edge > 0; ++k) {
      const int64_t off = eb + (int64_t)cell_e[3 * c + k] * ndof_edge;
      if (cell_ori[3 * c + k] > 0)
        for (int j = 0; j < ndof_edge; ++j) d[p++] = off + j;
      else
        for (int j = ndof_edge - 1; j >= 0; --j) d[p++] = off + j;
    }
    for (int j = 0; j < ndof_cell; ++j) d[p++] = cb + c * ndof_cell + j;
  }
  *ndofs = cb + ncells * ndof_cell;
  return LFGPU_OK;
}

// ===========================================================================
// mesh upload
// ===========================================================================
__global__ void k_mesh_check(const double *xy, const int32_t *cell_v, int64_t ncells, int64_t nv,
                             int *bad_cell) {
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncells) return;
  const int32_t v0 = cell_v[3 * c], v1 = cell_v[3 * c + 1], v2 = cell_v[3 * c + 2];
  bool ok = v0 >= 0 && v1 >= 0 && v2 >= 0 && v0 < nv && v1 < nv && v2 < nv;
  if (ok) {
    ok = tria_ok(xy[2 * v0], xy[2 * v0 + 1], xy[2 * v1], xy[2 * v1 + 1], xy[2 * v2],
                 xy[2 * v2 + 1]);
  }
  if (!ok) atomicMin(bad_cell, (int)c);
}

__global__ void k_narrow_dofs(const int64_t *in, int32_t *out, int64_t n, int64_t ndofs,
                              int *bad) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const int64_t d = in[t];
  if (d < 0 || d >= ndofs) {
    atomicMin(bad, (int)(t >> 4));
    out[t] = 0;
  } else {
    out[t] = (int32_t)d;
  }
}

extern "C" int lfgpu_mesh_destroy(lfgpu_mesh *m) {
  if (!m) return LFGPU_OK;
  if (m->ctx) (void)hipStreamSynchronize(m->ctx->stream);
  if (!m->dofs_alias) dfree(m->cell_dofs);
  dfree(m->cell_v);
  dfree(m->xy);
  delete m;
  return LFGPU_OK;
}

extern "C" int lfgpu_mesh_create(lfgpu_ctx *ctx, int64_t nv, const double *xy, int64_t ncells,
                                 const int32_t *cell_v, int nloc, const int64_t *cell_dofs,
                                 int64_t ndofs, lfgpu_mesh **out) {
  LFG_REQUIRE(ctx && out && xy && cell_v, LFGPU_EINVAL, "NULL argument to lfgpu_mesh_create");
  *out = nullptr;
  LFG_REQUIRE(nv > 0 && ncells > 0, LFGPU_EINVAL, "empty mesh");
  LFG_REQUIRE(nloc == 3 || nloc == 10, LFGPU_ENORULE,
              "only P1 (nloc=3) and P3 (nloc=10) Lagrange triangles are supported");
  LFG_REQUIRE(nloc == 3 || cell_dofs, LFGPU_EINVAL, "P3 needs a cell DOF table");
  LFG_REQUIRE(ndofs > 0 && ndofs < 2147483647LL && ncells * nloc < 2147483647LL, LFGPU_ERANGE,
              "problem exceeds the int32 device index range");
  if (!cell_dofs) LFG_REQUIRE(ndofs == nv, LFGPU_EINVAL, "P1 without DOF table needs ndofs == nv");
  lfgpu_mesh *m = new lfgpu_mesh();
  m->ctx = ctx;
  m->nv = nv;
  m->ncells = ncells;
  m->ndofs = ndofs;
  m->nloc = nloc;
  hipStream_t s = ctx->stream;
  int rc = dalloc(&m->xy, (size_t)(2 * nv), "vertex coordinates");
  if (rc == LFGPU_OK) rc = dalloc(&m->cell_v, (size_t)(3 * ncells), "cell vertices");
  if (rc != LFGPU_OK) {
    lfgpu_mesh_destroy(m);
    return rc;
  }
  hipError_t e = hipMemcpyAsync(m->xy, xy, 2 * nv * sizeof(double), hipMemcpyHostToDevice, s);
  if (e == hipSuccess)
    e = hipMemcpyAsync(m->cell_v, cell_v, 3 * ncells * sizeof(int32_t), hipMemcpyHostToDevice, s);
  int *dbad = nullptr;
  if (e == hipSuccess) {
    rc = dalloc(&dbad, 2, "status word");
    if (rc != LFGPU_OK) {
      lfgpu_mesh_destroy(m);
      return rc;
    }
  }
  int hbad[2] = {2147483647, 2147483647};
  if (e == hipSuccess) e = hipMemcpyAsync(dbad, hbad, 8, hipMemcpyHostToDevice, s);
  if (e == hipSuccess) {
    hipLaunchKernelGGL(k_mesh_check, dim3(grid_for(ncells, 256)), dim3(256), 0, s, m->xy,
                       m->cell_v, ncells, nv, dbad);
    e = hipGetLastError();
  }
  if (e == hipSuccess && cell_dofs) {
    int64_t *tmp = nullptr;
    rc = dalloc(&tmp, (size_t)(ncells * nloc), "DOF staging");
    if (rc == LFGPU_OK) rc = dalloc(&m->cell_dofs, (size_t)(ncells * nloc), "cell DOFs");
    if (rc != LFGPU_OK) {
      dfree(tmp);
      dfree(dbad);
      lfgpu_mesh_destroy(m);
      return rc;
    }
    e = hipMemcpyAsync(tmp, cell_dofs, ncells * nloc * sizeof(int64_t), hipMemcpyHostToDevice, s);
    if (e == hipSuccess) {
      hipLaunchKernelGGL(k_narrow_dofs, dim3(grid_for(ncells * nloc, 256)), dim3(256), 0, s, tmp,
                         m->cell_dofs, ncells * nloc, ndofs, dbad + 1);
      e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    dfree(tmp);
  } else {
    m->cell_dofs = m->cell_v;
    m->dofs_alias = true;
  }
  if (e == hipSuccess) e = hipMemcpyAsync(hbad, dbad, 8, hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  dfree(dbad);
  if (e != hipSuccess) {
    lfgpu_mesh_destroy(m);
    return hip_fail(e, "lfgpu_mesh_create");
  }
  if (hbad[0] != 2147483647) {
    lfgpu_mesh_destroy(m);
    set_error("degenerate or invalid triangle: cell " + std::to_string(hbad[0]) +
              " (assertNonDegenerateTriangle, tol 1e-8)");
    return LFGPU_EDEGENERATE;
  }
  if (hbad[1] != 2147483647) {
    lfgpu_mesh_destroy(m);
    set_error("cell DOF index out of [0, ndofs) near cell " + std::to_string(hbad[1] * 16 / nloc));
    return LFGPU_ERANGE;
  }
  *out = m;
  return LFGPU_OK;
}

// ===========================================================================
// reference FE tables (host)
// ===========================================================================
namespace {

// quad/quad_rules_tria.cc:35-51 (degree 2) and :113-156 (degree 6); these
// literals are data and must match bit for bit.
const double kRule2Pts[6] = {0.1666666666666666019,  0.16666666666666679619,
                             0.66666666666666651864, 0.16666666666666665741,
                             0.1666666666666668517,  0.66666666666666674068};
const double kRule2W[3] = {0.16666666666666665741, 0.16666666666666665741,
                           0.16666666666666665741};
const double kRule6Pts[24] = {
    0.21942998254978290928,  0.21942998254978310357,  0.56114003490043395939,
    0.2194299825497829648,   0.21942998254978302031,  0.56114003490043418143,
    0.48013796411221504323,  0.039724071775569962106, 0.48013796411221509874,
    0.48013796411221504323,  0.039724071775569885778, 0.48013796411221526528,
    0.14161901592396813898,  0.019371724361240811457, 0.8390092597147910114,
    0.14161901592396811123,  0.019371724361240960643, 0.83900925971479112242,
    0.83900925971479112242,  0.019371724361240811457, 0.14161901592396813898,
    0.83900925971479123344,  0.019371724361240558188, 0.14161901592396844429};
const double kRule6W[12] = {
    0.085666562076490496147, 0.085666562076490496147, 0.085666562076490496147,
    0.040365544796515481774, 0.040365544796515481774, 0.040365544796515481774,
    0.020317279896830325869, 0.020317279896830325869, 0.020317279896830325869,
    0.020317279896830325869, 0.020317279896830325869, 0.020317279896830325869};

// FeLagrangeO1Tria values/gradients (uscalfe/lagr_fe.h:104-129).
void lagrange_o1(double a, double b, double *N, double *gx, double *gy) {
  N[0] = 1.0 - a - b;
  N[1] = a;
  N[2] = b;
  gx[0] = -1.0; gy[0] = -1.0;
  gx[1] = 1.0;  gy[1] = 0.0;
  gx[2] = 0.0;  gy[2] = 1.0;
}

// FeLagrangeO3Tria values/gradients (uscalfe/lagr_fe.h:1010-1101), written
// with the same operand order so the tables round identically.
void lagrange_o3(double a, double b, double *N, double *gx, double *gy) {
  const double t1 = 1 / 3.0, t2 = 2 / 3.0;
  const double l1 = a, l2 = b, l0 = 1 - a - b;
  N[0] = 4.5 * l0 * (l0 - t1) * (l0 - t2);
  N[1] = 4.5 * l1 * (l1 - t1) * (l1 - t2);
  N[2] = 4.5 * l2 * (l2 - t1) * (l2 - t2);
  N[3] = 13.5 * l0 * l1 * (l0 - t1);
  N[4] = 13.5 * l0 * l1 * (l1 - t1);
  N[5] = 13.5 * l1 * l2 * (l1 - t1);
  N[6] = 13.5 * l1 * l2 * (l2 - t1);
  N[7] = 13.5 * l2 * l0 * (l2 - t1);
  N[8] = 13.5 * l2 * l0 * (l0 - t1);
  N[9] = 27.0 * l0 * l1 * l2;
  const double q0 = (l0 - t1) * (l0 - t2) + l0 * (l0 - t2) + l0 * (l0 - t1);
  gx[0] = -4.5 * q0;
  gy[0] = -4.5 * q0;
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

}  // namespace

extern "C" int lfgpu_fe_destroy(lfgpu_fe *fe) {
  if (!fe) return LFGPU_OK;
  if (fe->ctx) (void)hipStreamSynchronize(fe->ctx->stream);
  dfree(fe->tab);
  dfree(fe->rten);
  delete fe;
  return LFGPU_OK;
}

extern "C" int lfgpu_fe_create(lfgpu_ctx *ctx, int degree, int nqp, const double *pts,
                               const double *w, lfgpu_fe **out) {
  LFG_REQUIRE(ctx && out, LFGPU_EINVAL, "NULL argument to lfgpu_fe_create");
  *out = nullptr;
  LFG_REQUIRE(degree == 1 || degree == 3, LFGPU_ENORULE,
              "No local shape function information for this degree (only 1 and 3)");
  const int nloc = degree == 1 ? 3 : 10;
  std::vector<double> P, W;
  if (!pts || !w) {
    // make_QuadRule(kTria, 2 * degree)
    if (degree == 1) {
      P.assign(kRule2Pts, kRule2Pts + 6);
      W.assign(kRule2W, kRule2W + 3);
    } else {
      P.assign(kRule6Pts, kRule6Pts + 24);
      W.assign(kRule6W, kRule6W + 12);
    }
  } else {
    LFG_REQUIRE(nqp > 0, LFGPU_ENORULE, "empty quadrature rule");
    P.assign(pts, pts + 2 * nqp);
    W.assign(w, w + nqp);
  }
  const int q = (int)W.size();
  // kernels are instantiated for the default point counts only
  LFG_REQUIRE((degree == 1 && q == 3) || (degree == 3 && q == 12), LFGPU_ENORULE,
              "quadrature rule size not supported by the kernels (P1: 3, P3: 12 points)");
  std::vector<double> tab(q + 2 * q + nloc * q + 2 * nloc * q);
  for (int k = 0; k < q; ++k) tab[k] = W[k];
  for (int k = 0; k < 2 * q; ++k) tab[q + k] = P[k];
  double *Nt = tab.data() + 3 * q, *dN = Nt + nloc * q;
  for (int k = 0; k < q; ++k) {
    double N[10], gx[10], gy[10];
    if (degree == 1)
      lagrange_o1(P[2 * k], P[2 * k + 1], N, gx, gy);
    else
      lagrange_o3(P[2 * k], P[2 * k + 1], N, gx, gy);
    for (int i = 0; i < nloc; ++i) {
      Nt[i + nloc * k] = N[i];
      dN[i + nloc * (2 * k)] = gx[i];
      dN[i + nloc * (2 * k + 1)] = gy[i];
    }
  }
  // Reference tensors for the contraction form of the element matrix.  With
  // B = (J^T J)^{-1} = M^T M, M = J^{-T}, the integrand of
  // loc_comp_ellbvp.h:324-337 is
  //   alpha (B00 dx_i dx_j + B11 dy_i dy_j + B01 (dx_i dy_j + dy_i dx_j))
  //   + gamma N_i N_j,
  // so K_e = sum over (point, term) of a cell coefficient times a fixed
  // reference matrix.
  const int ne = nloc * nloc, nep = (ne + 15) / 16 * 16;
  std::vector<double> rten((size_t)(4 + 4 * q) * nep, 0.0);
  for (int k = 0; k < q; ++k)
    for (int i = 0; i < nloc; ++i)
      for (int j = 0; j < nloc; ++j) {
        const double w = W[k];
        const double dxi = dN[i + nloc * (2 * k)], dyi = dN[i + nloc * (2 * k + 1)];
        const double dxj = dN[j + nloc * (2 * k)], dyj = dN[j + nloc * (2 * k + 1)];
        const double t[4] = {w * (dxi * dxj), w * (dyi * dyj), w * (dxi * dyj + dyi * dxj),
                             w * (Nt[i + nloc * k] * Nt[j + nloc * k])};
        for (int c = 0; c < 4; ++c) {
          rten[(size_t)(4 + 4 * k + c) * nep + i * nloc + j] = t[c];
          rten[(size_t)c * nep + i * nloc + j] += t[c];
        }
      }
  lfgpu_fe *fe = new lfgpu_fe();
  fe->ctx = ctx;
  fe->degree = degree;
  fe->nloc = nloc;
  fe->nqp = q;
  fe->tab_len = (int)tab.size();
  fe->nep = nep;
  int rc = dalloc(&fe->tab, tab.size(), "FE tables");
  if (rc == LFGPU_OK) rc = dalloc(&fe->rten, rten.size(), "reference tensors");
  if (rc != LFGPU_OK) {
    lfgpu_fe_destroy(fe);
    return rc;
  }
  hipError_t e = hipMemcpyAsync(fe->tab, tab.data(), tab.size() * sizeof(double),
                                hipMemcpyHostToDevice, ctx->stream);
  if (e == hipSuccess)
    e = hipMemcpyAsync(fe->rten, rten.data(), rten.size() * sizeof(double),
                       hipMemcpyHostToDevice, ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  if (e != hipSuccess) {
    lfgpu_fe_destroy(fe);
    return hip_fail(e, "lfgpu_fe_create");
  }
  *out = fe;
  return LFGPU_OK;
}

extern "C" int lfgpu_fe_info(const lfgpu_fe *fe, int *nloc, int *nqp) {
  LFG_REQUIRE(fe, LFGPU_EINVAL, "NULL fe");
  if (nloc) *nloc = fe->nloc;
  if (nqp) *nqp = fe->nqp;
  return LFGPU_OK;
}

// ===========================================================================
// element-level kernels
// ===========================================================================
__global__ void k_geometry(const double *xy, const int32_t *cell_v, int64_t ncells, double *J,
                           double *JinvT, double *det) {
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncells) return;
  const CellXY p = load_cell(xy, cell_v, c);
  Geo g;
  double Jl[4];
  tria_geo(p.x0, p.y0, p.x1, p.y1, p.x2, p.y2, g, Jl);
  for (int q = 0; q < 4; ++q) J[4 * c + q] = Jl[q];
  JinvT[4 * c + 0] = g.jit00;
  JinvT[4 * c + 1] = g.jit01;
  JinvT[4 * c + 2] = g.jit10;
  JinvT[4 * c + 3] = g.jit11;
  det[c] = g.det;
}

extern "C" int lfgpu_geometry(lfgpu_ctx *ctx, const lfgpu_mesh *m, lfgpu_vec *J, lfgpu_vec *JinvT,
                              lfgpu_vec *det) {
  LFG_REQUIRE(ctx && m && J && JinvT && det, LFGPU_EINVAL, "NULL argument to lfgpu_geometry");
  LFG_REQUIRE(J->n == 4 * m->ncells && JinvT->n == 4 * m->ncells && det->n == m->ncells,
              LFGPU_EINVAL, "geometry output sizes must be 4*ncells, 4*ncells, ncells");
  hipLaunchKernelGGL(k_geometry, dim3(grid_for(m->ncells, 256)), dim3(256), 0, ctx->stream, m->xy,
                     m->cell_v, m->ncells, J->d, JinvT->d, det->d);
  LFG_HIP(hipGetLastError());
  return LFGPU_OK;
}

// One thread per (cell, local row i): row i of the element matrix.
template <int NLOC, int NQP>
__global__ __launch_bounds__(256) void k_element_matrices(const double *xy, const int32_t *cell_v,
                                                          int64_t ncells, const double *tab,
                                                          CoeffD alpha, CoeffD gamma,
                                                          double *out) {
  __shared__ double s_tab[Tab<NLOC, NQP>::kLen];
  for (int t = threadIdx.x; t < Tab<NLOC, NQP>::kLen; t += blockDim.x) s_tab[t] = tab[t];
  __syncthreads();
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= ncells * NLOC) return;
  const int64_t cell = t / NLOC;
  const int i = (int)(t - cell * NLOC);
  const Tab<NLOC, NQP> T{s_tab};
  const CellXY p = load_cell(xy, cell_v, cell);
  Geo g;
  tria_geo(p.x0, p.y0, p.x1, p.y1, p.x2, p.y2, g);
  double ak[NQP], gk[NQP];
  coeff_qp<NQP>(alpha, cell, p, s_tab, ak);
  coeff_qp<NQP>(gamma, cell, p, s_tab, gk);
  double row[NLOC];
  element_row<NLOC, NQP>(T, g, ak, gk, i, row);
  double *o = out + t * NLOC;
#pragma unroll
  for (int j = 0; j < NLOC; ++j) o[j] = row[j];
}

template <int NLOC, int NQP>
__global__ __launch_bounds__(256) void k_element_vectors(const double *xy, const int32_t *cell_v,
                                                         int64_t ncells, const double *tab,
                                                         CoeffD f, double *out) {
  __shared__ double s_tab[Tab<NLOC, NQP>::kLen];
  for (int t = threadIdx.x; t < Tab<NLOC, NQP>::kLen; t += blockDim.x) s_tab[t] = tab[t];
  __syncthreads();
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= ncells * NLOC) return;
  const int64_t cell = t / NLOC;
  const int i = (int)(t - cell * NLOC);
  const Tab<NLOC, NQP> T{s_tab};
  const CellXY p = load_cell(xy, cell_v, cell);
  Geo g;
  tria_geo(p.x0, p.y0, p.x1, p.y1, p.x2, p.y2, g);
  double fk[NQP];
  coeff_qp<NQP>(f, cell, p, s_tab, fk);
  out[t] = element_load<NLOC, NQP>(T, g, fk, i);
}

static int check_fe_mesh(const lfgpu_mesh *m, const lfgpu_fe *fe) {
  LFG_REQUIRE(m && fe, LFGPU_EINVAL, "NULL mesh or fe");
  LFG_REQUIRE(fe->nloc == m->nloc, LFGPU_EINVAL,
              "finite element and DOF layout disagree on the number of local DOFs");
  return LFGPU_OK;
}

extern "C" int lfgpu_element_matrices(lfgpu_ctx *ctx, const lfgpu_mesh *m, const lfgpu_fe *fe,
                                      const lfgpu_coeff *alpha, const lfgpu_coeff *gamma,
                                      lfgpu_vec *out) {
  LFG_REQUIRE(ctx && out, LFGPU_EINVAL, "NULL argument to lfgpu_element_matrices");
  LFG_TRY(check_fe_mesh(m, fe));
  LFG_REQUIRE(out->n == m->ncells * m->nloc * m->nloc, LFGPU_EINVAL,
              "output must hold ncells*nloc*nloc values");
  CoeffD a, g;
  LFG_TRY(resolve_coeff(alpha, m->ncells, fe->nqp, false, &a, "alpha"));
  LFG_TRY(resolve_coeff(gamma, m->ncells, fe->nqp, false, &g, "gamma"));
  const int64_t nt = m->ncells * m->nloc;
  if (fe->nloc == 3)
    hipLaunchKernelGGL((k_element_matrices<3, 3>), dim3(grid_for(nt, 256)), dim3(256), 0,
                       ctx->stream, m->xy, m->cell_v, m->ncells, fe->tab, a, g, out->d);
  else
    hipLaunchKernelGGL((k_element_matrices<10, 12>), dim3(grid_for(nt, 256)), dim3(256), 0,
                       ctx->stream, m->xy, m->cell_v, m->ncells, fe->tab, a, g, out->d);
  LFG_HIP(hipGetLastError());
  return LFGPU_OK;
}

extern "C" int lfgpu_element_vectors(lfgpu_ctx *ctx, const lfgpu_mesh *m, const lfgpu_fe *fe,
                                     const lfgpu_coeff *f, lfgpu_vec *out) {
  LFG_REQUIRE(ctx && out, LFGPU_EINVAL, "NULL argument to lfgpu_element_vectors");
  LFG_TRY(check_fe_mesh(m, fe));
  LFG_REQUIRE(out->n == m->ncells * m->nloc, LFGPU_EINVAL, "output must hold ncells*nloc values");
  CoeffD fc;
  LFG_TRY(resolve_coeff(f, m->ncells, fe->nqp, false, &fc, "f"));
  const int64_t nt = m->ncells * m->nloc;
  if (fe->nloc == 3)
    hipLaunchKernelGGL((k_element_vectors<3, 3>), dim3(grid_for(nt, 256)), dim3(256), 0,
                       ctx->stream, m->xy, m->cell_v, m->ncells, fe->tab, fc, out->d);
  else
    hipLaunchKernelGGL((k_element_vectors<10, 12>), dim3(grid_for(nt, 256)), dim3(256), 0,
                       ctx->stream, m->xy, m->cell_v, m->ncells, fe->tab, fc, out->d);
  LFG_HIP(hipGetLastError());
  return LFGPU_OK;
}
