"""ctypes binding of the liblfgpu C ABI (include/lfgpu.h).

This is the Python side of the boundary used by the test suite and the
benchmark; C++ callers include lehrfempp_amd/include/lf_amd/lf_amd.h instead.
Loading fails loudly when the HIP library has not been built: there is no
CPU fallback anywhere in the product.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liblfgpu.so")

OK, EINVAL, EDEGENERATE, ENOCONV, EOOM, EHIP, ENORULE, ERANGE = range(8)
COEFF_CONST, COEFF_CELL, COEFF_QP, COEFF_SINSIN, COEFF_EXPCELL = range(5)
PCG_FP64, PCG_MIXED, PCG_MIXED_MATRIX = 0, 1, 2
STORAGE_ELL, STORAGE_DIA, STORAGE_BLOCK_DIA, STORAGE_ELL16 = 0, 1, 2, 3

_P = C.c_void_p
_d = C.POINTER(C.c_double)
_i32 = C.POINTER(C.c_int32)
_i64 = C.POINTER(C.c_int64)
_u8 = C.POINTER(C.c_uint8)
_i8 = C.POINTER(C.c_int8)


class Coeff(C.Structure):
    _fields_ = [("kind", C.c_int), ("value", C.c_double), ("data", C.c_void_p)]


class PcgOpts(C.Structure):
    _fields_ = [("rtol", C.c_double), ("maxit", C.c_int), ("mode", C.c_int),
                ("check_every", C.c_int), ("refine_delta", C.c_double),
                ("refine_max_interval", C.c_int), ("single_reduction", C.c_int),
                ("graph", C.c_int)]


class PcgInfo(C.Structure):
    _fields_ = [("iters", C.c_int), ("converged", C.c_int), ("refinements", C.c_int),
                ("relres", C.c_double), ("ms", C.c_double), ("syncs", C.c_int)]


class StrangInfo(C.Structure):
    _fields_ = [("diffusion_steps", C.c_int), ("solves", C.c_int), ("iters", C.c_int64),
                ("max_iters", C.c_int), ("converged", C.c_int), ("worst_relres", C.c_double),
                ("ms", C.c_double)]


# name -> (restype, argtypes); the list is also the export check of the tests
SIGNATURES = {
    "lfgpu_last_error": (C.c_char_p, []),
    "lfgpu_ctx_create": (C.c_int, [C.c_int, C.POINTER(_P)]),
    "lfgpu_ctx_destroy": (C.c_int, [_P]),
    "lfgpu_ctx_sync": (C.c_int, [_P]),
    "lfgpu_ctx_last_ms": (C.c_int, [_P, _d]),
    "lfgpu_vec_create": (C.c_int, [_P, C.c_int64, _d, C.POINTER(_P)]),
    "lfgpu_vec_destroy": (C.c_int, [_P]),
    "lfgpu_vec_size": (C.c_int, [_P, _i64]),
    "lfgpu_vec_upload": (C.c_int, [_P, _d, C.c_int64]),
    "lfgpu_vec_download": (C.c_int, [_P, _d, C.c_int64]),
    "lfgpu_vec_fill": (C.c_int, [_P, C.c_double]),
    "lfgpu_vec_copy": (C.c_int, [_P, _P, _P]),
    "lfgpu_vec_copy_nt": (C.c_int, [_P, _P, _P]),
    "lfgpu_mesh_create": (C.c_int, [_P, C.c_int64, _d, C.c_int64, _i32, C.c_int, _i64,
                                    C.c_int64, C.POINTER(_P)]),
    "lfgpu_mesh_destroy": (C.c_int, [_P]),
    "lfgpu_tp_triag_mesh": (C.c_int, [C.c_int32, C.c_int32, C.c_double, C.c_double,
                                      C.c_double, C.c_double, _d, _i32, _i32]),
    "lfgpu_mesh_edges": (C.c_int, [C.c_int64, C.c_int64, _i32, C.c_int64, _i32, C.c_int64,
                                   _i32, _i32, _i8, _i64]),
    "lfgpu_dof_table": (C.c_int, [C.c_int64, C.c_int64, C.c_int64, _i32, _i32, _i8, C.c_int,
                                  C.c_int, _i64, _i64]),
    "lfgpu_fe_create": (C.c_int, [_P, C.c_int, C.c_int, _d, _d, C.POINTER(_P)]),
    "lfgpu_fe_destroy": (C.c_int, [_P]),
    "lfgpu_fe_info": (C.c_int, [_P, C.POINTER(C.c_int), C.POINTER(C.c_int)]),
    "lfgpu_geometry": (C.c_int, [_P, _P, _P, _P, _P]),
    "lfgpu_element_matrices": (C.c_int, [_P, _P, _P, C.POINTER(Coeff), C.POINTER(Coeff), _P]),
    "lfgpu_element_matrices_mfma": (C.c_int, [_P, _P, _P, C.POINTER(Coeff), C.POINTER(Coeff),
                                              _P]),
    "lfgpu_element_vectors": (C.c_int, [_P, _P, _P, C.POINTER(Coeff), _P]),
    "lfgpu_plan_create": (C.c_int, [_P, _P, _u8, C.POINTER(_P)]),
    "lfgpu_edges_create": (C.c_int, [_P, _P, C.c_int, C.c_int64, _i32, _i64, C.POINTER(_P)]),
    "lfgpu_edges_destroy": (C.c_int, [_P]),
    "lfgpu_edge_matrices": (C.c_int, [_P, _P, C.POINTER(Coeff), _P]),
    "lfgpu_edge_vectors": (C.c_int, [_P, _P, C.POINTER(Coeff), _P]),
    "lfgpu_plan_set_edges": (C.c_int, [_P, _P, _u8, _u8]),
    "lfgpu_assemble_galerkin_bd": (C.c_int, [_P, _P, _P, C.POINTER(Coeff), C.POINTER(Coeff),
                                             C.POINTER(Coeff), C.POINTER(Coeff),
                                             C.POINTER(Coeff), _P, _P]),
    "lfgpu_plan_destroy": (C.c_int, [_P]),
    "lfgpu_plan_csr": (C.c_int, [_P, C.POINTER(_P)]),
    "lfgpu_assemble_galerkin": (C.c_int, [_P, _P, _P, C.POINTER(Coeff), C.POINTER(Coeff),
                                          C.POINTER(Coeff), _P, _P]),
    "lfgpu_assemble_galerkin_mfma": (C.c_int, [_P, _P, _P, C.POINTER(Coeff), C.POINTER(Coeff),
                                               C.POINTER(Coeff), _P, _P]),
    "lfgpu_csr_info": (C.c_int, [_P, _i32, _i32, _i64]),
    "lfgpu_csr_download": (C.c_int, [_P, _i32, _i32, _d]),
    "lfgpu_coo_to_csr": (C.c_int, [_P, C.c_int64, _i32, _i32, _d, C.c_int32, C.c_int32,
                                   C.POINTER(_P)]),
    "lfgpu_csr_destroy": (C.c_int, [_P]),
    "lfgpu_spmv": (C.c_int, [_P, _P, _P, _P]),
    "lfgpu_spmv_timed": (C.c_int, [_P, _P, _P, _P, C.c_int, _d]),
    "lfgpu_pcg": (C.c_int, [_P, _P, _P, _P, C.POINTER(PcgOpts), C.POINTER(PcgInfo)]),
    "lfgpu_pcg_noalloc": (C.c_int, [_P, _P, _P, _P, C.POINTER(PcgOpts), C.POINTER(PcgInfo)]),
    "lfgpu_solver_storage": (C.c_int, [_P, C.POINTER(C.c_int), C.POINTER(C.c_int),
                                        C.POINTER(C.c_int)]),
    "lfgpu_pcg_sweep_times": (C.c_int, [_P, _P, _P, _P, C.POINTER(PcgOpts), C.c_int,
                                         _d, C.POINTER(PcgInfo)]),
    "lfgpu_batch_create": (C.c_int, [_P, _P, C.c_int, _i32, C.POINTER(_P)]),
    "lfgpu_batch_destroy": (C.c_int, [_P]),
    "lfgpu_batch_storage": (C.c_int, [_P, C.POINTER(C.c_int)]),
    "lfgpu_batch_assemble_pcg": (C.c_int, [_P, _P, _P, _P, _P, _P, _P, _P, _P,
                                           C.POINTER(PcgOpts), C.POINTER(PcgInfo),
                                           C.POINTER(PcgInfo)]),
    "lfgpu_batch_pcg": (C.c_int, [_P, _P, _P, _P, C.POINTER(PcgOpts), C.POINTER(PcgInfo),
                                  C.POINTER(PcgInfo)]),
    "lfgpu_error_norms": (C.c_int, [_P, _P, _P, _P, C.c_int, _P, _P, C.POINTER(C.c_double),
                                    C.POINTER(C.c_double)]),
    "lfgpu_adjoint_gradient": (C.c_int, [_P, _P, _P, _P, _P, _P, _P]),
    "lfgpu_strang_create": (C.c_int, [_P, _P, _P, C.c_double, C.c_int, C.c_double,
                                      C.POINTER(PcgOpts), C.POINTER(_P)]),
    "lfgpu_strang_destroy": (C.c_int, [_P]),
    "lfgpu_strang_diffusion": (C.c_int, [_P, _P, C.c_int, _P, _P, C.POINTER(StrangInfo)]),
    "lfgpu_strang_evolution": (C.c_int, [_P, _P, _P, _P, _P, C.POINTER(StrangInfo)]),
    "lfgpu_csr_clone": (C.c_int, [_P, _P, C.POINTER(_P)]),
    "lfgpu_csr_axpby": (C.c_int, [_P, C.c_double, _P, C.c_double, _P, C.POINTER(_P)]),
    "lfgpu_nonlocal_gain": (C.c_int, [_P, _P, C.POINTER(C.c_uint8), C.c_double, C.c_double,
                                      _P]),
    "lfgpu_nonlocal_loss": (C.c_int, [_P, _P, C.POINTER(C.c_uint8), C.c_double, C.c_double,
                                      C.POINTER(C.c_int32), C.POINTER(C.c_int32), _P]),
    "lfgpu_csr_set_dense_block": (C.c_int, [_P, _P, C.c_int32, C.POINTER(C.c_int32), _P,
                                            C.c_double]),
    "lfgpu_csr_dense_block": (C.c_int, [_P, C.POINTER(C.c_int32), C.POINTER(C.c_double)]),
    "lfgpu_axpby": (C.c_int, [_P, C.c_double, _P, C.c_double, _P]),
    "lfgpu_dot": (C.c_int, [_P, _P, _P, _d]),
}

_LIB = None


class LfgpuError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("lfgpu error %d: %s" % (code, msg))
        self.code = code


def lib():
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("liblfgpu.so is not built; run `python -m lehrfempp_amd.build` "
                               "(there is no CPU fallback)")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _LIB = L
    return _LIB


def check(rc):
    if rc != OK:
        raise LfgpuError(rc, lib().lfgpu_last_error().decode())


def _ptr(a, ct):
    return a.ctypes.data_as(C.POINTER(ct))


class Context:
    def __init__(self, device=0):
        self.h = _P()
        check(lib().lfgpu_ctx_create(device, C.byref(self.h)))

    def close(self):
        if self.h:
            lib().lfgpu_ctx_destroy(self.h)
            self.h = _P()

    def sync(self):
        check(lib().lfgpu_ctx_sync(self.h))

    def last_ms(self):
        v = C.c_double()
        check(lib().lfgpu_ctx_last_ms(self.h, C.byref(v)))
        return v.value

    # -- factories ------------------------------------------------------------
    def vec(self, n_or_array):
        return Vec(self, n_or_array)

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


class Vec:
    def __init__(self, ctx: Context, n_or_array):
        self.ctx = ctx
        self.h = _P()
        if np.isscalar(n_or_array):
            self.n = int(n_or_array)
            check(lib().lfgpu_vec_create(ctx.h, self.n, None, C.byref(self.h)))
        else:
            a = np.ascontiguousarray(n_or_array, dtype=np.float64).ravel()
            self.n = a.size
            check(lib().lfgpu_vec_create(ctx.h, self.n, _ptr(a, C.c_double), C.byref(self.h)))

    def numpy(self):
        out = np.empty(self.n)
        check(lib().lfgpu_vec_download(self.h, _ptr(out, C.c_double), self.n))
        return out

    def upload(self, a):
        a = np.ascontiguousarray(a, dtype=np.float64).ravel()
        check(lib().lfgpu_vec_upload(self.h, _ptr(a, C.c_double), a.size))

    def fill(self, v):
        check(lib().lfgpu_vec_fill(self.h, float(v)))

    def close(self):
        if self.h:
            lib().lfgpu_vec_destroy(self.h)
            self.h = _P()


class Mesh:
    """Device mesh + DOF table (lfgpu_mesh_create)."""

    def __init__(self, ctx: Context, xy, cells, cell_dofs=None, ndofs=None):
        self.ctx = ctx
        self.xy = np.ascontiguousarray(xy, dtype=np.float64)
        self.cells = np.ascontiguousarray(cells, dtype=np.int32)
        self.nv = self.xy.shape[0]
        self.ne = self.cells.shape[0]
        if cell_dofs is None:
            self.nloc, self.ndofs, cd = 3, self.nv, None
        else:
            cd = np.ascontiguousarray(cell_dofs, dtype=np.int64)
            self.nloc = cd.shape[1]
            self.ndofs = int(ndofs)
        self.h = _P()
        check(lib().lfgpu_mesh_create(ctx.h, self.nv, _ptr(self.xy, C.c_double), self.ne,
                                      _ptr(self.cells, C.c_int32), self.nloc,
                                      None if cd is None else _ptr(cd, C.c_int64),
                                      self.ndofs, C.byref(self.h)))

    def close(self):
        if self.h:
            lib().lfgpu_mesh_destroy(self.h)
            self.h = _P()


class FE:
    def __init__(self, ctx: Context, degree, pts=None, w=None):
        self.ctx = ctx
        self.h = _P()
        if pts is None:
            check(lib().lfgpu_fe_create(ctx.h, degree, 0, None, None, C.byref(self.h)))
        else:
            p = np.ascontiguousarray(pts, dtype=np.float64).ravel()
            ww = np.ascontiguousarray(w, dtype=np.float64)
            check(lib().lfgpu_fe_create(ctx.h, degree, ww.size, _ptr(p, C.c_double),
                                        _ptr(ww, C.c_double), C.byref(self.h)))
        nl, nq = C.c_int(), C.c_int()
        check(lib().lfgpu_fe_info(self.h, C.byref(nl), C.byref(nq)))
        self.nloc, self.nqp = nl.value, nq.value

    def close(self):
        if self.h:
            lib().lfgpu_fe_destroy(self.h)
            self.h = _P()


def coeff(kind=COEFF_CONST, value=1.0, data: Vec | None = None):
    return Coeff(kind, float(value), data.h if data is not None else None)


class Csr:
    def __init__(self, ctx, h, owned):
        self.ctx, self.h, self.owned = ctx, h, owned

    def info(self):
        nr, nc, nnz = C.c_int32(), C.c_int32(), C.c_int64()
        check(lib().lfgpu_csr_info(self.h, C.byref(nr), C.byref(nc), C.byref(nnz)))
        return nr.value, nc.value, nnz.value

    def download(self):
        nr, _, nnz = self.info()
        rp = np.zeros(nr + 1, np.int32)
        col = np.zeros(max(nnz, 1), np.int32)
        val = np.zeros(max(nnz, 1))
        check(lib().lfgpu_csr_download(self.h, _ptr(rp, C.c_int32), _ptr(col, C.c_int32),
                                       _ptr(val, C.c_double)))
        return rp, col[:nnz], val[:nnz]

    def close(self):
        if self.h and self.owned:
            lib().lfgpu_csr_destroy(self.h)
        self.h = _P()


class Plan:
    def __init__(self, ctx: Context, mesh: Mesh, fixed_flags=None):
        self.ctx = ctx
        self.mesh = mesh
        self.h = _P()
        fl = None
        if fixed_flags is not None:
            fl = np.ascontiguousarray(fixed_flags, dtype=np.uint8)
            assert fl.size == mesh.ndofs
        check(lib().lfgpu_plan_create(ctx.h, mesh.h, None if fl is None else _ptr(fl, C.c_uint8),
                                      C.byref(self.h)))
        ch = _P()
        check(lib().lfgpu_plan_csr(self.h, C.byref(ch)))
        self.csr = Csr(ctx, ch, owned=False)

    def assemble(self, fe: FE, alpha: Coeff, gamma: Coeff, f: Coeff | None, rhs: Vec | None,
                 fixed_vals: Vec | None = None, mfma: bool = False):
        """lfgpu_assemble_galerkin, or lfgpu_assemble_galerkin_mfma with mfma=True.
        Returns the device time in ms."""
        fn = lib().lfgpu_assemble_galerkin_mfma if mfma else lib().lfgpu_assemble_galerkin
        check(fn(self.ctx.h, self.h, fe.h, C.byref(alpha), C.byref(gamma),
                 None if f is None else C.byref(f),
                 None if fixed_vals is None else fixed_vals.h,
                 None if rhs is None else rhs.h))
        return self.ctx.last_ms()

    def set_edges(self, edges, mat_flags=None, vec_flags=None):
        """lfgpu_plan_set_edges: edge mass term on mat_flags, edge load on vec_flags."""
        mf = None if mat_flags is None else np.ascontiguousarray(mat_flags, np.uint8)
        vf = None if vec_flags is None else np.ascontiguousarray(vec_flags, np.uint8)
        check(lib().lfgpu_plan_set_edges(self.h, edges.h,
                                         None if mf is None else _ptr(mf, C.c_uint8),
                                         None if vf is None else _ptr(vf, C.c_uint8)))

    def assemble_bd(self, fe: FE, alpha: Coeff, gamma: Coeff, f: Coeff | None,
                    eta: Coeff | None, h: Coeff | None, rhs: Vec | None,
                    fixed_vals: Vec | None = None):
        """lfgpu_assemble_galerkin_bd: cell terms plus the plan's edge terms."""
        check(lib().lfgpu_assemble_galerkin_bd(
            self.ctx.h, self.h, fe.h, C.byref(alpha), C.byref(gamma),
            None if f is None else C.byref(f), None if eta is None else C.byref(eta),
            None if h is None else C.byref(h), None if fixed_vals is None else fixed_vals.h,
            None if rhs is None else rhs.h))
        return self.ctx.last_ms()

    def close(self):
        if self.h:
            lib().lfgpu_plan_destroy(self.h)
            self.h = _P()


class Edges:
    """Codim-1 entities with the segment element (lfgpu_edges_create)."""

    def __init__(self, ctx: Context, mesh: Mesh, degree, edge_v, edge_dofs=None):
        self.ctx = ctx
        self.ev = np.ascontiguousarray(edge_v, dtype=np.int32)
        self.n = self.ev.shape[0]
        self.nle = degree + 1
        ed = None if edge_dofs is None else np.ascontiguousarray(edge_dofs, dtype=np.int64)
        self.h = _P()
        check(lib().lfgpu_edges_create(ctx.h, mesh.h, degree, self.n, _ptr(self.ev, C.c_int32),
                                       None if ed is None else _ptr(ed, C.c_int64),
                                       C.byref(self.h)))

    def close(self):
        if self.h:
            lib().lfgpu_edges_destroy(self.h)
            self.h = _P()


def edge_matrices(ctx, edges: Edges, eta: Coeff):
    out = Vec(ctx, edges.n * edges.nle * edges.nle)
    check(lib().lfgpu_edge_matrices(ctx.h, edges.h, C.byref(eta), out.h))
    a = out.numpy().reshape(edges.n, edges.nle, edges.nle)
    out.close()
    return a


def edge_vectors(ctx, edges: Edges, h: Coeff):
    out = Vec(ctx, edges.n * edges.nle)
    check(lib().lfgpu_edge_vectors(ctx.h, edges.h, C.byref(h), out.h))
    a = out.numpy().reshape(edges.n, edges.nle)
    out.close()
    return a


def element_matrices(ctx, mesh: Mesh, fe: FE, alpha: Coeff, gamma: Coeff, mfma: bool = False):
    out = Vec(ctx, mesh.ne * mesh.nloc * mesh.nloc)
    fn = lib().lfgpu_element_matrices_mfma if mfma else lib().lfgpu_element_matrices
    check(fn(ctx.h, mesh.h, fe.h, C.byref(alpha), C.byref(gamma), out.h))
    a = out.numpy().reshape(mesh.ne, mesh.nloc, mesh.nloc)
    out.close()
    return a


def element_vectors(ctx, mesh: Mesh, fe: FE, f: Coeff):
    out = Vec(ctx, mesh.ne * mesh.nloc)
    check(lib().lfgpu_element_vectors(ctx.h, mesh.h, fe.h, C.byref(f), out.h))
    a = out.numpy().reshape(mesh.ne, mesh.nloc)
    out.close()
    return a


def geometry(ctx, mesh: Mesh):
    J, Jit, det = Vec(ctx, 4 * mesh.ne), Vec(ctx, 4 * mesh.ne), Vec(ctx, mesh.ne)
    check(lib().lfgpu_geometry(ctx.h, mesh.h, J.h, Jit.h, det.h))
    res = (J.numpy().reshape(-1, 2, 2), Jit.numpy().reshape(-1, 2, 2), det.numpy())
    for v in (J, Jit, det):
        v.close()
    return res


def coo_to_csr(ctx, rows, cols, vals, nrows, ncols):
    r = np.ascontiguousarray(rows, np.int32)
    c = np.ascontiguousarray(cols, np.int32)
    v = np.ascontiguousarray(vals, np.float64)
    h = _P()
    check(lib().lfgpu_coo_to_csr(ctx.h, r.size, _ptr(r, C.c_int32), _ptr(c, C.c_int32),
                                 _ptr(v, C.c_double), nrows, ncols, C.byref(h)))
    return Csr(ctx, h, owned=True)


def spmv(ctx, A: Csr, x: Vec, y: Vec):
    check(lib().lfgpu_spmv(ctx.h, A.h, x.h, y.h))
    return ctx.last_ms()


def spmv_timed(ctx, A: Csr, x: Vec, y: Vec, reps: int):
    v = C.c_double()
    check(lib().lfgpu_spmv_timed(ctx.h, A.h, x.h, y.h, int(reps), C.byref(v)))
    return v.value


def pcg(ctx, A: Csr, b: Vec, x: Vec, rtol=1e-10, maxit=None, mode=PCG_FP64, check_every=64,
        raise_on_noconv=True, refine_delta=0.0, refine_max_interval=0, single_reduction=False,
        graph=0):
    """lfgpu_pcg.  graph: 0 = replay captured FP64 iteration chunks below 2^20
    rows, 1 = always, -1 = never."""
    n = A.info()[0]
    opts = PcgOpts(rtol, int(2 * n if maxit is None else maxit), mode, check_every,
                   float(refine_delta), int(refine_max_interval), int(bool(single_reduction)),
                   int(graph))
    info = PcgInfo()
    rc = lib().lfgpu_pcg(ctx.h, A.h, b.h, x.h, C.byref(opts), C.byref(info))
    if rc != OK and (rc != ENOCONV or raise_on_noconv):
        check(rc)
    return info


def solver_storage(A: Csr):
    """(storage, ndiag, sym) of the solver's copy of A (storage -1 before the
    first solve)."""
    st, nd, sy = C.c_int(), C.c_int(), C.c_int()
    check(lib().lfgpu_solver_storage(A.h, C.byref(st), C.byref(nd), C.byref(sy)))
    return st.value, nd.value, bool(sy.value)


def pcg_sweep_times(ctx, A: Csr, b: Vec, x: Vec, iters=200, mode=PCG_FP64, max_kernels=None):
    """Runs `iters` PCG iterations (rtol 0) with every sweep kernel timed by
    its own hipEvent pair; returns the median durations (ms) of the first
    sweep, the second sweep and the row-sweep reduction kernels."""
    n = A.info()[0]
    opts = PcgOpts(0.0, int(iters), mode, 64, 0.0, 0, 0, -1)
    out = (C.c_double * 3)()
    info = PcgInfo()
    rc = lib().lfgpu_pcg_sweep_times(ctx.h, A.h, b.h, x.h, C.byref(opts),
                                     int(max_kernels or 4 * iters + 8), out, C.byref(info))
    if rc != OK and rc != ENOCONV:
        check(rc)
    return list(out), info


def pcg_noalloc(ctx, A: Csr, b: Vec, x: Vec, rtol=1e-10, maxit=None, check_every=64,
                raise_on_noconv=True):
    """lfgpu_pcg_noalloc: the FP64 two-sweep PCG without any device allocation
    (fixed per-GPU work space, systems of up to 2^21 rows)."""
    n = A.info()[0]
    opts = PcgOpts(rtol, int(2 * n if maxit is None else maxit), PCG_FP64, check_every,
                   0.0, 0, 0, 0)
    info = PcgInfo()
    rc = lib().lfgpu_pcg_noalloc(ctx.h, A.h, b.h, x.h, C.byref(opts), C.byref(info))
    if rc != OK and (rc != ENOCONV or raise_on_noconv):
        check(rc)
    return info


class Batch:
    """Batched Jacobi-PCG over a block-diagonal matrix (lfgpu_batch_*)."""

    def __init__(self, ctx, A: Csr, row_offsets):
        self.ctx = ctx
        self.off = np.ascontiguousarray(row_offsets, dtype=np.int32)
        self.nprob = self.off.size - 1
        self.h = _P()
        check(lib().lfgpu_batch_create(ctx.h, A.h, self.nprob, _ptr(self.off, C.c_int32),
                                       C.byref(self.h)))

    def pcg(self, b: Vec, x: Vec, rtol=1e-10, maxit=100000, check_every=64,
            raise_on_noconv=True, mode=PCG_FP64, refine_delta=0.0, refine_max_interval=0,
            graph=0):
        """lfgpu_batch_pcg.  graph: 0 = replay the captured chunk, -1 = plain
        launches."""
        opts = PcgOpts(rtol, int(maxit), mode, check_every, refine_delta, refine_max_interval, 0,
                       int(graph))
        infos = (PcgInfo * self.nprob)()
        summ = PcgInfo()
        rc = lib().lfgpu_batch_pcg(self.ctx.h, self.h, b.h, x.h, C.byref(opts), infos,
                                   C.byref(summ))
        if rc != OK and (rc != ENOCONV or raise_on_noconv):
            check(rc)
        return summ, list(infos)

    def assemble_pcg(self, plan, fe, alpha, gamma, f, rhs: Vec, x: Vec, rtol=1e-10,
                     maxit=100000, check_every=64, raise_on_noconv=True, mode=PCG_FP64,
                     refine_delta=0.0, refine_max_interval=0):
        """lfgpu_batch_assemble_pcg: plan.assemble(fe, alpha, gamma, f, rhs) and
        pcg(rhs, x), replayed as one captured graph from the second call with
        the same arguments on."""
        opts = PcgOpts(rtol, int(maxit), mode, check_every, refine_delta, refine_max_interval, 0,
                       0)
        infos = (PcgInfo * self.nprob)()
        summ = PcgInfo()
        rc = lib().lfgpu_batch_assemble_pcg(
            self.ctx.h, self.h, plan.h, fe.h, C.byref(alpha), C.byref(gamma),
            None if f is None else C.byref(f), rhs.h, x.h, C.byref(opts), infos, C.byref(summ))
        if rc != OK and (rc != ENOCONV or raise_on_noconv):
            check(rc)
        return summ, list(infos)

    def storage(self):
        """lfgpu_batch_storage: STORAGE_ELL / _ELL16 / _DIA / _BLOCK_DIA (-1 before a solve)."""
        v = C.c_int()
        check(lib().lfgpu_batch_storage(self.h, C.byref(v)))
        return v.value

    def close(self):
        if self.h:
            lib().lfgpu_batch_destroy(self.h)
            self.h = _P()


EXACT_SINSIN, EXACT_QP = 0, 1


def error_norms(ctx, mesh: Mesh, fe: FE, uh: Vec, u_qp: Vec | None = None,
                grad_qp: Vec | None = None):
    """(L2 error^2, H1-seminorm error^2) of uh (lfgpu_error_norms): against
    sin(pi x) sin(pi y) without u_qp, else against the given point values."""
    l2, h1 = C.c_double(), C.c_double()
    kind = EXACT_SINSIN if u_qp is None else EXACT_QP
    check(lib().lfgpu_error_norms(ctx.h, mesh.h, fe.h, uh.h, kind,
                                  None if u_qp is None else u_qp.h,
                                  None if grad_qp is None else grad_qp.h,
                                  C.byref(l2), C.byref(h1)))
    return l2.value, h1.value


def adjoint_gradient(ctx, mesh: Mesh, fe: FE, theta: Vec, u: Vec, lam: Vec, grad: Vec):
    check(lib().lfgpu_adjoint_gradient(ctx.h, mesh.h, fe.h, theta.h, u.h, lam.h, grad.h))
    return ctx.last_ms()


def csr_clone(ctx, A: Csr) -> Csr:
    """lfgpu_csr_clone: an owned copy of A (pattern and values)."""
    h = _P()
    check(lib().lfgpu_csr_clone(ctx.h, A.h, C.byref(h)))
    return Csr(ctx, h, owned=True)


def nonlocal_gain(ctx, edges: Edges, flags, rmin=0.0, rmax=0.0) -> Vec:
    """lfgpu_nonlocal_gain: h at the segment points of the flagged edges, a
    device vector [nedges * nqe] for an LFGPU_COEFF_QP edge coefficient."""
    fl = np.ascontiguousarray(flags, dtype=np.uint8)
    assert fl.size == edges.n
    nqe = edges.nle  # degree + 1 points for degree 1 and 3
    eta = Vec(ctx, edges.n * nqe)
    check(lib().lfgpu_nonlocal_gain(ctx.h, edges.h, _ptr(fl, C.c_uint8), float(rmin),
                                    float(rmax), eta.h))
    return eta


def nonlocal_loss(ctx, edges: Edges, flags, rmin=0.0, rmax=0.0):
    """lfgpu_nonlocal_loss: (rows, L) with rows the DOFs of the flagged edges
    (ascending) and L a device vector holding the nb x nb block row-major."""
    fl = np.ascontiguousarray(flags, dtype=np.uint8)
    assert fl.size == edges.n
    nb = C.c_int32()
    check(lib().lfgpu_nonlocal_loss(ctx.h, edges.h, _ptr(fl, C.c_uint8), float(rmin),
                                    float(rmax), C.byref(nb), None, None))
    rows = np.zeros(max(nb.value, 1), np.int32)
    L = Vec(ctx, max(nb.value, 1) ** 2)
    check(lib().lfgpu_nonlocal_loss(ctx.h, edges.h, _ptr(fl, C.c_uint8), float(rmin),
                                    float(rmax), C.byref(nb), _ptr(rows, C.c_int32), L.h))
    return rows[:nb.value], L


def csr_set_dense_block(ctx, A: Csr, rows, B: Vec | None, scale=1.0):
    """lfgpu_csr_set_dense_block: A stands for its sparse part + scale * B on
    rows (ascending); rows of length 0 removes the block."""
    r = np.ascontiguousarray(rows, dtype=np.int32)
    check(lib().lfgpu_csr_set_dense_block(ctx.h, A.h, int(r.size),
                                          _ptr(r, C.c_int32) if r.size else None,
                                          None if B is None else B.h, float(scale)))


def csr_dense_block(A: Csr):
    """(nb, scale) of the dense block attached to A."""
    nb, sc = C.c_int32(), C.c_double()
    check(lib().lfgpu_csr_dense_block(A.h, C.byref(nb), C.byref(sc)))
    return nb.value, sc.value


def csr_axpby(ctx, a, X: Csr, b, Y: Csr) -> Csr:
    """lfgpu_csr_axpby: a X + b Y on the common pattern of X and Y."""
    h = _P()
    check(lib().lfgpu_csr_axpby(ctx.h, float(a), X.h, float(b), Y.h, C.byref(h)))
    return Csr(ctx, h, owned=True)


class Strang:
    """Strang-splitting time stepper (lfgpu_strang_*)."""

    def __init__(self, ctx, A: Csr, M: Csr, T, m, lam, rtol=1e-10, maxit=100000,
                 mode=PCG_FP64, check_every=0):
        self.ctx = ctx
        opts = PcgOpts(rtol, int(maxit), mode, check_every, 0.0, 0, 0)
        self.h = _P()
        check(lib().lfgpu_strang_create(ctx.h, A.h, M.h, float(T), int(m), float(lam),
                                        C.byref(opts), C.byref(self.h)))

    def diffusion(self, firstcall, mu: Vec, out: Vec, raise_on_noconv=True):
        info = StrangInfo()
        rc = lib().lfgpu_strang_diffusion(self.ctx.h, self.h, int(bool(firstcall)), mu.h, out.h,
                                          C.byref(info))
        if rc != OK and (rc != ENOCONV or raise_on_noconv):
            check(rc)
        return info

    def evolution(self, cap: Vec, mu: Vec, out: Vec, raise_on_noconv=True):
        info = StrangInfo()
        rc = lib().lfgpu_strang_evolution(self.ctx.h, self.h, cap.h, mu.h, out.h, C.byref(info))
        if rc != OK and (rc != ENOCONV or raise_on_noconv):
            check(rc)
        return info

    def close(self):
        if self.h:
            lib().lfgpu_strang_destroy(self.h)
            self.h = _P()


def copy(ctx, x: Vec, y: Vec):
    check(lib().lfgpu_vec_copy(ctx.h, x.h, y.h))


def copy_nt(ctx, x: Vec, y: Vec):
    check(lib().lfgpu_vec_copy_nt(ctx.h, x.h, y.h))


def dot(ctx, x: Vec, y: Vec):
    v = C.c_double()
    check(lib().lfgpu_dot(ctx.h, x.h, y.h, C.byref(v)))
    return v.value


def axpby(ctx, a, x: Vec, b, y: Vec):
    check(lib().lfgpu_axpby(ctx.h, float(a), x.h, float(b), y.h))


# -- host topology helpers (no GPU work) --------------------------------------
def tp_triag_mesh(nx, ny=None, box=(0.0, 0.0, 1.0, 1.0)):
    """TPTriagMeshBuilder: (xy, cells, edges) in the builder's order."""
    ny = nx if ny is None else ny
    nv, ne = (nx + 1) * (ny + 1), 2 * nx * ny
    E = nx * ny + (nx + 1) * ny + nx * (ny + 1)
    xy = np.zeros((nv, 2))
    cells = np.zeros((ne, 3), np.int32)
    edges = np.zeros((E, 2), np.int32)
    check(lib().lfgpu_tp_triag_mesh(nx, ny, box[0], box[1], box[2], box[3],
                                    _ptr(xy, C.c_double), _ptr(cells, C.c_int32),
                                    _ptr(edges, C.c_int32)))
    return xy, cells, edges


def batched_tp_mesh(sizes, box=(0.0, 0.0, 1.0, 1.0)):
    """Disjoint union of n x n TPTriagMeshBuilder meshes, in order: (xy, cells,
    boundary flags, vertex offsets [len(sizes)+1], cell offsets)."""
    xs, cs, fl, voff, coff = [], [], [], [0], [0]
    for n in sizes:
        xy, cells, _ = tp_triag_mesh(int(n), int(n), box)
        j, i = np.divmod(np.arange(xy.shape[0]), n + 1)
        fl.append(((i == 0) | (i == n) | (j == 0) | (j == n)).astype(np.uint8))
        xs.append(xy)
        cs.append(cells + voff[-1])
        voff.append(voff[-1] + xy.shape[0])
        coff.append(coff[-1] + cells.shape[0])
    return (np.concatenate(xs), np.ascontiguousarray(np.concatenate(cs), np.int32),
            np.concatenate(fl), np.array(voff, np.int64), np.array(coff, np.int64))


def mesh_edges(nv, cells, pre_edges=None):
    cells = np.ascontiguousarray(cells, np.int32)
    ne = cells.shape[0]
    pre = (np.zeros((0, 2), np.int32) if pre_edges is None
           else np.ascontiguousarray(pre_edges, np.int32))
    cap = pre.shape[0] + 3 * ne
    ev = np.zeros((cap, 2), np.int32)
    ce = np.zeros((ne, 3), np.int32)
    co = np.zeros((ne, 3), np.int8)
    E = C.c_int64()
    check(lib().lfgpu_mesh_edges(nv, ne, _ptr(cells, C.c_int32), pre.shape[0],
                                 _ptr(pre, C.c_int32) if pre.size else None, cap,
                                 _ptr(ev, C.c_int32), _ptr(ce, C.c_int32), _ptr(co, C.c_int8),
                                 C.byref(E)))
    return ev[:E.value].copy(), ce, co


def dof_table(nv, cells, nedges, cell_e, cell_ori, degree):
    cells = np.ascontiguousarray(cells, np.int32)
    ne = cells.shape[0]
    de, dc = (0, 0) if degree == 1 else (2, 1)
    nloc = 3 + 3 * de + dc
    cd = np.zeros((ne, nloc), np.int64)
    nd = C.c_int64()
    ce = np.ascontiguousarray(cell_e, np.int32)
    co = np.ascontiguousarray(cell_ori, np.int8)
    check(lib().lfgpu_dof_table(nv, ne, nedges, _ptr(cells, C.c_int32), _ptr(ce, C.c_int32),
                                _ptr(co, C.c_int8), de, dc, _ptr(cd, C.c_int64), C.byref(nd)))
    return cd, nd.value
