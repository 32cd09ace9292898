"""numpy front end of the CPU oracle (oracle/liblfref.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by ``__graft_entry__.smoke``
and by the CPU-baseline leg of ``bench.py``.  Never used by the product
library.  See oracle/lfref.h for the reference file:line each function
restates.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

_d = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
_i32 = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
_i64 = np.ctypeslib.ndpointer(np.int64, flags="C_CONTIGUOUS")
_u8 = np.ctypeslib.ndpointer(np.uint8, flags="C_CONTIGUOUS")
_i8 = np.ctypeslib.ndpointer(np.int8, flags="C_CONTIGUOUS")


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liblfref.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle not built: run `make -C oracle`")
        L = C.CDLL(path)
        sig = {
            "lfref_quad_rule": (C.c_int, [C.c_int, _d, _d]),
            "lfref_shape_functions": (C.c_int, [C.c_int, C.c_int, _d, _d, _d]),
            "lfref_tp_triag_mesh": (None, [C.c_int, C.c_int, C.c_double, C.c_double,
                                           C.c_double, C.c_double, _d, _i32, _i32]),
            "lfref_mesh_edges": (C.c_int, [C.c_int, C.c_int, _i32, C.c_int, _i32, C.c_int,
                                           _i32, _i32, _i8]),
            "lfref_dof_table": (C.c_int64, [C.c_int, C.c_int, C.c_int, _i32, _i32, _i8,
                                            C.c_int, C.c_int, _i64]),
            "lfref_boundary_flags": (None, [C.c_int, C.c_int, C.c_int, _i32, _i32, _u8, _u8]),
            "lfref_geometry": (C.c_int, [C.c_int, _d, _i32, _d, _d, _d]),
            "lfref_element_matrices": (C.c_int, [C.c_int, C.c_int, _d, _i32, _d, _d, _d]),
            "lfref_element_vectors": (C.c_int, [C.c_int, C.c_int, _d, _i32, C.c_int,
                                                C.c_double, C.c_void_p, _d]),
            "lfref_qp_coords": (C.c_int, [C.c_int, C.c_int, _d, _i32, _d]),
            "lfref_assemble_coo": (None, [C.c_int, C.c_int, _i64, _d, _i32, _i32, _d]),
            "lfref_assemble_vector": (None, [C.c_int, C.c_int, _i64, _d, C.c_int64, _d]),
            "lfref_coo_matvec": (None, [C.c_int64, _i32, _i32, _d, C.c_double, _d, _d]),
            "lfref_fix_flagged": (C.c_int64, [C.c_int64, _i32, _i32, _d, C.c_int64, _u8,
                                              _d, _d]),
            "lfref_set_from_triplets": (C.c_int64, [C.c_int64, _i32, _i32, _d, C.c_int32,
                                                    _i32, _i32, _d]),
            "lfref_band_cholesky_solve": (C.c_int, [C.c_int32, _i32, _i32, _d, _d, _d]),
            "lfref_csr_spmv": (None, [C.c_int32, _i32, _i32, _d, _d, _d]),
            "lfref_pcg": (C.c_int, [C.c_int32, _i32, _i32, _d, _d, _d, C.c_double, C.c_int,
                                    C.POINTER(C.c_double)]),
            "lfref_cpu_path_assemble": (C.c_int64, [C.c_int, C.c_int, _d, _i32, _i64,
                                                    C.c_int64, _d, C.c_double, C.c_int,
                                                    C.c_double, _u8, _i32, _i32, _d, _d]),
            "lfref_adjoint_gradient_p1": (C.c_int, [C.c_int, _d, _i32, _d, _d, _d, _d]),
            "lfref_error_norms": (C.c_int, [C.c_int, C.c_int, _d, _i32, _i64, _d, C.c_int,
                                            C.c_void_p, C.c_void_p, _d]),
            "lfref_segment_rule": (C.c_int, [C.c_int, _d, _d]),
            "lfref_edge_matrices": (C.c_int, [C.c_int, C.c_int, _d, _i32, C.c_int, C.c_double,
                                              C.c_void_p, _d]),
            "lfref_edge_vectors": (C.c_int, [C.c_int, C.c_int, _d, _i32, C.c_int, C.c_double,
                                             C.c_void_p, _d]),
            "lfref_nonlocal_gain": (C.c_int, [C.c_int, C.c_int, _d, _i32, _u8, C.c_double,
                                              C.c_double, _d]),
            "lfref_nonlocal_loss": (C.c_int, [C.c_int, C.c_int, _d, _i32, _i64, _u8,
                                              C.c_double, C.c_double, C.c_int, _i32, _d]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _LIB = L
    return _LIB


def nqp_for_degree(degree: int) -> int:
    return {1: 3, 3: 12}[degree]


def quad_rule(qdeg: int):
    pts = np.zeros(24)
    w = np.zeros(12)
    n = lib().lfref_quad_rule(qdeg, pts, w)
    if n < 0:
        raise ValueError("no triangle rule of degree %d in the oracle" % qdeg)
    return pts[: 2 * n].reshape(n, 2).copy(), w[:n].copy()


def shape_functions(degree: int, pts: np.ndarray):
    """Values [nsf x npts] and gradients [nsf x 2*npts] (Eigen column layout)."""
    pts = np.ascontiguousarray(pts, dtype=np.float64)
    npts = pts.shape[0]
    nsf = 3 if degree == 1 else 10
    N = np.zeros(nsf * npts)
    dN = np.zeros(nsf * 2 * npts)
    lib().lfref_shape_functions(degree, npts, pts.ravel(), N, dN)
    return N.reshape(npts, nsf).T.copy(), dN.reshape(2 * npts, nsf).T.copy()


class TriMesh:
    """Triangle mesh arrays: xy [nv,2], cells [ne,3], edges [E,2], cell_e, cell_ori."""

    def __init__(self, xy, cells, pre_edges=None):
        self.xy = np.ascontiguousarray(xy, dtype=np.float64)
        self.cells = np.ascontiguousarray(cells, dtype=np.int32)
        self.nv = self.xy.shape[0]
        self.ne = self.cells.shape[0]
        pre = (np.zeros((0, 2), np.int32) if pre_edges is None
               else np.ascontiguousarray(pre_edges, dtype=np.int32))
        cap = pre.shape[0] + 3 * self.ne
        ev = np.zeros(2 * cap, np.int32)
        ce = np.zeros(3 * self.ne, np.int32)
        co = np.zeros(3 * self.ne, np.int8)
        E = lib().lfref_mesh_edges(self.nv, self.ne, self.cells.ravel(), pre.shape[0],
                                   pre.ravel() if pre.size else np.zeros(2, np.int32),
                                   cap, ev, ce, co)
        assert E >= 0
        self.nedges = E
        self.edges = ev[: 2 * E].reshape(E, 2).copy()
        self.cell_e = ce.reshape(self.ne, 3)
        self.cell_ori = co.reshape(self.ne, 3)

    def boundary_flags(self):
        nf = np.zeros(self.nv, np.uint8)
        ef = np.zeros(self.nedges, np.uint8)
        lib().lfref_boundary_flags(self.nv, self.ne, self.nedges, self.cell_e.ravel(),
                                   self.edges.ravel(), nf, ef)
        return nf, ef

    def dofs(self, degree: int):
        """(cell_dofs [ne, nloc] int64, ndofs) of UniformFEDofHandler."""
        ndof_edge, ndof_cell = (0, 0) if degree == 1 else (2, 1)
        nloc = 3 + 3 * ndof_edge + ndof_cell
        cd = np.zeros(self.ne * nloc, np.int64)
        nd = lib().lfref_dof_table(self.nv, self.ne, self.nedges, self.cells.ravel(),
                                   self.cell_e.ravel(), self.cell_ori.ravel(), ndof_edge,
                                   ndof_cell, cd)
        return cd.reshape(self.ne, nloc), int(nd)

    def edge_dofs(self, degree: int):
        """GlobalDofIndices of every edge (dofhandler.cc:160-181): its two
        endpoint DOFs in the edge's orientation, then its interior DOFs."""
        E = self.nedges
        if degree == 1:
            return self.edges.astype(np.int64).copy()
        ed = np.zeros((E, 4), np.int64)
        ed[:, :2] = self.edges
        ed[:, 2] = self.nv + 2 * np.arange(E)
        ed[:, 3] = self.nv + 2 * np.arange(E) + 1
        return ed

    def dirichlet_flags(self, degree: int):
        """Boundary DOF flags: boundary nodes, plus interior DOFs of boundary edges."""
        nf, ef = self.boundary_flags()
        if degree == 1:
            return nf.copy()
        _, nd = self.dofs(degree)
        fl = np.zeros(nd, np.uint8)
        fl[: self.nv] = nf
        be = np.nonzero(ef)[0]
        fl[self.nv + 2 * be] = 1
        fl[self.nv + 2 * be + 1] = 1
        return fl


def tp_triag_mesh(nx: int, ny: int | None = None, box=(0.0, 0.0, 1.0, 1.0)) -> TriMesh:
    ny = nx if ny is None else ny
    nv = (nx + 1) * (ny + 1)
    ne = 2 * nx * ny
    E = nx * ny + (nx + 1) * ny + nx * (ny + 1)
    xy = np.zeros(2 * nv)
    cells = np.zeros(3 * ne, np.int32)
    edges = np.zeros(2 * E, np.int32)
    lib().lfref_tp_triag_mesh(nx, ny, box[0], box[1], box[2], box[3], xy, cells, edges)
    return TriMesh(xy.reshape(nv, 2), cells.reshape(ne, 3), edges.reshape(E, 2))


def geometry(mesh: TriMesh):
    J = np.zeros(4 * mesh.ne)
    Jit = np.zeros(4 * mesh.ne)
    det = np.zeros(mesh.ne)
    rc = lib().lfref_geometry(mesh.ne, mesh.xy.ravel(), mesh.cells.ravel(), J, Jit, det)
    if rc != 0:
        raise ValueError("degenerate cell %d" % (-rc - 1))
    return J.reshape(-1, 2, 2), Jit.reshape(-1, 2, 2), det


def _qp_array(v, ne, nqp):
    v = np.asarray(v, dtype=np.float64)
    if v.ndim == 0:
        return np.full(ne * nqp, float(v))
    if v.size == ne:
        return np.ascontiguousarray(np.repeat(v.ravel(), nqp))
    assert v.size == ne * nqp
    return np.ascontiguousarray(v.ravel())


def element_matrices(mesh: TriMesh, degree: int, alpha, gamma):
    q = nqp_for_degree(degree)
    nsf = 3 if degree == 1 else 10
    out = np.zeros(mesh.ne * nsf * nsf)
    rc = lib().lfref_element_matrices(degree, mesh.ne, mesh.xy.ravel(), mesh.cells.ravel(),
                                      _qp_array(alpha, mesh.ne, q),
                                      _qp_array(gamma, mesh.ne, q), out)
    if rc != 0:
        raise ValueError("element matrices failed: %d" % rc)
    return out.reshape(mesh.ne, nsf, nsf)


F_CONST, F_SINSIN, F_QP = 0, 1, 2


def element_vectors(mesh: TriMesh, degree: int, f_kind=F_CONST, f_const=1.0, f_qp=None):
    nsf = 3 if degree == 1 else 10
    out = np.zeros(mesh.ne * nsf)
    fq = None
    if f_kind == F_QP:
        fq = np.ascontiguousarray(f_qp, dtype=np.float64).ravel()
    rc = lib().lfref_element_vectors(degree, mesh.ne, mesh.xy.ravel(), mesh.cells.ravel(),
                                     f_kind, f_const,
                                     None if fq is None else fq.ctypes.data, out)
    if rc != 0:
        raise ValueError("element vectors failed: %d" % rc)
    return out.reshape(mesh.ne, nsf)


def qp_coords(mesh: TriMesh, degree: int):
    q = nqp_for_degree(degree)
    out = np.zeros(mesh.ne * q * 2)
    lib().lfref_qp_coords(degree, mesh.ne, mesh.xy.ravel(), mesh.cells.ravel(), out)
    return out.reshape(mesh.ne, q, 2)


def assemble_coo(cell_dofs, mats):
    ne, nloc = cell_dofs.shape
    nnz = ne * nloc * nloc
    r = np.zeros(nnz, np.int32)
    c = np.zeros(nnz, np.int32)
    v = np.zeros(nnz)
    lib().lfref_assemble_coo(ne, nloc, np.ascontiguousarray(cell_dofs, np.int64).ravel(),
                             np.ascontiguousarray(mats).ravel(), r, c, v)
    return r, c, v


def assemble_vector(cell_dofs, vecs, ndofs):
    ne, nloc = cell_dofs.shape
    b = np.zeros(ndofs)
    lib().lfref_assemble_vector(ne, nloc, np.ascontiguousarray(cell_dofs, np.int64).ravel(),
                                np.ascontiguousarray(vecs).ravel(), ndofs, b)
    return b


def coo_matvec(rows, cols, vals, alpha, x, y):
    lib().lfref_coo_matvec(len(rows), rows, cols, vals, alpha,
                           np.ascontiguousarray(x, np.float64), y)
    return y


def fix_flagged(rows, cols, vals, flags, fixed_vals, b):
    ndofs = len(flags)
    nnz = len(rows)
    cap = nnz + ndofs
    R = np.zeros(cap, np.int32)
    Cc = np.zeros(cap, np.int32)
    V = np.zeros(cap)
    R[:nnz], Cc[:nnz], V[:nnz] = rows, cols, vals
    b = np.array(b, dtype=np.float64)
    n = lib().lfref_fix_flagged(nnz, R, Cc, V, ndofs, np.ascontiguousarray(flags, np.uint8),
                                np.ascontiguousarray(fixed_vals, np.float64), b)
    return R[:n].copy(), Cc[:n].copy(), V[:n].copy(), b


def set_from_triplets(rows, cols, vals, nrows):
    nnz = len(rows)
    rp = np.zeros(nrows + 1, np.int32)
    col = np.zeros(max(nnz, 1), np.int32)
    val = np.zeros(max(nnz, 1))
    n = lib().lfref_set_from_triplets(nnz, np.ascontiguousarray(rows, np.int32),
                                      np.ascontiguousarray(cols, np.int32),
                                      np.ascontiguousarray(vals, np.float64), nrows, rp,
                                      col, val)
    return rp, col[:n].copy(), val[:n].copy()


def band_cholesky_solve(rp, col, val, b):
    x = np.zeros(len(rp) - 1)
    rc = lib().lfref_band_cholesky_solve(len(rp) - 1, rp, col, val,
                                         np.ascontiguousarray(b, np.float64), x)
    if rc != 0:
        raise ValueError("band Cholesky failed: %d" % rc)
    return x


def csr_spmv(rp, col, val, x):
    y = np.zeros(len(rp) - 1)
    lib().lfref_csr_spmv(len(rp) - 1, rp, col, val, np.ascontiguousarray(x, np.float64), y)
    return y


def pcg(rp, col, val, b, tol=1e-10, maxit=100000, x0=None):
    n = len(rp) - 1
    x = np.zeros(n) if x0 is None else np.array(x0, dtype=np.float64)
    err = C.c_double(0.0)
    it = lib().lfref_pcg(n, rp, col, val, np.ascontiguousarray(b, np.float64), x, tol,
                         maxit, C.byref(err))
    return x, it, err.value


_OMP = None


def _omp_lib():
    """oracle/liblfref_omp.so: the all-cores PCG, loaded only when timed."""
    global _OMP
    if _OMP is None:
        path = os.path.join(_HERE, "liblfref_omp.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle not built: run `make -C oracle`")
        L = C.CDLL(path)
        L.lfref_pcg_omp.restype = C.c_int
        L.lfref_pcg_omp.argtypes = [C.c_int32, _i32, _i32, _d, _d, _d, C.c_double, C.c_int,
                                    C.POINTER(C.c_double)]
        L.lfref_omp_threads.restype = C.c_int
        L.lfref_omp_threads.argtypes = []
        _OMP = L
    return _OMP


def omp_threads():
    return int(_omp_lib().lfref_omp_threads())


def pcg_omp(rp, col, val, b, tol=1e-10, maxit=100000, x0=None):
    """pcg on all cores (OpenMP loops and reductions; different summation
    order, so the last bits of the iterates differ from pcg)."""
    n = len(rp) - 1
    x = np.zeros(n) if x0 is None else np.array(x0, dtype=np.float64)
    err = C.c_double(0.0)
    it = _omp_lib().lfref_pcg_omp(n, rp, col, val, np.ascontiguousarray(b, np.float64), x,
                                  tol, maxit, C.byref(err))
    return x, it, err.value


def cpu_path_assemble(mesh: TriMesh, degree, cell_dofs, ndofs, alpha_cell, gamma_const,
                      f_kind, f_const, fixed):
    """Reference-faithful single-thread path: Eval -> triplets -> fix -> CSR."""
    nloc = cell_dofs.shape[1]
    cap = mesh.ne * nloc * nloc + ndofs
    rp = np.zeros(ndofs + 1, np.int32)
    col = np.zeros(cap, np.int32)
    val = np.zeros(cap)
    b = np.zeros(ndofs)
    n = lib().lfref_cpu_path_assemble(degree, mesh.ne, mesh.xy.ravel(), mesh.cells.ravel(),
                                      np.ascontiguousarray(cell_dofs, np.int64).ravel(),
                                      ndofs, np.ascontiguousarray(alpha_cell, np.float64),
                                      gamma_const, f_kind, f_const,
                                      np.ascontiguousarray(fixed, np.uint8), rp, col, val, b)
    if n < 0:
        raise ValueError("cpu path failed: %d" % n)
    return rp, col[:n].copy(), val[:n].copy(), b


def adjoint_gradient_p1(mesh: TriMesh, theta, u, lam):
    g = np.zeros(mesh.ne)
    rc = lib().lfref_adjoint_gradient_p1(mesh.ne, mesh.xy.ravel(), mesh.cells.ravel(),
                                         np.ascontiguousarray(theta, np.float64),
                                         np.ascontiguousarray(u, np.float64),
                                         np.ascontiguousarray(lam, np.float64), g)
    assert rc == 0
    return g


def error_norms(mesh: TriMesh, degree: int, uh, u_qp=None, g_qp=None):
    """Squared L2 and H1-seminorm errors of the FE function uh (lfref_error_norms):
    against u = sin(pi x) sin(pi y) when u_qp is None, else against the given
    per-point values u_qp [ne x nqp] and gradients g_qp [ne x nqp x 2]."""
    cd, nd = mesh.dofs(degree)
    out = np.zeros(2)
    if u_qp is None:
        rc = lib().lfref_error_norms(degree, mesh.ne, mesh.xy.ravel(), mesh.cells.ravel(), cd,
                                     np.ascontiguousarray(uh, np.float64), 0, None, None, out)
    else:
        uq = np.ascontiguousarray(u_qp, np.float64)
        gq = np.ascontiguousarray(g_qp, np.float64)
        rc = lib().lfref_error_norms(degree, mesh.ne, mesh.xy.ravel(), mesh.cells.ravel(), cd,
                                     np.ascontiguousarray(uh, np.float64), 1,
                                     uq.ctypes.data, gq.ctypes.data, out)
    assert rc == 0, rc
    return out[0], out[1]


def segment_rule(n: int):
    pts, w = np.zeros(n), np.zeros(n)
    assert lib().lfref_segment_rule(n, pts, w) == 0
    return pts, w


G_CONST, G_EDGE, G_QP, G_SINSIN = 0, 1, 2, 3


def _edge_call(fn, degree, xy, edge_v, kind, const, data, width):
    ev = np.ascontiguousarray(edge_v, np.int32)
    E = ev.shape[0]
    out = np.zeros(E * width)
    d = None
    if data is not None:
        d = np.ascontiguousarray(data, np.float64)
    rc = fn(degree, E, np.ascontiguousarray(xy, np.float64).ravel(), ev.ravel(), kind,
            float(const), None if d is None else d.ctypes.data, out)
    assert rc == 0, rc
    return out


def edge_matrices(xy, edge_v, degree, kind=G_CONST, const=1.0, data=None):
    """MassEdgeMatrixProvider::Eval for every edge (lfref_edge_matrices)."""
    n = degree + 1
    return _edge_call(lib().lfref_edge_matrices, degree, xy, edge_v, kind, const, data,
                      n * n).reshape(-1, n, n)


def edge_vectors(xy, edge_v, degree, kind=G_CONST, const=1.0, data=None):
    """ScalarLoadEdgeVectorProvider::Eval for every edge (lfref_edge_vectors)."""
    n = degree + 1
    return _edge_call(lib().lfref_edge_vectors, degree, xy, edge_v, kind, const, data,
                      n).reshape(-1, n)


def nonlocal_gain(xy, edge_v, degree, flags, rmin=0.0, rmax=0.0):
    """FisherKPP's gain h at the segment rule's points of every flagged edge
    ([nedges, nqp], zero elsewhere; lfref_nonlocal_gain)."""
    ev = np.ascontiguousarray(edge_v, np.int32)
    nq = degree + 1
    out = np.zeros(ev.shape[0] * nq)
    rc = lib().lfref_nonlocal_gain(degree, ev.shape[0], np.ascontiguousarray(xy, np.float64).ravel(),
                                   ev.ravel(), np.ascontiguousarray(flags, np.uint8),
                                   float(rmin), float(rmax), out)
    assert rc == 0, rc
    return out.reshape(-1, nq)


def nonlocal_loss(xy, edge_v, edge_dofs, degree, flags, rmin=0.0, rmax=0.0):
    """FisherKPP's loss matrix on the DOFs of the flagged edges: (rows, L)
    with L [nb, nb] (lfref_nonlocal_loss)."""
    ev = np.ascontiguousarray(edge_v, np.int32)
    ed = np.ascontiguousarray(edge_dofs, np.int64)
    fl = np.ascontiguousarray(flags, np.uint8)
    cap = int(len(np.unique(ed[fl != 0]))) if np.any(fl) else 0
    rows = np.zeros(max(cap, 1), np.int32)
    L = np.zeros(max(cap, 1) ** 2)
    nb = lib().lfref_nonlocal_loss(degree, ev.shape[0],
                                   np.ascontiguousarray(xy, np.float64).ravel(), ev.ravel(),
                                   ed.ravel(), fl, float(rmin), float(rmax), cap, rows, L)
    assert nb == cap, (nb, cap)
    return rows[:nb].copy(), L[:nb * nb].reshape(nb, nb)


def eta_sub(coef, sel):
    """(kind, const, data) of an edge coefficient restricted to the edges sel."""
    kind, const, data = coef
    if kind == G_EDGE:
        data = np.asarray(data)[sel]
    elif kind == G_QP:
        data = np.asarray(data)[sel]
    return kind, const, data


def _add_vector(b, edofs, evecs):
    """b[dof] += v in edge order (AssembleVectorLocally, assembler.h:298-327)."""
    b = b.copy()
    for e in range(edofs.shape[0]):
        for i in range(edofs.shape[1]):
            b[edofs[e, i]] += evecs[e, i]
    return b


def galerkin_system(mesh: TriMesh, degree: int, alpha, gamma, f_kind=F_CONST, f_const=1.0,
                    f_qp=None, flags=None, fixed_vals=None, edge_terms=None):
    """The homDir_linfe_demo.cc:130-170 pipeline: element matrices -> COO ->
    load vector -> FixFlaggedSolutionComponents -> setFromTriplets (CSR).
    edge_terms = (mat_flags, (kind, const, data), vec_flags, (kind, const,
    data)) adds MassEdgeMatrixProvider / ScalarLoadEdgeVectorProvider terms on
    the flagged edges (None flags: no term)."""
    cd, nd = mesh.dofs(degree)
    mats = element_matrices(mesh, degree, alpha, gamma)
    r, c, v = assemble_coo(cd, mats)
    vecs = element_vectors(mesh, degree, f_kind, f_const, f_qp)
    b = assemble_vector(cd, vecs, nd)
    if edge_terms is not None:
        # AssembleMatrixLocally(1, ...) and AssembleVectorLocally(1, ...) after
        # the cell terms (ellbvp_linfe_demo.cc:272-315): edge triplets follow
        # the cell triplets, edge loads are added after the cell loads
        ed = mesh.edge_dofs(degree)
        mat_on, eta, vec_on, h = edge_terms
        if mat_on is not None:
            sel = np.nonzero(mat_on)[0]
            em = edge_matrices(mesh.xy, mesh.edges[sel], degree, *eta_sub(eta, sel))
            re, ce, ve = assemble_coo(ed[sel], em)
            r, c, v = np.concatenate([r, re]), np.concatenate([c, ce]), np.concatenate([v, ve])
        if vec_on is not None:
            sel = np.nonzero(vec_on)[0]
            evv = edge_vectors(mesh.xy, mesh.edges[sel], degree, *eta_sub(h, sel))
            b = _add_vector(b, ed[sel], evv)
    if flags is None:
        flags = mesh.dirichlet_flags(degree)
    if fixed_vals is None:
        fixed_vals = np.zeros(nd)
    r, c, v, b = fix_flagged(r, c, v, flags, fixed_vals, b)
    rp, col, val = set_from_triplets(r, c, v, nd)
    return rp, col, val, b


# -- Fisher/KPP Strang splitting (projects/fisher_kpp) -------------------------
def fisher_kpp_matrices(mesh: TriMesh, c, h):
    """assembleGalerkinMatrices (projects/fisher_kpp/strangsplitting.h:123-174)
    with L = 0: A = P1 stiffness with diffusion coefficient c (alpha = c,
    gamma = 0) plus MassEdgeMatrixProvider(h) on the boundary edges, M = P1
    mass matrix (alpha = 0, gamma = 1), both as setFromTriplets CSR without
    Dirichlet rows.  c is a cell/point array or a constant; h = (kind, const,
    data) an edge coefficient.  Returns ((rp, col, val) of A, of M)."""
    nd = mesh.nv
    free = np.zeros(nd, np.uint8)
    _, ef = mesh.boundary_flags()
    rpA, colA, valA, _ = galerkin_system(mesh, 1, c, 0.0, flags=free,
                                         edge_terms=(ef, h, None, None))
    rpM, colM, valM, _ = galerkin_system(mesh, 1, 0.0, 1.0, flags=free)
    return (rpA, colA, valA), (rpM, colM, valM)


SDIRK2_KAPPA = 1.0 - 0.5 * np.sqrt(2.0)  # strangsplitting.h:194


class StrangSplit:
    """FisherKPP::StrangSplit (strangsplitting.h:28-76, :176-201;
    strangsplitting.cc:15-94) on scipy.sparse matrices, with SciPy's SuperLU
    (scipy.sparse.linalg.splu) in place of Eigen::SparseLU."""

    def __init__(self, A, M, T, m, lam):
        from scipy.sparse.linalg import splu
        self.A = A.tocsr()
        self.m = int(m)
        self.lam = float(lam)
        self.kappa = 1.0 - 0.5 * np.sqrt(2.0)
        self.tau = T / m
        # solver1 / solver2 of the constructor (strangsplitting.h:196-200)
        self.lu1 = splu((M + 0.5 * self.tau * self.kappa * A).tocsc())
        self.lu2 = splu((M + self.tau * self.kappa * A).tocsc())

    def diffusion(self, firstcall, mu):
        """diffusionEvolutionOperator (strangsplitting.cc:15-50)."""
        A, tau, kappa = self.A, self.tau, self.kappa
        rhs = -(A @ mu)
        if firstcall:
            k1 = self.lu1.solve(rhs)
            k2 = self.lu1.solve(rhs - 0.5 * tau * (1 - kappa) * (A @ k1))
            return mu + 0.5 * tau * (1 - kappa) * k1 + 0.5 * tau * kappa * k2
        k1 = self.lu2.solve(rhs)
        k2 = self.lu2.solve(rhs - tau * (1 - kappa) * (A @ k1))
        return mu + tau * (1 - kappa) * k1 + tau * kappa * k2

    @staticmethod
    def reaction(cap, u, e):
        """cap ./ (1 + (cap ./ u - 1) e) (strangsplitting.cc:75-77, :85-87)."""
        with np.errstate(divide="ignore", invalid="ignore"):
            return cap / (1.0 + (cap / u - 1.0) * e)

    def evolution(self, cap, mu):
        """Evolution (strangsplitting.cc:53-94): half a diffusion step, then
        for i = 1 .. m-2 a reaction over tau and a diffusion step, then a
        reaction over tau / 2."""
        import math
        cur = self.diffusion(True, mu)
        for _ in range(1, self.m - 1):
            cur = self.reaction(cap, cur, math.exp(-self.lam * self.tau))
            cur = self.diffusion(False, cur)
        return self.reaction(cap, cur, math.exp(-self.lam * self.tau / 2.))


def sdirk2_stability(z, kappa=SDIRK2_KAPPA):
    """R(z) = 1 + z b^T (I - z A_bt)^{-1} 1 of the SDIRK-2 Butcher tableau
    A_bt = [[kappa, 0], [1 - kappa, kappa]], b = (1 - kappa, kappa): the factor
    one step multiplies a mode of y' = (z / h) y by."""
    Abt = np.array([[kappa, 0.0], [1.0 - kappa, kappa]])
    b = np.array([1.0 - kappa, kappa])
    return 1.0 + z * b @ np.linalg.solve(np.eye(2) - z * Abt, np.ones(2))
