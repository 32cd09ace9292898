"""Writes the golden fixtures this directory holds next to the reference's own
test vectors (reference_tests.json), from the CPU oracle:

* ``c1_poisson_64.npz``: the 64 x 64 P1 Poisson problem (alpha = 1,
  gamma = 0, f = 1, zero Dirichlet data), i.e. the homDir_linfe_demo.cc:
  114-175 pipeline.  Arrays: row_ptr, col, val (the CSR matrix after
  FixFlaggedSolutionComponents and setFromTriplets), rhs, flags (boundary
  DOFs) and x (the FP64 direct solution, banded Cholesky).
* ``cells_c2_c3.npz``: the first and the last 64 cells of two large meshes.
  For each of them it stores the cell index, the three vertex coordinates,
  theta, alpha = exp(theta), the element matrix
  (ReactionDiffusionElementMatrixProvider, gamma = 1) and the element vector
  (ScalarLoadElementVectorProvider, f = sin(pi x) sin(pi y)).  The meshes are:
  - ``c2_``: the 2048 x 2048 P1 mesh with theta = default_rng(0).standard_normal(Ne);
  - ``c3_``: the 512 x 512 P3 mesh, with interior vertices moved by
    default_rng(0).uniform(-0.2 h, 0.2 h) (both coordinates, vertex order)
    and theta = default_rng(1).standard_normal(Ne).

Run from the repository root after building the oracle (make -C oracle):

    python tests/golden/make_fixtures.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

from oracle import lfref as L  # noqa: E402

NCELLS = 64  # cells taken from each end of a mesh


def structured_arrays(n):
    """Vertex coordinates and cell -> vertex table of TPTriagMeshBuilder on
    the unit square (the oracle's builder, without the edge table)."""
    nv, ne = (n + 1) * (n + 1), 2 * n * n
    E = n * n + 2 * n * (n + 1)
    xy = np.zeros(2 * nv)
    cells = np.zeros(3 * ne, np.int32)
    edges = np.zeros(2 * E, np.int32)
    L.lib().lfref_tp_triag_mesh(n, n, 0.0, 0.0, 1.0, 1.0, xy, cells, edges)
    return xy.reshape(nv, 2), cells.reshape(ne, 3)


def jitter(xy, n, seed, amount=0.2):
    """Interior vertices moved by U(-amount h, amount h) in each coordinate."""
    xy = xy.copy()
    h = 1.0 / n
    inner = (xy[:, 0] > 0) & (xy[:, 0] < 1) & (xy[:, 1] > 0) & (xy[:, 1] < 1)
    xy[inner] += np.random.default_rng(seed).uniform(-amount * h, amount * h,
                                                     size=(int(inner.sum()), 2))
    return xy


def end_cells(ne, k=NCELLS):
    return np.concatenate([np.arange(k), np.arange(ne - k, ne)]).astype(np.int64)


def cell_mesh(verts):
    """A mesh of separate cells with the given vertex triples."""
    m = verts.shape[0]
    return L.TriMesh(verts.reshape(3 * m, 2), np.arange(3 * m, dtype=np.int32).reshape(m, 3))


def cells_fixture(n, degree, theta_seed, jitter_seed=None):
    xy, cells = structured_arrays(n)
    if jitter_seed is not None:
        xy = jitter(xy, n, jitter_seed)
    theta_all = np.random.default_rng(theta_seed).standard_normal(cells.shape[0])
    sel = end_cells(cells.shape[0])
    verts = np.ascontiguousarray(xy[cells[sel]])
    theta = theta_all[sel]
    alpha = np.exp(theta)
    sub = cell_mesh(verts)
    return {"cells": sel, "verts": verts, "theta": theta, "alpha": alpha,
            "mats": L.element_matrices(sub, degree, alpha, 1.0),
            "vecs": L.element_vectors(sub, degree, L.F_SINSIN)}


def c1_fixture(n=64):
    m = L.tp_triag_mesh(n)
    flags = m.dirichlet_flags(1)
    rp, col, val, b = L.galerkin_system(m, 1, 1.0, 0.0, L.F_CONST, 1.0, flags=flags)
    x = L.band_cholesky_solve(rp, col, val, b)
    return {"row_ptr": rp, "col": col, "val": val, "rhs": b, "flags": flags, "x": x}


def main():
    np.savez_compressed(os.path.join(HERE, "c1_poisson_64.npz"), **c1_fixture())
    out = {}
    for prefix, (n, degree, ts, js) in {"c2_": (2048, 1, 0, None), "c3_": (512, 3, 1, 0)}.items():
        for k, v in cells_fixture(n, degree, ts, js).items():
            out[prefix + k] = v
    np.savez_compressed(os.path.join(HERE, "cells_c2_c3.npz"), **out)


if __name__ == "__main__":
    main()
