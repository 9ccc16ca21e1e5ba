"""Writes the golden fixtures in this folder (run in the build container).

Each fixture holds a mesh (nodes, cell->node CSR), per-face coefficients,
boundary types and values, a right-hand side and the solution of the
pressure-Poisson system A x = b computed by scipy's sparse LU (splu) -- the
exact-solve semantics of the reference's `lib eigen` backend
(src/Math/EigenSparseMatrixSolver.cpp:60-64: SparseLU.compute + solve).
A is assembled here with numpy from the textbook two-point flux
c_f = gamma_f * dot(d, S_f) / |d|^2, independently of the oracle, and the
faces are enumerated by scipy-free edge sets, so a fixture checks the
oracle's mesh, assembly and solve together.

Run: python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from phase_amd import meshgen as G  # noqa: E402


def textbook_system(xy, cptr, cind, gamma_of_edge, fixed_of_edge, phi_of_edge):
    n = len(cptr) - 1
    # cell area / centroid (shoelace)
    cen = np.zeros((n, 2))
    area = np.zeros(n)
    edges = {}
    for c in range(n):
        ids = cind[cptr[c]:cptr[c + 1]]
        p = xy[ids]
        q = np.roll(p, -1, axis=0)
        a = p[:, 0] * q[:, 1] - q[:, 0] * p[:, 1]
        A2 = a.sum()
        area[c] = A2 / 2
        cen[c] = [((p[:, 0] + q[:, 0]) * a).sum() / (3 * A2), ((p[:, 1] + q[:, 1]) * a).sum() / (3 * A2)]
        for k in range(len(ids)):
            e = (min(ids[k], ids[(k + 1) % len(ids)]), max(ids[k], ids[(k + 1) % len(ids)]))
            edges.setdefault(e, []).append(c)
    rows, cols, vals = [], [], []
    b_add = np.zeros(n)
    for (a, bnode), cells in edges.items():
        mid = 0.5 * (xy[a] + xy[bnode])
        t = xy[bnode] - xy[a]
        s = np.array([t[1], -t[0]])
        g = gamma_of_edge(mid)
        if len(cells) == 2:
            l, r = cells
            d = cen[r] - cen[l]
            if np.dot(mid - cen[l], s) < 0:
                s = -s
            c = g * np.dot(d, s) / np.dot(d, d)
            rows += [l, r, l, r]
            cols += [r, l, l, r]
            vals += [c, c, -c, -c]
        elif fixed_of_edge(mid):
            l = cells[0]
            d = mid - cen[l]
            if np.dot(d, s) < 0:
                s = -s
            c = g * np.dot(d, s) / np.dot(d, d)
            rows.append(l)
            cols.append(l)
            vals.append(-c)
            b_add[l] += c * phi_of_edge(mid)
    A = sp.csr_matrix((vals, (rows, cols)), shape=(n, n))
    return A, b_add


def make(name, xy, cptr, cind, sigma, seed):
    rng = np.random.default_rng(seed)
    # a smooth lognormal-like coefficient field evaluated at face midpoints
    kx, ky, ph = rng.uniform(1, 6, 3)

    def gamma(mid):
        return float(np.exp(sigma * np.sin(kx * mid[0] + ph) * np.cos(ky * mid[1])))

    def fixed(mid):
        return mid[1] == 0.0

    def phi(mid):
        return float(np.sin(3 * mid[0]))

    A, b_add = textbook_system(xy, cptr, cind, gamma, fixed, phi)
    src = rng.standard_normal(len(cptr) - 1)
    src -= src.mean()
    # laplacian(gamma, p) == src  ->  stored rhs = b_add - src ; A x = src - b_add
    b = src - b_add
    x = spla.splu(A.tocsc()).solve(b)
    np.savez_compressed(os.path.join(HERE, name + ".npz"), xy=xy, cptr=cptr, cind=cind, src=src,
                        coeff=np.array([sigma, kx, ky, ph]), x=x, b=b)
    print(name, len(cptr) - 1, "cells, |x| =", np.linalg.norm(x))


if __name__ == "__main__":
    make("rect8", *G.rectilinear(8, 8), sigma=0.0, seed=1)
    make("rect32", *G.rectilinear(32, 32), sigma=0.5, seed=2)
    make("cavity128", *G.rectilinear(128, 128), sigma=0.0, seed=3)
    make("tri2k", *G.delaunay_square(20, G.delaunay_cells_for(2000, 20), seed=5), sigma=1.0, seed=4)
