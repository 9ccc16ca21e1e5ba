"""Shared problem builders for the tests (CPU side, numpy + oracle)."""
import numpy as np

from phase_amd import meshgen as G


def rect_case(nx, ny):
    xy, cptr, cind = G.rectilinear(nx, ny)
    return xy, cptr, cind


def delaunay_case(n_cells, k, seed=1):
    ni = G.delaunay_cells_for(n_cells, k)
    return G.delaunay_square(k, ni, seed=seed)


def bc_from_patch(O, mesh, xy, fixed_patches=(2,)):
    f = mesh.faces()
    pid = G.boundary_patch(f["lnode"], f["rnode"], f["rcell"], xy)
    bc = np.full(mesh.n_faces, O.BC_NORMAL_GRADIENT, np.int32)
    for p in fixed_patches:
        bc[pid == p] = O.BC_FIXED
    return bc, pid


def pressure_system(O, xy, cptr, cind, sigma=1.0, fixed_patches=(2,), seed_g=2, seed_b=3):
    """Lognormal-coefficient pressure Poisson (field-gamma laplacian, pCorr=0
    on the fixed patches), returned as (mesh, padded CSR, rhs b, gamma, bc)."""
    m = O.Mesh(xy, cptr, cind)
    bc, _ = bc_from_patch(O, m, xy, fixed_patches)
    g = G.lognormal_face_coeffs(m.n_faces, sigma, seed=seed_g)
    e = O.laplacian(m, 3, gamma_face=g, bc_type=bc, phi_face=np.zeros(m.n_faces))
    rp, ci, v, rhs = e.get()
    b = G.zero_mean_normal(m.n_cells, seed=seed_b)
    return m, (rp, ci, v), b, g, bc


def compact_sorted(rp, ci, v):
    """Drop -1 slots and sort each row by column (scipy CSR)."""
    import scipy.sparse as sp

    n = len(rp) - 1
    rows = np.repeat(np.arange(n), np.diff(rp))
    keep = ci >= 0
    A = sp.csr_matrix((v[keep], (rows[keep], ci[keep])), shape=(n, n))
    A.sort_indices()
    return A


def permuted(A, perm):
    P = A[perm][:, perm].tocsr()
    P.sort_indices()
    return P
