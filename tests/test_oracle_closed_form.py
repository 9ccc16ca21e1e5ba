"""The oracle pinned by what the mathematics fixes (no reference golden
vectors exist: the reference's own tests check inline arithmetic only).

* Dyadic rectilinear meshes (h = 2^-k): areas, centroids, normals and every
  laplacian coefficient are exact binary numbers (interior c = gamma, FIXED
  boundary c = 2 gamma on a square cell); checked against those closed forms.
* Unstructured meshes: geometric invariants (sum of outward normals per cell
  vanishes, areas sum to the domain, Sum_f dot(c_f, S_f) = 2V), symmetric
  pressure Laplacian, zero row sums on Neumann rows, diag = -sum offdiag -
  sum FIXED coefficients, and a dense brute-force assembly written here from
  the textbook two-point flux.
* Delaunay topology against scipy.spatial (face count from Euler's formula,
  every triangle edge a face).
"""
import numpy as np
import pytest

from phase_amd import meshgen as G
from tests import helpers as H


def test_rectilinear_geometry_closed_form(oracle):
    nx = ny = 8
    h = 1.0 / 8
    xy, cptr, cind = G.rectilinear(nx, ny)
    m = oracle.Mesh(xy, cptr, cind)
    assert m.n_cells == 64 and m.n_faces == 2 * 8 * 9 and m.n_interior == 2 * 8 * 7
    vol, cen = m.cells()
    assert np.all(vol == h * h)
    i, j = np.arange(64) % 8, np.arange(64) // 8
    assert np.array_equal(cen[:, 0], (i + 0.5) * h) and np.array_equal(cen[:, 1], (j + 0.5) * h)
    f = m.faces()
    assert np.all(np.abs(f["sf"]).sum(axis=1) == h)
    # first face of cell 0 is its bottom edge (nodes 0->1), normal (0,-h)
    assert f["lnode"][0] == 0 and f["rnode"][0] == 1 and tuple(f["sf"][0]) == (0.0, -h)


@pytest.mark.parametrize("variant", [1, 3])
def test_rectilinear_laplacian_closed_form(oracle, variant):
    nx, ny = 16, 8
    xy, cptr, cind = G.rectilinear(nx, ny, 1.0, 0.5)   # square cells h = 1/16
    m = oracle.Mesh(xy, cptr, cind)
    bc, pid = H.bc_from_patch(oracle, m, xy, fixed_patches=(2,))
    gamma = 0.75
    gf = np.full(m.n_faces, gamma)
    phi_f = np.where(pid == 2, 3.0, 0.0)
    e = oracle.laplacian(m, variant, gamma=gamma, gamma_face=gf, bc_type=bc, phi_face=phi_f)
    rp, ci, v, rhs = e.get()
    A = oracle.to_dense(rp, ci, v)
    n = nx * ny
    for c in range(n):
        i, j = c % nx, c // nx
        nbs = [(i - 1, j), (i + 1, j), (i, j - 1), (i, j + 1)]
        expect_diag = 0.0
        for (a, b) in nbs:
            if 0 <= a < nx and 0 <= b < ny:
                assert A[c, b * nx + a] == gamma
                expect_diag -= gamma
        if j == 0:  # FIXED y-: c = gamma * (h/2 * h) / (h/2)^2 = 2 gamma
            expect_diag -= 2 * gamma
            assert rhs[c] == 2 * gamma * 3.0
        else:
            assert rhs[c] == 0.0
        assert A[c, c] == expect_diag
    # padded layout: 5 slots per row, first-touch column order
    assert np.array_equal(rp, 5 * np.arange(n + 1))
    row0 = ci[0:5]
    if variant == 1:   # (P,nb) before (P,P): neighbour right (face 1), then P
        assert list(row0) == [1, 0, nx, -1, -1]
    else:              # (P,P) first
        assert list(row0) == [0, 1, nx, -1, -1]


def test_unstructured_invariants(oracle):
    xy, cptr, cind = H.delaunay_case(4000, 24, seed=2)
    m = oracle.Mesh(xy, cptr, cind)
    vol, cen = m.cells()
    f = m.faces()
    L = m.links()
    assert abs(vol.sum() - 1.0) < 1e-12
    assert np.all(vol > 0)
    sx = np.zeros(m.n_cells)
    sy = np.zeros(m.n_cells)
    div_x = np.zeros(m.n_cells)
    for c in range(m.n_cells):
        faces = list(L["iface"][L["iptr"][c]:L["iptr"][c + 1]]) + list(L["bface"][L["bptr"][c]:L["bptr"][c + 1]])
        for fi in faces:
            n = f["sf"][fi]
            if np.dot(f["centroid"][fi] - cen[c], n) <= 0:
                n = -n
            sx[c] += n[0]
            sy[c] += n[1]
            div_x[c] += np.dot(f["centroid"][fi], n)
    assert np.max(np.abs(sx)) < 1e-15 and np.max(np.abs(sy)) < 1e-15
    assert np.max(np.abs(div_x - 2 * vol)) < 1e-15
    # links: interior links ascending face id then boundary links ascending
    for c in range(0, m.n_cells, 97):
        fi = L["iface"][L["iptr"][c]:L["iptr"][c + 1]]
        assert np.all(np.diff(fi) > 0)
    # Euler: F = 3T/2 + h/2 for a triangulation with h boundary edges
    assert 2 * m.n_faces == 3 * m.n_cells + m.n_boundary


def test_unstructured_laplacian_brute_force(oracle):
    xy, cptr, cind = H.delaunay_case(600, 10, seed=4)
    m = oracle.Mesh(xy, cptr, cind)
    bc, pid = H.bc_from_patch(oracle, m, xy, fixed_patches=(2, 3))
    rng = np.random.default_rng(0)
    gf = np.exp(rng.standard_normal(m.n_faces))
    phi_f = rng.standard_normal(m.n_faces)
    e = oracle.laplacian(m, 3, gamma_face=gf, bc_type=bc, phi_face=phi_f)
    rp, ci, v, rhs = e.get()
    A = oracle.to_dense(rp, ci, v)
    # textbook two-point flux, written independently of the oracle
    vol, cen = m.cells()
    f = m.faces()
    B = np.zeros_like(A)
    r = np.zeros(m.n_cells)
    for fi in range(m.n_faces):
        l, rr = f["lcell"][fi], f["rcell"][fi]
        n = f["sf"][fi]
        if rr >= 0:
            d = cen[rr] - cen[l]
            nl = n if np.dot(f["centroid"][fi] - cen[l], n) > 0 else -n
            c = gf[fi] * np.dot(d, nl) / np.dot(d, d)
            B[l, rr] += c
            B[rr, l] += c
            B[l, l] -= c
            B[rr, rr] -= c
        elif bc[fi] == oracle.BC_FIXED:
            d = f["centroid"][fi] - cen[l]
            nl = n if np.dot(d, n) > 0 else -n
            c = gf[fi] * np.dot(d, nl) / np.dot(d, d)
            B[l, l] -= c
            r[l] += c * phi_f[fi]
    assert np.allclose(A, B, rtol=1e-13, atol=1e-13)
    assert np.allclose(rhs, r, rtol=1e-13, atol=1e-13)
    assert np.array_equal(A, A.T)                     # symmetric, bitwise (same coefficient both sides)
    neumann = np.ones(m.n_cells, bool)
    neumann[f["lcell"][bc == oracle.BC_FIXED]] = False
    assert np.max(np.abs(A[neumann].sum(axis=1))) < 1e-12


def test_delaunay_topology_matches_scipy(oracle):
    from scipy.spatial import Delaunay

    xy, cptr, cind = H.delaunay_case(2000, 16, seed=6)
    tri = Delaunay(xy)
    assert len(cptr) - 1 == len(tri.simplices)
    assert len(cptr) - 1 == 2 * len(xy) - 2 - 4 * 16
    m = oracle.Mesh(xy, cptr, cind)
    edges = set()
    for t in tri.simplices:
        for a, b in ((t[0], t[1]), (t[1], t[2]), (t[2], t[0])):
            edges.add((min(a, b), max(a, b)))
    f = m.faces()
    got = set(zip(np.minimum(f["lnode"], f["rnode"]), np.maximum(f["lnode"], f["rnode"])))
    assert got == edges


def test_vector_laplacian_symmetry_bc(oracle):
    """Vector2D laplacian with SYMMETRY walls: outer(tw,tw) on the diagonal
    block (Laplacian.cpp:36-43); on an axis-aligned wall tw = (+-1, 0) or
    (0, +-1) so only xx or yy picks up theta*c, the cross terms are dropped
    because they are exactly zero."""
    xy, cptr, cind = G.rectilinear(4, 4)
    m = oracle.Mesh(xy, cptr, cind)
    bc, pid = H.bc_from_patch(oracle, m, xy, fixed_patches=())
    bc[pid >= 0] = oracle.BC_SYMMETRY
    e = oracle.laplacian(m, 0, ncomp=2, gamma=1.0, theta=1.0, bc_type=bc,
                         phi0_cell=np.zeros(2 * 16), phi0_face=np.zeros(2 * m.n_faces))
    rp, ci, v, rhs = e.get()
    A = oracle.to_dense(rp, ci, v)
    n = 16
    # interior c = 1, wall c = 2.  Cell 0: two interior links, x- wall
    # (tw = (0,-1): +2 on yy) and y- wall (tw = (1,0): +2 on xx).
    assert A[0, 0] == -2.0 - 4.0 + 2.0 and A[n, n] == -2.0 - 4.0 + 2.0
    # cell 1: three interior links and the y- wall only -> xx gains, yy does not
    assert A[1, 1] == -3.0 - 2.0 + 2.0 and A[n + 1, n + 1] == -3.0 - 2.0
    assert A[0, n] == 0.0 and A[n, 0] == 0.0
    assert not np.any(A[:n, n:]) and not np.any(A[n:, :n])


def test_ddt_div_src_closed_form(oracle):
    xy, cptr, cind = G.rectilinear(4, 4)
    m = oracle.Mesh(xy, cptr, cind)
    vol, cen = m.cells()
    phi0 = np.arange(16, dtype=float)
    e = oracle.ddt(m, 0.5, phi0)
    rp, ci, v, rhs = e.get()
    A = oracle.to_dense(rp, ci, v)
    assert np.array_equal(np.diag(A), vol / 0.5) and np.array_equal(rhs, -vol * phi0 / 0.5)
    # uniform u = (1, 0): src::div = 0 in every cell (closed surfaces)
    uf = np.tile([1.0, 0.0], m.n_faces)
    assert np.max(np.abs(oracle.src_div(m, uf))) == 0.0
    # upwind div with theta=1: interior flux |Sf| = h; outflow face goes to diag
    bc = np.full(m.n_faces, oracle.BC_NORMAL_GRADIENT, np.int32)
    ed = oracle.div(m, uf, theta=1.0, bc_type=bc)
    D = oracle.to_dense(*ed.get()[:3])
    h = 0.25
    assert np.allclose(D.sum(axis=1), 0.0)  # sum of outward fluxes of a divergence-free u
    # cell 0: outflow h through its right face, inflow -h through the x- wall
    # (NORMAL_GRADIENT puts theta*flux on the diagonal); cell 1: upwind -h to cell 0
    assert D[0, 0] == 0.0 and D[1, 1] == h and D[1, 0] == -h and D[0, 1] == 0.0
    # theta = 0: every implicit coefficient is +-0 and operator+= drops them all
    e0 = oracle.div(m, uf, theta=0.0, bc_type=bc)
    z = oracle.Eqn.create(16)
    z += e0
    assert np.all(z.get()[1] >= -1) and z.get()[0][-1] == 0


def test_gradient_linear_field_exact(oracle):
    """Face gradients of a linear field are exact along rc; on a rectilinear
    mesh FACE_TO_CELL recovers grad(phi) = (2, -3) up to rounding."""
    xy, cptr, cind = G.rectilinear(8, 8)
    m = oracle.Mesh(xy, cptr, cind)
    vol, cen = m.cells()
    f = m.faces()
    phi = 2 * cen[:, 0] - 3 * cen[:, 1]
    phif = 2 * f["centroid"][:, 0] - 3 * f["centroid"][:, 1]
    gf, gc = oracle.grad(m, phi, phif)
    assert np.allclose(gc, [[2.0, -3.0]] * 64, atol=1e-12)
