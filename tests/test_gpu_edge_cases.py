"""Edge cases of the device path against the oracle: a ragged mesh (quads,
triangles and six-sided cells with collinear nodes, so cell-to-node lists of
3, 4 and 6 and equation rows that grow past five entries); the smallest
meshes (one and two cells); a zero right-hand side."""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from phase_amd import meshgen as G

pytestmark = pytest.mark.gpu


def ragged_mesh(nx, ny, seed):
    """Rectilinear nodes; each quad kept, split into two triangles, or merged
    with its right neighbour into one six-node cell (counter-clockwise)."""
    xy, _, _ = G.rectilinear(nx, ny)
    rng = np.random.default_rng(seed)
    nnx = nx + 1
    cells = []
    for j in range(ny):
        i = 0
        while i < nx:
            a, b, c, d = j * nnx + i, j * nnx + i + 1, (j + 1) * nnx + i + 1, (j + 1) * nnx + i
            u = rng.random()
            if u < 0.15 and i + 1 < nx:
                cells.append([a, b, b + 1, c + 1, c, d])
                i += 2
                continue
            if u < 0.5:
                cells.extend([[a, b, c], [a, c, d]])
            else:
                cells.append([a, b, c, d])
            i += 1
    cptr = np.cumsum([0] + [len(c) for c in cells]).astype(np.int32)
    cind = np.concatenate(cells).astype(np.int32)
    return xy, cptr, cind


def _csr(rp, ci, v, n):
    rows = np.repeat(np.arange(n), np.diff(rp))
    keep = ci >= 0
    return sp.csr_matrix((v[keep], (rows[keep], ci[keep])), shape=(n, n))


@pytest.fixture(scope="module")
def ragged(oracle, hip):
    xy, cptr, cind = ragged_mesh(14, 11, seed=5)
    assert set(np.diff(cptr)) == {3, 4, 6}
    return xy, cptr, cind, oracle.Mesh(xy, cptr, cind), hip.HipMesh(xy, cptr, cind)


def test_ragged_mesh_topology_and_geometry(ragged):
    xy, cptr, cind, om, hm = ragged
    of, hf = om.faces(), hm.faces()
    for k in ("lnode", "rnode", "lcell", "rcell"):
        assert np.array_equal(of[k], hf[k])
    vol, cc, fc, sf = hm.geometry()
    ovol, occ = om.cells()
    assert np.array_equal(vol, ovol) and np.array_equal(cc, occ)
    assert np.array_equal(fc, of["centroid"]) and np.array_equal(sf, of["sf"])
    assert abs(vol.sum() - 1.0) < 1e-14


def test_ragged_rows_grow_and_solve(oracle, hip, ragged):
    xy, cptr, cind, om, hm = ragged
    N, F = om.n_cells, om.n_faces
    rng = np.random.default_rng(7)
    f = om.faces()
    bc = np.where(f["rcell"] < 0, oracle.BC_FIXED, oracle.BC_NORMAL_GRADIENT).astype(np.int32)
    g, pf, b = np.exp(rng.standard_normal(F)), rng.standard_normal(F), rng.standard_normal(N)
    oe = oracle.laplacian(om, 3, gamma_face=g, bc_type=bc, phi_face=pf)
    oe -= b
    he = hip.HipEquation(hm, 1).laplacian(3, gamma_face=hip.DeviceArray(g), bc_type=hip.DeviceArray(bc),
                                          phi_face=hip.DeviceArray(pf))
    he.sub_vector(hip.DeviceArray(b))
    rp, ci, v, rhs = he.export()
    orp, oci, ov, orhs = oe.get()
    assert np.array_equal(rp, orp) and np.array_equal(ci, oci) and np.array_equal(v, ov)
    assert np.array_equal(rhs, orhs)
    assert np.diff(orp).max() > 5  # the six-sided cells' rows grew
    xs = spla.spsolve(_csr(orp, oci, ov, N).tocsc(), -orhs)
    for mode, solver in (("fp64", "bicgstab"), ("mixed", "cg")):
        s = hip.HipSparseMatrixSolver(0, tolerance=1e-12, maxIters=5000, mode=mode, solver=solver)
        s.set_equation(he)
        s.solve()
        assert np.linalg.norm(s.x() - xs) <= 1e-8 * np.linalg.norm(xs)
    # ddt + upwind div merged: zero-dropped compact rows of every width
    u = rng.standard_normal(2 * F)
    p0 = rng.standard_normal(N)
    oe = oracle.ddt(om, 0.1, p0)
    oe += oracle.div(om, u, u, theta=1.0, bc_type=bc, phi_face=pf, phi0_cell=p0, phi0_face=pf)
    he = hip.HipEquation(hm, 1).ddt(0.1, hip.DeviceArray(p0))
    he += hip.HipEquation(hm, 1).div(hip.DeviceArray(u), hip.DeviceArray(u), theta=1.0,
                                     bc_type=hip.DeviceArray(bc), phi_face=hip.DeviceArray(pf),
                                     phi0_cell=hip.DeviceArray(p0), phi0_face=hip.DeviceArray(pf))
    got, ref = he.export(), oe.get()
    assert all(np.array_equal(x, y) for x, y in zip(got, ref))


@pytest.mark.parametrize("nx", [1, 2])
def test_smallest_meshes(oracle, hip, nx):
    """One cell (and two) with FIXED walls: A x = b against the dense solve."""
    xy, cptr, cind = G.rectilinear(nx, 1)
    om, hm = oracle.Mesh(xy, cptr, cind), hip.HipMesh(xy, cptr, cind)
    N, F = om.n_cells, om.n_faces
    bc = np.full(F, oracle.BC_FIXED, np.int32)
    pf = np.linspace(0.5, 1.5, F)
    oe = oracle.laplacian(om, 1, gamma=1.0, bc_type=bc, phi_face=pf)
    rp, ci, v, rhs = oe.get()
    xs = np.linalg.solve(_csr(rp, ci, v, N).toarray(), -rhs)
    he = hip.HipEquation(hm, 1).laplacian(1, gamma=1.0, bc_type=hip.DeviceArray(bc), phi_face=hip.DeviceArray(pf))
    s = hip.HipSparseMatrixSolver(0, tolerance=1e-12, maxIters=100)
    s.set_equation(he)
    s.solve()
    assert np.allclose(s.x(), xs, rtol=1e-12, atol=1e-14)


def test_zero_rhs_gives_zero(hip, ragged):
    xy, cptr, cind, om, hm = ragged
    F = om.n_faces
    f = om.faces()
    bc = np.where(f["rcell"] < 0, hip.BC_FIXED, hip.BC_NORMAL_GRADIENT).astype(np.int32)
    he = hip.HipEquation(hm, 1).laplacian(3, gamma_face=hip.DeviceArray(np.ones(F)), bc_type=hip.DeviceArray(bc),
                                          phi_face=hip.DeviceArray(np.zeros(F)))
    s = hip.HipSparseMatrixSolver(0, tolerance=1e-10, maxIters=100)
    s.set_equation(he)
    s.solve()
    assert np.all(s.x() == 0.0)
