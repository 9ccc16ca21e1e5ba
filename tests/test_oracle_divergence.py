"""The remaining convection operators of the oracle pinned by closed forms:
fv::divc (central) is conservative and reproduces src::div on a uniform field;
fv::dive (explicit) is theta c0 div(u0) + (1 - theta) c1 div(u1) on uniform
old fields; fv::div(phiU, gradU, u) with the upwind cell's gradient is exact
for a linear field (each row equals sum_f F_f u(x_f))."""
import numpy as np
import pytest

from phase_amd import meshgen as G


@pytest.fixture(scope="module")
def tri(oracle):
    xy, cptr, cind = G.delaunay_square(12, G.delaunay_cells_for(500, 12), seed=8)
    return oracle.Mesh(xy, cptr, cind)


def _boundary_flux(m, u):
    """Outward flux per boundary face and its lCell."""
    f = m.faces()
    _, cc = m.cells()
    bnd = np.where(f["rcell"] < 0)[0]
    l = f["lcell"][bnd]
    sf = f["sf"][bnd]
    sign = np.where(np.sum((f["centroid"][bnd] - cc[l]) * sf, axis=1) > 0, 1.0, -1.0)
    return bnd, l, np.sum(u.reshape(-1, 2)[bnd] * sf, axis=1) * sign


def _apply(e, x):
    rp, ci, v, rhs = e.get()
    rows = np.repeat(np.arange(len(rp) - 1), np.diff(rp))
    keep = ci >= 0
    y = np.zeros(len(rp) - 1)
    np.add.at(y, rows[keep], v[keep] * x[ci[keep]])
    return y, rhs


def test_divc_conserves_and_uniform(oracle, tri):
    m = tri
    rng = np.random.default_rng(1)
    u = rng.standard_normal(2 * m.n_faces)
    phi = rng.standard_normal(m.n_cells)
    y, rhs = _apply(oracle.divc(m, u, theta=1.0), phi)
    bnd, l, fb = _boundary_flux(m, u)
    scale = np.abs(y).sum()
    assert abs(y.sum() + rhs.sum() - np.sum(fb * phi[l])) <= 1e-13 * scale
    y1, _ = _apply(oracle.divc(m, u, theta=1.0), np.full(m.n_cells, 2.5))
    assert np.allclose(y1, 2.5 * oracle.src_div(m, u), rtol=0, atol=1e-13 * np.abs(u).max())


def test_dive_uniform_old_fields(oracle, tri):
    m = tri
    rng = np.random.default_rng(2)
    u0, u1 = rng.standard_normal(2 * m.n_faces), rng.standard_normal(2 * m.n_faces)
    N, F = m.n_cells, m.n_faces
    theta = 0.3
    f = m.faces()
    bc = np.where(f["rcell"] < 0, oracle.BC_FIXED, oracle.BC_NORMAL_GRADIENT).astype(np.int32)
    e = oracle.dive(m, u0, u1, theta, np.full(N, 2.0), np.full(N, -3.0), np.full(F, 2.0), np.full(F, -3.0),
                    bc_type=bc)
    rp, ci, v, rhs = e.get()
    assert np.all(ci == -1)  # sources only
    expect = theta * 2.0 * oracle.src_div(m, u0) + (1 - theta) * -3.0 * oracle.src_div(m, u1)
    assert np.allclose(rhs, expect, rtol=0, atol=1e-13 * (np.abs(u0).max() + np.abs(u1).max()))


@pytest.mark.parametrize("fixed", [True, False])
def test_div_grad_exact_for_linear_fields(oracle, tri, fixed):
    m = tri
    _, cc = m.cells()
    f = m.faces()
    a = np.array([0.3, -1.1])
    B = np.array([[1.7, -0.4], [2.3, 0.9]])  # not symmetric: a transposed dot would show
    lin = lambda x: a + x @ B.T
    rng = np.random.default_rng(3)
    phiu = rng.standard_normal(2 * m.n_faces)
    grad = np.tile(B.reshape(-1), m.n_cells)
    bc = np.full(m.n_faces, oracle.BC_FIXED if fixed else oracle.BC_NORMAL_GRADIENT, np.int32)
    uf = lin(f["centroid"]).reshape(-1)
    e = oracle.div_grad(m, phiu, grad, bc_type=bc, u_face=uf)
    N = m.n_cells
    ucell = lin(cc)
    x = np.r_[ucell[:, 0], ucell[:, 1]]  # IndexMap block layout
    y, rhs = _apply(e, x)
    # sum over every face of the cell of F_f u(x_f)
    expect = np.zeros((N, 2))
    n_out = lambda k, c: f["sf"][k] * (1.0 if np.dot(f["centroid"][k] - cc[c], f["sf"][k]) > 0 else -1.0)
    for k in range(m.n_faces):
        for c in (f["lcell"][k], f["rcell"][k]):
            if c >= 0:
                expect[c] += np.dot(phiu.reshape(-1, 2)[k], n_out(k, c)) * lin(f["centroid"][k])
    got = (y + rhs).reshape(2, N).T
    assert np.abs(got - expect).max() <= 1e-12 * np.abs(expect).max()


def test_ddt_density_forms(oracle):
    """fv::ddt with rho = 1 is the plain form bit for bit; a field density puts
    rho on the diagonal and rho0 in the source (dyadic values: exact); a cell
    group leaves the other rows empty."""
    xy, cptr, cind = G.rectilinear(8, 8)
    m = oracle.Mesh(xy, cptr, cind)
    rng = np.random.default_rng(4)
    phi0 = rng.standard_normal(2 * m.n_cells)
    a = oracle.ddt(m, 0.5, phi0, ncomp=2).get()
    b = oracle.ddt_rho(m, 0.5, phi0, rho=1.0, ncomp=2).get()
    assert all(np.array_equal(x, y) for x, y in zip(a, b))
    rho = np.full(m.n_cells, 2.0)
    rho0 = np.full(m.n_cells, 4.0)
    phi = np.full(m.n_cells, 0.25)
    rp, ci, v, rhs = oracle.ddt_rho(m, 0.5, phi, rho_cell=rho, rho0_cell=rho0).get()
    diag = v[rp[:-1]]
    assert np.all(ci[rp[:-1]] == np.arange(m.n_cells))
    assert np.all(diag == 2.0 * (1 / 64) / 0.5) and np.all(rhs == -4.0 * (1 / 64) * 0.25 / 0.5)
    mask = (np.arange(m.n_cells) % 3 == 0).astype(np.int32)
    rp, ci, v, rhs = oracle.ddt_rho(m, 0.5, phi, rho=3.0, cell_mask=mask).get()
    rows = np.repeat(np.arange(m.n_cells), np.diff(rp))
    assert set(rows[ci >= 0]) == set(np.where(mask == 1)[0])
    assert np.all(rhs[mask == 0] == 0.0)


def test_face_to_cell_linear_field(oracle):
    """faceToCell on a uniform rectilinear mesh with dx != dy: the x component
    sees only the east and west faces (|S_f|.x), the y component only north and
    south, so a linear vector field's face values average to its centroid value
    (different components, so a swapped |S| component shows); the weighted form
    with constant weights is cw / fw times the plain one (dyadic: exact)."""
    xy, cptr, cind = G.rectilinear(8, 4, width=1.0, height=2.0)
    m = oracle.Mesh(xy, cptr, cind)
    f = m.faces()
    _, cc = m.cells()
    B = np.array([[1.5, -0.75], [0.25, 2.0]])
    lin = lambda x: np.array([0.5, -1.0]) + x @ B.T
    fv = lin(f["centroid"]).reshape(-1)
    got = oracle.face_to_cell(m, fv)
    assert np.abs(got - lin(cc)).max() <= 1e-14
    w = oracle.face_to_cell(m, fv, np.full(m.n_cells, 4.0), np.full(m.n_faces, 2.0))
    assert np.array_equal(w, 2.0 * got)
    with pytest.raises(ValueError):
        oracle.face_to_cell(m, fv, np.ones(m.n_cells), None)


def test_face_to_cell_weighted_brute_force(oracle):
    """On an unstructured mesh: a face loop that scatters each face to both of
    its cells (the oracle gathers per cell) gives the same result."""
    xy, cptr, cind = G.delaunay_square(10, G.delaunay_cells_for(300, 10), seed=5)
    m = oracle.Mesh(xy, cptr, cind)
    f = m.faces()
    rng = np.random.default_rng(6)
    N, F = m.n_cells, m.n_faces
    fv = rng.standard_normal((F, 2))
    cw, fw = np.exp(rng.standard_normal(N)), np.exp(rng.standard_normal(F))
    t = np.zeros((N, 2))
    s = np.zeros((N, 2))
    for k in range(F):
        a = np.abs(f["sf"][k])
        for c in (f["lcell"][k], f["rcell"][k]):
            if c >= 0:
                t[c] += fv[k] * a / fw[k]
                s[c] += a
    expect = cw[:, None] * t / s
    got = oracle.face_to_cell(m, fv.reshape(-1), cw, fw)
    assert np.abs(got - expect).max() <= 1e-13 * np.abs(expect).max()
