"""The CICSAM restatement (oracle cicsam_*, vof_properties) pinned by what the
mathematics fixes: hand-worked values of the HC and UQ limiters, face weights
on a rectilinear mesh where donor, acceptor, upwind value and Courant number
are known exactly, the weight of a uniform field, conservation of the
discretised divergence (every interior face enters its two cells with equal
and opposite flux), and closed forms of the mixture properties."""
import numpy as np
import pytest

from phase_amd import meshgen as G


def test_limiters_known_values(oracle):
    # hc: min(1, g / co) inside [0, 1], identity outside
    assert oracle.cicsam_hc(0.5, 0.25) == 1.0
    assert oracle.cicsam_hc(0.1, 0.5) == 0.2
    assert oracle.cicsam_hc(1.5, 0.3) == 1.5 and oracle.cicsam_hc(-0.25, 0.3) == -0.25
    # uq: min((8 co g + (1 - co)(6 g + 3)) / 8, hc) inside [0, 1]
    assert oracle.cicsam_uq(0.5, 0.25) == 0.6875  # (1 + 0.75 * 6) / 8
    assert oracle.cicsam_uq(0.1, 0.5) == 0.2      # hc (0.2) below (0.4 + 1.8) / 8 = 0.275
    assert oracle.cicsam_uq(2.0, 0.5) == 2.0


def _rect(oracle, nx=8, ny=4):
    xy, cptr, cind = G.rectilinear(nx, ny)
    return oracle.Mesh(xy, cptr, cind)


def _face_between(m, a, b):
    f = m.faces()
    hit = np.where(((f["lcell"] == a) & (f["rcell"] == b)) | ((f["lcell"] == b) & (f["rcell"] == a)))[0]
    assert len(hit) == 1
    return hit[0]


@pytest.mark.parametrize("dt, grad, expect", [
    # gradient along the face normal (psi = 1): beta = (hc - g~) / (1 - g~)
    (0.01, (4.0, 0.0), 1.0),    # Co = 0.08: hc = min(1, 0.5 / 0.08) = 1
    (0.1, (4.0, 0.0), 0.25),    # Co = 0.8:  hc = 0.625 -> (0.625 - 0.5) / 0.5
    # gradient at 45 degrees: psi = 1/2, uq = (0.32 + 0.92 * 6) / 8 = 0.73
    (0.01, (4.0, 4.0), 0.73),
])
def test_weights_known_values(oracle, dt, grad, expect):
    """Face between cells 3 and 4 of row 0 on an 8 x 4 unit-square mesh
    (h_x = 1/8, h_y = 1/4), u = (1, 0) everywhere: donor 3 (gamma 0.5),
    acceptor 4 (gamma 1), upwind value 1 - 2 h_x grad_x = 0, so g~_D = 0.5;
    the donor's only outflow face gives Co = u h_y / (h_x h_y) dt = 8 dt."""
    m = _rect(oracle)
    u = np.tile([1.0, 0.0], m.n_faces)
    gamma = np.zeros(m.n_cells)
    gamma[3], gamma[4] = 0.5, 1.0
    gradc = np.tile(grad, m.n_cells).astype(float)
    beta = oracle.cicsam_weights(m, u, gamma, gradc, dt)
    f = _face_between(m, 3, 4)
    assert abs(beta[f] - expect) <= 1e-12
    assert np.all(beta[m.faces()["rcell"] < 0] == 0.0)
    assert np.all((beta >= 0.0) & (beta <= 1.0))


def test_weights_uniform_field_is_upwind(oracle):
    xy, cptr, cind = G.delaunay_square(12, G.delaunay_cells_for(600, 12), seed=3)
    m = oracle.Mesh(xy, cptr, cind)
    rng = np.random.default_rng(0)
    u = rng.standard_normal(2 * m.n_faces)
    beta = oracle.cicsam_weights(m, u, np.full(m.n_cells, 0.3), np.zeros(2 * m.n_cells), 0.01)
    assert np.all(beta == 0.0)


@pytest.mark.parametrize("theta", [1.0, 0.5])
def test_div_conserves(oracle, theta):
    """Interior faces cancel between their two cells, so the sum over all rows
    of A gamma + rhs is the boundary flux alone (zero-gradient boundaries:
    F_b gamma of the boundary cell, implicit and explicit parts)."""
    xy, cptr, cind = G.delaunay_square(12, G.delaunay_cells_for(600, 12), seed=3)
    m = oracle.Mesh(xy, cptr, cind)
    rng = np.random.default_rng(1)
    u = rng.standard_normal(2 * m.n_faces)
    g0 = rng.random(m.n_cells)
    g = rng.random(m.n_cells)
    grad = rng.standard_normal(2 * m.n_cells)
    beta = oracle.cicsam_weights(m, u, g0, grad, 0.01)
    assert np.any((beta > 0) & (beta < 1))
    e = oracle.cicsam_div(m, u, beta, theta, gamma0_cell=g0)
    rp, ci, v, rhs = e.get()
    rows = np.repeat(np.arange(m.n_cells), np.diff(rp))
    keep = ci >= 0
    total = np.sum(v[keep] * g[ci[keep]]) + rhs.sum()
    f = m.faces()
    _, cc = m.cells()
    bnd = np.where(f["rcell"] < 0)[0]
    l = f["lcell"][bnd]
    sf = f["sf"][bnd]
    sign = np.where(np.sum((f["centroid"][bnd] - cc[l]) * sf, axis=1) > 0, 1.0, -1.0)
    fb = np.sum(u.reshape(-1, 2)[bnd] * sf, axis=1) * sign
    expect = np.sum(theta * fb * g[l] + (1 - theta) * fb * g0[l])
    scale = np.sum(np.abs(v[keep])) + np.abs(rhs).sum()
    assert abs(total - expect) <= 1e-13 * scale
    # each interior link puts its donor coefficient first (first touch)
    assert np.all(rows[keep] >= 0)


def test_momentum_flux_and_properties(oracle):
    m = _rect(oracle, 6, 6)
    rng = np.random.default_rng(2)
    u = rng.standard_normal(2 * m.n_faces)
    gc = rng.random(m.n_cells)
    gf = rng.random(m.n_faces)
    beta = np.zeros(m.n_faces)  # upwind: the donor's value
    rhou = oracle.cicsam_momentum_flux(m, 1000.0, 1.0, u, gc, gf, beta).reshape(-1, 2)
    f = m.faces()
    _, cc = m.cells()
    for k in range(m.n_faces):
        l, r = f["lcell"][k], f["rcell"][k]
        if r < 0:
            g = gf[k]
        else:
            n = f["sf"][k] * (1.0 if np.dot(f["centroid"][k] - cc[l], f["sf"][k]) > 0 else -1.0)
            g = gc[l] if np.dot(u.reshape(-1, 2)[k], n) > 0 else gc[r]
        assert np.allclose(rhou[k], (1000.0 + g * (1.0 - 1000.0)) * u.reshape(-1, 2)[k], rtol=1e-15, atol=0)
    g = np.array([0.0, 1.0, 0.5, -0.2, 1.3])
    m1 = _rect(oracle, 5, 1)
    rc, rf, mc, mf = oracle.vof_properties(m1, 1000.0, 1.0, 1e-2, 1e-4, g, np.zeros(m1.n_faces))
    gcl = np.clip(g, 0.0, 1.0)
    assert np.allclose(rc, 1000.0 + gcl * (1.0 - 1000.0), rtol=1e-15)
    # mu = rho / (rho1/mu1 + g (rho2/mu2 - rho1/mu1)): the pure phases give mu1 and mu2
    assert abs(mc[0] - 1e-2) <= 1e-17 and abs(mc[1] - 1e-4) <= 1e-19
    assert np.all(rf == 1000.0) and np.all(np.abs(mf - 1e-2) <= 1e-17)
