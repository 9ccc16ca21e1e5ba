"""The CPU restatement of the CELESTE surface-tension model
(oracle/celeste_oracle.py) pinned by what the mathematics fixes: the
five-column least-squares stencils are exact for quadratic fields (gradient,
weighted; divergence, unweighted), so a dropped column, a swapped row scale or
a transposed read fails; the normalised smoothing kernel keeps a constant and,
by symmetry, a linear field; the curvature of a disc with a smooth edge is 1 / R
up to the discretisation; a 90 degree contact angle makes the boundary normal
tangential to the wall; the diagonal links of a structured quad mesh are the
four corner cells."""
import numpy as np
import pytest

from phase_amd import meshgen as G


@pytest.fixture(scope="module")
def co():
    from oracle.celeste_oracle import CelesteOracle
    return CelesteOracle


def _quad(x):
    return 0.3 + 1.1 * x[..., 0] - 0.7 * x[..., 1] + 0.9 * x[..., 0] ** 2 - 1.3 * x[..., 0] * x[..., 1] + 0.4 * x[..., 1] ** 2


def _quad_grad(x):
    return np.stack([1.1 + 1.8 * x[..., 0] - 1.3 * x[..., 1], -0.7 - 1.3 * x[..., 0] + 0.8 * x[..., 1]], -1)


@pytest.mark.parametrize("kind", ["rect", "tri"])
def test_stencils_exact_for_quadratics(oracle, co, kind):
    if kind == "rect":
        xy, cptr, cind = G.rectilinear(10, 7, width=1.0, height=0.8)
    else:
        xy, cptr, cind = G.delaunay_square(8, G.delaunay_cells_for(200, 8), seed=3)
    m = oracle.Mesh(xy, cptr, cind)
    c = co(m, cptr, cind, kernel_width=0.2)
    f = m.faces()
    _, cc = m.cells()
    g = c.grad(_quad(cc), _quad(f["centroid"]))
    assert np.abs(g - _quad_grad(cc)).max() <= 1e-9
    # u = (q(x), q(x + shift)): div u = dq/dx + dq/dy (shifted)
    shift = np.array([0.2, -0.1])
    uc = np.stack([_quad(cc), _quad(cc + shift)], -1)
    uf = np.stack([_quad(f["centroid"]), _quad(f["centroid"] + shift)], -1)
    d = c.div(uc, uf)
    expect = _quad_grad(cc)[:, 0] + _quad_grad(cc + shift)[:, 1]
    assert np.abs(d - expect).max() <= 1e-9


def test_diagonals_of_a_quad_mesh(oracle, co):
    xy, cptr, cind = G.rectilinear(5, 5)
    m = oracle.Mesh(xy, cptr, cind)
    c = co(m, cptr, cind, kernel_width=0.3)
    centre = 2 * 5 + 2
    assert sorted(c.diag[centre]) == sorted([centre - 6, centre - 4, centre + 4, centre + 6])
    assert len(c.diag[0]) == 1


def test_smoothing_keeps_constant_and_linear(oracle, co):
    n = 24
    xy, cptr, cind = G.rectilinear(n, n)
    m = oracle.Mesh(xy, cptr, cind)
    h = 1.0 / n
    for kind in ("pow8", "pow6", "peskin"):
        c = co(m, cptr, cind, kernel_width=3.1 * h, kernel_type=kind)
        _, cc = m.cells()
        gt, gtf = c.smooth(np.full(m.n_cells, 0.75))
        assert np.abs(gt - 0.75).max() <= 1e-14
        lin = 0.2 + 0.5 * cc[:, 0] - 0.3 * cc[:, 1]
        gt, gtf = c.smooth(lin)
        inner = (cc.min(1) > 3.5 * h) & (cc.max(1) < 1 - 3.5 * h)
        assert np.abs(gt - lin)[inner].max() <= 1e-13
        f = m.faces()
        b = f["rcell"] < 0
        assert np.array_equal(gtf[b], gt[f["lcell"][b]])


def test_disc_curvature_and_contact_angle(oracle, co):
    n = 64
    xy, cptr, cind = G.rectilinear(n, n)
    m = oracle.Mesh(xy, cptr, cind)
    F = m.n_faces
    _, cc = m.cells()
    f = m.faces()
    R = 0.3
    rr = np.hypot(cc[:, 0] - 0.5, cc[:, 1] - 0.45)
    gamma = 0.5 * (1.0 - np.tanh((rr - R) / (1.5 / n)))  # a disc with a smooth edge
    c = co(m, cptr, cind, kernel_width=2.5 / n)
    gf = oracle.interpolate_faces(m, gamma, np.zeros(F))
    gface, _ = oracle.grad(m, gamma, gf)
    ones, zeros = np.ones(F), np.zeros(F)
    th = np.full(F, np.pi / 2)
    r = c.face_forces(gamma, gface, 0.07, np.cos(th), np.sin(th), zeros)
    band = np.abs(r["gamma_tilde"] - 0.5) < 0.25
    assert band.sum() > 50
    kap = r["kappa_cell"][band]
    assert np.all(kap != 0)
    assert abs(np.median(kap) * R - 1.0) < 0.02
    assert np.all(np.abs(np.percentile(kap * R, [10, 90]) - 1.0) < 0.08)
    # fst = sigma kappa_f grad gamma_f, pointing into the disc
    assert np.array_equal(r["fst_face"], (0.07 * r["kappa_face"])[:, None] * gface)
    # the normal is the outward unit normal of the disc near the interface
    nC = r["n_cell"][band]
    radial = cc[band] - np.array([0.5, 0.45])
    radial /= np.linalg.norm(radial, axis=1)[:, None]
    assert np.median(np.sum(nC * radial, 1)) > 0.99
    # contact angle 90 degrees: boundary normals (where set) lie along the wall
    gamma2 = (np.hypot(cc[:, 0] - 0.5, cc[:, 1]) < R).astype(float)  # a drop on the wall y = 0
    gf2 = oracle.interpolate_faces(m, gamma2, np.zeros(F))
    g2, _ = oracle.grad(m, gamma2, gf2)
    r2 = c.face_forces(gamma2, g2, 0.07, np.cos(th), np.sin(th), zeros)
    bot = np.where((f["rcell"] < 0) & (np.abs(f["centroid"][:, 1]) < 1e-12))[0]
    nb = r2["n_face"][bot]
    set_ = np.abs(nb).sum(1) > 0
    assert set_.sum() >= 4
    assert np.abs(nb[set_, 1]).max() <= 1e-12 and np.allclose(np.abs(nb[set_, 0]), 1.0)
