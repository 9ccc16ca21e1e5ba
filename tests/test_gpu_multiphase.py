"""The device multiphase driver (FractionalStepMultiphase without surface
tension: CICSAM volume fraction, mixture properties, gravity source,
momentum, pressure projection, all fields in HBM) against the CPU
restatement with direct solves, three steps of a heavy drop falling in a
closed box on a rectilinear and a triangle mesh; the device solves stop at a
relative residual of 1e-12, so the fields agree to about that."""
import numpy as np
import pytest

from tests import helpers as H

pytestmark = pytest.mark.gpu


def _case(oracle, kind):
    if kind == "rect":
        xy, cptr, cind = H.rect_case(12, 24)
        xy = xy * np.array([1.0, 2.0])
    else:
        xy, cptr, cind = H.delaunay_case(600, 12, seed=9)
        xy = xy * np.array([1.0, 2.0])
    om = oracle.Mesh(xy, cptr, cind)
    f = om.faces()
    F = om.n_faces
    a, b = xy[f["lnode"]], xy[f["rnode"]]
    top = (f["rcell"] < 0) & (np.abs(a[:, 1] - 2.0) < 1e-12) & (np.abs(b[:, 1] - 2.0) < 1e-12)
    pbc = np.where(top, oracle.BC_FIXED, oracle.BC_NORMAL_GRADIENT).astype(np.int32)
    ubc = np.full(F, oracle.BC_FIXED, np.int32)
    _, cc = om.cells()
    r = np.hypot(cc[:, 0] - 0.5, cc[:, 1] - 1.2)
    gam = np.clip(0.5 - (r - 0.25) * 12, 0.0, 1.0)
    return xy, cptr, cind, om, ubc, pbc, gam


@pytest.mark.parametrize("kind,correctors", [("rect", 1), ("tri", 1), ("rect", 2)])
def test_multiphase_steps_match_oracle(oracle, hip, kind, correctors):
    from oracle.multiphase_oracle import MultiphaseOracle
    xy, cptr, cind, om, ubc, pbc, gam = _case(oracle, kind)
    F = om.n_faces
    assert pbc.min() == oracle.BC_FIXED  # the top is a pressure outlet: no pin needed
    args = (1.0, 1000.0, 0.01, 1.0, (0.0, -9.81))
    ref = MultiphaseOracle(om, *args, ubc, np.zeros(2 * F), pbc, np.zeros(F), gam)
    hm = hip.HipMesh(xy, cptr, cind)
    dev = hip.HipMultiphase(hm, *args, ubc, np.zeros(2 * F), gam, p_bc=pbc, p_face0=np.zeros(F))
    for k in ("gamma", "u", "p"):
        dev.setup(k + ".tolerance", 1e-12)
        dev.setup(k + ".maxIters", 5000)
    dev.setup("numPressureCorrections", correctors)
    dt = 2e-3
    for step in range(3):
        ref.step(dt, correctors)
        hist = dev.advance(dt)
        u, p, uf, g = dev.fields()
        print("step", step, "iters", hist[0, :3], "div", hist[0, 3], flush=True)
        assert np.abs(g - ref.gC).max() <= 1e-9
        su, sp = np.abs(ref.uC).max(), np.abs(ref.pC).max()
        assert su > 0 and sp > 0
        assert np.abs(u - ref.uC).max() <= 1e-7 * su
        assert np.abs(uf - ref.uF).max() <= 1e-7 * su
        assert np.abs(p - ref.pC).max() <= 1e-7 * sp
        assert hist[0, 3] <= 1e-9 * np.abs(uf).max()


def test_multiphase_rejects_bad_input(hip):
    xy, cptr, cind = H.rect_case(4, 4)
    hm = hip.HipMesh(xy, cptr, cind)
    F = hm.n_faces
    ubc = np.zeros(F, np.int32)
    with pytest.raises(hip.PhaseHipError):
        hip.HipMultiphase(hm, -1.0, 1.0, 1.0, 1.0, (0, -1), ubc, np.zeros(2 * F), np.zeros(hm.n_cells))
    d = hip.HipMultiphase(hm, 1.0, 2.0, 1.0, 1.0, (0, -1), ubc, np.zeros(2 * F), np.zeros(hm.n_cells), pin_cell=0)
    with pytest.raises(hip.PhaseHipError):
        d.setup("q.tolerance", 1e-8)
    for bad in ("0", "11", "2x", ""):
        with pytest.raises(hip.PhaseHipError):
            d.setup("numPressureCorrections", bad)
    with pytest.raises(hip.PhaseHipError):
        d.advance(-1.0)


@pytest.mark.parametrize("kind", ["rect", "tri"])
def test_surface_tension_matches_oracle(oracle, hip, kind):
    """CELESTE on the device: the smoothed volume fraction bit-identical
    (same kernel weights and summation order); normals, curvature and force
    within 1e-9 (the stencils' pseudo-inverses come from different
    factorisations); then three steps of a drop with surface tension and
    gravity against the restatement with direct solves."""
    from oracle.celeste_oracle import CelesteOracle
    from oracle.multiphase_oracle import MultiphaseOracle
    xy, cptr, cind, om, ubc, pbc, _ = _case(oracle, kind)
    F = om.n_faces
    _, cc = om.cells()
    r = np.hypot(cc[:, 0] - 0.5, cc[:, 1] - 1.0)
    h = 1.0 / 12
    gam = 0.5 * (1.0 - np.tanh((r - 0.3) / (1.5 * h)))
    f = om.faces()
    ca = np.where(f["rcell"] < 0, np.deg2rad(60.0), np.pi / 2)
    args = (1.0, 10.0, 0.05, 0.1, (0.0, -1.0))
    model = CelesteOracle(om, cptr, cind, kernel_width=2.5 * h)
    ref = MultiphaseOracle(om, *args, ubc, np.zeros(2 * F), pbc, np.zeros(F), gam,
                           surface_tension=(model, 0.07, ca))
    hm = hip.HipMesh(xy, cptr, cind)
    dev = hip.HipMultiphase(hm, *args, ubc, np.zeros(2 * F), gam, p_bc=pbc, p_face0=np.zeros(F))
    dev.surface_tension(0.07, 2.5 * h, "pow8", contact_angle=ca)
    d, o = dev.surface_fields(), ref.st_fields
    assert same(d["gamma_tilde"], o["gamma_tilde"])
    assert np.abs(d["n_cell"] - o["n_cell"]).max() <= 1e-9
    assert np.abs(d["n_face"] - o["n_face"]).max() <= 1e-9
    assert np.array_equal(d["kappa_cell"] != 0, o["kappa_cell"] != 0)
    sk = np.abs(o["kappa_cell"]).max()
    assert sk > 0
    assert np.abs(d["kappa_cell"] - o["kappa_cell"]).max() <= 1e-9 * sk
    assert np.abs(d["kappa_face"] - o["kappa_face"]).max() <= 1e-9 * sk
    sf = np.abs(o["fst_face"]).max()
    assert np.abs(d["fst_face"] - o["fst_face"]).max() <= 1e-9 * sf
    assert np.abs(d["fst_cell"] - ref.fstC).max() <= 1e-9 * sf
    for k in ("gamma", "u", "p"):
        dev.setup(k + ".tolerance", 1e-12)
        dev.setup(k + ".maxIters", 5000)
    for step in range(3):
        ref.step(2e-3)
        hist = dev.advance(2e-3)
        u, p, uf, g = dev.fields()
        print("step", step, "iters", hist[0, :3], "div", hist[0, 3], flush=True)
        assert np.abs(g - ref.gC).max() <= 1e-9
        su, sp = np.abs(ref.uC).max(), np.abs(ref.pC).max()
        assert np.abs(u - ref.uC).max() <= 1e-7 * su
        assert np.abs(uf - ref.uF).max() <= 1e-7 * su
        assert np.abs(p - ref.pC).max() <= 1e-7 * sp
    sf = np.abs(ref.fstF).max()
    assert np.abs(dev.surface_fields()["fst_face"] - ref.fstF).max() <= 1e-8 * sf


def same(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return a.shape == b.shape and np.array_equal(a, b)
