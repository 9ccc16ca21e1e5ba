"""Ghost-cell immersed-boundary rows on the device against the CPU
restatement: cell types identical; the FIXED, NORMAL_GRADIENT and velocity
rows with identical columns and sources and values within 1e-12 (the 4x4
interpolation inverses come from different eliminations); a Poisson
problem around a circular hole assembled on the device over the fluid cells
and solved through the solver plug-in agrees with the restatement's direct
solve and with the harmonic solution."""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from phase_amd import meshgen as G

pytestmark = pytest.mark.gpu


def _csr(rp, ci, v, n):
    rows = np.repeat(np.arange(n), np.diff(rp))
    keep = ci >= 0
    return sp.csr_matrix((v[keep], (rows[keep], ci[keep])), shape=(n, n))


@pytest.fixture(scope="module")
def case(oracle, hip):
    from oracle.ghostcell_oracle import GhostCellOracle
    n = 48
    xy, cptr, cind = G.rectilinear(n, n)
    om = oracle.Mesh(xy, cptr, cind)
    hm = hip.HipMesh(xy, cptr, cind)
    circles = [(0.5, 0.5, 0.2), (0.82, 0.2, 0.09)]
    return om, hm, GhostCellOracle(om, xy, cptr, cind, circles), hip.HipGhostCellIB(hm, circles)


def test_cell_types(case):
    om, hm, og, dg = case
    t, o = dg.cell_types()
    assert np.array_equal(t, og.types) and np.array_equal(o, og.owner)
    assert (t == 1).sum() > 50 and (t == 2).sum() > 50


@pytest.mark.parametrize("kind", ["fixed", "ng", "velocity"])
def test_rows_match_oracle(oracle, hip, case, kind):
    om, hm, og, dg = case
    if kind == "velocity":
        U = [(0.3, -0.2), (0.0, 0.5)]
        ref = og.velocity_bcs(U)
        he = hip.HipEquation(hm, 2).gcib_velocity_bcs(dg, U)
    else:
        bc = oracle.BC_FIXED if kind == "fixed" else oracle.BC_NORMAL_GRADIENT
        ref = og.bcs(bc, 0.75)
        he = hip.HipEquation(hm, 1).gcib_bcs(dg, bc, 0.75)
    rp, ci, v, rhs = he.export()
    orp, oci, ov, orhs = ref.get()
    assert np.array_equal(rp, orp) and np.array_equal(ci, oci) and np.array_equal(rhs, orhs)
    assert np.abs(v - ov).max() <= 1e-12 * max(1.0, np.abs(ov).max())


@pytest.mark.parametrize("solver", [{}, {"solver": "gmres"}, {"solver": "gmres", "preconditioner": "multigrid"},
                                    {"solver": "gmres", "gmresRestart": 8},
                                    {"preconditioner": "multigrid"}, {"mode": "mixed"},
                                    {"mode": "mixed", "solver": "gmres", "preconditioner": "multigrid"}])
def test_poisson_around_hole_on_device(oracle, hip, case, solver):
    """The ghost-cell rows make the system non-symmetric: BiCGStab (the
    reference's default) and flexible GMRES (its multigrid backend's), with
    ILU(0) or the multigrid cycle, in FP64 or with FP32 inner iterations
    inside FP64 refinement, all agree with the direct solve."""
    om, hm, og, dg = case
    N, F = om.n_cells, om.n_faces
    f = om.faces()
    _, cc = om.cells()
    exact = lambda x: np.log(np.hypot(x[..., 0] - 0.5, x[..., 1] - 0.5) / 0.8) / np.log(0.2 / 0.8)
    bc = np.where(f["rcell"] < 0, oracle.BC_FIXED, oracle.BC_NORMAL_GRADIENT).astype(np.int32)
    pf = exact(f["centroid"])
    # oracle: fluid rows of the laplacian + ghost rows, direct solve
    e = oracle.laplacian(om, 1, gamma=1.0, bc_type=bc, phi_face=pf)
    e.remove_cells(N, (og.types != 0).astype(np.int32))
    e += og.bcs(oracle.BC_FIXED, 1.0)
    rp, ci, v, rhs = e.get()
    xs = spla.spsolve(_csr(rp, ci, v, N).tocsc(), -rhs)
    # device: the same expression, solved by the plug-in
    mask = hip.DeviceArray.zeros(N, np.int32)
    dg.cell_types_device(mask)
    he = hip.HipEquation(hm, 1).laplacian(1, gamma=1.0, bc_type=hip.DeviceArray(bc), phi_face=hip.DeviceArray(pf))
    he.remove_cells(mask)
    he += hip.HipEquation(hm, 1).gcib_bcs(dg, hip.BC_FIXED, 1.0)
    s = hip.HipSparseMatrixSolver(0, tolerance=1e-12, maxIters=5000, **solver)
    s.set_equation(he)
    assert s.solve() <= 1e-12
    x = s.x()
    assert np.linalg.norm(x - xs) <= 1e-8 * np.linalg.norm(xs)
    fluid = og.types == 0
    # the hole is the first object; the second (a FIXED 1 body as well) makes
    # the field differ from the one-hole harmonic away from it only
    far = fluid & (np.hypot(cc[:, 0] - 0.82, cc[:, 1] - 0.2) > 0.3)
    assert np.abs(x - exact(cc))[far].max() < 0.2


def test_invalid_inputs(hip, case):
    om, hm, og, dg = case
    with pytest.raises(hip.PhaseHipError):
        hip.HipEquation(hm, 1).gcib_bcs(dg, hip.BC_SYMMETRY, 0.0)
    with pytest.raises(hip.PhaseHipError):
        hip.HipEquation(hm, 2).gcib_bcs(dg, hip.BC_FIXED, 0.0)
    xy, cptr, cind = G.delaunay_square(12, G.delaunay_cells_for(600, 12), seed=3)
    with pytest.raises(hip.PhaseHipError):  # nodes of a triangle mesh do not have four cells
        hip.HipGhostCellIB(hip.HipMesh(xy, cptr, cind), [(0.5, 0.5, 0.2)])


@pytest.mark.parametrize("omega,correctors,psolver", [(None, 1, {}), (2.0, 1, {}), (None, 2, {}),
                                                     (None, 2, {"solver": "gmres", "preconditioner": "multigrid"})])
def test_fractional_step_past_a_cylinder(oracle, hip, omega, correctors, psolver):
    """FractionalStepGCIB on the device: a channel with inflow u = 1, a fixed
    cylinder of radius 0.25 (ghost-cell velocity and NORMAL_GRADIENT pressure
    rows), outlet p = 0; four steps against the restatement with direct solves:
    velocities and pressure within 1e-7, solid cells at the body's velocity
    (at rest, or spinning at omega about its centre: Motion::velocity rows in
    the time step), fluid faces divergence-free; one or two pressure-corrector
    passes per step; the pressure by BiCGStab + ILU(0), or by flexible GMRES +
    multigrid on the scaled ghost rows."""
    from oracle.fstep_oracle import FractionalStepOracle
    from oracle.ghostcell_oracle import GhostCellOracle
    xy, cptr, cind = G.rectilinear(48, 32, width=3.0, height=2.0)
    om, hm = oracle.Mesh(xy, cptr, cind), hip.HipMesh(xy, cptr, cind)
    N, F = om.n_cells, om.n_faces
    f = om.faces()
    a, b = xy[f["lnode"]], xy[f["rnode"]]
    bnd = f["rcell"] < 0
    left = bnd & (a[:, 0] == 0.0) & (b[:, 0] == 0.0)
    right = bnd & (a[:, 0] == 3.0) & (b[:, 0] == 3.0)
    ubc = np.where(right, oracle.BC_NORMAL_GRADIENT, oracle.BC_FIXED).astype(np.int32)
    ub = np.zeros((F, 2))
    ub[left, 0] = 1.0
    pbc = np.where(right, oracle.BC_FIXED, oracle.BC_NORMAL_GRADIENT).astype(np.int32)
    circles = [(1.0, 1.0, 0.25)]
    og = GhostCellOracle(om, xy, cptr, cind, circles)
    omegas = None if omega is None else [omega]
    ref = FractionalStepOracle(om, 1.0, 0.01, ubc, ub, p_bc=pbc, pin=-1,
                               immersed=(og, [(0.0, 0.0)], oracle.BC_NORMAL_GRADIENT, 0.0, omegas))
    dg = hip.HipGhostCellIB(hm, circles)
    dev = hip.HipFractionalStep(hm, 1.0, 0.01, ubc, ub, p_bc=pbc, pin_cell=-1)
    dev.immersed_boundary(dg, [(0.0, 0.0)], hip.BC_NORMAL_GRADIENT, 0.0, omegas=omegas)
    for k in ("u", "p"):
        dev.setup(k + ".tolerance", 1e-12)
        dev.setup(k + ".maxIters", 20000)
    dev.setup("numPressureCorrections", correctors)
    for k, v in psolver.items():
        dev.setup("p." + k, v)
    solid = og.types == 2
    _, cc = om.cells()
    w = 0.0 if omega is None else omega
    u_body = np.c_[-(cc[:, 1] - 1.0) * w, (cc[:, 0] - 1.0) * w]
    for step in range(4):
        ref.step(0.01, correctors)
        hist = dev.advance(0.01)
        u, p, uf = dev.fields()
        print("step", step, "iters", hist[0, :2], "div", hist[0, 2], flush=True)
        su, sp_ = np.abs(ref.uC).max(), np.abs(ref.pC).max()
        assert np.abs(u - ref.uC).max() <= 1e-7 * su
        assert np.abs(uf - ref.uF).max() <= 1e-7 * su
        assert np.abs(p - ref.pC).max() <= 1e-7 * sp_
        assert np.abs(u[solid] - u_body[solid]).max() <= 1e-9 * su
        assert hist[0, 2] <= 1e-9 * np.abs(uf).max()


def test_hydrodynamic_force_matches_oracle(oracle, hip, case):
    """applyHydrodynamicForce on device fields against the restatement (the
    4x4 inverses differ in rounding: 1e-10), with gravity."""
    om, hm, og, dg = case
    N = om.n_cells
    _, cc = om.cells()
    rng = np.random.default_rng(17)
    u = np.stack([np.sin(3 * cc[:, 0]) * cc[:, 1], np.cos(2 * cc[:, 1]) + cc[:, 0] ** 2], -1) \
        + 0.01 * rng.standard_normal((N, 2))
    p = 1.0 + cc[:, 0] - 2.0 * cc[:, 1] ** 2
    ref, _ = og.hydrodynamic_force(1.3, 0.02, u, p, g=(0.0, -9.81))
    got = dg.hydrodynamic_force(1.3, 0.02, hip.DeviceArray(u.reshape(-1)), hip.DeviceArray(p), g=(0.0, -9.81))
    assert got.shape == ref.shape
    assert np.abs(got - ref).max() <= 1e-10 * max(1.0, np.abs(ref).max())


@pytest.mark.parametrize("kind", ["fixed", "ng", "velocity", "force"])
def test_box_and_circle_bodies_match_oracle(oracle, hip, kind):
    """A circle and a box (Box.cpp) together: cell types identical, rows with
    identical columns and sources and values within 1e-12, forces within 1e-10."""
    from oracle.ghostcell_oracle import GhostCellOracle
    xy, cptr, cind = G.rectilinear(44, 40)
    om, hm = oracle.Mesh(xy, cptr, cind), hip.HipMesh(xy, cptr, cind)
    circles, boxes = [(0.3, 0.35, 0.15)], [(0.55, 0.42, 0.83, 0.71)]
    og = GhostCellOracle(om, xy, cptr, cind, circles, boxes=boxes)
    dg = hip.HipGhostCellIB(hm, circles, boxes=boxes)
    t, o = dg.cell_types()
    assert np.array_equal(t, og.types) and np.array_equal(o, og.owner)
    assert set(np.unique(o[t > 0])) == {0, 1}
    if kind == "force":
        _, cc = om.cells()
        u = np.stack([np.sin(3 * cc[:, 0]) * cc[:, 1], cc[:, 0] ** 2 - cc[:, 1]], -1)
        p = 1.0 + cc[:, 0] - 2.0 * cc[:, 1] ** 2
        ref, _ = og.hydrodynamic_force(1.0, 0.01, u, p)
        got = dg.hydrodynamic_force(1.0, 0.01, hip.DeviceArray(u.reshape(-1)), hip.DeviceArray(p))
        assert np.abs(got - ref).max() <= 1e-10 * max(1.0, np.abs(ref).max())
        return
    if kind == "velocity":  # moving bodies: translation and rotation
        U, W = [(0.2, 0.0), (0.0, -0.4)], [0.8, -1.3]
        ref = og.velocity_bcs(U, omegas=W)
        he = hip.HipEquation(hm, 2).gcib_velocity_bcs(dg, U, omegas=W)
    else:
        bc = oracle.BC_FIXED if kind == "fixed" else oracle.BC_NORMAL_GRADIENT
        ref = og.bcs(bc, 0.25)
        he = hip.HipEquation(hm, 1).gcib_bcs(dg, bc, 0.25)
    rp, ci, v, rhs = he.export()
    orp, oci, ov, orhs = ref.get()
    assert np.array_equal(rp, orp) and np.array_equal(ci, oci) and np.array_equal(rhs, orhs)
    assert np.abs(v - ov).max() <= 1e-12 * max(1.0, np.abs(ov).max())
