"""Volume-of-fluid kernels on the device against the FP64 oracle: CICSAM face
weights, the cicsam::div equation (reference row layout, values, rhs),
the momentum flux and the mixture properties bit-identical; one implicit
volume-fraction step (fv::ddt + cicsam::div, theta = 1/2) solved through the
solver plugin agrees with the oracle's assembly and a direct solve, and keeps
the total volume of the dispersed phase in a closed box."""
import numpy as np
import pytest
import scipy.sparse.linalg as spla

from tests import helpers as H

pytestmark = pytest.mark.gpu


def same(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return a.shape == b.shape and np.array_equal(a, b)


def closed_box_velocity(fc):
    """Divergence-free, no flow through the unit square's walls."""
    x, y = fc[:, 0], fc[:, 1]
    return np.c_[np.sin(np.pi * x) * np.cos(np.pi * y), -np.cos(np.pi * x) * np.sin(np.pi * y)].reshape(-1)


@pytest.fixture(scope="module", params=["rect", "tri"])
def case(request, oracle, hip):
    if request.param == "rect":
        xy, cptr, cind = H.rect_case(24, 24)
    else:
        xy, cptr, cind = H.delaunay_case(2000, 20, seed=6)
    om = oracle.Mesh(xy, cptr, cind)
    hm = hip.HipMesh(xy, cptr, cind)
    f = om.faces()
    _, cc = om.cells()
    u = closed_box_velocity(f["centroid"])
    # a disc of the dispersed phase, smeared over about a cell
    r = np.hypot(cc[:, 0] - 0.5, cc[:, 1] - 0.7)
    h = 1.0 / np.sqrt(om.n_cells)
    gamma = np.clip(0.5 - (r - 0.2) / (2 * h), 0.0, 1.0)
    return om, hm, u, gamma


def test_weights_flux_properties_bitwise(oracle, hip, case):
    om, hm, u, gamma = case
    N, F = om.n_cells, om.n_faces
    rng = np.random.default_rng(3)
    g = gamma + 0.05 * rng.standard_normal(N)  # also outside [0, 1]: the clamps
    bc = np.full(F, hip.BC_NORMAL_GRADIENT, np.int32)
    gf = oracle.interpolate_faces(om, g, np.zeros(F), bc_type=bc)
    _, gradc = oracle.grad(om, g, gf)
    for dt in (1e-3, 2e-2):
        ref = oracle.cicsam_weights(om, u, g, gradc, dt)
        d_beta = hip.DeviceArray.zeros(F)
        hm.cicsam_weights(hip.DeviceArray(u), hip.DeviceArray(g), hip.DeviceArray(gradc.reshape(-1)), dt, d_beta)
        beta = d_beta.get()
        assert same(beta, ref)
        assert np.any((beta > 0) & (beta < 1)) and np.all((beta >= 0) & (beta <= 1))
    ref = oracle.cicsam_momentum_flux(om, 1000.0, 1.0, u, g, gf, beta)
    out = hip.DeviceArray.zeros(2 * F)
    hm.cicsam_momentum_flux(1000.0, 1.0, hip.DeviceArray(u), hip.DeviceArray(g), hip.DeviceArray(gf),
                            hip.DeviceArray(beta), out)
    assert same(out.get(), ref)
    rc, rf, mc, mf = oracle.vof_properties(om, 1000.0, 1.0, 1e-2, 1e-4, g, gf)
    d = [hip.DeviceArray.zeros(n) for n in (N, F, N, F)]
    hm.vof_properties(1000.0, 1.0, 1e-2, 1e-4, hip.DeviceArray(g), hip.DeviceArray(gf), *d)
    for a, b in zip(d, (rc, rf, mc, mf)):
        assert same(a.get(), b)


@pytest.mark.parametrize("theta", [0.5, 1.0])
def test_cicsam_div_bitwise(oracle, hip, case, theta):
    om, hm, u, gamma = case
    F = om.n_faces
    f = om.faces()
    rng = np.random.default_rng(4)
    # every boundary type the operator takes, on random boundary faces
    bc = np.full(F, hip.BC_NORMAL_GRADIENT, np.int32)
    bnd = np.where(f["rcell"] < 0)[0]
    pick = rng.integers(0, 3, len(bnd))
    bc[bnd[pick == 0]] = hip.BC_FIXED
    bc[bnd[pick == 2]] = hip.BC_SYMMETRY
    uu = u + 0.3 * rng.standard_normal(2 * F)  # flow through the walls as well
    gf, g0f = rng.random(F), rng.random(F)
    beta = rng.random(F)
    beta[f["rcell"] < 0] = 0.0
    ref = oracle.cicsam_div(om, uu, beta, theta, bc_type=bc, gamma_face=gf, gamma0_cell=gamma, gamma0_face=g0f)
    he = hip.HipEquation(hm, 1).cicsam_div(hip.DeviceArray(uu), hip.DeviceArray(beta), theta,
                                           bc_type=hip.DeviceArray(bc), gamma_face=hip.DeviceArray(gf),
                                           gamma0_cell=hip.DeviceArray(gamma), gamma0_face=hip.DeviceArray(g0f))
    rp, ci, v, rhs = he.export()
    orp, oci, ov, orhs = ref.get()
    assert same(rp, orp) and same(ci, oci) and same(v, ov) and same(rhs, orhs)


def test_unsupported_boundary_raises(hip, case):
    om, hm, u, gamma = case
    bc = np.full(om.n_faces, hip.BC_OUTFLOW, np.int32)
    with pytest.raises(hip.PhaseHipError):
        hip.HipEquation(hm, 1).cicsam_div(hip.DeviceArray(u), hip.DeviceArray.zeros(om.n_faces), 0.5,
                                          bc_type=hip.DeviceArray(bc), gamma0_cell=hip.DeviceArray(gamma))


def test_volume_fraction_step(oracle, hip, case):
    """gammaEqn = fv::ddt(gamma, dt) + cicsam::div(u, gamma, beta, 0.5) == 0
    (FractionalStepMultiphase.cpp:78-88), three steps on the device (weights,
    assembly and solve each step, fields in HBM) against the oracle with
    direct solves; the closed box keeps sum V gamma."""
    om, hm, u, gamma = case
    N, F = om.n_cells, om.n_faces
    vol, _ = om.cells()
    bc = np.full(F, hip.BC_NORMAL_GRADIENT, np.int32)
    dt = 0.2 / np.sqrt(N)
    d_u, d_bc = hip.DeviceArray(u), hip.DeviceArray(bc)
    d_g, d_gf = hip.DeviceArray(gamma), hip.DeviceArray.zeros(F)
    d_gradf, d_gradc, d_beta = hip.DeviceArray.zeros(2 * F), hip.DeviceArray.zeros(2 * N), hip.DeviceArray.zeros(F)
    s = hip.HipSparseMatrixSolver(0, tolerance=1e-13, maxIters=2000)
    g_ref = gamma.copy()
    mass0 = np.dot(vol, gamma)
    for _ in range(3):
        # oracle step
        gf = oracle.interpolate_faces(om, g_ref, np.zeros(F), bc_type=bc)
        _, gradc = oracle.grad(om, g_ref, gf)
        beta = oracle.cicsam_weights(om, u, g_ref, gradc, dt)
        e = oracle.ddt(om, dt, g_ref)
        e += oracle.cicsam_div(om, u, beta, 0.5, bc_type=bc, gamma_face=gf, gamma0_cell=g_ref, gamma0_face=gf)
        rp, ci, v, rhs = e.get()
        g_ref = spla.spsolve(H.compact_sorted(rp, ci, v).tocsc(), -rhs)
        # device step
        hm.interpolate_faces(d_g, d_gf, bc_type=d_bc)
        hm.grad(d_g, d_gf, d_gradf, d_gradc)
        hm.cicsam_weights(d_u, d_g, d_gradc, dt, d_beta)
        assert same(d_beta.get(), beta)
        he = hip.HipEquation(hm, 1).ddt(dt, d_g)
        he += hip.HipEquation(hm, 1).cicsam_div(d_u, d_beta, 0.5, bc_type=d_bc, gamma_face=d_gf, gamma0_cell=d_g,
                                                gamma0_face=d_gf)
        s.set_equation(he)
        s.solve()
        s.x_device(d_g)
        g = d_g.get()
        assert np.abs(g - g_ref).max() <= 1e-9
        g_ref = g  # both continue from the device field
    # conservative up to the solves and the walls' round-off normal flux
    assert abs(np.dot(vol, g) - mass0) <= 1e-9 * mass0
    assert g.min() > -0.05 and g.max() < 1.05
