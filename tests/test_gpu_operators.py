"""fv:: operators, CrsEquation algebra and field kernels on the device against
the FP64 oracle: bit-identical rows (reference column order and padding),
values and rhs; fields bit-identical; a device-assembled equation solved
through the solver plugin agrees with a direct solve."""
import numpy as np
import pytest
import scipy.sparse.linalg as spla

from phase_amd import meshgen as G
from tests import helpers as H

pytestmark = pytest.mark.gpu


def same(a, b):
    """Bit-identical arrays (NaN-free; +0 and -0 compare equal)."""
    a, b = np.asarray(a), np.asarray(b)
    return a.shape == b.shape and np.array_equal(a, b)


@pytest.fixture(scope="module", params=["rect", "tri"])
def meshes(request, oracle, hip):
    if request.param == "rect":
        xy, cptr, cind = H.rect_case(8, 8)
    else:
        xy, cptr, cind = H.delaunay_case(800, 12, seed=4)
    om = oracle.Mesh(xy, cptr, cind)
    hm = hip.HipMesh(xy, cptr, cind)
    f = om.faces()
    pid = G.boundary_patch(f["lnode"], f["rnode"], f["rcell"], xy)
    # a mix of every supported boundary type: x- FIXED, x+ NORMAL_GRADIENT,
    # y- SYMMETRY, y+ FIXED
    bc = np.full(om.n_faces, hip.BC_NORMAL_GRADIENT, np.int32)
    bc[pid == 0] = hip.BC_FIXED
    bc[pid == 2] = hip.BC_SYMMETRY
    bc[pid == 3] = hip.BC_FIXED
    return om, hm, bc


def fields(om, seed):
    rng = np.random.default_rng(seed)
    N, F = om.n_cells, om.n_faces
    return dict(
        g=np.exp(rng.standard_normal(F)), g0=np.exp(rng.standard_normal(F)),
        s1c=rng.standard_normal(N), s1f=rng.standard_normal(F), s0c=rng.standard_normal(N),
        s0f=rng.standard_normal(F),
        v1c=rng.standard_normal(2 * N), v1f=rng.standard_normal(2 * F), v0c=rng.standard_normal(2 * N),
        v0f=rng.standard_normal(2 * F),
        u=rng.standard_normal(2 * F), u0=rng.standard_normal(2 * F))


def dev(hip, **arrays):
    return {k: hip.DeviceArray(v) for k, v in arrays.items()}


def check_eq(he, oe):
    rp, ci, v, rhs = he.export()
    orp, oci, ov, orhs = oe.get()
    assert same(rp, orp)
    assert same(ci, oci)
    assert same(v, ov)
    assert same(rhs, orhs)


@pytest.mark.parametrize("variant", [0, 1, 2, 3])
@pytest.mark.parametrize("ncomp", [1, 2])
def test_laplacian_matches_oracle(hip, oracle, meshes, variant, ncomp):
    om, hm, bc = meshes
    fl = fields(om, 10 + variant)
    key = "s" if ncomp == 1 else "v"
    phiF, phi0C, phi0F = fl[key + "1f"], fl[key + "0c"], fl[key + "0f"]
    theta = 0.625
    oe = oracle.laplacian(om, variant, ncomp=ncomp, gamma=0.75, gamma_face=fl["g"], gamma0_face=fl["g0"],
                          theta=theta, bc_type=bc, phi_face=phiF, phi0_cell=phi0C, phi0_face=phi0F)
    d = dev(hip, g=fl["g"], g0=fl["g0"], bc=bc, phiF=phiF, phi0C=phi0C, phi0F=phi0F)
    he = hip.HipEquation(hm, ncomp)
    he.laplacian(variant, gamma=0.75, gamma_face=d["g"], gamma0_face=d["g0"], theta=theta, bc_type=d["bc"],
                 phi_face=d["phiF"], phi0_cell=d["phi0C"], phi0_face=d["phi0F"])
    check_eq(he, oe)


@pytest.mark.parametrize("ncomp", [1, 2])
def test_ddt_matches_oracle(hip, oracle, meshes, ncomp):
    om, hm, bc = meshes
    fl = fields(om, 3)
    phi0 = fl["s0c"] if ncomp == 1 else fl["v0c"]
    oe = oracle.ddt(om, 0.0125, phi0, ncomp=ncomp)
    he = hip.HipEquation(hm, ncomp)
    he.ddt(0.0125, hip.DeviceArray(phi0))
    check_eq(he, oe)


@pytest.mark.parametrize("form", ["scalar", "field", "group"])
@pytest.mark.parametrize("ncomp", [1, 2])
def test_ddt_density_matches_oracle(hip, oracle, meshes, form, ncomp):
    om, hm, bc = meshes
    fl = fields(om, 4)
    phi0 = fl["s0c"] if ncomp == 1 else fl["v0c"]
    rho, rho0 = np.exp(fl["s1c"]), np.exp(fl["s0c"])
    mask = (np.arange(om.n_cells) % 3 != 1).astype(np.int32)
    kw = dict(scalar=dict(rho=1.5), field=dict(rho_cell=rho, rho0_cell=rho0),
              group=dict(rho=0.75, cell_mask=mask))[form]
    oe = oracle.ddt_rho(om, 0.0125, phi0, ncomp=ncomp, **kw)
    dkw = {k: (hip.DeviceArray(v) if isinstance(v, np.ndarray) else v) for k, v in kw.items()}
    he = hip.HipEquation(hm, ncomp).ddt_rho(0.0125, hip.DeviceArray(phi0), **dkw)
    check_eq(he, oe)


@pytest.mark.parametrize("theta", [1.0, 0.5, 0.0])
@pytest.mark.parametrize("ncomp", [1, 2])
def test_div_matches_oracle(hip, oracle, meshes, theta, ncomp):
    om, hm, bc = meshes
    fl = fields(om, 5)
    key = "s" if ncomp == 1 else "v"
    phiF, phi0C, phi0F = fl[key + "1f"], fl[key + "0c"], fl[key + "0f"]
    oe = oracle.div(om, fl["u"], fl["u0"], theta=theta, bc_type=bc, phi_face=phiF, phi0_cell=phi0C,
                    phi0_face=phi0F, ncomp=ncomp)
    d = dev(hip, u=fl["u"], u0=fl["u0"], bc=bc, phiF=phiF, phi0C=phi0C, phi0F=phi0F)
    he = hip.HipEquation(hm, ncomp)
    he.div(d["u"], d["u0"], theta=theta, bc_type=d["bc"], phi_face=d["phiF"], phi0_cell=d["phi0C"],
           phi0_face=d["phi0F"])
    check_eq(he, oe)


@pytest.mark.parametrize("theta", [1.0, 0.5])
@pytest.mark.parametrize("ncomp", [1, 2])
def test_divc_matches_oracle(hip, oracle, meshes, theta, ncomp):
    om, hm, bc = meshes
    fl = fields(om, 6)
    key = "s" if ncomp == 1 else "v"
    phiF, phi0C, phi0F = fl[key + "1f"], fl[key + "0c"], fl[key + "0f"]
    oe = oracle.divc(om, fl["u"], fl["u0"], theta=theta, bc_type=bc, phi_face=phiF, phi0_cell=phi0C,
                     phi0_face=phi0F, ncomp=ncomp)
    d = dev(hip, u=fl["u"], u0=fl["u0"], bc=bc, phiF=phiF, phi0C=phi0C, phi0F=phi0F)
    he = hip.HipEquation(hm, ncomp).divc(d["u"], d["u0"], theta=theta, bc_type=d["bc"], phi_face=d["phiF"],
                                          phi0_cell=d["phi0C"], phi0_face=d["phi0F"])
    check_eq(he, oe)


@pytest.mark.parametrize("ncomp", [1, 2])
def test_dive_matches_oracle(hip, oracle, meshes, ncomp):
    om, hm, bc = meshes
    fl = fields(om, 7)
    key = "s" if ncomp == 1 else "v"
    a = dict(p0c=fl[key + "0c"], p1c=fl[key + "1c"], p0f=fl[key + "0f"], p1f=fl[key + "1f"])
    oe = oracle.dive(om, fl["u0"], fl["u"], 0.5, a["p0c"], a["p1c"], a["p0f"], a["p1f"], bc_type=bc, ncomp=ncomp)
    d = dev(hip, u0=fl["u0"], u1=fl["u"], bc=bc, **a)
    he = hip.HipEquation(hm, ncomp).dive(d["u0"], d["u1"], 0.5, d["p0c"], d["p1c"], d["p0f"], d["p1f"],
                                          bc_type=d["bc"])
    check_eq(he, oe)


def test_div_grad_matches_oracle(hip, oracle, meshes):
    om, hm, bc = meshes
    fl = fields(om, 8)
    J = np.random.default_rng(9).standard_normal(4 * om.n_cells)
    oe = oracle.div_grad(om, fl["u"], J, bc_type=bc, u_face=fl["v1f"])
    d = dev(hip, u=fl["u"], J=J, bc=bc, uf=fl["v1f"])
    he = hip.HipEquation(hm, 2).div_grad(d["u"], d["J"], bc_type=d["bc"], u_face=d["uf"])
    check_eq(he, oe)


def test_algebra_matches_oracle(hip, oracle, meshes):
    """ddt + div(theta=0: every coefficient 0, dropped) == laplacian - v, then
    *=, ==scalar, relax: the reference's row rebuild with zero-drop."""
    om, hm, bc = meshes
    fl = fields(om, 7)
    N = om.n_cells
    # momentum-like vector equation
    oe = oracle.ddt(om, 0.01, fl["v0c"], ncomp=2)
    oe += oracle.div(om, fl["u"], fl["u0"], theta=0.0, bc_type=bc, phi_face=fl["v1f"], phi0_cell=fl["v0c"],
                     phi0_face=fl["v0f"], ncomp=2)
    ol = oracle.laplacian(om, 0, ncomp=2, gamma=0.01, theta=0.5, bc_type=bc, phi_face=fl["v1f"],
                          phi0_cell=fl["v0c"], phi0_face=fl["v0f"])
    oe -= ol
    src = np.random.default_rng(8).standard_normal(2 * N)
    oe -= src
    oe.scale(0.5)
    oe.eq(0.25)
    oe.relax(0.7, fl["v1c"].reshape(-1, 2).T.reshape(-1))  # block layout of the current field

    d = dev(hip, u=fl["u"], u0=fl["u0"], bc=bc, v1f=fl["v1f"], v0c=fl["v0c"], v0f=fl["v0f"], src=src,
            cur=fl["v1c"].reshape(-1, 2).T.reshape(-1))
    he = hip.HipEquation(hm, 2).ddt(0.01, d["v0c"])
    t = hip.HipEquation(hm, 2).div(d["u"], d["u0"], theta=0.0, bc_type=d["bc"], phi_face=d["v1f"],
                                   phi0_cell=d["v0c"], phi0_face=d["v0f"])
    he += t
    t.laplacian(0, gamma=0.01, theta=0.5, bc_type=d["bc"], phi_face=d["v1f"], phi0_cell=d["v0c"],
                phi0_face=d["v0f"])
    he -= t
    he.sub_vector(d["src"]).scale(0.5).eq_scalar(0.25).relax(0.7, d["cur"])
    check_eq(he, oe)


def test_scale_cells_and_remove_match_oracle(hip, oracle, meshes):
    """rho * (ddt + laplacian) with a cell field, then remove(cell) on a
    cell group and a merge that drops the removed columns."""
    om, hm, bc = meshes
    fl = fields(om, 12)
    N = om.n_cells
    rho = np.exp(fl["s1c"])
    mask = (np.arange(N) % 5 == 2).astype(np.int32)
    oe = oracle.ddt(om, 0.02, fl["v0c"], ncomp=2)
    oe += oracle.laplacian(om, 1, ncomp=2, gamma=0.4, bc_type=bc, phi_face=fl["v1f"])
    oe.scale_cells(N, rho)
    oe.remove_cells(N, mask)
    d = dev(hip, v0c=fl["v0c"], bc=bc, v1f=fl["v1f"], rho=rho, mask=mask)
    he = hip.HipEquation(hm, 2).ddt(0.02, d["v0c"])
    he += hip.HipEquation(hm, 2).laplacian(1, gamma=0.4, bc_type=d["bc"], phi_face=d["v1f"])
    he.scale_cells(d["rho"]).remove_cells(d["mask"])
    check_eq(he, oe)
    oe += oracle.ddt(om, 0.02, fl["v0c"], ncomp=2)
    he += hip.HipEquation(hm, 2).ddt(0.02, d["v0c"])
    check_eq(he, oe)


def test_unsupported_boundary_raises(hip, oracle, meshes):
    om, hm, bc = meshes
    bad = bc.copy()
    bad[bc == hip.BC_SYMMETRY] = hip.BC_OUTFLOW
    with pytest.raises(RuntimeError):
        oracle.laplacian(om, 1, bc_type=bad)
    he = hip.HipEquation(hm, 1)
    with pytest.raises(hip.PhaseHipError) as ei:
        he.laplacian(1, bc_type=hip.DeviceArray(bad), phi_face=hip.DeviceArray(np.zeros(om.n_faces)))
    assert "boundary" in str(ei.value)


def test_fields_match_oracle(hip, oracle, meshes):
    om, hm, bc = meshes
    fl = fields(om, 9)
    N, F = om.n_cells, om.n_faces
    d = dev(hip, u=fl["u"], s1c=fl["s1c"], s1f=fl["s1f"], v1c=fl["v1c"], bc=bc)
    out = hip.DeviceArray.zeros(N)
    hm.src_div(d["u"], out)
    assert same(out.get(), oracle.src_div(om, fl["u"]))
    out2 = hip.DeviceArray.zeros(2 * N)
    hm.src_src(d["v1c"], out2, ncomp=2)
    assert same(out2.get(), oracle.src_src(om, fl["v1c"], ncomp=2))
    gf, gc = hip.DeviceArray.zeros(2 * F), hip.DeviceArray.zeros(2 * N)
    hm.grad(d["s1c"], d["s1f"], gf, gc)
    ogf, ogc = oracle.grad(om, fl["s1c"], fl["s1f"])
    assert same(gf.get(), ogf.reshape(-1))
    assert same(gc.get(), ogc.reshape(-1))
    for ncomp, key in ((1, "s1c"), (2, "v1c")):
        start = np.random.default_rng(ncomp).standard_normal(ncomp * F)
        pf = hip.DeviceArray(start)
        hm.interpolate_faces(d[key], pf, bc_type=d["bc"], ncomp=ncomp)
        assert same(pf.get(), oracle.interpolate_faces(om, fl[key], start, bc_type=bc, ncomp=ncomp))


def test_equation_solve(hip, oracle, meshes):
    """A device-assembled pressure-like equation handed to the solver plugin
    (A x = -rhs) agrees with the oracle's direct solve; a second hand-over with
    new values and the same pattern reuses the ordering."""
    om, hm, bc = meshes
    fl = fields(om, 11)
    N = om.n_cells
    d = dev(hip, g=fl["g"], bc=bc, phiF=fl["s1f"])
    he = hip.HipEquation(hm, 1).laplacian(3, gamma_face=d["g"], bc_type=d["bc"], phi_face=d["phiF"])
    b = np.random.default_rng(2).standard_normal(N)
    he.sub_vector(hip.DeviceArray(b))
    oe = oracle.laplacian(om, 3, gamma_face=fl["g"], bc_type=bc, phi_face=fl["s1f"])
    oe -= b
    rp, ci, v, rhs = oe.get()
    xs = spla.spsolve(H.compact_sorted(rp, ci, v).tocsc(), -rhs)
    s = hip.HipSparseMatrixSolver(0, tolerance=1e-12, maxIters=5000)
    s.set_equation(he)
    s.solve()
    x = s.x()
    assert np.linalg.norm(x - xs) <= 1e-8 * np.linalg.norm(xs)
    xd = hip.DeviceArray.zeros(N)
    s.x_device(xd)
    assert same(xd.get(), x)
    # new values, same pattern
    he.scale(2.0)
    s.set_equation(he)
    s.solve()
    assert np.linalg.norm(s.x() - xs) <= 1e-8 * np.linalg.norm(xs)


def test_union_pattern_solves(hip, oracle, meshes):
    """pattern = union: ddt + upwind div drops the zero outflow coefficients,
    so reversing the flow changes the pattern; the solver keeps one ordering
    for the union of the patterns (explicit zeros) and every solve still
    agrees with a direct solve of that step's exported equation."""
    om, hm, bc = meshes
    N, F = om.n_cells, om.n_faces
    fl = fields(om, 21)
    d = dev(hip, bc=bc, phiF=fl["s1f"], phi0C=fl["s0c"], phi0F=fl["s0f"])
    s = hip.HipSparseMatrixSolver(0, tolerance=1e-12, maxIters=5000, pattern="union")
    counts = []
    for sign in (1.0, -1.0, 1.0):
        u = hip.DeviceArray(sign * fl["u"])
        he = hip.HipEquation(hm, 1).ddt(0.05, d["phi0C"])
        he += hip.HipEquation(hm, 1).div(u, u, theta=1.0, bc_type=d["bc"], phi_face=d["phiF"], phi0_cell=d["phi0C"],
                                         phi0_face=d["phi0F"])
        rp, ci, v, rhs = he.export()
        counts.append(int((ci >= 0).sum()))
        xs = spla.spsolve(H.compact_sorted(rp, ci, v).tocsc(), -rhs)
        s.set_equation(he)
        s.solve()
        assert np.linalg.norm(s.x() - xs) <= 1e-9 * np.linalg.norm(xs)
    assert counts[0] < N + 2 * om.n_interior  # entries were dropped
    # a full Laplacian pattern after the union of both flow directions
    he = hip.HipEquation(hm, 1).laplacian(3, gamma_face=hip.DeviceArray(fl["g"]), bc_type=d["bc"],
                                          phi_face=d["phiF"])
    he.sub_vector(hip.DeviceArray(fl["s1c"]))
    rp, ci, v, rhs = he.export()
    xs = spla.spsolve(H.compact_sorted(rp, ci, v).tocsc(), -rhs)
    s.set_equation(he)
    s.solve()
    assert np.linalg.norm(s.x() - xs) <= 1e-9 * np.linalg.norm(xs)
    with pytest.raises(hip.PhaseHipError):
        s.setup("pattern", "sometimes")


def test_face_to_cell_matches_oracle(hip, oracle, meshes):
    """VectorFiniteVolumeField::faceToCell, plain and weighted, bit-identical;
    one weight without the other is rejected like the reference's overloads."""
    om, hm, _ = meshes
    N, F = om.n_cells, om.n_faces
    rng = np.random.default_rng(12)
    fv = rng.standard_normal(2 * F)
    cw, fw = np.exp(rng.standard_normal(N)), np.exp(rng.standard_normal(F))
    d_fv, d_out = hip.DeviceArray(fv), hip.DeviceArray.zeros(2 * N)
    hm.face_to_cell(d_fv, d_out)
    assert same(d_out.get(), oracle.face_to_cell(om, fv).reshape(-1))
    hm.face_to_cell(d_fv, d_out, hip.DeviceArray(cw), hip.DeviceArray(fw))
    assert same(d_out.get(), oracle.face_to_cell(om, fv, cw, fw).reshape(-1))
    with pytest.raises(hip.PhaseHipError):
        hm.face_to_cell(d_fv, d_out, hip.DeviceArray(cw), None)


def test_copy_device(hip):
    """phase_hip_copy_device: a byte-exact copy (with a ragged tail of the
    grid-stride loop) and EINVAL for sizes that are not 16-byte multiples."""
    src = np.random.default_rng(5).standard_normal(1_000_002)
    a, b = hip.DeviceArray(src), hip.DeviceArray.zeros(len(src))
    b.copy_from(a)
    assert same(b.get(), src)
    c = hip.DeviceArray(np.zeros(3))
    with pytest.raises(hip.PhaseHipError):
        hip.DeviceArray(np.zeros(3)).copy_from(c)  # 24 bytes
