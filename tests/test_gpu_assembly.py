"""GPU mesh geometry and fv::laplacian assembly, bit-exact against the oracle.

The device computes geometry and coefficients with the reference's formulas
in the reference's operation order (no contraction), so the padded CSR, its
column order, values and rhs must equal the oracle's restatement of
Laplacian.h:7-167 bit for bit.
"""
import numpy as np
import pytest

from phase_amd import meshgen as G
from tests import helpers as H

pytestmark = pytest.mark.gpu


def _meshes():
    yield "rect", H.rect_case(13, 9)
    yield "tri", H.delaunay_case(2500, 20, seed=3)


@pytest.mark.parametrize("name,case", list(_meshes()))
def test_geometry_bitwise(hip, oracle, name, case):
    xy, cptr, cind = case
    om = oracle.Mesh(xy, cptr, cind)
    hm = hip.HipMesh(xy, cptr, cind)
    assert (hm.n_cells, hm.n_faces, hm.n_interior, hm.n_boundary) == (om.n_cells, om.n_faces, om.n_interior,
                                                                        om.n_boundary)
    of = om.faces()
    hf = hm.faces()
    for k in ("lnode", "rnode", "lcell", "rcell"):
        assert np.array_equal(of[k], hf[k])
    vol, cc, fc, sf = hm.geometry()
    ovol, occ = om.cells()
    assert np.array_equal(vol, ovol)
    assert np.array_equal(cc, occ)
    assert np.array_equal(fc, of["centroid"])
    assert np.array_equal(sf, of["sf"])


@pytest.mark.parametrize("variant", [0, 1, 2, 3])
@pytest.mark.parametrize("name,case", list(_meshes()))
def test_laplacian_padded_bitwise(hip, oracle, name, case, variant):
    xy, cptr, cind = case
    om = oracle.Mesh(xy, cptr, cind)
    hm = hip.HipMesh(xy, cptr, cind)
    rng = np.random.default_rng(variant + 10)
    F, N = om.n_faces, om.n_cells
    bc, pid = H.bc_from_patch(oracle, om, xy, fixed_patches=(2, 1))
    bc[pid == 3] = oracle.BC_SYMMETRY
    gf = np.exp(rng.standard_normal(F))
    g0 = np.exp(rng.standard_normal(F))
    phi_f = rng.standard_normal(F)
    phi0_c = rng.standard_normal(N)
    phi0_f = rng.standard_normal(F)
    kw = dict(gamma=0.37, gamma_face=gf, gamma0_face=g0, theta=0.6, bc_type=bc, phi_face=phi_f,
              phi0_cell=phi0_c, phi0_face=phi0_f)
    ref = oracle.laplacian(om, variant, **kw).get()
    s = hip.HipSparseMatrixSolver(0)
    s.assemble_laplacian(hm, variant, **kw)
    got = s.export_padded()
    for a, b in zip(got, ref):
        assert np.array_equal(a, b)


def test_pressure_equation_solve(hip, oracle):
    """fv::laplacian(gamma_f, p) == src  assembled and solved on the device."""
    xy, cptr, cind = H.delaunay_case(5000, 30, seed=9)
    om = oracle.Mesh(xy, cptr, cind)
    hm = hip.HipMesh(xy, cptr, cind)
    bc, pid = H.bc_from_patch(oracle, om, xy, fixed_patches=(2,))
    gf = G.lognormal_face_coeffs(om.n_faces, 1.0, seed=2)
    src = G.zero_mean_normal(om.n_cells, seed=3)
    e = oracle.laplacian(om, 3, gamma_face=gf, bc_type=bc, phi_face=np.zeros(om.n_faces))
    e.eq(src)
    rp, ci, v, rhs = e.get()
    s = hip.HipSparseMatrixSolver(0, tolerance=1e-9, maxIters=5000)
    s.assemble_laplacian(hm, 3, gamma_face=gf, bc_type=bc, phi_face=np.zeros(om.n_faces))
    s.sub_rhs(src)
    grp, gci, gv, grhs = s.export_padded()
    assert np.array_equal(grhs, rhs)
    s.solve()
    x = s.x()
    A = H.compact_sorted(rp, ci, v)
    assert np.linalg.norm(-rhs - A @ x) / np.linalg.norm(rhs) <= 1.5e-9


def test_unsupported_boundary(hip, oracle):
    xy, cptr, cind = H.rect_case(4, 4)
    hm = hip.HipMesh(xy, cptr, cind)
    bc = np.full(hm.n_faces, hip.BC_OUTFLOW, np.int32)
    s = hip.HipSparseMatrixSolver(0)
    with pytest.raises(hip.PhaseHipError) as ei:
        s.assemble_laplacian(hm, 1, bc_type=bc, phi_face=np.zeros(hm.n_faces))
    assert ei.value.code == -6
