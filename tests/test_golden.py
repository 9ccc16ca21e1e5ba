"""Golden fixtures (tests/golden/*.npz, written by tests/golden/make_golden.py
from an independent numpy two-point-flux assembly and scipy's sparse LU):
the oracle and the device path must both reproduce the direct solution."""
import os

import numpy as np
import pytest

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
NAMES = ["rect8", "rect32", "cavity128", "tri2k"]


def load(name):
    with np.load(os.path.join(HERE, name + ".npz"), allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def face_inputs(O, fx, faces):
    sigma, kx, ky, ph = fx["coeff"]
    mid = faces["centroid"]
    gamma = np.exp(sigma * np.sin(kx * mid[:, 0] + ph) * np.cos(ky * mid[:, 1]))
    bc = np.where((faces["rcell"] < 0) & (mid[:, 1] == 0.0), O.BC_FIXED, O.BC_NORMAL_GRADIENT).astype(np.int32)
    phi = np.sin(3 * mid[:, 0])
    return gamma, bc, phi


@pytest.mark.parametrize("name", NAMES)
def test_oracle_reproduces_direct_solution(oracle, name):
    fx = load(name)
    m = oracle.Mesh(fx["xy"], fx["cptr"], fx["cind"])
    gamma, bc, phi = face_inputs(oracle, fx, m.faces())
    e = oracle.laplacian(m, 3, gamma_face=gamma, bc_type=bc, phi_face=phi)
    e.eq(fx["src"])
    rp, ci, v, rhs = e.get()
    assert np.allclose(-rhs, fx["b"], rtol=0, atol=1e-12 * np.abs(fx["b"]).max())
    x, it, rr, rc = oracle.bicgstab_ilu0(rp, ci, v, -rhs, tol=1e-12, max_iters=20000)
    assert rc == 0 and rr <= 1e-12
    err = np.linalg.norm(x - fx["x"]) / np.linalg.norm(fx["x"])
    assert err < 1e-8, err


@pytest.mark.parametrize("name", NAMES)
def test_oracle_omp_solver_reproduces_direct_solution(oracle, name):
    """The multi-threaded CPU baseline solves the same systems."""
    fx = load(name)
    m = oracle.Mesh(fx["xy"], fx["cptr"], fx["cind"])
    gamma, bc, phi = face_inputs(oracle, fx, m.faces())
    e = oracle.laplacian(m, 3, gamma_face=gamma, bc_type=bc, phi_face=phi)
    e.eq(fx["src"])
    rp, ci, v, rhs = e.get()
    x, it, rr, rc = oracle.bicgstab_ilu0_omp(rp, ci, v, -rhs, tol=1e-12, max_iters=20000, threads=4)
    assert rc == 0 and rr <= 1e-12
    assert np.linalg.norm(x - fx["x"]) / np.linalg.norm(fx["x"]) < 1e-8


@pytest.mark.parametrize("name", NAMES)
def test_oracle_ilut_reproduces_direct_solution(oracle, name):
    fx = load(name)
    m = oracle.Mesh(fx["xy"], fx["cptr"], fx["cind"])
    gamma, bc, phi = face_inputs(oracle, fx, m.faces())
    e = oracle.laplacian(m, 3, gamma_face=gamma, bc_type=bc, phi_face=phi)
    e.eq(fx["src"])
    rp, ci, v, rhs = e.get()
    x, it, rr, rc = oracle.bicgstab_ilut(rp, ci, v, -rhs, tol=1e-12, max_iters=20000)
    assert rc == 0 and rr <= 1e-12
    assert np.linalg.norm(x - fx["x"]) / np.linalg.norm(fx["x"]) < 1e-8


def test_ilu0_reproduces_a_on_pattern(oracle):
    """ILU(0) is exact on the pattern: (L U)_ij = a_ij for every (i,j) in A."""
    fx = load("tri2k")
    m = oracle.Mesh(fx["xy"], fx["cptr"], fx["cind"])
    gamma, bc, phi = face_inputs(oracle, fx, m.faces())
    rp, ci, v, rhs = oracle.laplacian(m, 3, gamma_face=gamma, bc_type=bc, phi_face=phi).get()
    from tests import helpers as H

    A = H.compact_sorted(rp, ci, v)
    fac = oracle.ilu0(A.indptr.astype(np.int32), A.indices.astype(np.int32), A.data)
    n = A.shape[0]
    F = A.copy()
    F.data = fac
    Fd = F.toarray()
    L = np.tril(Fd, -1) + np.eye(n)
    U = np.triu(Fd)
    LU = L @ U
    Ad = A.toarray()
    mask = Ad != 0
    assert np.allclose(LU[mask], Ad[mask], rtol=1e-12, atol=1e-12 * np.abs(Ad).max())


@pytest.mark.gpu
@pytest.mark.parametrize("name", NAMES)
@pytest.mark.parametrize("mode", ["fp64", "mixed"])
def test_device_reproduces_direct_solution(hip, oracle, name, mode):
    fx = load(name)
    om = oracle.Mesh(fx["xy"], fx["cptr"], fx["cind"])
    gamma, bc, phi = face_inputs(oracle, fx, om.faces())
    hm = hip.HipMesh(fx["xy"], fx["cptr"], fx["cind"])
    s = hip.HipSparseMatrixSolver(0, tolerance=1e-11, maxIters=20000, mode=mode)
    s.assemble_laplacian(hm, 3, gamma_face=gamma, bc_type=bc, phi_face=phi)
    s.sub_rhs(fx["src"])
    rel = s.solve()
    assert rel <= 1e-11
    x = s.x()
    assert np.linalg.norm(x - fx["x"]) / np.linalg.norm(fx["x"]) < 1e-7
