"""Flexible restarted GMRES (setup "solver" = "GMRES", the Krylov method of
the reference's MueLu backend, src/Math/TrilinosMueluSparseMatrixSolver.cpp:47)
on the GPU: with ILU(0) and with the multigrid cycle (K cycles, which vary
between applications: the flexible form is what makes that valid), restart
lengths from 5 to 64 (30 and 64 with ILU(0) alone), the FP64 true relative residual reaches the goal and the
solution agrees with a direct solve (scipy splu) within cond(A) * 1.5e-8; the
restarted form keeps the residual it reports; repeated solves are
bit-identical; bad restart lengths are rejected.  Mode mixed (FP32 cycles
restarted from the FP64 residual) reaches the same FP64 goal, and continues
in FP64 when its FP32 cycles cannot run."""
import numpy as np
import pytest
import scipy.sparse.linalg as spla

from tests import helpers as H

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def sys1(oracle):
    xy, cptr, cind = H.delaunay_case(6000, 32, seed=7)
    m, (rp, ci, v), b, g, bc = H.pressure_system(oracle, xy, cptr, cind, sigma=2.0)
    A = H.compact_sorted(rp, ci, v)
    S = (-A).tocsc()
    lmax = spla.eigsh(S, k=1, which="LA", return_eigenvectors=False)[0]
    lmin = spla.eigsh(S, k=1, sigma=0.0, which="LM", return_eigenvectors=False)[0]
    return rp, ci, v, b, A, lmax / lmin, spla.splu(A.tocsc()).solve(b)


def _solve(hip, rp, ci, v, b, **kw):
    s = hip.HipSparseMatrixSolver(0, tolerance=1e-8, maxIters=20000, solver="gmres", **kw)
    s.set(rp, ci, v)
    s.set_rhs(b)
    rel = s.solve()
    return s, rel, s.x()


# GMRES(5) with ILU(0) alone stagnates on this system (so does a textbook
# FGMRES(5) with scipy's ILU(0): 3.9e-2 after 20,000 iterations)
@pytest.mark.parametrize("prec,restart,mode", [("ilu", 30, "fp64"), ("ilu", 64, "fp64"), ("multigrid", 5, "fp64"),
                                               ("multigrid", 30, "fp64"), ("multigrid", 64, "fp64"),
                                               ("ilu", 30, "mixed"), ("multigrid", 5, "mixed"),
                                               ("multigrid", 30, "mixed")])
def test_gmres_matches_direct(hip, sys1, prec, restart, mode):
    rp, ci, v, b, A, kappa, x_ref = sys1
    s, rel, x = _solve(hip, rp, ci, v, b, preconditioner=prec, gmresRestart=restart, mode=mode)
    assert s.stats()["fallbacks"] == 0
    assert rel <= 1e-8
    true = np.linalg.norm(b - A @ x) / np.linalg.norm(b)
    assert true <= 1.5e-8 and abs(true - rel) <= 1e-3 * rel + 1e-14
    assert np.linalg.norm(x - x_ref) <= kappa * 1.5e-8 * np.linalg.norm(x_ref)
    print(prec, restart, mode, "iterations", s.n_iters(), flush=True)


def test_mixed_gmres_continues_in_fp64(hip, sys1):
    """A system scaled by 1e40 is the same system in FP64, but its FP32
    factor overflows: the FP32 cycles break down at once and the solve
    continues in FP64 to the same goal, counting one fallback."""
    rp, ci, v, b, A, kappa, x_ref = sys1
    for prec in ("ilu", "multigrid"):
        s, rel, x = _solve(hip, rp, ci, v * 1e40, b * 1e40, preconditioner=prec, mode="mixed")
        assert rel <= 1e-8 and s.stats()["fallbacks"] == 1
        assert np.linalg.norm(b - A @ x) <= 1.5e-8 * np.linalg.norm(b)


def test_gmres_iteration_counts(hip, sys1):
    """Multigrid cuts the GMRES count against ILU(0) alone, and with it a
    longer restart needs no more iterations than a short one."""
    rp, ci, v, b, A, kappa, x_ref = sys1
    it = {}
    for prec, r in (("ilu", 60), ("multigrid", 10), ("multigrid", 60)):
        it[prec, r] = _solve(hip, rp, ci, v, b, preconditioner=prec, gmresRestart=r)[0].n_iters()
    assert 3 * it["multigrid", 60] <= it["ilu", 60], it
    assert it["multigrid", 60] <= it["multigrid", 10], it


def test_gmres_repeatable_and_rejects_bad_restart(hip, sys1):
    rp, ci, v, b, A, kappa, x_ref = sys1
    for mode in ("fp64", "mixed"):
        s1, _, x1 = _solve(hip, rp, ci, v, b, preconditioner="multigrid", mode=mode)
        s2, _, x2 = _solve(hip, rp, ci, v, b, preconditioner="multigrid", mode=mode)
        assert np.array_equal(x1, x2) and s1.n_iters() == s2.n_iters()
    for bad in (0, 65):
        with pytest.raises(hip.PhaseHipError):
            hip.HipSparseMatrixSolver(0, solver="gmres", gmresRestart=bad)
