"""The learned-preconditioner workload at a size the oracle covers: the
pressure-correction Poisson on a ~6k-cell Delaunay mesh of the unit square with
lognormal face coefficients of sigma = 2, pCorr = 0 on y-.  Every method
(FP64 BiCGStab + ILU(0), FP64 CG, FP32-inner CG with FP64 refinement, learned)
reaches the FP64 true relative residual 1e-8 and agrees with a direct solve
(scipy splu) within cond(A) * 1.5e-8, with cond(A) computed from the extreme
eigenvalues of this system.  The learned network is trained on pairs captured
from this system's own FP64 solves and used on a held-out right-hand side."""
import numpy as np
import pytest
import scipy.sparse.linalg as spla

from tests import helpers as H

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def c4(oracle):
    xy, cptr, cind = H.delaunay_case(6000, 32, seed=31)
    m, (rp, ci, v), b, g, bc = H.pressure_system(oracle, xy, cptr, cind, sigma=2.0, seed_g=32, seed_b=33)
    A = H.compact_sorted(rp, ci, v)
    S = (-A).tocsc()  # symmetric positive definite
    lmax = spla.eigsh(S, k=1, which="LA", return_eigenvectors=False)[0]
    lmin = spla.eigsh(S, k=1, sigma=0.0, which="LM", return_eigenvectors=False)[0]
    x_ref = spla.splu(A.tocsc()).solve(b)
    return rp, ci, v, b, A, lmax / lmin, x_ref


def _check(A, b, x, x_ref, kappa):
    assert np.linalg.norm(b - A @ x) <= 1.5e-8 * np.linalg.norm(b)
    assert np.linalg.norm(x - x_ref) <= kappa * 1.5e-8 * np.linalg.norm(x_ref)


@pytest.mark.parametrize("mode,solver", [("fp64", "bicgstab"), ("fp64", "cg"), ("mixed", "cg"),
                                         ("mixed", "bicgstab")])
def test_sigma2_methods_match_direct_solve(hip, c4, mode, solver):
    rp, ci, v, b, A, kappa, x_ref = c4
    assert kappa > 1e4  # sigma = 2 coefficients: a badly conditioned system
    s = hip.HipSparseMatrixSolver(0, tolerance=1e-8, maxIters=40000, mode=mode, solver=solver)
    s.set(rp, ci, v)
    s.set_rhs(b)
    assert s.solve() <= 1e-8
    _check(A, b, s.x(), x_ref, kappa)


def test_sigma2_learned_on_held_out_rhs(hip, c4):
    rp, ci, v, b, A, kappa, x_ref = c4
    n = len(b)
    s = hip.HipSparseMatrixSolver(0, tolerance=1e-8, maxIters=40000, mode="fp64")
    s.set(rp, ci, v)
    s.learned_capture(25, 300)
    rng = np.random.default_rng(41)
    for _ in range(8):
        bb = rng.standard_normal(n)
        s.set_rhs(bb - bb.mean())
        s.set_guess(np.zeros(n))
        s.solve()
    assert s.learned_n_pairs() >= 40
    first = s.learned_train(100, lr=0.0, seed=3)
    s.learned_train(400, lr=1e-3, seed=1)
    last = s.learned_train(100, lr=0.0, seed=3)
    assert np.isfinite(first) and last < first
    s.setup("mode", "learned")
    s.set_rhs(b)  # not among the captured right-hand sides
    s.set_guess(np.zeros(n))
    assert s.solve() <= 1e-8
    _check(A, b, s.x(), x_ref, kappa)
