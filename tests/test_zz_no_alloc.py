"""The allocation-free SpMV entry point (phase_hip_solver_spmv_noalloc), called
through the C ABI, against phase_hip_solver_spmv on the same system: identical
bits in FP64 and FP32, on two calls in a row, the second on a smaller system
(the shared scratch keeps the first call's contents, which must not leak into
the second), and the documented errors."""
import ctypes

import numpy as np
import pytest

from tests import helpers as H

pytestmark = pytest.mark.gpu


def _system(oracle, cells, seed):
    xy, cptr, cind = H.delaunay_case(cells, 24, seed=seed)
    m, (rp, ci, v), b, g, bc = H.pressure_system(oracle, xy, cptr, cind, sigma=1.0, seed_g=seed + 1)
    return rp, ci, v


def _noalloc(hip, s, x, single):
    lib = hip.load_library()
    y = np.full(s.rank, np.nan)
    rc = lib.phase_hip_solver_spmv_noalloc(s._h, int(single), np.ascontiguousarray(x, np.float64), y)
    return rc, y


def test_noalloc_spmv_matches_allocating_entry_point(hip, oracle):
    rng = np.random.default_rng(17)
    for cells, seed in ((6000, 51), (1500, 53)):  # second call: smaller input
        rp, ci, v = _system(oracle, cells, seed)
        s = hip.HipSparseMatrixSolver(0)
        s.set(rp, ci, v)
        x = rng.standard_normal(s.rank)
        rc, y = _noalloc(hip, s, x, single=False)
        assert rc == 0, hip.load_library().phase_hip_last_error(s._h)
        assert np.array_equal(y, s.spmv(x))
        assert np.array_equal(y, oracle.spmv(rp, ci, v, x))
        y32 = s.spmv(x, single=True)  # builds the FP32 copy
        rc, z = _noalloc(hip, s, x, single=True)
        assert rc == 0 and np.array_equal(z, y32)
        s.close()


def test_noalloc_spmv_errors(hip, oracle):
    rp, ci, v = _system(oracle, 1500, 55)
    s = hip.HipSparseMatrixSolver(0)
    s.set(rp, ci, v)
    x = np.ones(s.rank)
    rc, _ = _noalloc(hip, s, x, single=True)  # no FP32 copy yet
    assert rc == -3  # PHASE_HIP_ESTATE
    s.close()
    big = hip.HipSparseMatrixSolver(0)
    big.set_rank((1 << 21) + 1)
    lib = hip.load_library()
    y = np.zeros(4)
    rc = lib.phase_hip_solver_spmv_noalloc(big._h, 0, np.zeros(4), y)
    assert rc == -1  # PHASE_HIP_EINVAL: beyond the scratch, rejected before anything else
    assert b"exceeds the allocation-free scratch" in lib.phase_hip_last_error(big._h)
    big.close()
