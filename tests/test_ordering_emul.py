"""Host logic of the device solver, checked on the CPU.

tests/emul/solver_emul.cpp replays the gfx950 kernels' index arithmetic and
schedule on the product's ordering (bounds and dependency order asserted on
every access); its SpMV, ILU(0) factor and forward sweep must equal the oracle's bit for bit
on the same permuted matrix, the backward sweep to rounding.
"""
import ctypes

import numpy as np
import pytest

from tests import helpers as H


def _run(emul, rp, ci, v, rhs_solver, x_solver, block_rows=64, xy=None):
    n = len(rp) - 1
    cap = 2 * len(v) + n + 16
    perm = np.empty(n, np.int32)
    frp = np.empty(n + 1, np.int32)
    fci = np.empty(cap, np.int32)
    fv = np.empty(cap)
    y, x, sp = np.empty(n), np.empty(n), np.empty(n)
    info = np.zeros(8, np.int32)
    err = ctypes.create_string_buffer(256)
    xyp = None if xy is None else np.ascontiguousarray(xy, np.float64)
    rc = emul.emul_run(n, rp, ci, v, block_rows, perm, frp, fci, fv, cap, rhs_solver, y, x, x_solver, sp, info,
                       err, 256, None if xyp is None else xyp.ctypes.data)
    assert rc == 0, err.value.decode()
    return dict(perm=perm, frp=frp, fci=fci[:frp[-1]], fv=fv[:frp[-1]], y=y, x=x, spmv=sp, info=info)


def _check(oracle, emul, rp, ci, v, block_rows=64, seed=0, xy=None):
    n = len(rp) - 1
    rng = np.random.default_rng(seed)
    # first pass only to learn the permutation
    dummy = np.zeros(n)
    r0 = _run(emul, rp, ci, v, dummy, dummy, block_rows, xy)
    perm = r0["perm"]
    assert sorted(perm.tolist()) == list(range(n))
    x_orig = rng.standard_normal(n)
    b_solver = rng.standard_normal(n)
    r = _run(emul, rp, ci, v, b_solver, x_orig[perm], block_rows, xy)
    # SpMV: same rows, same in-row order -> bitwise equal
    y_ref = oracle.spmv(rp, ci, v, x_orig)
    assert np.array_equal(r["spmv"], y_ref[perm])
    # ILU(0) of the permuted matrix -> bitwise equal factors
    A = H.permuted(H.compact_sorted(rp, ci, v), perm)
    fac = oracle.ilu0(A.indptr.astype(np.int32), A.indices.astype(np.int32), A.data)
    assert np.array_equal(r["frp"], A.indptr)
    assert np.array_equal(r["fci"], A.indices)
    assert np.array_equal(r["fv"], fac)
    # forward sweep -> bitwise equal (same column order); backward sweep: the
    # device multiplies by the inverse diagonal and subtracts the later-block
    # terms as one sum, so it agrees with the textbook back substitution to
    # rounding (a wrong or dropped term would be O(1))
    y, x = oracle.lu_solve(A.indptr, A.indices, fac, b_solver)
    assert np.array_equal(r["y"], y)
    assert np.linalg.norm(r["x"] - x) <= 1e-12 * np.linalg.norm(x)
    return r


def test_emul_rectilinear(oracle, emul):
    xy, cptr, cind = H.rect_case(24, 20)
    m, (rp, ci, v), b, g, bc = H.pressure_system(oracle, xy, cptr, cind)
    r = _check(oracle, emul, rp, ci, v)
    assert r["info"][0] >= (24 * 20) // 64


@pytest.mark.parametrize("block_rows", [1, 7, 32, 64])
def test_emul_delaunay_block_sizes(oracle, emul, block_rows):
    xy, cptr, cind = H.delaunay_case(3000, 24, seed=5)
    m, (rp, ci, v), b, g, bc = H.pressure_system(oracle, xy, cptr, cind, sigma=2.0)
    r = _check(oracle, emul, rp, ci, v, block_rows=block_rows, seed=block_rows)
    nb, ncol = r["info"][0], r["info"][1]
    # aggregates of <= block_rows rows, packed per colour into blocks of <= 64
    assert nb >= 3000 // 64
    assert ncol >= 2


@pytest.mark.parametrize("case", ["delaunay", "rectilinear"])
def test_emul_brick_ordering(oracle, emul, case):
    """With cell centroids the rows fall into running-bond brick aggregates in
    exactly three colours; layout, factor and sweeps keep the bit-exact checks."""
    if case == "delaunay":
        xy, cptr, cind = H.delaunay_case(6000, 32, seed=9)
    else:
        xy, cptr, cind = H.rect_case(60, 50)
    m, (rp, ci, v), b, g, bc = H.pressure_system(oracle, xy, cptr, cind, sigma=1.0)
    _, cc = m.cells()
    r = _check(oracle, emul, rp, ci, v, block_rows=32, seed=11, xy=cc.reshape(-1))
    assert r["info"][1] == 3


def test_emul_brick_fallback(oracle, emul):
    """Centroids that do not belong to the rows (shuffled) couple same-colour
    bricks: the ordering falls back to graph aggregates and stays correct."""
    xy, cptr, cind = H.delaunay_case(3000, 24, seed=5)
    m, (rp, ci, v), b, g, bc = H.pressure_system(oracle, xy, cptr, cind, sigma=1.0)
    _, cc = m.cells()
    cc = cc[np.random.default_rng(1).permutation(len(cc))]
    r = _check(oracle, emul, rp, ci, v, block_rows=32, seed=12, xy=cc.reshape(-1))
    assert r["info"][1] >= 2


def test_emul_nonsymmetric_pattern(oracle, emul):
    """A random nonsymmetric, diagonally dominant pattern with empty slots."""
    rng = np.random.default_rng(11)
    n, per = 700, 6
    rp = np.arange(n + 1, dtype=np.int32) * per
    ci = np.full(n * per, -1, np.int32)
    v = np.zeros(n * per)
    for i in range(n):
        cols = {i}
        for _ in range(rng.integers(0, 4)):
            cols.add(int(rng.integers(max(0, i - 40), min(n, i + 40))))
        cols = list(cols)
        rng.shuffle(cols)
        for k, c in enumerate(cols):
            ci[i * per + k] = c
            v[i * per + k] = 10.0 if c == i else rng.uniform(-1, 1)
    _check(oracle, emul, rp, ci, v, seed=3)


def test_emul_rejects_out_of_range(emul):
    rp = np.array([0, 1, 2], np.int32)
    ci = np.array([0, 5], np.int32)
    v = np.ones(2)
    n = 2
    perm = np.empty(n, np.int32)
    frp = np.empty(n + 1, np.int32)
    fci = np.empty(16, np.int32)
    fv = np.empty(16)
    z = np.zeros(n)
    info = np.zeros(8, np.int32)
    err = ctypes.create_string_buffer(256)
    rc = emul.emul_run(n, rp, ci, v, 64, perm, frp, fci, fv, 16, z, z.copy(), z.copy(), z, z.copy(), info, err, 256,
                       None)
    assert rc == 1 and b"out of range" in err.value
