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


def _hierarchy(emul, rp, ci, xy, fine_rows, coarse_rows, dense_max, path):
    n = len(rp) - 1
    err = ctypes.create_string_buffer(256)
    xyp = None if xy is None else np.ascontiguousarray(xy, np.float64).ctypes.data_as(ctypes.c_void_p)
    emul.emul_hierarchy.restype = ctypes.c_int
    rc = emul.emul_hierarchy(ctypes.c_int(n), np.ascontiguousarray(rp, np.int32).ctypes.data_as(ctypes.c_void_p),
                             np.ascontiguousarray(ci, np.int32).ctypes.data_as(ctypes.c_void_p), xyp,
                             ctypes.c_int(fine_rows), ctypes.c_int(coarse_rows), ctypes.c_int(dense_max),
                             str(path).encode(), err, ctypes.c_int(256))
    assert rc >= 0, err.value.decode()
    d = np.fromfile(str(path), np.int32)
    o, levels = 0, []
    for _ in range(rc):
        n = d[o]; o += 1
        perm = d[o:o + n]; o += n
        na = d[o]; o += 1
        agg_ptr = d[o:o + na + 1]; o += na + 1
        crp = d[o:o + na + 1]; o += na + 1
        nnz = d[o]; o += 1
        cci = d[o:o + nnz]; o += nnz
        gp = d[o:o + nnz + 1]; o += nnz + 1
        gs = d[o:o + gp[-1]]; o += gp[-1]
        levels.append(dict(n=n, perm=perm, aggPtr=agg_ptr, crp=crp, cci=cci, gp=gp, gs=gs))
    last = d[o]
    return levels, last


@pytest.mark.parametrize("geometric", [True, False])
def test_emul_multigrid_hierarchy_is_galerkin(oracle, emul, tmp_path, geometric):
    """The multigrid hierarchy the product builds (ordering aggregates, then
    buildCoarse per level): every level coarsens, the last is within the dense
    limit, aggregates are contiguous solver-row ranges covering every row once,
    and the coarse matrices summed through the Galerkin lists equal P^T A P
    with P the aggregate indicator (computed here with scipy)."""
    import scipy.sparse as sp

    xy, cptr, cind = H.delaunay_case(20000, 64, seed=12)
    m, (rp, ci, v), b, g, bc = H.pressure_system(oracle, xy, cptr, cind, sigma=1.0)
    _, cc = m.cells()
    levels, last = _hierarchy(emul, rp, ci, cc if geometric else None, 32, 16, 64, tmp_path / "h.bin")
    assert len(levels) >= 2 and last <= 64
    vals, lrp, lci = v, rp, ci
    for L in levels:
        n = L["n"]
        perm = L["perm"]
        assert sorted(perm.tolist()) == list(range(n))
        ap = L["aggPtr"]
        assert ap[0] == 0 and ap[-1] == n and np.all(np.diff(ap) > 0)
        na = len(ap) - 1
        assert na < n
        iperm = np.empty(n, np.int64)
        iperm[perm] = np.arange(n)
        agg_of_solver_row = np.repeat(np.arange(na), np.diff(ap))
        P = sp.csr_matrix((np.ones(n), (np.arange(n), agg_of_solver_row[iperm])), shape=(n, na))
        rows = np.repeat(np.arange(n), np.diff(lrp))
        keep = lci >= 0
        A = sp.csr_matrix((vals[keep], (rows[keep], lci[keep])), shape=(n, n))
        Ac = (P.T @ A @ P).tocsr()
        Ac.sort_indices()
        cvals = np.array([vals[L["gs"][L["gp"][q]:L["gp"][q + 1]]].sum() for q in range(len(L["cci"]))])
        got = sp.csr_matrix((cvals, L["cci"], L["crp"]), shape=(na, na))
        got.sort_indices()
        assert np.array_equal(got.indptr, Ac.indptr) and np.array_equal(got.indices, Ac.indices)
        assert np.allclose(got.data, Ac.data, rtol=1e-12, atol=1e-12 * np.abs(Ac.data).max())
        vals, lrp, lci = cvals, L["crp"], L["cci"]


def test_emul_wide_coarse_rows(oracle, emul, tmp_path):
    """Coarse multigrid matrices from graph aggregates have rows that couple to
    more rows of their own or of later ILU blocks than the sweep kernels keep
    in registers (8); the layout still reproduces the oracle's SpMV, ILU(0)
    factor and sweeps on such a matrix."""
    xy, cptr, cind = H.delaunay_case(250000, 128, seed=14)
    m, (rp, ci, v), b, g, bc = H.pressure_system(oracle, xy, cptr, cind, sigma=1.0)
    levels, last = _hierarchy(emul, rp, ci, None, 32, 16, 256, tmp_path / "h.bin")
    widest = 0
    vals, lrp, lci = v, rp, ci
    for L in levels:
        cvals = np.array([vals[L["gs"][L["gp"][q]:L["gp"][q + 1]]].sum() for q in range(len(L["cci"]))])
        vals, lrp, lci = cvals, L["crp"].astype(np.int32), L["cci"].astype(np.int32)
        if len(lrp) - 1 > 256:
            r = _check(oracle, emul, lrp, lci, vals, block_rows=16, seed=3)
            widest = max(widest, int(r["info"][6]))
    assert widest > 8


def test_emul_block_jacobi_ordering(oracle, emul):
    """The block-Jacobi ordering of coarse multigrid levels: one colour, the
    ILU(0) factor of the block-diagonal part only (couplings between blocks
    dropped), the SpMV still over the whole matrix; the factor and both sweeps
    equal the oracle's on the block-diagonal restriction of the permuted
    matrix."""
    import scipy.sparse as sp

    xy, cptr, cind = H.delaunay_case(3000, 24, seed=15)
    m, (rp, ci, v), b, g, bc = H.pressure_system(oracle, xy, cptr, cind, sigma=2.0)
    n = len(rp) - 1
    emul.emul_set_block_jacobi.restype = None
    emul.emul_set_block_jacobi(1)
    try:
        perm = _run(emul, rp, ci, v, np.zeros(n), np.zeros(n), 16)["perm"]
        rng = np.random.default_rng(4)
        x_orig, b_solver = rng.standard_normal(n), rng.standard_normal(n)
        r = _run(emul, rp, ci, v, b_solver, x_orig[perm], 16)
    finally:
        emul.emul_set_block_jacobi(0)
    assert r["info"][1] == 1
    assert np.array_equal(r["spmv"], oracle.spmv(rp, ci, v, x_orig)[perm])
    A = H.permuted(H.compact_sorted(rp, ci, v), perm)
    pat = sp.csr_matrix((np.ones(len(r["fci"])), r["fci"], r["frp"]), shape=(n, n))
    Ab = A.multiply(pat).tocsr()
    Ab.sort_indices()
    assert np.array_equal(Ab.indptr, r["frp"]) and np.array_equal(Ab.indices, r["fci"])
    assert Ab.nnz < A.nnz  # couplings between blocks were dropped
    fac = oracle.ilu0(Ab.indptr.astype(np.int32), Ab.indices.astype(np.int32), Ab.data)
    assert np.array_equal(r["fv"], fac)
    y, x = oracle.lu_solve(Ab.indptr, Ab.indices, fac, b_solver)
    assert np.array_equal(r["y"], y)
    assert np.linalg.norm(r["x"] - x) <= 1e-12 * np.linalg.norm(x)
