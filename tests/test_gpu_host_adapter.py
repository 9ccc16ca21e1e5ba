"""The C++ plugin adapter (phase_amd/host/HipSparseMatrixSolver) created by
name as a std::shared_ptr to the plugin base class and driven the way
FiniteVolumeEquation<T>::solve drives a SparseMatrixSolver, on the GPU; the
same system then goes through every other set() overload of the interface
(CoefficientList, SparseEntry list, triplets) and solve(x0)."""
import os
import subprocess

import numpy as np
import pytest

from tests import helpers as H

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "tests", "host", "_build", "host_driver")


def test_driver_is_built():
    assert os.path.exists(DRIVER), "build() must produce tests/host/_build/host_driver"


@pytest.mark.gpu
def test_cpp_adapter_solves_padded_equation(hip, oracle, tmp_path):
    xy, cptr, cind = H.delaunay_case(3000, 24, seed=12)
    m, (rp, ci, v), b, g, bc = H.pressure_system(oracle, xy, cptr, cind)
    rhs = -b  # stored equation A x + rhs = 0
    n, ns = len(rp) - 1, len(ci)
    inp = tmp_path / "in.bin"
    out = tmp_path / "out.bin"
    with open(inp, "wb") as fh:
        np.array([n, ns], np.int32).tofile(fh)
        rp.astype(np.int32).tofile(fh)
        ci.astype(np.int32).tofile(fh)
        v.astype(np.float64).tofile(fh)
        rhs.astype(np.float64).tofile(fh)
    r = subprocess.run([DRIVER, str(inp), str(out)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.count("exception ok") == 5 and "iterations =" in r.stdout
    data = np.fromfile(out, np.float64).reshape(4, n + 1)
    A = H.compact_sorted(rp, ci, v)
    for form in range(4):  # padded CSR, CoefficientList, SparseEntry, triplets
        err, x = data[form, 0], data[form, 1:]
        assert err <= 1e-10, form
        assert np.linalg.norm(b - A @ x) / np.linalg.norm(b) <= 1.5e-10, form
