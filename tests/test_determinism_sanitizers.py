"""Repeatability and memory safety of the CPU-side code.

* The oracle's assembly and BiCGStab + ILU(0), and the emulator of the device
  ordering and sweeps, give bit-identical results when run twice on the same
  inputs (fixed summation orders; nothing depends on addresses or timing).
* oracle/phase_oracle.cpp and tests/emul/solver_emul.cpp (with the product's
  ordering.cpp) built with AddressSanitizer and UndefinedBehaviorSanitizer run
  a Delaunay case cleanly in a child process (oracle/Makefile target `asan`).
"""
import os
import subprocess
import sys

import numpy as np
import pytest

from tests import helpers as H
from tests.test_ordering_emul import _run

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_oracle_repeatable(oracle):
    xy, cptr, cind = H.delaunay_case(3000, 24, seed=9)
    outs = []
    for _ in range(2):
        m, (rp, ci, v), b, g, bc = H.pressure_system(oracle, xy, cptr, cind, sigma=1.0)
        x, it, rr, rc = oracle.bicgstab_ilu0(rp, ci, v, b, tol=1e-10)
        outs.append((rp, ci, v, x, it, rr))
    for a, c in zip(outs[0], outs[1]):
        assert np.array_equal(np.asarray(a), np.asarray(c))


def test_emulator_repeatable(oracle, emul):
    xy, cptr, cind = H.delaunay_case(3000, 24, seed=11)
    m, (rp, ci, v), b, g, bc = H.pressure_system(oracle, xy, cptr, cind, sigma=2.0)
    n = len(rp) - 1
    rng = np.random.default_rng(3)
    bs, xs = rng.standard_normal(n), rng.standard_normal(n)
    r1 = _run(emul, rp, ci, v, bs, xs, 32)
    r2 = _run(emul, rp, ci, v, bs, xs, 32)
    for k in r1:
        assert np.array_equal(r1[k], r2[k]), k


ASAN_SCRIPT = r"""
import ctypes, sys
import numpy as np
sys.path.insert(0, %(root)r)
from oracle import oracle as O
from tests import helpers as H
O.use_library(%(oracle)r)
xy, cptr, cind = H.delaunay_case(2000, 20, seed=13)
m, (rp, ci, v), b, g, bc = H.pressure_system(O, xy, cptr, cind, sigma=1.0)
x, it, rr, rc = O.bicgstab_ilu0(rp, ci, v, b, tol=1e-10)
assert rc == 0 and rr <= 1e-10, (rc, rr)
O.grad(m, np.arange(m.n_cells, dtype=float), np.zeros(m.n_faces))
lib = ctypes.CDLL(%(emul)r)
from tests.test_ordering_emul import _run
n = len(rp) - 1
i32 = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
f64 = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
lib.emul_run.restype = ctypes.c_int
lib.emul_run.argtypes = [ctypes.c_int, i32, i32, f64, ctypes.c_int, i32, i32, i32, f64, ctypes.c_int, f64,
                         f64, f64, f64, f64, i32, ctypes.c_char_p, ctypes.c_int, ctypes.c_void_p]
r = _run(lib, rp, ci, v, np.ones(n), np.ones(n), 32)
print("sanitized run ok", it)
"""


@pytest.mark.slow
def test_oracle_and_emulator_under_asan_ubsan(tmp_path):
    subprocess.run(["make", "-C", os.path.join(ROOT, "oracle"), "asan"], check=True, capture_output=True)
    build = os.path.join(ROOT, "oracle", "_build")
    oracle_so = os.path.join(build, "libphase_oracle_asan.so")
    emul_so = os.path.join(build, "libemul_asan.so")
    asan_rt = subprocess.run(["g++", "-print-file-name=libasan.so"], capture_output=True, text=True).stdout.strip()
    ubsan_rt = subprocess.run(["g++", "-print-file-name=libubsan.so"], capture_output=True,
                              text=True).stdout.strip()
    script = tmp_path / "asan_run.py"
    script.write_text(ASAN_SCRIPT % dict(root=ROOT, oracle=oracle_so, emul=emul_so))
    env = dict(os.environ, LD_PRELOAD=asan_rt + " " + ubsan_rt,
               ASAN_OPTIONS="detect_leaks=0:abort_on_error=1",
               UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1")
    r = subprocess.run([sys.executable, str(script)], env=env, capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-4000:]
    assert "sanitized run ok" in r.stdout
