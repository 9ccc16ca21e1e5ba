import os
import subprocess
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")
    config.addinivalue_line("markers", "slow: larger CPU cases")


@pytest.fixture(scope="session")
def oracle():
    from oracle import oracle as O

    O.build()
    return O


@pytest.fixture(scope="session")
def emul():
    """CPU emulator of the device solver layout (tests/emul/solver_emul.cpp)."""
    import ctypes

    src = os.path.join(ROOT, "tests", "emul")
    out = os.path.join(src, "_build")
    os.makedirs(out, exist_ok=True)
    so = os.path.join(out, "libemul.so")
    deps = [os.path.join(src, "solver_emul.cpp"), os.path.join(ROOT, "phase_amd", "csrc", "ordering.cpp"),
            os.path.join(ROOT, "phase_amd", "csrc", "phase_internal.h")]
    if not os.path.exists(so) or any(os.path.getmtime(d) > os.path.getmtime(so) for d in deps):
        subprocess.run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-ffp-contract=off",
                        "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include", "-o", so, deps[0], deps[1]],
                       check=True)
    lib = ctypes.CDLL(so)
    i32 = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
    f64 = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
    lib.emul_run.restype = ctypes.c_int
    lib.emul_run.argtypes = [ctypes.c_int, i32, i32, f64, ctypes.c_int, i32, i32, i32, f64, ctypes.c_int, f64,
                             f64, f64, f64, f64, i32, ctypes.c_char_p, ctypes.c_int, ctypes.c_void_p]
    return lib


def gpu_available():
    try:
        from phase_amd import hip

        return hip.device_count() > 0
    except Exception:
        return False


@pytest.fixture(scope="session")
def hip():
    """The product library on a GPU; fails loudly when the build is missing."""
    from phase_amd import hip as H

    H.load_library()
    if H.device_count() == 0:
        pytest.fail("a gpu test ran without a visible HIP device")
    return H
