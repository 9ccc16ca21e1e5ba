"""The C ABI library loads without a GPU and exports every entry point that
include/*.h declares; the Python binding declares nothing the header lacks."""
import ctypes
import glob
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names |= set(re.findall(r"\b(phase_hip_\w+)\s*\(", text))
    return names


def test_header_symbols_exported():
    from phase_amd import hip

    lib = hip.load_library()
    names = declared()
    assert len(names) >= 25
    missing = [n for n in sorted(names) if not hasattr(lib, n)]
    assert not missing, missing


def test_binding_matches_header():
    from phase_amd import hip

    assert set(hip._SIGS) <= declared()


def test_no_device_reports_cleanly():
    from phase_amd import hip

    lib = hip.load_library()
    c = ctypes.c_int(-1)
    assert lib.phase_hip_device_count(ctypes.byref(c)) == 0
    if c.value == 0:
        h = ctypes.c_void_p()
        rc = lib.phase_hip_solver_create(ctypes.byref(h), 0)
        assert rc == -2 and not h.value
        assert b"no HIP device" in lib.phase_hip_last_error(None)


def test_missing_library_fails_loudly(tmp_path, monkeypatch):
    import importlib

    from phase_amd import hip

    monkeypatch.setattr(hip, "_LIB", str(tmp_path / "absent.so"))
    monkeypatch.setattr(hip, "_lib", None)
    try:
        hip.load_library()
    except ImportError as e:
        assert "not built" in str(e)
    else:
        raise AssertionError("load_library must fail without the built library")
    importlib.reload(hip)
