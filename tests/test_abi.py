"""The C-ABI library loads and exports every entry point include/b2r_hip.h
declares (no device calls: runs without a GPU)."""
import ctypes
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    src = open(os.path.join(ROOT, "include", "b2r_hip.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(b2r_[a-z0-9_]+)\s*\(", src)))


def test_header_declares_entry_points():
    names = declared()
    assert "b2r_boundary_updating" in names and "b2r_heat_run" in names and len(names) > 20


def test_library_exports_every_declared_symbol():
    from b2r_amd import _abi
    lib = ctypes.CDLL(_abi.LIB_PATH)
    missing = [n for n in declared() if not hasattr(lib, n)]
    assert not missing, missing


def test_binding_covers_header():
    from b2r_amd import _abi
    assert set(declared()) <= set(_abi.EXPORTED)


def test_error_string_without_device():
    from b2r_amd import _abi
    # geometry validation runs before any device call
    g = _abi.Geom(10, 10, 10, 3, 1, 1, 1, 1, 0, 4)
    h = ctypes.c_void_p()
    rc = _abi.lib().b2r_create(ctypes.byref(g), 0, ctypes.byref(h))
    assert rc == _abi.B2R_E_GEOM and "divisible" in _abi.last_error()
    g = _abi.Geom(8, 8, 8, 1, 1, 1, 3, 2, 0, 4)  # R*D = 6 > min(B)/2 = 4
    rc = _abi.lib().b2r_create(ctypes.byref(g), 0, ctypes.byref(h))
    assert rc == _abi.B2R_E_GEOM and "min(B)/2" in _abi.last_error()
