"""The C-ABI library loads and exports every entry point include/dg3d.h declares (no device calls)."""
import ctypes
import os
import re

from helpers import dg3d

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "dg3d.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(dg3d_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(dg3d.LIB_PATH)
    syms = declared_symbols()
    assert len(syms) >= 25
    missing = [s for s in syms if not hasattr(lib, s)]
    assert not missing, missing


def test_python_binding_covers_header():
    assert set(declared_symbols()) == set(dg3d._SIGS)


def test_abi_version_and_create_error_without_device():
    L = dg3d.lib()
    assert L.dg3d_abi_version() == 1
    # invalid descriptors are rejected on the host before any device call
    d = dg3d.Desc()
    d.order = 0
    h = ctypes.c_void_p()
    assert L.dg3d_create(ctypes.byref(d), ctypes.byref(h)) == dg3d.E_UNSUPPORTED
    assert b"order" in L.dg3d_create_error()
    assert not h.value
