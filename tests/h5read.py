"""Minimal HDF5 reader over the C library (ctypes), for checking the driver's output files."""
import ctypes
import os

import numpy as np

_CANDIDATES = [os.environ.get("DG3D_HDF5_LIB"), "libhdf5.so", "libhdf5.so.103", "/opt/conda/lib/libhdf5.so"]
_lib = None


def lib():
    global _lib
    if _lib is None:
        for c in _CANDIDATES:
            if not c:
                continue
            try:
                _lib = ctypes.CDLL(c)
                break
            except OSError:
                continue
        if _lib is None:
            return None
        _lib.H5open()
        _lib.H5Fopen.restype = ctypes.c_int64
        _lib.H5Fopen.argtypes = [ctypes.c_char_p, ctypes.c_uint, ctypes.c_int64]
        _lib.H5Dopen2.restype = ctypes.c_int64
        _lib.H5Dopen2.argtypes = [ctypes.c_int64, ctypes.c_char_p, ctypes.c_int64]
        _lib.H5Dget_space.restype = ctypes.c_int64
        _lib.H5Dget_space.argtypes = [ctypes.c_int64]
        _lib.H5Sget_simple_extent_ndims.argtypes = [ctypes.c_int64]
        _lib.H5Sget_simple_extent_dims.argtypes = [ctypes.c_int64, ctypes.POINTER(ctypes.c_ulonglong), ctypes.c_void_p]
        _lib.H5Dread.argtypes = [ctypes.c_int64] * 5 + [ctypes.c_void_p]
        for f in ("H5Dclose", "H5Sclose", "H5Fclose"):
            getattr(_lib, f).argtypes = [ctypes.c_int64]
    return _lib


def available():
    return lib() is not None


def read(path, name):
    L = lib()
    f = L.H5Fopen(str(path).encode(), 0, 0)
    assert f >= 0, path
    d = L.H5Dopen2(f, name.encode(), 0)
    assert d >= 0, name
    s = L.H5Dget_space(d)
    nd = L.H5Sget_simple_extent_ndims(s)
    dims = (ctypes.c_ulonglong * nd)()
    L.H5Sget_simple_extent_dims(s, dims, None)
    arr = np.zeros(tuple(dims), dtype=np.float64)
    dbl = ctypes.c_int64.in_dll(L, "H5T_NATIVE_DOUBLE_g").value
    rc = L.H5Dread(d, dbl, 0, 0, 0, arr.ctypes.data_as(ctypes.c_void_p))
    assert rc >= 0
    L.H5Sclose(s)
    L.H5Dclose(d)
    L.H5Fclose(f)
    return arr
