"""Minimal HDF5 1-D dataset I/O through ctypes (h5py is not part of this image).

Used to write the ``Temp.h5`` forcing input and to read the drop-in's QoI and
field files back.  Same C API as cupf_amd/host/h5io.cpp (HDF5 1.10, 64-bit hid_t).
"""
from __future__ import annotations

import ctypes

import numpy as np

_NAMES = ["libhdf5.so.103", "/opt/conda/lib/libhdf5.so.103", "libhdf5.so", "/opt/conda/lib/libhdf5.so"]
_lib = None
hid_t = ctypes.c_int64


def lib():
    global _lib
    if _lib is None:
        for n in _NAMES:
            try:
                _lib = ctypes.CDLL(n)
                break
            except OSError:
                continue
        if _lib is None:
            raise OSError("libhdf5 not found")
        _lib.H5open()
        _lib.H5Fcreate.restype = hid_t
        _lib.H5Fopen.restype = hid_t
        _lib.H5Screate_simple.restype = hid_t
        _lib.H5Dcreate2.restype = hid_t
        _lib.H5Dopen2.restype = hid_t
        _lib.H5Dget_space.restype = hid_t
        for f in ("H5Fcreate", "H5Fopen", "H5Dcreate2", "H5Dopen2", "H5Dwrite", "H5Dread"):
            getattr(_lib, f).argtypes = None
    return _lib


def available() -> bool:
    try:
        lib()
        return True
    except OSError:
        return False


def _type(dtype):
    L = lib()
    name = "H5T_NATIVE_INT_g" if np.dtype(dtype) == np.int32 else "H5T_NATIVE_FLOAT_g"
    return hid_t.in_dll(L, name).value


def write(path: str, datasets: dict) -> None:
    L = lib()
    f = L.H5Fcreate(path.encode(), ctypes.c_uint(2), hid_t(0), hid_t(0))
    if f < 0:
        raise OSError(f"H5Fcreate {path}")
    try:
        for name, arr in datasets.items():
            a = np.ascontiguousarray(arr)
            if a.dtype not in (np.int32, np.float32):
                a = a.astype(np.float32)
            dim = (ctypes.c_ulonglong * 1)(a.size)
            sp = L.H5Screate_simple(ctypes.c_int(1), dim, None)
            t = _type(a.dtype)
            ds = L.H5Dcreate2(hid_t(f), name.encode(), hid_t(t), hid_t(sp), hid_t(0), hid_t(0), hid_t(0))
            L.H5Dwrite(hid_t(ds), hid_t(t), hid_t(0), hid_t(0), hid_t(0), a.ctypes.data_as(ctypes.c_void_p))
            L.H5Dclose(hid_t(ds))
            L.H5Sclose(hid_t(sp))
    finally:
        L.H5Fclose(hid_t(f))


def read(path: str, name: str, dtype=np.float32) -> np.ndarray:
    L = lib()
    f = L.H5Fopen(path.encode(), ctypes.c_uint(0), hid_t(0))
    if f < 0:
        raise OSError(f"H5Fopen {path}")
    try:
        ds = L.H5Dopen2(hid_t(f), name.encode(), hid_t(0))
        if ds < 0:
            raise KeyError(name)
        sp = L.H5Dget_space(hid_t(ds))
        dims = (ctypes.c_ulonglong * 8)()
        nd = L.H5Sget_simple_extent_dims(hid_t(sp), dims, None)
        n = int(np.prod([dims[q] for q in range(nd)])) if nd > 0 else 1
        out = np.empty(n, dtype=dtype)
        if n:
            L.H5Dread(hid_t(ds), hid_t(_type(dtype)), hid_t(0), hid_t(0), hid_t(0),
                      out.ctypes.data_as(ctypes.c_void_p))
        L.H5Sclose(hid_t(sp))
        L.H5Dclose(hid_t(ds))
        return out
    finally:
        L.H5Fclose(hid_t(f))
