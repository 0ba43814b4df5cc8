"""A small reader of HDF5 files through the HDF5 C library itself (ctypes), used
by the tests to check the files libdgpu_io.so writes independently of it."""
from __future__ import annotations

import ctypes
import os

import numpy as np

LIB = os.environ.get("DGPU_HDF5_LIB", "/opt/conda/lib/libhdf5.so.103")
_h = None
hid_t = ctypes.c_int64
hsize_t = ctypes.c_uint64


def available() -> bool:
    return os.path.exists(LIB)


def lib():
    global _h
    if _h is None:
        L = ctypes.CDLL(LIB)
        L.H5open()
        L.H5Eset_auto2(hid_t(0), None, None)
        for fn in ("H5Fopen", "H5Gopen2", "H5Dopen2", "H5Aopen_by_name", "H5Aget_space", "H5Aget_type",
                   "H5Dget_space", "H5Dget_type", "H5Dget_create_plist", "H5Screate"):
            getattr(L, fn).restype = hid_t
        L.H5Fopen.argtypes = [ctypes.c_char_p, ctypes.c_uint, hid_t]
        L.H5Gopen2.argtypes = [hid_t, ctypes.c_char_p, hid_t]
        L.H5Dopen2.argtypes = [hid_t, ctypes.c_char_p, hid_t]
        L.H5Aopen_by_name.argtypes = [hid_t, ctypes.c_char_p, ctypes.c_char_p, hid_t, hid_t]
        L.H5Aread.argtypes = [hid_t, hid_t, ctypes.c_void_p]
        L.H5Dread.argtypes = [hid_t, hid_t, hid_t, hid_t, hid_t, ctypes.c_void_p]
        for fn in ("H5Aget_space", "H5Aget_type", "H5Dget_space", "H5Dget_type", "H5Dget_create_plist",
                   "H5Fclose", "H5Gclose", "H5Dclose", "H5Aclose", "H5Sclose", "H5Tclose", "H5Pclose",
                   "H5Sget_simple_extent_ndims", "H5Tget_size", "H5Tget_class"):
            getattr(L, fn).argtypes = [hid_t]
        L.H5Sget_simple_extent_dims.argtypes = [hid_t, ctypes.POINTER(hsize_t), ctypes.POINTER(hsize_t)]
        L.H5Pget_chunk.argtypes = [hid_t, ctypes.c_int, ctypes.POINTER(hsize_t)]
        L.H5Lexists.argtypes = [hid_t, ctypes.c_char_p, hid_t]
        L.H5Tequal.argtypes = [hid_t, hid_t]
        _h = L
    return _h


def _glob(name: str) -> int:
    return hid_t.in_dll(lib(), name).value


def native_double() -> int:
    return _glob("H5T_NATIVE_DOUBLE_g")


def native_int() -> int:
    return _glob("H5T_NATIVE_INT_g")


def std_i32le() -> int:
    return _glob("H5T_STD_I32LE_g")


def ieee_f64le() -> int:
    return _glob("H5T_IEEE_F64LE_g")


class File:
    def __init__(self, path: str):
        self.f = lib().H5Fopen(path.encode(), 0, 0)  # H5F_ACC_RDONLY, H5P_DEFAULT
        if self.f < 0:
            raise OSError(f"cannot open {path}")

    def close(self):
        lib().H5Fclose(self.f)

    def exists(self, path: str) -> bool:
        return lib().H5Lexists(self.f, path.encode(), 0) > 0

    def attr(self, group: str, name: str, kind: str = "f64"):
        """(values, stored type id equal to the expected file type, dims)."""
        L = lib()
        obj = b"." if group in ("", "/") else group.encode()
        a = L.H5Aopen_by_name(self.f, obj, name.encode(), 0, 0)
        if a < 0:
            raise KeyError(f"{group}@{name}")
        sp = L.H5Aget_space(a)
        nd = L.H5Sget_simple_extent_ndims(sp)
        dims = (hsize_t * max(nd, 1))()
        if nd > 0:
            L.H5Sget_simple_extent_dims(sp, dims, None)
        n = int(np.prod([dims[i] for i in range(nd)])) if nd > 0 else 1
        ty = L.H5Aget_type(a)
        if kind == "str":
            buf = ctypes.create_string_buffer(n * L.H5Tget_size(ty) + 1)
            L.H5Aread(a, ty, buf)
            out = buf.raw[:n * L.H5Tget_size(ty)]
            same = True
        else:
            want_file = ieee_f64le() if kind == "f64" else std_i32le()
            same = L.H5Tequal(ty, want_file) > 0
            arr = np.zeros(n, np.float64 if kind == "f64" else np.int32)
            L.H5Aread(a, native_double() if kind == "f64" else native_int(), arr.ctypes.data_as(ctypes.c_void_p))
            out = arr
        L.H5Tclose(ty)
        L.H5Sclose(sp)
        L.H5Aclose(a)
        return out, same, tuple(dims[i] for i in range(nd))

    def dataset(self, path: str):
        """(data as float64 C-order array, chunk dims or None, stored as IEEE f64 LE)."""
        L = lib()
        d = L.H5Dopen2(self.f, path.encode(), 0)
        if d < 0:
            raise KeyError(path)
        sp = L.H5Dget_space(d)
        nd = L.H5Sget_simple_extent_ndims(sp)
        dims = (hsize_t * nd)()
        L.H5Sget_simple_extent_dims(sp, dims, None)
        shape = tuple(int(dims[i]) for i in range(nd))
        out = np.zeros(shape, np.float64)
        L.H5Dread(d, native_double(), 0, 0, 0, out.ctypes.data_as(ctypes.c_void_p))
        pl = L.H5Dget_create_plist(d)
        ch = (hsize_t * nd)()
        r = L.H5Pget_chunk(pl, nd, ch)
        chunk = tuple(int(ch[i]) for i in range(nd)) if r > 0 else None
        ty = L.H5Dget_type(d)
        f64 = L.H5Tequal(ty, ieee_f64le()) > 0
        for close, obj in ((L.H5Tclose, ty), (L.H5Pclose, pl), (L.H5Sclose, sp), (L.H5Dclose, d)):
            close(obj)
        return out, chunk, f64
