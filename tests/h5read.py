"""Minimal HDF5 reader over the image's libhdf5 (ctypes), for checking the driver's output layout.
h5py is not installed in this image.  Supports: listing a group, reading f64/int datasets, the
compound {r, i} complex datasets, and f64 attributes."""
from __future__ import annotations

import ctypes
import os

import numpy as np

HDF5_DIR = os.environ.get("HDF5_DIR", "/opt/conda")
_lib = None
hid_t = ctypes.c_int64
hsize_t = ctypes.c_uint64


def lib():
    global _lib
    if _lib is None:
        L = ctypes.CDLL(os.path.join(HDF5_DIR, "lib", "libhdf5.so"))
        L.H5open()
        L.H5Fopen.restype = hid_t
        L.H5Fopen.argtypes = [ctypes.c_char_p, ctypes.c_uint, hid_t]
        L.H5Dopen2.restype = hid_t
        L.H5Dopen2.argtypes = [hid_t, ctypes.c_char_p, hid_t]
        L.H5Gopen2.restype = hid_t
        L.H5Gopen2.argtypes = [hid_t, ctypes.c_char_p, hid_t]
        L.H5Dget_space.restype = hid_t
        L.H5Dget_space.argtypes = [hid_t]
        L.H5Sget_simple_extent_ndims.argtypes = [hid_t]
        L.H5Sget_simple_extent_dims.argtypes = [hid_t, ctypes.POINTER(hsize_t), ctypes.c_void_p]
        L.H5Dread.argtypes = [hid_t, hid_t, hid_t, hid_t, hid_t, ctypes.c_void_p]
        L.H5Dget_type.restype = hid_t
        L.H5Dget_type.argtypes = [hid_t]
        L.H5Tget_class.argtypes = [hid_t]
        L.H5Tcreate.restype = hid_t
        L.H5Tcreate.argtypes = [ctypes.c_int, ctypes.c_size_t]
        L.H5Tinsert.argtypes = [hid_t, ctypes.c_char_p, ctypes.c_size_t, hid_t]
        L.H5Tget_nmembers.argtypes = [hid_t]
        L.H5Tget_member_name.restype = ctypes.c_void_p
        L.H5Tget_member_name.argtypes = [hid_t, ctypes.c_uint]
        L.H5Aopen.restype = hid_t
        L.H5Aopen.argtypes = [hid_t, ctypes.c_char_p, hid_t]
        L.H5Aread.argtypes = [hid_t, hid_t, ctypes.c_void_p]
        L.H5Lexists.argtypes = [hid_t, ctypes.c_char_p, hid_t]
        L.H5Gget_info.argtypes = [hid_t, ctypes.c_void_p]
        L.H5Lget_name_by_idx.restype = ctypes.c_ssize_t
        L.H5Lget_name_by_idx.argtypes = [hid_t, ctypes.c_char_p, ctypes.c_int, ctypes.c_int, hsize_t, ctypes.c_char_p,
                                         ctypes.c_size_t, hid_t]
        for f in ("H5Fclose", "H5Dclose", "H5Sclose", "H5Tclose", "H5Aclose", "H5Gclose"):
            getattr(L, f).argtypes = [hid_t]
        _lib = L
    return _lib


def _native(name):
    return hid_t.in_dll(lib(), name).value


class File:
    def __init__(self, path):
        self.f = lib().H5Fopen(path.encode(), 0, 0)  # H5F_ACC_RDONLY, H5P_DEFAULT
        if self.f < 0:
            raise OSError(f"cannot open {path}")

    def close(self):
        lib().H5Fclose(self.f)

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def exists(self, name):
        return lib().H5Lexists(self.f, name.encode(), 0) > 0

    def list(self, group="/"):
        L = lib()
        g = L.H5Gopen2(self.f, group.encode(), 0)

        class Info(ctypes.Structure):
            _fields_ = [("storage_type", ctypes.c_int), ("nlinks", hsize_t), ("max_corder", ctypes.c_int64),
                        ("mounted", ctypes.c_bool)]

        info = Info()
        L.H5Gget_info(g, ctypes.byref(info))
        names = []
        buf = ctypes.create_string_buffer(256)
        for i in range(info.nlinks):
            L.H5Lget_name_by_idx(g, b".", 0, 0, i, buf, 256, 0)  # H5_INDEX_NAME, H5_ITER_INC
            names.append(buf.value.decode())
        L.H5Gclose(g)
        return names

    def shape(self, name):
        L = lib()
        d = L.H5Dopen2(self.f, name.encode(), 0)
        sp = L.H5Dget_space(d)
        nd = L.H5Sget_simple_extent_ndims(sp)
        dims = (hsize_t * nd)()
        L.H5Sget_simple_extent_dims(sp, dims, None)
        L.H5Sclose(sp)
        L.H5Dclose(d)
        return tuple(int(x) for x in dims)

    def compound_members(self, name):
        L = lib()
        d = L.H5Dopen2(self.f, name.encode(), 0)
        t = L.H5Dget_type(d)
        out = [ctypes.string_at(L.H5Tget_member_name(t, i)).decode() for i in range(L.H5Tget_nmembers(t))]
        L.H5Tclose(t)
        L.H5Dclose(d)
        return out

    def read(self, name, kind="f64"):
        L = lib()
        shp = self.shape(name)
        d = L.H5Dopen2(self.f, name.encode(), 0)
        if kind == "complex":
            out = np.empty(shp, dtype=np.complex128)
            t = L.H5Tcreate(6, 16)  # H5T_COMPOUND
            L.H5Tinsert(t, b"r", 0, _native("H5T_NATIVE_DOUBLE_g"))
            L.H5Tinsert(t, b"i", 8, _native("H5T_NATIVE_DOUBLE_g"))
            L.H5Dread(d, t, 0, 0, 0, out.ctypes.data)
            L.H5Tclose(t)
        elif kind == "int":
            out = np.empty(shp, dtype=np.int32)
            L.H5Dread(d, _native("H5T_NATIVE_INT_g"), 0, 0, 0, out.ctypes.data)
        else:
            out = np.empty(shp, dtype=np.float64)
            L.H5Dread(d, _native("H5T_NATIVE_DOUBLE_g"), 0, 0, 0, out.ctypes.data)
        L.H5Dclose(d)
        return out

    def attr(self, group, name):
        L = lib()
        g = L.H5Gopen2(self.f, group.encode(), 0)
        a = L.H5Aopen(g, name.encode(), 0)
        v = ctypes.c_double()
        L.H5Aread(a, _native("H5T_NATIVE_DOUBLE_g"), ctypes.byref(v))
        L.H5Aclose(a)
        L.H5Gclose(g)
        return v.value
