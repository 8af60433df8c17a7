"""Minimal HDF5 access through ctypes on the HDF5 C library (test
infrastructure: the image has the library but no h5py).  Writes the
reference's flat pre-inverted table layout and reads datasets, groups and
attributes back."""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

PREFIX = os.environ.get("CCAKE_HDF5_PREFIX", "/opt/conda")
LIB = os.path.join(PREFIX, "lib", "libhdf5.so")
hid_t = C.c_int64
H5P_DEFAULT = 0
H5S_ALL = 0
H5F_ACC_RDONLY = 0
H5F_ACC_TRUNC = 2

_L = None


def available():
    return os.path.exists(LIB)


def lib():
    global _L
    if _L is None:
        L = C.CDLL(LIB)
        L.H5open()
        for name, res, args in [
            ("H5Fcreate", hid_t, [C.c_char_p, C.c_uint, hid_t, hid_t]),
            ("H5Fopen", hid_t, [C.c_char_p, C.c_uint, hid_t]),
            ("H5Fclose", C.c_int, [hid_t]),
            ("H5Gcreate2", hid_t, [hid_t, C.c_char_p, hid_t, hid_t, hid_t]),
            ("H5Gopen2", hid_t, [hid_t, C.c_char_p, hid_t]),
            ("H5Gclose", C.c_int, [hid_t]),
            ("H5Gget_num_objs", C.c_int, [hid_t, C.POINTER(C.c_uint64)]),
            ("H5Gget_objname_by_idx", C.c_ssize_t, [hid_t, C.c_uint64, C.c_char_p, C.c_size_t]),
            ("H5Screate_simple", hid_t, [C.c_int, C.POINTER(C.c_uint64), C.c_void_p]),
            ("H5Sclose", C.c_int, [hid_t]),
            ("H5Sget_simple_extent_npoints", C.c_int64, [hid_t]),
            ("H5Dcreate2", hid_t, [hid_t, C.c_char_p, hid_t, hid_t, hid_t, hid_t, hid_t]),
            ("H5Dopen2", hid_t, [hid_t, C.c_char_p, hid_t]),
            ("H5Dwrite", C.c_int, [hid_t, hid_t, hid_t, hid_t, hid_t, C.c_void_p]),
            ("H5Dread", C.c_int, [hid_t, hid_t, hid_t, hid_t, hid_t, C.c_void_p]),
            ("H5Dget_space", hid_t, [hid_t]),
            ("H5Dclose", C.c_int, [hid_t]),
            ("H5Aopen", hid_t, [hid_t, C.c_char_p, hid_t]),
            ("H5Aread", C.c_int, [hid_t, hid_t, C.c_void_p]),
            ("H5Aget_type", hid_t, [hid_t]),
            ("H5Aclose", C.c_int, [hid_t]),
            ("H5Tclose", C.c_int, [hid_t]),
            ("H5Dvlen_reclaim", C.c_int, [hid_t, hid_t, hid_t, C.c_void_p]),
        ]:
            f = getattr(L, name)
            f.restype, f.argtypes = res, args
        _L = L
    return _L


def native(name):
    return hid_t.in_dll(lib(), name).value


def write_flat_table(path, Ns, NB, arrays, dims_shape=2):
    """A flat (s, nB) table file: /dimensions = [Ns, NB] (int) and
    /data/<name> (Ns x NB doubles) for the 12 names in order."""
    L = lib()
    names = ["s", "nB", "T", "muB", "muS", "muQ", "e", "p", "cs2", "chiTT", "chiTB", "chiBB"]
    f = L.H5Fcreate(path.encode(), H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT)
    assert f >= 0
    d1 = (C.c_uint64 * 1)(dims_shape)
    sp = L.H5Screate_simple(1, d1, None)
    ds = L.H5Dcreate2(f, b"dimensions", native("H5T_NATIVE_INT_g"), sp, 0, 0, 0)
    buf = np.array([Ns, NB, 0][:dims_shape], dtype=np.int32)
    assert L.H5Dwrite(ds, native("H5T_NATIVE_INT_g"), H5S_ALL, H5S_ALL, H5P_DEFAULT, buf.ctypes.data) >= 0
    L.H5Dclose(ds); L.H5Sclose(sp)
    g = L.H5Gcreate2(f, b"data", 0, 0, 0)
    d2 = (C.c_uint64 * 2)(Ns, NB)
    sp = L.H5Screate_simple(2, d2, None)
    for name, a in zip(names, arrays):
        a = np.ascontiguousarray(a, dtype=np.float64)
        ds = L.H5Dcreate2(g, name.encode(), native("H5T_NATIVE_DOUBLE_g"), sp, 0, 0, 0)
        assert L.H5Dwrite(ds, native("H5T_NATIVE_DOUBLE_g"), H5S_ALL, H5S_ALL, H5P_DEFAULT, a.ctypes.data) >= 0
        L.H5Dclose(ds)
    L.H5Sclose(sp); L.H5Gclose(g); L.H5Fclose(f)


class File:
    def __init__(self, path):
        self.f = lib().H5Fopen(path.encode(), H5F_ACC_RDONLY, H5P_DEFAULT)
        assert self.f >= 0, path

    def close(self):
        lib().H5Fclose(self.f)

    def members(self, group):
        L = lib()
        g = L.H5Gopen2(self.f, group.encode(), 0)
        n = C.c_uint64()
        L.H5Gget_num_objs(g, C.byref(n))
        out = []
        for k in range(n.value):
            buf = C.create_string_buffer(256)
            L.H5Gget_objname_by_idx(g, k, buf, 256)
            out.append(buf.value.decode())
        L.H5Gclose(g)
        return out

    def dataset(self, path, dtype=np.float64):
        L = lib()
        ds = L.H5Dopen2(self.f, path.encode(), 0)
        assert ds >= 0, path
        sp = L.H5Dget_space(ds)
        n = L.H5Sget_simple_extent_npoints(sp)
        L.H5Sclose(sp)
        out = np.zeros(n, dtype=dtype)
        t = native("H5T_NATIVE_DOUBLE_g" if dtype == np.float64 else "H5T_NATIVE_INT_g")
        assert L.H5Dread(ds, t, H5S_ALL, H5S_ALL, H5P_DEFAULT, out.ctypes.data) >= 0
        L.H5Dclose(ds)
        return out

    def double_attribute(self, obj_path, name):
        L = lib()
        g = L.H5Gopen2(self.f, obj_path.encode(), 0)
        a = L.H5Aopen(g, name.encode(), 0)
        v = C.c_double()
        assert L.H5Aread(a, native("H5T_NATIVE_DOUBLE_g"), C.byref(v)) >= 0
        L.H5Aclose(a); L.H5Gclose(g)
        return v.value

    def units(self, ds_path):
        """The variable-length string attribute "Units" of a dataset."""
        L = lib()
        ds = L.H5Dopen2(self.f, ds_path.encode(), 0)
        a = L.H5Aopen(ds, b"Units", 0)
        t = L.H5Aget_type(a)
        p = C.c_char_p()
        assert L.H5Aread(a, t, C.byref(p)) >= 0
        s = p.value.decode()
        L.H5Tclose(t); L.H5Aclose(a); L.H5Dclose(ds)
        return s
