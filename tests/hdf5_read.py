"""Minimal reader over the HDF5 C library (ctypes), to read back what the
output writers wrote: datasets and attributes with their stored type and
shape.  Test infrastructure only."""
import ctypes
import os

import numpy as np

HDF5_DIR = os.environ.get("HDF5_DIR", "/opt/conda")
_h5 = None
hid_t = ctypes.c_int64
H5T_INTEGER, H5T_FLOAT, H5T_STRING = 0, 1, 3
H5S_SCALAR = 0


def h5():
    global _h5
    if _h5 is None:
        L = ctypes.CDLL(os.path.join(HDF5_DIR, "lib", "libhdf5.so"))
        for f in ("H5Fopen", "H5Dopen2", "H5Dget_space", "H5Dget_type", "H5Aopen_by_name", "H5Aget_space",
                  "H5Aget_type", "H5Tget_native_type", "H5Gopen2"):
            getattr(L, f).restype = hid_t
        L.H5Fopen.argtypes = [ctypes.c_char_p, ctypes.c_uint, hid_t]
        L.H5Dopen2.argtypes = [hid_t, ctypes.c_char_p, hid_t]
        L.H5Gopen2.argtypes = [hid_t, ctypes.c_char_p, hid_t]
        L.H5Aopen_by_name.argtypes = [hid_t, ctypes.c_char_p, ctypes.c_char_p, hid_t, hid_t]
        for f in ("H5Dget_space", "H5Dget_type", "H5Aget_space", "H5Aget_type", "H5Tget_class", "H5Tget_size",
                  "H5Tget_sign", "H5Sget_simple_extent_ndims", "H5Sget_simple_extent_type", "H5Dclose", "H5Sclose",
                  "H5Tclose", "H5Aclose", "H5Fclose", "H5Gclose"):
            getattr(L, f).argtypes = [hid_t]
        L.H5Tget_size.restype = ctypes.c_size_t
        L.H5Tget_native_type.argtypes = [hid_t, ctypes.c_int]
        L.H5Sget_simple_extent_dims.argtypes = [hid_t, ctypes.POINTER(ctypes.c_ulonglong), ctypes.c_void_p]
        L.H5Dread.argtypes = [hid_t, hid_t, hid_t, hid_t, hid_t, ctypes.c_void_p]
        L.H5Aread.argtypes = [hid_t, hid_t, ctypes.c_void_p]
        L.H5Lexists.argtypes = [hid_t, ctypes.c_char_p, hid_t]
        L.H5open()
        _h5 = L
    return _h5


def _dtype(L, t):
    cls, size = L.H5Tget_class(t), L.H5Tget_size(t)
    if cls == H5T_FLOAT:
        return np.dtype(np.float64 if size == 8 else np.float32)
    if cls == H5T_INTEGER:
        signed = L.H5Tget_sign(t) == 1
        return np.dtype(("i" if signed else "u") + str(size))
    raise TypeError(f"unsupported HDF5 type class {cls}")


def _shape(L, sp):
    nd = L.H5Sget_simple_extent_ndims(sp)
    dims = (ctypes.c_ulonglong * max(nd, 1))()
    L.H5Sget_simple_extent_dims(sp, dims, None)
    return tuple(int(dims[i]) for i in range(nd))


class File:
    def __init__(self, fname):
        self.L = h5()
        self.f = self.L.H5Fopen(os.fsencode(fname), 0, 0)
        if self.f < 0:
            raise OSError(f"cannot open {fname}")

    def close(self):
        if self.f >= 0:
            self.L.H5Fclose(self.f)
            self.f = -1

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def exists(self, path):
        return self.L.H5Lexists(self.f, path.encode(), 0) > 0

    def dataset(self, path):
        """(array, stored dtype)"""
        L = self.L
        d = L.H5Dopen2(self.f, path.encode(), 0)
        if d < 0:
            raise KeyError(path)
        t = L.H5Dget_type(d)
        sp = L.H5Dget_space(d)
        dt = _dtype(L, t)
        out = np.empty(_shape(L, sp), dtype=dt)
        mt = L.H5Tget_native_type(t, 0)
        st = L.H5Dread(d, mt, 0, 0, 0, out.ctypes.data_as(ctypes.c_void_p))
        for h in (mt, t):
            L.H5Tclose(h)
        L.H5Sclose(sp)
        L.H5Dclose(d)
        if st < 0:
            raise OSError(f"cannot read {path}")
        return out, dt

    def attr(self, obj, name):
        """value (scalar, array or str) and (dtype or 'str', shape or None for a scalar)"""
        L = self.L
        a = L.H5Aopen_by_name(self.f, obj.encode(), name.encode(), 0, 0)
        if a < 0:
            raise KeyError(f"{obj}:{name}")
        t = L.H5Aget_type(a)
        sp = L.H5Aget_space(a)
        scalar = L.H5Sget_simple_extent_type(sp) == H5S_SCALAR
        if L.H5Tget_class(t) == H5T_STRING:
            buf = ctypes.create_string_buffer(L.H5Tget_size(t) + 1)
            L.H5Aread(a, t, buf)
            val, kind = buf.value.decode(), ("str", None)
        else:
            dt = _dtype(L, t)
            shape = () if scalar else _shape(L, sp)
            out = np.empty(shape, dtype=dt)
            mt = L.H5Tget_native_type(t, 0)
            L.H5Aread(a, mt, out.ctypes.data_as(ctypes.c_void_p))
            L.H5Tclose(mt)
            val, kind = (out[()] if scalar else out), (dt, None if scalar else shape)
        L.H5Tclose(t)
        L.H5Sclose(sp)
        L.H5Aclose(a)
        return val, kind
