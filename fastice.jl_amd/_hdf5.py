"""Minimal ctypes binding of the HDF5 C library (1.10 ABI: hid_t is int64).

Only what the writers need: create / open a file, write and read one
contiguous dataset of doubles or floats, query its dimensions.  The library
is located through FASTICE_HDF5_LIB, then the loader's search path, then
the conda prefix this image ships (/opt/conda/lib).
"""
import ctypes as C
import ctypes.util
import glob
import os

import numpy as np

_lib = None
hid_t = C.c_int64
herr_t = C.c_int
hsize_t = C.c_ulonglong

H5F_ACC_RDONLY = 0x0000
H5F_ACC_TRUNC = 0x0002
H5P_DEFAULT = 0
H5S_ALL = 0


def _candidates():
    env = os.environ.get("FASTICE_HDF5_LIB")
    if env:
        yield env
    name = ctypes.util.find_library("hdf5")
    if name:
        yield name
    for prefix in (os.environ.get("CONDA_PREFIX"), "/opt/conda"):
        if prefix:
            for p in sorted(glob.glob(os.path.join(prefix, "lib", "libhdf5.so*"))):
                yield p


def lib():
    """The loaded HDF5 library; raises RuntimeError when none is found."""
    global _lib
    if _lib is not None:
        return _lib
    errors = []
    for path in _candidates():
        try:
            h = C.CDLL(path)
        except OSError as e:
            errors.append("%s: %s" % (path, e))
            continue
        _declare(h)
        if h.H5open() < 0:
            raise RuntimeError("H5open failed in %s" % path)
        _lib = h
        return h
    raise RuntimeError("HDF5 C library not found (set FASTICE_HDF5_LIB); tried: %s" % "; ".join(errors))


def _declare(h):
    h.H5open.restype = herr_t
    h.H5Fcreate.argtypes = [C.c_char_p, C.c_uint, hid_t, hid_t]
    h.H5Fcreate.restype = hid_t
    h.H5Fopen.argtypes = [C.c_char_p, C.c_uint, hid_t]
    h.H5Fopen.restype = hid_t
    h.H5Fclose.argtypes = [hid_t]
    h.H5Fclose.restype = herr_t
    h.H5Screate_simple.argtypes = [C.c_int, C.POINTER(hsize_t), C.POINTER(hsize_t)]
    h.H5Screate_simple.restype = hid_t
    h.H5Sclose.argtypes = [hid_t]
    h.H5Sclose.restype = herr_t
    h.H5Sget_simple_extent_ndims.argtypes = [hid_t]
    h.H5Sget_simple_extent_ndims.restype = C.c_int
    h.H5Sget_simple_extent_dims.argtypes = [hid_t, C.POINTER(hsize_t), C.POINTER(hsize_t)]
    h.H5Sget_simple_extent_dims.restype = C.c_int
    h.H5Dcreate2.argtypes = [hid_t, C.c_char_p, hid_t, hid_t, hid_t, hid_t, hid_t]
    h.H5Dcreate2.restype = hid_t
    h.H5Dopen2.argtypes = [hid_t, C.c_char_p, hid_t]
    h.H5Dopen2.restype = hid_t
    h.H5Dwrite.argtypes = [hid_t, hid_t, hid_t, hid_t, hid_t, C.c_void_p]
    h.H5Dwrite.restype = herr_t
    h.H5Dread.argtypes = [hid_t, hid_t, hid_t, hid_t, hid_t, C.c_void_p]
    h.H5Dread.restype = herr_t
    h.H5Dget_space.argtypes = [hid_t]
    h.H5Dget_space.restype = hid_t
    h.H5Dget_type.argtypes = [hid_t]
    h.H5Dget_type.restype = hid_t
    h.H5Dclose.argtypes = [hid_t]
    h.H5Dclose.restype = herr_t
    h.H5Tget_size.argtypes = [hid_t]
    h.H5Tget_size.restype = C.c_size_t
    h.H5Tclose.argtypes = [hid_t]
    h.H5Tclose.restype = herr_t


def _native(dtype):
    h = lib()
    name = {np.dtype(np.float64): "H5T_NATIVE_DOUBLE_g", np.dtype(np.float32): "H5T_NATIVE_FLOAT_g"}.get(np.dtype(dtype))
    if name is None:
        raise TypeError("HDF5 datasets of %s are not supported (float64 or float32)" % dtype)
    return hid_t.in_dll(h, name).value


def _ok(rc, what):
    if rc < 0:
        raise RuntimeError("HDF5 %s failed" % what)
    return rc


class File:
    """An HDF5 file opened for writing (mode "w", truncates) or reading ("r")."""

    def __init__(self, path, mode="w"):
        h = lib()
        p = os.fsencode(path)
        if mode == "w":
            self.id = _ok(h.H5Fcreate(p, H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT), "create %s" % path)
        elif mode == "r":
            self.id = _ok(h.H5Fopen(p, H5F_ACC_RDONLY, H5P_DEFAULT), "open %s" % path)
        else:
            raise ValueError("mode must be 'w' or 'r'")

    def write(self, name, array):
        """Dataset `name` holding a C-contiguous array (its dims in C order)."""
        h = lib()
        a = np.ascontiguousarray(array)
        dims = (hsize_t * a.ndim)(*a.shape)
        space = _ok(h.H5Screate_simple(a.ndim, dims, None), "dataspace")
        try:
            t = _native(a.dtype)
            d = _ok(h.H5Dcreate2(self.id, name.encode(), t, space, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT),
                    "create dataset %s" % name)
            try:
                _ok(h.H5Dwrite(d, t, H5S_ALL, H5S_ALL, H5P_DEFAULT, a.ctypes.data), "write %s" % name)
            finally:
                h.H5Dclose(d)
        finally:
            h.H5Sclose(space)

    def read(self, name):
        """Dataset `name` as a C-order array (float64, or float32 if stored so)."""
        h = lib()
        d = _ok(h.H5Dopen2(self.id, name.encode(), H5P_DEFAULT), "open dataset %s" % name)
        try:
            space = _ok(h.H5Dget_space(d), "dataspace of %s" % name)
            try:
                nd = _ok(h.H5Sget_simple_extent_ndims(space), "rank")
                dims = (hsize_t * max(nd, 1))()
                _ok(h.H5Sget_simple_extent_dims(space, dims, None), "dims")
                shape = tuple(int(dims[q]) for q in range(nd))
            finally:
                h.H5Sclose(space)
            ft = h.H5Dget_type(d)
            size = h.H5Tget_size(ft)
            h.H5Tclose(ft)
            dtype = np.float32 if size == 4 else np.float64
            out = np.empty(shape, dtype=dtype)
            _ok(h.H5Dread(d, _native(dtype), H5S_ALL, H5S_ALL, H5P_DEFAULT, out.ctypes.data), "read %s" % name)
            return out
        finally:
            h.H5Dclose(d)

    def close(self):
        if getattr(self, "id", None) is not None:
            lib().H5Fclose(self.id)
            self.id = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()
