"""Reads and writes HDF5 files with the HDF5 C library (libhdf5 1.10) through
ctypes, to check kgeo_amd.h5lite against the library itself.  Test helper
only; find() returns None where no libhdf5 is installed."""

import ctypes
import ctypes.util
import os

import numpy as np

CANDIDATES = ["/opt/conda/lib/libhdf5.so.103", "/opt/conda/lib/libhdf5.so"]

H5F_ACC_RDONLY = 0
H5F_ACC_TRUNC = 2
H5P_DEFAULT = 0
H5S_ALL = 0
H5S_SCALAR = 0
H5_INDEX_NAME = 0
H5_ITER_INC = 0


class H5GInfo(ctypes.Structure):
    _fields_ = [("storage_type", ctypes.c_int), ("nlinks", ctypes.c_uint64), ("max_corder", ctypes.c_int64),
                ("mounted", ctypes.c_bool)]


_LIB = None


def find():
    for p in CANDIDATES + [ctypes.util.find_library("hdf5") or ""]:
        if p and os.path.exists(p):
            return p
    return None


def load():
    global _LIB
    if _LIB is None:
        path = find()
        if path is None:
            raise RuntimeError("no libhdf5")
        lib = ctypes.CDLL(path, mode=ctypes.RTLD_GLOBAL)
        hid = ctypes.c_int64
        for name, res, args in [
            ("H5open", ctypes.c_int, []),
            ("H5Fopen", hid, [ctypes.c_char_p, ctypes.c_uint, hid]),
            ("H5Fcreate", hid, [ctypes.c_char_p, ctypes.c_uint, hid, hid]),
            ("H5Fclose", ctypes.c_int, [hid]),
            ("H5Dopen2", hid, [hid, ctypes.c_char_p, hid]),
            ("H5Dcreate2", hid, [hid, ctypes.c_char_p, hid, hid, hid, hid, hid]),
            ("H5Dget_space", hid, [hid]),
            ("H5Dread", ctypes.c_int, [hid, hid, hid, hid, hid, ctypes.c_void_p]),
            ("H5Dwrite", ctypes.c_int, [hid, hid, hid, hid, hid, ctypes.c_void_p]),
            ("H5Dclose", ctypes.c_int, [hid]),
            ("H5Screate", hid, [ctypes.c_int]),
            ("H5Screate_simple", hid, [ctypes.c_int, ctypes.POINTER(ctypes.c_uint64), ctypes.c_void_p]),
            ("H5Sget_simple_extent_ndims", ctypes.c_int, [hid]),
            ("H5Sget_simple_extent_dims", ctypes.c_int, [hid, ctypes.POINTER(ctypes.c_uint64), ctypes.c_void_p]),
            ("H5Sclose", ctypes.c_int, [hid]),
            ("H5Gget_info", ctypes.c_int, [hid, ctypes.POINTER(H5GInfo)]),
            ("H5Lget_name_by_idx", ctypes.c_ssize_t, [hid, ctypes.c_char_p, ctypes.c_int, ctypes.c_int, ctypes.c_uint64,
                                                      ctypes.c_char_p, ctypes.c_size_t, hid]),
        ]:
            fn = getattr(lib, name)
            fn.restype, fn.argtypes = res, args
        if lib.H5open() < 0:
            raise RuntimeError("H5open failed")
        _LIB = lib
    return _LIB


def _ok(v, what):
    if v < 0:
        raise RuntimeError(f"{what} failed")
    return v


def read(path):
    """{name: array} of the root group's datasets, read by the library."""
    lib = load()
    native = ctypes.c_int64.in_dll(lib, "H5T_NATIVE_DOUBLE_g").value
    f = _ok(lib.H5Fopen(path.encode(), H5F_ACC_RDONLY, H5P_DEFAULT), "H5Fopen")
    out = {}
    try:
        info = H5GInfo()
        _ok(lib.H5Gget_info(f, ctypes.byref(info)), "H5Gget_info")
        for i in range(info.nlinks):
            buf = ctypes.create_string_buffer(256)
            _ok(lib.H5Lget_name_by_idx(f, b".", H5_INDEX_NAME, H5_ITER_INC, i, buf, 256, H5P_DEFAULT), "name")
            name = buf.value.decode()
            d = _ok(lib.H5Dopen2(f, name.encode(), H5P_DEFAULT), "H5Dopen2")
            s = _ok(lib.H5Dget_space(d), "H5Dget_space")
            rank = _ok(lib.H5Sget_simple_extent_ndims(s), "ndims")
            dims = (ctypes.c_uint64 * max(rank, 1))()
            _ok(lib.H5Sget_simple_extent_dims(s, dims, None), "dims")
            shape = tuple(int(dims[k]) for k in range(rank))
            arr = np.empty(shape, dtype=np.float64)
            _ok(lib.H5Dread(d, native, H5S_ALL, H5S_ALL, H5P_DEFAULT, arr.ctypes.data), "H5Dread")
            lib.H5Sclose(s)
            lib.H5Dclose(d)
            out[name] = arr
    finally:
        lib.H5Fclose(f)
    return out


def write(path, datasets):
    """Write {name: array-like} as IEEE float64 datasets with the library's
    default layout (what h5py's create_dataset writes)."""
    lib = load()
    native = ctypes.c_int64.in_dll(lib, "H5T_NATIVE_DOUBLE_g").value
    f64le = ctypes.c_int64.in_dll(lib, "H5T_IEEE_F64LE_g").value
    f = _ok(lib.H5Fcreate(path.encode(), H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT), "H5Fcreate")
    try:
        for name, v in datasets.items():
            arr = np.array(v, dtype=np.float64, order="C")
            if arr.ndim == 0:
                s = _ok(lib.H5Screate(H5S_SCALAR), "H5Screate")
            else:
                dims = (ctypes.c_uint64 * arr.ndim)(*arr.shape)
                s = _ok(lib.H5Screate_simple(arr.ndim, dims, None), "H5Screate_simple")
            d = _ok(lib.H5Dcreate2(f, name.encode(), f64le, s, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), "H5Dcreate2")
            _ok(lib.H5Dwrite(d, native, H5S_ALL, H5S_ALL, H5P_DEFAULT, arr.ctypes.data), "H5Dwrite")
            lib.H5Dclose(d)
            lib.H5Sclose(s)
    finally:
        lib.H5Fclose(f)
