"""HDF5 file layer: ctypes binding of libhpdbscan_amd_h5.so
(include/hpdbscan_h5.h).

* :func:`read_hdf5` -- a 2-D dataset in its stored element type
  (the reference's IO::read_hdf5, include/io.h:35-111)
* :func:`write_hdf5` -- labels as a 1-D native-long dataset, into a file
  that is created if missing (IO::write_hdf5, include/io.h:113-194)
* :func:`cluster_file` -- HPDBSCAN::cluster(path, dataset[, threads])
  (include/hpdbscan.h:300-356): read, then cluster on the GPU
* :func:`read_clusters`, :func:`write_matrix` -- reading labels back and
  writing input datasets (no counterpart in the reference, whose scripts use
  h5py for both)

Errors: the reference's std::invalid_argument cases raise ValueError, its
std::runtime_error cases RuntimeError, with the reference's messages.
There is no fallback: without the library every call raises RuntimeError.
"""
from __future__ import annotations

import ctypes
import os
import threading

import numpy as np

from . import _lib

H5_LIB_PATH = os.path.join(_lib.HERE, "libhpdbscan_amd_h5.so")
HPD_EIO = 6

_h5 = None
_h5_lock = threading.Lock()

_DTYPES = {code: dt for dt, code in _lib.DTYPE_CODES.items()}


def load() -> ctypes.CDLL:
    """Load libhpdbscan_amd_h5.so (after libhpdbscan_amd.so, which it uses)."""
    global _h5
    if _h5 is not None:
        return _h5
    with _h5_lock:
        if _h5 is not None:
            return _h5
        _lib.load()
        if not os.path.exists(H5_LIB_PATH):
            raise RuntimeError(f"{H5_LIB_PATH} is missing: run `python -m hpdbscan_amd.build` "
                               "where libhdf5 1.10 is installed")
        lib = ctypes.CDLL(H5_LIB_PATH)
        c = ctypes
        i64p = c.POINTER(c.c_int64)
        i32p = c.POINTER(c.c_int32)
        sigs = {
            "hpd_h5_last_error": (c.c_char_p, []),
            "hpd_h5_matrix_info": (c.c_int, [c.c_char_p, c.c_char_p, i64p, i32p, i32p]),
            "hpd_h5_read_matrix": (c.c_int, [c.c_char_p, c.c_char_p, c.c_int32, c.c_void_p,
                                             c.c_int64, c.c_int32]),
            "hpd_cluster_h5": (c.c_int, [c.c_void_p, c.c_char_p, c.c_char_p, c.c_float,
                                         c.c_int64, c.c_void_p, c.c_int64]),
            "hpd_h5_write_clusters": (c.c_int, [c.c_char_p, c.c_char_p, c.c_void_p, c.c_int64]),
            "hpd_h5_clusters_length": (c.c_int, [c.c_char_p, c.c_char_p, i64p]),
            "hpd_h5_read_clusters": (c.c_int, [c.c_char_p, c.c_char_p, c.c_void_p, c.c_int64]),
            "hpd_h5_write_matrix": (c.c_int, [c.c_char_p, c.c_char_p, c.c_int32, c.c_void_p,
                                              c.c_int64, c.c_int32]),
        }
        for name, (res, args) in sigs.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _h5 = lib
    return _h5


def _check(code: int) -> None:
    if code == _lib.HPD_OK:
        return
    msg = (load().hpd_h5_last_error() or b"").decode(errors="replace")
    if code == _lib.HPD_EINVAL:
        raise ValueError(msg)
    if code == _lib.HPD_ENOMEM:
        raise MemoryError(msg)
    if code == _lib.HPD_EUNSUPPORTED:
        raise NotImplementedError(msg)
    if code == HPD_EIO:
        raise RuntimeError(msg)
    raise _lib.HPDError(code, msg)


def _b(s) -> bytes:
    if isinstance(s, bytes):
        return s
    if not isinstance(s, str):
        raise TypeError(f"expected a str, got {type(s).__name__}")
    return os.fsencode(s)


def matrix_info(path, dataset) -> tuple[int, int, np.dtype]:
    """(rows, cols, dtype) of a 2-D dataset."""
    rows, cols, code = ctypes.c_int64(), ctypes.c_int32(), ctypes.c_int32()
    _check(load().hpd_h5_matrix_info(_b(path), _b(dataset), ctypes.byref(rows),
                                     ctypes.byref(cols), ctypes.byref(code)))
    return rows.value, cols.value, _DTYPES[code.value]


def read_hdf5(path, dataset, dtype=None) -> np.ndarray:
    """The 2-D dataset as an array of its stored element type (or `dtype`)."""
    rows, cols, stored = matrix_info(path, dataset)
    dt = np.dtype(dtype) if dtype is not None else stored
    if dt not in _lib.DTYPE_CODES:
        raise TypeError(f"unsupported dtype {dt}")
    out = np.empty((rows, cols), dtype=dt)
    _check(load().hpd_h5_read_matrix(_b(path), _b(dataset), _lib.DTYPE_CODES[dt],
                                     out.ctypes.data, rows, cols))
    return out


def write_hdf5(path, dataset, clusters) -> None:
    """Labels as the 1-D native-long dataset `dataset` of `path`."""
    a = np.ascontiguousarray(np.asarray(clusters, dtype=np.int64).reshape(-1))
    _check(load().hpd_h5_write_clusters(_b(path), _b(dataset), a.ctypes.data, a.size))


def read_clusters(path, dataset) -> np.ndarray:
    """A 1-D dataset (the labels) as int64."""
    n = ctypes.c_int64()
    _check(load().hpd_h5_clusters_length(_b(path), _b(dataset), ctypes.byref(n)))
    out = np.empty(n.value, dtype=np.int64)
    _check(load().hpd_h5_read_clusters(_b(path), _b(dataset), out.ctypes.data, n.value))
    return out


def write_matrix(path, dataset, data) -> None:
    """A new 2-D dataset holding `data` in its own element type."""
    a = np.ascontiguousarray(data)
    if a.ndim != 2:
        raise ValueError(f"expected a 2-D array, got {a.ndim} dimensions")
    if a.dtype not in _lib.DTYPE_CODES:
        raise TypeError(f"unsupported dtype {a.dtype}")
    _check(load().hpd_h5_write_matrix(_b(path), _b(dataset), _lib.DTYPE_CODES[a.dtype],
                                      a.ctypes.data, a.shape[0], a.shape[1]))


def cluster_file(path, dataset, eps: float, min_points: int, ctx=None) -> np.ndarray:
    """Labels (file order) of the 2-D dataset, clustered in its stored type."""
    rows, _, _ = matrix_info(path, dataset)
    ctx = ctx or _lib.default_context()
    out = np.empty(rows, dtype=np.int64)
    _check(load().hpd_cluster_h5(ctx.handle, _b(path), _b(dataset), float(eps), int(min_points),
                                 out.ctypes.data, rows))
    return out
