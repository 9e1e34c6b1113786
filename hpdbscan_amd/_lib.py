"""ctypes binding of libhpdbscan_amd.so (declared in include/hpdbscan_capi.h).

The library is loaded from this package directory.  There is no fallback: if
it is missing or cannot be loaded, every call raises ``RuntimeError``.
"""
from __future__ import annotations

import ctypes
import os
import threading

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libhpdbscan_amd.so")

HPD_OK, HPD_EINVAL, HPD_ENOMEM, HPD_EDEVICE, HPD_EUNSUPPORTED = 0, 1, 2, 3, 4

BORDER_MIN_CLUSTER = 0
BORDER_ONE_THREAD = 1

DTYPE_CODES = {
    np.dtype(np.float32): 0,
    np.dtype(np.float64): 1,
    np.dtype(np.int8): 2,
    np.dtype(np.int16): 3,
    np.dtype(np.int32): 4,
    np.dtype(np.int64): 5,
    np.dtype(np.uint8): 6,
    np.dtype(np.uint16): 7,
    np.dtype(np.uint32): 8,
    np.dtype(np.uint64): 9,
}


class RunStats(ctypes.Structure):
    """Mirror of hpd_run_stats."""

    _fields_ = [
        ("n", ctypes.c_int64),
        ("d", ctypes.c_int32),
        ("key_bits", ctypes.c_int32),
        ("total_cells", ctypes.c_uint64),
        ("occupied_cells", ctypes.c_uint64),
        ("core_points", ctypes.c_uint64),
        ("noise_points", ctypes.c_uint64),
        ("ambiguous_borders", ctypes.c_uint64),
        ("sum_candidates", ctypes.c_uint64),
        ("ms_grid", ctypes.c_double),
        ("ms_sort", ctypes.c_double),
        ("ms_index", ctypes.c_double),
        ("ms_core", ctypes.c_double),
        ("ms_union", ctypes.c_double),
        ("ms_border", ctypes.c_double),
        ("ms_total", ctypes.c_double),
        ("core_launches", ctypes.c_int32),
        ("ms_core_kernel", ctypes.c_double),
        ("boxes", ctypes.c_uint64),
        ("box_union", ctypes.c_int32),
    ]

    def as_dict(self) -> dict:
        return {name: getattr(self, name) for name, _ in self._fields_}


# every function of include/hpdbscan_capi.h: name -> (restype, argtypes)
_vp = ctypes.c_void_p
SIGNATURES = {
    "hpd_last_error": (ctypes.c_char_p, []),
    "hpd_version": (ctypes.c_char_p, []),
    "hpd_device_count": (ctypes.c_int, []),
    "hpd_context_create": (ctypes.c_int, [ctypes.c_int, ctypes.POINTER(_vp)]),
    "hpd_context_destroy": (ctypes.c_int, [_vp]),
    "hpd_context_set_border_rule": (ctypes.c_int, [_vp, ctypes.c_int]),
    "hpd_context_set_timing": (ctypes.c_int, [_vp, ctypes.c_int]),
    "hpd_context_get_stats": (ctypes.c_int, [_vp, ctypes.POINTER(RunStats)]),
    "hpd_cluster": (ctypes.c_int, [_vp, ctypes.c_int, _vp, ctypes.c_int64, ctypes.c_int32,
                                   ctypes.c_float, ctypes.c_uint64, _vp]),
    "hpd_cluster_device": (ctypes.c_int, [_vp, ctypes.c_int, _vp, ctypes.c_int64, ctypes.c_int32,
                                          ctypes.c_float, ctypes.c_uint64, _vp, _vp]),
    "hpd_device_alloc": (ctypes.c_int, [_vp, ctypes.c_size_t, ctypes.POINTER(_vp)]),
    "hpd_device_free": (ctypes.c_int, [_vp, _vp]),
    "hpd_memcpy_to_device": (ctypes.c_int, [_vp, _vp, _vp, ctypes.c_size_t]),
    "hpd_memcpy_to_host": (ctypes.c_int, [_vp, _vp, _vp, ctypes.c_size_t]),
    "hpd_context_stream": (ctypes.c_int, [_vp, ctypes.POINTER(_vp)]),
    "hpd_context_synchronize": (ctypes.c_int, [_vp]),
}

_lib = None
_lock = threading.Lock()


def load() -> ctypes.CDLL:
    """Load the HIP library (once per process)."""
    global _lib
    with _lock:
        if _lib is None:
            if not os.path.exists(LIB_PATH):
                raise RuntimeError(
                    f"{LIB_PATH} is not built; run `python -m hpdbscan_amd.build` first")
            lib = ctypes.CDLL(LIB_PATH)
            for name, (res, args) in SIGNATURES.items():
                fn = getattr(lib, name)
                fn.restype = res
                fn.argtypes = args
            _lib = lib
    return _lib


class HPDError(RuntimeError):
    def __init__(self, code: int, message: str):
        super().__init__(message)
        self.code = code


def check(code: int) -> None:
    """Raise the Python exception matching an hpd_status."""
    if code == HPD_OK:
        return
    msg = (load().hpd_last_error() or b"").decode(errors="replace")
    if code == HPD_EINVAL:
        raise ValueError(msg)
    if code == HPD_ENOMEM:
        raise MemoryError(msg)
    if code == HPD_EUNSUPPORTED:
        raise NotImplementedError(msg)
    raise HPDError(code, msg)


class Context:
    """An hpd_context: one HIP stream and a device workspace on one GPU."""

    def __init__(self, device: int = 0):
        lib = load()
        h = _vp()
        check(lib.hpd_context_create(device, ctypes.byref(h)))
        self._h = h
        self.device = device

    @property
    def handle(self):
        return self._h

    def close(self) -> None:
        if getattr(self, "_h", None):
            check(load().hpd_context_destroy(self._h))
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_border_rule(self, rule: int) -> None:
        check(load().hpd_context_set_border_rule(self._h, rule))

    def set_timing(self, enabled: bool) -> None:
        check(load().hpd_context_set_timing(self._h, 1 if enabled else 0))

    def synchronize(self) -> None:
        check(load().hpd_context_synchronize(self._h))

    def stream(self) -> int:
        p = _vp()
        check(load().hpd_context_stream(self._h, ctypes.byref(p)))
        return p.value or 0

    def stats(self) -> dict:
        st = RunStats()
        check(load().hpd_context_get_stats(self._h, ctypes.byref(st)))
        return st.as_dict()

    def cluster(self, data: np.ndarray, eps: float, min_points: int) -> np.ndarray:
        """Labels of the rows of a C-contiguous 2-D array (any of the ten dtypes)."""
        if data.ndim != 2:
            raise TypeError("data must be 2-D")
        if data.dtype not in DTYPE_CODES:
            raise TypeError(f"unsupported dtype {data.dtype}")
        a = np.ascontiguousarray(data)
        n, d = a.shape
        out = np.empty(n, dtype=np.int64)
        check(load().hpd_cluster(self._h, DTYPE_CODES[a.dtype], a.ctypes.data, n, d,
                                 ctypes.c_float(eps), int(min_points), out.ctypes.data))
        return out

    def cluster_device(self, data_ptr: int, n: int, d: int, dtype, eps: float, min_points: int,
                       labels_ptr: int, stream: int = 0) -> None:
        """Device-resident variant: pointers are device addresses (e.g. a torch
        tensor's data_ptr()); `stream` is a hipStream_t or 0."""
        code = DTYPE_CODES[np.dtype(dtype)]
        check(load().hpd_cluster_device(self._h, code, _vp(data_ptr), n, d, ctypes.c_float(eps),
                                        int(min_points), _vp(labels_ptr),
                                        _vp(stream) if stream else None))


class DeviceBuffer:
    """Device memory owned through a context (freed on close / GC)."""

    def __init__(self, ctx: Context, nbytes: int):
        self.ctx = ctx
        self.nbytes = int(nbytes)
        p = _vp()
        check(load().hpd_device_alloc(ctx.handle, self.nbytes, ctypes.byref(p)))
        self.ptr = p.value or 0

    @classmethod
    def from_array(cls, ctx: Context, a: np.ndarray) -> "DeviceBuffer":
        a = np.ascontiguousarray(a)
        buf = cls(ctx, a.nbytes)
        check(load().hpd_memcpy_to_device(ctx.handle, _vp(buf.ptr), a.ctypes.data, a.nbytes))
        return buf

    def to_array(self, dtype, count: int) -> np.ndarray:
        out = np.empty(count, dtype=dtype)
        check(load().hpd_memcpy_to_host(self.ctx.handle, out.ctypes.data, _vp(self.ptr),
                                        out.nbytes))
        return out

    def close(self) -> None:
        if self.ptr:
            check(load().hpd_device_free(self.ctx.handle, _vp(self.ptr)))
            self.ptr = 0

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


_default: dict[int, Context] = {}


def default_context(device: int | None = None) -> Context:
    if device is None:
        device = int(os.environ.get("HPD_DEVICE", "0"))
    with _lock:
        ctx = _default.get(device)
    if ctx is None:
        ctx = Context(device)
        with _lock:
            _default.setdefault(device, ctx)
            ctx = _default[device]
    return ctx
