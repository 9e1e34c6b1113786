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

HPD_OK, HPD_EINVAL, HPD_ENOMEM, HPD_EDEVICE, HPD_EUNSUPPORTED, HPD_ETIMEOUT = 0, 1, 2, 3, 4, 5

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
        ("ms_hook_kernel", ctypes.c_double),
        ("ms_union_kernel", ctypes.c_double),
        ("ms_border_kernel", ctypes.c_double),
        ("boxes", ctypes.c_uint64),
        ("box_union", ctypes.c_int32),
        ("neighbor_rows", ctypes.c_uint64),
        ("open_box_pairs", ctypes.c_uint64),
        ("ms_core_kernel", ctypes.c_double),
        ("core_tests", ctypes.c_uint64),
    ]

    def as_dict(self) -> dict:
        return {name: getattr(self, name) for name, _ in self._fields_}


class ServerConfig(ctypes.Structure):
    """Mirror of hpd_server_config."""

    _fields_ = [
        ("device", ctypes.c_int),
        ("lanes", ctypes.c_int),
        ("hbm_budget_bytes", ctypes.c_size_t),
        ("bytes_per_row_estimate", ctypes.c_size_t),
        ("max_batch_jobs", ctypes.c_int32),
        ("max_batch_rows", ctypes.c_int64),
    ]


class RequestInfo(ctypes.Structure):
    _fields_ = [("rows", ctypes.c_int64), ("batch_jobs", ctypes.c_int32),
                ("queue_ms", ctypes.c_double), ("latency_ms", ctypes.c_double)]

    def as_dict(self) -> dict:
        return {name: getattr(self, name) for name, _ in self._fields_}


class ServerStats(ctypes.Structure):
    _fields_ = [
        ("completed", ctypes.c_uint64), ("failed", ctypes.c_uint64), ("batches", ctypes.c_uint64),
        ("mean_batch_jobs", ctypes.c_double), ("rows", ctypes.c_int64),
        ("p50_ms", ctypes.c_double), ("p99_ms", ctypes.c_double), ("max_ms", ctypes.c_double),
        ("wall_s", ctypes.c_double), ("jobs_per_s", ctypes.c_double),
        ("mrows_per_s", ctypes.c_double), ("queued", ctypes.c_uint64),
        ("charged_bytes", ctypes.c_size_t),
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
    "hpd_context_set_work_split": (ctypes.c_int, [_vp, ctypes.c_int32, ctypes.c_int]),
    "hpd_context_set_open_pair_limit": (ctypes.c_int, [_vp, ctypes.c_uint64]),
    "hpd_context_set_box_split": (ctypes.c_int, [_vp, ctypes.c_int32]),
    "hpd_context_get_box_split": (ctypes.c_int, [_vp, _vp]),
    "hpd_context_get_work_split": (ctypes.c_int, [_vp, _vp, _vp]),
    "hpd_context_set_timing": (ctypes.c_int, [_vp, ctypes.c_int]),
    "hpd_context_set_graph": (ctypes.c_int, [_vp, ctypes.c_int]),
    "hpd_context_get_graph_stats": (ctypes.c_int, [_vp, ctypes.POINTER(ctypes.c_uint64),
                                                   ctypes.POINTER(ctypes.c_uint64),
                                                   ctypes.POINTER(ctypes.c_uint64)]),
    "hpd_context_get_stats": (ctypes.c_int, [_vp, ctypes.POINTER(RunStats)]),
    "hpd_cluster": (ctypes.c_int, [_vp, ctypes.c_int, _vp, ctypes.c_int64, ctypes.c_int32,
                                   ctypes.c_float, ctypes.c_uint64, _vp]),
    "hpd_cluster_device": (ctypes.c_int, [_vp, ctypes.c_int, _vp, ctypes.c_int64, ctypes.c_int32,
                                          ctypes.c_float, ctypes.c_uint64, _vp, _vp]),
    "hpd_cluster_device_noalloc": (ctypes.c_int, [_vp, ctypes.c_int, _vp, ctypes.c_int64,
                                                  ctypes.c_int32, ctypes.c_float, ctypes.c_uint64,
                                                  _vp, _vp]),
    "hpd_cluster_batch": (ctypes.c_int, [_vp, ctypes.c_int, _vp, _vp, ctypes.c_int32, ctypes.c_int32,
                                         _vp, _vp, _vp]),
    "hpd_cluster_batch_device": (ctypes.c_int, [_vp, ctypes.c_int, _vp, _vp, ctypes.c_int32,
                                                ctypes.c_int32, _vp, _vp, _vp, _vp]),
    "hpd_check_neighbor_rows": (ctypes.c_int, [ctypes.c_uint64]),
    "hpd_device_alloc": (ctypes.c_int, [_vp, ctypes.c_size_t, ctypes.POINTER(_vp)]),
    "hpd_device_free": (ctypes.c_int, [_vp, _vp]),
    "hpd_memcpy_to_device": (ctypes.c_int, [_vp, _vp, _vp, ctypes.c_size_t]),
    "hpd_memcpy_to_host": (ctypes.c_int, [_vp, _vp, _vp, ctypes.c_size_t]),
    "hpd_context_stream": (ctypes.c_int, [_vp, ctypes.POINTER(_vp)]),
    "hpd_context_synchronize": (ctypes.c_int, [_vp]),
    "hpd_server_default_config": (None, [ctypes.POINTER(ServerConfig)]),
    "hpd_server_create": (ctypes.c_int, [ctypes.POINTER(ServerConfig), ctypes.POINTER(_vp)]),
    "hpd_server_submit": (ctypes.c_int, [_vp, ctypes.c_int, _vp, ctypes.c_int64, ctypes.c_int32,
                                         ctypes.c_float, ctypes.c_uint64,
                                         ctypes.POINTER(ctypes.c_uint64)]),
    "hpd_server_submit_device": (ctypes.c_int, [_vp, ctypes.c_int, _vp, ctypes.c_int64,
                                                ctypes.c_int32, ctypes.c_float, ctypes.c_uint64,
                                                _vp, ctypes.POINTER(ctypes.c_uint64)]),
    "hpd_server_wait": (ctypes.c_int, [_vp, ctypes.c_uint64, _vp, ctypes.c_int64, ctypes.c_double,
                                       ctypes.POINTER(RequestInfo)]),
    "hpd_server_get_stats": (ctypes.c_int, [_vp, ctypes.POINTER(ServerStats)]),
    "hpd_server_destroy": (ctypes.c_int, [_vp]),
    "hpd_server_last_error": (ctypes.c_char_p, []),
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

    def set_work_split(self, sparse_max: int, sparse_core: bool = False) -> None:
        """Cells with at most sparse_max rows go one thread per row (border
        pass; core count too with sparse_core).  Results do not change."""
        check(load().hpd_context_set_work_split(self._h, int(sparse_max), 1 if sparse_core else 0))

    def set_open_pair_limit(self, limit: int) -> None:
        """Most open box pairs the union lists hold (0: 2n); beyond it the
        union walks every box's neighbourhood.  Results do not change."""
        check(load().hpd_context_set_open_pair_limit(self._h, int(limit)))

    def set_box_split(self, parts: int) -> None:
        """Sub-boxes per cell and dimension (0: 2 up to 3-D, 3 above; 1-8
        set it).  Results do not change."""
        check(load().hpd_context_set_box_split(self._h, int(parts)))

    def box_split(self) -> int:
        k = ctypes.c_int32()
        check(load().hpd_context_get_box_split(self._h, ctypes.byref(k)))
        return int(k.value)

    def work_split(self) -> tuple[int, bool]:
        sm, sc = ctypes.c_int32(), ctypes.c_int()
        check(load().hpd_context_get_work_split(self._h, ctypes.byref(sm), ctypes.byref(sc)))
        return int(sm.value), bool(sc.value)

    def set_timing(self, enabled) -> None:
        """True / 1: phase events; 2: also count the core count's executed
        distance tests (an instrumented, slower kernel)."""
        check(load().hpd_context_set_timing(self._h, int(enabled)))

    def set_graph(self, enabled: bool) -> None:
        """Graph mode: floating-point runs take the static pipeline, captured
        in a hipGraph once per shape and replayed (same labels)."""
        check(load().hpd_context_set_graph(self._h, 1 if enabled else 0))

    def graph_stats(self) -> dict:
        v = [ctypes.c_uint64() for _ in range(3)]
        check(load().hpd_context_get_graph_stats(self._h, *(ctypes.byref(x) for x in v)))
        return dict(zip(("launches", "captures", "fallbacks"), (int(x.value) for x in v)))

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

    def cluster_device_noalloc(self, data_ptr: int, n: int, d: int, dtype, eps: float,
                               min_points: int, labels_ptr: int, stream: int = 0) -> None:
        """cluster_device with every workspace buffer in the library's static
        device scratch (no device allocation); inputs too large for it raise
        NotImplementedError."""
        code = DTYPE_CODES[np.dtype(dtype)]
        check(load().hpd_cluster_device_noalloc(self._h, code, _vp(data_ptr), n, d,
                                                ctypes.c_float(eps), int(min_points),
                                                _vp(labels_ptr), _vp(stream) if stream else None))


def _batch_args(offsets, eps, min_points):
    off = np.ascontiguousarray(offsets, dtype=np.int64)
    nj = len(off) - 1
    e = np.ascontiguousarray(np.broadcast_to(np.asarray(eps, dtype=np.float32), (nj,)))
    m = np.ascontiguousarray(np.broadcast_to(np.asarray(min_points, dtype=np.uint64), (nj,)))
    return off, nj, e, m


def concat_jobs(clouds):
    """(rows, offsets) of a list of 2-D arrays of one dtype and width."""
    clouds = [np.ascontiguousarray(c) for c in clouds]
    off = np.zeros(len(clouds) + 1, dtype=np.int64)
    off[1:] = np.cumsum([len(c) for c in clouds])
    if not clouds:
        return np.zeros((0, 1), np.float32), off
    return np.ascontiguousarray(np.concatenate(clouds)), off


def _cluster_batch(self, data: np.ndarray, offsets, eps, min_points) -> np.ndarray:
    """Labels of several clouds at once: rows [offsets[j], offsets[j+1]) form
    job j with eps[j] / min_points[j] (scalars apply to every job)."""
    a = np.ascontiguousarray(data)
    if a.ndim != 2:
        raise TypeError("data must be 2-D")
    if a.dtype not in DTYPE_CODES:
        raise TypeError(f"unsupported dtype {a.dtype}")
    off, nj, e, m = _batch_args(offsets, eps, min_points)
    if nj > 0 and off[-1] != a.shape[0]:
        raise ValueError("offsets[-1] must equal the number of rows")
    out = np.empty(a.shape[0], dtype=np.int64)
    check(load().hpd_cluster_batch(self._h, DTYPE_CODES[a.dtype], a.ctypes.data, off.ctypes.data, nj,
                                   a.shape[1], e.ctypes.data, m.ctypes.data, out.ctypes.data))
    return out


def _cluster_batch_device(self, data_ptr: int, offsets, d: int, dtype, eps, min_points,
                          labels_ptr: int, stream: int = 0) -> None:
    off, nj, e, m = _batch_args(offsets, eps, min_points)
    check(load().hpd_cluster_batch_device(self._h, DTYPE_CODES[np.dtype(dtype)], _vp(data_ptr),
                                          off.ctypes.data, nj, d, e.ctypes.data, m.ctypes.data,
                                          _vp(labels_ptr), _vp(stream) if stream else None))


Context.cluster_batch = _cluster_batch
Context.cluster_batch_device = _cluster_batch_device


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

    def copy_from(self, a: np.ndarray) -> None:
        """Copy a host array into the buffer (synchronous)."""
        a = np.ascontiguousarray(a)
        if a.nbytes > self.nbytes:
            raise ValueError("array larger than the buffer")
        check(load().hpd_memcpy_to_device(self.ctx.handle, _vp(self.ptr), a.ctypes.data, a.nbytes))

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


def _check_server(code: int) -> None:
    if code == HPD_OK:
        return
    msg = (load().hpd_server_last_error() or b"").decode(errors="replace")
    if code == HPD_EINVAL:
        raise ValueError(msg)
    if code == HPD_ENOMEM:
        raise MemoryError(msg)
    if code == HPD_ETIMEOUT:
        raise TimeoutError(msg)
    raise HPDError(code, msg)


class Server:
    """Continuous-batching clustering server (hpd_server_*)."""

    def __init__(self, device: int = 0, lanes: int = 3, hbm_budget_bytes: int | None = None,
                 bytes_per_row_estimate: int | None = None, max_batch_jobs: int | None = None,
                 max_batch_rows: int | None = None):
        lib = load()
        cfg = ServerConfig()
        lib.hpd_server_default_config(ctypes.byref(cfg))
        cfg.device = device
        cfg.lanes = lanes
        if hbm_budget_bytes is not None:
            cfg.hbm_budget_bytes = hbm_budget_bytes
        if bytes_per_row_estimate is not None:
            cfg.bytes_per_row_estimate = bytes_per_row_estimate
        if max_batch_jobs is not None:
            cfg.max_batch_jobs = max_batch_jobs
        if max_batch_rows is not None:
            cfg.max_batch_rows = max_batch_rows
        h = _vp()
        _check_server(lib.hpd_server_create(ctypes.byref(cfg), ctypes.byref(h)))
        self._h = h
        self._n = {}

    def submit(self, data: np.ndarray, eps: float, min_points: int) -> int:
        a = np.ascontiguousarray(data)
        if a.ndim != 2 or a.dtype not in DTYPE_CODES:
            raise TypeError("data must be a 2-D array of a supported dtype")
        t = ctypes.c_uint64()
        _check_server(load().hpd_server_submit(self._h, DTYPE_CODES[a.dtype], a.ctypes.data,
                                               a.shape[0], a.shape[1], ctypes.c_float(eps),
                                               int(min_points), ctypes.byref(t)))
        self._n[t.value] = a.shape[0]
        return t.value

    def submit_device(self, data_ptr: int, n: int, d: int, dtype, eps: float, min_points: int,
                      labels_ptr: int) -> int:
        """A request whose rows (n x d of dtype) and labels (n int64) are in
        device memory; both must stay valid until wait returns."""
        t = ctypes.c_uint64()
        _check_server(load().hpd_server_submit_device(self._h, DTYPE_CODES[np.dtype(dtype)],
                                                      data_ptr, int(n), int(d), ctypes.c_float(eps),
                                                      int(min_points), labels_ptr, ctypes.byref(t)))
        self._n[t.value] = None
        return t.value

    def wait(self, ticket: int, timeout_ms: float = -1.0):
        """(labels, info dict) of a ticket; labels None for a device request
        (they are in its labels buffer)."""
        n = self._n[ticket]
        out = None if n is None else np.empty(n, dtype=np.int64)
        info = RequestInfo()
        _check_server(load().hpd_server_wait(self._h, ticket, None if out is None else out.ctypes.data,
                                             0 if n is None else n, timeout_ms, ctypes.byref(info)))
        del self._n[ticket]
        return out, info.as_dict()

    def stats(self) -> dict:
        st = ServerStats()
        _check_server(load().hpd_server_get_stats(self._h, ctypes.byref(st)))
        return st.as_dict()

    def close(self) -> None:
        if getattr(self, "_h", None):
            load().hpd_server_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
