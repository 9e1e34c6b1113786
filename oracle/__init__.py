"""CPU oracle for HPDBSCAN's single-node clustering path (test infrastructure).

Only ``tests/``, ``__graft_entry__.smoke()`` and the CPU-baseline leg of
``bench.py`` import this package, and only as the checker or as the timed CPU
reference.  The product (``hpdbscan_amd``) never imports it.

Two CPU implementations are reachable from here:

* ``liboracle.so`` -- ``oracle.c``, a sequential restatement of the reference
  at one thread (see the header of ``oracle.c`` for the file:line map).
* ``_ref/libhpdbscan_ref.so`` -- the reference itself, compiled from its own
  headers with its own flags by ``oracle/Makefile`` when the reference sources
  are mounted.  On a machine without them the ``ref_*`` functions raise
  ``RuntimeError`` and ``ref_available()`` is False.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "liboracle.so")
REF_SO = os.path.join(HERE, "_ref", "libhpdbscan_ref.so")
REF_RUN = os.path.join(HERE, "_ref", "hpdbscan_ref_run")

# dtype codes shared with oracle.c
DTYPES = {
    np.dtype(np.float32): (0, "f32"),
    np.dtype(np.float64): (1, "f64"),
    np.dtype(np.int8): (2, "i8"),
    np.dtype(np.int16): (3, "i16"),
    np.dtype(np.int32): (4, "i32"),
    np.dtype(np.int64): (5, "i64"),
    np.dtype(np.uint8): (6, "u8"),
    np.dtype(np.uint16): (7, "u16"),
    np.dtype(np.uint32): (8, "u32"),
    np.dtype(np.uint64): (9, "u64"),
}


class OracleStats(ctypes.Structure):
    _fields_ = [
        ("sum_candidates", ctypes.c_uint64),
        ("occupied_cells", ctypes.c_uint64),
        ("total_cells", ctypes.c_uint64),
        ("core_points", ctypes.c_uint64),
        ("noise_points", ctypes.c_uint64),
        ("ambiguous_borders", ctypes.c_uint64),
        ("swapped", ctypes.c_int32 * 16),
        ("cells_per_dim", ctypes.c_uint64 * 16),
    ]


_oracle = None
_ref = None


def build(quiet: bool = True) -> None:
    """Compile liboracle.so and, where the reference is mounted, oracle/_ref."""
    subprocess.run(["make", "-C", HERE, "-B", "all"], check=True,
                   stdout=subprocess.DEVNULL if quiet else None)


def _lib():
    global _oracle
    if _oracle is None:
        if not os.path.exists(ORACLE_SO):
            raise RuntimeError(f"{ORACLE_SO} is not built; run `make -C oracle`")
        lib = ctypes.CDLL(ORACLE_SO)
        vp, i64, i32, u64 = ctypes.c_void_p, ctypes.c_int64, ctypes.c_int32, ctypes.c_uint64
        lib.hpd_oracle_cluster.argtypes = [ctypes.c_int, vp, i64, i32, ctypes.c_float, u64,
                                           vp, vp, vp, vp, vp, ctypes.POINTER(OracleStats)]
        lib.hpd_oracle_cluster.restype = ctypes.c_int
        lib.hpd_oracle_check.argtypes = [ctypes.c_int, vp, i64, i32, ctypes.c_float, u64, vp,
                                         ctypes.POINTER(u64), ctypes.POINTER(u64)]
        lib.hpd_oracle_check.restype = ctypes.c_int
        lib.hpd_oracle_check_parallel.argtypes = [ctypes.c_int, vp, i64, i32, ctypes.c_float, u64,
                                                  vp, ctypes.c_int, ctypes.POINTER(u64),
                                                  ctypes.POINTER(u64), ctypes.POINTER(u64),
                                                  ctypes.POINTER(u64)]
        lib.hpd_oracle_check_parallel.restype = ctypes.c_int
        _oracle = lib
    return _oracle


def _prep(data):
    a = np.ascontiguousarray(data)
    if a.ndim != 2:
        raise ValueError("data must be 2-D")
    if a.dtype not in DTYPES:
        raise TypeError(f"unsupported dtype {a.dtype}")
    return a, DTYPES[a.dtype]


def cluster(data, eps: float, min_pts: int, full: bool = False):
    """One-thread reference labels of ``data`` (input order).

    With ``full=True`` returns a dict with ``labels``, ``sorted_pos``,
    ``cell_of_sorted``, ``core``, ``ambiguous`` and ``stats``.
    """
    a, (code, _) = _prep(data)
    n, d = a.shape
    labels = np.empty(n, np.int64)
    pos = np.empty(n, np.int64)
    cells = np.empty(n, np.uint64)
    core = np.empty(n, np.uint8)
    amb = np.empty(n, np.uint8)
    st = OracleStats()
    rc = _lib().hpd_oracle_cluster(code, a.ctypes.data, n, d, ctypes.c_float(eps), min_pts,
                                   labels.ctypes.data, pos.ctypes.data, cells.ctypes.data,
                                   core.ctypes.data, amb.ctypes.data, ctypes.byref(st))
    if rc == 1:
        raise ValueError("oracle: invalid arguments")
    if rc != 0:
        raise MemoryError("oracle: out of memory")
    if not full:
        return labels
    stats = {f: getattr(st, f) for f, _ in OracleStats._fields_ if f not in ("swapped", "cells_per_dim")}
    stats["swapped"] = list(st.swapped)[:d]
    stats["cells_per_dim"] = list(st.cells_per_dim)[:d]
    return {"labels": labels, "sorted_pos": pos, "cell_of_sorted": cells, "core": core.astype(bool),
            "ambiguous": amb.astype(bool), "stats": stats}


def check(data, eps: float, min_pts: int, got) -> tuple[int, int]:
    """(mismatches outside ambiguous borders, ambiguous borders not within eps
    of a core of their assigned cluster) for candidate labels ``got``."""
    a, (code, _) = _prep(data)
    n, d = a.shape
    g = np.ascontiguousarray(got, dtype=np.int64)
    if g.shape != (n,):
        raise ValueError("labels must have one entry per row")
    m, b = ctypes.c_uint64(), ctypes.c_uint64()
    rc = _lib().hpd_oracle_check(code, a.ctypes.data, n, d, ctypes.c_float(eps), min_pts,
                                 g.ctypes.data, ctypes.byref(m), ctypes.byref(b))
    if rc != 0:
        raise ValueError(f"oracle check failed with code {rc}")
    return int(m.value), int(b.value)


def check_parallel(data, eps: float, min_pts: int, got, threads: int = 0) -> dict:
    """Like ``check`` for inputs too large for the one-thread restatement:
    ``got`` is compared with the closed form of the reference's result
    (core flags, components of cores, cluster ids, noise, one-cluster
    borders exact; an ambiguous border must carry a cluster with a core
    within eps), computed with ``threads`` OpenMP threads (0: all).
    Returns {"mismatches", "bad_ambiguous", "cores", "ambiguous"}."""
    a, (code, _) = _prep(data)
    n, d = a.shape
    g = np.ascontiguousarray(got, dtype=np.int64)
    if g.shape != (n,):
        raise ValueError("labels must have one entry per row")
    m, b, c, am = (ctypes.c_uint64() for _ in range(4))
    rc = _lib().hpd_oracle_check_parallel(code, a.ctypes.data, n, d, ctypes.c_float(eps), min_pts,
                                          g.ctypes.data, int(threads), ctypes.byref(m),
                                          ctypes.byref(b), ctypes.byref(c), ctypes.byref(am))
    if rc != 0:
        raise ValueError(f"oracle parallel check failed with code {rc}")
    return {"mismatches": int(m.value), "bad_ambiguous": int(b.value), "cores": int(c.value),
            "ambiguous": int(am.value)}


# --------------------------------------------------------------------------
# the compiled reference (only where oracle/_ref was built)
# --------------------------------------------------------------------------

def ref_available() -> bool:
    return os.path.exists(REF_SO)


def _ref_lib():
    global _ref
    if _ref is None:
        if not ref_available():
            raise RuntimeError("oracle/_ref is not built (reference sources absent)")
        lib = ctypes.CDLL(REF_SO)
        vp = ctypes.c_void_p
        for _, sfx in DTYPES.values():
            f = getattr(lib, f"hpd_ref_cluster_{sfx}")
            f.argtypes = [vp, ctypes.c_int, ctypes.c_int, ctypes.c_float, ctypes.c_uint64,
                          ctypes.c_int, vp, ctypes.c_char_p, ctypes.c_int]
            f.restype = ctypes.c_int
            g = getattr(lib, f"hpd_ref_spatial_{sfx}")
            g.argtypes = [vp, ctypes.c_int, ctypes.c_int, ctypes.c_float, vp, vp]
            g.restype = ctypes.c_int
            h = getattr(lib, f"hpd_ref_candidates_{sfx}")
            h.argtypes = [vp, ctypes.c_int, ctypes.c_int, ctypes.c_float, ctypes.POINTER(ctypes.c_uint64)]
            h.restype = ctypes.c_int
        _ref = lib
    return _ref


def ref_cluster(data, eps: float, min_pts: int, threads: int = 1):
    """Labels from the compiled reference, HPDBSCAN(eps, minPts).cluster<T>(...)."""
    a, (_, sfx) = _prep(data)
    n, d = a.shape
    out = np.empty(n, np.int64)
    err = ctypes.create_string_buffer(256)
    rc = getattr(_ref_lib(), f"hpd_ref_cluster_{sfx}")(a.ctypes.data, n, d, ctypes.c_float(eps),
                                                        min_pts, threads, out.ctypes.data, err, 256)
    if rc == 1:
        raise ValueError(err.value.decode())
    if rc != 0:
        raise RuntimeError(err.value.decode())
    return out


def ref_spatial(data, eps: float):
    """(cell id per sorted position, sorted position of every input row)."""
    a, (_, sfx) = _prep(data)
    n, d = a.shape
    cells = np.empty(n, np.uint64)
    pos = np.empty(n, np.int64)
    getattr(_ref_lib(), f"hpd_ref_spatial_{sfx}")(a.ctypes.data, n, d, ctypes.c_float(eps),
                                                  cells.ctypes.data, pos.ctypes.data)
    return cells, pos


def ref_candidates(data, eps: float) -> int:
    """Σ_p |get_neighbors(cell_of(p))| of the reference."""
    a, (_, sfx) = _prep(data)
    n, d = a.shape
    t = ctypes.c_uint64()
    getattr(_ref_lib(), f"hpd_ref_candidates_{sfx}")(a.ctypes.data, n, d, ctypes.c_float(eps),
                                                     ctypes.byref(t))
    return int(t.value)
