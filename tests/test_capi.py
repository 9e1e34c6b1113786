"""CPU checks of the drop-in boundary: the C-ABI library loads and exports
every symbol the public headers declare, and argument errors behave like the
reference's (no GPU is touched)."""
from __future__ import annotations

import ctypes
import os
import re

import numpy as np
import pytest

from conftest import ROOT

INCLUDE = os.path.join(ROOT, "include")


# the HDF5 file layer is a library of its own (libhpdbscan_amd_h5.so),
# checked in test_h5.py
H5_HEADER = "hpdbscan_h5.h"


def declared_functions(headers=None):
    names = []
    for fn in sorted(os.listdir(INCLUDE)):
        if not fn.endswith(".h"):
            continue
        if headers is None and fn == H5_HEADER:
            continue
        if headers is not None and fn not in headers:
            continue
        text = open(os.path.join(INCLUDE, fn)).read()
        if 'extern "C"' not in text:
            continue
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"^\s*(?:const\s+)?[A-Za-z_][\w\s\*]*?\b(hpd_\w+)\s*\(", text, flags=re.M):
            names.append(m.group(1))
    return sorted(set(names))


def test_headers_declare_entry_points():
    names = declared_functions()
    for required in ("hpd_cluster", "hpd_cluster_device", "hpd_context_create",
                     "hpd_context_destroy", "hpd_last_error"):
        assert required in names


def test_library_exports_every_declared_symbol():
    from hpdbscan_amd import _lib
    lib = _lib.load()
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing
    # the Python binding covers every C entry point too
    assert set(declared_functions()) <= set(_lib.SIGNATURES)


def test_version_and_null_context():
    from hpdbscan_amd import _lib
    lib = _lib.load()
    assert b"gfx950" in lib.hpd_version()
    x = np.zeros((3, 2), np.float32)
    out = np.zeros(3, np.int64)
    rc = lib.hpd_cluster(None, 0, x.ctypes.data, 3, 2, ctypes.c_float(0.5), 2, out.ctypes.data)
    assert rc == _lib.HPD_EINVAL
    assert b"context" in lib.hpd_last_error()


def test_hpdbscan_constructor_errors():
    import hpdbscan_amd as h
    with pytest.raises(ValueError, match="epsilon needs to be positive"):
        h.HPDBSCAN(0.0, 4)
    with pytest.raises(ValueError, match="epsilon needs to be positive"):
        h.HPDBSCAN(-0.5, 4)
    with pytest.raises(OverflowError):
        h.HPDBSCAN(0.3, -1)
    with pytest.raises(TypeError):
        h.HPDBSCAN(0.3, 2.5)
    with pytest.raises(TypeError):
        h.HPDBSCAN("0.3", 4)
    with pytest.raises(OverflowError):
        h.HPDBSCAN(1e300, 4)
    obj = h.HPDBSCAN(epsilon=0.3, min_points=4)
    assert obj.min_points == 4 and abs(obj.epsilon - 0.3) < 1e-7


def test_cluster_argument_errors_before_device():
    import hpdbscan_amd as h
    obj = h.HPDBSCAN(0.3, 4)
    with pytest.raises(TypeError, match="2 dimensions"):
        obj.cluster(np.zeros(5))
    with pytest.raises(TypeError):
        obj.cluster(np.zeros((2, 2, 2)))
    with pytest.raises(ValueError, match="^Could not open /nonexistent/data.h5$"):
        obj.cluster("/nonexistent/data.h5", "DATA")
    with pytest.raises(TypeError):
        obj.cluster("data.h5", "DATA", 2.5)
    with pytest.raises(TypeError):
        obj.cluster()


def test_no_device_here_is_reported():
    from hpdbscan_amd import _lib
    lib = _lib.load()
    if lib.hpd_device_count() > 0:
        pytest.skip("a GPU is visible")
    with pytest.raises(_lib.HPDError):
        _lib.Context(0)


def test_cxx_header_compiles():
    """include/hpdbscan.h (the drop-in C++ class) compiles as C++11 and links."""
    import shutil
    import subprocess
    import tempfile
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ missing")
    src = r'''
    #include "hpdbscan.h"
    #include <cstdio>
    int main() {
        try { HPDBSCAN bad(0.0f, 4); return 1; } catch (const std::invalid_argument& e) {
            if (std::string(e.what()) != "epsilon needs to be positive") return 2; }
        HPDBSCAN h(0.3f, 4);
        (void)h;
        return 0;
    }'''
    with tempfile.TemporaryDirectory() as td:
        cpp = os.path.join(td, "t.cpp")
        open(cpp, "w").write(src)
        exe = os.path.join(td, "t")
        lib_dir = os.path.join(ROOT, "hpdbscan_amd")
        r = subprocess.run([gxx, "-std=c++11", "-I" + INCLUDE, cpp, "-o", exe, "-L" + lib_dir,
                            "-lhpdbscan_amd", "-Wl,-rpath," + lib_dir], capture_output=True, text=True)
        assert r.returncode == 0, r.stderr
        r = subprocess.run([exe], capture_output=True, text=True)
        assert r.returncode == 0, (r.returncode, r.stdout, r.stderr)


def test_batch_argument_errors():
    from hpdbscan_amd import _lib
    lib = _lib.load()
    x = np.zeros((4, 2), np.float32)
    off = np.array([0, 2, 4], np.int64)
    eps = np.array([0.5, 0.0], np.float32)
    mp = np.array([2, 2], np.uint64)
    out = np.zeros(4, np.int64)
    rc = lib.hpd_cluster_batch(None, 0, x.ctypes.data, off.ctypes.data, 2, 2, eps.ctypes.data,
                               mp.ctypes.data, out.ctypes.data)
    assert rc == _lib.HPD_EINVAL


def test_neighbor_row_total_guard():
    """The neighbour table of a run is indexed with 32-bit offsets; a 64-bit
    total above 2^31-1 is refused before any table is sized from it."""
    from hpdbscan_amd import _lib
    lib = _lib.load()
    assert lib.hpd_check_neighbor_rows(0) == _lib.HPD_OK
    assert lib.hpd_check_neighbor_rows(2**31 - 1) == _lib.HPD_OK
    # 8-D: one million occupied cells with all 3^7 rows non-empty
    for total in (2**31, 2**32, 2**32 + 5, 1_000_000 * 3**7, 2**64 - 1):
        assert lib.hpd_check_neighbor_rows(total) == _lib.HPD_EUNSUPPORTED
        msg = lib.hpd_last_error().decode()
        assert str(total) in msg and "2^31-1" in msg


def test_reference_module_name():
    """`import hpdbscan` works as with the reference's SWIG module
    (swig/hpdbscan.i:1, README.md:69-72) and exposes the same class."""
    import importlib
    import hpdbscan
    import hpdbscan_amd.hpdbscan as impl
    importlib.reload(hpdbscan)
    assert hpdbscan.HPDBSCAN is impl.HPDBSCAN
    c = hpdbscan.HPDBSCAN(epsilon=0.3, min_points=4)
    assert c.min_points == 4
    with pytest.raises(ValueError, match="epsilon needs to be positive"):
        hpdbscan.HPDBSCAN(0, 4)
    with pytest.raises(TypeError, match="2 dimensions"):
        c.cluster(np.zeros(5))
