"""The drop-in C++ class (include/hpdbscan.h, include/hpdbscan_dataset.h) on
the GPU: a C++11 program built with g++ against the headers calls
HPDBSCAN::cluster<T>(T*, int, int[, int]) and cluster<T>(Dataset&[, int])
(the reference's hpdbscan.h:358-359, 475-486) on golden fixtures written by
the compiled reference, and the labels are compared with the fixtures."""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

import numpy as np
import pytest

import oracle
from conftest import GOLDEN, ROOT

pytestmark = pytest.mark.gpu

INCLUDE = os.path.join(ROOT, "include")
LIB_DIR = os.path.join(ROOT, "hpdbscan_amd")
H5_LINKS = os.path.join(LIB_DIR, "_h5lib")
HDF5_INCLUDE = os.path.join(os.environ.get("HDF5_PREFIX", "/opt/conda"), "include")

DRIVER = r'''
#include "hpdbscan.h"
#include "hpdbscan_dataset.h"

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

template <typename T>
static int run(const char* in, int n, int d, float eps, size_t mp, hid_t type, const char* out) {
    std::vector<T> rows(static_cast<size_t>(n) * d);
    FILE* f = std::fopen(in, "rb");
    if (!f || std::fread(rows.data(), sizeof(T), rows.size(), f) != rows.size()) return 10;
    std::fclose(f);
    const std::vector<T> copy = rows;
    HPDBSCAN h(eps, mp);
    const Clusters a = h.cluster<T>(rows.data(), n, d);
    const Clusters b = h.cluster(rows.data(), n, d, 4);
    if (a != b) return 11;
    if (std::memcmp(rows.data(), copy.data(), rows.size() * sizeof(T)) != 0) return 12;
    const hsize_t shape[2] = {static_cast<hsize_t>(n), static_cast<hsize_t>(d)};
    Dataset ds(rows.data(), shape, type);
    const Clusters c = h.cluster<T>(ds);
    const Clusters e = h.cluster<T>(ds, 2);
    if (a != c || a != e) return 13;
    if (std::memcmp(ds.m_p, copy.data(), rows.size() * sizeof(T)) != 0) return 14;
    f = std::fopen(out, "wb");
    if (!f || std::fwrite(a.data(), sizeof(Cluster), a.size(), f) != a.size()) return 15;
    std::fclose(f);
    return 0;
}

int main(int argc, char** argv) {
    if (argc != 8) return 2;
    const std::string t = argv[1];
    const int n = std::atoi(argv[3]), d = std::atoi(argv[4]);
    const float eps = std::strtof(argv[5], nullptr);
    const size_t mp = std::strtoull(argv[6], nullptr, 10);
    try {
        if (t == "f32") return run<float>(argv[2], n, d, eps, mp, H5T_NATIVE_FLOAT, argv[7]);
        if (t == "f64") return run<double>(argv[2], n, d, eps, mp, H5T_NATIVE_DOUBLE, argv[7]);
        if (t == "i16") return run<int16_t>(argv[2], n, d, eps, mp, H5T_NATIVE_SHORT, argv[7]);
        if (t == "u8") return run<uint8_t>(argv[2], n, d, eps, mp, H5T_NATIVE_UCHAR, argv[7]);
    } catch (const std::exception& e) {
        std::fprintf(stderr, "%s\n", e.what());
        return 20;
    }
    return 3;
}
'''

CASES = [("blobs_f32_d3.npz", "f32"), ("blobs_f32_d2.npz", "f32"), ("blobs_f64_d3.npz", "f64"),
         ("edge_int_int16_eps3.npz", "i16"), ("edge_int_uint8_eps3.npz", "u8")]


@pytest.fixture(scope="module")
def driver(tmp_path_factory):
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ missing")
    if not os.path.exists(os.path.join(HDF5_INCLUDE, "hdf5.h")):
        pytest.skip("hdf5.h missing")
    td = tmp_path_factory.mktemp("cxx")
    cpp = td / "driver.cpp"
    cpp.write_text(DRIVER)
    exe = str(td / "driver")
    r = subprocess.run([gxx, "-std=c++11", "-O1", "-Wall", "-I" + INCLUDE, "-I" + HDF5_INCLUDE,
                        str(cpp), "-o", exe, "-L" + LIB_DIR, "-lhpdbscan_amd", "-L" + H5_LINKS,
                        "-lhdf5", "-Wl,-rpath," + LIB_DIR + ":" + H5_LINKS,
                        "-Wl,-rpath-link,/opt/rocm/lib"], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    return exe, td


@pytest.mark.parametrize("name,code", CASES)
def test_cxx_class_in_memory_and_dataset(driver, name, code):
    exe, td = driver
    z = np.load(os.path.join(GOLDEN, name))
    x, eps, mp = np.ascontiguousarray(z["data"]), float(z["eps"]), int(z["min_pts"])
    n, d = x.shape
    inp, out = str(td / f"{name}.bin"), str(td / f"{name}.labels")
    x.tofile(inp)
    print(f"[gpu-case] C++ class {name} n={n} d={d} {code}", flush=True)
    sys.stdout.flush()
    r = subprocess.run([exe, code, inp, str(n), str(d), repr(eps), str(mp), out],
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, (r.returncode, r.stdout, r.stderr)
    got = np.fromfile(out, dtype=np.int64)
    want = z["labels"]
    if not np.array_equal(got, want):
        mism, bad = oracle.check(x, eps, mp, got)
        assert mism == 0 and bad == 0, f"{mism} mismatches, {bad} bad ambiguous borders"
    np.testing.assert_array_equal(got < 0, want < 0)
    np.testing.assert_array_equal(got == 0, want == 0)
