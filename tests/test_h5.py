"""The HDF5 file layer and the command-line tool.

CPU tests (no GPU is touched): the library loads and exports what
include/hpdbscan_h5.h declares; datasets of all ten element types round-trip;
the reference's error messages and its write semantics (create the file if
missing, write an existing dataset in place); the fixture written by the
reference's own tool (tests/golden/cli_blobs.h5, made by
make_h5_golden.py) reads back and matches the oracle; and every failing
command line whose failure comes before any clustering prints what the
reference's tool prints (tests/golden/cli_golden.json).

GPU tests: the tool, the Python class and the C++ class cluster the
fixture's datasets to the reference tool's labels, bit for bit, and print
its summary block.
"""
from __future__ import annotations

import json
import os
import shutil
import subprocess

import numpy as np
import pytest

from conftest import GOLDEN, ROOT
from test_capi import declared_functions

H5_FIXTURE = os.path.join(GOLDEN, "cli_blobs.h5")
CLI_GOLDEN = os.path.join(GOLDEN, "cli_golden.json")
CLI = os.path.join(ROOT, "hpdbscan_amd", "bin", "hpdbscan")
INCLUDE = os.path.join(ROOT, "include")

DTYPES = [np.float32, np.float64, np.int8, np.int16, np.int32, np.int64, np.uint8, np.uint16,
          np.uint32, np.uint64]


def golden():
    with open(CLI_GOLDEN) as f:
        return json.load(f)


@pytest.fixture
def h5():
    from hpdbscan_amd import h5 as mod
    mod.load()
    return mod


@pytest.fixture
def fixture_copy(tmp_path):
    p = tmp_path / "cli_blobs.h5"
    shutil.copy(H5_FIXTURE, p)
    return str(p)


def run_cli(args, cwd):
    assert os.path.exists(CLI), "hpdbscan_amd/bin/hpdbscan is not built"
    return subprocess.run([CLI, *args], capture_output=True, text=True, cwd=cwd, timeout=300)


# --------------------------------------------------------------- library (CPU)

def test_h5_library_exports_every_declared_symbol(h5):
    names = declared_functions({"hpdbscan_h5.h"})
    assert "hpd_cluster_h5" in names and "hpd_h5_write_clusters" in names
    lib = h5.load()
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


@pytest.mark.parametrize("dt", DTYPES, ids=lambda d: np.dtype(d).name)
def test_matrix_round_trip(h5, tmp_path, dt):
    path = str(tmp_path / "m.h5")
    info = np.iinfo(dt) if np.issubdtype(dt, np.integer) else np.finfo(dt)
    a = np.array([[0, 1, 2], [info.max, info.min, 7], [3, 4, 5], [1, 1, 1]], dtype=dt)
    h5.write_matrix(path, "DATA", a)
    rows, cols, stored = h5.matrix_info(path, "DATA")
    assert (rows, cols, stored) == (4, 3, np.dtype(dt))
    b = h5.read_hdf5(path, "DATA")
    assert b.dtype == np.dtype(dt) and np.array_equal(a, b)


def test_read_converts_to_a_requested_type(h5, tmp_path):
    path = str(tmp_path / "m.h5")
    a = np.arange(-6, 6, dtype=np.int16).reshape(4, 3)
    h5.write_matrix(path, "DATA", a)
    assert np.array_equal(h5.read_hdf5(path, "DATA", np.float64), a.astype(np.float64))


def test_empty_matrix(h5, tmp_path):
    path = str(tmp_path / "m.h5")
    h5.write_matrix(path, "DATA", np.zeros((0, 3), np.float32))
    assert h5.read_hdf5(path, "DATA").shape == (0, 3)


def test_read_errors_match_the_reference(h5, tmp_path):
    # io.h:49-66 and its exception kinds (std::invalid_argument -> ValueError)
    path = str(tmp_path / "m.h5")
    h5.write_matrix(path, "DATA", np.ones((3, 2), np.float32))
    h5.write_hdf5(path, "VEC", [1, 2, 3])
    missing = str(tmp_path / "missing.h5")
    with pytest.raises(ValueError, match=f"^Could not open {missing}$"):
        h5.read_hdf5(missing, "DATA")
    with pytest.raises(ValueError, match="^Could not open dataset NOPE$"):
        h5.read_hdf5(path, "NOPE")
    with pytest.raises(ValueError, match="^Data in dataset VEC should be a matrix$"):
        h5.read_hdf5(path, "VEC")
    not_h5 = tmp_path / "text.h5"
    not_h5.write_text("not an HDF5 file")
    with pytest.raises(ValueError, match="^Could not open"):
        h5.read_hdf5(str(not_h5), "DATA")


def test_write_creates_the_file_and_keeps_other_datasets(h5, tmp_path):
    path = str(tmp_path / "new.h5")
    h5.write_hdf5(path, "CLUSTERS", [0, -1, 2])
    assert np.array_equal(h5.read_clusters(path, "CLUSTERS"), [0, -1, 2])
    h5.write_matrix(path, "DATA", np.ones((3, 2), np.float64))
    h5.write_hdf5(path, "OTHER", [5, 6, 7])
    assert np.array_equal(h5.read_clusters(path, "CLUSTERS"), [0, -1, 2])
    assert np.array_equal(h5.read_hdf5(path, "DATA"), np.ones((3, 2)))


def test_write_into_an_existing_dataset(h5, tmp_path):
    # io.h:170-190: an existing dataset is opened and written in place; a
    # shorter label vector overwrites its head, a longer one fails
    path = str(tmp_path / "m.h5")
    h5.write_hdf5(path, "CLUSTERS", [1, 2, 3, 4])
    h5.write_hdf5(path, "CLUSTERS", [-9, -8, -7, -6])
    assert np.array_equal(h5.read_clusters(path, "CLUSTERS"), [-9, -8, -7, -6])
    h5.write_hdf5(path, "CLUSTERS", [5, 5])
    assert np.array_equal(h5.read_clusters(path, "CLUSTERS"), [5, 5, -7, -6])
    with pytest.raises(RuntimeError, match="^Failed to write data, target data set is probably "
                                           "of wrong size or type$"):
        h5.write_hdf5(path, "CLUSTERS", [1, 2, 3, 4, 5])
    with pytest.raises(ValueError, match="^Could not open "):
        h5.write_hdf5(str(tmp_path / "no_dir" / "x.h5"), "CLUSTERS", [1])


def test_labels_are_native_longs(h5, tmp_path):
    path = str(tmp_path / "m.h5")
    big = [np.iinfo(np.int64).min + 1, np.iinfo(np.int64).max, 0]
    h5.write_hdf5(path, "CLUSTERS", big)
    assert h5.read_clusters(path, "CLUSTERS").tolist() == big


def test_matrix_dataset_exists(h5, tmp_path):
    path = str(tmp_path / "m.h5")
    h5.write_matrix(path, "DATA", np.ones((2, 2), np.float32))
    with pytest.raises(ValueError, match="exists already"):
        h5.write_matrix(path, "DATA", np.ones((2, 2), np.float32))


# ------------------------------------------------- reference fixture (CPU)

def test_fixture_holds_the_reference_tool_labels(h5):
    """The labels the reference's tool wrote into the fixture are the
    oracle's labels of the same datasets (one thread), so the fixture, the
    reader and the oracle agree."""
    import oracle
    for run in golden()["runs"]:
        x = h5.read_hdf5(H5_FIXTURE, run["input_dataset"])
        want = h5.read_clusters(H5_FIXTURE, run["output_dataset"])
        got = oracle.cluster(x, run["eps"], run["min_points"])
        assert np.array_equal(got, want), run["input_dataset"]


def test_fixture_dtypes(h5):
    assert h5.matrix_info(H5_FIXTURE, "DATA")[2] == np.float32
    assert h5.matrix_info(H5_FIXTURE, "DATA_F64")[2] == np.float64
    assert h5.matrix_info(H5_FIXTURE, "DATA_I16")[2] == np.int16


def test_reference_tool_reads_our_files(h5, tmp_path):
    """Where the reference's tool is built (oracle/_ref), it reads a file this
    library wrote and writes labels this library reads back."""
    import oracle
    ref_cli = os.path.join(ROOT, "oracle", "_ref", "hpdbscan_ref_cli")
    if not os.path.exists(ref_cli):
        pytest.skip("oracle/_ref/hpdbscan_ref_cli not built")
    rng = np.random.default_rng(11)
    x = np.concatenate([rng.normal(c, 0.2, (200, 2)) for c in rng.uniform(0, 5, (4, 2))])
    x = x.astype(np.float64)
    path = str(tmp_path / "ours.h5")
    h5.write_matrix(path, "POINTS", x)
    r = subprocess.run([ref_cli, "-i", path, "-o", path, "--input-dataset", "POINTS", "-e", "0.25",
                        "-m", "6", "-t", "1"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0 and not r.stderr, r.stderr
    labels = h5.read_clusters(path, "CLUSTERS")
    assert np.array_equal(labels, oracle.cluster(x, 0.25, 6))


def summary_lines(labels) -> list[str]:
    """The reference's summary block (hpdbscan.h:213-284) of a label vector,
    counted with numpy as a check of the tool's own counting."""
    labels = np.asarray(labels)
    noise = int((labels == 0).sum())
    uniq = np.unique(np.abs(labels)).size
    return ["Summary...", f"\tClusters:       {uniq - 1 if noise else uniq}",
            f"\tCluster points: {int((labels != 0).sum())}", f"\tNoise points:   {noise}",
            f"\tCore points:    {int((labels < 0).sum())}"]


def test_fixture_summaries(h5):
    for run in golden()["runs"]:
        labels = h5.read_clusters(H5_FIXTURE, run["output_dataset"])
        assert summary_lines(labels) == run["summary"]


# ------------------------------------------------ command line (CPU part)

def _expand(args, tmp):
    return [a.replace("{tmp}", tmp) for a in args]


def test_cli_help(tmp_path):
    r = run_cli(["--help"], str(tmp_path))
    assert r.returncode == 0
    for opt in ("--minPoints", "--epsilon", "--threads", "--input", "--output",
                "--input-dataset", "--output-dataset"):
        assert opt in r.stdout


def test_cli_errors_before_clustering_match_the_reference(tmp_path, fixture_copy, h5):
    """Every failing command line of the fixture whose failure comes before
    any clustering: same stdout, stderr and exit status as the reference's
    tool (the failure after clustering is checked on the GPU)."""
    tmp = str(tmp_path)
    h5.write_hdf5(fixture_copy, "VECTOR", np.arange(4))
    checked = 0
    for case in golden()["errors"]:
        if any("no_dir" in a for a in case["args"]):
            continue
        r = run_cli(_expand(case["args"], tmp), tmp)
        assert (r.returncode, r.stdout, r.stderr.replace(tmp, "{tmp}")) == \
            (case["returncode"], case["stdout"], case["stderr"]), case["args"]
        checked += 1
    assert checked == len(golden()["errors"]) - 1


# ------------------------------------------------------------------ GPU parity

@pytest.mark.gpu
@pytest.mark.parametrize("run", golden()["runs"], ids=lambda r: r["input_dataset"])
def test_cli_matches_the_reference_tool(tmp_path, fixture_copy, h5, run):
    out = str(tmp_path / "out.h5")
    r = run_cli(["-i", fixture_copy, "-o", out, "--input-dataset", run["input_dataset"],
                 "-e", repr(run["eps"]), "-m", str(run["min_points"]), "--summary"],
                str(tmp_path))
    assert r.returncode == 0 and not r.stderr, r.stderr
    got = h5.read_clusters(out, "CLUSTERS")
    want = h5.read_clusters(H5_FIXTURE, run["output_dataset"])
    assert np.array_equal(got, want)
    lines = r.stdout.splitlines()
    i = lines.index("Summary...")
    assert lines[i:i + 5] == run["summary"]
    assert lines[i + 5].startswith("Total time: ")


@pytest.mark.gpu
def test_cli_writes_into_the_input_file_by_name(tmp_path, fixture_copy, h5):
    run = golden()["runs"][0]
    r = run_cli(["--input=" + fixture_copy, "--output=" + fixture_copy,
                 "--output-dataset=MINE", "-e" + repr(run["eps"]), "-m" + str(run["min_points"])],
                str(tmp_path))
    assert r.returncode == 0 and not r.stderr and not r.stdout, (r.stdout, r.stderr)
    assert np.array_equal(h5.read_clusters(fixture_copy, "MINE"),
                          h5.read_clusters(H5_FIXTURE, run["output_dataset"]))
    # the other datasets of the file are untouched
    assert np.array_equal(h5.read_hdf5(fixture_copy, "DATA"), h5.read_hdf5(H5_FIXTURE, "DATA"))


@pytest.mark.gpu
def test_cli_defaults_write_data_h5(tmp_path, h5):
    # -i and -o both default to data.h5 in the working directory (cli.cpp:49-50)
    run = golden()["runs"][0]
    shutil.copy(H5_FIXTURE, tmp_path / "data.h5")
    r = run_cli(["-e", repr(run["eps"]), "-m", str(run["min_points"])], str(tmp_path))
    assert r.returncode == 0 and not r.stderr, r.stderr
    assert np.array_equal(h5.read_clusters(str(tmp_path / "data.h5"), "CLUSTERS"),
                          h5.read_clusters(H5_FIXTURE, "CLUSTERS"))


@pytest.mark.gpu
def test_cli_output_error_matches_the_reference(tmp_path, fixture_copy):
    tmp = str(tmp_path)
    case = next(c for c in golden()["errors"] if any("no_dir" in a for a in c["args"]))
    r = run_cli(_expand(case["args"], tmp), tmp)
    assert (r.returncode, r.stdout, r.stderr.replace(tmp, "{tmp}")) == \
        (case["returncode"], case["stdout"], case["stderr"])


@pytest.mark.gpu
@pytest.mark.parametrize("run", golden()["runs"], ids=lambda r: r["input_dataset"])
def test_python_cluster_path(run, h5):
    from hpdbscan_amd import HPDBSCAN
    labels = HPDBSCAN(run["eps"], run["min_points"]).cluster(H5_FIXTURE, run["input_dataset"], 4)
    assert isinstance(labels, tuple)
    assert np.array_equal(np.array(labels), h5.read_clusters(H5_FIXTURE, run["output_dataset"]))


@pytest.mark.gpu
def test_python_cluster_path_errors():
    from hpdbscan_amd import HPDBSCAN
    h = HPDBSCAN(0.15, 5)
    with pytest.raises(ValueError, match="^Could not open dataset NOPE$"):
        h.cluster(H5_FIXTURE, "NOPE")
    with pytest.raises(TypeError):
        h.cluster(H5_FIXTURE, "DATA", 1.5)


@pytest.mark.gpu
def test_cxx_class_cluster_path(tmp_path, h5):
    """HPDBSCAN::cluster(path, dataset[, threads]) of include/hpdbscan.h and
    write_hdf5, compiled as C++11 against both libraries."""
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ missing")
    src = r'''
    #include "hpdbscan.h"
    #include <cstdio>
    int main(int argc, char** argv) {
        HPDBSCAN h(0.15f, 5);
        Clusters c = h.cluster(argv[1], "DATA_F64");
        Clusters d = h.cluster(argv[1], "DATA", 2);
        if (c != d) return 3;
        try { h.cluster(argv[1], "NOPE"); return 4; } catch (const std::invalid_argument& e) {
            if (std::string(e.what()) != "Could not open dataset NOPE") return 5; }
        hpdbscan_amd::write_hdf5(argv[2], "CLUSTERS", c);
        return 0;
    }'''
    cpp = tmp_path / "t.cpp"
    cpp.write_text(src)
    exe = str(tmp_path / "t")
    lib_dir = os.path.join(ROOT, "hpdbscan_amd")
    r = subprocess.run([gxx, "-std=c++11", "-I" + INCLUDE, str(cpp), "-o", exe, "-L" + lib_dir,
                        "-lhpdbscan_amd_h5", "-lhpdbscan_amd", "-Wl,-rpath," + lib_dir,
                        "-Wl,-rpath-link,/opt/rocm/lib:" + os.path.join(lib_dir, "_h5lib")],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    out = str(tmp_path / "out.h5")
    r = subprocess.run([exe, H5_FIXTURE, out], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, (r.returncode, r.stdout, r.stderr)
    assert np.array_equal(h5.read_clusters(out, "CLUSTERS"),
                          h5.read_clusters(H5_FIXTURE, "CLUSTERS_F64"))
