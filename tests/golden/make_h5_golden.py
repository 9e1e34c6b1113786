"""Writes the HDF5 fixtures of tests/golden/ with the reference's own
command-line tool.

    make -C oracle all && python -m hpdbscan_amd.build && python tests/golden/make_h5_golden.py

cli_blobs.h5 holds three input datasets of one point cloud (1 600 rows,
5 blobs + noise, 3-D):
  DATA      float32 coordinates
  DATA_F64  the same coordinates as float64
  DATA_I16  the coordinates x 100 rounded to int16
and, next to each, the labels that the reference's tool
(oracle/_ref/hpdbscan_ref_cli, src/cli.cpp built as is) wrote for it at one
thread: CLUSTERS, CLUSTERS_F64, CLUSTERS_I16.  The input datasets are written
with this repository's hpd_h5_write_matrix (the image has no h5py); the
reference's reader reading them is part of what the fixture pins.

cli_golden.json holds, per case, the arguments, the eps/minPts of the
labels datasets, the summary block that the reference's tool prints when
built WITH_OUTPUT, and what the tool prints (stdout, stderr) and returns on
failing command lines.
"""
from __future__ import annotations

import json
import os
import shutil
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from hpdbscan_amd import h5  # noqa: E402

REF_CLI = os.path.join(ROOT, "oracle", "_ref", "hpdbscan_ref_cli")
REF_CLI_OUT = os.path.join(ROOT, "oracle", "_ref", "hpdbscan_ref_cli_out")
H5_OUT = os.path.join(HERE, "cli_blobs.h5")
JSON_OUT = os.path.join(HERE, "cli_golden.json")

RUNS = [
    # input dataset, output dataset, eps, minPts
    ("DATA", "CLUSTERS", 0.15, 5),
    ("DATA_F64", "CLUSTERS_F64", 0.15, 5),
    ("DATA_I16", "CLUSTERS_I16", 15.0, 5),
]


def points() -> np.ndarray:
    rng = np.random.default_rng(5)
    centres = rng.uniform(0, 3, (5, 3))
    x = np.concatenate([rng.normal(c, 0.1, (300, 3)) for c in centres]
                       + [rng.uniform(0, 3, (100, 3))])
    rng.shuffle(x)
    return x.astype(np.float32)


def run(exe: str, args: list[str], cwd: str) -> dict:
    r = subprocess.run([exe, *args], capture_output=True, text=True, cwd=cwd, timeout=120)
    return {"returncode": r.returncode, "stdout": r.stdout, "stderr": r.stderr}


def summary_block(stdout: str) -> list[str]:
    lines = stdout.splitlines()
    i = lines.index("Summary...")
    return lines[i:i + 5]


def main() -> None:
    for exe in (REF_CLI, REF_CLI_OUT):
        if not os.path.exists(exe):
            raise SystemExit(f"{exe} is missing: run `make -C oracle all` first")
    x = points()
    golden = {"runs": [], "errors": []}
    with tempfile.TemporaryDirectory() as tmp:
        path = os.path.join(tmp, "cli_blobs.h5")
        h5.write_matrix(path, "DATA", x)
        h5.write_matrix(path, "DATA_F64", x.astype(np.float64))
        h5.write_matrix(path, "DATA_I16", np.rint(x * 100).astype(np.int16))
        for ds_in, ds_out, eps, mp in RUNS:
            args = ["-i", path, "-o", path, "--input-dataset", ds_in, "--output-dataset", ds_out,
                    "-e", repr(eps), "-m", str(mp), "-t", "1"]
            r = run(REF_CLI, args, tmp)
            assert r["returncode"] == 0 and not r["stderr"], r
            # the WITH_OUTPUT build writes into a scratch copy
            scratch = os.path.join(tmp, "scratch.h5")
            shutil.copy(path, scratch)
            out_args = ["-i", scratch, "-o", scratch, "--input-dataset", ds_in,
                        "--output-dataset", ds_out + "_X", "-e", repr(eps), "-m", str(mp),
                        "-t", "1"]
            ro = run(REF_CLI_OUT, out_args, tmp)
            os.remove(scratch)
            golden["runs"].append({"input_dataset": ds_in, "output_dataset": ds_out, "eps": eps,
                                   "min_points": mp, "summary": summary_block(ro["stdout"])})
        # a 1-D dataset where a matrix is expected
        h5.write_hdf5(path, "VECTOR", np.arange(4))
        bad = [
            ["-i", path, "-o", path, "-e", "-0.5", "-m", "4"],
            ["-i", path, "-o", path, "-e", "0", "-m", "4"],
            ["-i", os.path.join(tmp, "missing.h5"), "-o", path],
            ["-i", path, "-o", path, "--input-dataset", "NOPE"],
            ["-i", path, "-o", path, "--input-dataset", "VECTOR"],
            ["-i", path, "-o", os.path.join(tmp, "no_dir", "out.h5")],
            ["--bogus"],
            ["-m", "abc", "-i", path],
        ]
        for args in bad:
            r = run(REF_CLI, args, tmp)
            golden["errors"].append({"args": [a.replace(tmp, "{tmp}") for a in args],
                                     "returncode": r["returncode"],
                                     "stdout": r["stdout"].replace(tmp, "{tmp}"),
                                     "stderr": r["stderr"].replace(tmp, "{tmp}")})
        shutil.copy(path, H5_OUT)
    with open(JSON_OUT, "w") as f:
        json.dump(golden, f, indent=1)
        f.write("\n")
    print(f"wrote {H5_OUT} and {JSON_OUT}")


if __name__ == "__main__":
    main()
