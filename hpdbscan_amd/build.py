"""Builds libhpdbscan_amd.so (the HIP kernels and the C-ABI) for gfx950.

    python -m hpdbscan_amd.build [--jobs N] [--verbose]

Objects go to hpdbscan_amd/_build/, the library next to this file so that it
travels with the source tree, and so does `copy_probe` (a plain copy kernel
of known size that calibrates the HBM traffic counters).  Compiler flags that matter for the results:

  -ffp-contract=off                no fused multiply-add: the squared distance
                                   is evaluated with separate roundings, as
                                   the reference's SSE2 build does
  -fdenormal-fp-math=preserve-sign (device) f32 and f64 denormals flush to
                                   zero on input and output, as under the
                                   reference's FTZ/DAZ floating-point mode
  -target-feature -packed-fp32-ops no packed FP32 instructions
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libhpdbscan_amd.so")
COPY_PROBE = os.path.join(HERE, "copy_probe")
INCLUDE = os.path.join(os.path.dirname(HERE), "include")

SOURCES = ["capi.hip", "scan.hip", "radix_sort.hip", "grid.hip", "index.hip", "dbscan.hip",
           "boxes.hip", "server.cpp"]

ARCH = os.environ.get("HPD_OFFLOAD_ARCH", "gfx950")


def hipcc() -> str:
    cand = os.environ.get("HIPCC") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(cand):
        raise RuntimeError(f"hipcc not found at {cand}")
    return cand


def compile_flags() -> list[str]:
    return [
        f"--offload-arch={ARCH}",
        "-O3",
        "-std=c++17",
        "-fPIC",
        "-ffp-contract=off",
        "-Xarch_device", "-fdenormal-fp-math=preserve-sign",
        "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
        "-Wall",
        "-Wno-unused-function",
        f"-I{INCLUDE}",
    ]


def _compile(src: str, verbose: bool) -> str:
    obj = os.path.join(BUILD, os.path.splitext(src)[0] + ".o")
    lang = ["-x", "hip"] if src.endswith(".hip") else []
    flags = compile_flags() if src.endswith(".hip") else [
        "-O3", "-std=c++17", "-fPIC", "-Wall", f"-I{INCLUDE}", "-pthread"]
    cmd = [hipcc(), *flags, *lang, "-c", os.path.join(CSRC, src), "-o", obj]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    # the host pass does not know the AMDGPU target feature; that warning is noise
    err = "\n".join(l for l in r.stderr.splitlines() if "packed-fp32-ops" not in l)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed on {src}:\n{err}")
    if verbose and err.strip():
        print(err, file=sys.stderr)
    return obj


def build(jobs: int = 0, verbose: bool = False) -> str:
    """Compile every source and link the shared library; returns its path."""
    os.makedirs(BUILD, exist_ok=True)
    jobs = jobs or min(len(SOURCES), max(1, (os.cpu_count() or 2)))
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(lambda s: _compile(s, verbose), SOURCES))
    tmp = LIB + ".tmp"
    cmd = [hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-pthread", "-o", tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr}")
    os.replace(tmp, LIB)
    build_copy_probe(verbose)
    return LIB


def build_copy_probe(verbose: bool = False) -> str:
    """The copy kernel of known size that calibrates the HBM traffic counters."""
    tmp = COPY_PROBE + ".tmp"
    cmd = [hipcc(), *compile_flags(), "-x", "hip", os.path.join(CSRC, "copy_probe.cpp"), "-o", tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed on copy_probe.cpp:\n{r.stderr}")
    os.replace(tmp, COPY_PROBE)
    return COPY_PROBE


def main() -> None:
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--jobs", type=int, default=0)
    ap.add_argument("--verbose", action="store_true")
    a = ap.parse_args()
    print(build(a.jobs, a.verbose))


if __name__ == "__main__":
    main()
