"""Builds libhpdbscan_amd.so (the HIP kernels and the C-ABI) for gfx950.

    python -m hpdbscan_amd.build [--jobs N] [--verbose]

Objects go to hpdbscan_amd/_build/, the library next to this file so that it
travels with the source tree.  Where libhdf5 1.10
is installed (HDF5_PREFIX, default /opt/conda) the HDF5 file layer
libhpdbscan_amd_h5.so and the command-line tool bin/hpdbscan are built too;
libhdf5 is linked through _h5lib/, a directory of links to that one library
(and its zlib), so that no other library of that prefix is picked up.
Compiler flags that matter for the results:

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
INCLUDE = os.path.join(os.path.dirname(HERE), "include")
H5_LIB = os.path.join(HERE, "libhpdbscan_amd_h5.so")
H5_LINKS = os.path.join(HERE, "_h5lib")
CLI = os.path.join(HERE, "bin", "hpdbscan")
HDF5_PREFIX = os.environ.get("HDF5_PREFIX", "/opt/conda")
ROCM_LIB = "/opt/rocm/lib"

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
    cmd = [hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-pthread",
           "-Wl,-soname,libhpdbscan_amd.so", "-o", tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr}")
    os.replace(tmp, LIB)
    if hdf5_available():
        build_h5(verbose)
        build_cli(verbose)
    else:
        print(f"libhdf5 not found under {HDF5_PREFIX}: skipping {os.path.basename(H5_LIB)} "
              "and bin/hpdbscan", file=sys.stderr)
    return LIB


def hdf5_available() -> bool:
    return (os.path.exists(os.path.join(HDF5_PREFIX, "include", "hdf5.h"))
            and os.path.exists(os.path.join(HDF5_PREFIX, "lib", "libhdf5.so.103")))


def _run(cmd: list[str], what: str, verbose: bool) -> None:
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"{what} failed:\n{r.stderr}")


def _h5_links() -> None:
    os.makedirs(H5_LINKS, exist_ok=True)
    for name, target in (("libhdf5.so.103", "libhdf5.so.103"), ("libhdf5.so", "libhdf5.so.103"),
                         ("libz.so.1", "libz.so.1")):
        link = os.path.join(H5_LINKS, name)
        want = os.path.join(HDF5_PREFIX, "lib", target)
        if os.path.islink(link) and os.readlink(link) == want:
            continue
        if os.path.lexists(link):
            os.remove(link)
        os.symlink(want, link)


def build_h5(verbose: bool = False) -> str:
    """The HDF5 file layer (include/hpdbscan_h5.h) over libhpdbscan_amd.so."""
    _h5_links()
    tmp = H5_LIB + ".tmp"
    _run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-Wall", f"-I{INCLUDE}",
          f"-I{HDF5_PREFIX}/include", os.path.join(CSRC, "h5io.cpp"), "-o", tmp,
          f"-L{HERE}", "-lhpdbscan_amd", f"-L{H5_LINKS}", "-lhdf5",
          "-Wl,-soname,libhpdbscan_amd_h5.so", "-Wl,-rpath,$ORIGIN:$ORIGIN/_h5lib",
          f"-Wl,-rpath-link,{ROCM_LIB}"], "g++ on h5io.cpp", verbose)
    os.replace(tmp, H5_LIB)
    return H5_LIB


def build_cli(verbose: bool = False) -> str:
    """bin/hpdbscan, the command-line tool (csrc/cli.cpp)."""
    os.makedirs(os.path.dirname(CLI), exist_ok=True)
    tmp = CLI + ".tmp"
    _run(["g++", "-O2", "-std=c++17", "-Wall", f"-I{INCLUDE}", os.path.join(CSRC, "cli.cpp"),
          "-o", tmp, f"-L{HERE}", "-lhpdbscan_amd_h5", "-lhpdbscan_amd",
          "-Wl,-rpath,$ORIGIN/..", f"-Wl,-rpath-link,{ROCM_LIB}:{H5_LINKS}"],
         "g++ on cli.cpp", verbose)
    os.replace(tmp, CLI)
    return CLI


def main() -> None:
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--jobs", type=int, default=0)
    ap.add_argument("--verbose", action="store_true")
    a = ap.parse_args()
    print(build(a.jobs, a.verbose))


if __name__ == "__main__":
    main()
