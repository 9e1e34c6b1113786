"""Work-split autotuner for the neighbour kernels.

    python -m hpdbscan_amd.autotune [--workload blobs_10m] [--n N] [--reps 3]
                                    [--batch] [--out split.json]

The core count and the border pass run one 64-lane wavefront per 64-row
chunk of a dense cell and one thread per row of a sparse cell.  Where the
line between the two pays off depends on the data (density spread, D,
eps), so this module times candidate settings on the GPU for a given
workload and returns the fastest.  Every candidate must reproduce the first
candidate's labels exactly; a mismatch raises (the split never changes
results).  `apply` sets a tuned split on a Context.

The library's defaults (sparse_max 8, core count not split) are what this
tuner picks on the 10M x 3-D workload.
"""
from __future__ import annotations

import argparse
import json
import time

import numpy as np

from . import _lib, datagen

# (sparse_max, sparse_core)
CANDIDATES = [(0, False), (4, False), (8, False), (16, False), (32, False), (8, True), (16, True)]


def _time_runs(fn, sync, reps: int) -> float:
    fn()  # warm-up (workspace growth, code load)
    sync()
    times = []
    for _ in range(reps):
        t0 = time.perf_counter()
        fn()
        sync()
        times.append(time.perf_counter() - t0)
    return float(np.median(times))


def _sweep(ctx, run, fetch, reps: int, candidates) -> dict:
    results = []
    ref = None
    for sm, sc in candidates:
        ctx.set_work_split(sm, sc)
        sec = _time_runs(run, ctx.synchronize, reps)
        labels = fetch()
        if ref is None:
            ref = labels
        elif not np.array_equal(labels, ref):
            raise RuntimeError(f"work split ({sm}, {sc}) changed the labels")
        results.append({"sparse_max": sm, "sparse_core": sc, "ms": round(sec * 1e3, 3)})
    best = min(results, key=lambda r: r["ms"])
    return {"best": {"sparse_max": best["sparse_max"], "sparse_core": best["sparse_core"]},
            "results": results}


def tune(data: np.ndarray, eps: float, min_points: int, ctx=None, reps: int = 3,
         candidates=CANDIDATES) -> dict:
    """Fastest work split for clustering `data` (2-D array) once."""
    own = ctx is None
    ctx = ctx or _lib.Context(0)
    a = np.ascontiguousarray(data)
    n, d = a.shape
    xd = _lib.DeviceBuffer.from_array(ctx, a)
    ld = _lib.DeviceBuffer(ctx, 8 * max(n, 1))
    try:
        out = _sweep(ctx,
                     lambda: ctx.cluster_device(xd.ptr, n, d, a.dtype, eps, min_points, ld.ptr),
                     lambda: ld.to_array(np.int64, n), reps, candidates)
    finally:
        xd.close()
        ld.close()
        if own:
            ctx.close()
    out.update({"n": int(n), "d": int(d), "eps": float(eps), "min_points": int(min_points)})
    return out


def tune_batch(data: np.ndarray, offsets, eps, min_points, ctx=None, reps: int = 3,
               candidates=CANDIDATES) -> dict:
    """Fastest work split for one batched call over the jobs of `offsets`."""
    own = ctx is None
    ctx = ctx or _lib.Context(0)
    a = np.ascontiguousarray(data)
    n, d = a.shape
    xd = _lib.DeviceBuffer.from_array(ctx, a)
    ld = _lib.DeviceBuffer(ctx, 8 * max(n, 1))
    try:
        out = _sweep(ctx,
                     lambda: ctx.cluster_batch_device(xd.ptr, offsets, d, a.dtype, eps, min_points,
                                                      ld.ptr),
                     lambda: ld.to_array(np.int64, n), reps, candidates)
    finally:
        xd.close()
        ld.close()
        if own:
            ctx.close()
    out.update({"n": int(n), "d": int(d), "jobs": int(len(offsets) - 1)})
    return out


def apply(ctx, tuned: dict) -> None:
    """Set a tuned split (the dict `tune` returns, or its "best") on ctx."""
    best = tuned.get("best", tuned)
    ctx.set_work_split(int(best["sparse_max"]), bool(best["sparse_core"]))


def main() -> None:
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--workload", default="blobs_10m", choices=sorted(datagen.WORKLOADS))
    ap.add_argument("--n", type=int, default=0, help="override the row count")
    ap.add_argument("--batch", action="store_true",
                    help="tune the batched workload (4096 clouds) instead")
    ap.add_argument("--batch-jobs", type=int, default=4096)
    ap.add_argument("--reps", type=int, default=3)
    ap.add_argument("--out", default="")
    a = ap.parse_args()
    if a.batch:
        x, off, eps, mps = datagen.batch_clouds(n_jobs=a.batch_jobs)
        res = tune_batch(x, off, eps, mps, reps=a.reps)
        res["workload"] = "batch_clouds"
    else:
        x, eps, mp = datagen.make(a.workload, a.n or None)
        res = tune(x, eps, mp, reps=a.reps)
        res["workload"] = a.workload
    text = json.dumps(res, indent=1)
    if a.out:
        with open(a.out, "w") as f:
            f.write(text + "\n")
    print(text)


if __name__ == "__main__":
    main()
