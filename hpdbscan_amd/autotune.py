"""Autotuner of the pair kernels' tile and work split.

    python -m hpdbscan_amd.autotune [--workload blobs_10m] [--n N] [--reps 3]
                                    [--batch] [--out split.json]

Two settings change the speed of a run and never its labels:

* the box split (hpd_context_set_box_split): every eps-cell is cut into
  parts^d sub-boxes, the tiles whose bounding boxes decide whole groups of
  distance tests.  More parts give tighter boxes (more groups decided
  without row tests) at the price of more boxes to walk and more key bits
  to sort;
* the work split (hpd_context_set_work_split): the core count and the
  border pass run one 64-lane wavefront per 64-row chunk of a dense cell
  and one thread per row of a sparse cell; the line between the two.

Where each pays off depends on the data (density spread, D, eps), so this
module times candidates on the GPU for a given workload: first the box
splits under the library's work split, then the work splits under the
fastest box split.  Every candidate must reproduce the first candidate's
labels exactly; a mismatch raises.  `apply` sets a tuned result on a
Context.

The library's defaults (box split 2 up to 3-D, sparse_max 8, core count not
split) are what this tuner picks on the 10M x 3-D workload.
"""
from __future__ import annotations

import argparse
import json
import time

import numpy as np

from . import _lib, datagen

# work splits (sparse_max, sparse_core)
CANDIDATES = [(0, False), (4, False), (8, False), (16, False), (32, False), (8, True), (16, True)]
DEFAULT_WORK_SPLIT = (8, False)
# box splits: 0 is the library's choice for the dimension
BOX_SPLITS = [0, 2, 3, 4]


def default_box_split(d: int) -> int:
    return 2 if d <= 3 else 3


def box_split_candidates(d: int) -> list:
    """The box splits worth timing for dimension d (the default once)."""
    out = []
    for k in BOX_SPLITS:
        if k != 0 and k == default_box_split(d):
            continue
        if k ** d > 1 << 24:  # more sub-boxes than a key digit holds
            continue
        out.append(k)
    return out


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


def _set(ctx, cand: dict) -> None:
    if "box_split" in cand:
        ctx.set_box_split(int(cand["box_split"]))
    if "sparse_max" in cand:
        ctx.set_work_split(int(cand["sparse_max"]), bool(cand["sparse_core"]))


def _sweep(ctx, run, fetch, reps: int, candidates, ref=None) -> dict:
    """Time every candidate setting (dicts of box_split and / or sparse_max
    and sparse_core); all must give the labels `ref` (the first candidate's
    when None)."""
    results = []
    for cand in candidates:
        _set(ctx, cand)
        sec = _time_runs(run, ctx.synchronize, reps)
        labels = fetch()
        if ref is None:
            ref = labels
        elif not np.array_equal(labels, ref):
            raise RuntimeError(f"setting {cand} changed the labels")
        results.append({**cand, "ms": round(sec * 1e3, 3)})
    best = min(results, key=lambda r: r["ms"])
    return {"best": {k: v for k, v in best.items() if k != "ms"}, "results": results, "ref": ref}


def _two_stage(ctx, run, fetch, reps: int, d: int, candidates) -> dict:
    sm, sc = DEFAULT_WORK_SPLIT
    ctx.set_work_split(sm, sc)
    boxes = _sweep(ctx, run, fetch, reps, [{"box_split": k} for k in box_split_candidates(d)])
    k = boxes["best"]["box_split"]
    ctx.set_box_split(k)
    works = _sweep(ctx, run, fetch, reps,
                   [{"sparse_max": a, "sparse_core": b} for a, b in candidates], ref=boxes["ref"])
    best = {"box_split": k, **works["best"]}
    return {"best": best, "box_split_results": boxes["results"], "results": works["results"]}


def tune(data: np.ndarray, eps: float, min_points: int, ctx=None, reps: int = 3,
         candidates=CANDIDATES) -> dict:
    """Fastest box split and work split for clustering `data` (2-D array)
    once."""
    own = ctx is None
    ctx = ctx or _lib.Context(0)
    a = np.ascontiguousarray(data)
    n, d = a.shape
    xd = _lib.DeviceBuffer.from_array(ctx, a)
    ld = _lib.DeviceBuffer(ctx, 8 * max(n, 1))
    try:
        out = _two_stage(ctx,
                         lambda: ctx.cluster_device(xd.ptr, n, d, a.dtype, eps, min_points, ld.ptr),
                         lambda: ld.to_array(np.int64, n), reps, d, candidates)
    finally:
        xd.close()
        ld.close()
        if own:
            ctx.close()
    out.update({"n": int(n), "d": int(d), "eps": float(eps), "min_points": int(min_points)})
    return out


def tune_batch(data: np.ndarray, offsets, eps, min_points, ctx=None, reps: int = 3,
               candidates=CANDIDATES) -> dict:
    """Fastest box split and work split for one batched call over the jobs
    of `offsets`."""
    own = ctx is None
    ctx = ctx or _lib.Context(0)
    a = np.ascontiguousarray(data)
    n, d = a.shape
    xd = _lib.DeviceBuffer.from_array(ctx, a)
    ld = _lib.DeviceBuffer(ctx, 8 * max(n, 1))
    try:
        out = _two_stage(ctx,
                         lambda: ctx.cluster_batch_device(xd.ptr, offsets, d, a.dtype, eps,
                                                          min_points, ld.ptr),
                         lambda: ld.to_array(np.int64, n), reps, d, candidates)
    finally:
        xd.close()
        ld.close()
        if own:
            ctx.close()
    out.update({"n": int(n), "d": int(d), "jobs": int(len(offsets) - 1)})
    return out


def apply(ctx, tuned: dict) -> None:
    """Set a tuned result (the dict `tune` returns, or its "best") on ctx;
    a result without a box split leaves the context's box split alone."""
    _set(ctx, tuned.get("best", tuned))


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
