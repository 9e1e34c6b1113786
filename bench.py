"""Benchmark: Mpoints/s of the HIP clustering path on the 10M x 3-D workload.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload blobs_10m]
                    [--no-cpu-baseline] [--cpu-threads T]

One step = one full clustering of the workload's rows (grid, radix sort,
neighbour ranges, core counts, union-find, border labels, labels written in
input order) with the rows already resident in device memory and the labels
left there (hpd_cluster_device).  The timed region is bracketed by stream
synchronisation; value = rows * steps / seconds / 1e6.

The JSON line also carries
  roofline      the longest neighbour-pair kernel: its compulsory HBM bytes
                over its duration (HIP events the library records on its own
                stream during the timed steps) against 8 TB/s, and the HBM
                traffic rocprofv3 counted for it (profiles/)
  cpu_baseline  the reference (oracle/_ref, compiled from the reference's own
                headers with its own flags) timed on this host on a bounded
                sample: the rows of the same workload inside a slab
                x0 < X, which keeps the density (and so the per-row work) of
                the full workload
  parity        GPU labels vs the reference's labels on that sample
  batched       the batched workload (4096 clouds, 1k-200k rows each) in one
                hpd_cluster_batch_device call vs a loop of one call per
                cloud: jobs/s of both, per-call p50/p99 of the loop, and
                equality of the labels; the work split of both comes from
                hpdbscan_amd.autotune run on the batch first; graph_* is the
                same batch through the static pipeline captured once in a
                hipGraph and replayed (hpd_context_set_graph)
  serving       a Poisson stream (200 requests/s, 10k-5M rows, d in 2..4)
                through the continuous-batching server: offered vs achieved
                load, p50/p99 latency; the same stream device-resident
                (hpd_server_submit_device) and at 2x and 4x its rate
Only one GPU is used; --gpus above 1 exits with a message.
"""
from __future__ import annotations

import argparse
import json
import os
import subprocess
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from hpdbscan_amd import _lib, autotune, datagen  # noqa: E402

METRIC = "Mpoints/s at 10M x 3-D (bit-exact labels); batched jobs/s + p99 vs plain"
PEAK_HBM_GBS = 8000.0  # MI355X HBM3E


# The kernels that compare rows with their neighbours, keyed by the
# per-kernel time the library reports (HIP events on its stream).
PAIR_KERNELS = {
    "ms_core_kernel": "k_core_boxes",
    "ms_hook_kernel": "k_box_hook+k_box_hook_small",
    "ms_union_kernel": "k_box_scan+k_box_union_list+k_box_union",
    "ms_border_kernel": "k_border_boxes+k_border_sparse",
}


def compulsory_bytes(kernel: str, st: dict, n: int, d: int, s: int) -> int:
    """Bytes a kernel must move at least once: every array it reads, once,
    and every array it writes, once (n rows, D dims, s-byte values, nb boxes,
    m occupied cells, R neighbour rows).

      k_core_boxes    rows n*D*s, box of row n*4, core flag written n;
                      per box first row 4, bounding box 2*D*s, flags 1;
                      per cell first row, ranges offset, own row, job (16);
                      per neighbour row its cells (8)
      k_box_hook*     (the hook forest) rows n*D*s, core flag n; per box
                      first row, cell, smallest core, parent written (16) +
                      bounding box; per cell ranges offset, first box, job
                      (12); 8 per row
      k_box_scan*     (the open-pair union) as the hook forest plus the
                      hook roots (4 per box)
      k_border_*      rows n*D*s, core flag n, input row 4n, box of row 4n,
                      labels written 8 per non-core row; per box first row,
                      cluster, cell (12) + bounding box; per cell 16; 8 per
                      neighbour row
    """
    nb, m, r = int(st["boxes"]), int(st["occupied_cells"]), int(st["neighbor_rows"])
    noncore = n - int(st["core_points"])
    if kernel == "k_core_boxes":
        return n * (d * s + 5) + nb * (2 * d * s + 5) + m * 16 + r * 8
    if kernel.startswith("k_box_hook"):
        return n * (d * s + 1) + nb * (2 * d * s + 16) + m * 12 + r * 8
    if kernel.startswith("k_box_scan"):
        return n * (d * s + 1) + nb * (2 * d * s + 20) + m * 12 + r * 8
    return n * (d * s + 9) + 8 * noncore + nb * (2 * d * s + 12) + m * 16 + r * 8


def measured_traffic(workload: str, kernel: str):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 counter
    summary of this workload (FETCH_SIZE + WRITE_SIZE, calibrated against a
    copy kernel of known size), or None."""
    path = os.path.join(ROOT, "profiles", f"hbm_traffic_{workload}.json")
    try:
        with open(path) as f:
            t = json.load(f)
    except (OSError, ValueError):
        return None
    # entries are keyed by the kernel's name with its template arguments;
    # "a+b" sums the kernels a and b
    want = set(kernel.split("+"))
    hits = [int(v["bytes_per_launch"]) for k, v in t.get("kernels", {}).items()
            if k.split("<")[0] in want]
    return sum(hits) if hits else None


def kernel_roofline(st: dict, med: dict, n: int, d: int, s: int, workload: str) -> dict:
    """Roofline of the longest neighbour-pair kernel: compulsory HBM bytes
    over its event-timed duration, against MI355X's 8 TB/s."""
    dom = max(PAIR_KERNELS, key=lambda k: med[k])
    kname = PAIR_KERNELS[dom]
    ms = med[dom]
    nbytes = compulsory_bytes(kname, st, n, d, s)
    achieved = nbytes / (ms * 1e-3) / 1e9 if ms > 0 else 0.0
    return {"bound": "hbm", "kernel": kname, "achieved": round(achieved, 1),
            "peak": PEAK_HBM_GBS, "unit": "GB/s", "frac": round(achieved / PEAK_HBM_GBS, 5),
            "traffic": measured_traffic(workload, kname), "algorithmic_bytes": nbytes,
            "kernel_ms": round(ms, 3),
            "note": "compulsory bytes (each array read once, each output written once); the "
                    "kernel's time goes to dependent loads and issue, not to bandwidth"}


# MI355X FP32 vector rate: 157.3 TFLOP/s counting an FMA as 2, so 78.6 T
# lane-ops/s of plain adds / multiplies / compares
PEAK_VALU_LANE_OPS = 78.6e12


def core_test_figures(st: dict, med: dict, d: int) -> dict:
    """The row-row distance tests the core count kernel executed (the rest
    of the reference's candidate tests are decided by bounding boxes), and
    their VALU rate: 3D lane-ops per test (D subtracts, D multiplies, D - 1
    adds, 1 compare) over the kernel's event-timed duration."""
    tests = int(st["core_tests"])
    ms = med["ms_core_kernel"]
    ops = tests * 3 * d
    rate = ops / (ms * 1e-3) if ms > 0 else 0.0
    return {"executed": tests,
            "fraction_of_sum_candidates": round(tests / max(1, int(st["sum_candidates"])), 6),
            "lane_ops_per_s": round(rate, 1), "peak_lane_ops_per_s": PEAK_VALU_LANE_OPS,
            "valu_frac": round(rate / PEAK_VALU_LANE_OPS, 6), "kernel_ms": round(ms, 3)}


def cpu_threads() -> int:
    """Every host CPU this process may run on."""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, n)


def cpu_model() -> str:
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def host_info() -> dict:
    """nproc, the CPUs this process may use, the CPU model and SMT state."""
    smt = None
    try:
        smt = open("/sys/devices/system/cpu/smt/active").read().strip() == "1"
    except OSError:
        pass
    return {"nproc": os.cpu_count(), "usable_cpus": cpu_threads(), "cpu": cpu_model(),
            "smt_active": smt}


def slab_sample(x: np.ndarray, target: int) -> np.ndarray:
    """Rows with x[:, 0] below the quantile that keeps about `target` rows."""
    if target >= len(x):
        return x
    cut = np.partition(x[:, 0], target)[target]
    return np.ascontiguousarray(x[x[:, 0] < cut])


def run_reference(sample: np.ndarray, eps: float, mp: int, threads: int, reps: int):
    exe = os.path.join(ROOT, "oracle", "_ref", "hpdbscan_ref_run")
    if not os.path.exists(exe):
        return None, None, "oracle/_ref/hpdbscan_ref_run not built"
    with tempfile.TemporaryDirectory(dir=os.environ.get("TMPDIR", "/tmp")) as td:
        inp = os.path.join(td, "in.bin")
        out = os.path.join(td, "labels.bin")
        sample.astype(np.float32).tofile(inp)
        cmd = [exe, "f32", inp, str(len(sample)), str(sample.shape[1]), repr(float(eps)), str(mp),
               str(threads), str(reps), out]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            return None, None, r.stderr.strip()[-300:]
        res = json.loads(r.stdout.strip().splitlines()[-1])
        labels = np.fromfile(out, dtype=np.int64)
    return res, labels, None


def batched_leg(a) -> dict:
    """All clouds of the batched workload in one call vs one call per cloud."""
    t0 = time.time()
    rows, off, eps, mps = datagen.batch_clouds(n_jobs=a.batch_jobs)
    nj = len(off) - 1
    print(f"[bench] batched: {nj} clouds, {len(rows)} rows (generated in {time.time() - t0:.1f}s)",
          file=sys.stderr, flush=True)
    ctx = _lib.Context(0)
    # the work split comes from the autotuner, run on this batch (every
    # candidate must give the same labels); the plain loop uses it too
    t0 = time.time()
    tuned = autotune.tune_batch(rows, off, eps, mps, ctx=ctx, reps=1) if a.autotune else None
    if tuned:
        autotune.apply(ctx, tuned)
        print(f"[bench] batched: work split {tuned['best']} (autotuned in {time.time() - t0:.1f}s)",
              file=sys.stderr, flush=True)
    xd = _lib.DeviceBuffer.from_array(ctx, rows)
    ld = _lib.DeviceBuffer(ctx, 8 * len(rows))
    # warm up both paths (allocations)
    ctx.cluster_batch_device(xd.ptr, off, 3, np.float32, eps, mps, ld.ptr)
    ctx.synchronize()
    reps = []
    for _ in range(a.batch_reps):
        t = time.perf_counter()
        ctx.cluster_batch_device(xd.ptr, off, 3, np.float32, eps, mps, ld.ptr)
        ctx.synchronize()
        reps.append(time.perf_counter() - t)
    batched_labels = ld.to_array(np.int64, len(rows))
    t_batch = float(np.median(reps))
    # the same batch through the static pipeline: captured in a hipGraph by
    # the first call, replayed by the others (no host read-back in between)
    graph = {}
    if a.graph:
        gctx = _lib.Context(0)
        gctx.set_work_split(*ctx.work_split())
        gctx.set_box_split(ctx.box_split())
        gctx.set_graph(True)
        gl = _lib.DeviceBuffer(gctx, 8 * len(rows))
        t = time.perf_counter()
        gctx.cluster_batch_device(xd.ptr, off, 3, np.float32, eps, mps, gl.ptr)
        gctx.synchronize()
        t_capture = time.perf_counter() - t
        greps = []
        for _ in range(a.batch_reps):
            t = time.perf_counter()
            gctx.cluster_batch_device(xd.ptr, off, 3, np.float32, eps, mps, gl.ptr)
            gctx.synchronize()
            greps.append(time.perf_counter() - t)
        g_labels = gl.to_array(np.int64, len(rows))
        t_graph = float(np.median(greps))
        small = _small_batch(ctx, gctx, rows, off, eps, mps, a.batch_reps)
        graph = {"graph_jobs_per_s": round(nj / t_graph, 1), "graph_s": round(t_graph, 4),
                 "graph_first_call_s": round(t_capture, 4),
                 "graph_labels_equal": bool(np.array_equal(g_labels, batched_labels)),
                 "graph_stats": gctx.graph_stats(),
                 "graph_small": small,
                 "graph_note": "static pipeline (worst-case tables, device-side counts and grids, "
                               "no host read-back) captured in a hipGraph by the first call and "
                               "replayed"}
        gl.close()
        gctx.close()
        del g_labels
    # the plain loop: one hpd_cluster_device per cloud, same device buffers
    lat = []
    for j in range(nj):
        n = int(off[j + 1] - off[j])
        t = time.perf_counter()
        ctx.cluster_device(xd.ptr + 12 * int(off[j]), n, 3, np.float32, float(eps[j]), int(mps[j]),
                           ld.ptr + 8 * int(off[j]))
        lat.append(time.perf_counter() - t)
    ctx.synchronize()
    plain_labels = ld.to_array(np.int64, len(rows))
    t_plain = float(sum(lat))
    equal = bool(np.array_equal(batched_labels, plain_labels))
    xd.close()
    ld.close()
    ctx.close()
    lat_ms = np.array(lat) * 1e3
    return {
        "workload": "batch_clouds", "jobs": nj, "rows": int(len(rows)),
        "batched_jobs_per_s": round(nj / t_batch, 1), "batched_s": round(t_batch, 4),
        "batched_mpoints_per_s": round(len(rows) / t_batch / 1e6, 1),
        "plain_jobs_per_s": round(nj / t_plain, 1), "plain_s": round(t_plain, 3),
        "plain_p50_ms": round(float(np.percentile(lat_ms, 50)), 3),
        "plain_p99_ms": round(float(np.percentile(lat_ms, 99)), 3),
        "batched_latency_ms": round(t_batch * 1e3, 2),
        "speedup_jobs_per_s": round(t_plain / t_batch, 2),
        "labels_equal": equal,
        "work_split": tuned["best"] if tuned else {"box_split": 0, "sparse_max": 8,
                                                   "sparse_core": False},
        "note": "device-resident rows; a batched job's latency is the whole batch's",
        **graph,
    }


def _small_batch(ctx, gctx, rows, off, eps, mps, reps, jobs=256) -> dict:
    """The `jobs` smallest clouds of the batch as one batch, where the plain
    pipeline's host read-backs weigh most: plain batched call vs the
    hipGraph replay, median over 5 * reps calls each."""
    pick = np.argsort(np.diff(off), kind="stable")[:jobs]
    x, o = _lib.concat_jobs([rows[off[j]:off[j + 1]] for j in pick])
    e, m = eps[pick], mps[pick]
    out = {"jobs": int(len(pick)), "rows": int(len(x))}
    labels = []
    for name, c in (("plain_batched", ctx), ("graph", gctx)):
        xd = _lib.DeviceBuffer.from_array(c, x)
        ld = _lib.DeviceBuffer(c, 8 * len(x))
        c.cluster_batch_device(xd.ptr, o, 3, np.float32, e, m, ld.ptr)
        c.synchronize()
        ts = []
        for _ in range(5 * reps):
            t = time.perf_counter()
            c.cluster_batch_device(xd.ptr, o, 3, np.float32, e, m, ld.ptr)
            c.synchronize()
            ts.append(time.perf_counter() - t)
        labels.append(ld.to_array(np.int64, len(x)))
        out[f"{name}_ms"] = round(float(np.median(ts)) * 1e3, 3)
        out[f"{name}_jobs_per_s"] = round(len(pick) / float(np.median(ts)), 1)
        xd.close()
        ld.close()
    out["labels_equal"] = bool(np.array_equal(labels[0], labels[1]))
    return out


def _serve_stream(a, stream, scale=1.0, device=None) -> dict:
    """Replay `stream` (arrival times multiplied by `scale`) into a fresh
    server; `device`: per request (rows, labels) DeviceBuffers, submitted
    with hpd_server_submit_device.  Returns the latency / throughput figures."""
    import threading
    srv = _lib.Server(lanes=a.serve_lanes)
    tickets = [None] * len(stream)
    start = time.perf_counter()

    def producer():
        for i, (arr, x, e, m) in enumerate(stream):
            delay = start + arr * scale - time.perf_counter()
            if delay > 0:
                time.sleep(delay)
            if device is None:
                tickets[i] = srv.submit(x, e, m)
            else:
                xd, ld = device[i]
                tickets[i] = srv.submit_device(xd.ptr, len(x), x.shape[1], x.dtype, e, m, ld.ptr)

    th = threading.Thread(target=producer)
    th.start()
    th.join()
    infos = []
    for t in tickets:
        _, info = srv.wait(t)
        infos.append(info)
    wall = time.perf_counter() - start
    st = srv.stats()
    srv.close()
    lat = np.array([i["latency_ms"] for i in infos])
    span = stream[-1][0] * scale if stream else 0.0
    rows = sum(len(r[1]) for r in stream)
    return {
        "offered_jobs_per_s": round(len(stream) / span, 1) if span > 0 else None,
        "offered_mpoints_per_s": round(rows / span / 1e6, 1) if span > 0 else None,
        "achieved_jobs_per_s": round(len(stream) / wall, 1),
        "achieved_mpoints_per_s": round(rows / wall / 1e6, 1),
        "p50_ms": round(float(np.percentile(lat, 50)), 2),
        "p99_ms": round(float(np.percentile(lat, 99)), 2),
        "max_ms": round(float(lat.max()), 2),
        "batches": int(st["batches"]), "mean_batch_jobs": round(st["mean_batch_jobs"], 2),
    }


def serving_leg(a) -> dict:
    """Poisson arrivals into the continuous-batching server: host requests
    at the workload's rate, the same stream device-resident (rows and labels
    in HBM, hpd_server_submit_device), and host requests at 2x and 4x the
    rate (the same arrivals compressed)."""
    t0 = time.time()
    stream = datagen.serving_stream(n_requests=a.serve_requests, rate=a.serve_rate)
    offered_rows = sum(len(r[1]) for r in stream)
    span = stream[-1][0] if stream else 0.0
    print(f"[bench] serving: {len(stream)} requests over {span:.2f}s, {offered_rows} rows "
          f"(generated in {time.time() - t0:.1f}s)", file=sys.stderr, flush=True)
    srv = _lib.Server(lanes=a.serve_lanes)
    # warm-up request per dimension (contexts, allocations)
    for d in (2, 3, 4):
        w = next(r for r in stream if r[1].shape[1] == d)
        srv.wait(srv.submit(w[1], w[2], w[3]))
    srv.close()
    out = {"workload": "serving_stream", "requests": len(stream), "lanes": a.serve_lanes,
           **_serve_stream(a, stream)}
    # the same stream with every request's rows and labels in device memory
    ctx = _lib.Context(0)
    device = [(_lib.DeviceBuffer.from_array(ctx, x), _lib.DeviceBuffer(ctx, 8 * len(x)))
              for _, x, _, _ in stream]
    sweep = []
    try:
        dev = _serve_stream(a, stream, device=device)
        # a sample of the device requests' labels against plain calls
        for i in range(0, len(stream), max(1, len(stream) // 8)):
            _, x, e, m = stream[i]
            if not np.array_equal(device[i][1].to_array(np.int64, len(x)), ctx.cluster(x, e, m)):
                dev["labels_equal_sample"] = False
                break
        else:
            dev["labels_equal_sample"] = True
        print(f"[bench] serving device-resident: p50 {dev['p50_ms']} ms, p99 {dev['p99_ms']} ms",
              file=sys.stderr, flush=True)
        for mult in a.serve_sweep:
            for kind, bufs in (("host", None), ("device", device)):
                r = _serve_stream(a, stream, scale=1.0 / mult, device=bufs)
                sweep.append({"rate_multiple": mult, "requests_in": kind, **r})
                print(f"[bench] serving {kind} at {mult}x: achieved {r['achieved_jobs_per_s']} of "
                      f"{r['offered_jobs_per_s']} jobs/s, p50 {r['p50_ms']} ms, "
                      f"p99 {r['p99_ms']} ms", file=sys.stderr, flush=True)
    finally:
        for xd, ld in device:
            xd.close()
            ld.close()
        ctx.close()
    out["device_resident"] = dev
    out["rate_sweep"] = sweep
    return out


def device_steps(ctx, x, eps, mp, steps, warmup, timing=True):
    """Clusterings of x with rows and labels resident in device memory
    (hpd_cluster_device), bracketed by stream synchronisation.  Returns
    (Mpoints/s, ms per step, per-step stats, labels)."""
    n, d = x.shape
    xd = _lib.DeviceBuffer.from_array(ctx, x)
    ld = _lib.DeviceBuffer(ctx, 8 * n)
    try:
        ctx.set_timing(timing)
        for _ in range(warmup):
            ctx.cluster_device(xd.ptr, n, d, x.dtype, eps, mp, ld.ptr)
        ctx.synchronize()
        phases = []
        t0 = time.perf_counter()
        for _ in range(steps):
            ctx.cluster_device(xd.ptr, n, d, x.dtype, eps, mp, ld.ptr)
            phases.append(ctx.stats())
        ctx.synchronize()
        elapsed = time.perf_counter() - t0
        ctx.set_timing(False)
        labels = ld.to_array(np.int64, n)
    finally:
        xd.close()
        ld.close()
    return n * steps / elapsed / 1e6, 1e3 * elapsed / steps, phases, labels


def host_to_host(ctx, x, eps, mp, reps) -> dict:
    """The reference's boundary: host rows in, host labels out (hpd_cluster,
    pageable numpy buffers), with the two copies timed on their own."""
    n, d = x.shape
    ctx.cluster(x, eps, mp)  # warm-up (workspace)
    t_all, t_h2d, t_d2h = [], [], []
    xb = _lib.DeviceBuffer(ctx, x.nbytes)
    lb = _lib.DeviceBuffer(ctx, 8 * n)
    try:
        for _ in range(reps):
            t = time.perf_counter()
            ctx.cluster(x, eps, mp)
            t_all.append(time.perf_counter() - t)
            t = time.perf_counter()
            xb.copy_from(x)
            t_h2d.append(time.perf_counter() - t)
            t = time.perf_counter()
            lb.to_array(np.int64, n)
            t_d2h.append(time.perf_counter() - t)
    finally:
        xb.close()
        lb.close()
    med = float(np.median(t_all))
    return {"value": round(n / med / 1e6, 2), "unit": "Mpoints/s", "ms_per_call": round(1e3 * med, 3),
            "h2d_ms": round(1e3 * float(np.median(t_h2d)), 3),
            "d2h_ms": round(1e3 * float(np.median(t_d2h)), 3),
            "h2d_bytes": int(x.nbytes), "d2h_bytes": 8 * n,
            "note": "hpd_cluster on pageable host arrays (rows copied in, labels copied out, "
                    "inside the call); h2d/d2h: the same copies timed alone"}


def reference_leg(sample, eps, mp, threads, reps, what) -> dict:
    res, labels, err = run_reference(sample, eps, mp, threads, reps)
    if res is None:
        return {"value": None, "error": err, "sample": what}
    return {"value": round(len(sample) / res["median_s"] / 1e6, 4), "unit": "Mpoints/s",
            "threads": threads, "rows": int(len(sample)), "times_s": res["times_s"],
            "sample": what, "labels": labels}


def parity_fields(gpu_labels, ref_labels) -> dict:
    return {"core_flags_equal": bool(np.array_equal(gpu_labels < 0, ref_labels < 0)),
            "noise_equal": bool(np.array_equal(gpu_labels == 0, ref_labels == 0)),
            "core_ids_equal": bool(np.array_equal(gpu_labels[gpu_labels < 0],
                                                  ref_labels[gpu_labels < 0])),
            "label_mismatches": int((gpu_labels != ref_labels).sum())}


def p1_leg(a) -> dict:
    """The 100k x 3-D workload: GPU (device-resident and host-to-host) and
    the reference on all host threads and on one, full size."""
    x, eps, mp = datagen.make("blobs_100k")
    ctx = _lib.Context(0)
    try:
        v, ms, _, labels = device_steps(ctx, x, eps, mp, 50, 5, timing=False)
        hh = host_to_host(ctx, x, eps, mp, 10)
    finally:
        ctx.close()
    out = {"workload": "blobs_100k", "n": len(x), "d": 3, "eps": eps, "min_points": mp,
           "gpu": {"value": round(v, 2), "unit": "Mpoints/s", "ms_per_step": round(ms, 4),
                   "input": "device-resident"},
           "gpu_host_to_host": hh}
    if not a.no_cpu_baseline:
        threads = a.cpu_threads or cpu_threads()
        print(f"[bench] blobs_100k: reference, {threads} threads and 1 thread", file=sys.stderr,
              flush=True)
        ref = reference_leg(x, eps, mp, threads, 3, "the full 100k workload")
        ref1 = reference_leg(x, eps, mp, 1, 1, "the full 100k workload")
        out["cpu"] = {k: v for k, v in ref.items() if k != "labels"}
        out["cpu_one_thread"] = {k: v for k, v in ref1.items() if k != "labels"}
        if ref1.get("labels") is not None:
            # at one thread the reference's labels are fixed: under the
            # one-thread border rule every label is equal; under the default
            # rule only border rows next to two clusters may differ
            out["parity_one_thread_ref"] = parity_fields(_one_thread_labels(x, eps, mp),
                                                         ref1["labels"])
            out["parity_default_rule"] = parity_fields(labels, ref1["labels"])
        if ref.get("value"):
            out["speedup_vs_cpu"] = round(v / ref["value"], 1)
    return out


def p3_leg(a) -> dict:
    """The 100M x 2-D skewed workload on the GPU; the reference on a
    density-preserving slab, its time scaled to the full workload by the
    reference's own work measure (candidate distance tests)."""
    t0 = time.time()
    x, eps, mp = datagen.make("skewed_2d")
    n = len(x)
    print(f"[bench] skewed_2d: n={n} (generated in {time.time() - t0:.1f}s)", file=sys.stderr,
          flush=True)
    ctx = _lib.Context(0)
    try:
        v, ms, phases, labels = device_steps(ctx, x, eps, mp, a.p3_steps, 1)
        st = phases[-1]
        all_core = bool(np.all(labels == -1))
        del labels
        out = {"workload": "skewed_2d", "n": n, "d": 2, "eps": eps, "min_points": mp,
               "gpu": {"value": round(v, 2), "unit": "Mpoints/s", "ms_per_step": round(ms, 3),
                       "input": "device-resident"},
               "phases_ms": {k: round(float(np.median([p[k] for p in phases])), 3)
                             for k in ("ms_grid", "ms_sort", "ms_index", "ms_core", "ms_union",
                                       "ms_border", "ms_total")},
               "sum_candidates": int(st["sum_candidates"]), "occupied_cells": int(st["occupied_cells"]),
               "labels_all_minus_one": all_core}
        if not a.no_cpu_baseline:
            lo = np.partition(x[:, 0], n // 2)[n // 2]
            hi = np.partition(x[:, 0], n // 2 + a.p3_cpu_sample)[n // 2 + a.p3_cpu_sample]
            slab = np.ascontiguousarray(x[(x[:, 0] >= lo) & (x[:, 0] < hi)])
            ctx.cluster(slab, eps, mp)
            slab_cand = int(ctx.stats()["sum_candidates"])
            threads = a.cpu_threads or cpu_threads()
            print(f"[bench] skewed_2d: reference on a {len(slab)}-row slab, {threads} threads",
                  file=sys.stderr, flush=True)
            ref = reference_leg(slab, eps, mp, threads, 1,
                                f"{len(slab)} rows of the skewed_2d workload with x0 between its "
                                f"0.5 and {0.5 + len(slab) / n:.4f} quantiles (same density)")
            ref.pop("labels", None)
            if ref.get("value"):
                t_slab = float(ref["times_s"][0])
                t_full = t_slab * st["sum_candidates"] / max(1, slab_cand)
                ref["slab_sum_candidates"] = slab_cand
                ref["full_s_extrapolated"] = round(t_full, 1)
                ref["full_value_extrapolated"] = round(n / t_full / 1e6, 6)
                ref["note"] = ("a full run is infeasible (Σcand of the full workload / Σcand of "
                               "the slab times the slab's time); the slab is thinner than eps, so "
                               "its per-row work is far below the full workload's")
                out["speedup_vs_cpu_extrapolated"] = round(v / ref["full_value_extrapolated"], 1)
            out["cpu"] = ref
    finally:
        ctx.close()
    del x
    return out


def main() -> None:
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="blobs_10m", choices=sorted(datagen.WORKLOADS))
    ap.add_argument("--n", type=int, default=0, help="override the row count")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--cpu-sample", type=int, default=500_000,
                    help="rows in the CPU-baseline slab")
    ap.add_argument("--cpu-reps", type=int, default=3)
    ap.add_argument("--no-cpu-full", action="store_true",
                    help="skip the reference's full-size run (about 100 s on 16 threads)")
    ap.add_argument("--cpu-one-thread-sample", type=int, default=100_000,
                    help="rows of the one-thread reference run")
    ap.add_argument("--no-host-to-host", action="store_true")
    ap.add_argument("--no-p1", action="store_true")
    ap.add_argument("--no-p3", action="store_true")
    ap.add_argument("--p3-steps", type=int, default=3)
    ap.add_argument("--p3-cpu-sample", type=int, default=100_000)
    ap.add_argument("--no-batched", action="store_true")
    ap.add_argument("--batch-jobs", type=int, default=4096)
    ap.add_argument("--batch-reps", type=int, default=3)
    ap.add_argument("--no-graph", dest="graph", action="store_false",
                    help="batched leg: skip the hipGraph-captured static pipeline")
    ap.add_argument("--no-autotune", dest="autotune", action="store_false",
                    help="batched leg: library default work split instead of the autotuned one")
    ap.add_argument("--no-serving", action="store_true")
    ap.add_argument("--serve-requests", type=int, default=400)
    ap.add_argument("--serve-rate", type=float, default=200.0)
    ap.add_argument("--serve-lanes", type=int, default=3)
    ap.add_argument("--no-core-tests", dest="core_tests", action="store_false",
                    help="skip the instrumented clustering that counts the core count's "
                         "executed distance tests (keeps kernel traces to the timed runs)")
    ap.add_argument("--serve-sweep", type=float, nargs="*", default=[2.0, 4.0],
                    help="serving: also replay the stream at these multiples of its rate")
    a = ap.parse_args()

    if a.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {a.gpus} is not supported (this build has no "
              "multi-GPU path)", file=sys.stderr)
        sys.exit(2)

    w = datagen.WORKLOADS[a.workload]
    t0 = time.time()
    x, eps, mp = datagen.make(a.workload, a.n or None)
    n, d = x.shape
    print(f"[bench] workload {a.workload}: n={n} d={d} eps={eps} minPts={mp} "
          f"(generated in {time.time() - t0:.1f}s)", file=sys.stderr, flush=True)

    ctx = _lib.Context(0)
    # phase times come from HIP events the library records on its own stream
    # (the one every kernel is launched on) during the timed steps themselves
    value, ms_per_step, phases, labels = device_steps(ctx, x, eps, mp, a.steps, a.warmup)
    print(f"[bench] {a.steps} steps: {ms_per_step:.2f} ms/step, {value:.1f} Mpoints/s",
          file=sys.stderr, flush=True)
    st = phases[-1]
    keys = ["ms_grid", "ms_sort", "ms_index", "ms_core", "ms_union", "ms_border", "ms_total",
            "ms_core_kernel", "ms_hook_kernel", "ms_union_kernel", "ms_border_kernel"]
    med = {k: float(np.median([p[k] for p in phases])) for k in keys}
    # one more (untimed) clustering with the instrumented count kernel: the
    # distance tests it executes
    st_tests = {"core_tests": 0}
    if a.core_tests:
        ctx.set_timing(2)
        ctx.cluster(x, eps, mp)
        st_tests = ctx.stats()
        ctx.set_timing(False)
    roofline = kernel_roofline(st, med, n, d, x.dtype.itemsize, a.workload)
    pair_ms = med["ms_core"] + med["ms_hook_kernel"] + med["ms_union_kernel"] + med["ms_border_kernel"]

    result = {
        "metric": METRIC,
        "value": round(value, 2),
        "unit": "Mpoints/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": a.workload, "n": n, "d": d, "eps": eps, "min_points": mp,
                   "seed": w["seed"], "input": "device-resident rows, labels left on device"},
        "roofline": roofline,
        "phases_ms": {k: round(v, 3) for k, v in med.items()},
        "reference_tests": {
            "sum_candidates": int(st["sum_candidates"]),
            "per_s": round(float(st["sum_candidates"]) / (pair_ms * 1e-3), 1) if pair_ms > 0 else None,
            "note": "the reference's distance tests (3^D candidate cells per row) divided by the "
                    "summed time of the core, hook, union and border kernels, which skip tests "
                    "that bounding boxes decide"},
        "core_tests": (core_test_figures({**st, "core_tests": st_tests["core_tests"]}, med, d)
                       if a.core_tests else None),
        "stats": {k: int(st[k]) for k in ("occupied_cells", "core_points", "noise_points",
                                          "ambiguous_borders", "sum_candidates", "boxes",
                                          "neighbor_rows", "open_box_pairs")},
    }
    if not a.no_host_to_host:
        result["host_to_host"] = host_to_host(ctx, x, eps, mp, 5)
        print(f"[bench] host to host: {result['host_to_host']}", file=sys.stderr, flush=True)

    if not a.no_cpu_baseline:
        threads = a.cpu_threads or cpu_threads()
        result["host"] = host_info()
        sample = slab_sample(x, a.cpu_sample)
        print(f"[bench] cpu baseline: reference on {len(sample)} rows, {threads} threads",
              file=sys.stderr, flush=True)
        ref = reference_leg(sample, eps, mp, threads, a.cpu_reps,
                            f"{len(sample)} rows of the {a.workload} workload with x0 below its "
                            f"{len(sample) / n:.3f} quantile (same density); median of "
                            f"{a.cpu_reps} runs of HPDBSCAN::cluster<float>(ptr, n, d, {threads})")
        ref_labels = ref.pop("labels", None)
        if ref.get("value") is None:
            result["cpu_baseline"] = {"value": None, "error": ref.get("error")}
        else:
            cpu_v = ref["value"]
            result["cpu_baseline"] = {
                "value": cpu_v, "unit": "Mpoints/s", "cores": threads, "kind": "reference",
                "sample": ref["sample"], "times_s": ref["times_s"], "cpu": cpu_model()}
            gl = ctx.cluster(sample, eps, mp)
            result["parity"] = dict(sample_rows=int(len(sample)), **parity_fields(gl, ref_labels),
                                    note="reference run with several threads; only border rows "
                                         "next to two clusters may differ")
            result["speedup_vs_cpu"] = round(value / cpu_v, 1) if cpu_v > 0 else None
        small = slab_sample(x, a.cpu_one_thread_sample)
        print(f"[bench] cpu baseline: reference on {len(small)} rows, 1 thread", file=sys.stderr,
              flush=True)
        ref1 = reference_leg(small, eps, mp, 1, 1,
                             f"{len(small)} rows of the {a.workload} workload with x0 below its "
                             f"{len(small) / n:.4f} quantile (same density), one thread")
        ref1_labels = ref1.pop("labels", None)
        if ref1_labels is not None:
            ref1["parity_gpu_one_thread_rule"] = parity_fields(
                _one_thread_labels(small, eps, mp), ref1_labels)
        result["cpu_one_thread"] = ref1
        if not a.no_cpu_full:
            print(f"[bench] cpu baseline: reference on all {n} rows, {threads} threads (one run)",
                  file=sys.stderr, flush=True)
            full = reference_leg(x, eps, mp, threads, 1,
                                 f"the full {a.workload} workload, one run")
            full_labels = full.pop("labels", None)
            if full_labels is not None:
                full["parity"] = parity_fields(labels, full_labels)
                full["speedup_vs_cpu"] = round(value / full["value"], 1)
            result["cpu_full_size"] = full

    ctx.close()
    del labels
    if not a.no_p1:
        result["p1"] = p1_leg(a)
    if not a.no_p3:
        result["p3"] = p3_leg(a)
    if not a.no_batched:
        result["batched"] = batched_leg(a)
    if not a.no_serving:
        result["serving"] = serving_leg(a)
    print(json.dumps(result), flush=True)


def _one_thread_labels(x, eps, mp):
    """GPU labels under the one-thread border rule (every label comparable
    with the reference run at one thread)."""
    c = _lib.Context(0)
    try:
        c.set_border_rule(_lib.BORDER_ONE_THREAD)
        return c.cluster(x, eps, mp)
    finally:
        c.close()


if __name__ == "__main__":
    main()
