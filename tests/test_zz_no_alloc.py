"""hpd_cluster_device_noalloc (every workspace buffer in the library's static
device scratch) against hpd_cluster_device, through the C-ABI: two calls in a
row, the second with a smaller input, give identical labels; device memory
does not change across the calls; an input beyond the scratch is refused with
a clear error before any work."""
from __future__ import annotations

import ctypes
import gc
import sys

import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu


def say(msg: str) -> None:
    print(f"[gpu-case] {msg}", flush=True)
    sys.stdout.flush()


def blobs(n, d, k, sigma, box, seed):
    rng = np.random.default_rng(seed)
    c = rng.uniform(0, box, (k, d))
    nb = int(n * 0.9)
    x = np.concatenate([c[rng.integers(0, k, nb)] + rng.normal(0, sigma, (nb, d)),
                        rng.uniform(0, box, (n - nb, d))])
    return np.ascontiguousarray(x[rng.permutation(n)].astype(np.float32))


def device_free_bytes() -> int:
    hip = ctypes.CDLL("libamdhip64.so")
    free, total = ctypes.c_size_t(), ctypes.c_size_t()
    assert hip.hipMemGetInfo(ctypes.byref(free), ctypes.byref(total)) == 0
    return int(free.value)


def run_both(ctx, x, eps, mp):
    from hpdbscan_amd import _lib
    n, d = x.shape
    xd = _lib.DeviceBuffer.from_array(ctx, x)
    a = _lib.DeviceBuffer(ctx, 8 * n)
    b = _lib.DeviceBuffer(ctx, 8 * n)
    try:
        ctx.cluster_device(xd.ptr, n, d, np.float32, eps, mp, a.ptr)
        want = a.to_array(np.int64, n)
        st_want = ctx.stats()
        gc.collect()  # no finaliser frees device memory while the call runs
        before = device_free_bytes()
        ctx.cluster_device_noalloc(xd.ptr, n, d, np.float32, eps, mp, b.ptr)
        got = b.to_array(np.int64, n)
        after = device_free_bytes()
        st_got = ctx.stats()
    finally:
        xd.close()
        a.close()
        b.close()
    return want, got, st_want, st_got, before, after


@pytest.mark.parametrize("rule", ["min_cluster", "one_thread"])
def test_noalloc_equals_device_entry_point(rule):
    from hpdbscan_amd import _lib
    ctx = _lib.Context(0)
    try:
        if rule == "one_thread":
            ctx.set_border_rule(_lib.BORDER_ONE_THREAD)
        for n, d, seed, eps, mp in ((200_000, 3, 1, 0.3, 8), (30_000, 2, 2, 0.2, 5)):
            x = blobs(n, d, 30, 0.3, 10.0, seed)
            say(f"noalloc rule={rule} n={n} d={d}")
            want, got, sw, sg, before, after = run_both(ctx, x, eps, mp)
            np.testing.assert_array_equal(got, want)
            assert sg["core_points"] == sw["core_points"]
            assert sg["noise_points"] == sw["noise_points"]
            assert after == before, (before, after)
            mism, bad = oracle.check(x, eps, mp, got)
            assert mism == 0 and bad == 0
    finally:
        ctx.close()


def test_noalloc_rejects_inputs_beyond_the_scratch():
    from hpdbscan_amd import _lib
    ctx = _lib.Context(0)
    n, d = 2_000_000, 3
    xd = _lib.DeviceBuffer(ctx, 4 * n * d)
    out = _lib.DeviceBuffer(ctx, 8 * n)
    try:
        say(f"noalloc oversize n={n} d={d}")
        with pytest.raises(NotImplementedError, match="static scratch holds 134217728 bytes"):
            ctx.cluster_device_noalloc(xd.ptr, n, d, np.float32, 0.3, 8, out.ptr)
        # the context still works afterwards, on both entry points
        x = blobs(5000, 3, 5, 0.3, 5.0, 3)
        small = _lib.DeviceBuffer.from_array(ctx, x)
        ctx.cluster_device_noalloc(small.ptr, len(x), 3, np.float32, 0.3, 8, out.ptr)
        got = out.to_array(np.int64, len(x))
        np.testing.assert_array_equal(got, ctx.cluster(x, 0.3, 8))
        small.close()
    finally:
        xd.close()
        out.close()
        ctx.close()
