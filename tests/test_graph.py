"""Graph mode (hpd_context_set_graph): the static pipeline, captured once in a
hipGraph and replayed, gives the plain pipeline's labels (and the oracle's),
on every replay, with new data, eps and minPts of the same shape, and falls
back to the plain pipeline where its flags say so."""
from __future__ import annotations

import sys

import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu


def say(msg: str) -> None:
    print(f"[gpu-case] {msg}", flush=True)
    sys.stdout.flush()


def blobs(n, d, k, sigma, box, seed, noise=0.1):
    rng = np.random.default_rng(seed)
    c = rng.uniform(0, box, (k, d))
    nb = int(n * (1 - noise))
    x = np.concatenate([c[rng.integers(0, k, nb)] + rng.normal(0, sigma, (nb, d)),
                        rng.uniform(0, box, (n - nb, d))])
    return x[rng.permutation(n)].astype(np.float32)


@pytest.fixture(scope="module")
def plain():
    from hpdbscan_amd import _lib
    c = _lib.Context(0)
    yield c
    c.close()


@pytest.fixture()
def graph():
    from hpdbscan_amd import _lib
    c = _lib.Context(0)
    c.set_graph(True)
    yield c
    c.close()


def batch(seed, jobs=20, d=3):
    from hpdbscan_amd import _lib
    rng = np.random.default_rng(seed)
    clouds, eps, mps = [], [], []
    for j in range(jobs):
        n = int(np.exp(rng.uniform(np.log(200), np.log(8000))))
        clouds.append(blobs(n, d, int(rng.integers(3, 12)), 0.2, 6.0, seed * 100 + j))
        eps.append(float(rng.uniform(0.2, 0.5)))
        mps.append(int(rng.choice([4, 8, 16])))
    x, off = _lib.concat_jobs(clouds)
    return clouds, x, off, np.array(eps, np.float32), np.array(mps, np.uint64)


def run_device(ctx, xd, off, d, eps, mps, ld, n):
    ctx.cluster_batch_device(xd.ptr, off, d, np.float32, eps, mps, ld.ptr)
    return ld.to_array(np.int64, n)


@pytest.mark.parametrize("d", [2, 3, 4])
def test_graph_batch_replays_equal_plain_and_oracle(plain, graph, d):
    from hpdbscan_amd import _lib
    clouds, x, off, eps, mps = batch(7 + d, d=d)
    n = len(x)
    say(f"graph batch d={d}: {len(clouds)} clouds, {n} rows")
    want = plain.cluster_batch(x, off, eps, mps)
    for j, c in enumerate(clouds):
        mism, bad = oracle.check(c, float(eps[j]), int(mps[j]), want[off[j]:off[j + 1]])
        assert mism == 0 and bad == 0
    xd = _lib.DeviceBuffer.from_array(graph, x)
    ld = _lib.DeviceBuffer(graph, 8 * n)
    other = blobs(3000, d, 4, 0.3, 5.0, 99)
    try:
        for rep in range(4):
            got = run_device(graph, xd, off, d, eps, mps, ld, n)
            np.testing.assert_array_equal(got, want)
            # the program's own eager work between replays
            plain.cluster(other, 0.3, 5)
        st = graph.graph_stats()
        assert st == {"launches": 4, "captures": 1, "fallbacks": 0}, st
        assert graph.stats()["core_points"] == int(np.count_nonzero(want < 0))
        # new data, eps and minPts of the same shape: replayed, not captured
        rng = np.random.default_rng(d)
        x2 = np.ascontiguousarray(x + rng.normal(0, 0.05, x.shape).astype(np.float32))
        eps2 = (eps * np.float32(1.1)).astype(np.float32)
        mps2 = np.where(mps > 4, mps - 2, mps).astype(np.uint64)
        xd.copy_from(x2)
        got2 = run_device(graph, xd, off, d, eps2, mps2, ld, n)
        np.testing.assert_array_equal(got2, plain.cluster_batch(x2, off, eps2, mps2))
        assert graph.graph_stats()["captures"] == 1
    finally:
        xd.close()
        ld.close()


def test_graph_plain_call_and_recapture(plain, graph):
    from hpdbscan_amd import _lib
    x = blobs(50_000, 3, 25, 0.3, 10.0, 4)
    say("graph plain call n=50000, then another shape")
    want = plain.cluster(x, 0.3, 6)
    xd = _lib.DeviceBuffer.from_array(graph, x)
    ld = _lib.DeviceBuffer(graph, 8 * len(x))
    try:
        for _ in range(3):
            graph.cluster_device(xd.ptr, len(x), 3, np.float32, 0.3, 6, ld.ptr)
            np.testing.assert_array_equal(ld.to_array(np.int64, len(x)), want)
        assert graph.graph_stats()["captures"] == 1
    finally:
        xd.close()
        ld.close()
    # the host entry point with another shape: captured anew
    y = blobs(20_000, 2, 10, 0.2, 6.0, 5)
    np.testing.assert_array_equal(graph.cluster(y, 0.25, 5), plain.cluster(y, 0.25, 5))
    assert graph.graph_stats()["captures"] == 2


def test_graph_falls_back_on_loose_boxes(plain, graph):
    """x near 6e6: loose sub-boxes; the static pipeline hands the run to the
    plain pipeline, which gives the same labels as a plain context."""
    rng = np.random.default_rng(3)
    x = np.stack([np.float32(6.0e6) + rng.integers(0, 4000, 6000).astype(np.float32) * np.float32(0.5),
                  (rng.integers(0, 3, 6000) * 0.1).astype(np.float32)], axis=1).astype(np.float32)
    x = np.concatenate([x, np.zeros((1, 2), np.float32)])
    say("graph fallback on loose boxes")
    np.testing.assert_array_equal(graph.cluster(x, 0.3, 3), plain.cluster(x, 0.3, 3))
    assert graph.graph_stats()["fallbacks"] == 1
    with pytest.raises(ValueError):
        bad = x.copy()
        bad[7, 0] = np.nan
        graph.cluster(bad, 0.3, 3)


def test_graph_batch_workload_sample(plain, graph):
    """256 clouds of the batched workload, the largest of them included."""
    from hpdbscan_amd import _lib, datagen
    rows, off, eps, mps = datagen.batch_clouds(n_jobs=256, seed=11)
    say(f"graph: batched workload sample, {len(off) - 1} clouds, {len(rows)} rows")
    want = plain.cluster_batch(rows, off, eps, mps)
    xd = _lib.DeviceBuffer.from_array(graph, rows)
    ld = _lib.DeviceBuffer(graph, 8 * len(rows))
    try:
        for _ in range(2):
            got = run_device(graph, xd, off, 3, eps, mps, ld, len(rows))
            np.testing.assert_array_equal(got, want)
        assert graph.graph_stats() == {"launches": 2, "captures": 1, "fallbacks": 0}
    finally:
        xd.close()
        ld.close()


@pytest.mark.parametrize("split", [(8, False), (0, False), (0, True)])
def test_graph_sparse_rows_beyond_the_likely_grid(plain, graph, split):
    """~1.5 rows per cell: more chunk items than the one-wave-per-item grid
    of a device-written count covers (an eighth of the rows), so the
    walking grid takes the rest (core count; border labels with every cell
    dense)."""
    from hpdbscan_amd import _lib
    rng = np.random.default_rng(21)
    x = rng.uniform(0, 100.0, (60_000, 2)).astype(np.float32)
    say(f"graph sparse rows, work split {split}")
    graph.set_work_split(*split)
    want = oracle.cluster(x, 0.5, 3)
    got = graph.cluster(x, 0.5, 3)
    np.testing.assert_array_equal(got, want)
    np.testing.assert_array_equal(got, plain.cluster(x, 0.5, 3))
    st = graph.graph_stats()
    assert st["captures"] == 1 and st["fallbacks"] == 0, st
    xd = _lib.DeviceBuffer.from_array(graph, x)
    ld = _lib.DeviceBuffer(graph, 8 * len(x))
    try:
        for _ in range(2):
            graph.cluster_device(xd.ptr, len(x), 2, np.float32, 0.5, 3, ld.ptr)
            np.testing.assert_array_equal(ld.to_array(np.int64, len(x)), want)
    finally:
        xd.close()
        ld.close()
