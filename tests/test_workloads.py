"""The benchmark workloads at their real sizes on the GPU, checked against the
oracle (the one-thread restatement where it finishes in seconds, its OpenMP
checker of the closed form, oracle.check_parallel, on larger inputs):

* skewed 2-D, 100M rows (densest cell ~4.75M rows): every row is core and
  the rows form one cluster at this density, so every label is -1; plus a
  density-preserving slab against the oracle;
* the batched workload (4096 clouds of 1k-200k rows): 32 clouds spanning the
  size range, the largest included, each checked against the oracle, and the
  whole batch equal to one call per cloud;
* the serving stream (10k-5M rows, d in {2, 3, 4}): every result of the
  server equal to clustering the request alone, and the requests of up to
  1.3M rows (at least two of 1M or more among them) checked against the
  oracle.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import sys

import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu

CHECK_THREADS = int(os.environ.get("HPD_CHECK_THREADS", "16"))


def say(msg: str) -> None:
    print(f"[gpu-case] {msg}", flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def ctx():
    from hpdbscan_amd import _lib
    c = _lib.Context(0)
    yield c
    c.close()


def assert_checked(x, eps, mp, got, what=""):
    r = oracle.check_parallel(x, eps, mp, got, threads=CHECK_THREADS)
    assert r["mismatches"] == 0 and r["bad_ambiguous"] == 0, (what, r)
    return r


def test_skewed_100m_full_size(ctx):
    from hpdbscan_amd import datagen
    x, eps, mp = datagen.make("skewed_2d")
    n = len(x)
    say(f"skewed_2d n={n} eps={eps} minPts={mp}: slab vs oracle")
    # a slab around the median of x0: the full workload's density, so every
    # cell is dense (~1000 rows per cell and up to 10^5 in the clusters)
    lo = np.partition(x[:, 0], n // 2)[n // 2]
    hi = np.partition(x[:, 0], n // 2 + 100_000)[n // 2 + 100_000]
    slab = np.ascontiguousarray(x[(x[:, 0] >= lo) & (x[:, 0] < hi)])
    got = ctx.cluster(slab, eps, mp)
    r = assert_checked(slab, eps, mp, got, "slab")
    assert r["cores"] == int(np.count_nonzero(got < 0))
    # the oracle's one-thread labels on the slab too (6 s)
    np.testing.assert_array_equal(got, oracle.cluster(slab, eps, mp))
    say("skewed_2d full size")
    a = ctx.cluster(x, eps, mp)
    st = ctx.stats()
    # at 10^8 rows in the unit square the noise alone puts >= 1500 rows
    # within eps of every row (a quarter disk in a corner), minPts is 32:
    # every row is core, and cores within eps chain through the noise into
    # one cluster whose smallest sorted position is 0
    assert st["core_points"] == n and st["noise_points"] == 0
    assert bool(np.all(a == -1))
    b = ctx.cluster(x, eps, mp)
    np.testing.assert_array_equal(a, b)
    c = ctx.cluster_batch(x, [0, n], eps, mp)
    np.testing.assert_array_equal(a, c)


@pytest.fixture(scope="module")
def batch():
    from hpdbscan_amd import datagen
    return datagen.batch_clouds()


def test_batch_workload_clouds_vs_oracle(ctx, batch):
    """32 clouds of the batched workload spanning its size range (the largest
    included), clustered in one batched call, each equal to the oracle."""
    from hpdbscan_amd import _lib
    rows, off, eps, mps = batch
    sizes = np.diff(off)
    order = np.argsort(sizes, kind="stable")
    pick = sorted(set(order[np.linspace(0, len(order) - 1, 32).round().astype(int)].tolist()))
    assert int(order[-1]) in pick and len(pick) == 32
    clouds = [rows[off[j]:off[j + 1]] for j in pick]
    x, o = _lib.concat_jobs(clouds)
    say(f"batch workload: {len(pick)} clouds of {sizes[pick].min()}-{sizes[pick].max()} rows")
    got = ctx.cluster_batch(x, o, eps[pick], mps[pick])
    with cf.ThreadPoolExecutor(max_workers=4) as ex:
        res = list(ex.map(
            lambda i: oracle.check_parallel(clouds[i], float(eps[pick[i]]), int(mps[pick[i]]),
                                            got[o[i]:o[i + 1]], threads=max(1, CHECK_THREADS // 4)),
            range(len(pick))))
    bad = [(pick[i], r) for i, r in enumerate(res) if r["mismatches"] or r["bad_ambiguous"]]
    assert not bad, bad
    # the smaller clouds also against the one-thread labels
    for i, j in enumerate(pick):
        if sizes[j] <= 20_000:
            want = oracle.cluster(clouds[i], float(eps[j]), int(mps[j]))
            mism, amb_bad = oracle.check(clouds[i], float(eps[j]), int(mps[j]), got[o[i]:o[i + 1]])
            assert mism == 0 and amb_bad == 0
            assert np.count_nonzero(got[o[i]:o[i + 1]] != want) <= res[i]["ambiguous"]


def test_batch_workload_full_equals_per_job_loop(ctx, batch):
    rows, off, eps, mps = batch
    nj = len(off) - 1
    say(f"batch workload: all {nj} clouds, {len(rows)} rows, batched vs one call per cloud")
    got = ctx.cluster_batch(rows, off, eps, mps)
    st = ctx.stats()
    assert st["n"] == len(rows)
    for j in range(nj):
        want = ctx.cluster(rows[off[j]:off[j + 1]], float(eps[j]), int(mps[j]))
        if not np.array_equal(got[off[j]:off[j + 1]], want):
            raise AssertionError(f"cloud {j} ({off[j + 1] - off[j]} rows) differs")


def test_serving_stream_alone_in_the_mix_and_vs_oracle():
    from hpdbscan_amd import _lib, datagen
    stream = datagen.serving_stream(n_requests=60)
    reqs = [(x, e, m) for _, x, e, m in stream[:36]]
    # two requests of at least 1M rows that the oracle checks: prefixes of
    # the stream's largest requests, as the stream itself forms requests
    big = sorted(stream, key=lambda r: -len(r[1]))[:2]
    reqs += [(big[0][1][:1_100_000], big[0][2], big[0][3]),
             (big[1][1][:1_000_000], big[1][2], big[1][3])]
    sizes = [len(x) for x, _, _ in reqs]
    say(f"serving: {len(reqs)} requests of {min(sizes)}-{max(sizes)} rows, d in "
        f"{sorted(set(x.shape[1] for x, _, _ in reqs))}")
    with _lib.Server(lanes=3) as srv:
        tickets = [srv.submit(x, e, m) for x, e, m in reqs]
        outs = [srv.wait(t)[0] for t in tickets]
        st = srv.stats()
    assert st["completed"] == len(reqs) and st["failed"] == 0
    ctx = _lib.Context(0)
    try:
        for (x, e, m), got in zip(reqs, outs):
            np.testing.assert_array_equal(got, ctx.cluster(x, e, m))
    finally:
        ctx.close()
    checked = 0
    for (x, e, m), got in zip(reqs, outs):
        if len(x) <= 1_300_000:
            say(f"serving: oracle check of {len(x)} rows d={x.shape[1]}")
            assert_checked(x, e, m, got, f"{len(x)} rows")
            checked += 1
    assert checked >= 20 and sum(1 for s in sizes if 1_000_000 <= s <= 1_300_000) >= 2
