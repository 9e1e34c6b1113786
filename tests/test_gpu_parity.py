"""Parity of the HIP path (libhpdbscan_amd.so through the C-ABI) with the
reference, on an MI355X.

Bar: core/noise flags, core cluster ids and one-cluster border labels are
identical to the reference's one-thread labels; a border row next to cores
of two clusters carries the id of one of those clusters (the reference's own
multi-thread behaviour), checked by oracle.check.  With the one-thread border
rule every label equals the reference's one-thread label.
"""
from __future__ import annotations

import os
import sys

import numpy as np
import pytest

import oracle
from conftest import GOLDEN, golden_files

pytestmark = pytest.mark.gpu


def say(msg: str) -> None:
    print(f"[gpu-case] {msg}", flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def ctx():
    from hpdbscan_amd import _lib
    c = _lib.Context(0)
    c.set_timing(True)
    yield c
    c.close()


def assert_parity(x, eps, mp, got, want=None):
    if want is not None:
        diff = np.nonzero(got != want)[0]
        if len(diff) == 0:
            return
    mism, bad = oracle.check(x, eps, mp, got)
    assert mism == 0 and bad == 0, f"{mism} exact mismatches, {bad} bad ambiguous borders"


FIXTURES = [f for f in golden_files() if not f.startswith("sumtree_")]


@pytest.mark.parametrize("name", FIXTURES)
def test_golden_fixture(ctx, name):
    z = np.load(os.path.join(GOLDEN, name))
    x, eps, mp = z["data"], float(z["eps"]), int(z["min_pts"])
    say(f"fixture {name} n={len(x)} d={x.shape[1]} dtype={x.dtype}")
    got = ctx.cluster(x, eps, mp)
    assert_parity(x, eps, mp, got, z["labels"])
    # the reference's core flags and noise are exact
    np.testing.assert_array_equal(got < 0, z["labels"] < 0)
    np.testing.assert_array_equal(got == 0, z["labels"] == 0)


@pytest.mark.parametrize("name", golden_files("sumtree_"))
def test_summation_order_fixture(ctx, name):
    z = np.load(os.path.join(GOLDEN, name))
    say(f"sumtree {name}")
    for x, e, want in zip(z["data"], z["eps"], z["labels"]):
        np.testing.assert_array_equal(ctx.cluster(x, float(e), 2), want)


def blobs(n, d, k, sigma, box, seed, dtype, noise=0.1):
    rng = np.random.default_rng(seed)
    c = rng.uniform(0, box, (k, d))
    nb = int(n * (1 - noise))
    x = np.concatenate([c[rng.integers(0, k, nb)] + rng.normal(0, sigma, (nb, d)),
                        rng.uniform(0, box, (n - nb, d))])
    return x[rng.permutation(n)].astype(dtype)


CASES = [(d, dt, seed) for d in range(1, 9) for dt in (np.float32, np.float64) for seed in (0, 1)]


@pytest.mark.parametrize("d,dtype,seed", CASES)
def test_random_blobs_vs_oracle(ctx, d, dtype, seed):
    rng = np.random.default_rng(seed * 100 + d)
    n = int(rng.integers(2000, 12000)) if d <= 4 else int(rng.integers(1000, 4000))
    x = blobs(n, d, int(rng.integers(3, 30)), float(rng.uniform(0.1, 0.4)), 3.0 + d, seed + 7 * d, dtype)
    eps = float(rng.uniform(0.15, 0.5)) * (1.0 if d <= 4 else 2.0)
    mp = int(rng.integers(1, 20))
    say(f"random d={d} dtype={np.dtype(dtype).name} seed={seed} n={n} eps={eps:.3f} mp={mp}")
    want = oracle.cluster(x, eps, mp)
    got = ctx.cluster(x, eps, mp)
    assert_parity(x, eps, mp, got, want)
    st = ctx.stats()
    full = oracle.cluster(x, eps, mp, full=True)
    assert st["core_points"] == full["stats"]["core_points"]
    assert st["sum_candidates"] == full["stats"]["sum_candidates"]
    assert st["occupied_cells"] == full["stats"]["occupied_cells"]


@pytest.mark.parametrize("dtype", [np.int8, np.int16, np.int32, np.int64,
                                   np.uint8, np.uint16, np.uint32, np.uint64])
def test_integer_dtypes_vs_oracle(ctx, dtype):
    rng = np.random.default_rng(5)
    lo, hi = (0, 200) if np.dtype(dtype).kind == "u" else (-100, 100)
    x = rng.integers(lo, hi, size=(4000, 3)).astype(dtype)
    for eps, mp in ((0.5, 1), (4.0, 3), (9.5, 12)):
        say(f"int {np.dtype(dtype).name} eps={eps} mp={mp}")
        assert_parity(x, eps, mp, ctx.cluster(x, eps, mp), oracle.cluster(x, eps, mp))


def test_dropin_python_float64(ctx):
    """HPDBSCAN(eps, minPts).cluster(ndarray) runs in float64 like the SWIG
    module and returns a tuple; `import hpdbscan` is that module."""
    import hpdbscan as h
    x32 = blobs(5000, 3, 12, 0.25, 6.0, 3, np.float32)
    say("dropin float64 via HPDBSCAN.cluster")
    res = h.HPDBSCAN(0.3, 5).cluster(x32)
    assert isinstance(res, tuple) and len(res) == len(x32)
    x64 = x32.astype(np.float64)
    assert_parity(x64, 0.3, 5, np.array(res, dtype=np.int64), oracle.cluster(x64, 0.3, 5))
    # keyword arguments and list input
    res2 = h.HPDBSCAN(epsilon=0.3, min_points=5).cluster(x32.tolist())
    assert res2 == res


def test_determinism_and_empty(ctx):
    x = blobs(20000, 3, 30, 0.3, 10.0, 9, np.float32)
    say("determinism n=20000")
    a = ctx.cluster(x, 0.3, 6)
    b = ctx.cluster(x, 0.3, 6)
    np.testing.assert_array_equal(a, b)
    assert ctx.cluster(np.zeros((0, 3), np.float32), 0.3, 6).shape == (0,)


def test_device_entry_point(ctx):
    from hpdbscan_amd import _lib
    x = blobs(30000, 3, 25, 0.3, 10.0, 4, np.float32)
    say("device entry point n=30000")
    xd = _lib.DeviceBuffer.from_array(ctx, x)
    out = _lib.DeviceBuffer(ctx, 8 * len(x))
    ctx.cluster_device(xd.ptr, len(x), 3, np.float32, 0.3, 6, out.ptr)
    got = out.to_array(np.int64, len(x))
    np.testing.assert_array_equal(got, ctx.cluster(x, 0.3, 6))
    # explicit stream argument: the context's own stream
    ctx.cluster_device(xd.ptr, len(x), 3, np.float32, 0.3, 6, out.ptr, ctx.stream())
    np.testing.assert_array_equal(out.to_array(np.int64, len(x)), got)
    xd.close()
    out.close()


def test_nonfinite_input_rejected(ctx):
    x = blobs(100, 2, 2, 0.3, 3.0, 1, np.float32)
    x[5, 1] = np.nan
    say("nan input")
    with pytest.raises(ValueError):
        ctx.cluster(x, 0.3, 3)
    x[5, 1] = np.inf
    with pytest.raises(ValueError):
        ctx.cluster(x, 0.3, 3)


def test_loose_boxes_fall_back_to_point_union(ctx):
    """x near 6e6 with one row at 0 and eps 0.3: (x - min) * (1/eps) exceeds
    2^24, its float spacing is 2, cells are 2 eps wide and some sub-boxes are
    not tight; the slot-level union must take over and stay exact."""
    rng = np.random.default_rng(3)
    base = np.float32(6.0e6)
    x = np.stack([base + rng.integers(0, 4000, 6000).astype(np.float32) * np.float32(0.5),
                  (rng.integers(0, 3, 6000) * 0.1).astype(np.float32)], axis=1).astype(np.float32)
    x = np.concatenate([x, np.zeros((1, 2), np.float32)])
    say("loose boxes n=6001")
    got = ctx.cluster(x, 0.3, 3)
    assert ctx.stats()["box_union"] == 0
    assert_parity(x, 0.3, 3, got, oracle.cluster(x, 0.3, 3))


def test_box_union_used_and_exact(ctx):
    x = blobs(40000, 3, 20, 0.15, 8.0, 21, np.float32)
    say("box union n=40000")
    got = ctx.cluster(x, 0.3, 8)
    st = ctx.stats()
    assert st["box_union"] == 1 and st["boxes"] > 0
    assert_parity(x, 0.3, 8, got, oracle.cluster(x, 0.3, 8))
    # the instrumented core count (timing level 2): same labels; it executes
    # some of the reference's candidate tests, the bounding boxes decide the
    # others
    ctx.set_timing(2)
    try:
        np.testing.assert_array_equal(ctx.cluster(x, 0.3, 8), got)
        st2 = ctx.stats()
    finally:
        ctx.set_timing(True)
    assert 0 < st2["core_tests"] < st2["sum_candidates"]


@pytest.mark.parametrize("d", [2, 3])
def test_open_pair_limit_takes_the_full_join(ctx, d):
    """With the open-pair lists limited to one pair, the union walks every
    box's neighbourhood instead (decided on the device): same labels as with
    the lists, and the oracle's."""
    from hpdbscan_amd import _lib
    x = blobs(30000, d, 20, 0.15, {2: 6.0, 3: 8.0}[d], 40 + d, np.float32)
    eps, mp = {2: (0.1, 8), 3: (0.3, 8)}[d]
    say(f"open-pair limit 1, d={d}")
    want = ctx.cluster(x, eps, mp)
    assert ctx.stats()["open_box_pairs"] > 1
    c = _lib.Context(0)
    try:
        c.set_open_pair_limit(1)
        got = c.cluster(x, eps, mp)
        assert c.stats()["box_union"] == 1
    finally:
        c.close()
    np.testing.assert_array_equal(got, want)
    assert_parity(x, eps, mp, got, oracle.cluster(x, eps, mp))


@pytest.mark.parametrize("d,dtype,mp", [(2, np.float32, 2000), (3, np.float64, 400)])
def test_big_boxes_gathered_in_parts(ctx, d, dtype, mp):
    """A sub-box of 70k rows (over the 65536 from which a box's rows are
    gathered in parts across all such boxes) in the grid's corner, inside
    sparse rows whose neighbour counts sit around minPts: core flags,
    clusters and borders next to the big box equal the oracle's, and the
    static pipeline's (whose gather takes every box whole)."""
    from hpdbscan_amd import _lib
    rng = np.random.default_rng(70 + d)
    big = rng.uniform(0.0, 0.45, (70_000, d))
    sparse = rng.uniform(0.0, 6.0, (20_000, d))
    x = np.concatenate([big, sparse])[rng.permutation(90_000)].astype(dtype)
    lo = x.min(axis=0)
    assert np.count_nonzero(np.all(x - lo < 0.5, axis=1)) >= 70_000  # one sub-box
    say(f"big boxes d={d} {np.dtype(dtype).name} n=90000 eps=1 minPts={mp}")
    got = ctx.cluster(x, 1.0, mp)
    r = oracle.check_parallel(x, 1.0, mp, got, threads=16)
    assert r["mismatches"] == 0 and r["bad_ambiguous"] == 0, r
    cores = int(np.count_nonzero(got < 0))
    assert r["cores"] == cores and 70_000 < cores < 90_000
    assert (got > 0).any() and (got == 0).any()
    g = _lib.Context(0)
    try:
        g.set_graph(True)
        np.testing.assert_array_equal(g.cluster(x, 1.0, mp), got)
        assert g.graph_stats()["fallbacks"] == 0
    finally:
        g.close()


@pytest.mark.parametrize("d,dtype", [(1, np.float32), (2, np.float32), (3, np.float32),
                                     (4, np.float64), (6, np.float32)])
def test_box_split_never_changes_labels(d, dtype):
    """Every box split (1 to 5 sub-boxes per cell and dimension; 1: boxes
    are cells, wider than eps from 2-D on, which takes the row-level union)
    gives the labels of the default split, which equal the oracle's; with
    the one-thread border rule every label is the oracle's; the static
    pipeline takes the split too."""
    from hpdbscan_amd import _lib
    box, eps, mp = {1: (40.0, 0.02, 5), 2: (6.0, 0.1, 8), 3: (8.0, 0.3, 8), 4: (8.0, 0.5, 8),
                    6: (6.0, 0.9, 6)}[d]
    x = blobs(20000, d, 15, 0.15, box, 80 + d, dtype)
    say(f"box splits d={d} {np.dtype(dtype).name} n=20000 eps={eps} minPts={mp}")
    want = oracle.cluster(x, eps, mp)
    c = _lib.Context(0)
    c1 = _lib.Context(0)
    g = _lib.Context(0)
    try:
        c1.set_border_rule(_lib.BORDER_ONE_THREAD)
        g.set_graph(True)
        base = c.cluster(x, eps, mp)
        assert_parity(x, eps, mp, base, want)
        for k in (1, 2, 3, 4, 5):
            if k ** d > 1 << 24:
                continue
            c.set_box_split(k)
            assert c.box_split() == k
            np.testing.assert_array_equal(c.cluster(x, eps, mp), base, err_msg=f"split {k}")
            c1.set_box_split(k)
            np.testing.assert_array_equal(c1.cluster(x, eps, mp), want, err_msg=f"split {k}")
        g.set_box_split(3)
        np.testing.assert_array_equal(g.cluster(x, eps, mp), base)
        with pytest.raises(ValueError):
            c.set_box_split(9)
    finally:
        c.close()
        c1.close()
        g.close()


def dense_blobs(n, d, seed):
    """Four tight blobs on separated centres with a little noise around them:
    few occupied cells, each holding many rows."""
    rng = np.random.default_rng(seed)
    c = np.zeros((4, d))
    for i in range(4):
        c[i, i % d] = 1.0 + i // d
    nb = int(n * 0.97)
    x = np.concatenate([c[rng.integers(0, 4, nb)] + rng.normal(0, 0.05, (nb, d)),
                        c[rng.integers(0, 4, n - nb)] + rng.uniform(-0.25, 0.25, (n - nb, d))])
    return x[rng.permutation(n)].astype(np.float32)


@pytest.mark.parametrize("d", [2, 3, 4])
def test_dense_cells_nearest_first_rows(ctx, d):
    """With at least 3^(D-1) rows per occupied cell the neighbour rows are
    kept from the count pass and stored nearest first (own row, then by the
    number of offset dimensions); the labels equal the oracle's."""
    n = 30000 if d == 2 else 50000
    x = dense_blobs(n, d, 50 + d)
    say(f"dense cells d={d} n={n} eps=0.1 minPts=20")
    got = ctx.cluster(x, 0.1, 20)
    st = ctx.stats()
    assert st["occupied_cells"] * 3 ** (d - 1) <= n  # the nearest-first path ran
    assert_parity(x, 0.1, 20, got, oracle.cluster(x, 0.1, 20))
    assert len(np.unique(got[got < 0])) == 4 and (got > 0).any() and (got == 0).any()


def test_skewed_2d_sample(ctx):
    from hpdbscan_amd import datagen
    x = datagen.skewed_2d(n=200_000, seed=3)
    say("skewed 2-D sample n=200000 eps=0.01 minPts=32")
    got = ctx.cluster(x, 0.01, 32)
    assert_parity(x, 0.01, 32, got, oracle.cluster(x, 0.01, 32))


def test_blobs_10m_slab_and_full_size(ctx):
    """The headline workload: a 100k-row slab (rows with x0 below a quantile,
    same density) against the oracle, and at the full 10M rows the
    size-independent properties: two runs are identical, a batch of one
    equals the plain call, every core label is negative and every cluster id
    k has a core labelled -k."""
    from hpdbscan_amd import _lib, datagen
    x, eps, mp = datagen.make("blobs_10m")
    cut = np.partition(x[:, 0], 100_000)[100_000]
    slab = np.ascontiguousarray(x[x[:, 0] < cut])
    say(f"blobs_10m slab n={len(slab)} vs oracle")
    got = ctx.cluster(slab, eps, mp)
    assert_parity(slab, eps, mp, got, oracle.cluster(slab, eps, mp))
    say("blobs_10m full size, properties")
    a = ctx.cluster(x, eps, mp)
    st = ctx.stats()
    b = ctx.cluster(x, eps, mp)
    np.testing.assert_array_equal(a, b)
    c = ctx.cluster_batch(x, [0, len(x)], eps, mp)
    np.testing.assert_array_equal(a, c)
    assert int(np.count_nonzero(a < 0)) == st["core_points"]
    assert int(np.count_nonzero(a == 0)) == st["noise_points"]
    ids = np.unique(np.abs(a[a != 0]))
    core_ids = np.unique(-a[a < 0])
    np.testing.assert_array_equal(ids, core_ids)  # every cluster has a core


def test_batch_matches_single_jobs(ctx):
    """One batched call returns, for every job, the labels of clustering the
    job alone (and the reference's labels)."""
    from hpdbscan_amd import _lib
    rng = np.random.default_rng(17)
    clouds, eps, mps = [], [], []
    for j in range(24):
        n = int(np.exp(rng.uniform(np.log(50), np.log(6000))))
        if j == 5:
            n = 0
        if j == 9:
            n = 1
        clouds.append(blobs(n, 3, int(rng.integers(2, 10)), 0.2, 6.0, 1000 + j, np.float32)
                      if n > 1 else rng.uniform(0, 5, (n, 3)).astype(np.float32))
        eps.append(float(rng.uniform(0.2, 0.5)))
        mps.append(int(rng.choice([4, 8, 16])))
    x, off = _lib.concat_jobs(clouds)
    say(f"batch of {len(clouds)} jobs, {len(x)} rows")
    got = ctx.cluster_batch(x, off, eps, mps)
    for j, c in enumerate(clouds):
        g = got[off[j]:off[j + 1]]
        if len(c) == 0:
            continue
        np.testing.assert_array_equal(g, ctx.cluster(c, eps[j], mps[j]))
        assert_parity(c, eps[j], mps[j], g, oracle.cluster(c, eps[j], mps[j]))


@pytest.mark.parametrize("dtype", [np.float64, np.int32])
def test_batch_other_dtypes(ctx, dtype):
    from hpdbscan_amd import _lib
    rng = np.random.default_rng(5)
    clouds = []
    for j in range(6):
        if np.dtype(dtype).kind == "f":
            clouds.append(blobs(int(rng.integers(500, 3000)), 2, 5, 0.3, 5.0, j, dtype))
        else:
            clouds.append(rng.integers(0, 60, size=(int(rng.integers(500, 3000)), 2)).astype(dtype))
    eps = [0.4, 0.3, 2.5, 0.35, 3.0, 0.5]
    x, off = _lib.concat_jobs(clouds)
    say(f"batch dtype={np.dtype(dtype).name}")
    got = ctx.cluster_batch(x, off, eps, 5)
    for j, c in enumerate(clouds):
        assert_parity(c, eps[j], 5, got[off[j]:off[j + 1]], oracle.cluster(c, eps[j], 5))


# ---------------------------------------------------------------------------
# HPD_BORDER_ONE_THREAD: every label, ambiguous borders included, equals the
# reference's at one thread.
# ---------------------------------------------------------------------------
@pytest.fixture(scope="module")
def ctx1():
    from hpdbscan_amd import _lib
    c = _lib.Context(0)
    c.set_border_rule(_lib.BORDER_ONE_THREAD)
    yield c
    c.close()


@pytest.mark.parametrize("name", FIXTURES)
def test_one_thread_rule_golden(ctx1, name):
    z = np.load(os.path.join(GOLDEN, name))
    x, eps, mp = z["data"], float(z["eps"]), int(z["min_pts"])
    say(f"one-thread rule fixture {name}")
    np.testing.assert_array_equal(ctx1.cluster(x, eps, mp), z["labels"])


# (n, d, clusters, sigma, box, seed, dtype, eps, minPts): overlapping blobs
# with 30 % uniform noise, chosen so that borders next to two clusters occur
# and the smallest-cluster choice differs from the one-thread choice
ONE_THREAD_CASES = [(1500, 2, 40, 0.35, 6.0, s, np.float32, 0.25, 6) for s in range(6)] + [
    (20000, 2, 60, 0.3, 10.0, 1, np.float32, 0.12, 8),
    (20000, 3, 60, 0.3, 8.0, 1, np.float64, 0.2, 10),
    (30000, 4, 80, 0.3, 7.0, 1, np.float32, 0.35, 12),
    (8000, 1, 30, 0.4, 40.0, 2, np.float32, 0.02, 4),
    (6000, 6, 40, 0.5, 5.0, 3, np.float64, 0.9, 8),
]


@pytest.mark.parametrize("case", ONE_THREAD_CASES, ids=lambda c: f"n{c[0]}d{c[1]}s{c[5]}")
def test_one_thread_rule_exact(ctx1, case):
    n, d, k, sig, box, seed, dt, eps, mp = case
    x = blobs(n, d, k, sig, box, seed, dt, noise=0.3)
    say(f"one-thread rule n={n} d={d} seed={seed} dtype={np.dtype(dt).name}")
    full = oracle.cluster(x, eps, mp, full=True)
    got = ctx1.cluster(x, eps, mp)
    np.testing.assert_array_equal(got, full["labels"])
    assert ctx1.stats()["ambiguous_borders"] == full["stats"]["ambiguous_borders"]


def test_one_thread_rule_discriminates(ctx, ctx1):
    """The cases above hold borders where the two rules choose differently."""
    differ = 0
    for case in ONE_THREAD_CASES[:6]:
        n, d, k, sig, box, seed, dt, eps, mp = case
        x = blobs(n, d, k, sig, box, seed, dt, noise=0.3)
        differ += int(np.count_nonzero(ctx.cluster(x, eps, mp) != ctx1.cluster(x, eps, mp)))
    say(f"rows where the border rules differ: {differ}")
    assert differ > 0


def test_one_thread_rule_point_union_and_integers(ctx1):
    rng = np.random.default_rng(3)
    base = np.float32(6.0e6)
    x = np.stack([base + rng.integers(0, 4000, 6000).astype(np.float32) * np.float32(0.5),
                  (rng.integers(0, 3, 6000) * 0.1).astype(np.float32)], axis=1).astype(np.float32)
    x = np.concatenate([x, np.zeros((1, 2), np.float32)])
    say("one-thread rule, loose boxes")
    np.testing.assert_array_equal(ctx1.cluster(x, 0.3, 3), oracle.cluster(x, 0.3, 3))
    assert ctx1.stats()["box_union"] == 0
    xi = (blobs(6000, 2, 30, 3.0, 100.0, 4, np.float64, noise=0.3)).astype(np.int32)
    say("one-thread rule, int32")
    np.testing.assert_array_equal(ctx1.cluster(xi, 2.5, 5), oracle.cluster(xi, 2.5, 5))


def lattice_blobs(n, d, k_per_dim, spacing, sigma, noise_frac, seed):
    """Tight blobs on a k^d lattice with uniform noise between them: few
    occupied cells with many rows each, and border rows between blobs that
    lie within eps of cores of two clusters."""
    rng = np.random.default_rng(seed)
    g = np.stack(np.meshgrid(*[np.arange(k_per_dim)] * d, indexing="ij"), -1).reshape(-1, d)
    g = g * spacing
    nb = int(n * (1 - noise_frac))
    x = np.concatenate([g[rng.integers(0, len(g), nb)] + rng.normal(0, sigma, (nb, d)),
                        rng.uniform(-spacing / 2, spacing * (k_per_dim - 0.5), (n - nb, d))])
    return x[rng.permutation(n)].astype(np.float32)


# (d, n, blobs per dim, spacing, sigma, noise, eps, minPts): every case has
# m * 3^(d-1) <= n (neighbour rows kept from the count pass, nearest first)
# and border rows where the one-thread choice differs from the smallest
# cluster (94 rows in 3-D, 9 in 4-D, counted on the CPU)
DENSE_ONE_THREAD = [(3, 60000, 3, 0.38, 0.03, 0.1, 0.1, 40),
                    (4, 200000, 3, 0.4, 0.03, 0.05, 0.12, 40)]


@pytest.mark.parametrize("case", DENSE_ONE_THREAD, ids=lambda c: f"d{c[0]}")
def test_one_thread_rule_dense_nearest_first(ctx, ctx1, case):
    """The border relaxation reads neighbour rows stored nearest first (not
    ascending); every label equals the reference's one-thread label."""
    d, n, k, sp, sig, nf, eps, mp = case
    x = lattice_blobs(n, d, k, sp, sig, nf, 7)
    say(f"one-thread rule, dense cells d={d} n={n}")
    full = oracle.cluster(x, eps, mp, full=True)
    got = ctx1.cluster(x, eps, mp)
    st = ctx1.stats()
    assert st["occupied_cells"] * 3 ** (d - 1) <= n
    assert st["ambiguous_borders"] == full["stats"]["ambiguous_borders"] > 0
    np.testing.assert_array_equal(got, full["labels"])
    # the default rule chooses differently on some of these rows
    assert int(np.count_nonzero(ctx.cluster(x, eps, mp) != got)) > 0


def test_loose_boxes_dense_cells_3d(ctx, ctx1):
    """Loose boxes (x0 near 6e6, float spacing 0.5 > eps) in 3-D with dense
    cells: the slot-level hook and union read neighbour rows stored nearest
    first; every label equals the oracle's under both border rules."""
    rng = np.random.default_rng(11)
    n = 40000
    x = np.stack([np.float32(6.0e6) + rng.integers(0, 200, n).astype(np.float32) * np.float32(0.5),
                  rng.uniform(0, 1.0, n).astype(np.float32),
                  rng.uniform(0, 1.0, n).astype(np.float32)], axis=1).astype(np.float32)
    x = np.concatenate([x, np.zeros((1, 3), np.float32)])
    eps, mp = 0.3, 60  # 186 clusters, 113 border rows next to two of them
    say(f"loose boxes, dense 3-D cells n={len(x)}")
    want = oracle.cluster(x, eps, mp)
    got = ctx.cluster(x, eps, mp)
    st = ctx.stats()
    assert st["box_union"] == 0
    assert st["occupied_cells"] * 9 <= len(x)
    assert_parity(x, eps, mp, got, want)
    np.testing.assert_array_equal(ctx1.cluster(x, eps, mp), want)


def test_one_thread_rule_batch(ctx1):
    from hpdbscan_amd import _lib
    clouds = [blobs(1500, 2, 40, 0.35, 6.0, s, np.float32, noise=0.3) for s in range(4)]
    x, off = _lib.concat_jobs(clouds)
    say("one-thread rule, batch of 4")
    got = ctx1.cluster_batch(x, off, 0.25, 6)
    for j, c in enumerate(clouds):
        np.testing.assert_array_equal(got[off[j]:off[j + 1]], oracle.cluster(c, 0.25, 6))


def test_server_results_equal_alone_and_in_the_mix(ctx):
    """Requests of mixed sizes and dimensions submitted at once come back
    identical to clustering each request alone."""
    from hpdbscan_amd import _lib
    rng = np.random.default_rng(23)
    reqs = []
    for i in range(30):
        d = int(rng.choice([2, 3, 4]))
        n = int(np.exp(rng.uniform(np.log(200), np.log(20000))))
        x = blobs(n, d, int(rng.integers(5, 20)), 0.2, 10.0, 300 + i, np.float32)
        reqs.append((x, float(rng.uniform(0.2, 0.5)), int(rng.integers(4, 33))))
    say(f"server: {len(reqs)} requests")
    with _lib.Server(lanes=3) as srv:
        tickets = [srv.submit(x, e, m) for x, e, m in reqs]
        outs = [srv.wait(t) for t in tickets]
        st = srv.stats()
    assert st["completed"] == len(reqs) and st["failed"] == 0
    assert st["batches"] < len(reqs)  # requests were batched
    for (x, e, m), (got, info) in zip(reqs, outs):
        np.testing.assert_array_equal(got, ctx.cluster(x, e, m))
        assert_parity(x, e, m, got, oracle.cluster(x, e, m))


def test_server_device_and_host_requests_in_one_stream(ctx):
    """Device-resident requests (rows and labels in HBM) mixed with host
    requests of the same dimension: every device request's labels land in
    its own device buffer, equal to clustering it alone, as do the host
    requests'; the lanes' grow-only buffers serve growing batches."""
    from hpdbscan_amd import _lib
    rng = np.random.default_rng(29)
    reqs = []
    for i in range(24):
        n = int(np.exp(rng.uniform(np.log(300), np.log(60000))))
        x = blobs(n, 3, int(rng.integers(5, 20)), 0.2, 10.0, 500 + i, np.float32)
        reqs.append((x, float(rng.uniform(0.2, 0.5)), int(rng.integers(4, 33)), i % 2 == 0))
    say(f"server: {len(reqs)} requests, half device-resident")
    bufs = []
    try:
        with _lib.Server(lanes=2) as srv:
            tickets = []
            for x, e, m, dev in reqs:
                if dev:
                    xd = _lib.DeviceBuffer.from_array(ctx, x)
                    ld = _lib.DeviceBuffer(ctx, 8 * len(x))
                    bufs.append((xd, ld))
                    tickets.append(srv.submit_device(xd.ptr, len(x), 3, np.float32, e, m, ld.ptr))
                else:
                    bufs.append(None)
                    tickets.append(srv.submit(x, e, m))
            outs = [srv.wait(t) for t in tickets]
            st = srv.stats()
        assert st["completed"] == len(reqs) and st["failed"] == 0
        for (x, e, m, dev), (got, info), b in zip(reqs, outs, bufs):
            if dev:
                assert got is None
                got = b[1].to_array(np.int64, len(x))
            np.testing.assert_array_equal(got, ctx.cluster(x, e, m))
    finally:
        for b in bufs:
            if b:
                b[0].close()
                b[1].close()


def test_server_admission_rejects_oversize(ctx):
    from hpdbscan_amd import _lib
    say("server admission")
    with _lib.Server(lanes=1, hbm_budget_bytes=1 << 20) as srv:
        x = np.zeros((100000, 3), np.float32)
        with pytest.raises(MemoryError):
            srv.submit(x, 0.3, 4)
        small = blobs(500, 3, 3, 0.2, 3.0, 1, np.float32)
        t = srv.submit(small, 0.3, 4)
        got, info = srv.wait(t)
        np.testing.assert_array_equal(got, ctx.cluster(small, 0.3, 4))
