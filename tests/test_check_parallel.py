"""Pins oracle.check_parallel (the OpenMP checker of the closed form that the
GPU tests use on inputs too large for the one-thread restatement) against
oracle.check and the reference's fixtures: on every input both checkers
count the same mismatches and bad ambiguous borders, for the reference's own
labels and for labels with planted mistakes."""
from __future__ import annotations

import os

import numpy as np
import pytest

import oracle
from conftest import GOLDEN, golden_files

FIXTURES = [f for f in golden_files() if not f.startswith("sumtree_")]


@pytest.mark.parametrize("name", FIXTURES)
def test_fixture_labels_pass(name):
    z = np.load(os.path.join(GOLDEN, name))
    x, eps, mp, want = z["data"], float(z["eps"]), int(z["min_pts"]), z["labels"]
    r = oracle.check_parallel(x, eps, mp, want, threads=4)
    assert r["mismatches"] == 0 and r["bad_ambiguous"] == 0
    assert r["cores"] == int(np.count_nonzero(want < 0))
    full = oracle.cluster(x, eps, mp, full=True)
    assert r["ambiguous"] == full["stats"]["ambiguous_borders"]


def blobs(n, d, k, sigma, box, seed, dtype, noise=0.3):
    rng = np.random.default_rng(seed)
    c = rng.uniform(0, box, (k, d))
    nb = int(n * (1 - noise))
    x = np.concatenate([c[rng.integers(0, k, nb)] + rng.normal(0, sigma, (nb, d)),
                        rng.uniform(0, box, (n - nb, d))])
    return x[rng.permutation(n)].astype(dtype)


def planted(labels, amb, rng):
    """Copies of `labels` with one planted mistake of each kind."""
    out = []
    core = np.nonzero(labels < 0)[0]
    noise = np.nonzero(labels == 0)[0]
    border = np.nonzero((labels > 0) & ~amb)[0]
    ids = np.unique(-labels[core])
    if len(core):
        g = labels.copy()
        g[rng.choice(core)] *= -1  # a core reported as a border
        out.append(g)
    if len(core) and len(ids) > 1:
        g = labels.copy()
        i = rng.choice(core)
        g[i] = -ids[ids != -labels[i]][0]  # a core in another cluster
        out.append(g)
        g = labels.copy()
        g[labels == -ids[1]] = -ids[0]  # two clusters merged
        g[labels == ids[1]] = ids[0]
        out.append(g)
    if len(noise):
        g = labels.copy()
        g[rng.choice(noise)] = ids[0] if len(ids) else 1  # noise reported as a border
        out.append(g)
    if len(border):
        g = labels.copy()
        g[rng.choice(border)] = 0  # a border reported as noise
        out.append(g)
    a = np.nonzero(amb)[0]
    if len(a):
        g = labels.copy()
        g[a[0]] = ids.max() + 7  # an ambiguous border in a cluster that does not exist
        out.append(g)
        g = labels.copy()
        g[a] = 0
        out.append(g)
    return out


CASES = [(n, d, dt, s) for (n, d) in ((3000, 1), (4000, 2), (5000, 3), (3000, 4), (1500, 6))
         for dt in (np.float32, np.float64) for s in (0, 1)]


@pytest.mark.parametrize("n,d,dtype,seed", CASES)
def test_agrees_with_check(n, d, dtype, seed):
    rng = np.random.default_rng(seed)
    # 14-80 clusters with noise and border rows in every case
    box = {1: 60.0, 2: 12.0, 3: 7.0, 4: 6.0, 6: 6.0}[d]
    x = blobs(n, d, 25, 0.35, box, seed + 10 * d, dtype)
    eps, mp = {1: (0.05, 4), 2: (0.25, 6), 3: (0.35, 8), 4: (0.6, 8), 6: (1.2, 6)}[d]
    full = oracle.cluster(x, eps, mp, full=True)
    want, amb = full["labels"], full["ambiguous"]
    r = oracle.check_parallel(x, eps, mp, want, threads=3)
    assert (r["mismatches"], r["bad_ambiguous"]) == (0, 0)
    assert r["cores"] == full["stats"]["core_points"]
    assert r["ambiguous"] == full["stats"]["ambiguous_borders"]
    kinds = planted(want, amb, rng)
    assert len(kinds) >= 5
    for g in kinds:
        got = oracle.check_parallel(x, eps, mp, g, threads=3)
        ref = oracle.check(x, eps, mp, g)
        assert (got["mismatches"], got["bad_ambiguous"]) == ref
        assert ref[0] + ref[1] > 0


@pytest.mark.parametrize("dtype", [np.int16, np.uint8, np.int64])
def test_integer_dtypes(dtype):
    rng = np.random.default_rng(2)
    lo, hi = (0, 120) if np.dtype(dtype).kind == "u" else (-60, 60)
    x = rng.integers(lo, hi, size=(3000, 2)).astype(dtype)
    for eps, mp in ((0.5, 1), (4.0, 5), (6.5, 14)):
        full = oracle.cluster(x, eps, mp, full=True)
        r = oracle.check_parallel(x, eps, mp, full["labels"], threads=2)
        assert (r["mismatches"], r["bad_ambiguous"]) == (0, 0)
        assert r["cores"] == full["stats"]["core_points"]
        for g in planted(full["labels"], full["ambiguous"], rng):
            got = oracle.check_parallel(x, eps, mp, g, threads=2)
            assert (got["mismatches"], got["bad_ambiguous"]) == oracle.check(x, eps, mp, g)


def test_thread_counts_agree():
    x = blobs(20000, 2, 40, 0.3, 10.0, 5, np.float32)
    want = oracle.cluster(x, 0.2, 6)
    g = want.copy()
    g[::997] = 0
    res = {t: oracle.check_parallel(x, 0.2, 6, g, threads=t) for t in (1, 2, 8)}
    assert res[1] == res[2] == res[8]
    assert (res[1]["mismatches"], res[1]["bad_ambiguous"]) == oracle.check(x, 0.2, 6, g)


def test_empty_and_single_row():
    r = oracle.check_parallel(np.zeros((0, 3), np.float32), 0.5, 2, np.zeros(0, np.int64))
    assert r == {"mismatches": 0, "bad_ambiguous": 0, "cores": 0, "ambiguous": 0}
    r = oracle.check_parallel(np.zeros((1, 3), np.float32), 0.5, 1, np.array([-1]))
    assert r["mismatches"] == 0 and r["cores"] == 1
    r = oracle.check_parallel(np.zeros((1, 3), np.float32), 0.5, 2, np.array([-1]))
    assert r["mismatches"] == 1 and r["cores"] == 0
