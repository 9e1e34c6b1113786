"""Pins the CPU oracle (oracle/oracle.c) before anything is checked against it.

* every golden fixture written by the compiled reference (tests/golden/)
* the compiled reference itself on fresh random inputs, where oracle/_ref is
  built (this container)
* an independent closed-form brute force (tests/brute.py) on small inputs
* the summation-order and reciprocal fixtures really discriminate: the
  alternatives a port might use disagree with the reference on them
"""
from __future__ import annotations

import os

import numpy as np
import pytest

import oracle
from conftest import GOLDEN, golden_files

import brute

FIXTURES = [f for f in golden_files() if not f.startswith("sumtree_")]
SUMTREE = golden_files("sumtree_")


def load(name):
    z = np.load(os.path.join(GOLDEN, name))
    return z["data"], float(z["eps"]), int(z["min_pts"]), z


def test_fixture_inventory():
    # the families the generator writes
    names = set(golden_files())
    assert "p1_blobs_100k.npz" in names
    for tag in ("f32", "f64"):
        for d in range(1, 9):
            assert f"blobs_{tag}_d{d}.npz" in names
        for d in range(3, 9):
            assert f"sumtree_{tag}_d{d}.npz" in names
    assert any(n.startswith("edge_int_") for n in names)
    assert "lattice_f32.npz" in names


@pytest.mark.parametrize("name", FIXTURES)
def test_oracle_matches_reference_fixture(name):
    x, eps, mp, z = load(name)
    got = oracle.cluster(x, eps, mp)
    np.testing.assert_array_equal(got, z["labels"])


@pytest.mark.parametrize("name", SUMTREE)
def test_oracle_summation_order(name):
    z = np.load(os.path.join(GOLDEN, name))
    for x, e, want in zip(z["data"], z["eps"], z["labels"]):
        np.testing.assert_array_equal(oracle.cluster(x, float(e), 2), want)


def _alt_left_to_right(x):
    t = (x[0] - x[1]) * (x[0] - x[1])
    s = t.dtype.type(0)
    for v in t:
        s = s + v
    return s


@pytest.mark.parametrize("name", [n for n in SUMTREE if int(n.split("_d")[1][0]) >= 4])
def test_sumtree_fixtures_discriminate(name):
    """A left-to-right sum (the naive port) must get some of these pairs wrong."""
    z = np.load(os.path.join(GOLDEN, name))
    wrong = 0
    for x, e, want in zip(z["data"], z["eps"], z["labels"]):
        eps2 = np.float32(e * e)
        eps2 = eps2 if x.dtype == np.float32 else np.float64(eps2)
        naive_within = _alt_left_to_right(x) <= eps2
        wrong += int(naive_within != (want[0] < 0))
    assert wrong > 0


def test_lattice_discriminates_reciprocal():
    """floor(x / eps) differs from the reference's floor(x * (1/eps)) here."""
    x, eps, _, _ = load("lattice_f32.npz")
    e = np.float32(eps)
    mn = x.min(axis=0)
    div = np.floor((x - mn) / e)
    rcp = np.floor((x - mn) * (np.float32(1) / e))
    assert np.any(div != rcp)
    # and the reference's sorted order follows the reciprocal form
    cells, pos = oracle.ref_spatial(x, eps) if oracle.ref_available() else (None, None)
    if pos is not None:
        _, bpos = brute.sorted_positions(x, eps)
        np.testing.assert_array_equal(pos, bpos)


@pytest.mark.parametrize("seed", range(6))
@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_oracle_matches_brute_force(seed, dtype):
    rng = np.random.default_rng(seed)
    d = int(rng.integers(1, 6))
    n = int(rng.integers(200, 1500))
    k = int(rng.integers(2, 8))
    c = rng.uniform(0, 3, (k, d))
    x = np.concatenate([c[rng.integers(0, k, n - n // 5)] + rng.normal(0, 0.2, (n - n // 5, d)),
                        rng.uniform(0, 3, (n // 5, d))]).astype(dtype)
    x = x[rng.permutation(n)]
    eps = float(rng.uniform(0.15, 0.4))
    mp = int(rng.integers(1, 12))
    want = brute.labels(x, eps, mp)
    got = oracle.cluster(x, eps, mp)
    np.testing.assert_array_equal(got, want)


def test_oracle_thread_order_check():
    """With 8 threads the reference may change ambiguous border labels only."""
    x, eps, mp, z = load("ambiguous_f32.npz")
    full = oracle.cluster(x, eps, mp, full=True)
    assert full["stats"]["ambiguous_borders"] > 0
    mism, bad = oracle.check(x, eps, mp, z["labels_t8"])
    assert mism == 0 and bad == 0
    # a label moved to a cluster with no core nearby is caught
    lab = z["labels"].copy()
    i = int(np.nonzero(full["ambiguous"])[0][0])
    lab[i] = lab[i] + 10 ** 6
    mism, bad = oracle.check(x, eps, mp, lab)
    assert bad == 1


def test_oracle_invalid_arguments():
    x = np.zeros((4, 2), np.float32)
    with pytest.raises(ValueError):
        oracle.cluster(x, 0.0, 2)
    with pytest.raises(ValueError):
        oracle.cluster(x, -1.0, 2)


@pytest.mark.skipif(not oracle.ref_available(), reason="oracle/_ref not built (reference absent)")
@pytest.mark.parametrize("seed", range(8))
def test_oracle_matches_compiled_reference(seed):
    rng = np.random.default_rng(100 + seed)
    dtype = [np.float32, np.float64, np.int16, np.uint8][seed % 4]
    d = int(rng.integers(1, 7))
    n = int(rng.integers(500, 4000))
    if np.dtype(dtype).kind == "f":
        x = (rng.normal(0, 1, (n, d)) * rng.uniform(0.5, 3)).astype(dtype)
        eps = float(rng.uniform(0.1, 0.6))
    else:
        x = rng.integers(0, 60, (n, d)).astype(dtype)
        eps = float(rng.uniform(0.5, 6))
    mp = int(rng.integers(1, 10))
    np.testing.assert_array_equal(oracle.cluster(x, eps, mp), oracle.ref_cluster(x, eps, mp, 1))
    cells, pos = oracle.ref_spatial(x, eps)
    full = oracle.cluster(x, eps, mp, full=True)
    np.testing.assert_array_equal(full["sorted_pos"], pos)
    np.testing.assert_array_equal(full["cell_of_sorted"], cells)
    assert full["stats"]["sum_candidates"] == oracle.ref_candidates(x, eps)
