"""Seeded synthetic inputs of the benchmark workloads (numpy PCG64 streams).

Each generator returns C-contiguous float32 rows in shuffled order.

* ``blobs_100k``      100k x 3-D: 20 isotropic Gaussian blobs (centres uniform
                      in [0,10]^3, sigma 0.2, equal weights) + 10 % uniform
                      noise over [0,10]^3; eps 0.3, minPoints 4.
* ``blobs_10m``       10M x 3-D: 1000 anisotropic Gaussian blobs (centres
                      uniform in [0,100]^3, per-axis sigma log-uniform in
                      [0.1, 1.0], weights Dirichlet(1)) + 5 % uniform noise;
                      eps 0.5, minPoints 16.
* ``skewed_2d``       100M x 2-D in [0,1]^2: 80 % in 500 Gaussian clusters
                      (sigma log-normal(ln 0.005, 0.7), Zipf(1.2) weights),
                      20 % uniform noise; eps 0.01, minPoints 32.
* ``cloud``           one point cloud of the batched / serving workloads: a
                      mixture of K Gaussian blobs + 10 % noise in [0,10]^d.

``n`` can be scaled down for tests; the density then drops with it.
"""
from __future__ import annotations

import numpy as np

WORKLOADS = {
    "blobs_100k": {"n": 100_000, "d": 3, "eps": 0.3, "min_points": 4, "seed": 1},
    "blobs_10m": {"n": 10_000_000, "d": 3, "eps": 0.5, "min_points": 16, "seed": 2},
    "skewed_2d": {"n": 100_000_000, "d": 2, "eps": 0.01, "min_points": 32, "seed": 3},
}


def _finish(x: np.ndarray, rng: np.random.Generator) -> np.ndarray:
    perm = rng.permutation(x.shape[0])
    return np.ascontiguousarray(x[perm], dtype=np.float32)


def blobs_100k(n: int = 100_000, seed: int = 1) -> np.ndarray:
    rng = np.random.default_rng(seed)
    centres = rng.uniform(0.0, 10.0, size=(20, 3))
    n_noise = n // 10
    n_blob = n - n_noise
    which = rng.integers(0, 20, size=n_blob)
    pts = centres[which] + rng.normal(0.0, 0.2, size=(n_blob, 3))
    noise = rng.uniform(0.0, 10.0, size=(n_noise, 3))
    return _finish(np.concatenate([pts, noise]), rng)


def blobs_10m(n: int = 10_000_000, seed: int = 2) -> np.ndarray:
    rng = np.random.default_rng(seed)
    k = 1000
    centres = rng.uniform(0.0, 100.0, size=(k, 3))
    sigma = np.exp(rng.uniform(np.log(0.1), np.log(1.0), size=(k, 3)))
    weights = rng.dirichlet(np.ones(k))
    n_noise = n // 20
    counts = rng.multinomial(n - n_noise, weights)
    which = np.repeat(np.arange(k), counts)
    out = np.empty((n, 3), dtype=np.float32)
    out[: n - n_noise] = centres[which] + rng.normal(0.0, 1.0, size=(n - n_noise, 3)) * sigma[which]
    out[n - n_noise:] = rng.uniform(0.0, 100.0, size=(n_noise, 3))
    return _finish(out, rng)


def skewed_2d(n: int = 100_000_000, seed: int = 3, chunk: int = 10_000_000) -> np.ndarray:
    rng = np.random.default_rng(seed)
    k = 500
    centres = rng.uniform(0.0, 1.0, size=(k, 2))
    sigma = rng.lognormal(np.log(0.005), 0.7, size=k)
    w = np.arange(1, k + 1, dtype=np.float64) ** -1.2
    w /= w.sum()
    n_noise = n // 5
    counts = rng.multinomial(n - n_noise, w)
    out = np.empty((n, 2), dtype=np.float32)
    pos = 0
    for i in np.nonzero(counts)[0]:
        c = int(counts[i])
        done = 0
        while done < c:
            m = min(chunk, c - done)
            blk = centres[i] + rng.normal(0.0, sigma[i], size=(m, 2))
            out[pos:pos + m] = np.clip(blk, 0.0, 1.0)
            pos += m
            done += m
    done = 0
    while done < n_noise:
        m = min(chunk, n_noise - done)
        out[pos:pos + m] = rng.uniform(0.0, 1.0, size=(m, 2))
        pos += m
        done += m
    # a row of two float32 is one 8-byte word: shuffling the words is the
    # same permutation as shuffling the rows (same draws), 30x faster
    rng.shuffle(out.view(np.int64).reshape(-1))
    return out


def cloud(n: int, d: int, k: int, rng: np.random.Generator, sigma: float = 0.2,
          box: float = 10.0, noise: float = 0.1) -> np.ndarray:
    """K Gaussian blobs (centres uniform in [0, box]^d, isotropic sigma) plus a
    `noise` fraction of uniform points; float32, shuffled."""
    centres = rng.uniform(0.0, box, size=(k, d)).astype(np.float32)
    n_noise = int(round(n * noise))
    x = np.empty((n, d), dtype=np.float32)
    blk = x[: n - n_noise]
    np.multiply(rng.standard_normal((n - n_noise, d), dtype=np.float32), np.float32(sigma), out=blk)
    blk += centres[rng.integers(0, k, size=n - n_noise)]
    np.multiply(rng.random((n_noise, d), dtype=np.float32), np.float32(box), out=x[n - n_noise:])
    return x[rng.permutation(n)]


def make(name: str, n: int | None = None) -> tuple[np.ndarray, float, int]:
    """(rows, eps, min_points) of a named workload, optionally with fewer rows."""
    w = WORKLOADS[name]
    fn = {"blobs_100k": blobs_100k, "blobs_10m": blobs_10m, "skewed_2d": skewed_2d}[name]
    x = fn(n=n or w["n"], seed=w["seed"])
    return x, w["eps"], w["min_points"]


def batch_clouds(n_jobs: int = 4096, seed: int = 4, n_min: int = 1_000, n_max: int = 200_000):
    """The batched workload: n_jobs clouds in 3-D, sizes log-uniform in
    [n_min, n_max], each 5-50 Gaussian blobs (centres uniform in [0,10]^3,
    sigma 0.2, equal weights) + 10 % uniform noise in [0,10]^3; eps uniform
    in [0.2, 0.5], minPoints in {4, 8, 16}.  Within a cloud the blob rows
    come first, then the noise rows.  Returns (rows, offsets, eps,
    min_points)."""
    rng = np.random.default_rng(seed)
    sizes = np.round(np.exp(rng.uniform(np.log(n_min), np.log(n_max), size=n_jobs))).astype(np.int64)
    offsets = np.zeros(n_jobs + 1, dtype=np.int64)
    offsets[1:] = np.cumsum(sizes)
    rows = np.empty((int(offsets[-1]), 3), dtype=np.float32)
    for j in range(n_jobs):
        n = int(sizes[j])
        nn = int(round(n * 0.1))
        nb = n - nn
        k = int(rng.integers(5, 51))
        centres = rng.uniform(0.0, 10.0, size=(k, 3)).astype(np.float32)
        o = int(offsets[j])
        blk = rows[o:o + nb]
        np.multiply(rng.standard_normal((nb, 3), dtype=np.float32), np.float32(0.2), out=blk)
        blk += centres[rng.integers(0, k, size=nb)]
        np.multiply(rng.random((nn, 3), dtype=np.float32), np.float32(10.0), out=rows[o + nb:o + n])
    eps = rng.uniform(0.2, 0.5, size=n_jobs).astype(np.float32)
    mps = rng.choice(np.array([4, 8, 16], dtype=np.uint64), size=n_jobs)
    return rows, offsets, eps, mps


def serving_stream(n_requests: int = 400, rate: float = 200.0, seed: int = 5,
                   n_min: int = 10_000, n_max: int = 5_000_000, pool_per_d: int = 6):
    """The serving workload: Poisson arrivals at `rate` per second; sizes
    log-uniform in [n_min, n_max]; d uniform in {2, 3, 4}; eps uniform in
    [0.2, 0.5]; minPoints uniform in 4..32.  Each request's rows are the
    first n rows of one of `pool_per_d` shuffled blob mixtures (5-50 blobs,
    sigma 0.2, 10 % noise in [0,10]^d) of its dimension.
    Returns a list of (arrival_s, rows, eps, min_points)."""
    rng = np.random.default_rng(seed)
    pools = {d: [cloud(n_max, d, int(rng.integers(5, 51)), rng) for _ in range(pool_per_d)]
             for d in (2, 3, 4)}
    t = np.cumsum(rng.exponential(1.0 / rate, size=n_requests))
    out = []
    for i in range(n_requests):
        d = int(rng.choice([2, 3, 4]))
        n = int(round(np.exp(rng.uniform(np.log(n_min), np.log(n_max)))))
        base = pools[d][int(rng.integers(0, pool_per_d))]
        out.append((float(t[i]), base[:n], float(rng.uniform(0.2, 0.5)), int(rng.integers(4, 33))))
    return out
