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
    rng.shuffle(out, axis=0)
    return out


def cloud(n: int, d: int, k: int, rng: np.random.Generator, sigma: float = 0.2,
          box: float = 10.0, noise: float = 0.1) -> np.ndarray:
    """K Gaussian blobs (centres uniform in [0, box]^d, isotropic sigma) plus a
    `noise` fraction of uniform points; float32, shuffled."""
    centres = rng.uniform(0.0, box, size=(k, d))
    n_noise = int(round(n * noise))
    which = rng.integers(0, k, size=n - n_noise)
    pts = centres[which] + rng.normal(0.0, sigma, size=(n - n_noise, d))
    x = np.concatenate([pts, rng.uniform(0.0, box, size=(n_noise, d))])
    return _finish(x, rng)


def make(name: str, n: int | None = None) -> tuple[np.ndarray, float, int]:
    """(rows, eps, min_points) of a named workload, optionally with fewer rows."""
    w = WORKLOADS[name]
    fn = {"blobs_100k": blobs_100k, "blobs_10m": blobs_10m, "skewed_2d": skewed_2d}[name]
    x = fn(n=n or w["n"], seed=w["seed"])
    return x, w["eps"], w["min_points"]
