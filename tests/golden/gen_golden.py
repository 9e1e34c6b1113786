"""Writes the golden fixtures of tests/golden/ from the compiled reference.

    make -C oracle all && python tests/golden/gen_golden.py

Every fixture is an .npz holding the inputs and the labels that the
reference (Markus-Goetz/hpdbscan, compiled from its own headers by
oracle/Makefile with its own -Ofast flags) returns from
HPDBSCAN(eps, minPts).cluster<T>(data, n, d, threads) with threads = 1
(key ``labels``) and, for some fixtures, threads = 8 (``labels_t8``).
The reference is only run here, when the fixtures are (re)generated; the
tests read the .npz files.

Fixture families:
  blobs_<dt>_d<D>     Gaussian blobs + noise, D = 1..8, float32 / float64
  p1_blobs_100k       the 100k x 3-D workload (20 blobs, eps 0.3, minPts 4)
  sumtree_<dt>_d<D>   two-point inputs at minPts 2 whose eps is chosen so
                      that the squared distance evaluated in the reference
                      build's summation order and in other plausible orders
                      fall on different sides of EPS2
  lattice_f32         coordinates at exact multiples of eps, where
                      floor(x / eps) and floor(x * (1/eps)) differ
  edge_*              FLT_MIN max quirk, flush-to-zero cell counts,
                      duplicates, minPts 0/1/2, n = 1/2, one dense cell,
                      all noise, integer dtypes (EPS2 truncation, wrap)
  ambiguous_f32       border points next to two clusters, labels at 1 and
                      8 threads
"""
from __future__ import annotations

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import oracle  # noqa: E402
from hpdbscan_amd import datagen  # noqa: E402


def save(name: str, data: np.ndarray, eps: float, min_pts: int, threads8: bool = False,
         **extra) -> None:
    data = np.ascontiguousarray(data)
    labels = oracle.ref_cluster(data, eps, min_pts, threads=1)
    out = {"data": data, "eps": np.float32(eps), "min_pts": np.uint64(min_pts), "labels": labels}
    if threads8:
        out["labels_t8"] = oracle.ref_cluster(data, eps, min_pts, threads=8)
    out.update(extra)
    np.savez_compressed(os.path.join(HERE, name + ".npz"), **out)


def blobs(n, d, k, sigma, box, seed, dtype, noise=0.1):
    rng = np.random.default_rng(seed)
    c = rng.uniform(0, box, (k, d))
    nb = int(n * (1 - noise))
    x = c[rng.integers(0, k, nb)] + rng.normal(0, sigma, (nb, d))
    y = rng.uniform(0, box, (n - nb, d))
    a = np.concatenate([x, y])
    return a[rng.permutation(n)].astype(dtype)


# --------------------------------------------------------------------------
# summation-order discriminators
# --------------------------------------------------------------------------

def tree_f32(t):
    """The reference build's float order (see oracle/oracle.c dist2_f32)."""
    t = [np.float32(v) for v in t]
    d = len(t)
    s = np.float32(0)
    base = 0
    if d >= 4:
        a = list(t[:4])
        for k in range(1, d // 4):
            for j in range(4):
                a[j] = np.float32(a[j] + t[4 * k + j])
        s = np.float32(np.float32(a[0] + a[2]) + np.float32(a[1] + a[3]))
        base = (d // 4) * 4
    if d - base >= 2:
        s = np.float32(s + np.float32(t[base] + t[base + 1]))
        base += 2
    if d - base == 1:
        s = np.float32(s + t[base])
    return s


def tree_f64(t):
    t = [np.float64(v) for v in t]
    d = len(t)
    if d == 1:
        return t[0]
    a0, a1 = t[0], t[1]
    for k in range(1, d // 2):
        a0 = a0 + t[2 * k]
        a1 = a1 + t[2 * k + 1]
    s = a0 + a1
    if d % 2:
        s = s + t[d - 1]
    return s


def alternatives(t, ft):
    """Other summation orders a port might plausibly use."""
    t = [ft(v) for v in t]
    out = []
    s = ft(0)
    for v in t:  # left to right
        s = ft(s + v)
    out.append(s)
    s = ft(0)
    for v in reversed(t):  # right to left
        s = ft(s + v)
    out.append(s)
    # pairwise
    level = list(t)
    while len(level) > 1:
        nxt = [ft(level[i] + level[i + 1]) for i in range(0, len(level) - 1, 2)]
        if len(level) % 2:
            nxt.append(level[-1])
        level = nxt
    out.append(level[0])
    # even/odd lanes (2-wide) and 4-lane without the remainder pair
    for w in (2, 4, 8):
        lanes = [ft(0)] * w
        for i, v in enumerate(t):
            lanes[i % w] = ft(lanes[i % w] + v)
        s = ft(0)
        for v in lanes:
            s = ft(s + v)
        out.append(s)
    return out


def sumtree_cases(d, dtype, rng, want=12):
    """Pairs whose within-eps decision depends on the summation order.

    One squared term is E = e*e with e a 12-bit float32 (so E is exactly the
    reference's EPS2 = f32(eps*eps) for eps = e); the other terms are exact
    squares of tiny deltas worth 0.1 - 0.45 units in the last place of E, so
    the sum is E or the next value up depending on how they are grouped."""
    ft = np.float32 if dtype == np.float32 else np.float64
    ours = tree_f32 if dtype == np.float32 else tree_f64
    datas, epss = [], []
    tries = 0
    while len(datas) < want and tries < 200000:
        tries += 1
        e = np.float32(rng.integers(2048, 4096) * 2.0 ** -14)  # 12-bit significand, [0.125, 0.25)
        E = np.float64(e) * np.float64(e)
        ulp = np.spacing(ft(E))
        j0 = int(rng.integers(0, d))
        delta = np.zeros(d, dtype=np.float64)
        delta[j0] = np.float64(e)
        for j in range(d):
            if j == j0:
                continue
            frac = rng.uniform(0.05, min(0.45, 1.2 / (d - 1)))
            # tiny delta with a 10-bit significand: its square is exact in ft
            tgt = np.sqrt(frac * np.float64(ulp))
            ex = np.floor(np.log2(tgt)) - 9
            delta[j] = np.round(tgt / 2.0 ** ex) * 2.0 ** ex
        delta = delta.astype(dtype)
        tt = [ft(v) * ft(v) for v in delta]
        s_ours = ours(tt)
        alts = alternatives(tt, ft)
        if all(a == s_ours for a in alts):
            continue
        # for d >= 4 the first half must catch the most likely mistake, a
        # left-to-right sum (for d <= 3 that is the reference's own order)
        if d >= 4 and len(datas) < want // 2 and alts[0] == s_ours:
            continue
        if not (s_ours == ft(E) or any(a == ft(E) for a in alts)):
            continue
        p = np.zeros(d, dtype=dtype)
        q = (-delta).astype(dtype)
        datas.append(np.stack([p, q]))
        epss.append(e)
    return datas, epss


def gen_sumtree():
    for dtype, tag in ((np.float32, "f32"), (np.float64, "f64")):
        for d in range(2, 9):
            rng = np.random.default_rng(1000 + d + (0 if tag == "f32" else 100))
            datas, epss = sumtree_cases(d, dtype, rng)
            if not datas:
                continue
            labels = np.stack([oracle.ref_cluster(x, float(e), 2, threads=1)
                               for x, e in zip(datas, epss)])
            np.savez_compressed(os.path.join(HERE, f"sumtree_{tag}_d{d}.npz"),
                                data=np.stack(datas), eps=np.array(epss, dtype=np.float32),
                                min_pts=np.uint64(2), labels=labels)
            print(f"sumtree {tag} d={d}: {len(datas)} pairs, within={int((labels[:, 0] < 0).sum())}")


def gen_lattice():
    # coordinates k * eps for integer k: x * (1/eps) and x / eps round apart
    eps = np.float32(0.3)
    k = np.arange(0, 2000, dtype=np.float32)
    xs = (k * eps).astype(np.float32)
    rng = np.random.default_rng(7)
    pts = np.stack([xs, rng.permutation(xs), np.float32(0.5) * rng.permutation(xs)], axis=1)
    save("lattice_f32", pts, float(eps), 3)


def gen_edges():
    rng = np.random.default_rng(11)
    # all coordinates of dimension 1 negative: max starts at FLT_MIN
    a = blobs(1500, 3, 5, 0.2, 4.0, 12, np.float32)
    a[:, 1] = -np.abs(a[:, 1]) - 0.5
    save("edge_negdim_f32", a, 0.3, 4)
    # an all-zero dimension with eps > 1: FLT_MIN * (1/eps) flushes to zero
    b = blobs(1500, 3, 5, 1.0, 30.0, 13, np.float32)
    b[:, 2] = 0.0
    save("edge_zerodim_eps2_f32", b, 2.0, 4)
    c = b.copy()
    save("edge_zerodim_eps05_f32", c, 0.5, 4)
    # duplicates
    base = blobs(300, 2, 4, 0.3, 5.0, 14, np.float32)
    dup = np.concatenate([base, base, base[:100]])
    save("edge_duplicates_f32", dup[rng.permutation(len(dup))], 0.2, 3)
    x = blobs(2000, 3, 6, 0.25, 6.0, 15, np.float32)
    for mp in (0, 1, 2):
        save(f"edge_minpts{mp}_f32", x, 0.3, mp)
    save("edge_n1_f32", np.array([[1.5, -2.0, 3.25]], dtype=np.float32), 0.5, 1)
    save("edge_n1_noise_f32", np.array([[1.5, -2.0, 3.25]], dtype=np.float32), 0.5, 2)
    save("edge_n2_f32", np.array([[0.0, 0.0], [0.3, 0.4]], dtype=np.float32), 0.5, 2)
    dense = rng.uniform(0.0, 0.05, size=(3000, 3)).astype(np.float32)
    save("edge_dense_cell_f32", dense, 1.0, 10)
    save("edge_all_noise_f32", x, 0.3, 10 ** 6)
    save("edge_large_eps_f32", x, 1000.0, 5)
    # integer dtypes: eps < 1 truncates EPS2 to 0; uint8 distances wrap
    for dt in (np.int8, np.int16, np.int32, np.int64, np.uint8, np.uint16, np.uint32, np.uint64):
        lo, hi = (0, 100) if np.dtype(dt).kind == "u" else (-50, 50)
        xi = rng.integers(lo, hi, size=(1200, 3)).astype(dt)
        name = np.dtype(dt).name
        save(f"edge_int_{name}_eps3", xi, 3.0, 4)
        save(f"edge_int_{name}_eps05", xi, 0.5, 2)
    xw = rng.integers(0, 256, size=(1500, 2)).astype(np.uint8)
    save("edge_int_uint8_wrap", xw, 12.0, 5)


def main():
    if not oracle.ref_available():
        sys.exit("oracle/_ref is not built: run `make -C oracle all` where the reference is mounted")
    for dt, tag in ((np.float32, "f32"), (np.float64, "f64")):
        for d in range(1, 9):
            n = 3000 if d <= 4 else 1500
            x = blobs(n, d, 8, 0.25, 4.0 + d, 100 * d + (1 if tag == "f64" else 0), dt)
            save(f"blobs_{tag}_d{d}", x, 0.35 if d <= 4 else 0.6, 4)
    x = datagen.blobs_100k()
    save("p1_blobs_100k", x, 0.3, 4)
    gen_sumtree()
    gen_lattice()
    gen_edges()
    amb = blobs(20000, 2, 40, 0.3, 12.0, 99, np.float32, noise=0.2)
    save("ambiguous_f32", amb, 0.25, 8, threads8=True)
    total = sum(os.path.getsize(os.path.join(HERE, f)) for f in os.listdir(HERE) if f.endswith(".npz"))
    print(f"fixtures: {total / 1e6:.2f} MB")


if __name__ == "__main__":
    main()
