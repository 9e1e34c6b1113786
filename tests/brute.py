"""Closed-form DBSCAN labels by brute force (numpy + scipy), for small inputs.

An independent formulation of what the reference computes, used to pin the
oracle (oracle/oracle.c) and the HIP path on inputs of a few thousand rows:

  cells      idx_d = floor((x_d - min_d) * f32(1/eps)); dimension order by
             ascending cell count (stable); the cell id orders the rows
             (stable), giving every row its sorted position
  candidate  rows whose cell indices differ by at most one in every dimension
  within     squared distance in the reference build's summation order
             <= f32(eps*eps)
  core       |candidates within eps| >= minPts
  cluster    connected components of the core-within-core graph; id = 1 +
             smallest sorted position of a core in the component
  border     non-core row within eps of a core.  With one thread the
             reference assigns the cluster of the adjacent core c with the
             smallest f(c), where f(c) = min(c, f(q) over adjacent cores q
             that precede c in sorted order)
"""
from __future__ import annotations

import numpy as np
from scipy.sparse import coo_matrix
from scipy.sparse.csgraph import connected_components

F32_MIN = np.float32(np.finfo(np.float32).tiny)


def _tree(t: np.ndarray, dtype) -> np.ndarray:
    """Sum t[..., d] over the last axis in the reference build's order."""
    d = t.shape[-1]
    if dtype == np.float32:
        s = np.zeros(t.shape[:-1], dtype=np.float32)
        base = 0
        if d >= 4:
            a = [t[..., j].copy() for j in range(4)]
            for k in range(1, d // 4):
                for j in range(4):
                    a[j] = a[j] + t[..., 4 * k + j]
            s = (a[0] + a[2]) + (a[1] + a[3])
            base = (d // 4) * 4
        if d - base >= 2:
            s = s + (t[..., base] + t[..., base + 1])
            base += 2
        if d - base == 1:
            s = s + t[..., base]
        return s
    if d == 1:
        return t[..., 0]
    a0, a1 = t[..., 0].copy(), t[..., 1].copy()
    for k in range(1, d // 2):
        a0 = a0 + t[..., 2 * k]
        a1 = a1 + t[..., 2 * k + 1]
    s = a0 + a1
    if d % 2:
        s = s + t[..., d - 1]
    return s


def sorted_positions(x: np.ndarray, eps: float):
    """(cell index per row and dimension, sorted position per row)."""
    eps32 = np.float32(eps)
    if x.dtype == np.float32:
        rcp = np.float32(1.0) / eps32
        mn = x.min(axis=0)
        mx = np.maximum(x.max(axis=0), F32_MIN)
        cells = (np.ceil((mx - mn) * rcp)).astype(np.uint64) + 1
        idx = np.floor((x - mn) * rcp).astype(np.uint64)
    else:
        rcp = 1.0 / np.float64(eps32)
        mn = x.min(axis=0)
        mx = np.maximum(x.max(axis=0), np.finfo(np.float64).tiny)
        cells = (np.ceil((mx - mn) * rcp)).astype(np.uint64) + 1
        idx = np.floor((x - mn) * rcp).astype(np.uint64)
    order = np.argsort(cells, kind="stable")
    key = np.zeros(len(x), dtype=np.uint64)
    acc = np.uint64(1)
    for dim in order:
        key += idx[:, dim] * acc
        acc *= cells[dim]
    srt = np.argsort(key, kind="stable")
    pos = np.empty(len(x), dtype=np.int64)
    pos[srt] = np.arange(len(x))
    return idx, pos


def labels(x: np.ndarray, eps: float, min_pts: int) -> np.ndarray:
    n, d = x.shape
    idx, pos = sorted_positions(x, eps)
    eps2 = np.float32(np.float32(eps) * np.float32(eps))
    eps2 = eps2 if x.dtype == np.float32 else np.float64(eps2)
    diff = x[:, None, :] - x[None, :, :]
    within = _tree(diff * diff, x.dtype) <= eps2
    adj_cell = np.all(np.abs(idx[:, None, :].astype(np.int64) - idx[None, :, :].astype(np.int64)) <= 1,
                      axis=2)
    nb = within & adj_cell
    core = nb.sum(axis=1) >= min_pts
    cc = nb & core[:, None] & core[None, :]
    ii, jj = np.nonzero(cc)
    g = coo_matrix((np.ones(len(ii)), (ii, jj)), shape=(n, n))
    _, comp = connected_components(g, directed=False)
    big = np.iinfo(np.int64).max
    comp_min = np.full(comp.max() + 1 if n else 0, big, dtype=np.int64)
    np.minimum.at(comp_min, comp[core], pos[core])
    out = np.zeros(n, dtype=np.int64)
    out[core] = -(comp_min[comp[core]] + 1)
    # one-thread border rule: f over cores in sorted order
    f = np.full(n, big, dtype=np.int64)
    by_pos = np.argsort(pos)
    for i in by_pos:
        if not core[i]:
            continue
        prev = np.nonzero(cc[i] & (pos < pos[i]))[0]
        f[i] = min(pos[i], f[prev].min()) if len(prev) else pos[i]
    for i in np.nonzero(~core)[0]:
        adj = np.nonzero(nb[i] & core)[0]
        if len(adj) == 0:
            continue
        c = adj[np.argmin(f[adj])]
        out[i] = comp_min[comp[c]] + 1
    return out
