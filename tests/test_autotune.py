"""The work-split autotuner's selection logic (CPU, with a stand-in context)
and, on the GPU, a real sweep whose candidates must all give equal labels."""
from __future__ import annotations

import numpy as np
import pytest

from hpdbscan_amd import autotune


class FakeContext:
    """Records the split; 'runs' take a time that depends on the split."""

    def __init__(self, cost, labels_for=None):
        self.cost = cost
        self.labels_for = labels_for or (lambda sm, sc: np.arange(5))
        self.split = (8, False)
        self.clock = 0.0

    def set_work_split(self, sm, sc):
        self.split = (sm, sc)

    def synchronize(self):
        pass


def test_sweep_picks_fastest(monkeypatch):
    ctx = FakeContext(cost={(0, False): 3.0, (8, False): 1.0, (16, True): 2.0})
    t = [0.0]
    monkeypatch.setattr(autotune.time, "perf_counter", lambda: t[0])

    def run():
        t[0] += ctx.cost[ctx.split]

    res = autotune._sweep(ctx, run, lambda: np.arange(5), 3, [(0, False), (8, False), (16, True)])
    assert res["best"] == {"sparse_max": 8, "sparse_core": False}
    assert [r["ms"] for r in res["results"]] == [3000.0, 1000.0, 2000.0]


def test_sweep_rejects_label_change():
    ctx = FakeContext(cost={})
    calls = iter([np.arange(5), np.arange(5) + 1])
    with pytest.raises(RuntimeError, match="changed the labels"):
        autotune._sweep(ctx, lambda: None, lambda: next(calls), 1, [(0, False), (8, False)])


def test_apply_accepts_result_or_best():
    ctx = FakeContext(cost={})
    autotune.apply(ctx, {"best": {"sparse_max": 16, "sparse_core": True}})
    assert ctx.split == (16, True)
    autotune.apply(ctx, {"sparse_max": 4, "sparse_core": False})
    assert ctx.split == (4, False)


@pytest.mark.gpu
def test_autotune_on_gpu():
    from hpdbscan_amd import _lib, datagen
    x = datagen.blobs_100k(n=60_000, seed=7)
    print("[gpu-case] autotune blobs 60k", flush=True)
    ctx = _lib.Context(0)
    try:
        res = autotune.tune(x, 0.3, 4, ctx=ctx, reps=1)
        assert (res["best"]["sparse_max"], res["best"]["sparse_core"]) in autotune.CANDIDATES
        autotune.apply(ctx, res)
        assert ctx.work_split() == (res["best"]["sparse_max"], res["best"]["sparse_core"])
    finally:
        ctx.close()
