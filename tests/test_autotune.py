"""The autotuner's selection logic (CPU, with a stand-in context) and, on
the GPU, a real sweep whose candidates must all give equal labels."""
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
        self.box = 0
        self.clock = 0.0

    def set_work_split(self, sm, sc):
        self.split = (sm, sc)

    def set_box_split(self, k):
        self.box = k

    def synchronize(self):
        pass


def test_sweep_picks_fastest(monkeypatch):
    ctx = FakeContext(cost={(0, False): 3.0, (8, False): 1.0, (16, True): 2.0})
    t = [0.0]
    monkeypatch.setattr(autotune.time, "perf_counter", lambda: t[0])

    def run():
        t[0] += ctx.cost[ctx.split]

    cands = [{"sparse_max": a, "sparse_core": b} for a, b in [(0, False), (8, False), (16, True)]]
    res = autotune._sweep(ctx, run, lambda: np.arange(5), 3, cands)
    assert res["best"] == {"sparse_max": 8, "sparse_core": False}
    assert [r["ms"] for r in res["results"]] == [3000.0, 1000.0, 2000.0]


def test_sweep_rejects_label_change():
    ctx = FakeContext(cost={})
    calls = iter([np.arange(5), np.arange(5) + 1])
    with pytest.raises(RuntimeError, match="changed the labels"):
        autotune._sweep(ctx, lambda: None, lambda: next(calls), 1,
                        [{"sparse_max": 0, "sparse_core": False},
                         {"sparse_max": 8, "sparse_core": False}])


def test_two_stage_box_split_then_work_split(monkeypatch):
    """Box splits are timed under the default work split, then work splits
    under the fastest box split; both stages must give the same labels."""
    cost = {(0, (8, False)): 2.0, (3, (8, False)): 1.0, (4, (8, False)): 4.0,
            (3, (0, False)): 0.5, (3, (16, True)): 0.7}
    ctx = FakeContext(cost={})
    t = [0.0]
    seen = []
    monkeypatch.setattr(autotune.time, "perf_counter", lambda: t[0])

    def run():
        seen.append((ctx.box, ctx.split))
        t[0] += cost[(ctx.box, ctx.split)]

    res = autotune._two_stage(ctx, run, lambda: np.arange(5), 1, 3,
                              [(0, False), (8, False), (16, True)])
    assert res["best"] == {"box_split": 3, "sparse_max": 0, "sparse_core": False}
    assert [r["box_split"] for r in res["box_split_results"]] == [0, 3, 4]
    assert all(box == 3 for box, _ in seen[-6:])
    calls = iter([np.arange(5)] * 3 + [np.arange(5) + 1])
    with pytest.raises(RuntimeError, match="changed the labels"):
        autotune._two_stage(ctx, lambda: None, lambda: next(calls), 1, 3, [(0, False)])


def test_box_split_candidates():
    assert autotune.box_split_candidates(2) == [0, 3, 4]
    assert autotune.box_split_candidates(4) == [0, 2, 4]
    assert all(k ** 8 <= 1 << 24 for k in autotune.box_split_candidates(8))


def test_apply_accepts_result_or_best():
    ctx = FakeContext(cost={})
    autotune.apply(ctx, {"best": {"box_split": 3, "sparse_max": 16, "sparse_core": True}})
    assert ctx.split == (16, True) and ctx.box == 3
    autotune.apply(ctx, {"sparse_max": 4, "sparse_core": False})
    assert ctx.split == (4, False) and ctx.box == 3


@pytest.mark.gpu
def test_autotune_on_gpu():
    from hpdbscan_amd import _lib, datagen
    x = datagen.blobs_100k(n=60_000, seed=7)
    print("[gpu-case] autotune blobs 60k", flush=True)
    ctx = _lib.Context(0)
    try:
        res = autotune.tune(x, 0.3, 4, ctx=ctx, reps=1)
        assert (res["best"]["sparse_max"], res["best"]["sparse_core"]) in autotune.CANDIDATES
        assert res["best"]["box_split"] in autotune.box_split_candidates(3)
        autotune.apply(ctx, res)
        assert ctx.work_split() == (res["best"]["sparse_max"], res["best"]["sparse_core"])
        assert ctx.box_split() == res["best"]["box_split"]
    finally:
        ctx.close()
