"""The launch autotuner's host logic: replica orders, the choice by median
device time over interleaved repeats, the cache, and the serving-chunk
choice, on stand-in batches and servers (no GPU)."""
import numpy as np
import pytest

from kmos_amd import autotune


def test_replica_orders_are_permutations():
    vol = [800, 32768, 8192, 8192, 800, 32768, 8192, 8192]
    for how in autotune.ORDERS:
        ids = autotune.replica_order(vol, how)
        assert sorted(ids.tolist()) == list(range(len(vol))), how
    assert autotune.replica_order(vol, "given").tolist() == list(range(8))
    assert autotune.replica_order(vol, "size_desc").tolist() == [1, 5, 2, 3, 6, 7, 0, 4]
    # round robin from the largest size: 32768, 8192, 800, 32768, 8192, 800, 8192, 8192
    assert autotune.replica_order(vol, "interleave").tolist() == [1, 2, 0, 5, 3, 4, 6, 7]
    with pytest.raises(ValueError):
        autotune.replica_order(vol, "random")


class FakeBatch:
    """Device time = a per-shape cost + a drift that grows with every run
    (the probe's state moves on), so interleaving the repeats matters."""

    def __init__(self, cost):
        self.volumes = np.array([100, 400, 100, 400])
        self.cost = cost
        self.shape = None
        self.runs = []
        self.closed = False
        self.drift = 0.0

    def set_launch(self, lds, wpg):
        self.shape = (lds, wpg)

    def run(self, steps, replicas=None):
        self.runs.append((self.shape, None if replicas is None else list(replicas), steps))
        self.drift += 0.01
        order = "given" if replicas is None or list(replicas) == [0, 1, 2, 3] else "other"
        self.ms = self.cost[self.shape] + self.drift + (0.0 if order == "given" else 0.5) \
            if self.shape else 0.0

    def last_run_ms(self):
        return self.ms

    def close(self):
        self.closed = True


def test_tune_launch_picks_fastest_and_closes_probe(tmp_path):
    cost = {(-1, 1): 5.0, (-1, 2): 3.0, (-1, 4): 4.0, (0, 1): 9.0, (0, 4): 8.0}
    made = []

    def make():
        made.append(FakeBatch(cost))
        return made[-1]
    res = autotune.tune_launch(make, steps=100, warmup=10, repeats=3)
    assert (res["lattice_in_lds"], res["waves_per_group"], res["order"]) == (-1, 2, "given")
    assert len(made) == 1 and made[0].closed
    b = made[0]
    assert b.runs[0] == (None, None, 10)                        # warm-up on the default shape
    ncand = len(autotune.SHAPES) * len(autotune.ORDERS)
    assert len(b.runs) == 1 + 3 * ncand
    # repeats interleave the candidates: each candidate once per pass
    first_pass = [r[0] for r in b.runs[1:1 + ncand]]
    assert sorted(set(first_pass)) == sorted(cost)
    assert len(res["table"]) == ncand and all(len(t["runs"]) == 3 for t in res["table"])
    # the cache returns the stored choice without timing
    path = str(tmp_path / "tune.json")
    r1 = autotune.tune_launch(make, steps=100, warmup=0, repeats=1, cache_path=path, key="k")
    n = len(made)
    r2 = autotune.tune_launch(make, steps=100, warmup=0, repeats=1, cache_path=path, key="k")
    assert len(made) == n and r2["cached"] and not r1["cached"]
    assert (r2["lattice_in_lds"], r2["waves_per_group"], r2["order"]) == \
        (r1["lattice_in_lds"], r1["waves_per_group"], r1["order"])


def test_apply_sets_shape_and_returns_order():
    b = FakeBatch({})
    ids = autotune.apply(b, {"lattice_in_lds": 0, "waves_per_group": 4, "order": "size_desc"})
    assert b.shape == (0, 4) and ids.tolist() == [1, 3, 0, 2]


def test_tune_serving_picks_lowest_p99(monkeypatch):
    from kmos_amd import serve
    p99 = {1: 0.9, 2: 0.5, 4: 0.5, 8: 2.0}
    tput = {1: 1.0, 2: 2.0, 4: 3.0, 8: 4.0}

    class FakeServer:
        def __init__(self, k):
            self.k = k

        def serve(self, reqs, realtime=True):
            return [self.k]

        def close(self):
            pass
    monkeypatch.setattr(serve, "latency_summary",
                        lambda res: {"p99_latency_s": p99[res[0]], "steps_per_s": tput[res[0]],
                                     "p50_latency_s": 0.0})
    out = autotune.tune_serving(FakeServer, lambda: [], candidates=(1, 2, 4, 8))
    assert out["intervals_per_launch"] == 4            # tie on p99, more steps/s
    assert [r["intervals_per_launch"] for r in out["table"]] == [1, 2, 4, 8]
