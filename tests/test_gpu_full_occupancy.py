"""Parity at full occupancy: the ZGB workload (4096 replicas of 128^2), the
RuO2 mix (2048 replicas of 20^2 / 64^2 / 128^2, random starting coverage)
and the pairwise lat_int workload (512 replicas of 256^2, the one-pass
lat_int kernel) run as whole batches in the launch shape the benchmark
uses, and 32 replicas of each, drawn at random over the batch, equal the
oracle bit for bit after 2000 steps (lattice, both avail tables, counters,
clock)."""
import numpy as np
import pytest

from kmos_amd import workloads as W
from kmos_amd.compiler import compile_model
from paritylib import compare_state, oracle_replica

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name", ["zgb", "ruo2", "pairwise"])
def test_full_batch_sampled_replicas_match_oracle(engine, oracle, name):
    wl = W.WORKLOADS[name]()
    print("case", wl.name, flush=True)
    cm = compile_model(wl.project, wl.backend)
    b = engine.Batch(engine.Model(cm), [list(s) for s in wl.sizes])
    b.init(wl.seeds, wl.streams)
    b.set_rates(wl.rates)
    if wl.configs is not None:
        b.set_config(wl.configs)
    b.run(1000)
    b.run(1000)
    _, k_after, st = b.scalars_all()
    pick = np.sort(np.random.default_rng(5).choice(wl.n, 32, replace=False))
    for r in pick:
        o = oracle_replica(oracle, cm, tuple(wl.sizes[r][:2]), int(wl.seeds[r]), int(wl.streams[r]),
                           wl.rates[r], None if wl.configs is None else wl.configs[r])
        o.run(int(k_after[r]))
        assert k_after[r] == 2000 or st[r] != 0
        compare_state(b, int(r), o, where="%s replica %d" % (wl.name, r))
    b.close()
