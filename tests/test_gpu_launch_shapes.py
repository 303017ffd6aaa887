"""Trajectories do not depend on how the engine is launched: the same
replicas run under every launch shape (lattice in LDS or in HBM, one to
four replica waves per workgroup) and replica order give the oracle's
state bit for bit, and the autotuner's choice runs like the default."""
import numpy as np
import pytest

from kmos_amd import autotune
from kmos_amd import workloads as W
from kmos_amd.compiler import compile_model
from paritylib import compare_state, oracle_replica

pytestmark = pytest.mark.gpu

SIZES = [(6, 5), (20, 20), (9, 9), (20, 20), (6, 5), (13, 4)]


def _setup(engine, cm, rng_seed):
    g = np.random.default_rng(rng_seed)
    P = cm.nr_of_proc
    rates = 10.0 ** g.uniform(-2, 6, (len(SIZES), P))
    m = engine.Model(cm)
    b = engine.Batch(m, [list(s) + [1] for s in SIZES])
    seeds = np.arange(len(SIZES), dtype=np.uint32) + 31
    b.init(seeds, np.zeros(len(SIZES), np.uint32))
    b.set_rates(rates)
    ids = sorted(cm.project.species_ids().values())
    configs = [g.choice(ids, size=int(v)).astype(np.int8) for v in b.volumes]
    b.set_config(configs)
    return b, rates, seeds, configs


@pytest.mark.parametrize("backend", ["local_smart", "lat_int"])
def test_every_launch_shape_gives_the_same_trajectories(engine, oracle, backend):
    cm = compile_model(W.ruo2_project(), backend)
    b, rates, seeds, configs = _setup(engine, cm, 5)
    orcs = [oracle_replica(oracle, cm, s, int(seeds[i]), 0, rates[i], configs[i]) for i, s in enumerate(SIZES)]
    done = 0
    # both step kernel forms (select after the replay, or overlapped with
    # its reads) on every shape and order
    for ov in (0, 1):
        b.set_overlap(ov)
        for shape in autotune.SHAPES:
            for order in autotune.ORDERS:
                print("case", backend, "overlap", ov, shape, order, flush=True)
                b.set_launch(*shape)
                ids = autotune.replica_order(b.volumes, order)
                b.run(150, ids)
                for o in orcs:
                    o.run(150)
                done += 150
                for r, o in enumerate(orcs):
                    compare_state(b, r, o, where="%s ov%d %s %s r%d after %d" % (backend, ov, shape, order, r, done))
    b.close()


def test_tuned_launch_matches_default(engine):
    cm = compile_model(W.ruo2_project(), "local_smart")
    m = engine.Model(cm)
    best = autotune.tune_launch(lambda: _setup(engine, cm, 9)[0], steps=200, warmup=100, repeats=1)
    assert best["waves_per_group"] in (1, 2, 4) and best["ms"] > 0
    a, _, _, _ = _setup(engine, cm, 9)
    c, _, _, _ = _setup(engine, cm, 9)
    ids = autotune.apply(c, best)
    a.run(700)
    c.run(700, ids)
    for r in range(len(SIZES)):
        np.testing.assert_array_equal(a.lattice(r), c.lattice(r))
        assert a.scalars(r) == c.scalars(r)
        for x, y in zip(a.avail(r), c.avail(r)):
            np.testing.assert_array_equal(x, y)
    a.close()
    c.close()
    del m


def test_lattice_too_large_for_lds_runs_from_hbm(engine, oracle):
    """A lattice larger than the LDS holds (RuO2 1024 x 512 cells, 2^20
    sites, 256 KiB packed) makes the automatic launch keep it in HBM; the
    trajectory is the oracle's."""
    cm = compile_model(W.ruo2_project(), "local_smart")
    g = np.random.default_rng(8)
    rates = 10.0 ** g.uniform(-1, 4, (1, cm.nr_of_proc))
    m = engine.Model(cm)
    b = engine.Batch(m, [[1024, 512, 1]])
    b.init(np.array([77], np.uint32), np.zeros(1, np.uint32))
    b.set_rates(rates)
    ids = sorted(cm.project.species_ids().values())
    config = g.choice(ids, size=int(b.volumes[0])).astype(np.int8)
    b.set_config([config])
    b.run(20000)
    o = oracle_replica(oracle, cm, (1024, 512), 77, 0, rates[0], config)
    o.run(20000)
    compare_state(b, 0, o, where="ruo2 1024x512 from HBM")
    b.close()
