"""Runs replayed through captured HIP graphs give what direct launches
give: twin batches, one launching directly and one through graphs, run the
same sequence of runs (the same replica list and steps, so the graph is
replayed, with budgets and rate changes that differ between the replays,
eager work and synchronisations in between, a smaller replica list that
captures a second graph, recorded runs), and every replay is compared
with the direct run bit for bit; replica 0 also against the oracle."""
import numpy as np
import pytest

from kmos_amd import workloads as W
from kmos_amd.compiler import compile_model
from paritylib import compare_state, oracle_replica

pytestmark = pytest.mark.gpu

SIZES = [(6, 5), (20, 20), (9, 9), (13, 4)]


def _twin(engine, m, cm, rates, configs):
    b = engine.Batch(m, [list(s) + [1] for s in SIZES])
    b.init(np.arange(len(SIZES), dtype=np.uint32) + 41, np.zeros(len(SIZES), np.uint32))
    b.set_rates(rates)
    b.set_config(configs)
    return b


def _same(a, b, where):
    for r in range(len(SIZES)):
        assert a.scalars(r) == b.scalars(r), where
        np.testing.assert_array_equal(a.lattice(r), b.lattice(r), err_msg=where)
        for x, y in zip(a.avail(r), b.avail(r)):
            np.testing.assert_array_equal(x, y, err_msg=where)
        np.testing.assert_array_equal(a.procstat(r), b.procstat(r), err_msg=where)
        np.testing.assert_array_equal(a.integ_rates(r), b.integ_rates(r), err_msg=where)


def test_graph_replays_match_direct_launches(engine, oracle):
    cm = compile_model(W.ruo2_project(), "local_smart")
    g = np.random.default_rng(3)
    P = cm.nr_of_proc
    rates = 10.0 ** g.uniform(-2, 6, (len(SIZES), P))
    ids = sorted(cm.project.species_ids().values())
    m = engine.Model(cm)
    configs = [g.choice(ids, size=int(v)).astype(np.int8) for v in
               (np.array([s[0] * s[1] for s in SIZES]) * cm.spuck)]
    a = _twin(engine, m, cm, rates, configs)
    b = _twin(engine, m, cm, rates, configs)
    b.set_graph(True)
    o = oracle_replica(oracle, cm, SIZES[0], 41, 0, rates[0], configs[0])
    allr = np.arange(len(SIZES), dtype=np.int32)
    for it in range(6):
        print("case replay", it, flush=True)
        bud = g.integers(50, 400, len(SIZES))
        ba = a.run(300, allr, bud)
        bb = b.run(300, allr, bud)
        np.testing.assert_array_equal(ba, bb)
        o.run(int(min(300, bud[0])))
        _same(a, b, "replay %d" % it)
        compare_state(b, 0, o, where="graph replica 0 after replay %d" % it)
        if it == 2:
            # eager work between replays: new rates for replica 2 on both
            new = rates[2] * 1.5
            a.set_rates(new[None], r0=2)
            b.set_rates(new[None], r0=2)
    # a smaller replica list captures a second graph; the first is kept
    sub = np.array([3, 1], np.int32)
    for it in range(3):
        print("case subset", it, flush=True)
        a.run(150, sub)
        b.run(150, sub)
        _same(a, b, "subset %d" % it)
        a.run(120, allr)
        b.run(120, allr)
        _same(a, b, "full after subset %d" % it)
        o.run(120)
        compare_state(b, 0, o, where="graph replica 0 after subset %d" % it)
    # recorded runs replay too: records and budgets as direct
    for it in range(3):
        print("case recorded", it, flush=True)
        bud = np.array([400, 250, 400, 90])
        wa = a.run_recorded(400, allr, bud, 100, 4)
        wb = b.run_recorded(400, allr, bud, 100, 4)
        (ba, ra), (bb, rb) = wa(), wb()
        np.testing.assert_array_equal(ba, bb)
        for x, y in zip(ra, rb):
            for k in x:
                np.testing.assert_array_equal(x[k], y[k], err_msg="records %s" % k)
        _same(a, b, "recorded %d" % it)
        o.run(int(bud[0]))
        compare_state(b, 0, o, where="graph replica 0 after recorded %d" % it)
    # the select-overlapped kernel form is part of the graph key: toggling
    # it between replays captures (or reuses) the graph of that form
    for it, ov in enumerate((1, 0, 1, 0)):
        print("case overlap", ov, flush=True)
        a.set_overlap(ov)
        b.set_overlap(ov)
        a.run(200, allr)
        b.run(200, allr)
        _same(a, b, "overlap %d, run %d" % (ov, it))
        o.run(200)
        compare_state(b, 0, o, where="graph replica 0 overlap %d run %d" % (ov, it))
    a.set_overlap(-1)
    b.set_overlap(-1)
    b.set_graph(False)
    a.run(100)
    b.run(100)
    _same(a, b, "after graphs off")
    a.close()
    b.close()
