"""Threading contract of the C ABI: one handle is driven by one host thread
at a time, and different handles may be driven from different threads at
once (the entry points release the GIL through ctypes, keep no shared
mutable state and keep their error text per thread).

Two batches -- RuO2 local_smart and pairwise lat_int -- are advanced from
two threads at the same time, while a third thread keeps reading the
getters of a third batch and provokes argument errors on it.  Every
trajectory equals the oracle bit for bit, the third batch's state reads
the same throughout, each thread's error text is its own, and the runs
did overlap in time.
"""
import threading
import time

import numpy as np
import pytest

from kmos_amd import _lib
from kmos_amd import workloads as W
from kmos_amd.compiler import compile_model
from paritylib import compare_state, oracle_replica

pytestmark = pytest.mark.gpu


def _setup(engine, cm, sizes, seed, rates):
    b = engine.Batch(engine.Model(cm), sizes)
    n = len(sizes)
    b.init(np.full(n, seed, np.uint32), np.arange(n, dtype=np.uint32))
    b.set_rates(rates)
    return b


def test_two_threads_two_batches_and_a_reader(engine, oracle):
    g = np.random.default_rng(11)
    cm_ls = compile_model(W.ruo2_project(), "local_smart")
    cm_li = compile_model(W.pairwise_project(), "lat_int")
    pt = W.pairwise_project()
    jobs = {
        "ruo2": (cm_ls, [(24, 20, 1), (16, 16, 1), (32, 12, 1)], 5,
                 10.0 ** g.uniform(-2, 7, (3, cm_ls.nr_of_proc)), 4000),
        "pairwise": (cm_li, [(24, 24, 1), (16, 20, 1)], 6,
                     np.stack([W.pairwise_rates(pt, e) for e in (-0.1, 0.15)]), 3000),
    }
    batches = {k: _setup(engine, *v[:4]) for k, v in jobs.items()}
    reader = _setup(engine, cm_ls, [(12, 12, 1)], 9, 10.0 ** g.uniform(-2, 7, (1, cm_ls.nr_of_proc)))
    reader.run(500)
    snap = (reader.lattice(0).copy(), reader.procstat(0).copy(), reader.scalars(0))

    start = threading.Barrier(3)
    stop = threading.Event()
    spans, errors, failures = {}, [], []

    def runner(key):
        try:
            b, steps = batches[key], jobs[key][4]
            start.wait()
            t0 = time.perf_counter()
            for _ in range(4):
                b.run(steps // 4)
            spans[key] = (t0, time.perf_counter())
            # a failing call here leaves this thread's own error text
            with pytest.raises(_lib.KmcgError, match="replica index out of range"):
                b.get_rates(50 + len(key))
        except Exception as e:          # reported by the main thread
            failures.append((key, e))

    def read():
        try:
            start.wait()
            k = 0
            while not stop.is_set() or k < 5:
                lat, ps, sc = reader.lattice(0), reader.procstat(0), reader.scalars(0)
                assert np.array_equal(lat, snap[0]) and np.array_equal(ps, snap[1]) and sc == snap[2]
                try:
                    reader.set_config([np.zeros(5, np.int8)])
                except ValueError as e:
                    errors.append(str(e))
                rc = _lib.load().kmcg_set_config(reader.h, 7, 1, None)
                assert rc != 0 and _lib.load().kmcg_last_error() == b"bad replica range"
                k += 1
        except Exception as e:
            failures.append(("reader", e))

    threads = [threading.Thread(target=runner, args=(k,)) for k in jobs] + [threading.Thread(target=read)]
    for t in threads:
        t.start()
    for t in threads[:2]:
        t.join(timeout=300)
    stop.set()
    threads[2].join(timeout=60)
    assert not any(t.is_alive() for t in threads), "a thread did not finish"
    assert not failures, failures
    assert errors and all("sites" in e for e in errors)
    # the two runs overlapped in time
    (a0, a1), (b0, b1) = spans["ruo2"], spans["pairwise"]
    assert max(a0, b0) < min(a1, b1)

    for key, (cm, sizes, seed, rates, steps) in jobs.items():
        b = batches[key]
        for r, size in enumerate(sizes):
            o = oracle_replica(oracle, cm, size, seed, r, rates[r])
            o.run(steps)
            compare_state(b, r, o, where="thread %s replica %d" % (key, r))
        b.close()
    reader.close()
