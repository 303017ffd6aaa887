"""Continuous batching: every request's samples are the same alone and in
the mix, and match the oracle replica run chunk by chunk."""
import numpy as np
import pytest

from kmos_amd import serve
from paritylib import compare_clock, oracle_replica

pytestmark = pytest.mark.gpu


def _small_stream():
    reqs = serve.poisson_stream(rate=2000.0, n=40, seed=3, sizes=(20, 64),
                                steps_range=(1e3, 2e4), sample_every=1000)
    return reqs


def test_same_alone_and_in_the_mix(engine):
    reqs = _small_stream()
    sizes = [(20, 20, 1), (64, 64, 1)]
    srv = serve.Server(hbm_budget=0, sizes=sizes, slots_per_size=4, intervals_per_launch=3)
    # room for a handful of small pages only: requests wait for slots and reuse them
    srv.hbm_budget = 2 * srv.models["ruo2"].footprint([list(z) for z in sizes] * 4)
    print("serving", len(reqs), "requests", flush=True)
    res = srv.serve(reqs, realtime=True)
    assert srv.stats["peak_bytes"] <= srv.hbm_budget
    assert len(res) == len(reqs)
    for r, q in zip(res, reqs):
        assert r.rid == q.rid
        assert r.status in (0, 1)
        if r.status == 0:
            assert r.steps_done == q.steps
        assert r.finished >= r.arrival
    print(serve.latency_summary(res), srv.stats, flush=True)
    for q in reqs[::7]:
        alone = srv.run_alone(q)
        mix = res[q.rid]
        assert alone.status == mix.status and alone.steps_done == mix.steps_done
        for k in mix.samples:
            np.testing.assert_array_equal(alone.samples[k], mix.samples[k], err_msg="rid %d %s" % (q.rid, k))
    srv.close()


def test_samples_match_oracle(engine, oracle):
    reqs = [q for q in _small_stream() if q.model != "quickstart"][:3]
    srv = serve.Server()
    res = srv.serve(reqs, realtime=False)
    for q, r in zip(reqs, res):
        cm = srv.compiled[q.model]
        o = oracle.Oracle(cm.to_blob(), q.size[:2])
        o.init(q.seed, 0, q.config)
        o.set_rates(q.rates)
        done = 0
        for i in range(len(r.samples["kmc_step"])):
            n = min(q.sample_every, q.steps - done)
            done += o.run(n)
            assert r.samples["kmc_step"][i] == o.kmc_step == done
            np.testing.assert_array_equal(r.samples["procstat"][i], o.procstat())
            np.testing.assert_array_equal(r.samples["occupation"][i], o.occupation())
            compare_clock(r.samples["kmc_time"][i], r.samples["integ_rates"][i], o.kmc_time,
                          o.integ_rates(), "served sample %d" % i)
            if o.status:
                break
    srv.close()
