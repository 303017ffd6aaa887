"""The debug list check (kmcg_check_lists): zero violations after runs of
every model and backend (ragged sizes, random starting lattices), and it
finds a registration that was removed consistently (the list invariant
still holds, but a process whose conditions hold is missing at one site)
and one that was added where the conditions fail."""
import numpy as np
import pytest

from kmos_amd import workloads as W
from kmos_amd.compiler import compile_model

pytestmark = pytest.mark.gpu

CASES = [(W.ruo2_project, "local_smart"), (W.ruo2_project, "lat_int"), (W.zgb_project, "local_smart"),
         (W.pairwise_project, "lat_int"), (W.pairwise_project, "local_smart")]


def _batch(engine, cm, seed):
    sizes = [(6, 5), (17, 9), (20, 20)]
    g = np.random.default_rng(seed)
    m = engine.Model(cm)
    b = engine.Batch(m, [list(s) + [1] for s in sizes])
    b.init(np.arange(3, dtype=np.uint32) + 3, np.zeros(3, np.uint32))
    b.set_rates(10.0 ** g.uniform(-2, 4, (3, cm.nr_of_proc)))
    ids = sorted(cm.project.species_ids().values())
    b.set_config([g.choice(ids, size=int(v)).astype(np.int8) for v in b.volumes])
    return m, b


@pytest.mark.parametrize("factory,backend", CASES, ids=["%s-%s" % (f.__name__, b) for f, b in CASES])
def test_lists_consistent_after_runs(engine, factory, backend):
    print("case", factory.__name__, backend, flush=True)
    cm = compile_model(factory(), backend)
    m, b = _batch(engine, cm, 11)
    np.testing.assert_array_equal(b.check_lists(), 0)
    for chunk in (1, 999, 2000):
        b.run(chunk)
        np.testing.assert_array_equal(b.check_lists(), 0)
    b.close()


def test_check_finds_missing_and_extra_registrations(engine):
    cm = compile_model(W.ruo2_project(), "local_smart")
    m, b = _batch(engine, cm, 12)
    b.run(500)
    st = b.get_state(1)
    nr = st["nr_of_sites"]
    p = int(np.flatnonzero(nr >= 1)[0])
    # remove the last registration of process p, keeping the list invariant
    bad = {k: np.array(v, copy=True) if isinstance(v, np.ndarray) else v for k, v in st.items()}
    k = int(nr[p])
    s = int(bad["avail_sites"][p, k - 1])
    bad["avail_sites"][p, k - 1] = 0
    bad["avail_sites_back"][p, s - 1] = 0
    bad["nr_of_sites"][p] -= 1
    b.set_state(1, bad)
    np.testing.assert_array_equal(b.check_lists(), [0, 1, 0])
    # register it again, and also a site of p's type where p is not enabled
    b.set_state(1, st)
    np.testing.assert_array_equal(b.check_lists(), 0)
    extra = {k: np.array(v, copy=True) if isinstance(v, np.ndarray) else v for k, v in st.items()}
    spuck = cm.spuck
    n_exec = (s - 1) % spuck + 1
    free = [q for q in range(n_exec, int(b.volumes[1]) + 1, spuck) if extra["avail_sites_back"][p, q - 1] == 0]
    q = free[0]
    extra["nr_of_sites"][p] += 1
    extra["avail_sites"][p, k] = q
    extra["avail_sites_back"][p, q - 1] = k + 1
    b.set_state(1, extra)
    np.testing.assert_array_equal(b.check_lists(), [0, 1, 0])
    b.close()
