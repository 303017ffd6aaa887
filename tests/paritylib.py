"""Shared helpers of the parity tests: run the same replicas on the engine
and on the oracle and compare their full state."""
import numpy as np

from kmos_amd.compiler import compile_model


def oracle_replica(pyoracle, cm, size, seed, stream, rates, config=None):
    o = pyoracle.Oracle(cm.to_blob(), size)
    o.init(seed, stream)
    o.set_rates(rates)
    if config is not None:
        o.set_config(config)
    return o


def compare_state(batch, r, o, time_rtol=1e-12, where=""):
    """Bit-exact lattice, bookkeeping and counters; clock within time_rtol."""
    t, dt, ks, ctr, st = batch.scalars(r)
    assert st == o.status, "%s status %d vs %d" % (where, st, o.status)
    assert ks == o.kmc_step, "%s kmc_step %d vs %d" % (where, ks, o.kmc_step)
    assert ctr == o.rng_ctr, "%s rng counter" % where
    np.testing.assert_array_equal(batch.lattice(r), o.lattice(), err_msg="%s lattice" % where)
    np.testing.assert_array_equal(batch.nr_of_sites(r), o.nr_of_sites(), err_msg="%s nr_of_sites" % where)
    s_gpu, b_gpu = batch.avail(r)
    s_orc, b_orc = o.avail()
    np.testing.assert_array_equal(s_gpu, s_orc, err_msg="%s avail sites" % where)
    np.testing.assert_array_equal(b_gpu, b_orc, err_msg="%s avail back-index" % where)
    np.testing.assert_array_equal(batch.procstat(r), o.procstat(), err_msg="%s procstat" % where)
    np.testing.assert_allclose(batch.integ_rates(r), o.integ_rates(), rtol=1e-9, atol=0,
                               err_msg="%s integ_rates" % where)
    np.testing.assert_allclose(t, o.kmc_time, rtol=time_rtol, atol=0, err_msg="%s kmc_time" % where)
