"""Shared helpers of the parity tests: run the same replicas on the engine
and on the oracle and compare their full state."""
import os

import numpy as np

from kmos_amd.compiler import compile_model


GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
DATA = os.path.join(os.path.dirname(GOLDEN), os.pardir, "kmos_amd", "data")


def model_file(name):
    """A model file of the tests: tests/golden, or the package's data
    directory for the models the package itself ships (the RuO2 model)."""
    p = os.path.join(GOLDEN, name)
    return p if os.path.exists(p) else os.path.normpath(os.path.join(DATA, name))


def oracle_replica(pyoracle, cm, size, seed, stream, rates, config=None):
    o = pyoracle.Oracle(cm.to_blob(), size)
    o.init(seed, stream)
    o.set_rates(rates)
    if config is not None:
        o.set_config(config)
    return o


def compare_clock(t, integ, t_ref, integ_ref, where=""):
    """kmc_time and integ_rates of a replica against the oracle's.

    Both sides compute integ_rates(p) += (nr(p) * rate(p)) * dt with the
    same dt = -log(u) / total and the same operation order, so they can
    only differ through the device's log.  When kmc_time, the sum of every
    dt, is bit-identical, integ_rates must be bit-identical too; otherwise
    both are held to 1e-12 relative (about 4500 ulp), the bound kmc_time
    is held to."""
    if t == t_ref:
        np.testing.assert_array_equal(integ, integ_ref, err_msg="%s integ_rates" % where)
    else:
        np.testing.assert_allclose(t, t_ref, rtol=1e-12, atol=0, err_msg="%s kmc_time" % where)
        np.testing.assert_allclose(integ, integ_ref, rtol=1e-12, atol=0, err_msg="%s integ_rates" % where)


def compare_state(batch, r, o, where=""):
    """Bit-exact lattice, bookkeeping and counters; clock and integrated
    rates bit-identical, or within 1e-12 if the device's log differs."""
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
    compare_clock(t, batch.integ_rates(r), o.kmc_time, o.integ_rates(), where)
