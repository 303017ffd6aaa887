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


def random_config(pt, V, rng):
    """A random lattice of V sites.  One layer: any species on every site.
    Several layers (kmos multi-lattice models): per cell one layer chosen
    at random holds random species, the other layers' sites the null
    species, so that cells of different layers meet."""
    ids = sorted(pt.species_ids().values())
    if len(pt.layers) == 1:
        return rng.choice(ids, size=V).astype(np.int8)
    layers = [lname for lname, _ in pt.site_list()]
    names = [layer.name for layer in pt.layers]
    spuck = len(layers)
    out = np.full(V, pt.null_species(), np.int8)
    for c in range(V // spuck):
        keep = names[int(rng.integers(len(names)))]
        for n in range(spuck):
            if layers[n] == keep:
                out[c * spuck + n] = rng.choice(ids)
    return out


def pdopd_project():
    """The reference's two-layer test model (25 sites per cell, 46
    processes with create/annihilate actions)."""
    from kmos_amd.model import load_project
    return load_project(model_file("pdopd.xml"))


def oracle_replica(pyoracle, cm, size, seed, stream, rates, config=None):
    o = pyoracle.Oracle(cm.to_blob(), size)
    o.init(seed, stream)
    o.set_rates(rates)
    if config is not None:
        o.set_config(config)
    return o


def compare_clock(t, integ, t_ref, integ_ref, where=""):
    """kmc_time and integ_rates of a replica against the oracle's.

    Both sides compute dt = -log(u) / total with the same u and total, then
    kmc_time += dt and integ_rates(p) += (nr(p) * rate(p)) * dt in the same
    order.  The device's log may differ from glibc's by an ulp, so a step's
    dt may differ by an ulp; kmc_time's own rounding usually absorbs that
    (dt is far below kmc_time), an integ_rate's rounding now and then does
    not.  Both are held to 1e-12 relative (about 4500 ulp); the per-step
    bound of one ulp on dt is tested on its own
    (test_kmc_time_within_one_ulp_per_step)."""
    np.testing.assert_allclose(t, t_ref, rtol=1e-12, atol=0, err_msg="%s kmc_time" % where)
    np.testing.assert_allclose(integ, integ_ref, rtol=1e-12, atol=0, err_msg="%s integ_rates" % where)


def compare_state(batch, r, o, where=""):
    """Bit-exact lattice, bookkeeping and counters; clock and integrated
    rates within 1e-12 relative (see compare_clock)."""
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
