"""Particle tracking on the engine against the oracle: particle ids and
sites, displacements and trajectory events bit for bit, and the ACF bins
(counters, buffers, properties exact; bin contents, which carry kmc_time,
to 1e-12), on a pure-hop model with hops across cell borders and on the
same model with adsorption and desorption (no-hop processes), ragged sizes
in one batch, for both backends; and the kmos.run.acf-style functions on a KMC_Model of the
reference's 2d_grid test model.
"""
import os

import numpy as np
import pytest

from kmos_amd.acf import displacement_table
from kmos_amd.compiler import compile_model
from paritylib import compare_state, oracle_replica
from test_acf import hop_project

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _compare_msd(b, r, o, where):
    g, c = b.msd_state(r), o.msd_state()
    assert g["nr_of_ions"] == c["nr_of_ions"], where
    np.testing.assert_array_equal(g["site_arr"], c["site_arr"], err_msg=where)
    np.testing.assert_array_equal(g["id_arr"], c["id_arr"], err_msg=where)
    np.testing.assert_array_equal(g["displacement"], c["displacement"], err_msg=where)
    np.testing.assert_array_equal(g["events"], c["events"], err_msg=where)


@pytest.mark.parametrize("with_adsorption,backend", [(False, "local_smart"), (True, "local_smart"),
                                                     (False, "lat_int"), (True, "lat_int")])
def test_tracking_matches_oracle(engine, oracle, with_adsorption, backend):
    print("case adsorption=%s %s" % (with_adsorption, backend), flush=True)
    cm = compile_model(hop_project(with_adsorption), backend)
    hops, disp = displacement_table(cm)
    P = cm.nr_of_proc
    sizes = [(7, 5), (12, 9), (20, 20)]
    g = np.random.default_rng(4)
    rates = 10.0 ** g.uniform(-1, 1, (len(sizes), P))
    ion, empty = cm.species_names.index("ion"), cm.species_names.index("empty")
    m = engine.Model(cm)
    b = engine.Batch(m, [list(s) + [1] for s in sizes])
    seeds = np.array([3, 4, 5], np.uint32)
    b.init(seeds, np.zeros(3, np.uint32))
    b.set_rates(rates)
    configs = [np.where(g.random(int(v)) < 0.3, ion, empty).astype(np.int8) for v in b.volumes]
    b.set_config(configs)
    b.msd_init(ion, hops, disp)
    b.msd_trajectory(2500)
    orcs = []
    for r, s in enumerate(sizes):
        o = oracle_replica(oracle, cm, s, int(seeds[r]), 0, rates[r], configs[r])
        o.msd_init(ion, hops, disp)
        o.msd_trajectory(2500)
        orcs.append(o)
    for r, o in enumerate(orcs):
        _compare_msd(b, r, o, "r%d after init" % r)
    done = 0
    for chunk in (1, 999, 1500):
        b.run_displacement(chunk, traj_on=True)
        for o in orcs:
            o.run_displacement(chunk, traj_on=True)
        done += chunk
        for r, o in enumerate(orcs):
            compare_state(b, r, o, where="r%d after %d" % (r, done))
            _compare_msd(b, r, o, "r%d after %d" % (r, done))
    # a plain run does not track; tracking resumes after it, trajectory full
    before = [b.msd_state(r) for r in range(3)]
    b.run(300)
    for o in orcs:
        o.run(300)
    for r in range(3):
        np.testing.assert_array_equal(b.msd_state(r)["displacement"], before[r]["displacement"])
    b.run_displacement(200, traj_on=False)
    b.msd_zero(1, 1)
    for r, o in enumerate(orcs):
        o.run_displacement(200, traj_on=False)
        if r == 1:
            o.msd_zero()
        compare_state(b, r, o, where="r%d final" % r)
        _compare_msd(b, r, o, "r%d final" % r)


def test_msd_functions_on_kmc_model(engine, oracle, tmp_path):
    from kmos_amd import acf
    from kmos_amd.export import export_source
    from kmos_amd.model import evaluate_rate_expression
    from kmos_amd.run import KMC_Model
    d = export_source(os.path.join(GOLDEN, "2d_grid.xml"), str(tmp_path / "grid"), "local_smart")
    model = KMC_Model(export_dir=d, size=[10, 10], random_seed=3)
    cm = model.compiled
    rates = np.array([evaluate_rate_expression(model.settings.rate_constants[p][0], model.settings_parameters)
                      for p in cm.process_names])
    o = oracle_replica(oracle, cm, (10, 10), 3, 0, rates)
    hops, disp = displacement_table(cm)
    ion = cm.species_names.index("ion")
    acf.initialize_msd(model, "ion")
    o.msd_init(ion, hops, disp)
    acf.allocate_trajectory(model, 400)
    o.msd_trajectory(400)
    acf.do_kmc_steps_displacement(model, 250, traj_on=True)
    o.run_displacement(250, traj_on=True)
    st = o.msd_state()
    n = st["nr_of_ions"]
    assert n == 200                                    # a_1, a_2 start as ion in 100 cells
    np.testing.assert_array_equal(acf.get_displacement(model)[0], st["displacement"][1:n + 1])
    ref_msd = 0.0
    for k in range(1, n + 1):
        dk = st["displacement"][k]
        ref_msd = ref_msd + ((dk[0] ** 2 + dk[1] ** 2) + dk[2] ** 2)
    assert acf.calc_msd(model) == ref_msd / n
    traj = acf.get_trajectory(model)
    assert traj.shape == (1, n, 401)
    np.testing.assert_array_equal(traj[0, :, 250], st["site_arr"][1:n + 1])
    assert not traj[0, :, 251:].any()                  # steps not run yet
    acf.set_displacement_to_zero(model)
    assert not acf.get_displacement(model).any()
    compare_state(model._kmc.batch, 0, o, where="kmc_model after tracking")


def _compare_acf(b, r, o, where):
    """Counters, properties, buffers and bin indices bit for bit; the bins
    and t_o carry kmc_time, whose log may differ by an ulp from glibc's, so
    they are compared like kmc_time (1e-12 relative)."""
    g, c = b.acf_state(r), o.acf_state()
    for k in ("counter_write_in_bin", "buffer_acf", "property_o"):
        np.testing.assert_array_equal(g[k], c[k], err_msg="%s %s" % (where, k))
    np.testing.assert_allclose(g["config_bin"], c["config_bin"], rtol=1e-12, atol=0,
                               err_msg="%s config_bin" % where)
    assert g["t_o"] == pytest.approx(c["t_o"], rel=1e-12, abs=0), where
    for k in ("sum_buffer", "wrap_count", "t_bin", "bin_index", "nr_of_bins", "extended_nr_of_bins",
              "nr_of_ions_before", "nr_of_ions"):
        assert g[k] == c[k], "%s %s: %s vs %s" % (where, k, g[k], c[k])


@pytest.mark.parametrize("with_adsorption,backend", [(False, "local_smart"), (True, "local_smart"),
                                                     (True, "lat_int")])
def test_acf_matches_oracle(engine, oracle, with_adsorption, backend):
    """Two property types, bins short enough for several wraps and steps
    that spill into the extended bins, ragged sizes in one batch."""
    print("case acf adsorption=%s %s" % (with_adsorption, backend), flush=True)
    cm = compile_model(hop_project(with_adsorption), backend)
    hops, _ = displacement_table(cm)
    P = cm.nr_of_proc
    sizes = [(8, 6), (12, 9), (20, 16)]
    g = np.random.default_rng(6)
    rates = 10.0 ** g.uniform(-1, 1, (len(sizes), P))
    ion, empty = cm.species_names.index("ion"), cm.species_names.index("empty")
    m = engine.Model(cm)
    b = engine.Batch(m, [list(s) + [1] for s in sizes])
    seeds = np.array([7, 8, 9], np.uint32)
    b.init(seeds, np.zeros(3, np.uint32))
    b.set_rates(rates)
    configs = [np.where(g.random(int(v)) < 0.35, ion, empty).astype(np.int8) for v in b.volumes]
    b.set_config(configs)
    t_bin, t_f = 0.004, 0.03
    b.acf_alloc(2, t_bin, t_f)
    pas = []
    for v in b.volumes:
        pa = np.zeros(int(v) + 1, np.int32)
        pa[1:] = [1 + (s % cm.spuck) for s in range(int(v))]       # site a: type 1, site b: type 2
        pas.append(pa)
    pp = np.zeros((3, 3))
    pp[1:, 1:] = np.outer([2.0, -1.0], [2.0, -1.0])
    b.acf_set_properties(pas, pp)
    # a call rejected for a type out of range in its last replica changes
    # neither the products nor the types of the replicas before it: the
    # comparisons below use pas and pp
    from kmos_amd._lib import KmcgError
    bad = [np.where(pa > 0, 3 - pa, 0).astype(np.int32) for pa in pas]
    bad[-1][1] = 5
    with pytest.raises(KmcgError):
        b.acf_set_properties(bad, 3.0 * pp)
    b.acf_init(ion, hops)
    b.msd_trajectory(3000)
    orcs = []
    for r, s in enumerate(sizes):
        o = oracle_replica(oracle, cm, s, int(seeds[r]), 0, rates[r], configs[r])
        o.acf_alloc(2, t_bin, t_f)
        o.acf_set(pas[r], pp)
        o.acf_init(ion, hops)
        o.msd_trajectory(3000)
        orcs.append(o)
    for r, o in enumerate(orcs):
        _compare_acf(b, r, o, "r%d after init" % r)
    done = 0
    for chunk in (1, 1999, 1000):
        b.run_acf(chunk, traj_on=True)
        for o in orcs:
            o.run_acf(chunk, traj_on=True)
        done += chunk
        for r, o in enumerate(orcs):
            compare_state(b, r, o, where="r%d after %d" % (r, done))
            _compare_acf(b, r, o, "r%d after %d" % (r, done))
            _compare_msd(b, r, o, "r%d after %d" % (r, done))
    assert min(b.acf_state(r)["wrap_count"] for r in range(3)) >= 2
    b.acf_zero(0, 3)
    for o in orcs:
        o.acf_zero()
    b.run_acf(500)
    for r, o in enumerate(orcs):
        o.run_acf(500)
        _compare_acf(b, r, o, "r%d after zero" % r)


def test_acf_bins_on_kmc_model(engine, oracle, tmp_path):
    from kmos_amd import acf
    from kmos_amd.export import export_source
    from kmos_amd.model import evaluate_rate_expression
    from kmos_amd.run import KMC_Model
    d = export_source(os.path.join(GOLDEN, "2d_grid.xml"), str(tmp_path / "grid_acf"), "local_smart")
    model = KMC_Model(export_dir=d, size=[8, 8], random_seed=5)
    cm = model.compiled
    rates = np.array([evaluate_rate_expression(model.settings.rate_constants[p][0], model.settings_parameters)
                      for p in cm.process_names])
    acf.allocate_acf(model, 2, 1e-3, 5e-3)
    acf.set_property_acf(model, "default_a_1", 1)
    acf.set_property_acf(model, "default_a_2", 1)
    acf.set_property_acf(model, "default_b_1", 2)
    acf.set_property_acf(model, "default_b_2", 2)
    acf.set_types_acf(model, 1.0)
    acf.set_types_acf(model, -1.0)
    acf.calc_product_property(model)
    np.testing.assert_array_equal(acf.get_product_property(model)[0], [[1, -1], [-1, 1]])
    acf.initialize_acf(model, "ion")
    acf.do_kmc_steps_acf(model, 3000)
    o = oracle_replica(oracle, cm, (8, 8), 5, 0, rates)
    o.acf_alloc(2, 1e-3, 5e-3)
    pa = np.zeros(8 * 8 * 4 + 1, np.int32)
    pa[1:] = [1 if s % 4 < 2 else 2 for s in range(8 * 8 * 4)]
    np.testing.assert_array_equal(acf.get_property_acf(model)[0], pa[1:])
    o.acf_set(pa, np.array([[0, 0, 0], [0, 1, -1], [0, -1, 1]], float))
    o.acf_init(cm.species_names.index("ion"), displacement_table(cm)[0])
    o.run_acf(3000)
    st = o.acf_state()
    ref = st["config_bin"][:st["nr_of_bins"]] / (st["wrap_count"] * st["t_bin"])
    np.testing.assert_allclose(acf.get_acf(model), ref, rtol=1e-12, atol=0)
    np.testing.assert_allclose(acf.get_acf(model, normalization=True), ref / ref[0], rtol=1e-12, atol=0)
    np.testing.assert_array_equal(acf.get_buffer_acf(model)[0], st["buffer_acf"][1:st["nr_of_ions"] + 1])
    compare_state(model._kmc.batch, 0, o, where="kmc_model after acf")
