"""Displacement tracking on the engine against the oracle, bit for bit:
particle ids and sites, displacements and trajectory events, on a pure-hop
model with hops across cell borders and on the same model with adsorption
and desorption (no-hop processes), ragged sizes in one batch; and the
kmos.run.acf-style functions on a KMC_Model of the reference's 2d_grid
test model.
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


@pytest.mark.parametrize("with_adsorption", [False, True])
def test_tracking_matches_oracle(engine, oracle, with_adsorption):
    print("case adsorption=%s" % with_adsorption, flush=True)
    cm = compile_model(hop_project(with_adsorption), "local_smart")
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


def test_acf_functions_on_kmc_model(engine, oracle, tmp_path):
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
