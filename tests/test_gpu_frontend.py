"""The kmos caller path on the engine: export -> kmc_model -> KMC_Model.

A model exported by kmos_amd.export runs through the f2py-shaped module and
the KMC_Model front-end exactly like the oracle replica set up the same way
(proclist.init with the settings' seed, rate constants from the settings'
expressions, do_steps, set/get configuration, sampling).
"""
import os

import numpy as np
import pytest

from kmos_amd.export import export_source
from kmos_amd.model import evaluate_rate_expression
from kmos_amd.run import KMC_Model
from paritylib import compare_state, model_file, oracle_replica

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


@pytest.mark.parametrize("backend", ["local_smart", "lat_int"])
def test_kmc_model_matches_oracle(engine, oracle, tmp_path, backend):
    d = export_source(model_file("ruo2_default.xml"), str(tmp_path / backend), backend)
    model = KMC_Model(export_dir=d, size=[10, 8], random_seed=42)
    cm = model.compiled
    rates = np.array([evaluate_rate_expression(model.settings.rate_constants[p][0], model.settings_parameters)
                      for p in cm.process_names])
    o = oracle_replica(oracle, cm, (10, 8), 42, 0, rates)
    model.do_steps(1234.7)                 # truncated like f2py's int conversion
    o.run(1234)
    compare_state(model._kmc.batch, 0, o, where="KMC_Model %s" % backend)
    assert model.base.get_kmc_step() == 1234
    assert model.base.get_procstat(1) == o.procstat()[0]
    # rate constants through the parameters interface
    model.rate_constants.set("co_adsorption_*", 5.0)
    rates[[cm.process_names.index("co_adsorption_bridge"), cm.process_names.index("co_adsorption_cus")]] = 5.0
    o.set_rates(rates)
    model.do_steps(500)
    o.run(500)
    compare_state(model._kmc.batch, 0, o, where="after set()")
    # configuration round trip and _adjust_database
    cfg = model._get_configuration()
    g = np.random.default_rng(0)
    new = g.integers(0, 3, cfg.shape).astype(np.int8)
    model._set_configuration(new)
    o.set_config(new.transpose(2, 1, 0, 3).reshape(-1))
    np.testing.assert_array_equal(model._get_configuration(), new)
    model.do_steps(300)
    o.run(300)
    compare_state(model._kmc.batch, 0, o, where="after set_configuration")
    # sampling interface
    out = model.get_std_sampled_data(4, 400, output="dict")
    assert out["kmc_steps"] == 400
    occ = [v for k, v in out.items() if k.endswith("_ruo2_bridge") or k.endswith("_ruo2_cus")]
    assert abs(sum(occ) - 2.0) < 1e-9      # occupations of each site type sum to one
    model.deallocate()
    assert not model.base.is_allocated()


def test_f2py_surface_names(engine, tmp_path):
    d = export_source(os.path.join(GOLDEN, "zgb_model.ini"), str(tmp_path / "zgb"), "local_smart")
    import importlib.util
    spec = importlib.util.spec_from_file_location("kmc_model_zgb", os.path.join(d, "kmc_model.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    base, lattice, proclist = mod.base, mod.lattice, mod.proclist
    assert proclist.nr_of_proc == 10 and proclist.co_adsorption == 1
    assert proclist.co == 0 and proclist.o == 1 and proclist.empty == 2
    assert lattice.sc_site == 1 and lattice.spuck == 1
    proclist.init([16, 16], "zgb", lattice.default_layer, 7, True)
    for p in range(1, proclist.nr_of_proc + 1):
        base.set_rate_const(p, 1.0)
    proclist.do_kmc_steps(200)
    assert base.get_kmc_step() == 200
    occ = proclist.get_occupation()
    assert occ.shape == (3, 1) and abs(occ.sum() - 1.0) < 1e-12
    p, s = proclist.get_next_kmc_step()
    assert 1 <= p <= 10 and 1 <= s <= 256
    assert lattice.get_species([0, 0, 0, 1]) in (0, 1, 2)
    base.deallocate_system()
