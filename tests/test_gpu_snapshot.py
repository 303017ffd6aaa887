"""State snapshots on the engine: kmcg_set_state and the .reload files.

* a replica loaded from a file the reference's save_system wrote continues
  like the reference after its own reload_system (tests/golden/
  ref_reload_*: make_ref_reload.py);
* a replica saved to a file and restored into another replica with the same
  seed and stream continues bit for bit like the original, and both agree
  with the oracle that ran the steps without a break;
* lists that break the add_proc/del_proc invariant are rejected before any
  byte reaches the device;
* base.save_system / base.reload_system of the kmc_model module.
"""
import glob
import os

import numpy as np
import pytest

from kmos_amd import workloads as W
from kmos_amd._lib import KmcgError
from kmos_amd.compiler import compile_model
from kmos_amd.snapshot import read_reload, write_reload
from paritylib import compare_state, model_file, oracle_replica

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
CASES = sorted(glob.glob(os.path.join(GOLDEN, "ref_reload_*.reload")))


def _assert_state(b, r, ref, where):
    t, dt, ks, ctr, st = b.scalars(r)
    assert st == 0, where
    assert ks == ref["kmc_step"], where
    np.testing.assert_allclose(t, ref["kmc_time"], rtol=1e-12, atol=0, err_msg=where)
    np.testing.assert_array_equal(b.lattice(r), ref["lattice"], err_msg=where)
    np.testing.assert_array_equal(b.nr_of_sites(r), ref["nr"], err_msg=where)
    np.testing.assert_array_equal(b.procstat(r), ref["procstat"], err_msg=where)
    np.testing.assert_allclose(b.integ_rates(r), ref["integ"], rtol=1e-9, atol=0, err_msg=where)
    sites, back = b.avail(r)
    np.testing.assert_array_equal(sites, ref["sites"], err_msg=where)
    np.testing.assert_array_equal(back, ref["back"], err_msg=where)


@pytest.mark.parametrize("path", CASES, ids=[os.path.basename(p)[11:-7] for p in CASES])
def test_engine_continues_reference_reload(engine, path):
    print("case", os.path.basename(path), flush=True)
    d = np.load(path[:-7] + ".npz")
    backend = "lat_int" if "lat_int" in path else "local_smart"
    m = engine.Model(compile_model(W.ruo2_project(), backend))
    b = engine.Batch(m, [list(d["size"]) + [1]])
    b.init([int(d["seed"])], [int(d["stream"])])
    b.set_state(0, read_reload(path))
    saved = {k: d["saved_" + k] for k in ("lattice", "nr", "procstat", "sites", "back")}
    saved.update(kmc_step=int(d["saved_kmc_step"]), kmc_time=float(d["saved_kmc_time"]),
                 integ=np.zeros(b.P))
    _assert_state(b, 0, saved, "loaded")
    assert b.scalars(0)[3] == int(d["steps"])         # the stream continues at kmc_step
    b.run(int(d["cont"]))
    cont = {k: d["cont_" + k] for k in ("lattice", "nr", "procstat", "integ", "sites", "back")}
    cont.update(kmc_step=int(d["cont_kmc_step"]), kmc_time=float(d["cont_kmc_time"]))
    _assert_state(b, 0, cont, "continued")


SAVE_CASES = [
    ("ruo2_local_smart", W.ruo2_project, "local_smart", (20, 20)),
    ("ruo2_lat_int", W.ruo2_project, "lat_int", (20, 20)),
    ("zgb_local_smart", W.zgb_project, "local_smart", (32, 32)),
    ("pairwise_lat_int", W.pairwise_project, "lat_int", (12, 10)),
]


@pytest.mark.parametrize("name,factory,backend,size", SAVE_CASES, ids=[c[0] for c in SAVE_CASES])
def test_save_restore_continues_exactly(engine, oracle, tmp_path, name, factory, backend, size):
    print("case", name, flush=True)
    cm = compile_model(factory(), backend)
    P = cm.nr_of_proc
    rates = 10.0 ** np.random.default_rng(3).uniform(-2, 6, P)
    m = engine.Model(cm)
    b = engine.Batch(m, [list(size) + [1]] * 2)
    b.init(np.array([9, 9], np.uint32), np.array([4, 4], np.uint32))
    b.set_rates(np.stack([rates, np.ones(P)]))
    b.run(2500, replicas=[0])
    st = b.get_state(0)
    f = str(tmp_path / "r0.reload")
    write_reload(f, st)
    b.set_state(1, read_reload(f))
    np.testing.assert_array_equal(b.get_rates(1), rates)
    for r in (0, 1):
        assert b.scalars(r)[:4] == b.scalars(0)[:4]
    b.run(2500)
    o = oracle_replica(oracle, cm, size, 9, 4, rates)
    o.run(5000)
    compare_state(b, 0, o, where="%s uninterrupted" % name)
    compare_state(b, 1, o, where="%s restored" % name)
    s0, s1 = b.get_state(0), b.get_state(1)
    for k in ("kmc_time", "kmc_time_step", "kmc_step", "rng_counter"):
        assert s0[k] == s1[k], k
    for k in ("lattice", "nr_of_sites", "procstat", "integ_rates", "avail_sites", "avail_sites_back"):
        np.testing.assert_array_equal(s0[k], s1[k], err_msg=k)


def test_set_state_rejects_broken_lists(engine):
    cm = compile_model(W.ruo2_project(), "local_smart")
    m = engine.Model(cm)
    b = engine.Batch(m, [[8, 8, 1]])
    b.init([1], [0])
    b.set_rates(np.ones((1, b.P)))
    b.run(300)
    good = b.get_state(0)
    p = int(np.argmax(good["nr_of_sites"]))
    assert good["nr_of_sites"][p] >= 2

    def broken(edit):
        st = {k: (v.copy() if isinstance(v, np.ndarray) else v) for k, v in good.items()}
        edit(st)
        return st
    site0 = int(good["avail_sites"][p, 0])
    bad = [
        (lambda s: s["avail_sites_back"].__setitem__((p, site0 - 1), 2), "disagree"),
        (lambda s: s["nr_of_sites"].__setitem__(p, 10**6), "nr_of_sites out of range"),
        (lambda s: s["nr_of_sites"].__setitem__(p, s["nr_of_sites"][p] - 1), "beyond nr_of_sites"),
        (lambda s: s["avail_sites"].__setitem__((p, 0), 10**6), "disagree"),
    ]
    for edit, msg in bad:
        with pytest.raises(KmcgError, match=msg):
            b.set_state(0, broken(edit))
    # an unknown species is refused by the host wrapper before the call
    with pytest.raises(ValueError, match="species ids"):
        b.set_state(0, broken(lambda s: s["lattice"].__setitem__(0, 9)))
    after = b.get_state(0)
    for k in ("lattice", "nr_of_sites", "avail_sites", "avail_sites_back", "procstat"):
        np.testing.assert_array_equal(after[k], good[k], err_msg=k)
    b.run(100)                                        # the replica is still sound
    assert b.scalars(0)[4] == 0


def test_kmc_model_save_and_reload_system(engine, oracle, tmp_path, monkeypatch):
    from kmos_amd.export import export_source
    from kmos_amd.model import evaluate_rate_expression
    from kmos_amd.run import KMC_Model
    d = export_source(model_file("ruo2_default.xml"), str(tmp_path / "exp"), "local_smart")
    model = KMC_Model(export_dir=d, size=[10, 8], random_seed=42)
    cm = model.compiled
    rates = np.array([evaluate_rate_expression(model.settings.rate_constants[p][0], model.settings_parameters)
                      for p in cm.process_names])
    monkeypatch.chdir(tmp_path)
    model.do_steps(1000)
    model.base.set_system_name("snap")
    model.base.save_system()
    assert os.path.exists(tmp_path / "snap.reload")
    model.base.deallocate_system()
    assert not model.base.reload_system("missing")
    assert model.base.reload_system("snap")
    assert model.base.get_kmc_step() == 1000
    np.testing.assert_array_equal(model.lattice.system_size, [10, 8, 1])
    model.do_steps(700)
    o = oracle_replica(oracle, cm, (10, 8), 42, 0, rates)
    o.run(1700)
    compare_state(model._kmc.batch, 0, o, where="kmc_model reload")
