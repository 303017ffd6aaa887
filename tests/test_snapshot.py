"""The .reload snapshot format against the reference's own writer and reader.

* tests/golden/ref_reload_*.reload were written by the reference's
  save_system (tests/golden/make_ref_reload.py); the reader must give back
  the state the reference held when it wrote them (its binary dump, in the
  matching .npz), to the digits the reference's formats keep;
* the oracle, loaded from such a file, must continue exactly as the
  reference continued after its own reload_system (same .npz);
* files this package writes round-trip every field bit for bit, an oracle
  restored from one continues exactly like the oracle it was saved from, and
  the live reference build (when present) reads them and continues the same
  way;
* malformed files raise.
"""
import glob
import os

import numpy as np
import pytest

from kmos_amd import workloads as W
from kmos_amd.compiler import compile_model
from kmos_amd.snapshot import ReloadError, read_reload, write_reload
from paritylib import oracle_replica

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
CASES = sorted(glob.glob(os.path.join(GOLDEN, "ref_reload_*.reload")))
IDS = [os.path.basename(p)[11:-7] for p in CASES]


def _backend(path):
    return "lat_int" if "lat_int" in path else "local_smart"


def _oracle_state_fields(o):
    sites, back = o.avail()
    return {"kmc_step": o.kmc_step, "kmc_time": o.kmc_time, "lattice": o.lattice(),
            "nr_of_sites": o.nr_of_sites(), "procstat": o.procstat(),
            "integ_rates": o.integ_rates(), "avail_sites": sites, "avail_sites_back": back}


def _assert_same_state(a, b, where=""):
    for k in ("kmc_step", "kmc_time"):
        assert a[k] == b[k], "%s %s" % (where, k)
    for k in ("lattice", "nr_of_sites", "procstat", "integ_rates", "avail_sites", "avail_sites_back"):
        np.testing.assert_array_equal(a[k], b[k], err_msg="%s %s" % (where, k))


@pytest.mark.parametrize("path", CASES, ids=IDS)
def test_reader_matches_reference_writer(path):
    d = np.load(path[:-7] + ".npz")
    st = read_reload(path)
    V = int(np.prod(d["size"])) * 2
    assert st["nr_of_proc"] == 36 and st["volume"] == V
    assert st["kmc_step"] == int(d["saved_kmc_step"])
    # es22.15 keeps 16 significant digits, es14.7 keeps 8
    assert abs(st["kmc_time"] - float(d["saved_kmc_time"])) <= 5e-16 * abs(float(d["saved_kmc_time"]))
    np.testing.assert_allclose(st["rates"], d["rates"], rtol=5e-8, atol=0)
    np.testing.assert_array_equal(st["lattice"], d["saved_lattice"])
    np.testing.assert_array_equal(st["nr_of_sites"], d["saved_nr"])
    np.testing.assert_array_equal(st["procstat"], d["saved_procstat"])
    np.testing.assert_array_equal(st["avail_sites"], d["saved_sites"])
    np.testing.assert_array_equal(st["avail_sites_back"], d["saved_back"])
    for key in ("integ_rates", "rng_counter", "rng_seed", "system_size", "kmc_time_step"):
        assert key not in st                       # the reference's file has none of them


@pytest.mark.parametrize("path", CASES, ids=IDS)
def test_oracle_continues_like_reference_reload(oracle, path):
    """reload_system then `cont` steps, on the reference and on the oracle."""
    d = np.load(path[:-7] + ".npz")
    cm = compile_model(W.ruo2_project(), _backend(path))
    st = read_reload(path)
    o = oracle_replica(oracle, cm, tuple(d["size"]), int(d["seed"]), int(d["stream"]), d["rates"])
    o.set_state(st)                                # rng continues at kmc_step
    assert o.rng_ctr == int(d["steps"])
    k, p, s, t = o.run(int(d["cont"]), trace=True)
    np.testing.assert_array_equal(p, d["cont_proc"])
    np.testing.assert_array_equal(s, d["cont_site"])
    np.testing.assert_array_equal(t, d["cont_time"])
    ref = {"kmc_step": int(d["cont_kmc_step"]), "kmc_time": float(d["cont_kmc_time"]),
           "lattice": d["cont_lattice"], "nr_of_sites": d["cont_nr"], "procstat": d["cont_procstat"],
           "integ_rates": d["cont_integ"], "avail_sites": d["cont_sites"], "avail_sites_back": d["cont_back"]}
    _assert_same_state(_oracle_state_fields(o), ref, "after reload")


def _awkward_rates(g, P):
    rates = 10.0 ** g.uniform(-2, 9, P)
    rates[::5] = 1.0 / 3.0                        # no short decimal form
    rates[1] = 0.0
    return rates


@pytest.mark.parametrize("backend", ["local_smart", "lat_int"])
def test_round_trip_and_oracle_continuation(oracle, tmp_path, backend):
    cm = compile_model(W.ruo2_project(), backend)
    g = np.random.default_rng(5)
    rates = _awkward_rates(g, cm.nr_of_proc)
    size = (7, 5)
    a = oracle_replica(oracle, cm, size, 11, 2, rates)
    a.run(1500)
    st = a.get_state(rates)
    st.update(rng_seed=11, rng_stream=2, system_size=[7, 5, 1], walltime=1.25)
    f = str(tmp_path / "a.reload")
    write_reload(f, st)
    back = read_reload(f)
    for k in ("kmc_time", "kmc_time_step", "kmc_step", "rng_counter", "rng_seed", "rng_stream"):
        assert back[k] == st[k], k
    for k in ("lattice", "nr_of_sites", "procstat", "integ_rates", "rates", "avail_sites",
              "avail_sites_back", "system_size"):
        np.testing.assert_array_equal(back[k], st[k], err_msg=k)
    # a second oracle restored from the file continues like the first
    b = oracle_replica(oracle, cm, size, 11, 2, np.ones(cm.nr_of_proc))
    b.set_state(back)
    a.run(1500)
    b.run(1500)
    _assert_same_state(_oracle_state_fields(b), _oracle_state_fields(a), "continued")


@pytest.mark.parametrize("backend", ["local_smart", "lat_int"])
def test_reference_reads_our_file(oracle, tmp_path, backend):
    """The reference's reload_system loads a file of write_reload (skipping
    its extra records) and continues like the oracle."""
    from refbin import driver_path, run_reference
    name = "ruo2_%s" % backend
    if driver_path(name) is None:
        pytest.skip("reference build %s absent (oracle/build_ref.py needs /root/reference)" % name)
    cm = compile_model(W.ruo2_project(), backend)
    g = np.random.default_rng(8)
    rates = _awkward_rates(g, cm.nr_of_proc)
    size = (6, 8)
    o = oracle_replica(oracle, cm, size, 21, 6, rates)
    o.run(1200)
    st = o.get_state(rates)
    st.update(rng_seed=21, rng_stream=6, system_size=[6, 8, 1])
    f = str(tmp_path / "ours.reload")
    write_reload(f, st)
    ref = run_reference(name, size, 21, 6, 800, np.ones(cm.nr_of_proc), cm.nr_of_proc, mode=3,
                        reload_text=open(f).read(), rng_step=o.rng_ctr)
    st_ref_view = dict(st)
    st_ref_view.pop("integ_rates")                # reload_system leaves them at zero
    o.set_state(st_ref_view)
    k, p, s, t = o.run(800, trace=True)
    np.testing.assert_array_equal(p, ref["proc"])
    np.testing.assert_array_equal(s, ref["site"])
    np.testing.assert_array_equal(t, ref["time"])
    mine = _oracle_state_fields(o)
    theirs = {"kmc_step": ref["kmc_step"], "kmc_time": ref["kmc_time"], "lattice": ref["lattice"],
              "nr_of_sites": ref["nr"], "procstat": ref["procstat"], "integ_rates": ref["integ"],
              "avail_sites": ref["sites"], "avail_sites_back": ref["back"]}
    _assert_same_state(mine, theirs, "reference after reading our file")


def _minimal(tmp_path, lines):
    f = str(tmp_path / "m.reload")
    with open(f, "w") as fh:
        fh.write("\n".join(lines) + "\n")
    return f


def test_reader_fortran_forms_and_unknown_labels(tmp_path):
    f = _minimal(tmp_path, [
        "# comment", "  kmc_time   1.5D-03", " kmc_step 12", " nr_of_proc 2", " volume 3",
        "future_label 1 2 3", "procstat 4 5", "nr_of_sites 1 0", "rates 1.0000000E+00 2.5d0",
        "lattice 0 1 0", "avail_sites 1 2 0 0", "avail_sites 2 0 0 0",
        "avail_sites_back 1 0 1 0", "avail_sites_back 2 0 0 0"])
    st = read_reload(f)
    assert st["kmc_time"] == 1.5e-3 and st["kmc_step"] == 12
    np.testing.assert_array_equal(st["rates"], [1.0, 2.5])
    np.testing.assert_array_equal(st["avail_sites"], [[2, 0, 0], [0, 0, 0]])
    np.testing.assert_array_equal(st["avail_sites_back"], [[0, 1, 0], [0, 0, 0]])
    assert "future_label" not in st


@pytest.mark.parametrize("broken, msg", [
    ([" kmc_time 0.0", " kmc_step 0", " nr_of_proc 1", "procstat 0", "nr_of_sites 0", "rates 1",
      "lattice 0"], "volume are required"),
    ([" kmc_time 0.0", " kmc_step 0", " nr_of_proc 1", " volume 2", "procstat 0", "nr_of_sites 0",
      "rates 1", "lattice 0"], "lattice needs 2"),
    ([" kmc_time 0.0", " kmc_step 0", " nr_of_proc 1", " volume 1", "procstat 0", "nr_of_sites 0",
      "rates 1", "lattice 0", "avail_sites 2 0"], "outside 1..1"),
    ([" kmc_time x", " kmc_step 0", " nr_of_proc 1", " volume 1", "procstat 0", "nr_of_sites 0",
      "rates 1", "lattice 0"], "cannot read kmc_time"),
    ([" kmc_step 0", " nr_of_proc 1", " volume 1", "procstat 0", "nr_of_sites 0", "rates 1",
      "lattice 0"], "kmc_time is missing"),
])
def test_reader_rejects_malformed(tmp_path, broken, msg):
    with pytest.raises(ReloadError, match=msg):
        read_reload(_minimal(tmp_path, broken))
