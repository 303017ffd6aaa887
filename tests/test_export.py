"""Export backend on the CPU: kmc_settings.py keys, op program round trip,
and the project XML written into the settings re-compiles to the same
program (what kmc_model does when it loads an export directory)."""
import os

import numpy as np
import pytest

from kmos_amd.compiler import compile_model
from kmos_amd.export import export_source
from kmos_amd.kmc_model import compiled_for_export, load_settings
from kmos_amd.model import load_project, to_xml
from paritylib import model_file

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
KEYS = ["model_name", "simulation_size", "random_seed", "setup_model", "parameters", "rate_constants",
        "site_names", "representations", "lattice_representation", "species_tags", "tof_count", "xml"]


@pytest.mark.parametrize("src,backend", [("ruo2_default.xml", "local_smart"), ("ruo2_default.xml", "lat_int"),
                                         ("zgb_model.ini", "local_smart"), ("ab_model.ini", "lat_int")])
def test_export_directory(tmp_path, src, backend):
    d = export_source(model_file(src), str(tmp_path / "exp"), backend)
    s = load_settings(d)
    for k in KEYS:
        assert hasattr(s, k), k
    assert s.backend == backend
    pt = load_project(model_file(src))
    assert sorted(s.rate_constants) == sorted(p.name for p in pt.processes)
    cm = compiled_for_export(d)
    np.testing.assert_array_equal(cm.to_blob(), compile_model(pt, backend).to_blob())
    np.testing.assert_array_equal(np.load(os.path.join(d, "kmcg_model.npy")), cm.to_blob())


def test_export_from_xml_text(tmp_path):
    pt = load_project(model_file("ruo2_default.xml"))
    d = export_source(to_xml(pt), str(tmp_path / "x"), "local_smart")
    np.testing.assert_array_equal(compiled_for_export(d).to_blob(), compile_model(pt, "local_smart").to_blob())


def test_tampered_program_is_rejected(tmp_path):
    d = export_source(os.path.join(GOLDEN, "zgb_model.ini"), str(tmp_path / "z"), "local_smart")
    blob = np.load(os.path.join(d, "kmcg_model.npy"))
    blob[-1] ^= 1
    np.save(os.path.join(d, "kmcg_model.npy"), blob)
    with pytest.raises(RuntimeError):
        compiled_for_export(d)


@pytest.mark.parametrize("backend", ["local_smart", "lat_int"])
def test_multi_layer_model_exports(tmp_path, backend):
    """The reference's two-layer pdopd.xml (Pd100 and PdO layers, 25 sites
    per cell, create/annihilate actions, kmos/io/__init__.py:375-404,
    2445-2560) exports and re-compiles from its settings to the same
    program; its null species is nr_of_species and only the initial
    layer's sites start populated."""
    pt = load_project(os.path.join(GOLDEN, "pdopd.xml"))
    assert len(pt.layers) == 2
    d = export_source(os.path.join(GOLDEN, "pdopd.xml"), str(tmp_path / backend), backend)
    cm = compiled_for_export(d)
    np.testing.assert_array_equal(cm.to_blob(), compile_model(pt, backend).to_blob())
    assert cm.null_species == len(pt.species) == 4
    layers = [lname for lname, _ in pt.site_list()]
    assert [n for n in range(cm.spuck) if cm.site_default[n] != cm.null_species] == \
        [n for n in range(cm.spuck) if layers[n] == pt.default_layer] == \
        [n for n in range(cm.spuck) if cm.site_init[n]]


def test_otf_model_exports(tmp_path):
    """An otf export: the settings name the backend, the stored program is
    the compiled one and survives the XML round trip (bystanders and
    otf_rate included)."""
    from kmos_amd.kmc_model import compiled_for_export, load_settings
    d = export_source(os.path.join(GOLDEN, "intzgb_otf.xml"), str(tmp_path / "otf"), "otf")
    s = load_settings(d)
    assert s.backend == "otf"
    cm = compiled_for_export(d)
    assert cm.backend_name == "otf" and cm.nr_of_proc == 10
    assert cm.param_names == ["J_CO_CO", "J_O_CO", "J_O_O", "kDes", "yCO"]
    assert any(r.table_size > 1 for r in cm.otf_rates)
    stub = open(os.path.join(d, "kmc_model.py")).read()
    assert "proclist_pars" in stub
