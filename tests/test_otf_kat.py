"""otf export backend against the reference's generated otf Fortran.

The op programs of every run_proc_<proc> routine and of touchup_cell, the
nr_vars order of every gr_<proc> and its bystander counts are compared with
what the reference's otf writer produced for two models it ships as
fixtures (tests/export_test/reference_export_otf: the RuO2 model, 36
processes, no bystanders; reference_export_intZGB_otf: ZGB with lateral
interactions through bystanders and otf_rate expressions), extracted to
tests/golden/ref_ops_{ruo2,intzgb}_otf.json by extract_reference_ops.py.
The rate programs are checked by value: the table of a CO oxidation
process against the closed form of its expression.
"""
import json
import os

import numpy as np
import pytest

from kmos_amd import otf, workloads as W
from kmos_amd.compiler import compile_model
from kmos_amd.model import UNITS, load_project

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _project(name):
    if name == "ruo2":
        # the fixture was exported with shortened names
        # (tests/export_test/test_import_export.py:93)
        pt = W.ruo2_project()
        pt.shorten_names(max_length=35)
        assert [p.name for p in pt.processes] == sorted(p.name for p in pt.processes)
        return pt
    return load_project(os.path.join(GOLDEN, "intzgb_otf.xml"))


def _as_golden(cm, op):
    sp = cm.species_names
    names = cm.process_names
    if op.kind == otf.OP_DEL:
        return ["del", names[op.proc], list(op.off)]
    if op.kind == otf.OP_UPD:
        return ["upd", names[op.proc], list(op.off)]
    if op.kind == otf.OP_WRITE:
        return ["write", list(op.off) + [op.site], sp[op.old], sp[op.new]]
    return ["add", names[op.proc], list(op.off), [list(c) + [sp[s]] for c, s in op.conds]]


@pytest.mark.parametrize("name", ["ruo2", "intzgb"])
def test_run_proc_programs_match_reference(name):
    ref = json.load(open(os.path.join(GOLDEN, "ref_ops_%s_otf.json" % name)))
    cm = compile_model(_project(name), "otf")
    assert sorted(ref["run_proc"]) == sorted(cm.process_names)
    total = 0
    for p, pname in enumerate(cm.process_names):
        mine = [_as_golden(cm, op) for op in cm.otf_programs[p]]
        assert mine == ref["run_proc"][pname], pname
        total += len(mine)
    assert total > 500
    assert [_as_golden(cm, op) for op in cm.otf_touchup] == ref["touchup"]


@pytest.mark.parametrize("name", ["ruo2", "intzgb"])
def test_rate_functions_match_reference(name):
    ref = json.load(open(os.path.join(GOLDEN, "ref_ops_%s_otf.json" % name)))["rates"]
    pt = _project(name)
    cm = compile_model(pt, "otf")
    for p, proc in enumerate(pt.processes):
        r = cm.otf_rates[p]
        g = ref[proc.name]
        assert r.var_names == g["vars"], proc.name
        mine = []
        for b in proc.bystanders:
            n = pt.site_index(b.coord.layer, b.coord.name)
            mine.append([list(b.coord.offset) + [n],
                         {sp: [r.var_names.index("nr_%s_%s" % (sp, f)) + 1 for f in b.flag.split()]
                          for sp in b.allowed_species}])
        assert mine == g["bystanders"], proc.name
        # the table strides: one bystander's species moves the index by the
        # stride of every var it counts
        for (site, inc), b in zip(r.bystanders, proc.bystanders):
            for sp in b.allowed_species:
                exp = sum(r.strides[r.used.index(v - 1)] for v in
                          [r.var_names.index("nr_%s_%s" % (sp, f)) + 1 for f in b.flag.split()]
                          if v - 1 in r.used)
                assert inc[cm.species_names.index(sp)] == exp


def test_rate_table_closed_form():
    """CO_oxidation_00: base*(J_O_O**(-nOO))*(J_CO_CO**(-nCC))*(J_O_CO**(-nOC-nCO))
    on every count combination, against the same products formed with
    Python's float powers (to a few ulp: binary
    powering and pow round differently in the last places)."""
    pt = _project("intzgb")
    cm = compile_model(pt, "otf")
    p = cm.process_names.index("CO_oxidation_00")
    r = cm.otf_rates[p]
    par = dict(zip(cm.param_names, [1.3, 0.8, 1.7, 1e-10, 0.45]))
    base = 2.5
    t = otf.eval_rate_table(r, base, [par[k] for k in cm.param_names])
    assert t.shape == (256,)
    v = dict(zip(r.var_names, range(4)))
    for idx in range(256):
        c = [(idx // r.strides[i]) % (r.maxc[i] + 1) for i in range(4)]
        n = {name: c[r.used.index(v[name])] for name in r.var_names}
        want = base * par["J_O_O"] ** (-n["nr_O_Onn"]) * par["J_CO_CO"] ** (-n["nr_CO_COnn"]) \
            * par["J_O_CO"] ** (-n["nr_O_COnn"] - n["nr_CO_Onn"])
        assert abs(t[idx] - want) <= 2e-15 * want, (idx, c)
    # no environment: the base rate itself
    assert t[0] == base


def test_powi_matches_the_reference_compiler_form():
    """x**n as binary powering from the low bit with 1/r for negative n:
    hand-checked cases (bit patterns from the reference compiler, flang -O3,
    for x = 1.2345678901)."""
    x = 1.2345678901
    assert otf._powi(x, 0) == 1.0
    assert otf._powi(x, 1) == x
    assert otf._powi(x, 2) == x * x
    assert otf._powi(x, 3) == x * (x * x)
    assert otf._powi(x, 5) == x * ((x * x) * (x * x))
    assert otf._powi(x, -3) == 1.0 / (x * (x * x))
    assert otf._powi(x, 3).hex() == float.fromhex("0x1.e1b58af07447ap+0").hex()
    assert otf._powi(x, -3).hex() == float.fromhex("0x1.1019096106013p-1").hex()


def test_fortran_kinds_in_rate_programs():
    """Single-precision literals stay single until they meet a double;
    integer division truncates; integer powers of integers stay integers."""
    pt = _project("intzgb")
    proc = pt.processes[1]
    proc.otf_rate = "base_rate*(0.1*0.2)"
    r = otf.compile_rate(pt, proc, UNITS)
    t = otf.eval_rate_table(r, 3.0, [1.0] * 5)
    assert t[0] == 3.0 * float(np.float32(np.float32(0.1) * np.float32(0.2)))
    proc.otf_rate = "base_rate*(7/2)*(2**3)"
    t = otf.eval_rate_table(otf.compile_rate(pt, proc, UNITS), 1.0, [1.0] * 5)
    assert t[0] == 24.0
    proc.otf_rate = "aux = base_rate*2\\notf_rate = aux + 1"
    t = otf.eval_rate_table(otf.compile_rate(pt, proc, UNITS), 1.5, [1.0] * 5)
    assert t[0] == 4.0
    proc.otf_rate = "base_rate*mu_CO"
    with pytest.raises(NotImplementedError):
        otf.compile_rate(pt, proc, UNITS)
