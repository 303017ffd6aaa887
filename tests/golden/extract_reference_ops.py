#!/usr/bin/env python3
"""Extract golden op sequences from the reference's generated Fortran.

Reads the Fortran that the reference ships as export fixtures (under
/root/reference/tests) and writes, as JSON data, the ordered bookkeeping
operations each generated routine performs:

local_smart (proclist.f90)
  routines  put_/take_<species>_<layer>_<site>, touchup_<layer>_<site> and
            (multi-layer models) create_/annihilate_<layer>_<site>:
            ["del", proc, [dx,dy,dz,n]] in text order, then
            ["add", proc, [dx,dy,dz,n], [[dx,dy,dz,n,species], ...]] in text
            order, the list being the select-case path that guards the add
  run_proc  per process: [[routine, [dx,dy,dz,n]], ...], with the species
            argument appended for create_/annihilate_ calls
lat_int (run_proc_*.f90, nli_*.f90, proclist.f90)
  groups    per group: dels [[group, [dx,dy,dz]]], updates
            [[[dx,dy,dz,n], [old species], new]], adds (same as dels)
  nli       per group: leaves [[[[dx,dy,dz,n,[species]]...], result]] in
            text order (result 0 for the default branches)
  dispatch  process -> group
otf (run_proc_*.f90, proclist.f90, proclist_pars.f90)
  run_proc  per process, in text order: ["del", proc, [dx,dy,dz]],
            ["write", [dx,dy,dz,n], old, new], ["upd", proc, [dx,dy,dz]] and
            ["add", proc, [dx,dy,dz], [[dx,dy,dz,n,species], ...]] (the
            select-case path guarding the add); processes sit on site 1 of
            the cell, so del/upd/add carry the cell offset only
  touchup   the adds of touchup_cell (its generic delete loop aside)
  rates     per process: "vars", the nr_<species>_<flag> names in nr_vars
            order (byst_<proc>); "bystanders", [[dx,dy,dz,n], {species:
            [1-based nr_vars indices]}] in text order; "expr", the rate_<proc>
            right-hand side with continuation lines joined

Coordinates are made absolute: n is the 1-based site index in layer order.
Usage: python tests/golden/extract_reference_ops.py [--ref /root/reference]
"""
import argparse
import glob
import json
import os
import re
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))

FIXTURES = [
    # (output name, fixture dir, model file, backend)
    ("ruo2_local_smart", "tests/export_test/reference_export", "ruo2_default.xml", "local_smart"),
    ("ruo2_lat_int", "tests/export_test/reference_export_lat_int", "ruo2_default.xml", "lat_int"),
    ("cli_local_smart", "tests/test_cli_generated_model", "cli_model.xml", "local_smart"),
    ("pdopd_local_smart", "tests/export_test/reference_pdopd_local_smart", "pdopd.xml", "local_smart"),
    ("pdopd_lat_int", "tests/export_test/reference_pdopd_lat_int", "pdopd.xml", "lat_int"),
    ("ruo2_otf", "tests/export_test/reference_export_otf", "ruo2_default.xml", "otf"),
    ("intzgb_otf", "tests/export_test/reference_export_intZGB_otf", "intzgb_otf.xml", "otf"),
]

VEC = re.compile(r"\(/\s*([^/]*)/\)")


def site_names(model_file):
    sys.path.insert(0, REPO)
    from kmos_amd.model import load_project
    path = os.path.join(HERE, model_file)
    if not os.path.exists(path):        # the models the package ships
        path = os.path.join(REPO, "kmos_amd", "data", model_file)
    pt = load_project(path)
    return ["%s_%s" % ls for ls in pt.site_list()], pt


def parse_coord(text, own_n, names):
    """'site' / 'site + (/dx, dy, dz, X/)' -> [dx, dy, dz, n]."""
    m = VEC.search(text)
    if not m:
        return [0, 0, 0, own_n]
    parts = [p.strip() for p in m.group(1).split(",")]
    dx, dy, dz = (int(p) for p in parts[:3])
    last = parts[3]
    if last == "0":
        n = own_n
    elif " - " in last:
        n = names.index(last.split(" - ")[0].strip()) + 1
    elif last in names:
        n = names.index(last) + 1
    else:
        n = int(last)
    return [dx, dy, dz, n]


def subroutines(text):
    out = {}
    for m in re.finditer(r"^subroutine (\w+)\((.*?)\)\s*$(.*?)^end subroutine \1", text, re.M | re.S):
        out[m.group(1)] = m.group(3)
    return out


def parse_local_smart(fdir, names):
    text = open(os.path.join(fdir, "proclist.f90")).read()
    subs = subroutines(text)
    routines = {}
    for name, body in subs.items():
        if not name.startswith(("put_", "take_", "touchup_", "create_", "annihilate_")):
            continue
        own = None
        for i, nm in enumerate(names):
            if name.endswith("_" + nm):
                own = i + 1
        ops = []
        path = []            # stack of [coord, species or None]
        pending_coord = []
        for line in body.splitlines():
            s = line.strip()
            if not s or s.startswith("!"):
                continue
            m = re.match(r"call del_proc\((\w+),\s*(.*)\)$", s)
            if m:
                ops.append(["del", m.group(1), parse_coord(m.group(2), own, names)])
                continue
            m = re.match(r"select case\(get_species\((.*)\)\)$", s)
            if m:
                pending_coord.append(parse_coord(m.group(1), own, names))
                path.append(None)
                continue
            m = re.match(r"case\((\w+)\)$", s)
            if m and pending_coord:
                path[-1] = pending_coord[-1] + [m.group(1)]
                continue
            if s == "end select":
                pending_coord.pop()
                path.pop()
                continue
            m = re.match(r"call add_proc\((\w+),\s*(.*)\)$", s)
            if m:
                conds = [list(p) for p in path if p is not None]
                ops.append(["add", m.group(1), parse_coord(m.group(2), own, names), conds])
        routines[name] = ops
    run_proc = {}
    body = subs["run_proc_nr"]
    cur = None
    for line in body.splitlines():
        s = line.strip()
        m = re.match(r"case\((\w+)\)$", s)
        if m:
            cur = m.group(1)
            run_proc[cur] = []
            continue
        m = re.match(r"call ((?:put|take)_\w+)\((.*)\)$", s)
        if m and cur:
            rname = m.group(1)
            own = [i + 1 for i, nm in enumerate(names) if rname.endswith("_" + nm)][0]
            run_proc[cur].append([rname, parse_coord(m.group(2), own, names)])
            continue
        m = re.match(r"call ((?:create|annihilate)_\w+)\((.*), (\w+)\)$", s)
        if m and cur:
            rname = m.group(1)
            own = [i + 1 for i, nm in enumerate(names) if rname.endswith("_" + nm)][0]
            run_proc[cur].append([rname, parse_coord(m.group(2), own, names), m.group(3)])
    return {"routines": routines, "run_proc": run_proc}


def parse_lat_int(fdir, names):
    groups = {}
    for path in sorted(glob.glob(os.path.join(fdir, "run_proc_*.f90"))):
        text = open(path).read()
        m = re.search(r"subroutine run_proc_(\w+)\(cell\)", text)
        gname = m.group(1)
        dels, adds, updates = [], [], []
        for line in text.splitlines():
            s = line.strip()
            m = re.match(r"call (del_proc|add_proc)\(nli_(\w+)\(cell \+ \(/([^/]*)/\)\), cell", s)
            if m:
                off = [int(v) for v in m.group(3).split(",")[:3]]
                (dels if m.group(1) == "del_proc" else adds).append([m.group(2), off])
                continue
            m = re.match(r"call replace_species\(cell(.*?), (\w+), (\w+)\)$", s)
            if m:
                updates.append([parse_coord(m.group(1), 1, names), [m.group(2)], m.group(3)])
        groups[gname] = {"dels": dels, "updates": updates, "adds": adds}
    nli = {}
    for path in sorted(glob.glob(os.path.join(fdir, "nli_*.f90"))):
        text = open(path).read()
        m = re.search(r"function nli_(\w+)\(cell\)", text)
        gname = m.group(1)
        leaves = []
        stack = []
        pending = []
        for line in text.splitlines():
            s = line.strip()
            m = re.match(r"select case\(get_species\(cell(.*)\)\)$", s)
            if m:
                pending.append(parse_coord(m.group(1), 1, names))
                stack.append(None)
                continue
            m = re.match(r"case\((.*)\)$", s)
            if m and pending:
                stack[-1] = pending[-1] + [[x.strip() for x in m.group(1).split(",")]]
                continue
            if s == "case default" and pending:
                stack[-1] = pending[-1] + [["default"]]
                continue
            if s == "end select":
                pending.pop()
                stack.pop()
                continue
            m = re.match(r"nli_\w+ = (\w+); return$", s)
            if m:
                leaves.append([[list(p) for p in stack if p is not None], m.group(1)])
        nli[gname] = leaves
    text = open(os.path.join(fdir, "proclist.f90")).read()
    body = subroutines(text)["run_proc_nr"]
    dispatch = {}
    body = body.replace("&\n", " ")
    for m in re.finditer(r"case\(([^)]*)\)\s*\n\s*call run_proc_(\w+)\(cell\)", body):
        for p in m.group(1).replace("&", " ").split(","):
            dispatch[p.strip()] = m.group(2)
    tu = subroutines(text)["touchup_cell"]
    touch_groups = re.findall(r"call add_proc\(nli_(\w+)\(cell\), site\)", tu)
    return {"groups": groups, "nli": nli, "dispatch": dispatch, "touchup_groups": touch_groups}


def _otf_ops(body, names):
    """The ordered bookkeeping ops of one otf run_proc_<proc> or touchup_cell."""
    ops = []
    path, pending = [], []
    for line in body.replace("&\n", " ").splitlines():
        s = re.sub(r"\s+", " ", line.strip())
        if not s or s.startswith("!"):
            continue
        m = re.match(r"call del_proc\((\w+), ?cell \+ \(/([^/]*)/\)\)$", s)
        if m:
            ops.append(["del", m.group(1), [int(v) for v in m.group(2).split(",")[:3]]])
            continue
        m = re.match(r"call replace_species\(cell(.*?), ?(\w+), ?(\w+)\)$", s)
        if m:
            ops.append(["write", parse_coord(m.group(1), 1, names), m.group(2), m.group(3)])
            continue
        m = re.match(r"call update_rates_matrix\((\w+), ?cell \+ \(/([^/]*)/\), ?gr_", s)
        if m:
            ops.append(["upd", m.group(1), [int(v) for v in m.group(2).split(",")[:3]]])
            continue
        m = re.match(r"select case\(get_species\(cell(.*)\)\)$", s)
        if m:
            pending.append(parse_coord(m.group(1), 1, names))
            path.append(None)
            continue
        m = re.match(r"case\((\w+)\)$", s)
        if m and pending:
            path[-1] = pending[-1] + [m.group(1)]
            continue
        if s == "end select":
            pending.pop()
            path.pop()
            continue
        m = re.match(r"call add_proc\((\w+), ?cell \+ \(/([^/]*)/\), ?gr_", s)
        if m:
            ops.append(["add", m.group(1), [int(v) for v in m.group(2).split(",")[:3]],
                        [list(p) for p in path if p is not None]])
    return ops


def _functions(text):
    out = {}
    for m in re.finditer(r"^function (\w+)\((.*?)\)\s*$(.*?)^end function \1", text, re.M | re.S):
        out[m.group(1)] = m.group(3)
    return out


def parse_otf(fdir, names):
    run_proc = {}
    for path in sorted(glob.glob(os.path.join(fdir, "run_proc_*.f90"))):
        for name, body in subroutines(open(path).read()).items():
            if name.startswith("run_proc_"):
                run_proc[name[len("run_proc_"):]] = _otf_ops(body, names)
    text = open(os.path.join(fdir, "proclist.f90")).read()
    touchup = _otf_ops(subroutines(text)["touchup_cell"], names)
    pars = open(os.path.join(fdir, "proclist_pars.f90")).read()
    joined = re.sub(r"&\s*\n\s*&", "", pars)
    funcs = _functions(joined)
    rates = {}
    for m in re.finditer(r'byst_(\w+) = "([^"]*)"', joined):
        proc = m.group(1)
        entry = {"vars": m.group(2).split(), "bystanders": []}
        cur = None
        for line in funcs["gr_" + proc].splitlines():
            s = re.sub(r"\s+", " ", line.strip())
            mm = re.match(r"select case\(get_species\(cell(.*)\)\)$", s)
            if mm:
                entry["bystanders"].append([parse_coord(mm.group(1), 1, names), {}])
                continue
            mm = re.match(r"case\((\w+)\)$", s)
            if mm:
                cur = mm.group(1)
                entry["bystanders"][-1][1][cur] = []
                continue
            mm = re.match(r"nr_vars\((\d+)\) = nr_vars\(\d+\) \+ 1$", s)
            if mm:
                entry["bystanders"][-1][1][cur].append(int(mm.group(1)))
        body = funcs["rate_" + proc]
        mm = re.search(r"^\s*rate_%s = (.*)$" % proc, body, re.M)
        entry["expr"] = re.sub(r"\s+", " ", mm.group(1)).strip()
        rates[proc] = entry
    return {"run_proc": run_proc, "touchup": touchup, "rates": rates}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--ref", default="/root/reference")
    a = ap.parse_args()
    for out, fdir, model, backend in FIXTURES:
        names, pt = site_names(model)
        src = os.path.join(a.ref, fdir)
        parse = {"local_smart": parse_local_smart, "lat_int": parse_lat_int, "otf": parse_otf}[backend]
        data = parse(src, names)
        data["model"] = model
        data["backend"] = backend
        data["site_names"] = names
        data["source"] = fdir
        path = os.path.join(HERE, "ref_ops_%s.json" % out)
        with open(path, "w") as fh:
            json.dump(data, fh, indent=0, sort_keys=True)
        print(path)


if __name__ == "__main__":
    main()
