"""The export backend reproduces the reference generator's op order.

Golden data: tests/golden/ref_ops_*.json, extracted from the Fortran the
reference ships as export fixtures (tests/golden/extract_reference_ops.py).
The compiled op programs must list the same deletions, additions (with the
same guarding conditions, in the same if-tree text order), lattice updates,
nli case trees and process dispatch.
"""
import json
import os

import pytest

from kmos_amd.compiler import compile_model, iftree_order, multilattice_routine, _most_common
from kmos_amd.model import load_project
from paritylib import model_file

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _golden(name):
    with open(os.path.join(GOLDEN, "ref_ops_%s.json" % name)) as fh:
        return json.load(fh)


def _local_smart_ops(cm, r):
    procs = cm.process_names
    sp = cm.species_names
    ops = [["del", procs[p - 1], list(s)] for p, s in r.disables]
    for e in r.enables:
        conds = [list(c) + [sp[s]] for c, s in e.conds]
        ops.append(["add", procs[e.proc - 1], list(e.site), conds])
    return ops


LOCAL_SMART = ["ruo2_local_smart", "cli_local_smart", "pdopd_local_smart"]


@pytest.mark.parametrize("name", LOCAL_SMART)
def test_local_smart_routines_match_reference(name):
    ref = _golden(name)
    pt = load_project(model_file(ref["model"]))
    cm = compile_model(pt, "local_smart")
    # put/take/touchup routines; create/annihilate ones (named with their
    # species here) are checked below
    mine = {r.name: _local_smart_ops(cm, r) for r in cm.routines if ":" not in r.name}
    theirs = {k: v for k, v in ref["routines"].items() if not k.startswith(("create_", "annihilate_"))}
    # the reference only writes put/take routines for non-default species
    assert set(theirs) == set(mine)
    for rname, ops in theirs.items():
        assert mine[rname] == ops, rname


def test_multilattice_routines_match_reference():
    """Every create_/annihilate_<layer>_<site> the reference writes for the
    25-site, two-layer pdopd model (kmos/io/__init__.py:2445-2560): the op
    list does not depend on the species argument, so it is compared for
    each species; the compiled model carries exactly the (site, species)
    pairs its processes' '^'/'$' actions use."""
    ref = _golden("pdopd_local_smart")
    pt = load_project(model_file(ref["model"]))
    cm = compile_model(pt, "local_smart")
    multi = {k: v for k, v in ref["routines"].items() if k.startswith(("create_", "annihilate_"))}
    assert len(multi) == 2 * len(pt.site_list())
    sid = pt.species_ids()
    null = pt.null_species()
    assert null == len(pt.species) == 4
    for rname, ops in multi.items():
        op, rest = rname.split("_", 1)
        lname, sname = [ls for ls in pt.site_list() if "%s_%s" % ls == rest][0]
        for species in sorted(sid):
            r = multilattice_routine(pt, op, lname, sname, species)
            assert _local_smart_ops(cm, r) == ops, (rname, species)
            assert (r.old_species, r.new_species) == ((null, sid[species]) if op == "create" else (sid[species], null))
    used = {(r.name.split(":")[0], r.name.split(":")[1]) for r in cm.routines if ":" in r.name}
    calls = {(c[0], c[2]) for acts in ref["run_proc"].values() for c in acts if len(c) == 3}
    assert used == calls


@pytest.mark.parametrize("name", LOCAL_SMART)
def test_local_smart_run_proc_matches_reference(name):
    ref = _golden(name)
    pt = load_project(model_file(ref["model"]))
    cm = compile_model(pt, "local_smart")
    for p, acts in enumerate(cm.proc_actions):
        mine = []
        for r, s in acts:
            rname = cm.routines[r].name
            if ":" in rname:            # create/annihilate: routine, site, species
                mine.append([rname.split(":")[0], list(s), rname.split(":")[1]])
            else:
                mine.append([rname, list(s)])
        assert mine == ref["run_proc"][cm.process_names[p]], cm.process_names[p]


def _nli_leaves(cm, g, sp):
    """Non-zero leaves of a compiled case tree, in tree order, as paths of
    (coordinate, species set)."""
    nodes = cm.groups[g].nli_nodes
    out = []

    def walk(i, path):
        nd = nodes[i]
        if nd[4] == -1:
            out.append((path, cm.process_names[nd[5] - 1]))
            return
        coord = list(nd[0:4])
        for k in range(nd[4]):
            mask, kind, tgt = nd[5 + 3 * k:8 + 3 * k]
            names = sorted(sp[b] for b in range(len(sp)) if (mask >> b) & 1)
            p2 = path + [coord + [names]]
            if kind == 1:
                out.append((p2, cm.process_names[tgt - 1]))
            else:
                walk(tgt, p2)
    walk(0, [])
    return out


@pytest.mark.parametrize("name", ["ruo2_lat_int", "pdopd_lat_int"])
def test_lat_int_groups_match_reference(name):
    ref = _golden(name)
    pt = load_project(model_file(ref["model"]))
    cm = compile_model(pt, "lat_int")
    gnames = [g.name for g in cm.groups]
    assert sorted(gnames) == sorted(ref["groups"])
    assert gnames == ref["touchup_groups"]
    for p, g in enumerate(cm.proc_group):
        assert ref["dispatch"][cm.process_names[p]] == gnames[g]
    # the null species of a multi-layer model is written by its name
    sp = cm.species_names + (["null_species"] if cm.null_species >= 0 else [])
    for gi, g in enumerate(cm.groups):
        rg = ref["groups"][g.name]
        mods = [[gnames[og], list(off)] for og, off in g.modified]
        assert mods == rg["dels"], g.name
        assert mods == rg["adds"], g.name
        ups = [[list(s), [sp[o] for o in olds], sp[new]] for s, olds, new in g.updates]
        assert ups == rg["updates"], g.name
        # nli: the non-default leaves in text order
        ref_leaves = [(path, res) for path, res in ref["nli"][g.name] if res != "0"]
        mine = _nli_leaves(cm, gi, cm.species_names)
        assert [(p, r) for p, r in mine] == [([c[:4] + [sorted(c[4])] for c in p], r)
                                             for p, r in ref_leaves], g.name


def test_most_common_helper():
    """kmos/io tests (tests/test_io_utils.py) pin _most_common: highest
    count, ties to the earliest first occurrence."""
    assert _most_common([1, 2, 2, 3, 3]) == 2
    assert _most_common(["b", "a", "a", "b"]) == "b"
    assert _most_common([5]) == 5


def test_iftree_order_small():
    """A hand-worked tree: most common coordinate first, answers sorted,
    unconditional items emitted first, remaining items after the select."""
    items = [
        ([("x", "b")], "p1"),
        ([], "p0"),
        ([("x", "a"), ("y", "a")], "p2"),
        ([("y", "b")], "p3"),
    ]
    order = [p for p, _ in iftree_order(items)]
    # x is asked twice (first at index 0), y twice (first at index 2): x wins
    assert order == ["p0", "p2", "p1", "p3"]


def _chain_project():
    """A small lat_int model whose groups have several members and
    bystanders (single site 'a', species A and empty; names chosen so that
    the reference's string sort differs from file order: 'H' < 'd' < 'z')."""
    from kmos_amd.model import CondAct, Coord, Layer, Process, Project, Site

    def c(sp, off=(0, 0, 0)):
        return CondAct(sp, Coord("a", off, "sq"))
    pt = Project(model_name="chain", model_dimension=2, species=["A", "empty"],
                 default_species="empty", layers=[Layer("sq", [Site("a")])],
                 default_layer="sq", substrate_layer="sq")
    pt.processes = [
        Process("z_ads", "1.", [c("empty")], [c("A")]),
        Process("des_00", "1.", [c("A"), c("empty", (0, 1, 0)), c("empty", (1, 0, 0))], [c("empty")]),
        Process("des_11", "1.", [c("A"), c("A", (0, 1, 0)), c("A", (1, 0, 0))], [c("empty")]),
        Process("Hop_0", "1.", [c("A"), c("empty", (1, 0, 0)), c("empty", (-1, 0, 0))],
                [c("empty"), c("A", (1, 0, 0))]),
        Process("Hop_1", "1.", [c("A"), c("empty", (1, 0, 0)), c("A", (-1, 0, 0))],
                [c("empty"), c("A", (1, 0, 0))]),
    ]
    pt.validate()
    return pt


def test_lat_int_multi_member_groups_hand_worked():
    """Known answer worked out by hand from the reference generator's rules,
    for groups with several members and bystanders (none of the reference's
    lat_int fixtures has one).

    Grouping (kmos/io/__init__.py:1497-1624): a condition whose species an
    action on the same coordinate changes is a true condition, every other
    condition a bystander; processes join the first group (in file order)
    with equal sorted true conditions and actions and equal sorted
    bystander coordinates.  So des_11 joins des_00 and Hop_1 joins Hop_0;
    touchup_cell adds the groups in insertion order z_ads, des_00, Hop_0
    (io:1788-1789).

    modified_procs (io:1825-1852): for every action of the group's first
    process and every (true condition or bystander) of every group's first
    process on the same site name, the pair (group, action offset minus
    condition offset), deduplicated and sorted by the text
    "<name> (np.int64(x), np.int64(y), np.int64(z))".  The conditions of
    z_ads, des_00 and Hop_0 sit at offsets {0}, {0, (0,1,0), (1,0,0)} and
    {0, (1,0,0), (-1,0,0)}.  An action at 0 gives the negated offsets; the
    Hop action at (1,0,0) adds (1,0,0) minus each.  In the text sort 'H' <
    'd' < 'z' and '-' < '0' < '1' < '2'.

    nli (io:2027-2047): per member, its true conditions and bystanders sorted
    by cmp_coords (offset x, then y, then z) form one path of the case
    tree; leaves appear in member order.
    """
    pt = _chain_project()
    cm = compile_model(pt, "lat_int")
    gnames = [g.name for g in cm.groups]
    assert gnames == ["z_ads", "des_00", "Hop_0"]
    assert [gnames[g] for g in cm.proc_group] == ["z_ads", "des_00", "des_00", "Hop_0", "Hop_0"]
    assert [[cm.process_names[p - 1] for p in g.procs] for g in cm.groups] == \
        [["z_ads"], ["des_00", "des_11"], ["Hop_0", "Hop_1"]]

    single_action = [
        ["Hop_0", [-1, 0, 0]], ["Hop_0", [0, 0, 0]], ["Hop_0", [1, 0, 0]],
        ["des_00", [-1, 0, 0]], ["des_00", [0, -1, 0]], ["des_00", [0, 0, 0]],
        ["z_ads", [0, 0, 0]],
    ]
    hop = [
        ["Hop_0", [-1, 0, 0]], ["Hop_0", [0, 0, 0]], ["Hop_0", [1, 0, 0]], ["Hop_0", [2, 0, 0]],
        ["des_00", [-1, 0, 0]], ["des_00", [0, -1, 0]], ["des_00", [0, 0, 0]],
        ["des_00", [1, -1, 0]], ["des_00", [1, 0, 0]],
        ["z_ads", [0, 0, 0]], ["z_ads", [1, 0, 0]],
    ]
    expect_mod = {"z_ads": single_action, "des_00": single_action, "Hop_0": hop}
    expect_upd = {
        "z_ads": [[[0, 0, 0, 1], ["empty"], "A"]],
        "des_00": [[[0, 0, 0, 1], ["A"], "empty"]],
        "Hop_0": [[[0, 0, 0, 1], ["A"], "empty"], [[1, 0, 0, 1], ["empty"], "A"]],
    }
    expect_nli = {
        "z_ads": [([[0, 0, 0, 1, ["empty"]]], "z_ads")],
        "des_00": [([[0, 0, 0, 1, ["A"]], [0, 1, 0, 1, ["empty"]], [1, 0, 0, 1, ["empty"]]], "des_00"),
                   ([[0, 0, 0, 1, ["A"]], [0, 1, 0, 1, ["A"]], [1, 0, 0, 1, ["A"]]], "des_11")],
        "Hop_0": [([[-1, 0, 0, 1, ["empty"]], [0, 0, 0, 1, ["A"]], [1, 0, 0, 1, ["empty"]]], "Hop_0"),
                  ([[-1, 0, 0, 1, ["A"]], [0, 0, 0, 1, ["A"]], [1, 0, 0, 1, ["empty"]]], "Hop_1")],
    }
    sp = cm.species_names
    for gi, g in enumerate(cm.groups):
        assert [[gnames[og], list(off)] for og, off in g.modified] == expect_mod[g.name], g.name
        assert [[list(s), [sp[o] for o in olds], sp[new]] for s, olds, new in g.updates] \
            == expect_upd[g.name], g.name
        assert _nli_leaves(cm, gi, sp) == expect_nli[g.name], g.name
