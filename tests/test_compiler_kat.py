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

from kmos_amd.compiler import compile_model, iftree_order, _most_common
from kmos_amd.model import load_project

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


@pytest.mark.parametrize("name", ["ruo2_local_smart", "cli_local_smart"])
def test_local_smart_routines_match_reference(name):
    ref = _golden(name)
    pt = load_project(os.path.join(GOLDEN, ref["model"]))
    cm = compile_model(pt, "local_smart")
    mine = {r.name: _local_smart_ops(cm, r) for r in cm.routines}
    # the reference only writes put/take routines for non-default species
    assert set(ref["routines"]) == set(mine)
    for rname, ops in ref["routines"].items():
        assert mine[rname] == ops, rname


@pytest.mark.parametrize("name", ["ruo2_local_smart", "cli_local_smart"])
def test_local_smart_run_proc_matches_reference(name):
    ref = _golden(name)
    pt = load_project(os.path.join(GOLDEN, ref["model"]))
    cm = compile_model(pt, "local_smart")
    for p, acts in enumerate(cm.proc_actions):
        mine = [[cm.routines[r].name, list(s)] for r, s in acts]
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


def test_lat_int_groups_match_reference():
    ref = _golden("ruo2_lat_int")
    pt = load_project(os.path.join(GOLDEN, ref["model"]))
    cm = compile_model(pt, "lat_int")
    gnames = [g.name for g in cm.groups]
    assert sorted(gnames) == sorted(ref["groups"])
    assert gnames == ref["touchup_groups"]
    for p, g in enumerate(cm.proc_group):
        assert ref["dispatch"][cm.process_names[p]] == gnames[g]
    sp = cm.species_names
    for gi, g in enumerate(cm.groups):
        rg = ref["groups"][g.name]
        mods = [[gnames[og], list(off)] for og, off in g.modified]
        assert mods == rg["dels"], g.name
        assert mods == rg["adds"], g.name
        ups = [[list(s), [sp[o] for o in olds], sp[new]] for s, olds, new in g.updates]
        assert ups == rg["updates"], g.name
        # nli: the non-default leaves in text order
        ref_leaves = [(path, res) for path, res in ref["nli"][g.name] if res != "0"]
        mine = _nli_leaves(cm, gi, sp)
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
