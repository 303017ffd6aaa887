"""Export backend: compile a model into the op programs the engine executes.

kmos's exporter writes, per model, Fortran routines whose *order* of
``add_proc``/``del_proc`` calls fixes the layout of the available-site lists
and therefore which site a uniform picks.  This module derives the same
ordered op programs from the model, as data instead of Fortran text:

local_smart (kmos/io/__init__.py:305-465, 2219-2443, 2568-2655)
    * per process: its actions as ``take``/``put`` routine calls relative to
      the executing coordinate (write_proclist_run_proc_nr_smart);
    * per routine ``put_/take_<species>_<layer>_<site>``: the lattice write,
      the ordered disable list (process order x condition order) and the
      enable list in the text order of the generated if-tree
      (write_proclist_put_take, _write_optimal_iftree);
    * per site type: the touchup routine (write_proclist_touchup);
    * multi-layer models: per (site, species) used by a '^'/'$' action, the
      create/annihilate routine (write_proclist_multilattice): the write
      from/to the null species, then every process with a condition on
      that site disabled (annihilate) or enabled through the if-tree over
      all its conditions (create).

lat_int (kmos/io/__init__.py:1497-2057)
    * lateral-interaction groups (_get_lat_int_groups);
    * per group: the sorted ``modified_procs`` list, the lattice updates
      (multi-layer: '$' writes the null species, '^' writes over it) and
      the nli case tree that maps a cell to a process of the group or 0
      (write_proclist_lat_int_run_proc, write_proclist_lat_int_nli_casetree).

otf (kmos/io/__init__.py:2657-3723)
    * per process the ordered del/write/update/add program of its run_proc
      routine, the adds of touchup_cell and per process its rate table
      description (kmos_amd/otf.py).

``CompiledModel.to_blob()`` flattens everything into one int32 array; its
layout is documented in ``include/kmcg.h`` and read by the HIP library.
"""
from __future__ import annotations

import dataclasses
import functools
from typing import Dict, List, Optional, Sequence, Tuple

import numpy as np

from .model import CondAct, Coord, Process, Project, cmp_coords

BACKEND_LOCAL_SMART = 0
BACKEND_LAT_INT = 1
BACKEND_OTF = 2
BACKENDS = {"local_smart": BACKEND_LOCAL_SMART, "lat_int": BACKEND_LAT_INT, "otf": BACKEND_OTF}

BLOB_MAGIC = 0x4B4D4347  # 'KMCG'
BLOB_VERSION = 1
NO_WRITE = -1000          # routine with no lattice write (touchup)
HEADER_WORDS = 32

# A relative site: (dx, dy, dz, n) with n the absolute 1-based site type.
RelSite = Tuple[int, int, int, int]


@dataclasses.dataclass
class EnableItem:
    """``add_proc(proc, site + offset)`` guarded by ``conds``.

    ``conds`` lists (relative site, species id) pairs that must all hold.
    """
    proc: int
    site: RelSite
    conds: List[Tuple[RelSite, int]]


@dataclasses.dataclass
class Routine:
    """One generated put/take (or touchup) routine, as ordered ops."""
    name: str
    site_type: int                # n of the routine's own site
    old_species: int              # species the write expects (NO_WRITE: none)
    new_species: int
    disables: List[Tuple[int, RelSite]]   # (proc, relative site), in order
    enables: List[EnableItem]             # in if-tree text order


@dataclasses.dataclass
class LatIntGroup:
    name: str                     # name of the group's first process
    procs: List[int]              # member process ids
    modified: List[Tuple[int, Tuple[int, int, int]]]  # (group idx, cell offset)
    updates: List[Tuple[RelSite, List[int], int]]     # (site, allowed old, new)
    nli_nodes: List[list]         # flattened case tree (see _flatten_nli)


@dataclasses.dataclass
class CompiledModel:
    backend: int
    project: Project
    nr_of_proc: int
    nr_of_species: int
    null_species: int
    spuck: int
    dimension: int
    site_default: List[int]
    process_names: List[str]
    species_names: List[str]
    site_names: List[str]
    # 1 per site type that initialize_state touches up (the sites of the
    # initial layer; every site of a one-layer model)
    site_init: List[int] = dataclasses.field(default_factory=list)
    # local_smart
    proc_exec_site: List[int] = dataclasses.field(default_factory=list)
    proc_actions: List[List[Tuple[int, RelSite]]] = dataclasses.field(default_factory=list)
    routines: List[Routine] = dataclasses.field(default_factory=list)
    touchup_routine: List[int] = dataclasses.field(default_factory=list)
    # lat_int
    proc_group: List[int] = dataclasses.field(default_factory=list)
    groups: List[LatIntGroup] = dataclasses.field(default_factory=list)
    # otf
    otf_programs: List[list] = dataclasses.field(default_factory=list)
    otf_touchup: list = dataclasses.field(default_factory=list)
    otf_rates: list = dataclasses.field(default_factory=list)
    param_names: List[str] = dataclasses.field(default_factory=list)

    @property
    def backend_name(self) -> str:
        return {v: k for k, v in BACKENDS.items()}[self.backend]

    def _offsets(self):
        """Every cell offset the engine adds to a (wrapped) cell in one go."""
        out = []
        if self.backend == BACKEND_LOCAL_SMART:
            for acts in self.proc_actions:
                out += [s[:3] for _, s in acts]
            for r in self.routines:
                out += [s[:3] for _, s in r.disables]
                for e in r.enables:
                    out.append(e.site[:3])
                    out += [c[:3] for c, _ in e.conds]
        elif self.backend == BACKEND_LAT_INT:
            for g in self.groups:
                out += [tuple(off) for _, off in g.modified]
                out += [s[:3] for s, _, _ in g.updates]
                out += [tuple(nd[0:3]) for nd in g.nli_nodes if nd[4] != -1]
        else:
            # an op's cell offset plus the offset of each site it reads
            for prog in self.otf_programs + [self.otf_touchup]:
                for op in prog:
                    out.append(op.off)
                    for c, _ in op.conds:
                        out.append(c[:3])
                    if op.proc >= 0:
                        for b, _ in self.otf_rates[op.proc].bystanders:
                            out.append(tuple(op.off[i] + b[i] for i in range(3)))
        return out

    def min_size(self) -> int:
        """Smallest lattice extent (cells per periodic dimension) the engine
        accepts: each offset is applied to an already wrapped cell, so one
        periodic wrap must suffice."""
        return max([max(abs(o[0]), abs(o[1]), abs(o[2])) for o in self._offsets()] or [0]) or 1

    def consistent_size(self) -> int:
        """Smallest extent at which no step touches one physical site through
        two different offsets.  Below it the reference's own bookkeeping can
        register a process twice or delete an absent one (the engine then
        reproduces or flags that); the models' steady state is only
        meaningful above it."""
        return 2 * max([max(abs(o[0]), abs(o[1]), abs(o[2])) for o in self._offsets()] or [0]) + 1

    # -------------------------------------------------------------------
    def to_blob(self) -> np.ndarray:
        return _to_blob(self)


# ---------------------------------------------------------------------------
# coordinate helpers
# ---------------------------------------------------------------------------

def _rel(a: Coord, b: Coord, pt: Project) -> RelSite:
    """``a - b`` in kmos's sense (types.py Coord.__sub__): the cell offset
    difference, keeping a's site (absolute site index of a)."""
    return (a.offset[0] - b.offset[0], a.offset[1] - b.offset[1],
            a.offset[2] - b.offset[2], pt.site_index(a.layer, a.name))


def _most_common(seq: Sequence) -> object:
    """Highest count, ties to the earliest first occurrence
    (kmos/io/__init__.py:98-115 _most_common)."""
    stats: Dict[object, List[int]] = {}
    for i, x in enumerate(seq):
        if x in stats:
            stats[x][0] += 1
        else:
            stats[x] = [1, i]
    return max(stats.items(), key=lambda kv: (kv[1][0], -kv[1][1]))[0]


def iftree_order(items: List[Tuple[List[Tuple[object, str]], object]]) -> List[Tuple[object, Dict[object, str]]]:
    """Text order of the ``add_proc`` calls in the if-tree that
    kmos/io/__init__.py:2568-2655 (_write_optimal_iftree) writes.

    ``items`` are ``(conditions, payload)`` with conditions as
    ``(coordinate key, species name)``.  Returns ``(payload, checks)`` in
    emission order; ``checks`` maps each coordinate tested on the path to
    the species the case branch requires.
    """
    out: List[Tuple[object, Dict[object, str]]] = []

    def rec(items, path):
        for conds, payload in items:
            if not conds:
                out.append((payload, dict(path)))
        items = [it for it in items if it[0]]
        if not items:
            return
        mcc = _most_common([c for conds, _ in items for c, _ in conds])
        answers = sorted(set(sp for conds, _ in items for c, sp in conds if c == mcc))
        for ans in answers:
            nested = [it for it in items
                      if mcc in [c for c, _ in it[0]]
                      and [sp for c, sp in it[0] if c == mcc][0] == ans]
            pruned = [([(c, sp) for c, sp in conds if c != mcc], payload)
                      for conds, payload in nested]
            nested_ids = set(id(it) for it in nested)
            items = [it for it in items if id(it) not in nested_ids]
            rec(pruned, path + [(mcc, ans)])
        if items:
            rec(items, path)

    rec(list(items), [])
    return out


# ---------------------------------------------------------------------------
# local_smart
# ---------------------------------------------------------------------------

def multilattice_routine(pt: Project, op: str, lname: str, sname: str, species: str) -> Routine:
    """``create_`` / ``annihilate_<layer>_<site>(site, species)`` of a
    multi-layer model (kmos/io/__init__.py:2445-2560) for one species: the
    write from (create) or to (annihilate) the null species, then, for
    every process and every condition it has on this site type (process
    order, condition order), annihilate's ``del_proc`` of the process at its
    executing coordinate, or create's ``add_proc`` in the text order of the
    if-tree over all the process's conditions, this site's own included."""
    sid = pt.species_ids()
    pid = pt.process_ids()
    if species not in sid:
        raise ValueError("%s of unknown species %r" % (op, species))
    enabled, disabled = [], []
    for proc in pt.processes:
        ex = proc.executing_coord()
        for cond in proc.conditions:
            if (cond.coord.name, cond.coord.layer) != (sname, lname):
                continue
            if op == "create":
                conds = [(_rel(o.coord, cond.coord, pt), o.species) for o in proc.conditions]
                enabled.append((conds, (pid[proc.name], _rel(ex, cond.coord, pt))))
            else:
                disabled.append((pid[proc.name], _rel(ex, cond.coord, pt)))
    enables = [EnableItem(proc=payload[0], site=payload[1], conds=[(c, sid[sp]) for c, sp in checks.items()])
               for payload, checks in iftree_order(enabled)]
    null = pt.null_species()
    old, new = (null, sid[species]) if op == "create" else (sid[species], null)
    return Routine(name="%s_%s_%s:%s" % (op, lname, sname, species), site_type=pt.site_index(lname, sname),
                   old_species=old, new_species=new, disables=disabled, enables=enables)


def _compile_local_smart(pt: Project, cm: CompiledModel) -> None:
    sid = pt.species_ids()
    pid = pt.process_ids()
    default = pt.default_species
    null = pt.null_species()

    # routines: put/take per non-default species x site (io:2219-2409)
    routine_index: Dict[Tuple[str, str, str, str], int] = {}
    for species in pt.species:
        if species == default:
            continue
        for layer in pt.layers:
            for site in layer.sites:
                for op in ("put", "take"):
                    enabled, disabled = [], []
                    for proc in pt.processes:
                        ex = proc.executing_coord()
                        for cond in proc.conditions:
                            if (cond.coord.name, cond.coord.layer) != (site.name, layer.name):
                                continue
                            if (op == "put" and species == cond.species) or \
                               (op == "take" and cond.species == default):
                                others = [c for c in proc.conditions if c.coord != cond.coord]
                                conds = [(_rel(o.coord, cond.coord, pt), o.species) for o in others]
                                enabled.append((conds, (pid[proc.name], _rel(ex, cond.coord, pt))))
                            elif (op == "put" and cond.species == default) or \
                                 (op == "take" and species == cond.species):
                                disabled.append((pid[proc.name], _rel(ex, cond.coord, pt)))
                    ordered = iftree_order(enabled)
                    enables = [EnableItem(proc=payload[0], site=payload[1],
                                          conds=[(c, sid[sp]) for c, sp in checks.items()])
                               for payload, checks in ordered]
                    if op == "put":
                        old, new = sid[default], sid[species]
                    else:
                        old, new = sid[species], sid[default]
                    routine_index[(op, species, layer.name, site.name)] = len(cm.routines)
                    cm.routines.append(Routine(
                        name="%s_%s_%s_%s" % (op, species, layer.name, site.name),
                        site_type=pt.site_index(layer.name, site.name),
                        old_species=old, new_species=new,
                        disables=disabled, enables=enables))

    # create_/annihilate_<layer>_<site> (io:2445-2560), multi-layer models
    # only: kmos writes one routine per site taking the species as an
    # argument; here one routine per (site, species) that an action uses
    if len(pt.layers) > 1:
        used = sorted(set(("create" if ca.species[0] == "^" else "annihilate", ca.species[1:],
                           ca.coord.layer, ca.coord.name)
                          for p in pt.processes for ca in p.actions if ca.species[:1] in ("^", "$")))
        for op, species, lname, sname in used:
            routine_index[(op, species, lname, sname)] = len(cm.routines)
            cm.routines.append(multilattice_routine(pt, op, lname, sname, species))

    # touchups (io:2411-2443)
    for layer in pt.layers:
        for site in layer.sites:
            n = pt.site_index(layer.name, site.name)
            dels = [(pid[proc.name], (0, 0, 0, n)) for proc in pt.processes]
            items = []
            for proc in pt.processes:
                ex = proc.executing_coord()
                if (ex.layer, ex.name) == (layer.name, site.name):
                    conds = [(_rel(c.coord, ex, pt), c.species) for c in proc.conditions]
                    items.append((conds, (pid[proc.name], (0, 0, 0, n))))
            ordered = iftree_order(items)
            enables = [EnableItem(proc=payload[0], site=payload[1],
                                  conds=[(c, sid[sp]) for c, sp in checks.items()])
                       for payload, checks in ordered]
            cm.touchup_routine.append(len(cm.routines))
            cm.routines.append(Routine(
                name="touchup_%s_%s" % (layer.name, site.name), site_type=n,
                old_species=NO_WRITE, new_species=NO_WRITE,
                disables=dels, enables=enables))

    # run_proc_nr per process (io:305-465)
    for proc in pt.processes:
        ex = proc.executing_coord()
        cm.proc_exec_site.append(pt.site_index(ex.layer, ex.name))
        acts = []
        for action in proc.actions:
            rel = _rel(action.coord, ex, pt)
            key = (action.coord.layer, action.coord.name)
            if action.species[:1] == "^":
                acts.append((routine_index[("create", action.species[1:]) + key], rel))
                continue
            if action.species[:1] == "$":
                acts.append((routine_index[("annihilate", action.species[1:]) + key], rel))
                continue
            prev = [c.species for c in proc.conditions if c.coord == action.coord]
            if not prev:
                raise ValueError("process %s: action on %s has no condition on the same site"
                                 % (proc.name, action.coord))
            prev = prev[0]
            if action.species == default and action.species != prev:
                acts.append((routine_index[("take", prev) + key], rel))
            elif prev != action.species:
                if prev != default:
                    acts.append((routine_index[("take", prev) + key], rel))
                acts.append((routine_index[("put", action.species) + key], rel))
        cm.proc_actions.append(acts)


# ---------------------------------------------------------------------------
# lat_int
# ---------------------------------------------------------------------------

def _np2_tuple_str(t: Tuple[int, ...]) -> str:
    """``str(tuple(numpy_int_array))`` as numpy >= 2 prints it; the sort key
    of ``modified_procs`` (io:1850-1852) is built from this text."""
    return "(" + ", ".join("np.int64(%d)" % v for v in t) + ")"


def lat_int_groups(pt: Project) -> List[Tuple[str, List[dict]]]:
    """Split processes into true conditions/actions and bystanders and group
    them (kmos/io/__init__.py:1497-1624 _get_lat_int_groups)."""
    procs = []
    for p in pt.processes:
        true_c, true_a, byst = [], [], []
        for c in p.conditions:
            corr = [a for a in p.actions if c.coord == a.coord]
            if corr:
                if c.species != corr[0].species:
                    true_c.append(c)
                    true_a.append(corr[0])
                else:
                    byst.append(c)
            else:
                byst.append(c)
        # actions without a changing condition: multi-lattice create/annihilate
        # (io:1536-1551); any other is ignored, as kmos does after its warning
        for a in p.actions:
            if a not in true_a and a.species[:1] in ("^", "$"):
                true_a.append(a)
        procs.append(dict(name=p.name, conditions=true_c, actions=true_a, bystanders=byst))

    key = functools.cmp_to_key(lambda a, b: cmp_coords(a.coord, b.coord))
    groups: Dict[str, List[dict]] = {}
    for p in procs:
        for gname, members in groups.items():
            p0 = members[0]
            same = (sorted(p0["conditions"], key=key) == sorted(p["conditions"], key=key)
                    and sorted(p0["actions"], key=key) == sorted(p["actions"], key=key)
                    and [x.coord for x in sorted(p0["bystanders"], key=key)]
                    == [x.coord for x in sorted(p["bystanders"], key=key)])
            if same:
                members.append(p)
                break
        else:
            groups[p["name"]] = [p]
    return list(groups.items())


def _compile_lat_int(pt: Project, cm: CompiledModel) -> None:
    sid = pt.species_ids()
    pid = pt.process_ids()
    groups = lat_int_groups(pt)
    gindex = {name: i for i, (name, _) in enumerate(groups)}
    cm.proc_group = [0] * len(pt.processes)
    key = functools.cmp_to_key(lambda a, b: cmp_coords(a.coord, b.coord))
    for gi, (gname, members) in enumerate(groups):
        for m in members:
            cm.proc_group[pid[m["name"]] - 1] = gi
    for gi, (gname, members) in enumerate(groups):
        p0 = members[0]
        # modified_procs (io:1825-1852)
        modified = set()
        for action in p0["actions"]:
            for oname, others in groups:
                o0 = others[0]
                for cond in o0["conditions"] + o0["bystanders"]:
                    if action.coord.eq_mod_offset(cond.coord):
                        off = tuple(action.coord.offset[i] - cond.coord.offset[i] for i in range(3))
                        modified.add((oname, off))
        modified = sorted(modified, key=lambda x: "%s %s" % (x[0], _np2_tuple_str(x[1])))
        # lattice updates (io:1858-1965): changing conditions in order, then
        # the multi-lattice actions without a condition
        null = pt.null_species()
        updates = []
        matched = []

        def site_of(a):
            return (a.coord.offset[0], a.coord.offset[1], a.coord.offset[2],
                    pt.site_index(a.coord.layer, a.coord.name))

        for cond in p0["conditions"]:
            action = [a for a in p0["actions"] if cond.coord == a.coord][0]
            matched.append(action)
            olds = [sid[s.strip()] for s in cond.species.split(" or ")]
            if action.species[:1] == "$":
                new = null
            elif action.species[:1] == "^":
                # kmos writes a replace_species with the '^' species text here,
                # which does not compile
                raise NotImplementedError("%s: create action on a site with a condition" % gname)
            else:
                new = sid[action.species]
            updates.append((site_of(action), olds, new))
        for action in p0["actions"]:
            if action in matched:
                continue
            sp = action.species[1:]
            if action.species[:1] == "$":
                updates.append((site_of(action), [sid[s.strip()] for s in sp.split(" or ")], null))
            else:
                updates.append((site_of(action), [null], sid[sp]))
        # nli case tree (io:2027-2047): nested dicts in insertion order
        tree: Dict = {}
        for m in members:
            conds = [c for c in sorted(m["conditions"] + m["bystanders"], key=key) if not c.implicit]
            node = tree
            for c in conds:
                species_node = node.setdefault(("coord", c.coord), {})
                node = species_node.setdefault(("species", c.species), {})
                species_node.setdefault(("default",), 0)
            node[("leaf",)] = pid[m["name"]]
        nodes: List[list] = []
        _flatten_nli(tree, nodes, pt, sid)
        cm.groups.append(LatIntGroup(
            name=gname, procs=[pid[m["name"]] for m in members],
            modified=[(gindex[n], off) for n, off in modified],
            updates=updates, nli_nodes=nodes))


def _flatten_nli(tree: Dict, nodes: List[list], pt: Project, sid: Dict[str, int]) -> int:
    """Flatten the case tree into nodes ``[dx, dy, dz, n, n_cases,
    (species_mask, kind, target)*]``; kind 0 = child node, 1 = leaf value.

    The generated Fortran returns from every branch of the first
    ``select case`` it meets (io:42-66 _casetree_dict), so at each level only
    the first key in insertion order matters; a leaf key that comes first
    returns its process directly.
    """
    first = next(iter(tree))
    idx = len(nodes)
    if first[0] == "leaf":
        # degenerate: process with no remaining condition -> constant
        nodes.append([0, 0, 0, 1, -1, tree[first]])
        return idx
    coord = first[1]
    cases = tree[first]
    nodes.append(None)
    entries = []
    for k, v in cases.items():
        if k[0] == "default":
            continue
        mask = 0
        for s in k[1].split(" or "):
            mask |= 1 << sid[s.strip()]
        ck = next(iter(v))
        if ck[0] == "leaf":
            entries.append((mask, 1, v[ck]))
        else:
            child = _flatten_nli(v, nodes, pt, sid)
            entries.append((mask, 0, child))
    node = [coord.offset[0], coord.offset[1], coord.offset[2],
            pt.site_index(coord.layer, coord.name), len(entries)]
    for e in entries:
        node.extend(e)
    nodes[idx] = node
    return idx


# ---------------------------------------------------------------------------
# entry point
# ---------------------------------------------------------------------------

def compile_model(pt: Project, backend: str = "local_smart") -> CompiledModel:
    """Compile ``pt`` for ``backend`` ('local_smart' or 'lat_int')."""
    if backend not in BACKENDS:
        raise ValueError("unknown backend %r" % backend)
    pt.validate()
    species_sorted = sorted(pt.species)
    cm = CompiledModel(
        backend=BACKENDS[backend], project=pt,
        nr_of_proc=len(pt.processes), nr_of_species=len(pt.species),
        null_species=pt.null_species(), spuck=len(pt.site_list()),
        dimension=pt.model_dimension, site_default=pt.site_default_species(),
        process_names=[p.name for p in pt.processes],
        species_names=species_sorted,
        site_names=["%s_%s" % ls for ls in pt.site_list()],
        site_init=[1 if (len(pt.layers) == 1 or lname == pt.default_layer) else 0
                   for lname, _ in pt.site_list()])
    if len(pt.species) > 8:
        raise NotImplementedError("at most 8 species are supported")
    if backend == "local_smart":
        _compile_local_smart(pt, cm)
    elif backend == "lat_int":
        _compile_lat_int(pt, cm)
    else:
        _compile_otf(pt, cm)
    return cm


def _compile_otf(pt: Project, cm: CompiledModel) -> None:
    from . import otf
    from .model import UNITS
    otf.check_otf_model(pt)
    cm.param_names = sorted(pt.parameters)
    cm.otf_rates = [otf.compile_rate(pt, proc, UNITS) for proc in pt.processes]
    cm.otf_programs = [otf.run_proc_program(pt, proc) for proc in pt.processes]
    cm.otf_touchup = otf.touchup_program(pt)


# ---------------------------------------------------------------------------
# blob
# ---------------------------------------------------------------------------

def _to_blob(cm: CompiledModel) -> np.ndarray:
    """Serialise to int32 words.  Layout (word offsets are absolute):

    header[32]:
      0 magic, 1 version, 2 backend, 3 nr_of_proc, 4 nr_of_species,
      5 null_species, 6 spuck, 7 dimension,
      8 off_site_default, 9 off_proc_table, 10 n_routines, 11 off_routine_table,
      12 off_touchup (spuck words: routine id per site type),
      13 n_groups, 14 off_group_table, 15 max_offset, 16 total words,
      17 off_site_init (spuck words: 1 = initialize_state touches the site
         type up, proclist_generic_subroutines.mpy:279-300 runs the touchups
         of the initial layer only)
    local_smart:
      proc table [P][3]: exec_site, n_actions, off_actions -> [routine, dx, dy, dz, n]*
      routine table [R][7]: site_type, old, new, n_dis, off_dis, n_en, off_en
        dis: [proc, dx, dy, dz, n]*
        en:  [proc, dx, dy, dz, n, n_cond, off_cond]*, cond: [dx, dy, dz, n, species]*
    lat_int:
      proc table [P][1]: group
      group table [G][8]: n_mod, off_mod, n_upd, off_upd, off_nli, nli_words, n_procs, off_procs
        mod: [group, dx, dy, dz]*
        upd: [dx, dy, dz, n, old_mask, new]*
        nli nodes (offsets of child nodes are relative to off_nli)
    """
    words: List[int] = [0] * HEADER_WORDS
    hdr = words

    def here():
        return len(words)

    hdr[0] = BLOB_MAGIC
    hdr[1] = BLOB_VERSION
    hdr[2] = cm.backend
    hdr[3] = cm.nr_of_proc
    hdr[4] = cm.nr_of_species
    hdr[5] = cm.null_species
    hdr[6] = cm.spuck
    hdr[7] = cm.dimension
    hdr[8] = here()
    words.extend(cm.site_default)
    hdr[17] = here()
    words.extend(cm.site_init)
    if cm.backend == BACKEND_LOCAL_SMART:
        hdr[9] = here()
        ptab = here()
        words.extend([0] * (3 * cm.nr_of_proc))
        for p, acts in enumerate(cm.proc_actions):
            words[ptab + 3 * p + 0] = cm.proc_exec_site[p]
            words[ptab + 3 * p + 1] = len(acts)
            words[ptab + 3 * p + 2] = here()
            for r, s in acts:
                words.extend([r, s[0], s[1], s[2], s[3]])
        hdr[10] = len(cm.routines)
        hdr[11] = here()
        rtab = here()
        words.extend([0] * (7 * len(cm.routines)))
        for i, r in enumerate(cm.routines):
            base = rtab + 7 * i
            words[base + 0] = r.site_type
            words[base + 1] = r.old_species
            words[base + 2] = r.new_species
            words[base + 3] = len(r.disables)
            words[base + 4] = here()
            for proc, s in r.disables:
                words.extend([proc, s[0], s[1], s[2], s[3]])
            words[base + 5] = len(r.enables)
            en_off = here()
            words[base + 6] = en_off
            words.extend([0] * (7 * len(r.enables)))
            for j, e in enumerate(r.enables):
                eb = en_off + 7 * j
                words[eb:eb + 5] = [e.proc, e.site[0], e.site[1], e.site[2], e.site[3]]
                words[eb + 5] = len(e.conds)
                words[eb + 6] = here()
                for c, sp in e.conds:
                    words.extend([c[0], c[1], c[2], c[3], sp])
        hdr[12] = here()
        words.extend(cm.touchup_routine)
    else:
        hdr[9] = here()
        words.extend(cm.proc_group)
        hdr[13] = len(cm.groups)
        hdr[14] = here()
        gtab = here()
        words.extend([0] * (8 * len(cm.groups)))
        for gi, g in enumerate(cm.groups):
            base = gtab + 8 * gi
            words[base + 0] = len(g.modified)
            words[base + 1] = here()
            for og, off in g.modified:
                words.extend([og, off[0], off[1], off[2]])
            words[base + 2] = len(g.updates)
            words[base + 3] = here()
            for s, olds, new in g.updates:
                mask = 0
                for o in olds:
                    mask |= 1 << o
                words.extend([s[0], s[1], s[2], s[3], mask, new])
            words[base + 4] = here()
            # node offsets relative to the group's nli section
            node_pos = []
            pos = 0
            for node in g.nli_nodes:
                node_pos.append(pos)
                pos += len(node)
            for node in g.nli_nodes:
                if node[4] == -1:
                    words.extend(node)
                    continue
                out = list(node[:5])
                for k in range(node[4]):
                    mask, kind, target = node[5 + 3 * k: 8 + 3 * k]
                    out.extend([mask, kind, node_pos[target] if kind == 0 else target])
                words.extend(out)
            words[base + 5] = pos
            words[base + 6] = len(g.procs)
            words[base + 7] = here()
            words.extend(g.procs)
    if cm.backend == BACKEND_OTF:
        _otf_blob(cm, words, hdr)
    hdr[15] = cm.min_size()
    hdr[16] = len(words)
    return np.asarray(words, dtype=np.int32)


def _otf_ops_words(ops) -> List[int]:
    from . import otf
    out: List[int] = []
    for op in ops:
        if op.kind == otf.OP_WRITE:
            out += [op.kind, op.off[0], op.off[1], op.off[2], op.site, op.old, op.new]
        elif op.kind == otf.OP_ADD:
            out += [op.kind, op.proc, op.off[0], op.off[1], op.off[2], len(op.conds)]
            for c, sp in op.conds:
                out += [c[0], c[1], c[2], c[3], sp]
        else:
            out += [op.kind, op.proc, op.off[0], op.off[1], op.off[2]]
    return out


def _otf_blob(cm: CompiledModel, words: List[int], hdr: List[int]) -> None:
    """otf sections (header words 9-13, 18):
      9  proc table [P][2]: n_ops, off_ops -> run_proc program
      10 n_params (userpar entries, names in sorted order)
      11 rate table [P][9]: table_size, table_base, n_byst, off_byst, n_code,
         off_code, n_aux, n_used, off_maxc
           byst: [dx, dy, dz, n, inc[S]]*   code: rate program words
           maxc: [largest count]* per used nr var
      12 off_touchup program, 18 its n_ops
      13 table entries per replica
    program ops: DEL [0, q, dx, dy, dz], WRITE [1, dx, dy, dz, n, old, new],
      UPD [2, q, dx, dy, dz], ADD [3, q, dx, dy, dz, n_cond, (dx, dy, dz, n, species)*]
    """
    P = cm.nr_of_proc
    hdr[9] = len(words)
    ptab = len(words)
    words.extend([0] * (2 * P))
    for p, prog in enumerate(cm.otf_programs):
        words[ptab + 2 * p] = len(prog)
        words[ptab + 2 * p + 1] = len(words)
        words.extend(_otf_ops_words(prog))
    hdr[10] = len(cm.param_names)
    hdr[11] = len(words)
    rtab = len(words)
    words.extend([0] * (9 * P))
    base = 0
    for p, r in enumerate(cm.otf_rates):
        rec = rtab + 9 * p
        words[rec + 0] = r.table_size
        words[rec + 1] = base
        base += r.table_size
        words[rec + 2] = len(r.bystanders)
        words[rec + 3] = len(words)
        for site, inc in r.bystanders:
            words.extend(list(site) + list(inc))
        words[rec + 4] = len(r.code)
        words[rec + 5] = len(words)
        words.extend(r.code)
        words[rec + 6] = r.n_aux
        words[rec + 7] = len(r.maxc)
        words[rec + 8] = len(words)
        words.extend(r.maxc)
    hdr[12] = len(words)
    hdr[18] = len(cm.otf_touchup)
    words.extend(_otf_ops_words(cm.otf_touchup))
    hdr[13] = base
