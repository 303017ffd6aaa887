"""Export backend for otf models: per-site rates from the local environment.

kmos's otf writer (kmos/io/__init__.py:2657-3723) emits, per process, a
``run_proc_<proc>(cell)`` routine and a ``gr_<proc>(cell)`` rate function.
This module derives the same information as data:

* per process the ordered bookkeeping of its ``run_proc`` routine
  (write_proclist_run_proc_name_otf, io:3328-3596): the ``del_proc`` calls
  of every process a changed site inhibits, the lattice writes, the
  ``update_rates_matrix`` calls of every process that has a bystander on a
  changed site, and the ``add_proc`` calls in the text order of the if-tree
  (_write_optimal_iftree_otf, io:3598-3723);
* the adds of ``touchup_cell`` (write_proclist_touchup_otf, io:3240-3284);
* per process the bystander counts that ``gr_<proc>`` makes
  (write_proclist_pars_otf, io:2800-2900) and its rate expression compiled
  to a small stack program (the ``rate_<proc>`` function, io:3000-3122).

Processes sit on site 1 of their cell (the generated code passes
``cell + (/dx, dy, dz, 1/)``), so every list operation carries a cell
offset only.

Rates.  ``gr_<proc>`` counts, for every bystander, the species found on its
site into ``nr_<species>_<flag>`` and evaluates the expression.  The value
depends on the counts only, so the engine keeps one table per process and
replica with the expression's value for every combination of the counts it
uses; a site's rate is ``table[sum of per-bystander strides]``.  The table
is evaluated on the host by the library with Fortran's semantics: integer,
default-real (single) and double-precision operands keep their kinds,
``real ** integer`` is binary powering with 1/r for negative exponents (the
form the reference's compiler emits), the intrinsics go to the platform's
libm, and real literals are single-precision constants.
"""
from __future__ import annotations

import ast
import dataclasses
import fractions
import struct
from typing import Dict, List, Optional, Sequence, Tuple

import numpy as np

from .model import Coord, Process, Project

# ops of an otf program, as int32 words
OP_DEL, OP_WRITE, OP_UPD, OP_ADD = 0, 1, 2, 3

# rate program opcodes (the host evaluator in kmcg_otf.hip reads them)
R_END = 0
R_PUSH_I, R_PUSH_F4, R_PUSH_RATE, R_PUSH_PAR, R_PUSH_NR, R_LOAD, R_STORE = 1, 2, 3, 4, 5, 6, 7
R_I2F4, R_I2F8, R_F4F8 = 8, 9, 10
R_NEG = 11            # + kind (0 int, 1 single, 2 double)
R_ADD, R_SUB, R_MUL, R_DIV = 20, 23, 26, 29   # + kind
R_POW_II, R_POW_F4I, R_POW_F8I, R_POW_F4F4, R_POW_F8F8 = 32, 33, 34, 35, 36
R_EXP, R_LOG, R_SQRT = 40, 42, 44            # + kind - 1 (single 0, double 1)
R_PUSH_F8 = 46        # two words: the double's bits, low then high

INT, F4, F8 = 0, 1, 2

MAX_TABLE = 1 << 16        # entries per process


@dataclasses.dataclass
class OtfOp:
    kind: int
    proc: int = -1                          # 0-based target process
    off: Tuple[int, int, int] = (0, 0, 0)   # cell offset
    site: int = 0                           # write: 1-based site type
    old: int = -1                           # write: species ids
    new: int = -1
    conds: List[Tuple[Tuple[int, int, int, int], int]] = dataclasses.field(default_factory=list)


@dataclasses.dataclass
class OtfRate:
    """What ``gr_<proc>`` computes, in table form."""
    var_names: List[str]          # the nr_vars of byst_<proc>, in order
    used: List[int]               # indices into var_names the expression reads
    maxc: List[int]               # per used var: the largest count it can reach
    strides: List[int]            # per used var: its stride in the table index
    # per bystander: (dx, dy, dz, n) and, per species id, the index increment
    bystanders: List[Tuple[Tuple[int, int, int, int], List[int]]]
    code: List[int]               # rate program
    n_aux: int
    table_size: int
    expr: str


# ---------------------------------------------------------------------------
# Fortran literal kinds
# ---------------------------------------------------------------------------

def f4_literal(text: str) -> float:
    """The single-precision value of a default-real Fortran literal,
    rounded once from its decimal text (ties to even)."""
    q = fractions.Fraction(text)
    if q == 0:
        return 0.0
    d = float(q)
    f = struct.unpack("<f", struct.pack("<f", d))[0]
    # step to the float32 neighbour nearest to q (float(q) may sit on the
    # other side of a float32 midpoint than q itself)
    best = f
    for cand in (np.nextafter(np.float32(f), np.float32(-np.inf)),
                 np.nextafter(np.float32(f), np.float32(np.inf))):
        c = float(cand)
        if abs(fractions.Fraction(c) - q) < abs(fractions.Fraction(best) - q) or \
           (abs(fractions.Fraction(c) - q) == abs(fractions.Fraction(best) - q) and
            (struct.unpack("<I", struct.pack("<f", c))[0] & 1) == 0):
            best = c
    return best


def _f4_bits(v: float) -> int:
    return struct.unpack("<i", struct.pack("<f", v))[0]


# ---------------------------------------------------------------------------
# rate expressions
# ---------------------------------------------------------------------------

_FUNCS = {"exp": R_EXP, "log": R_LOG, "sqrt": R_SQRT}


class _RateCompiler:
    """Python-syntax otf_rate text -> typed stack program."""

    def __init__(self, proc: str, params: Sequence[str], nr_vars: Sequence[str], units: Dict[str, float]):
        self.proc = proc
        self.params = sorted(params)             # userpar index order (io:2709-2714)
        self.nr_vars = list(nr_vars)
        self.units = units
        self.aux: List[str] = []
        self.code: List[int] = []
        self.used_nr: List[int] = []

    def fail(self, what):
        raise NotImplementedError("otf_rate of %s: %s" % (self.proc, what))

    def emit_convert(self, have, want):
        if have == want:
            return
        if have == INT:
            self.code.append(R_I2F4 if want == F4 else R_I2F8)
        elif have == F4 and want == F8:
            self.code.append(R_F4F8)
        else:
            self.fail("narrowing conversion")

    def expr(self, node) -> int:
        if isinstance(node, ast.Constant):
            if isinstance(node.value, bool) or not isinstance(node.value, (int, float)):
                self.fail("constant %r" % (node.value,))
            if isinstance(node.value, int):
                self.code += [R_PUSH_I, int(node.value)]
                return INT
            text = ast.get_source_segment(self.src, node) or repr(node.value)
            self.code += [R_PUSH_F4, _f4_bits(f4_literal(text))]
            return F4
        if isinstance(node, ast.Name):
            name = node.id
            if name == "base_rate":
                self.code.append(R_PUSH_RATE)
                return F8
            if name in self.aux:
                self.code += [R_LOAD, self.aux.index(name)]
                return F8
            if name.startswith("nr_"):
                if name not in self.nr_vars:
                    self.fail("%s is not counted by any bystander" % name)
                i = self.nr_vars.index(name)
                if i not in self.used_nr:
                    self.used_nr.append(i)
                self.code += [R_PUSH_NR, i]
                return INT
            if name in self.params:
                self.code += [R_PUSH_PAR, self.params.index(name)]
                return F8
            if name in self.units:
                # written as `real(kind=rdouble), parameter :: <unit> = %.10e`
                # (io:2728-2732): a default-real literal
                self.code += [R_PUSH_F4, _f4_bits(f4_literal("%.10e" % self.units[name]))]
                return F4
            self.fail("unknown name %r" % name)
        if isinstance(node, ast.UnaryOp):
            k = self.expr(node.operand)
            if isinstance(node.op, ast.USub):
                self.code.append(R_NEG + k)
            elif not isinstance(node.op, ast.UAdd):
                self.fail("operator %s" % type(node.op).__name__)
            return k
        if isinstance(node, ast.BinOp):
            if isinstance(node.op, ast.Pow):
                return self.power(node)
            ops = {ast.Add: R_ADD, ast.Sub: R_SUB, ast.Mult: R_MUL, ast.Div: R_DIV}
            if type(node.op) not in ops:
                self.fail("operator %s" % type(node.op).__name__)
            ka = self.expr(node.left)
            mid = len(self.code)
            kb = self.expr(node.right)
            k = max(ka, kb)
            if kb != k:
                self.emit_convert(kb, k)
            if ka != k:
                # convert the left operand where it was pushed
                right = self.code[mid:]
                del self.code[mid:]
                self.emit_convert(ka, k)
                self.code += right
            self.code.append(ops[type(node.op)] + k)
            return k
        if isinstance(node, ast.Call):
            if not isinstance(node.func, ast.Name) or node.func.id not in _FUNCS or len(node.args) != 1:
                self.fail("call %s" % ast.dump(node.func))
            k = self.expr(node.args[0])
            if k == INT:
                self.fail("%s of an integer" % node.func.id)
            self.code.append(_FUNCS[node.func.id] + (k - 1))
            return k
        self.fail("syntax %s" % type(node).__name__)

    def power(self, node) -> int:
        ka = self.expr(node.left)
        mid = len(self.code)
        kb = self.expr(node.right)
        if kb == INT:
            self.code.append({INT: R_POW_II, F4: R_POW_F4I, F8: R_POW_F8I}[ka])
            return ka
        k = max(ka, kb)
        if kb != k:
            self.emit_convert(kb, k)
        if ka != k:
            right = self.code[mid:]
            del self.code[mid:]
            self.emit_convert(ka, k)
            self.code += right
        self.code.append(R_POW_F4F4 if k == F4 else R_POW_F8F8)
        return k

    def compile(self, text: Optional[str]) -> List[int]:
        """The program of ``rate_<proc>`` (io:3000-3054): lines split on
        newlines or a literal backslash-n; a single line without '=' is the
        rate itself; assignments to other names are double-precision
        auxiliaries; the value of ``otf_rate`` is the rate."""
        if not text:
            self.code += [R_PUSH_RATE, R_END]
            return self.code
        import re
        lines = [ln for ln in re.split("\n|\\\\n", text) if ln.strip()]
        if len(lines) == 1 and "=" not in lines[0]:
            lines = ["otf_rate = " + lines[0]]
        final = None
        for ln in lines:
            self.src = ln.strip()
            try:
                tree = ast.parse(self.src)
            except SyntaxError as e:
                self.fail("cannot parse %r (%s)" % (self.src, e))
            if len(tree.body) != 1 or not isinstance(tree.body[0], ast.Assign) or \
                    len(tree.body[0].targets) != 1 or not isinstance(tree.body[0].targets[0], ast.Name):
                self.fail("each line must assign one name: %r" % self.src)
            target = tree.body[0].targets[0].id
            k = self.expr(tree.body[0].value)
            self.emit_convert(k, F8)
            if target == "otf_rate":
                final = len(self.aux)
                if "otf_rate" not in self.aux:
                    self.aux.append("otf_rate")
                self.code += [R_STORE, self.aux.index("otf_rate")]
            else:
                if target.startswith("nr_"):
                    self.fail("assignment to %s" % target)
                if target not in self.aux:
                    self.aux.append(target)
                self.code += [R_STORE, self.aux.index(target)]
        if final is None:
            self.fail("no value assigned to otf_rate")
        self.code += [R_LOAD, self.aux.index("otf_rate"), R_END]
        return self.code


def nr_vars_of(proc: Process, used_names: Sequence[str]) -> List[str]:
    """The nr_vars list of gr_<proc> (io:2802-2827): the names the
    expression uses plus every nr_<species>_<flag> of a bystander flag,
    sorted by (flag, species)."""
    specs: Dict[str, List[str]] = {}
    for b in proc.bystanders:
        for flg in b.flag.split():
            specs.setdefault(flg, []).extend(b.allowed_species)
    names = list(dict.fromkeys(used_names))
    for flg in sorted(specs):
        for sp in sorted(specs[flg]):
            nv = "nr_%s_%s" % (sp, flg)
            if nv not in names:
                names.append(nv)
    return sorted(names, key=lambda x: (x.split("_")[2], x.split("_")[1]))


def _expr_names(text: Optional[str]) -> List[str]:
    import io
    import tokenize
    if not text:
        return []
    out = []
    for tok in tokenize.generate_tokens(io.StringIO(text.replace("\\n", "\n")).readline):
        if tok.type == tokenize.NAME and tok.string.startswith("nr_"):
            out.append(tok.string)
    return out


def compile_rate(pt: Project, proc: Process, units: Dict[str, float]) -> OtfRate:
    sid = pt.species_ids()
    names = nr_vars_of(proc, _expr_names(proc.otf_rate))
    rc = _RateCompiler(proc.name, list(pt.parameters), names, units)
    code = rc.compile(proc.otf_rate)
    used = sorted(rc.used_nr)
    # the largest count of each used var: bystanders that can increment it
    maxc = []
    for v in used:
        sp, flg = _split_nr(names[v])
        maxc.append(sum(1 for b in proc.bystanders if flg in b.flag.split() and sp in b.allowed_species))
    strides, size = [], 1
    for m in maxc:
        strides.append(size)
        size *= m + 1
    if size > MAX_TABLE:
        raise NotImplementedError("otf_rate of %s: %d count combinations (at most %d)"
                                  % (proc.name, size, MAX_TABLE))
    bys = []
    for b in proc.bystanders:
        inc = [0] * len(pt.species)
        for spname in b.allowed_species:
            for flg in b.flag.split():
                nv = "nr_%s_%s" % (spname, flg)
                v = names.index(nv)
                if v in used:
                    inc[sid[spname]] += strides[used.index(v)]
        n = pt.site_index(b.coord.layer, b.coord.name)
        bys.append(((b.coord.offset[0], b.coord.offset[1], b.coord.offset[2], n), inc))
    # the program reads nr vars by their position among the used ones
    code = _renumber_nr(code, used)
    return OtfRate(var_names=names, used=used, maxc=maxc, strides=strides, bystanders=bys,
                   code=code, n_aux=len(rc.aux), table_size=size, expr=proc.otf_rate or "base_rate")


def _split_nr(name: str) -> Tuple[str, str]:
    """nr_<species>_<flag> -> (species, flag), split on '_' as the
    reference's sort key does (io:2827)."""
    parts = name.split("_", 2)
    return parts[1], parts[2]


def _renumber_nr(code: List[int], used: List[int]) -> List[int]:
    out = list(code)
    i = 0
    while i < len(out):
        op = out[i]
        if op in (R_PUSH_I, R_PUSH_F4, R_PUSH_PAR, R_LOAD, R_STORE):
            i += 2
        elif op == R_PUSH_NR:
            out[i + 1] = used.index(out[i + 1])
            i += 2
        elif op == R_PUSH_F8:
            i += 3
        else:
            i += 1
    return out


def eval_rate_table(rate: OtfRate, base_rate: float, userpar: Sequence[float]) -> np.ndarray:
    """The table of one process, evaluated in Python with the same kinds
    and operation order as the host evaluator (used by tests)."""
    import math
    out = np.zeros(rate.table_size, np.float64)
    f32 = np.float32
    for t in range(rate.table_size):
        counts, rem = [], t
        for m in rate.maxc:
            counts.append(rem % (m + 1))
            rem //= m + 1
        st: list = []
        aux = [0.0] * rate.n_aux
        c = rate.code
        i = 0
        while True:
            op = c[i]
            if op == R_END:
                break
            if op == R_PUSH_I:
                st.append(int(c[i + 1])); i += 2; continue
            if op == R_PUSH_F4:
                st.append(f32(struct.unpack("<f", struct.pack("<i", c[i + 1]))[0])); i += 2; continue
            if op == R_PUSH_RATE:
                st.append(float(base_rate)); i += 1; continue
            if op == R_PUSH_PAR:
                st.append(float(userpar[c[i + 1]])); i += 2; continue
            if op == R_PUSH_NR:
                st.append(int(counts[c[i + 1]])); i += 2; continue
            if op == R_LOAD:
                st.append(aux[c[i + 1]]); i += 2; continue
            if op == R_STORE:
                aux[c[i + 1]] = float(st.pop()); i += 2; continue
            i += 1
            if op == R_I2F4:
                st.append(f32(st.pop()))
            elif op == R_I2F8 or op == R_F4F8:
                st.append(float(st.pop()))
            elif R_NEG <= op <= R_NEG + 2:
                st.append(-st.pop())
            elif R_ADD <= op < R_POW_II:
                b, a = st.pop(), st.pop()
                k = (op - R_ADD) % 3
                base = op - k
                if base == R_ADD:
                    r = a + b
                elif base == R_SUB:
                    r = a - b
                elif base == R_MUL:
                    r = a * b
                elif k == INT:
                    # Fortran integer division truncates toward zero
                    r = abs(a) // abs(b) * (1 if (a >= 0) == (b >= 0) else -1)
                else:
                    r = a / b
                st.append(f32(r) if k == F4 else r)
            elif op == R_POW_II:
                b, a = st.pop(), st.pop()
                st.append(_ipow_int(a, b))
            elif op in (R_POW_F4I, R_POW_F8I):
                b, a = st.pop(), st.pop()
                st.append(_powi(a, b, f4=(op == R_POW_F4I)))
            elif op == R_POW_F4F4:
                b, a = st.pop(), st.pop()
                st.append(f32(np.power(f32(a), f32(b))))
            elif op == R_POW_F8F8:
                b, a = st.pop(), st.pop()
                st.append(math.pow(a, b))
            elif op in (R_EXP, R_EXP + 1, R_LOG, R_LOG + 1, R_SQRT, R_SQRT + 1):
                a = st.pop()
                fn = {R_EXP: np.exp, R_LOG: np.log, R_SQRT: np.sqrt}[op - (op - R_EXP) % 2]
                st.append(f32(fn(f32(a))) if (op - R_EXP) % 2 == 0 else float(fn(np.float64(a))))
            else:
                raise ValueError("bad rate opcode %d" % op)
        out[t] = float(st[-1]) if st else float(aux[0])
    return out


def _powi(a, n, f4=False):
    """x ** n for an integer n: binary powering from the low bit, 1/r for a
    negative n (checked against the reference compiler's results)."""
    cast = np.float32 if f4 else float
    a = cast(a)
    u = -n if n < 0 else n
    r = cast(1.0)
    while u:
        if u & 1:
            r = cast(r * a)
        u >>= 1
        if u:
            a = cast(a * a)
    return cast(cast(1.0) / r) if n < 0 else r


def _ipow_int(a: int, n: int) -> int:
    if n < 0:
        return 1 if a == 1 else ((1 if n % 2 == 0 else -1) if a == -1 else 0)
    return a ** n


# ---------------------------------------------------------------------------
# programs
# ---------------------------------------------------------------------------

def _same_site(a: Coord, b: Coord) -> bool:
    return a.name == b.name and a.layer == b.layer


def _sub(a: Coord, b: Coord) -> Tuple[int, int, int]:
    return (a.offset[0] - b.offset[0], a.offset[1] - b.offset[1], a.offset[2] - b.offset[2])


def _iftree_adds(pt: Project, items):
    """Emission order of _write_optimal_iftree_otf (io:3598-3723): items
    without conditions first, then a select on the most common coordinate,
    its answers sorted, each answer's items pruned and recursed, then the
    items left over.  ``items`` are (conditions [(key, species)], payload)."""
    from .compiler import iftree_order
    return iftree_order(items)


def run_proc_program(pt: Project, exec_proc: Process) -> List[OtfOp]:
    """The ops of run_proc_<exec_proc> (io:3395-3593), in text order."""
    sid = pt.species_ids()
    procs = pt.processes
    nprocs = len(procs)
    inh: List[list] = [[] for _ in range(nprocs)]
    enh: List[list] = [[] for _ in range(nprocs)]
    aff: List[list] = [[] for _ in range(nprocs)]
    for act in exec_proc.actions:
        for ip, proc in enumerate(procs):
            for cond in proc.conditions:
                if _same_site(cond.coord, act.coord):
                    rel = _sub(act.coord, cond.coord)
                    (inh if cond.species != act.species else enh)[ip].append(rel)
            for b in proc.bystanders:
                if _same_site(b.coord, act.coord):
                    aff[ip].append(_sub(act.coord, b.coord))
    # duplicates dropped through a set (io:3442-3455): its iteration order
    # fixes the text order of the calls, so the same construction is kept
    for ip in range(nprocs):
        inh[ip] = [r for r in set(inh[ip])]
    for ip in range(nprocs):
        enh[ip] = [r for r in set(enh[ip]) if r not in inh[ip]]
        aff[ip] = [r for r in set(aff[ip]) if r not in inh[ip]]
    ops: List[OtfOp] = []
    for ip in range(nprocs):
        for rel in inh[ip]:
            ops.append(OtfOp(OP_DEL, proc=ip, off=rel))
    for act in exec_proc.actions:
        prev = [c for c in exec_proc.conditions if c.coord == act.coord]
        if len(prev) != 1:
            raise RuntimeError("process %s: %d conditions on the site of an action"
                               % (exec_proc.name, len(prev)))
        ops.append(OtfOp(OP_WRITE, off=tuple(act.coord.offset),
                         site=pt.site_index(act.coord.layer, act.coord.name),
                         old=sid[prev[0].species], new=sid[act.species]))
    for ip in range(nprocs):
        for rel in aff[ip]:
            ops.append(OtfOp(OP_UPD, proc=ip, off=rel))
    items = []
    for ip in range(nprocs):
        for rel in enh[ip]:
            others = []
            for cond in procs[ip].conditions:
                for act in exec_proc.actions:
                    if _same_site(act.coord, cond.coord) and rel == _sub(act.coord, cond.coord):
                        if act.species != cond.species:
                            raise RuntimeError("process %s: discrepancy in a process to enable"
                                               % exec_proc.name)
                        break
                else:
                    key = (cond.coord.offset[0] + rel[0], cond.coord.offset[1] + rel[1],
                           cond.coord.offset[2] + rel[2], pt.site_index(cond.coord.layer, cond.coord.name))
                    others.append((key, cond.species))
            items.append((others, (ip, rel)))
    for (ip, rel), checks in _iftree_adds(pt, items):
        ops.append(OtfOp(OP_ADD, proc=ip, off=rel,
                         conds=[(k, sid[sp]) for k, sp in checks.items()]))
    return ops


def touchup_program(pt: Project) -> List[OtfOp]:
    """The adds of touchup_cell (io:3240-3284): every process, at offset 0,
    guarded by all of its conditions, in if-tree order."""
    sid = pt.species_ids()
    items = []
    for ip, proc in enumerate(pt.processes):
        conds = [((c.coord.offset[0], c.coord.offset[1], c.coord.offset[2],
                   pt.site_index(c.coord.layer, c.coord.name)), c.species) for c in proc.conditions]
        items.append((conds, (ip, (0, 0, 0))))
    return [OtfOp(OP_ADD, proc=ip, off=rel, conds=[(k, sid[sp]) for k, sp in checks.items()])
            for (ip, rel), checks in _iftree_adds(pt, items)]


def check_otf_model(pt: Project) -> None:
    if len(pt.layers) > 1:
        raise NotImplementedError("the otf backend is built for one-layer models")
    for proc in pt.processes:
        for ca in proc.actions + proc.conditions:
            if ca.species[:1] in ("^", "$") or " or " in ca.species:
                raise NotImplementedError("process %s: %r is not supported by the otf backend"
                                          % (proc.name, ca.species))
