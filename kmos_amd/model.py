"""Model description for lattice kMC: species, lattice sites, processes.

A process is a list of conditions (species that must sit on given sites)
and actions (species placed on given sites when the process fires), the
same condition/action model kmos uses (kmos/types.py: Project, Process,
ConditionAction, Coord).  This module reads the two kmos project file
formats -- XML v0.2/v0.3 (kmos/types.py:915-1159, import_xml_file) and
INI (kmos/types.py:623-913, import_ini_file) -- into plain dataclasses, and
applies the numbering conventions of the generated Fortran:

* species are numbered 0-based in name order
  (kmos/fortran_src/proclist_constants.mpy:65-66);
* processes are numbered 1-based in file order
  (kmos/fortran_src/proclist_constants.mpy:75-77);
* sites are numbered 1-based in layer order (kmos/io/__init__.py:3847-3853).

Only what the local_smart and lat_int step loops need is kept; GUI,
representation strings and rate-expression evaluation are carried as text.
"""
from __future__ import annotations

import ast
import configparser
import dataclasses
import math
import os
import xml.etree.ElementTree as ET
from typing import Dict, List, Optional, Tuple

__all__ = [
    "Coord", "CondAct", "Process", "Site", "Layer", "Project",
    "load_project", "parse_xml", "parse_ini",
]


@dataclasses.dataclass(frozen=True)
class Coord:
    """A site relative to the cell a process is anchored at.

    Equality follows kmos/types.py Coord.__eq__: (layer, offset, name).
    """
    name: str
    offset: Tuple[int, int, int] = (0, 0, 0)
    layer: str = ""

    def full_name(self) -> str:
        """``layer_site`` as used for Fortran site constants."""
        return "%s_%s" % (self.layer, self.name) if self.layer else self.name

    def sort_key(self) -> str:
        """String key used to pick a process's executing coordinate
        (kmos/types.py Coord.sort_key): string order, so '-1' < '0'."""
        return "%s_%s_%s_%s_%s" % (self.layer, self.name, self.offset[0],
                                   self.offset[1], self.offset[2])

    def eq_mod_offset(self, other: "Coord") -> bool:
        return (self.layer, self.name) == (other.layer, other.name)


def cmp_coords(a: Coord, b: Coord) -> int:
    """Three-way comparison of kmos/types.py:1950 cmp_coords.

    Compares layer, then offsets; coordinates that differ only by site name
    compare equal, so a stable sort keeps their input order.
    """
    if a.layer != b.layer:
        return (a.layer > b.layer) - (a.layer < b.layer)
    for i in range(3):
        if a.offset[i] != b.offset[i]:
            return (a.offset[i] > b.offset[i]) - (a.offset[i] < b.offset[i])
    return 0


@dataclasses.dataclass(frozen=True)
class CondAct:
    """A condition or an action: ``species`` on ``coord``.

    ``species`` may carry the multi-lattice prefixes '^' (create) or '$'
    (annihilate), or an ``A or B`` alternative, exactly as in kmos files.
    """
    species: str
    coord: Coord
    implicit: bool = False


@dataclasses.dataclass(frozen=True)
class Bystander:
    """A site whose species enters an otf rate without being a condition
    (kmos/types.py Bystander): each species of ``allowed_species`` found
    on ``coord`` counts once towards ``nr_<species>_<flag>`` for every flag
    of the space-separated ``flag``."""
    allowed_species: Tuple[str, ...]
    coord: Coord
    flag: str = ""


@dataclasses.dataclass
class Process:
    name: str
    rate_constant: str = "0."
    conditions: List[CondAct] = dataclasses.field(default_factory=list)
    actions: List[CondAct] = dataclasses.field(default_factory=list)
    enabled: bool = True
    tof_count: Optional[Dict[str, float]] = None
    # otf backend: the rate expression in the local environment and the
    # bystanders it counts (kmos/types.py Process.otf_rate, bystander_list)
    otf_rate: Optional[str] = None
    bystanders: List[Bystander] = dataclasses.field(default_factory=list)

    def executing_coord(self) -> Coord:
        """Lowest ``sort_key`` among the actions (kmos/types.py:2094-2097)."""
        return sorted(self.actions, key=lambda a: a.coord.sort_key())[0].coord


@dataclasses.dataclass
class Site:
    name: str
    pos: Tuple[float, float, float] = (0.0, 0.0, 0.0)
    default_species: str = "default_species"


@dataclasses.dataclass
class Layer:
    name: str
    sites: List[Site] = dataclasses.field(default_factory=list)


@dataclasses.dataclass
class Project:
    model_name: str = "kmc_model"
    model_dimension: int = 2
    author: str = ""
    email: str = ""
    species: List[str] = dataclasses.field(default_factory=list)
    representations: Dict[str, str] = dataclasses.field(default_factory=dict)
    default_species: str = "empty"
    parameters: Dict[str, dict] = dataclasses.field(default_factory=dict)
    layers: List[Layer] = dataclasses.field(default_factory=list)
    default_layer: str = ""
    substrate_layer: str = ""
    cell: Tuple[Tuple[float, float, float], ...] = ((1.0, 0, 0), (0, 1.0, 0), (0, 0, 1.0))
    processes: List[Process] = dataclasses.field(default_factory=list)

    # ---- numbering -------------------------------------------------------
    def species_ids(self) -> Dict[str, int]:
        """0-based ids in name order (proclist_constants.mpy:65-66)."""
        return {name: i for i, name in enumerate(sorted(self.species))}

    def null_species(self) -> int:
        """-1 for one layer (base.mpy:85); nr_of_species for several
        (proclist_constants.mpy:67-68)."""
        return len(self.species) if len(self.layers) > 1 else -1

    def site_list(self) -> List[Tuple[str, str]]:
        """(layer, site) in site-index order (io/__init__.py:3847-3853)."""
        return [(layer.name, site.name) for layer in self.layers for site in layer.sites]

    def site_index(self, layer: str, site: str) -> int:
        """1-based site index within the unit cell."""
        return self.site_list().index((layer, site)) + 1

    def shorten_names(self, max_length: int = 15) -> Dict[str, str]:
        """Rename processes longer than ``max_length - 4`` characters to
        their first ``max_length - 4`` characters plus a 4-digit counter per
        stub (kmos/types.py:559-588; the CLI's export does this with
        ``--variable-length``, default 95).  Returns short -> long names."""
        if max_length < 5:
            raise ValueError("Max variable length has to be at least 5.")
        digits = 4
        stubs: Dict[str, int] = {}
        out = {}
        for proc in self.processes:
            if len(proc.name) > max_length - digits:
                stub = proc.name[:max_length - digits]
                short = "%s%04d" % (stub, stubs.get(stub, 0))
                stubs[stub] = stubs.get(stub, 0) + 1
                out[short] = proc.name
                proc.name = short
        return out

    def process_ids(self) -> Dict[str, int]:
        return {p.name: i + 1 for i, p in enumerate(self.processes)}

    def site_default_species(self) -> List[int]:
        """Species each site type starts with in ``initialize_state``
        (proclist_generic_subroutines.mpy:261-275).  Only sites of the default
        layer receive their default species; the others stay null."""
        ids = self.species_ids()
        out = []
        for layer in self.layers:
            for site in layer.sites:
                if layer.name != self.default_layer:
                    out.append(self.null_species())
                    continue
                sp = site.default_species
                if sp in ("default_species", "", None):
                    sp = self.default_species
                out.append(ids[sp])
        return out

    def tof_names(self) -> List[str]:
        """Sorted TOF labels (kmos/run/__init__.py:2452-2459)."""
        names = []
        for p in self.processes:
            for k in (p.tof_count or {}):
                if k not in names:
                    names.append(k)
        return sorted(names)

    def validate(self) -> None:
        names = set(self.species)
        if self.default_species not in names:
            raise ValueError("default species %r is not a species" % self.default_species)
        seen = set()
        for p in self.processes:
            if p.name in seen:
                raise ValueError("duplicate process name %r" % p.name)
            seen.add(p.name)
            if not p.actions:
                raise ValueError("process %r has no action" % p.name)
            for ca in p.conditions + p.actions:
                for sp in _species_alternatives(ca.species):
                    if sp not in names:
                        raise ValueError("process %r names unknown species %r" % (p.name, sp))
                if (ca.coord.layer, ca.coord.name) not in self.site_list():
                    raise ValueError("process %r names unknown site %s.%s"
                                     % (p.name, ca.coord.layer, ca.coord.name))


def _species_alternatives(species: str) -> List[str]:
    s = species.lstrip("^$")
    return [x.strip() for x in s.split(" or ")]


# ---------------------------------------------------------------------------
# XML (kmos/types.py:915-1159)
# ---------------------------------------------------------------------------

def _floats(text: str) -> List[float]:
    return [float(x) for x in text.split()]


def _cell_from(values: List[float]):
    if len(values) == 3:
        return ((values[0], 0.0, 0.0), (0.0, values[1], 0.0), (0.0, 0.0, values[2]))
    if len(values) == 9:
        return tuple(tuple(values[3 * i:3 * i + 3]) for i in range(3))
    raise ValueError("cell_size %r not understood" % values)


def _parse_tof(text):
    if text is None or text == "":
        return None
    val = ast.literal_eval(text) if isinstance(text, str) else text
    return dict(val) if val else {}


def parse_xml(path_or_text: str) -> Project:
    if os.path.exists(path_or_text):
        root = ET.parse(path_or_text).getroot()
    else:
        root = ET.fromstring(path_or_text)
    version = ast.literal_eval(root.attrib.get("version", "(0, 1)"))
    if tuple(version) not in ((0, 2), (0, 3)):
        raise ValueError("unsupported kmos XML version %r" % (version,))
    pt = Project()
    for child in root:
        if child.tag == "meta":
            pt.model_name = child.attrib.get("model_name", pt.model_name)
            pt.model_dimension = int(child.attrib.get("model_dimension", pt.model_dimension))
            pt.author = child.attrib.get("author", "")
            pt.email = child.attrib.get("email", "")
        elif child.tag == "species_list":
            pt.default_species = child.attrib.get("default_species", "")
            for sp in child:
                pt.species.append(sp.attrib["name"])
                pt.representations[sp.attrib["name"]] = sp.attrib.get("representation", "")
        elif child.tag == "parameter_list":
            for par in child:
                pt.parameters[par.attrib["name"]] = {
                    "value": par.attrib.get("value"),
                    "adjustable": par.attrib.get("adjustable", "False") == "True",
                    "min": float(par.attrib.get("min", 0.0)),
                    "max": float(par.attrib.get("max", 0.0)),
                    "scale": par.attrib.get("scale", "linear"),
                }
        elif child.tag == "lattice":
            pt.cell = _cell_from(_floats(child.attrib["cell_size"]))
            pt.default_layer = child.attrib["default_layer"]
            pt.substrate_layer = child.attrib.get("substrate_layer", pt.default_layer)
            for lay in child:
                if lay.tag != "layer":
                    continue
                layer = Layer(name=lay.attrib["name"])
                for s in lay:
                    pos = s.attrib.get("pos", s.attrib.get("vector", "0 0 0"))
                    layer.sites.append(Site(
                        name=s.attrib["type"],
                        pos=tuple(_floats(pos)),
                        default_species=s.attrib.get("default_species", "default_species")))
                pt.layers.append(layer)
        elif child.tag == "process_list":
            for pe in child:
                enabled = pe.attrib.get("enabled", "True")
                proc = Process(
                    name=pe.attrib["name"],
                    rate_constant=pe.attrib.get("rate_constant", "0."),
                    enabled=(enabled.strip() != "False"),
                    tof_count=_parse_tof(pe.attrib.get("tof_count")),
                    otf_rate=pe.attrib.get("otf_rate"),
                )
                for sub in pe:
                    if sub.tag not in ("condition", "action", "bystander"):
                        continue
                    coord = Coord(
                        name=sub.attrib["coord_name"],
                        offset=tuple(int(i) for i in sub.attrib["coord_offset"].split()),
                        layer=sub.attrib["coord_layer"])
                    if sub.tag == "bystander":
                        proc.bystanders.append(Bystander(
                            allowed_species=tuple(sub.attrib["allowed_species"].split()),
                            coord=coord, flag=sub.attrib.get("flag", "")))
                        continue
                    ca = CondAct(species=sub.attrib["species"], coord=coord,
                                 implicit=sub.attrib.get("implicit", "") == "True")
                    (proc.actions if sub.tag == "action" else proc.conditions).append(ca)
                pt.processes.append(proc)
    if not pt.default_layer and pt.layers:
        pt.default_layer = pt.layers[0].name
    if not pt.substrate_layer:
        pt.substrate_layer = pt.default_layer
    return pt


# ---------------------------------------------------------------------------
# INI (kmos/types.py:623-913)
# ---------------------------------------------------------------------------

def _parse_ini_coord(text: str, first_layer: str) -> Coord:
    parts = text.strip().split(".")
    if len(parts) == 3:
        name, off, layer = parts
    elif len(parts) == 2:
        (name, off), layer = parts, first_layer
    else:
        name, off, layer = parts[0], "(0, 0, 0)", first_layer
    off = tuple(int(v) for v in ast.literal_eval(off))
    off = tuple(list(off) + [0] * (3 - len(off)))
    return Coord(name=name.strip(), offset=off, layer=layer.strip())


def parse_ini(path_or_text: str) -> Project:
    cp = configparser.ConfigParser()
    cp.optionxform = str
    if os.path.exists(path_or_text):
        with open(path_or_text) as fh:
            text = fh.read()
    else:
        text = path_or_text
    if "#@" in text:
        raise ValueError("templated INI files are not supported")
    cp.read_string(text)
    pt = Project()
    layer_sections = [s for s in cp.sections() if s.startswith("Layer")]
    first_layer = layer_sections[0].split()[-1] if layer_sections else ""
    for sec in cp.sections():
        opts = cp[sec]
        if sec == "Meta":
            pt.model_name = opts.get("model_name", pt.model_name)
            pt.model_dimension = int(opts.get("model_dimension", pt.model_dimension))
            pt.author = opts.get("author", "")
            pt.email = opts.get("email", "")
        elif sec == "SpeciesList":
            pt.default_species = opts.get("default_species", "")
        elif sec.startswith("Species "):
            name = sec.split()[-1]
            pt.species.append(name)
            pt.representations[name] = opts.get("representation", "")
        elif sec.startswith("Parameter "):
            pt.parameters[sec.split()[-1]] = {
                "value": opts.get("value"),
                "adjustable": opts.get("adjustable", "False") == "True",
                "min": float(opts.get("min", 0.0)),
                "max": float(opts.get("max", 0.0)),
                "scale": opts.get("scale", "linear"),
            }
        elif sec == "Lattice":
            if "cell_size" in opts:
                pt.cell = _cell_from(_floats(opts["cell_size"]))
            pt.default_layer = opts.get("default_layer", pt.default_layer)
            pt.substrate_layer = opts.get("substrate_layer", pt.substrate_layer)
        elif sec.startswith("Layer "):
            layer = Layer(name=sec.split()[-1])
            for key in opts:
                if not key.startswith("site"):
                    continue
                fields = opts[key].split(";")
                pos = tuple(float(v) for v in ast.literal_eval(fields[0]))
                default = fields[1].strip() if len(fields) > 1 else "default_species"
                layer.sites.append(Site(name=key.split()[-1].strip(), pos=pos,
                                        default_species=default or "default_species"))
            pt.layers.append(layer)
        elif sec.startswith("Process "):
            proc = Process(
                name=sec.split()[-1],
                rate_constant=opts.get("rate_constant", "0."),
                enabled=opts.getboolean("enabled", True),
                tof_count=_parse_tof(opts.get("tof_count")) if "tof_count" in opts else None,
            )
            otf = opts.get("otf_rate")
            proc.otf_rate = None if otf is None or otf.strip() == "None" else otf
            # bystanders = ['A', 'B']@site.(dx, dy, dz)|flag + ...
            # (kmos/types.py:855-890)
            for item in (opts.get("bystanders") or "").split("+"):
                if not item.strip():
                    continue
                allowed, rest = item.strip().split("@")
                coord_text, flag = rest.split("|")
                proc.bystanders.append(Bystander(
                    allowed_species=tuple(ast.literal_eval(allowed.strip())),
                    coord=_parse_ini_coord(coord_text, first_layer), flag=flag.strip()))
            for item in opts["actions"].split("+"):
                sp, co = item.strip().split("@")
                proc.actions.append(CondAct(sp.strip(), _parse_ini_coord(co, first_layer)))
            for item in opts["conditions"].split("+"):
                sp, co = item.strip().split("@")
                proc.conditions.append(CondAct(sp.strip(), _parse_ini_coord(co, first_layer)))
            pt.processes.append(proc)
    if not pt.default_layer and pt.layers:
        pt.default_layer = pt.layers[0].name
    if not pt.substrate_layer:
        pt.substrate_layer = pt.default_layer
    return pt


def load_project(path: str) -> Project:
    """Read a kmos ``.xml`` or ``.ini`` project file, or kmos project XML
    text (what ``str(kmos.types.Project)`` returns)."""
    if path.lstrip().startswith("<"):
        pt = parse_xml(path)
    elif path.endswith(".ini"):
        pt = parse_ini(path)
    elif path.endswith(".xml"):
        pt = parse_xml(path)
    else:
        raise ValueError("don't know what to do with file ending of %s" % path)
    pt.validate()
    return pt


# ---------------------------------------------------------------------------
# rate expressions
# ---------------------------------------------------------------------------

_SAFE_FUNCS = {"exp": math.exp, "log": math.log, "sqrt": math.sqrt, "pi": math.pi,
               "sin": math.sin, "cos": math.cos}


def evaluate_numeric_rate(expr: str, parameters: Dict[str, float]) -> float:
    """Evaluate a purely numeric rate expression such as ``10**8`` or
    ``(1 - yCO)/2.`` with the given parameter values.

    kmos's full evaluator (kmos/__init__.py:67) also knows physical
    constants and JANAF chemical potentials; models used with this engine
    give numeric rates, so only arithmetic, parameters and exp/log/sqrt
    are accepted here.
    """
    tree = ast.parse(expr.strip(), mode="eval")
    allowed = (ast.Expression, ast.BinOp, ast.UnaryOp, ast.Constant, ast.Name,
               ast.Load, ast.Call, ast.Add, ast.Sub, ast.Mult, ast.Div, ast.Pow,
               ast.USub, ast.UAdd, ast.Mod)
    for node in ast.walk(tree):
        if not isinstance(node, allowed):
            raise ValueError("rate expression %r uses unsupported syntax" % expr)
        if isinstance(node, ast.Name) and node.id not in parameters and node.id not in _SAFE_FUNCS:
            raise ValueError("rate expression %r: unknown name %r" % (expr, node.id))
    env = dict(_SAFE_FUNCS)
    env.update({k: float(v) for k, v in parameters.items()})
    return float(eval(compile(tree, "<rate>", "eval"), {"__builtins__": {}}, env))


# ---------------------------------------------------------------------------
# XML writer (the layout kmos/types.py Project._get_xml_string produces)
# ---------------------------------------------------------------------------

def to_xml(pt: Project) -> str:
    """Serialise a project as a kmos XML (version 0.2) document."""
    root = ET.Element("kmc", version="(0, 2)")
    ET.SubElement(root, "meta", author=pt.author, email=pt.email, debug="0",
                  model_dimension=str(pt.model_dimension), model_name=pt.model_name)
    sl = ET.SubElement(root, "species_list", default_species=pt.default_species)
    for name in pt.species:
        ET.SubElement(sl, "species", name=name, color="#ffffff",
                      representation=pt.representations.get(name, ""))
    pl = ET.SubElement(root, "parameter_list")
    for name, par in pt.parameters.items():
        ET.SubElement(pl, "parameter", name=name, value=str(par.get("value")),
                      adjustable=str(bool(par.get("adjustable", False))),
                      min=str(par.get("min", 0.0)), max=str(par.get("max", 0.0)),
                      scale=str(par.get("scale", "linear")))
    cell = " ".join(str(float(v)) for row in pt.cell for v in row)
    lat = ET.SubElement(root, "lattice", cell_size=cell, default_layer=pt.default_layer,
                        substrate_layer=pt.substrate_layer, representation="")
    for layer in pt.layers:
        le = ET.SubElement(lat, "layer", name=layer.name, color="#ffffff")
        for s in layer.sites:
            ET.SubElement(le, "site", type=s.name, pos=" ".join(str(float(v)) for v in s.pos),
                          default_species=s.default_species)
    prl = ET.SubElement(root, "process_list")
    for p in pt.processes:
        attrs = dict(name=p.name, rate_constant=p.rate_constant, enabled=str(bool(p.enabled)))
        if p.tof_count:
            attrs["tof_count"] = repr(p.tof_count)
        if p.otf_rate:
            attrs["otf_rate"] = p.otf_rate
        pe = ET.SubElement(prl, "process", **attrs)
        for tag, lst in (("condition", p.conditions), ("action", p.actions)):
            for ca in lst:
                ET.SubElement(pe, tag, species=ca.species, coord_layer=ca.coord.layer,
                              coord_name=ca.coord.name,
                              coord_offset=" ".join(str(int(v)) for v in ca.coord.offset))
        for b in p.bystanders:
            attrs = dict(allowed_species=" ".join(b.allowed_species), coord_layer=b.coord.layer,
                         coord_name=b.coord.name,
                         coord_offset=" ".join(str(int(v)) for v in b.coord.offset))
            if b.flag:
                attrs["flag"] = b.flag
            ET.SubElement(pe, "bystander", **attrs)
    ET.SubElement(root, "output_list")
    ET.indent(root)
    return '<?xml version="1.0" ?>\n' + ET.tostring(root, encoding="unicode") + "\n"


# ---------------------------------------------------------------------------
# rate expressions in kmos's vocabulary (kmos/__init__.py:64-140, units.py)
# ---------------------------------------------------------------------------

UNITS = {
    "pi": 3.14159265358979323846, "c": 2.99792458e8, "h": 6.62606957e-34,
    "hbar": 1.054571726e-34, "eV": 1.602176565e-19, "kboltzmann": 1.3806488e-23,
    "umass": 1.660538921e-27, "angstrom": 1.0e-10, "bar": 1.0e5,
}


def evaluate_rate_expression(rate_expr: str, parameters: Dict[str, dict]) -> float:
    """Rate constant of a kmos rate expression.

    Supports numbers, the model's parameters (whose values may themselves
    be expressions), kmos's unit constants, the alias
    ``beta = 1/(kboltzmann*T)`` and exp/log/sqrt/sin/cos/pow.  Species
    masses (``m_X``) and JANAF chemical potentials (``mu_X``) need ASE and
    the JANAF tables, which this engine does not ship; such expressions
    raise ValueError.
    """
    if not rate_expr or not str(rate_expr).strip():
        return 0.0
    expr = str(rate_expr).replace("beta", "(1/(kboltzmann*T))")
    env: Dict[str, float] = dict(UNITS)
    env.update({"exp": math.exp, "log": math.log, "sqrt": math.sqrt, "sin": math.sin,
                "cos": math.cos, "pow": math.pow})
    resolving = set()

    def param_value(name):
        if name in resolving:
            raise ValueError("parameter %r is defined in terms of itself" % name)
        resolving.add(name)
        v = parameters[name]["value"] if isinstance(parameters[name], dict) else parameters[name]
        out = float(v) if isinstance(v, (int, float)) else evaluate_rate_expression(str(v).split()[0], parameters)
        resolving.discard(name)
        return out

    tree = ast.parse(expr.strip(), mode="eval")
    for node in ast.walk(tree):
        if isinstance(node, ast.Name) and node.id not in env:
            if node.id in parameters:
                env[node.id] = param_value(node.id)
            elif node.id.startswith("m_") or node.id.startswith("mu_"):
                raise ValueError("%s needs ASE/JANAF data, which is not available" % node.id)
            else:
                raise ValueError("rate expression %r: unknown name %r" % (rate_expr, node.id))
        if isinstance(node, (ast.Attribute, ast.Subscript, ast.Lambda, ast.comprehension)):
            raise ValueError("rate expression %r uses unsupported syntax" % rate_expr)
    return float(eval(compile(tree, "<rate>", "eval"), {"__builtins__": {}}, env))
