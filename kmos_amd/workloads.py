"""The benchmark workloads: models, lattice sizes, rate sets, step budgets.

quickstart  one replica, 20x20, 1 site/cell, CO adsorption/desorption at
            100 s^-1 each, starting empty, 1e6 steps (the `kmos benchmark`
            model, kmos/cli.py:293-312).
zgb         4096 replicas of the Ziff-Gulari-Barshad model on 128x128,
            y ~ U[0.35, 0.60] per replica, 1e5 steps, starting empty.
ruo2        2048 replicas of the RuO2(110) CO-oxidation model (2 sites/cell,
            36 processes; kmos_amd/data/ruo2_default.xml, the model of the
            reference's export fixtures), sizes 20^2 : 64^2 : 128^2 = 1 : 2 : 1, each rate
            log-uniform in [1e-2, 1e9], initial coverage 0.3 CO / 0.3 O,
            2e5 steps.
pairwise    512 replicas of a 256x256 lattice gas with nearest-neighbour
            interaction eps ~ U[-0.2, 0.2] eV, kT = 0.05 eV, run as lat_int:
            adsorption (1/s), 16 desorption variants (1/s * exp(n eps/kT)),
            4 x 8 diffusion variants (100/s * exp(n eps/kT)), 5e5 steps.
intzgb_otf  2048 replicas of the ZGB model with lateral interactions of the
            reference's otf fixtures (kmos_amd/data/intzgb_otf.xml: 10
            processes whose rates carry J**(+-n) factors of their
            bystanders), 64x64, y ~ U[0.35, 0.60], J_CO_CO, J_O_CO, J_O_O
            ~ U[0.8, 1.25], kDes 0.05, starting empty, 2e4 steps.

Parameters are drawn from numpy's Philox generator keyed by the workload
seed; the kMC uniform stream of replica r is keyed (seed, r).
"""
from __future__ import annotations

import dataclasses
import itertools
import math
import os
from typing import Callable, List, Optional

import numpy as np

from .model import CondAct, Coord, Layer, Process, Project, Site, load_project

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)
DATA = os.path.join(HERE, "data")
RUO2_XML = os.path.join(DATA, "ruo2_default.xml")
INTZGB_OTF_XML = os.path.join(DATA, "intzgb_otf.xml")


# ---------------------------------------------------------------------------
# models
# ---------------------------------------------------------------------------

def _c(species, name, off=(0, 0, 0), layer="default"):
    return CondAct(species, Coord(name=name, offset=tuple(off), layer=layer))


def quickstart_project() -> Project:
    """CO adsorption/desorption on one hollow site per cell."""
    pt = Project(model_name="quickstart", model_dimension=2,
                 species=["CO", "empty"], default_species="empty",
                 layers=[Layer("default", [Site("hollow", (0.5, 0.5, 0.5))])],
                 default_layer="default", substrate_layer="default")
    pt.parameters = {"k_ads": {"value": "100."}, "k_des": {"value": "100."}}
    pt.processes = [
        Process("CO_adsorption", "k_ads", [_c("empty", "hollow")], [_c("CO", "hollow")],
                tof_count={"CO_adsorption": 1}),
        Process("CO_desorption", "k_des", [_c("CO", "hollow")], [_c("empty", "hollow")],
                tof_count={"CO_desorption": 1}),
    ]
    pt.validate()
    return pt


def zgb_project() -> Project:
    """ZGB model without desorption (absorbing states possible)."""
    pt = Project(model_name="zgb", model_dimension=2, species=["CO", "O", "empty"],
                 default_species="empty",
                 layers=[Layer("sc", [Site("site", (0.5, 0.5, 0.5))])],
                 default_layer="sc", substrate_layer="sc")
    pt.parameters = {"yCO": {"value": "0.45"}}
    s = lambda sp, off=(0, 0, 0): _c(sp, "site", off, "sc")  # noqa: E731
    procs = [Process("CO_adsorption", "yCO", [s("empty")], [s("CO")])]
    for i, d in enumerate([(1, 0, 0), (0, 1, 0), (-1, 0, 0), (0, -1, 0)], 1):
        procs.append(Process("CO_oxidation%d" % i, "10**10", [s("CO"), s("O", d)],
                             [s("empty"), s("empty", d)], tof_count={"CO2": 1}))
    for i, d in enumerate([(1, 0, 0), (0, 1, 0)], 1):
        procs.append(Process("O2_adsorption%d" % i, "(1 - yCO)/2.", [s("empty"), s("empty", d)],
                             [s("O"), s("O", d)]))
    pt.processes = procs
    pt.validate()
    return pt


def ruo2_project() -> Project:
    return load_project(RUO2_XML)


NN = [(1, 0, 0), (0, 1, 0), (-1, 0, 0), (0, -1, 0)]
DIRS = {"right": (1, 0, 0), "up": (0, 1, 0), "left": (-1, 0, 0), "down": (0, -1, 0)}


def pairwise_project() -> Project:
    """Lattice gas A/empty with pairwise nearest-neighbour interaction.

    Every interaction configuration is its own process (the lat_int
    convention); ``pairwise_rates`` gives their rate constants.
    """
    pt = Project(model_name="pairwise", model_dimension=2, species=["A", "empty"],
                 default_species="empty",
                 layers=[Layer("sq", [Site("a", (0.0, 0.0, 0.0))])],
                 default_layer="sq", substrate_layer="sq")
    a = lambda sp, off=(0, 0, 0): _c(sp, "a", off, "sq")  # noqa: E731
    procs = [Process("A_adsorption", "1.", [a("empty")], [a("A")])]
    for conf in itertools.product(["empty", "A"], repeat=4):
        n = conf.count("A")
        name = "A_desorption_" + "".join("1" if c == "A" else "0" for c in conf)
        conds = [a(sp, d) for sp, d in zip(conf, NN)] + [a("A")]
        procs.append(Process(name, "1.*exp(%d*eps/kT)" % n, conds, [a("empty")]))
    for dname, d in DIRS.items():
        others = [o for o in NN if o != d]
        for conf in itertools.product(["empty", "A"], repeat=3):
            n = conf.count("A")
            name = "A_diffusion_%s_" % dname + "".join("1" if c == "A" else "0" for c in conf)
            conds = [a("A"), a("empty", d)] + [a(sp, o) for sp, o in zip(conf, others)]
            procs.append(Process(name, "100.*exp(%d*eps/kT)" % n, conds, [a("empty"), a("A", d)]))
    pt.processes = procs
    pt.validate()
    return pt


def pairwise_rates(pt: Project, eps: float, kT: float = 0.05) -> np.ndarray:
    """Rate constants of ``pairwise_project`` for one interaction energy."""
    out = np.zeros(len(pt.processes))
    for i, p in enumerate(pt.processes):
        if p.name == "A_adsorption":
            out[i] = 1.0
        else:
            n = p.name.rsplit("_", 1)[1].count("1")
            base = 1.0 if "desorption" in p.name else 100.0
            out[i] = base * math.exp(n * eps / kT)
    return out


# ---------------------------------------------------------------------------
# workloads
# ---------------------------------------------------------------------------

@dataclasses.dataclass
class Workload:
    name: str
    backend: str
    project: Project
    sizes: np.ndarray          # [n][3]
    rates: np.ndarray          # [n][P]
    steps: int
    seeds: np.ndarray          # kMC stream key (seed, stream) per replica
    streams: np.ndarray
    configs: Optional[List[np.ndarray]] = None   # initial lattices, or None = default
    userpar: Optional[np.ndarray] = None         # otf: [n][K] parameters in name order

    @property
    def n(self):
        return len(self.sizes)

    def describe(self):
        vols = self.sizes.prod(axis=1) * len(self.project.site_list())
        return {"workload": self.name, "backend": self.backend, "replicas": int(self.n),
                "steps_per_replica": int(self.steps), "sites_min": int(vols.min()),
                "sites_max": int(vols.max()), "processes": len(self.project.processes)}


def _gen(seed):
    return np.random.Generator(np.random.Philox(key=seed))


def quickstart(seed=0, steps=1_000_000) -> Workload:
    pt = quickstart_project()
    return Workload("quickstart_20x20", "local_smart", pt, np.array([[20, 20, 1]]),
                    np.array([[100.0, 100.0]]), steps,
                    np.array([seed], np.uint32), np.array([0], np.uint32))


def zgb(n=4096, size=128, steps=100_000, seed=0) -> Workload:
    pt = zgb_project()
    y = _gen(seed).uniform(0.35, 0.60, n)
    order = [p.name for p in pt.processes]
    rates = np.zeros((n, len(order)))
    for i, nm in enumerate(order):
        if nm == "CO_adsorption":
            rates[:, i] = y
        elif nm.startswith("CO_oxidation"):
            rates[:, i] = 1e10
        else:
            rates[:, i] = (1.0 - y) / 2.0
    return Workload("zgb_128x128_x4096", "local_smart", pt, np.tile([size, size, 1], (n, 1)),
                    rates, steps, np.full(n, seed, np.uint32), np.arange(n, dtype=np.uint32))


def ruo2(n=2048, steps=200_000, seed=0, backend="local_smart", sizes=(20, 64, 128),
         mix=(1, 2, 1), coverage=(0.3, 0.3)) -> Workload:
    pt = ruo2_project()
    g = _gen(seed)
    P = len(pt.processes)
    rates = 10.0 ** g.uniform(-2.0, 9.0, (n, P))
    w = np.array(mix, float) / sum(mix)
    counts = np.floor(w * n).astype(int)
    counts[-1] = n - counts[:-1].sum()
    side = np.repeat(np.array(sizes), counts)
    sz = np.stack([side, side, np.ones_like(side)], axis=1)
    ids = pt.species_ids()
    configs = []
    for i in range(n):
        V = int(side[i] * side[i] * 2)
        u = g.random(V)
        c = np.full(V, ids["empty"], np.int8)
        c[u < coverage[0]] = ids["co"]
        c[(u >= coverage[0]) & (u < coverage[0] + coverage[1])] = ids["oxygen"]
        configs.append(c)
    name = "ruo2_%s_mix_x%d" % (backend, n)
    return Workload(name, backend, pt, sz, rates, steps, np.full(n, seed, np.uint32),
                    np.arange(n, dtype=np.uint32), configs)


def pairwise(n=512, size=256, steps=500_000, seed=0, kT=0.05) -> Workload:
    pt = pairwise_project()
    eps = _gen(seed).uniform(-0.2, 0.2, n)
    rates = np.stack([pairwise_rates(pt, e, kT) for e in eps])
    return Workload("pairwise_lat_int_256x256_x%d" % n, "lat_int", pt,
                    np.tile([size, size, 1], (n, 1)), rates, steps,
                    np.full(n, seed, np.uint32), np.arange(n, dtype=np.uint32))


def intzgb_otf_project() -> Project:
    return load_project(INTZGB_OTF_XML)


def intzgb_otf(n=2048, size=64, steps=20_000, seed=0) -> Workload:
    pt = intzgb_otf_project()
    g = _gen(seed)
    y = g.uniform(0.35, 0.60, n)
    J = g.uniform(0.8, 1.25, (n, 3))
    names = sorted(pt.parameters)                  # J_CO_CO, J_O_CO, J_O_O, kDes, yCO
    par = {"J_CO_CO": J[:, 0], "J_O_CO": J[:, 1], "J_O_O": J[:, 2], "kDes": np.full(n, 0.05), "yCO": y}
    userpar = np.stack([par[k] for k in names], axis=1)
    rates = np.zeros((n, len(pt.processes)))
    for i, p in enumerate(pt.processes):
        rates[:, i] = {"yCO": y, "kDes": 0.05, "0.25*yCO": 0.25 * y,
                       "0.5*(1-yCO)": 0.5 * (1 - y)}[p.rate_constant]
    return Workload("intzgb_otf_%dx%d_x%d" % (size, size, n), "otf", pt, np.tile([size, size, 1], (n, 1)),
                    rates, steps, np.full(n, seed, np.uint32), np.arange(n, dtype=np.uint32),
                    userpar=userpar)


WORKLOADS: dict = {"quickstart": quickstart, "zgb": zgb, "ruo2": ruo2, "pairwise": pairwise,
                   "intzgb_otf": intzgb_otf}
