"""``kmc_model``-compatible module objects of an otf export.

An otf export of kmos builds into an f2py module with four submodules
(kmos/run/__init__.py:86-114): ``base`` (kmos/fortran_src/base_otf.f90),
``lattice``, ``proclist`` and ``proclist_pars`` (the user parameters the
otf_rate expressions read, io:2657-2780).  ``OtfKmcModule`` exposes the
names ``kmos.run.KMC_Model`` and user scripts call, backed by one replica
of an ``OtfBatch``:

    base.set_rate_const(proc, k)               proclist.init(size, name, layer, seed)
    proclist_pars.update_user_parameter(i, v)  proclist.recalculate_rates_matrix()
    base.get_rate(proc, site)                  proclist.do_kmc_steps(n)

Numbering follows the generated Fortran: processes 1..P, species 0..S-1 in
name order, sites 1..V, parameters 1..K in name order.
"""
from __future__ import annotations

import time
from typing import Sequence

import numpy as np

from .compiler import CompiledModel
from .otf_engine import OtfBatch, OtfModel


class _View:
    """The replica-0 getters KMC_Model reads from an engine batch."""

    def __init__(self, owner: "OtfKmcModule"):
        self._o = owner

    def scalars(self, r):
        t, dt, k, s = self._o.batch.scalars_full(r)
        return t, dt, k, 0, s

    def procstat(self, r):
        return self._o.batch.counters(r)["procstat"]

    def integ_rates(self, r):
        return self._o.batch.counters(r)["integ_rates"]

    def lattice(self, r):
        return self._o.batch.lattice(r)

    def set_config(self, flat):
        self._o.batch.set_config([np.asarray(flat)])


class _Base:
    """kmc_model.base of an otf export (base_otf.f90)."""

    def __init__(self, owner: "OtfKmcModule"):
        self._o = owner
        self.null_species = -1

    def set_rate_const(self, proc_nr, rate):
        """Takes effect at proclist.recalculate_rates_matrix (as in the
        reference, where only new entries see a new constant before it)."""
        self._o._rates[int(proc_nr) - 1] = float(rate)

    def get_rate(self, proc_nr, site_nr=-1):
        """rates_matrix(proc, site), or the list's sum for site -1
        (base_otf.f90:1121-1146)."""
        _, _, rm = self._o._b().avail(0)
        return float(rm[int(proc_nr) - 1, -1 if int(site_nr) == -1 else int(site_nr) - 1])

    def update_accum_rate(self):
        pass        # every step re-sums the lists itself

    def update_integ_rate(self):
        pass

    def get_accum_rate(self, proc_nr=0):
        acc = self._o._b().counters(0)["accum_rates"]
        return float(acc[-1] if int(proc_nr) == 0 else acc[int(proc_nr) - 1])

    def get_integ_rate(self, proc_nr=0):
        ig = self._o._b().counters(0)["integ_rates"]
        return float(ig[-1] if int(proc_nr) == 0 else ig[int(proc_nr) - 1])

    def get_kmc_time(self):
        return self._o._b().scalars_full(0)[0]

    def get_kmc_time_step(self):
        return self._o._b().scalars_full(0)[1]

    def get_kmc_step(self):
        return self._o._b().scalars_full(0)[2]

    def set_kmc_time(self, t):
        raise NotImplementedError("set_kmc_time is not provided by the otf engine")

    def get_walltime(self):
        return self._o._walltime

    def get_procstat(self, proc_nr):
        return int(self._o._b().counters(0)["procstat"][int(proc_nr) - 1])

    def get_nrofsites(self, proc_nr):
        return int(self._o._b().counters(0)["nr_of_sites"][int(proc_nr) - 1])

    def get_avail_site(self, proc_nr, field, switch):
        sites, back, _ = self._o._b().avail(0)
        arr = sites if int(switch) == 1 else back
        return int(arr[int(proc_nr) - 1, int(field) - 1])

    def get_volume(self):
        return int(self._o._b().volumes[0])

    def get_species(self, site):
        return int(self._o._b().lattice(0)[int(site) - 1])

    def get_null_species(self):
        return self.null_species

    def is_allocated(self):
        return self._o.batch is not None

    def deallocate_system(self):
        self._o._free()


class _Lattice:
    """kmc_model.lattice."""

    def __init__(self, owner: "OtfKmcModule"):
        self._o = owner
        pt = owner.compiled.project
        self.spuck = owner.compiled.spuck
        self.model_dimension = pt.model_dimension
        self.system_size = np.zeros(3, np.int32)
        self.unit_cell_size = np.array(pt.cell, dtype=np.float64)
        self.site_positions = np.array([s.pos for l in pt.layers for s in l.sites], dtype=np.float64)
        for i, layer in enumerate(pt.layers):
            setattr(self, layer.name.lower(), i)
        self.default_layer = getattr(self, pt.default_layer.lower())
        self.substrate_layer = getattr(self, pt.substrate_layer.lower())
        for n, (lay, site) in enumerate(pt.site_list(), 1):
            setattr(self, ("%s_%s" % (lay, site)).lower(), n)

    def calculate_lattice2nr(self, site):
        x, y, z, n = (int(v) for v in site)
        X, Y, Z = (int(v) for v in self.system_size)
        return self.spuck * ((x % X) + X * (y % Y) + X * Y * (z % Z)) + n

    def calculate_nr2lattice(self, nr):
        X, Y = int(self.system_size[0]), int(self.system_size[1])
        s0 = int(nr) - 1
        cell = s0 // self.spuck
        return np.array([cell % X, (cell // X) % Y, cell // (X * Y), s0 % self.spuck + 1], dtype=np.int32)

    def get_species(self, site):
        return int(self._o._b().lattice(0)[self.calculate_lattice2nr(site) - 1])

    def deallocate_system(self):
        self._o._free()


class _Proclist:
    """kmc_model.proclist of an otf export."""

    def __init__(self, owner: "OtfKmcModule"):
        self._o = owner
        cm = owner.compiled
        self.nr_of_proc = cm.nr_of_proc
        self.nr_of_species = cm.nr_of_species
        self.backend = "otf"
        self.seed = 0
        for i, name in enumerate(cm.process_names, 1):
            setattr(self, name.lower(), i)
        for i, name in enumerate(cm.species_names):
            setattr(self, name.lower(), i)
        self.default_species = cm.project.species_ids()[cm.project.default_species]

    def init(self, input_system_size, system_name="kmc_model", layer=0, seed_in=1, no_banner=True):
        self._o._allocate(input_system_size, int(seed_in))
        self._o.system_name = str(system_name).strip()
        self.seed = int(seed_in)

    def recalculate_rates_matrix(self):
        """Every listed entry takes its cell's rate under the current rate
        constants and parameters."""
        self._o._push_params()

    def do_kmc_steps(self, n):
        o = self._o
        t0 = time.perf_counter()
        o._b().run(int(n))
        o._walltime += time.perf_counter() - t0
        o._check_status()

    def do_kmc_step(self):
        self.do_kmc_steps(1)

    def get_occupation(self):
        """occupation(species, site) as a fraction of the cells."""
        o = self._o
        lat = o._b().lattice(0).astype(np.int64).reshape(-1, o.compiled.spuck)
        S = o.compiled.nr_of_species
        out = np.zeros((S, o.compiled.spuck))
        for n in range(o.compiled.spuck):
            out[:, n] = np.bincount(lat[:, n], minlength=S)[:S]
        return out / lat.shape[0]


class _ProclistPars:
    """kmc_model.proclist_pars: the user parameters, 1..K in name order."""

    def __init__(self, owner: "OtfKmcModule"):
        self._o = owner
        for i, name in enumerate(owner.compiled.param_names, 1):
            setattr(self, name.lower(), i)

    def update_user_parameter(self, param, val):
        self._o._userpar[int(param) - 1] = float(val)

    def get_user_parameter(self, param):
        return float(self._o._userpar[int(param) - 1])


class OtfKmcModule:
    """The f2py submodules of one otf model, backed by one replica."""

    def __init__(self, compiled: CompiledModel, device: int = 0):
        self.compiled = compiled
        self.device = device
        self.model = OtfModel(compiled)
        self.batch = None
        self._rates = np.zeros(compiled.nr_of_proc)
        self._userpar = np.zeros(max(len(compiled.param_names), 1))
        self._walltime = 0.0
        self._params_set = False
        self.system_name = "kmc_model"
        self.base = _Base(self)
        self.lattice = _Lattice(self)
        self.proclist = _Proclist(self)
        self.proclist_pars = _ProclistPars(self)
        self._view = _View(self)

    def _b(self) -> OtfBatch:
        if self.batch is None:
            raise RuntimeError("kmc_model: system not allocated (call proclist.init first)")
        return self.batch

    def view(self):
        """The replica-0 getters KMC_Model reads (scalars, procstat,
        integ_rates, lattice, set_config)."""
        self._b()
        return self._view

    def _allocate(self, size: Sequence[int], seed: int, stream: int = 0):
        if self.batch is not None:
            raise RuntimeError("kmc_model: tried to allocate the system twice, please deallocate first")
        dim = self.compiled.dimension
        size = [int(v) for v in np.atleast_1d(size)]
        full = (size + [1, 1, 1])[:3]
        for d in range(dim, 3):
            full[d] = 1
        self.batch = OtfBatch(self.model, [full], device=self.device)
        # like allocate_system, whose rates start at zero: until rate
        # constants and parameters are given, every entry's rate is 0
        if self._params_set:
            self._push_params()
        self.batch.init([seed & 0xFFFFFFFF], [stream & 0xFFFFFFFF])
        self.lattice.system_size[:] = full
        self._walltime = 0.0

    def _push_params(self):
        self._params_set = True
        if self.batch is not None:
            K = len(self.compiled.param_names)
            self.batch.set_params(self._rates[None, :], self._userpar[None, :K] if K else None)

    def set_rates(self, rates, userpar=None):
        """set_rate_const for every process and, with userpar,
        update_user_parameter for every parameter; then
        recalculate_rates_matrix."""
        self._rates[:] = rates
        if userpar is not None and len(self.compiled.param_names):
            self._userpar[:len(userpar)] = userpar
        self._push_params()

    def _free(self):
        if self.batch is not None:
            self.batch.close()
            self.batch = None

    def _check_status(self):
        st = self._b().scalars_full(0)[3]
        if st != 0:
            from ._lib import REPLICA_STATUS
            raise RuntimeError("kMC replica stopped: %s" % REPLICA_STATUS.get(st, st))
