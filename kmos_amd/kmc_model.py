"""``kmc_model``-compatible module objects over the engine's C ABI.

kmos builds its generated Fortran into an f2py extension ``kmc_model`` with
three submodules that ``kmos.run.KMC_Model`` drives
(kmos/run/__init__.py:86-114):  ``base``, ``lattice`` and ``proclist``.
``KmcModule`` exposes the same attribute and function names, backed by one
replica of an engine ``Batch`` on the GPU, so that code written against the
f2py module runs unchanged:

    base.set_rate_const(proc, k)      proclist.init(size, name, layer, seed, no_banner)
    base.get_kmc_time()               proclist.do_kmc_steps(n)
    base.get_procstat(proc)           proclist.get_occupation()
    lattice.get_species([x, y, z, n]) proclist.<process_name> -> number, ...

Numbering follows the generated Fortran: processes 1..P, species 0..S-1 in
name order, sites 1..V, site types 1..spuck.
"""
from __future__ import annotations

import time
from typing import Sequence

import numpy as np

from .compiler import CompiledModel
from .engine import Batch, Model
from .snapshot import ReloadError, read_reload, write_reload


class _Base:
    """kmc_model.base (kmos/fortran_src/base.mpy)."""

    def __init__(self, owner: "KmcModule"):
        self._o = owner
        self.null_species = owner.compiled.null_species

    # rates -----------------------------------------------------------------
    def set_rate_const(self, proc_nr, rate):
        o = self._o
        o._rates[int(proc_nr) - 1] = float(rate)
        if o.batch is not None:
            o.batch.set_rates(o._rates[None, :])

    def get_rate(self, proc_nr):
        return float(self._o._rates[int(proc_nr) - 1])

    def update_accum_rate(self):
        pass        # accum_rates are formed on demand, left to right

    def update_integ_rate(self):
        self._o._b().update_integ_rate(0)

    def get_accum_rate(self, proc_nr):
        return float(self._o._b().accum_rates(0)[int(proc_nr) - 1])

    def get_integ_rate(self, proc_nr):
        return float(self._o._b().integ_rates(0)[int(proc_nr) - 1])

    # clocks ------------------------------------------------------------------
    def get_kmc_time(self):
        return self._o._b().scalars(0)[0]

    def set_kmc_time(self, t):
        self._o._b().set_kmc_time(0, t)

    def get_kmc_time_step(self):
        return self._o._b().scalars(0)[1]

    def get_kmc_step(self):
        return self._o._b().scalars(0)[2]

    def get_walltime(self):
        return self._o._walltime

    # bookkeeping -------------------------------------------------------------
    def get_procstat(self, proc_nr):
        return int(self._o._b().procstat(0)[int(proc_nr) - 1])

    def get_nrofsites(self, proc_nr):
        return int(self._o._b().nr_of_sites(0)[int(proc_nr) - 1])

    def get_avail_site(self, proc_nr, field, switch):
        sites, back = self._o._b().avail(0)
        arr = sites if int(switch) == 1 else back
        return int(arr[int(proc_nr) - 1, int(field) - 1])

    def get_volume(self):
        return int(self._o._b().volumes[0])

    def get_species(self, site):
        return int(self._o._b().lattice(0)[int(site) - 1])

    def get_null_species(self):
        return self.null_species

    # snapshots ---------------------------------------------------------------
    def get_system_name(self):
        return self._o.system_name

    def set_system_name(self, name):
        self._o.system_name = str(name).strip()

    def save_system(self):
        """Write <system_name>.reload in the working directory (base.mpy:517-578)."""
        o = self._o
        st = o._b().get_state(0)
        st.update(walltime=o._walltime, rng_seed=o.proclist.seed, rng_stream=0,
                  system_size=o.lattice.system_size.copy())
        write_reload(o.system_name + ".reload", st)

    def reload_system(self, input_system_name):
        """Allocate the system and restore it from <name>.reload
        (base.mpy:363-514); False when the file does not exist.  Files
        without the engine's rng records continue the random stream of the
        seed last given to proclist.init (1 if none) from kmc_step; files
        without system_size take lattice.system_size when it matches the
        volume, else the square (2-D) or cube (3-D) lattice of that volume."""
        import os
        o = self._o
        name = str(input_system_name).strip()
        path = name + ".reload"
        if not os.path.exists(path):
            return False
        st = read_reload(path)
        cm = o.compiled
        if st["nr_of_proc"] != cm.nr_of_proc:
            raise ReloadError("%s holds %d processes, the model has %d" % (path, st["nr_of_proc"], cm.nr_of_proc))
        size = o._reload_size(st, path)
        seed = int(st.get("rng_seed", o.proclist.seed or 1))
        o.system_name = name
        o._allocate(size, seed, stream=int(st.get("rng_stream", 0)))
        try:
            o._rates[:] = st["rates"]
            o.batch.set_state(0, st)
        except Exception:
            o._free()
            raise
        o.proclist.seed = seed
        o._walltime = float(st.get("walltime", 0.0))
        return True

    def is_allocated(self):
        return self._o.batch is not None

    def deallocate_system(self):
        self._o._free()


class _Lattice:
    """kmc_model.lattice (kmos/fortran_src/lattice.mpy)."""

    def __init__(self, owner: "KmcModule"):
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
        n = s0 % self.spuck + 1
        cell = s0 // self.spuck
        return np.array([cell % X, (cell // X) % Y, cell // (X * Y), n], dtype=np.int32)

    def get_species(self, site):
        return int(self._o._b().lattice(0)[self.calculate_lattice2nr(site) - 1])

    def replace_species(self, site, old_species, new_species):
        self._o._b().replace_species(0, self.calculate_lattice2nr(site), int(old_species), int(new_species))

    def deallocate_system(self):
        self._o._free()


class _Proclist:
    """kmc_model.proclist (the generated proclist module)."""

    def __init__(self, owner: "KmcModule"):
        self._o = owner
        cm = owner.compiled
        self.nr_of_proc = cm.nr_of_proc
        self.nr_of_species = cm.nr_of_species
        self.backend = cm.backend_name
        self.seed = 0
        for i, name in enumerate(cm.process_names, 1):
            setattr(self, name.lower(), i)
        for i, name in enumerate(cm.species_names):
            setattr(self, name.lower(), i)
        self.default_species = cm.project.species_ids()[cm.project.default_species]
        if cm.backend == 0:
            for n, (lay, site) in enumerate(cm.project.site_list(), 1):
                setattr(self, "touchup_%s_%s" % (lay, site), self._touchup_site(n))
        else:
            self.touchup_cell = self._touchup_cell

    def _touchup_site(self, n):
        def touchup(site):
            x, y, z, _ = (int(v) for v in site)
            self._o._b().touchup(0, x, y, z, n)
        return touchup

    def _touchup_cell(self, cell):
        x, y, z = (int(v) for v in list(cell)[:3])
        self._o._b().touchup(0, x, y, z, 1)

    def init(self, input_system_size, system_name="kmc_model", layer=0, seed_in=1, no_banner=True):
        """allocate_system + initialize_state (proclist_generic_subroutines.mpy:166-304)."""
        self._o._allocate(input_system_size, int(seed_in))
        self._o.system_name = str(system_name).strip()
        self.seed = int(seed_in)

    def do_kmc_steps(self, n):
        t0 = time.perf_counter()
        self._o._b().run(int(n))
        self._o._walltime += time.perf_counter() - t0
        self._o._check_status()

    def do_kmc_step(self):
        self.do_kmc_steps(1)

    def get_next_kmc_step(self):
        return self._o._b().get_next_step(0)

    def run_proc_nr(self, proc, nr_site):
        self._o._b().run_proc_nr(0, int(proc), int(nr_site))
        self._o._check_status()

    def get_occupation(self):
        return self._o._b().occupation(0, 1)[0]


class KmcModule:
    """The three f2py submodules of one model, backed by one replica."""

    def __init__(self, compiled: CompiledModel, device: int = 0):
        self.compiled = compiled
        self.device = device
        self.model = Model(compiled)
        self.batch = None
        self._rates = np.zeros(compiled.nr_of_proc)
        self._walltime = 0.0
        self.system_name = "kmc_model"
        self.base = _Base(self)
        self.lattice = _Lattice(self)
        self.proclist = _Proclist(self)

    def _b(self) -> Batch:
        if self.batch is None:
            raise RuntimeError("kmc_model: system not allocated (call proclist.init first)")
        return self.batch

    def view(self):
        """The replica-0 getters KMC_Model reads (scalars, procstat,
        integ_rates, lattice, set_config)."""
        return self._b()

    def _allocate(self, size: Sequence[int], seed: int, stream: int = 0):
        if self.batch is not None:
            raise RuntimeError("kmc_model: tried to allocate the system twice, please deallocate first")
        dim = self.compiled.dimension
        size = [int(v) for v in np.atleast_1d(size)]
        full = (size + [1, 1, 1])[:3]
        for d in range(dim, 3):
            full[d] = 1
        self.batch = Batch(self.model, [full], device=self.device)
        self.batch.init([seed & 0xFFFFFFFF], [stream & 0xFFFFFFFF])
        self.batch.set_rates(self._rates[None, :])
        self.lattice.system_size[:] = full
        self._walltime = 0.0

    def set_rates(self, rates):
        """base.set_rate_const for every process at once."""
        self._rates[:] = rates
        if self.batch is not None:
            self.batch.set_rates(self._rates[None, :])

    def _reload_size(self, st: dict, path: str):
        spuck, dim = self.compiled.spuck, self.compiled.dimension
        V = st["volume"]
        if st.get("system_size") is not None:
            size = [int(v) for v in st["system_size"]]
        elif int(np.prod(self.lattice.system_size)) * spuck == V:
            size = [int(v) for v in self.lattice.system_size]
        else:
            side = int(round((V / spuck) ** (1.0 / dim)))
            size = [side if d < dim else 1 for d in range(3)]
        if int(np.prod(size)) * spuck != V:
            raise ReloadError("%s: volume %d is no lattice of this model's %d sites per cell; "
                              "set lattice.system_size first" % (path, V, spuck))
        return size

    def _free(self):
        if self.batch is not None:
            self.batch.close()
            self.batch = None

    def _check_status(self):
        st = self._b().scalars(0)[4]
        if st != 0:
            from ._lib import REPLICA_STATUS
            raise RuntimeError("kMC replica stopped: %s" % REPLICA_STATUS.get(st, st))


def load_settings(export_dir: str):
    """Import kmc_settings.py of an export directory as a module object."""
    import importlib.util
    import os
    path = os.path.join(export_dir, "kmc_settings.py")
    spec = importlib.util.spec_from_file_location("kmc_settings_%d" % abs(hash(path)), path)
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def compiled_for_export(export_dir: str) -> CompiledModel:
    """Recompile the model of an export directory (its kmc_settings.xml)
    and check it against the stored op program."""
    import os
    from .compiler import compile_model
    from .model import parse_xml
    settings = load_settings(export_dir)
    cm = compile_model(parse_xml(settings.xml), getattr(settings, "backend", "local_smart"))
    stored = np.load(os.path.join(export_dir, "kmcg_model.npy"))
    if not np.array_equal(stored, cm.to_blob()):
        raise RuntimeError("kmcg_model.npy does not match kmc_settings.xml in %s" % export_dir)
    return cm


_modules = {}


def module_for(compiled: CompiledModel, device: int = 0):
    """The engine module of a compiled model: KmcModule, or for the otf
    backend OtfKmcModule (which adds proclist_pars)."""
    from .compiler import BACKEND_OTF
    if compiled.backend == BACKEND_OTF:
        from .kmc_model_otf import OtfKmcModule
        return OtfKmcModule(compiled, device=device)
    return KmcModule(compiled, device=device)


def module_for_export(stub_file: str, device: int = 0):
    """(base, lattice, proclist, proclist_pars) for the export directory
    holding stub_file (proclist_pars None but for otf exports); one engine
    module per directory and process, like an f2py import."""
    import os
    d = os.path.dirname(os.path.abspath(stub_file))
    if d not in _modules:
        _modules[d] = module_for(compiled_for_export(d), device=device)
    m = _modules[d]
    return m.base, m.lattice, m.proclist, getattr(m, "proclist_pars", None)
