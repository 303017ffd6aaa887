"""KMC_Model: the kmos runtime front-end, driving the GPU engine.

Mirrors the part of ``kmos.run.KMC_Model`` (kmos/run/__init__.py:149-1640)
that calls the hot path: construction/``reset``, ``do_steps``,
``get_atoms(geometry=False)``, ``get_std_sampled_data``, ``parameters`` and
``rate_constants``, configuration get/set with ``_adjust_database``, and
``dump_config``/``load_config``.  Geometry output needs ASE, which this
engine does not ship; ``get_atoms(geometry=True)`` raises.

    model = KMC_Model(export_dir="my_export", size=[20, 20])
    model.parameters.yCO = 0.5
    model.do_steps(100000)
    print(model.get_std_sampled_data(10, 100000))
"""
from __future__ import annotations

import copy
import fnmatch

import numpy as np

from .kmc_model import compiled_for_export, load_settings, module_for
from .model import evaluate_rate_expression


class _Expando:
    pass


def parameters_repr(parameters) -> str:
    """The text of Model_Parameters.__repr__ (kmos/run/__init__.py:1670-1692)."""
    fixed = [n for n, p in parameters.items() if not p.get("adjustable", False)]
    res = "# kMC model parameters (%i, fixed %i)\n" % (len(parameters), len(fixed))
    res += "# --------------------\n"
    for n in sorted(parameters):
        res += "# %s = %s" % (n, parameters[n]["value"])
        res += "  # *\n" if parameters[n].get("adjustable", False) else "\n"
    res += "# --------------------\n"
    if len(fixed) != len(parameters):
        res += "# * adjustable parameters\n"
    return res


def rate_constants(settings, parameters, compiled) -> np.ndarray:
    """set_rate_constants (kmos/run/__init__.py:2369-2437): every process's
    rate constant from its expression, in process-number order."""
    names = [n.lower() for n in compiled.process_names]
    rates = np.zeros(len(names))
    for proc in sorted(settings.rate_constants):
        expr = settings.rate_constants[proc][0]
        k = evaluate_rate_expression(expr, parameters)
        if k < 0.0:
            raise UserWarning("%s = %s: Negative rate-constants do no make sense" % (expr, k))
        rates[names.index(proc.lower())] = k
    return rates


def tof_matrix(settings, compiled) -> np.ndarray:
    """[tof, process] counts of each TOF (kmos/run/__init__.py:262-267)."""
    tofs = compiled.project.tof_names()
    names = [n.lower() for n in compiled.process_names]
    m = np.zeros((len(tofs), compiled.nr_of_proc))
    for process, tof_count in sorted(settings.tof_count.items()):
        for tof, factor in tof_count.items():
            m[tofs.index(tof), names.index(process.lower())] += factor
    return m


def std_header(settings, compiled) -> str:
    """get_std_header (kmos/run/__init__.py:856-862): adjustable parameters,
    TOFs, species_site occupations, then the clocks."""
    p = settings.parameters
    params = " ".join(n for n in sorted(p) if p[n].get("adjustable", False))
    tofs = " ".join(compiled.project.tof_names())
    occ = " ".join("%s_%s" % (sp, site) for sp in sorted(settings.representations)
                   for site in settings.site_names)
    return "#%s %s %s kmc_time simulated_time kmc_steps\n" % (params, tofs, occ)


def adjustable_values(parameters) -> list:
    """atoms.params: the adjustable parameters' values, sorted by name."""
    return [float(parameters[p]["value"]) for p in sorted(parameters)
            if parameters[p].get("adjustable", False)]


def interval_tof(tof_matrix, d_procstat, d_integ, delta_t, ncells, tof_method):
    """get_atoms' tof_data / tof_integ over one interval
    (kmos/run/__init__.py:769-800)."""
    if tof_method == "procrates":
        return np.dot(tof_matrix, d_procstat / delta_t / ncells)
    if tof_method == "integ":
        return np.dot(tof_matrix, d_integ / delta_t / ncells)
    raise NotImplementedError('tof_method="%s" not supported' % tof_method)


def std_sampled_output(params, tofs, occs, delta_ts, step_ts, total_time, simulated_time,
                       total_steps, header, output="str"):
    """The averaging and formatting of get_std_sampled_data
    (kmos/run/__init__.py:925-966): occupations weighted by each interval's
    last kmc_time_step, TOFs by the interval's duration."""
    occs_mean = np.average(occs, axis=0, weights=step_ts)
    tof_mean = np.average(tofs, axis=0, weights=delta_ts)
    outdata = tuple(list(params) + list(np.asarray(tof_mean).flatten()) + list(np.asarray(occs_mean).flatten())
                    + [total_time, simulated_time, total_steps])
    if output == "str":
        return (" ".join(["%.5e"] * len(outdata)) + "\n") % outdata
    if output == "dict":
        return dict(zip(header[1:].split(), outdata))
    raise UserWarning("Output format %s not defined. I only know 'str' and 'dict'" % output)


class Model_Parameters:
    """model.parameters.<name> = value re-evaluates every rate constant
    (kmos/run/__init__.py:1643-1714)."""

    def __init__(self, model: "KMC_Model"):
        object.__setattr__(self, "_model", model)

    def __getattr__(self, attr):
        params = object.__getattribute__(self, "_model").settings_parameters
        if attr in params:
            return params[attr]
        raise AttributeError(attr)

    def __setattr__(self, attr, value):
        model = object.__getattribute__(self, "_model")
        if attr not in model.settings_parameters:
            raise AttributeError("don't know parameter '%s'" % attr)
        model.settings_parameters[attr]["value"] = value
        model.set_rate_constants()

    def names(self, pattern=None):
        return [n for n in sorted(self._model.settings_parameters)
                if pattern is None or fnmatch.fnmatch(n, pattern)]

    def __repr__(self):
        return parameters_repr(object.__getattribute__(self, "_model").settings_parameters)


class Model_Rate_Constants:
    """model.rate_constants (kmos/run/__init__.py:1717-1842)."""

    def __init__(self, model: "KMC_Model"):
        self._model = model

    def names(self, pattern=None):
        return [p for p in sorted(self._model.settings.rate_constants)
                if pattern is None or fnmatch.fnmatch(p, pattern)]

    def by_name(self, proc):
        expr = self._model.settings.rate_constants[proc][0]
        return evaluate_rate_expression(expr, self._model.settings_parameters)

    def set(self, pattern, rate_constant, parameters=None):
        if isinstance(rate_constant, str):
            rate_constant = evaluate_rate_expression(rate_constant, parameters or self._model.settings_parameters)
        rate_constant = float(rate_constant)
        for proc in sorted(self._model.settings.rate_constants):
            if pattern is None or fnmatch.fnmatch(proc, pattern):
                self._model.base.set_rate_const(getattr(self._model.proclist, proc.lower()), rate_constant)

    def __call__(self, pattern=None):
        res = ""
        for i, proc in enumerate(self.names(pattern)):
            k = self._model.base.get_rate(getattr(self._model.proclist, proc.lower()))
            res += "# %s: %s = %.2e s^{-1}\n" % (proc, self._model.settings.rate_constants[proc][0], k)
        return res


class KMC_Model:
    def __init__(self, export_dir=None, size=None, system_name="kmc_model", banner=False,
                 print_rates=False, random_seed=None, device=0, compiled=None, settings=None):
        if export_dir is not None:
            self.settings = load_settings(export_dir)
            compiled = compiled_for_export(export_dir)
        elif compiled is None or settings is None:
            raise ValueError("give an export directory, or both compiled and settings")
        else:
            self.settings = settings
        self.settings_parameters = copy.deepcopy(self.settings.parameters)
        self.compiled = compiled
        self._kmc = module_for(compiled, device=device)
        self.base, self.lattice, self.proclist = self._kmc.base, self._kmc.lattice, self._kmc.proclist
        # otf exports also carry the user parameters of their otf_rate expressions
        self.proclist_pars = getattr(self._kmc, "proclist_pars", None)
        self.system_name = system_name
        self.banner = banner
        self.print_rates = print_rates
        dim = compiled.dimension
        if size is None:
            size = self.settings.simulation_size
        if isinstance(size, (int, np.integer)):
            size = [int(size)] * dim
        self.size = np.array(size, dtype=np.int32)
        self.random_seed = self.settings.random_seed if random_seed is None else random_seed
        self.parameters = Model_Parameters(self)
        self.rate_constants = Model_Rate_Constants(self)
        self.reset()
        if hasattr(self.settings, "setup_model"):
            self.settings.setup_model(self)

    # -------------------------------------------------------------------
    def get_backend(self):
        return self.proclist.backend

    def set_rate_constants(self):
        """set_rate_constants (kmos/run/__init__.py:2369-2437)."""
        rates = rate_constants(self.settings, self.settings_parameters, self.compiled)
        if self.proclist_pars is None:
            self._kmc.set_rates(rates)
            return
        # otf: every parameter's value to update_user_parameter, then
        # recalculate_rates_matrix (kmos/run/__init__.py:2415-2432)
        userpar = [evaluate_rate_expression(str(self.settings_parameters[name]["value"]).split(" ")[0],
                                            self.settings_parameters)
                   for name in self.compiled.param_names]
        self._kmc.set_rates(rates, userpar)

    def reset(self):
        """kmos/run/__init__.py:243-339 without the ASE representations."""
        self.proclist.init(self.size, self.system_name, self.lattice.default_layer,
                           self.random_seed, not self.banner)
        self.tofs = self.compiled.project.tof_names()
        self.tof_matrix = tof_matrix(self.settings, self.compiled)
        self.procstat = np.zeros(self.proclist.nr_of_proc, dtype=np.int64)
        self.integ_rates = np.zeros(self.proclist.nr_of_proc)
        self.time = 0.0
        self.steps = 0
        self.set_rate_constants()
        self.base.update_accum_rate()
        self.base.update_integ_rate()

    def deallocate(self):
        if self.base.is_allocated():
            self.lattice.deallocate_system()

    def do_steps(self, n=10000, progress=False):
        """proclist.do_kmc_steps(n) (kmos/run/__init__.py:416-432); like the
        f2py call, a float n is truncated."""
        self.proclist.do_kmc_steps(int(n))

    # -------------------------------------------------------------------
    def get_atoms(self, geometry=True, tag=None, reset_time_overrun=False):
        """kmos/run/__init__.py:680-850 for geometry=False."""
        if geometry:
            raise NotImplementedError("geometry output needs ASE; call get_atoms(geometry=False)")
        atoms = _Expando()
        atoms.calc = None
        b = self._kmc.view()
        t, dt, ks, _, _ = b.scalars(0)
        atoms.kmc_time = t
        atoms.kmc_step = ks
        atoms.params = adjustable_values(self.settings_parameters)
        atoms.procstat = b.procstat(0).astype(np.float64)
        atoms.occupation = self.proclist.get_occupation()
        atoms.integ_rates = b.integ_rates(0)
        delta_t = atoms.kmc_time - self.time
        delta_steps = atoms.kmc_step - self.steps
        size = self.size.prod()
        if delta_steps == 0:
            atoms.tof_data = getattr(self, "tof_data", np.zeros_like(self.tof_matrix[:, 0]))
            atoms.tof_integ = getattr(self, "tof_integ", np.zeros_like(self.tof_matrix[:, 0]))
        elif delta_t == 0.0 and atoms.kmc_time > 0 and reset_time_overrun:
            self.base.set_kmc_time(0.0)
            atoms.tof_data = np.zeros_like(self.tof_matrix[:, 0])
            atoms.tof_integ = np.zeros_like(self.tof_matrix[:, 0])
        else:
            d_ps, d_ig = atoms.procstat - self.procstat, atoms.integ_rates - self.integ_rates
            atoms.tof_data = interval_tof(self.tof_matrix, d_ps, d_ig, delta_t, size, "procrates")
            atoms.tof_integ = interval_tof(self.tof_matrix, d_ps, d_ig, delta_t, size, "integ")
        atoms.delta_t = delta_t
        self.procstat[:] = atoms.procstat
        self.integ_rates[:] = atoms.integ_rates
        self.time = atoms.kmc_time
        self.steps = atoms.kmc_step
        self.tof_data = atoms.tof_data
        self.tof_integ = atoms.tof_integ
        return atoms

    def get_param_header(self):
        return " ".join(p for p in sorted(self.settings_parameters)
                        if self.settings_parameters[p].get("adjustable", False))

    def get_occupation_header(self):
        return " ".join("%s_%s" % (sp, site) for sp in sorted(self.settings.representations)
                        for site in self.settings.site_names)

    def get_tof_header(self):
        return " ".join(self.tofs)

    def get_std_header(self):
        return std_header(self.settings, self.compiled)

    def get_std_sampled_data(self, samples, sample_size, tof_method="integ", output="str",
                             show_progress=False):
        """kmos/run/__init__.py:865-966."""
        occs, tofs, delta_ts, step_ts = [], [], [], []
        t0 = self.base.get_kmc_time()
        step0 = self.base.get_kmc_step()
        self.get_atoms(geometry=False, reset_time_overrun=False)
        for _ in range(samples):
            self.do_steps(sample_size / samples)
            atoms = self.get_atoms(geometry=False, reset_time_overrun=False)
            delta_ts.append(atoms.delta_t)
            step_ts.append(self.base.get_kmc_time_step())
            occs.append(list(atoms.occupation.flatten()))
            if tof_method == "procrates":
                tofs.append(atoms.tof_data.flatten())
            elif tof_method == "integ":
                tofs.append(atoms.tof_integ.flatten())
            else:
                raise NotImplementedError('tof_method="%s" not supported' % tof_method)
        t_end = self.base.get_kmc_time()
        return std_sampled_output(atoms.params, tofs, occs, delta_ts, step_ts, t_end - t0, t_end,
                                  self.base.get_kmc_step() - step0, self.get_std_header(), output)

    # -------------------------------------------------------------------
    def _get_configuration(self):
        X, Y, Z = (int(v) for v in self.lattice.system_size)
        N = int(self.lattice.spuck)
        lat = self._kmc.view().lattice(0)
        return lat.reshape(Z, Y, X, N).transpose(2, 1, 0, 3).copy()

    def _set_configuration(self, config):
        """Replace species site by site, then _adjust_database
        (kmos/run/__init__.py:1411-1437)."""
        X, Y, Z = (int(v) for v in self.lattice.system_size)
        N = int(self.lattice.spuck)
        config = np.asarray(config)
        if config.shape != (X, Y, Z, N):
            raise ValueError("config shape %s does not match model shape %s" % (config.shape, [X, Y, Z, N]))
        flat = config.transpose(2, 1, 0, 3).reshape(-1)
        self._kmc.view().set_config(flat)

    def _adjust_database(self):
        """Touch up every site with x outermost (kmos/run/__init__.py:1439-1457)."""
        b = self._kmc.view()
        b.set_config(b.lattice(0))

    def dump_config(self, filename):
        np.save("%s.npy" % filename, self._get_configuration())

    def load_config(self, filename):
        """kmos/run/__init__.py:1628-1640: set the configuration, then touch
        the database up once more."""
        self._set_configuration(np.load("%s.npy" % filename))
        self._adjust_database()
