"""Parameter sweeps on the batched engine: ModelRunner.

kmos's ModelRunner (kmos/run/__init__.py:1898-2366) scans a regular grid of
model parameters: for every grid point it builds a KMC_Model, runs
`init_steps`, samples `sample_steps` with get_std_sampled_data and appends
the line to ``<runner_name>.dat``; points are spread over `cores` worker
processes and claimed through ``<runner_name>.lock``.

Here every grid point is one replica of one engine batch: the whole grid
advances in one kernel launch per phase, the sampling intervals are
recorded inside the kernel (kmcg_run_recorded), and the host turns the
records into the same get_std_sampled_data line that KMC_Model writes for
that point run alone (shared code in kmos_amd.run).  As in kmos, every
point uses the same seed (`random_seed`, else the model's), so a point's
line equals the line of ``KMC_Model(random_seed=seed)`` with those
parameters.  ``run_steady_state`` does the same for sample_steady_state
(kmos/run/steady_state.py:135-267): replicas that pass the EWMA test drop
out of the batch while the others continue.

    class ScanKinetics(ModelRunner):
        p_O2gas = PressureParameter(1)
        T = TemperatureParameter(600)
        p_COgas = PressureParameter(min=1, max=10, steps=40)

    ScanKinetics(export_dir="my_export").run(init_steps=1e7, sample_steps=1e7)
"""
from __future__ import annotations

import copy
import itertools
import os
from collections import OrderedDict

import numpy as np

from .engine import Batch, Model
from .kmc_model import compiled_for_export, load_settings
from .run import (adjustable_values, interval_tof, parameters_repr, rate_constants, std_header,
                  std_sampled_output, tof_matrix)
from .steady_state import SteadyStateMonitor


def T_grid(T_min, T_max, n):
    """n temperatures from T_min to T_max, evenly spaced in 1/T
    (kmos/utils/__init__.py:1420-1435)."""
    grid = list(np.linspace(T_max ** (-1.0), T_min ** (-1.0), n))
    grid.reverse()
    return [x ** (-1.0) for x in grid]


def p_grid(p_min, p_max, n):
    """n pressures from p_min to p_max, evenly spaced in log p
    (kmos/utils/__init__.py:1438-1446)."""
    return np.logspace(np.log10(p_min), np.log10(p_max), n)


class ModelParameter:
    """A parameter to scan; with `min` only it stays fixed
    (kmos/run/__init__.py:1898-1929)."""

    def __init__(self, min, max=None, steps=1, type=None, unit=""):
        self.min = min
        self.max = max if max is not None else min
        self.steps = steps
        self.type = type
        self.unit = unit

    def __repr__(self):
        return "[%s] min: %s, max: %s, steps: %s" % (self.type, self.min, self.max, self.steps)

    def get_grid(self):
        raise NotImplementedError("use a subclass of ModelParameter")


class PressureParameter(ModelParameter):
    """Grid regular in ln p."""

    def __init__(self, *args, **kwargs):
        kwargs["type"], kwargs["unit"] = "pressure", "bar"
        super().__init__(*args, **kwargs)

    def get_grid(self):
        return p_grid(self.min, self.max, self.steps)


class TemperatureParameter(ModelParameter):
    """Grid regular in 1/T."""

    def __init__(self, *args, **kwargs):
        kwargs["type"], kwargs["unit"] = "temperature", "K"
        super().__init__(*args, **kwargs)

    def get_grid(self):
        return T_grid(self.min, self.max, self.steps)


class LogParameter(ModelParameter):
    """Grid 10**min .. 10**max, like np.logspace."""

    def __init__(self, *args, **kwargs):
        kwargs["type"] = "log"
        super().__init__(*args, **kwargs)

    def get_grid(self):
        return np.logspace(self.min, self.max, self.steps)


class LinearParameter(ModelParameter):
    """Regular grid min .. max."""

    def __init__(self, *args, **kwargs):
        kwargs["type"] = "linear"
        super().__init__(*args, **kwargs)

    def get_grid(self):
        return np.linspace(self.min, self.max, self.steps)


class _ModelRunner(type):
    """Collects the class's ModelParameters in definition order and names
    the runner after the class (kmos/run/__init__.py:1991-2002)."""

    def __new__(mcs, name, bases, dct):
        obj = super().__new__(mcs, name, bases, dct)
        obj.runner_name = name
        obj.parameters = OrderedDict((k, v) for k, v in dct.items() if isinstance(v, ModelParameter))
        return obj


class ModelRunner(metaclass=_ModelRunner):
    """Subclass with ModelParameter attributes, then call run()."""

    def __init__(self, export_dir=".", size=None, device=0, max_replicas=None, directory="."):
        self.export_dir = export_dir
        self.settings = load_settings(export_dir)
        self.compiled = compiled_for_export(export_dir)
        dim = self.compiled.dimension
        size = self.settings.simulation_size if size is None else size
        if isinstance(size, (int, np.integer)):
            size = [int(size)] * dim
        self.size = [int(v) for v in size]
        self.device = device
        self.max_replicas = max_replicas
        self.directory = directory
        unknown = [k for k in self.parameters if k not in self.settings.parameters]
        if unknown:
            raise ValueError("%s scans parameters the model does not have: %s" % (self.runner_name, unknown))

    # --- grid ----------------------------------------------------------------
    def points(self):
        """Every grid point, the first parameter varying slowest."""
        return list(itertools.product(*[tuple(p.get_grid()) for p in self.parameters.values()]))

    def input_line(self, point) -> str:
        return "_".join(["%s"] * (len(self.parameters) + 1)) % tuple([self.runner_name] + list(point))

    def point_parameters(self, point) -> dict:
        params = copy.deepcopy(self.settings.parameters)
        for name, value in zip(self.parameters, point):
            params[name]["value"] = value
        return params

    def std_header(self) -> str:
        return std_header(self.settings, self.compiled)

    def _path(self, suffix):
        return os.path.join(self.directory, "%s%s" % (self.runner_name, suffix))

    def _claim(self, points):
        """The lock-file protocol: skip points already listed, list the rest."""
        lock = self._path(".lock")
        seen = set()
        if os.path.exists(lock):
            with open(lock) as f:
                seen = {line.strip() for line in f}
        todo = [pt for pt in points if self.input_line(pt) not in seen]
        with open(lock, "a") as f:
            for pt in todo:
                f.write("%s\n" % self.input_line(pt))
        return todo

    def _write(self, lines, first_params):
        out = self._path(".dat")
        if not lines:
            return
        if not os.path.exists(out):
            with open(out, "a") as f:
                f.write(self.std_header())
                f.write(parameters_repr(first_params))
                f.write("# If one or more parameters change between data lines\n"
                        "# the set above corresponds to the first line.\n")
        with open(out, "a") as f:
            for line in lines:
                f.write(line)

    def _batch(self, points, seed):
        """One replica per point, every replica with the same seed and stream
        (the sequence of a KMC_Model built with that seed)."""
        m = Model(self.compiled)
        n = len(points)
        b = Batch(m, [self.size + [1] * (3 - len(self.size))] * n, device=self.device)
        b.init(np.full(n, seed & 0xFFFFFFFF, np.uint32), np.zeros(n, np.uint32))
        rates = np.stack([rate_constants(self.settings, self.point_parameters(pt), self.compiled)
                          for pt in points])
        b.set_rates(rates)
        return m, b

    def _chunks(self, points):
        cap = self.max_replicas or len(points)
        if self.max_replicas is None:
            m = Model(self.compiled)
            per = m.footprint([self.size + [1] * (3 - len(self.size))])
            cap = max(1, int((64 << 30) // max(per, 1)))
        for i in range(0, len(points), cap):
            yield points[i:i + cap]

    def _sampled(self, b, ids, base, samples, interval, tof_method, tof_matrix, ncells, prev_tof):
        """get_std_sampled_data(samples, samples*interval) of replicas `ids`,
        from `base` (the last get_atoms state of each replica): run and record
        in one launch, then average on the host like KMC_Model."""
        bud = np.full(len(ids), samples * interval, np.int64)
        _, recs = b.run_recorded(samples * interval, ids, bud, max(interval, 1), samples)()
        out = []
        for j, rec in enumerate(recs):
            t_prev, s_prev = base["kmc_time"][j], base["kmc_step"][j]
            ps_prev, ig_prev = base["procstat"][j], base["integ_rates"][j]
            tof_last = prev_tof[j]
            tofs, occs, dts, step_ts = [], [], [], []
            for i in range(samples):
                k = min(i, len(rec["kmc_time"]) - 1)
                t, s = float(rec["kmc_time"][k]), int(rec["kmc_step"][k])
                ps = rec["procstat"][k].astype(np.float64)
                ig = rec["integ_rates"][k]
                if s != s_prev:
                    tof_last = interval_tof(tof_matrix, ps - ps_prev, ig - ig_prev, t - t_prev, ncells, tof_method)
                dts.append(t - t_prev)
                step_ts.append(float(rec["kmc_time_step"][k]))
                occs.append(list(rec["occupation"][k].flatten()))
                tofs.append(tof_last.flatten())
                t_prev, s_prev, ps_prev, ig_prev = t, s, ps, ig
            out.append((tofs, occs, dts, step_ts, t_prev, s_prev, ps_prev, ig_prev, tof_last,
                        int(rec["status"][-1])))
        return out

    def _baseline(self, b, tofm, ncells, tof_method):
        """The state of every replica at the get_atoms call after the
        initial steps, and the TOF that call reports (over the steps from
        the start; zero when there were none)."""
        t, k, _ = b.scalars_all()
        ps, ig = b.counters()
        ps = ps.astype(np.float64)
        tof = [interval_tof(tofm, ps[j], ig[j], t[j], ncells, tof_method) if k[j] != 0
               else np.zeros(tofm.shape[0]) for j in range(b.n)]
        return {"kmc_time": t, "kmc_step": k, "procstat": ps, "integ_rates": ig}, tof

    # --- the two drivers -----------------------------------------------------
    def run(self, init_steps=1e8, sample_steps=1e8, cores=4, samples=1, random_seed=None):
        """Scan the grid: init_steps, then get_std_sampled_data(samples,
        sample_steps) per point, one line each in <runner_name>.dat.  `cores`
        is accepted for compatibility; the GPU batch replaces the process pool.
        Returns {input_line: data line}."""
        seed = self.settings.random_seed if random_seed is None else random_seed
        todo = self._claim(self.points())
        tofm = tof_matrix(self.settings, self.compiled)
        ncells = int(np.prod(self.size))
        interval = int(int(sample_steps) / samples)
        results = OrderedDict()
        for chunk in self._chunks(todo):
            m, b = self._batch(chunk, seed)
            try:
                n = len(chunk)
                if int(init_steps) > 0:
                    b.run(int(init_steps))
                base, tof0 = self._baseline(b, tofm, ncells, "integ")
                t0, k0 = base["kmc_time"].copy(), base["kmc_step"].copy()
                res = self._sampled(b, np.arange(n, dtype=np.int32), base, samples, interval,
                                    "integ", tofm, ncells, tof0)
                lines = []
                for j, pt in enumerate(chunk):
                    tofs, occs, dts, step_ts, t_end, k_end, _, _, _, status = res[j]
                    if status != 0:
                        lines.append("# %s: replica stopped (status %d)\n" % (self.input_line(pt), status))
                    else:
                        lines.append(std_sampled_output(
                            adjustable_values(self.point_parameters(pt)), tofs, occs, dts, step_ts,
                            t_end - t0[j], t_end, k_end - k0[j], self.std_header()))
                    results[self.input_line(pt)] = lines[-1]
                self._write(lines, self.point_parameters(chunk[0]))
            finally:
                b.close()
        return results

    def run_steady_state(self, batch_size=1000000, L=4, alpha=0.05, bias_threshold=0.15,
                         tof_method="integ", warm_up=20, check_frequency=10, random_seed=None,
                         init_steps=0, max_batches=100000):
        """sample_steady_state for every grid point at once: each batch is
        get_std_sampled_data(100, batch_size) of the replicas not yet steady;
        a replica leaves when its EWMA test passes.  Writes and returns the
        'str' output of sample_steady_state per point."""
        seed = self.settings.random_seed if random_seed is None else random_seed
        todo = self._claim(self.points())
        tofm = tof_matrix(self.settings, self.compiled)
        ncells = int(np.prod(self.size))
        interval = int(batch_size / 100)
        header = self.std_header()
        results = OrderedDict()
        for chunk in self._chunks(todo):
            m, b = self._batch(chunk, seed)
            try:
                if int(init_steps) > 0:
                    b.run(int(init_steps))
                base, last_tof = self._baseline(b, tofm, ncells, tof_method)
                params = [adjustable_values(self.point_parameters(pt)) for pt in chunk]
                dets = [SteadyStateMonitor(L, alpha, bias_threshold, warm_up, check_frequency) for _ in chunk]
                stopped = {}
                live = list(range(len(chunk)))
                for _ in range(max_batches):
                    if not live:
                        break
                    sub = {k: v[live] for k, v in base.items()}
                    res = self._sampled(b, np.array(live, np.int32), sub, 100, interval, tof_method,
                                        tofm, ncells, [last_tof[r] for r in live])
                    still = []
                    for j, r in enumerate(live):
                        tofs, occs, dts, step_ts, t_end, k_end, ps_end, ig_end, tof_end, status = res[j]
                        data = std_sampled_output(params[r], tofs, occs, dts, step_ts, t_end - base["kmc_time"][r],
                                                  t_end, k_end - base["kmc_step"][r], header, "dict")
                        base["kmc_time"][r], base["kmc_step"][r] = t_end, k_end
                        base["procstat"][r], base["integ_rates"][r] = ps_end, ig_end
                        last_tof[r] = tof_end
                        if status != 0:
                            stopped[r] = status
                        elif not dets[r].add(data):
                            still.append(r)
                    live = still
                lines = []
                for r, pt in enumerate(chunk):
                    if r in stopped:
                        lines.append("# %s: replica stopped (status %d)\n" % (self.input_line(pt), stopped[r]))
                    elif not dets[r].done:
                        lines.append("# %s: no steady state after %d batches\n" % (self.input_line(pt), max_batches))
                    else:
                        lines.append(dets[r].summary(header, "str"))
                    results[self.input_line(pt)] = lines[-1]
                self._write(lines, self.point_parameters(chunk[0]))
            finally:
                b.close()
        return results
