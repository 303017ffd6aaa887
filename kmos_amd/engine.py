"""Python handles over the C ABI: a loaded model and a batch of replicas.

``Batch`` holds many independent replicas of one model on one GPU; every
method maps to one entry point of include/kmcg.h.  Arrays go in and out as
numpy arrays in the reference's numbering (processes 1..P, sites 1..V,
species 0..S-1).
"""
from __future__ import annotations

import ctypes
from typing import Optional, Sequence

import numpy as np

from . import _lib
from .compiler import CompiledModel

_vp = ctypes.c_void_p


def _ptr(a: Optional[np.ndarray]):
    return None if a is None else a.ctypes.data_as(_vp)


class Model:
    """A compiled model loaded into the engine."""

    def __init__(self, compiled: CompiledModel):
        self.compiled = compiled
        self.blob = np.ascontiguousarray(compiled.to_blob(), dtype=np.int32)
        lib = _lib.load()
        h = _vp()
        _lib.check(lib.kmcg_model_load(_ptr(self.blob), self.blob.size, ctypes.byref(h)))
        self.h = h
        info = np.zeros(5, np.int32)
        _lib.check(lib.kmcg_model_info(self.h, _ptr(info)))
        self.nr_of_proc, self.nr_of_species, self.spuck, self.backend, self.dimension = \
            (int(v) for v in info)
        # the species id of an unpopulated site: -1 for one layer (never
        # stored), nr_of_species for several layers (kmos' null_species)
        self.null_species = int(compiled.null_species)
        li = ctypes.c_int32(0)
        _lib.check(lib.kmcg_model_forms(self.h, ctypes.byref(li)))
        # the step kernel the library launches for this model's batches
        self.step_kernel = "kmc_steps_kernel_li" if li.value else "kmc_steps_kernel"

    def __del__(self):
        if getattr(self, "h", None) is not None and _lib._lib is not None:
            _lib._lib.kmcg_model_free(self.h)
            self.h = None

    def footprint(self, sizes: Sequence[Sequence[int]]) -> int:
        sz = _sizes(sizes)
        out = ctypes.c_int64(0)
        _lib.check(_lib.load().kmcg_batch_footprint(self.h, len(sz), _ptr(sz), ctypes.byref(out)))
        return out.value


def _sizes(sizes) -> np.ndarray:
    sz = np.zeros((len(sizes), 3), np.int32)
    for i, s in enumerate(sizes):
        s = list(s) + [1] * (3 - len(s))
        sz[i] = s[:3]
    return sz


class Batch:
    """Replicas of one model resident on one GPU."""

    def __init__(self, model: Model, sizes: Sequence[Sequence[int]], device: int = 0):
        self.model = model
        self.sizes = _sizes(sizes)
        self.n = len(self.sizes)
        self.P = model.nr_of_proc
        self.S = model.nr_of_species
        # largest species id a lattice site may hold (the null species of
        # a multi-layer model included)
        self.max_species = self.S if model.null_species >= 0 else self.S - 1
        self.spuck = model.spuck
        cells = self.sizes.copy()
        if model.dimension < 3:
            cells[:, 2] = 1
        if model.dimension < 2:
            cells[:, 1] = 1
        self.volumes = (cells.prod(axis=1) * self.spuck).astype(np.int64)
        h = _vp()
        _lib.check(_lib.load().kmcg_batch_create(model.h, self.n, _ptr(self.sizes), device,
                                                 ctypes.byref(h)))
        self.h = h

    def close(self):
        if getattr(self, "h", None) is not None and _lib._lib is not None:
            _lib._lib.kmcg_batch_free(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def _species_array(self, configs, r0: int, n: int) -> np.ndarray:
        """Lattices of replicas r0..r0+n-1 as one int8 array, after checking
        that every entry is a species id 0..S-1, or S (null) for a
        multi-layer model (before the cast, so that
        an out-of-range id cannot wrap into a valid one) and that the
        sizes match the replicas' volumes."""
        if r0 < 0 or n < 0 or r0 + n > self.n:
            raise ValueError("replicas %d..%d outside the batch of %d" % (r0, r0 + n - 1, self.n))
        if n != len(configs):
            raise ValueError("expected %d lattices, got %d" % (n, len(configs)))
        parts = [np.asarray(c).reshape(-1) for c in configs]
        flat = np.concatenate(parts) if parts else np.zeros(0, np.int64)
        if flat.size and not np.issubdtype(flat.dtype, np.integer):
            raise ValueError("lattices must hold integer species ids")
        for i, c in enumerate(parts):
            if c.size != int(self.volumes[r0 + i]):
                raise ValueError("lattice of replica %d has %d sites, expected %d"
                                 % (r0 + i, c.size, int(self.volumes[r0 + i])))
        if flat.size and (flat.min() < 0 or flat.max() > self.max_species):
            raise ValueError("species ids must lie in 0..%d, got %d..%d"
                             % (self.max_species, int(flat.min()), int(flat.max())))
        return np.ascontiguousarray(flat, dtype=np.int8)

    # ---- set-up -----------------------------------------------------------
    def init(self, seeds, streams=None, r0: int = 0, configs=None):
        """initialize_state for replicas r0.. (one per seed), from the
        default species or from the given lattices."""
        seeds = np.ascontiguousarray(np.atleast_1d(seeds), dtype=np.uint32)
        n = seeds.size
        st = None if streams is None else np.ascontiguousarray(np.atleast_1d(streams), dtype=np.uint32)
        cf = None
        if configs is not None:
            if isinstance(configs, np.ndarray) and configs.ndim == 1:
                configs = [configs]
            cf = self._species_array(configs, r0, n)
        _lib.check(_lib.load().kmcg_init(self.h, r0, n, _ptr(seeds), _ptr(st), _ptr(cf)))

    def set_rates(self, rates, r0: int = 0):
        r = np.ascontiguousarray(np.atleast_2d(rates), dtype=np.float64)
        if r.ndim != 2 or r.shape[1] != self.P:
            raise ValueError("rates: expected [n][%d] values, got shape %s" % (self.P, r.shape))
        _lib.check(_lib.load().kmcg_set_rates(self.h, r0, r.shape[0], _ptr(r)))

    def get_rates(self, r: int) -> np.ndarray:
        out = np.zeros(self.P, np.float64)
        _lib.check(_lib.load().kmcg_get_rates(self.h, r, _ptr(out)))
        return out

    def set_config(self, configs, r0: int = 0):
        """Set the lattice of replicas r0, r0+1, ... (one array per replica,
        or a single array for replica r0) and rebuild their bookkeeping."""
        if isinstance(configs, np.ndarray) and configs.ndim == 1:
            configs = [configs]
        n = len(configs)
        flat = self._species_array(configs, r0, n)
        _lib.check(_lib.load().kmcg_set_config(self.h, r0, n, _ptr(flat)))

    def set_launch(self, lattice_in_lds: int = -1, waves_per_group: int = 0):
        _lib.check(_lib.load().kmcg_set_launch(self.h, lattice_in_lds, waves_per_group))

    def set_overlap(self, mode: int = -1):
        """Step kernel form: 1 overlapped select, 0 plain, -1 automatic."""
        _lib.check(_lib.load().kmcg_set_overlap(self.h, int(mode)))

    def check_lists(self, r0: int = 0, n: Optional[int] = None) -> np.ndarray:
        """Violations per replica of the list invariant and of the
        registration against the conditions (kmcg_check_lists); all 0 for a
        consistent state."""
        n = self.n - r0 if n is None else n
        out = np.zeros(n, np.int32)
        _lib.check(_lib.load().kmcg_check_lists(self.h, r0, n, _ptr(out)))
        return out

    def set_graph(self, enable: bool = True):
        """Replay runs through captured HIP graphs (kmcg_set_graph)."""
        _lib.check(_lib.load().kmcg_set_graph(self.h, 1 if enable else 0))

    # ---- stepping ---------------------------------------------------------
    def run(self, steps: int, replicas=None, budget=None, sync: bool = True):
        """do_kmc_steps(steps) on the given replicas (all by default)."""
        ids = np.arange(self.n, dtype=np.int32) if replicas is None else \
            np.ascontiguousarray(replicas, dtype=np.int32)
        bud = None if budget is None else np.array(budget, dtype=np.int64)
        self._ids = ids
        self._bud = bud          # written back by the library at the next sync
        _lib.check(_lib.load().kmcg_run(self.h, _ptr(ids), ids.size, int(steps), _ptr(bud)))
        if sync:
            self.sync()
        return bud

    def idle(self) -> bool:
        """True when no queued work is left on the batch's stream."""
        d = ctypes.c_int32(0)
        _lib.check(_lib.load().kmcg_query(self.h, ctypes.byref(d)))
        return bool(d.value)

    def sync(self):
        _lib.check(_lib.load().kmcg_sync(self.h))

    def last_run_ms(self) -> float:
        v = ctypes.c_float(0)
        _lib.check(_lib.load().kmcg_last_run_ms(self.h, ctypes.byref(v)))
        return float(v.value)

    def phase_profile(self, r0: int = 0, n: Optional[int] = None) -> np.ndarray:
        """Step-loop phase counters of replicas r0.. ([n][8]: ticks in
        select, site read, post, replay, rest; steps), cleared after the
        read.  Needs the library built with KMCG_PHASES (loaded when the
        environment sets KMCG_PHASES=1)."""
        n = self.n - r0 if n is None else n
        out = np.zeros((n, 8), np.uint64)
        _lib.check(_lib.load().kmcg_phase_profile(self.h, r0, n, _ptr(out)))
        return out

    def get_next_step(self, r: int):
        p, s = ctypes.c_int32(0), ctypes.c_int32(0)
        _lib.check(_lib.load().kmcg_get_next_step(self.h, r, ctypes.byref(p), ctypes.byref(s)))
        return p.value, s.value

    def run_proc_nr(self, r: int, proc: int, site: int):
        _lib.check(_lib.load().kmcg_run_proc_nr(self.h, r, proc, site))

    def touchup(self, r: int, x: int, y: int, z: int, n: int):
        _lib.check(_lib.load().kmcg_touchup(self.h, r, x, y, z, n))

    def replace_species(self, r: int, site: int, old: int, new: int):
        _lib.check(_lib.load().kmcg_replace_species(self.h, r, site, old, new))

    # ---- state ------------------------------------------------------------
    def scalars(self, r: int):
        """(kmc_time, kmc_time_step, kmc_step, rng_counter, status)"""
        t = np.zeros(2, np.float64)
        s = np.zeros(2, np.int64)
        st = ctypes.c_int32(0)
        _lib.check(_lib.load().kmcg_get_scalars(self.h, r, _ptr(t), _ptr(s), ctypes.byref(st)))
        return float(t[0]), float(t[1]), int(s[0]), int(s[1]), st.value

    def scalars_all(self):
        t = np.zeros(self.n, np.float64)
        k = np.zeros(self.n, np.int64)
        st = np.zeros(self.n, np.int32)
        _lib.check(_lib.load().kmcg_get_scalars_all(self.h, _ptr(t), _ptr(k), _ptr(st)))
        return t, k, st

    def set_kmc_time(self, r: int, t: float):
        _lib.check(_lib.load().kmcg_set_kmc_time(self.h, r, float(t)))

    def nr_of_sites(self, r: int) -> np.ndarray:
        out = np.zeros(self.P, np.int32)
        _lib.check(_lib.load().kmcg_get_nr_of_sites(self.h, r, _ptr(out)))
        return out

    def procstat(self, r: int) -> np.ndarray:
        out = np.zeros(self.P, np.int64)
        _lib.check(_lib.load().kmcg_get_procstat(self.h, r, _ptr(out)))
        return out

    def integ_rates(self, r: int) -> np.ndarray:
        out = np.zeros(self.P, np.float64)
        _lib.check(_lib.load().kmcg_get_integ_rates(self.h, r, _ptr(out)))
        return out

    def update_integ_rate(self, r: int):
        _lib.check(_lib.load().kmcg_update_integ_rate(self.h, r))

    def accum_rates(self, r: int) -> np.ndarray:
        out = np.zeros(self.P, np.float64)
        _lib.check(_lib.load().kmcg_get_accum_rates(self.h, r, _ptr(out)))
        return out

    def lattice(self, r: int) -> np.ndarray:
        out = np.zeros(int(self.volumes[r]), np.int8)
        _lib.check(_lib.load().kmcg_get_lattice(self.h, r, _ptr(out)))
        return out

    def avail(self, r: int):
        V = int(self.volumes[r])
        sites = np.zeros((self.P, V), np.int32)
        back = np.zeros((self.P, V), np.int32)
        _lib.check(_lib.load().kmcg_get_avail(self.h, r, _ptr(sites), _ptr(back)))
        return sites, back

    def get_state(self, r: int) -> dict:
        """Every field of replica r that a .reload snapshot holds."""
        t, dt, step, ctr, status = self.scalars(r)
        sites, back = self.avail(r)
        return {"kmc_time": t, "kmc_time_step": dt, "kmc_step": step, "rng_counter": ctr,
                "status": status, "lattice": self.lattice(r), "nr_of_sites": self.nr_of_sites(r),
                "procstat": self.procstat(r), "integ_rates": self.integ_rates(r),
                "rates": self.get_rates(r), "avail_sites": sites, "avail_sites_back": back}

    def set_state(self, r: int, state: dict):
        """Overwrite replica r with a state as get_state returns it (integ_rates,
        kmc_time_step and rng_counter optional: 0, 0 and kmc_step); rates are
        applied when present."""
        V, P = int(self.volumes[r]), self.P

        def arr(key, dtype, shape):
            a = np.ascontiguousarray(state[key], dtype=dtype)
            if a.shape != shape:
                raise ValueError("state[%r] has shape %s, expected %s" % (key, a.shape, shape))
            return a
        lat = np.asarray(state["lattice"])
        if lat.shape != (V,) or not np.issubdtype(lat.dtype, np.integer) or \
                (V and (lat.min() < 0 or lat.max() > self.max_species)):
            raise ValueError("state['lattice'] must hold %d species ids in 0..%d" % (V, self.max_species))
        lat = np.ascontiguousarray(lat, dtype=np.int8)
        nr = arr("nr_of_sites", np.int32, (P,))
        ps = arr("procstat", np.int64, (P,))
        sites = arr("avail_sites", np.int32, (P, V))
        back = arr("avail_sites_back", np.int32, (P, V))
        ig = arr("integ_rates", np.float64, (P,)) if state.get("integ_rates") is not None else None
        step = int(state["kmc_step"])
        ctr = state.get("rng_counter")
        t2 = np.array([state["kmc_time"], state.get("kmc_time_step", 0.0)], np.float64)
        s2 = np.array([step, step if ctr is None else int(ctr)], np.int64)
        rates = arr("rates", np.float64, (P,)) if state.get("rates") is not None else None
        if rates is not None and not (np.all(np.isfinite(rates)) and np.all(rates >= 0.0)):
            raise ValueError("state['rates'] must be finite and >= 0")
        _lib.check(_lib.load().kmcg_set_state(self.h, r, _ptr(t2), _ptr(s2), _ptr(lat), _ptr(nr),
                                              _ptr(ps), _ptr(ig), _ptr(sites), _ptr(back)))
        if rates is not None:
            self.set_rates(rates[None, :], r0=r)

    # ---- displacement tracking ----------------------------------------------
    def msd_init(self, trace_species: int, hops, disp, r0: int = 0, n: Optional[int] = None):
        """Number the traced species' sites of replicas [r0, r0+n) and zero
        their displacements; hops/disp from acf.displacement_table."""
        n = self.n - r0 if n is None else n
        h = np.ascontiguousarray(hops, dtype=np.int32)
        d = np.ascontiguousarray(disp, dtype=np.float64)
        if h.shape != (self.P, 9) or d.shape != (self.P, 3):
            raise ValueError("hop table must be [P][9] and [P][3]")
        _lib.check(_lib.load().kmcg_msd_init(self.h, r0, n, int(trace_species), _ptr(h), _ptr(d)))

    def msd_trajectory(self, nr_of_steps: int):
        _lib.check(_lib.load().kmcg_msd_trajectory(self.h, int(nr_of_steps)))

    def msd_zero(self, r0: int = 0, n: Optional[int] = None):
        n = self.n - r0 if n is None else n
        _lib.check(_lib.load().kmcg_msd_zero(self.h, r0, n))

    def run_displacement(self, steps: int, replicas=None, traj_on: bool = False):
        """do_kmc_steps_displacement: run() with displacement tracking."""
        ids = np.arange(self.n, dtype=np.int32) if replicas is None else \
            np.ascontiguousarray(np.atleast_1d(replicas), dtype=np.int32)
        self._ids = ids
        _lib.check(_lib.load().kmcg_run_displacement(self.h, _ptr(ids), ids.size, int(steps), None,
                                                     1 if traj_on else 0))
        self.sync()

    def msd_state(self, r: int) -> dict:
        """nr_of_ions, site_arr[V+1], id_arr[V+1], displacement[V+1][3] and the
        recorded trajectory events [n][2] (particle, new site) of replica r."""
        V1 = int(self.volumes[r]) + 1
        ni = np.zeros(1, np.int32)
        ne = np.zeros(1, np.int64)
        sa = np.zeros(V1, np.int32)
        ia = np.zeros(V1, np.int32)
        dp = np.zeros((V1, 3), np.float64)
        L = _lib.load()
        _lib.check(L.kmcg_msd_get(self.h, r, _ptr(ni), _ptr(sa), _ptr(ia), _ptr(dp), _ptr(ne), None))
        ev = np.zeros((max(int(ne[0]), 1), 2), np.int32)
        if ne[0]:
            _lib.check(L.kmcg_msd_get(self.h, r, None, None, None, None, None, _ptr(ev)))
        return {"nr_of_ions": int(ni[0]), "site_arr": sa, "id_arr": ia, "displacement": dp,
                "events": ev[: int(ne[0])]}

    # ---- autocorrelation bins -----------------------------------------------
    def acf_alloc(self, nr_of_types: int, t_bin: float, t_f: float, safety_factor: int = 0,
                  extending_factor: int = 0):
        _lib.check(_lib.load().kmcg_acf_alloc(self.h, int(nr_of_types), float(t_bin), float(t_f),
                                              int(safety_factor), int(extending_factor)))

    def acf_set_properties(self, property_acf=None, product_property=None, r0: int = 0):
        """property_acf: list of per-replica arrays [V+1] (from r0), or None;
        product_property: [(T+1), (T+1)] or None."""
        pa = None
        n = 0
        if property_acf is not None:
            pa = np.ascontiguousarray(np.concatenate([np.asarray(a, np.int32) for a in property_acf]))
            n = len(property_acf)
        pp = None if product_property is None else np.ascontiguousarray(product_property, np.float64)
        _lib.check(_lib.load().kmcg_acf_set_properties(self.h, r0, n, _ptr(pa), _ptr(pp)))

    def acf_init(self, trace_species: int, hops, r0: int = 0, n: Optional[int] = None):
        n = self.n - r0 if n is None else n
        h = np.ascontiguousarray(hops, dtype=np.int32)
        _lib.check(_lib.load().kmcg_acf_init(self.h, r0, n, int(trace_species), _ptr(h)))

    def run_acf(self, steps: int, replicas=None, traj_on: bool = False):
        """do_kmc_steps_acf: run() with the ACF bins and particle tracking."""
        ids = np.arange(self.n, dtype=np.int32) if replicas is None else \
            np.ascontiguousarray(np.atleast_1d(replicas), dtype=np.int32)
        self._ids = ids
        _lib.check(_lib.load().kmcg_run_acf(self.h, _ptr(ids), ids.size, int(steps), None,
                                            1 if traj_on else 0))
        self.sync()

    def acf_zero(self, r0: int = 0, n: Optional[int] = None):
        n = self.n - r0 if n is None else n
        _lib.check(_lib.load().kmcg_acf_zero(self.h, r0, n))

    def acf_state(self, r: int) -> dict:
        i5 = np.zeros(5, np.int32)
        d4 = np.zeros(4, np.float64)
        L = _lib.load()
        _lib.check(L.kmcg_acf_get(self.h, r, None, None, _ptr(d4), _ptr(i5), None, None))
        ext = int(i5[2])
        V1 = int(self.volumes[r]) + 1
        cb = np.zeros(ext, np.float64)
        cc = np.zeros(ext, np.int32)
        buf = np.zeros(V1, np.float64)
        po = np.zeros(V1, np.int32)
        _lib.check(L.kmcg_acf_get(self.h, r, _ptr(cb), _ptr(cc), None, None, _ptr(buf), _ptr(po)))
        return {"config_bin": cb, "counter_write_in_bin": cc, "sum_buffer": float(d4[0]), "t_o": float(d4[1]),
                "wrap_count": float(d4[2]), "t_bin": float(d4[3]), "bin_index": int(i5[0]),
                "nr_of_bins": int(i5[1]), "extended_nr_of_bins": ext, "nr_of_ions_before": int(i5[3]),
                "nr_of_ions": int(i5[4]), "buffer_acf": buf, "property_o": po}

    def occupation(self, r0: int = 0, n: Optional[int] = None) -> np.ndarray:
        n = self.n - r0 if n is None else n
        out = np.zeros((n, self.S, self.spuck), np.float64)
        _lib.check(_lib.load().kmcg_get_occupation(self.h, r0, n, _ptr(out)))
        return out

    def _records(self, out):
        S, N, P = self.S, self.spuck, self.P
        return {
            "kmc_time": out[:, 0].copy(),
            "kmc_time_step": out[:, 1].copy(),
            "kmc_step": out[:, 2].astype(np.int64),
            "status": out[:, 3].astype(np.int32),
            "occupation": out[:, 4:4 + S * N].reshape(-1, S, N).copy(),
            "integ_rates": out[:, 4 + S * N:4 + S * N + P].copy(),
            "procstat": out[:, 4 + S * N + P:].astype(np.int64),
        }

    def sample(self, replicas):
        """One record per replica: dict of arrays kmc_time[n], kmc_time_step[n],
        kmc_step[n], status[n], occupation[n, S, spuck], integ_rates[n, P],
        procstat[n, P] (one device gather, one copy)."""
        ids = np.ascontiguousarray(np.atleast_1d(replicas), dtype=np.int32)
        out = np.zeros((ids.size, 4 + self.S * self.spuck + 2 * self.P), np.float64)
        _lib.check(_lib.load().kmcg_sample(self.h, _ptr(ids), ids.size, _ptr(out)))
        return self._records(out)

    def run_sampled(self, steps: int, replicas, budget=None):
        """run() then sample() of the same replicas, queued together and not
        waited for: returns a function that waits and gives (budget, records)."""
        ids = np.ascontiguousarray(np.atleast_1d(replicas), dtype=np.int32)
        bud = None if budget is None else np.array(budget, dtype=np.int64)
        out = np.zeros((ids.size, 4 + self.S * self.spuck + 2 * self.P), np.float64)
        self._ids, self._bud, self._out = ids, bud, out   # written at the next sync
        _lib.check(_lib.load().kmcg_run_sampled(self.h, _ptr(ids), ids.size, int(steps), _ptr(bud), _ptr(out)))

        def wait():
            self.sync()
            return bud, self._records(out)
        return wait

    def run_recorded(self, steps: int, replicas, budget, sample_every: int, max_records: int):
        """run() that records a sample every `sample_every` steps inside the
        kernel; not waited for: returns a function that waits and gives
        (budget, [records of replica i: dict of arrays over its samples])."""
        ids = np.ascontiguousarray(np.atleast_1d(replicas), dtype=np.int32)
        bud = np.array(budget, dtype=np.int64)
        before = bud.copy()
        w = 4 + self.S * self.spuck + 2 * self.P
        out = np.zeros((ids.size, max_records, w), np.float64)
        self._ids, self._bud, self._rec = ids, bud, out   # written at the next sync
        _lib.check(_lib.load().kmcg_run_recorded(self.h, _ptr(ids), ids.size, int(steps), _ptr(bud),
                                                 int(sample_every), int(max_records), _ptr(out)))

        def wait():
            self.sync()
            done = before - bud
            recs = []
            for i in range(ids.size):
                k = max(1, -(-int(done[i]) // sample_every))
                recs.append(self._records(out[i, :k]))
            return bud, recs
        return wait

    def counters(self, r0: int = 0, n: Optional[int] = None):
        n = self.n - r0 if n is None else n
        ps = np.zeros((n, self.P), np.int64)
        ig = np.zeros((n, self.P), np.float64)
        _lib.check(_lib.load().kmcg_get_counters(self.h, r0, n, _ptr(ps), _ptr(ig)))
        return ps, ig
