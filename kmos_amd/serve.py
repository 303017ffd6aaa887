"""Continuous-batching kMC server: parameter-sweep requests on one GPU.

A request is one kMC trajectory: a model, a lattice size, its rate
constants, a seed, a step budget and a sampling interval (the quantities a
``KMC_Model`` user sets before ``get_std_sampled_data``, kmos/run/__init__.py:
865-966).  The server keeps every admitted trajectory resident in HBM as a
*slot* of a *page* (an engine ``Batch`` of one model holding a number of
slots of every lattice size) and advances every page a few sampling
intervals at a time, each page at its own pace:

  loop:  admit arrived requests into free slots of their size (new pages
         while the HBM budget allows), initialise them on the device
         -> every idle page with active slots queues its next launch: up to
            `intervals_per_launch` intervals of every active slot, with the
            sample records written by the kernel itself at each interval's
            end (one interval and a record gather when the page's slots
            sample at different intervals); each page has its own stream,
            so pages overlap on the GPU
         -> pages whose launch finished hand over their records; finished
            slots retire and their requests complete now

Requests that arrive while others run join at the next tick, and a slot
freed by a finished request is reused at once: this is continuous batching,
not a queue in front of a fixed batch.  Every trajectory keeps its own
uniform stream (keyed by its seed), so its samples do not depend on what
else is in flight; ``run_alone`` and the tests check that bit for bit.

    srv = Server(hbm_budget=64 << 30)
    results = srv.serve(poisson_stream(rate=100.0, n=500, seed=1))
    print(latency_summary(results))
"""
from __future__ import annotations

import dataclasses
import time
from typing import Dict, List, Optional, Sequence, Tuple

import numpy as np

from . import engine
from . import workloads as W
from .compiler import CompiledModel, compile_model


@dataclasses.dataclass
class Request:
    rid: int
    model: str                   # key of Server.models
    size: Tuple[int, int, int]   # cells
    rates: np.ndarray            # [P] rate constants (s^-1)
    seed: int
    steps: int                   # step budget
    sample_every: int = 1000     # steps between samples
    config: Optional[np.ndarray] = None  # initial lattice [V] (species ids), None = defaults
    arrival: float = 0.0         # seconds after the start of the stream


@dataclasses.dataclass
class Result:
    rid: int
    status: int
    steps_done: int
    samples: Dict[str, np.ndarray]   # per sample: kmc_time, kmc_step, occupation, integ_rates, procstat
    arrival: float = 0.0
    admitted: float = 0.0
    finished: float = 0.0

    @property
    def latency(self) -> float:
        return self.finished - self.arrival


class _Page:
    """One engine batch of one model: `per_size` slots of each lattice size,
    so that a single launch advances every admitted trajectory of the model
    whatever its size (one stream per model keeps the pages' kernels on
    separate hardware queues)."""

    def __init__(self, model: engine.Model, sizes, per_size: int, device: int):
        layout = [list(sz) for sz in sizes for _ in range(per_size)]
        self.batch = engine.Batch(model, layout, device=device)
        self.slots = len(layout)
        self.free: Dict[tuple, List[int]] = {}
        for i, sz in enumerate(layout):
            self.free.setdefault(tuple(sz), []).append(i)
        for v in self.free.values():
            v.reverse()
        self.active: Dict[int, "_Flight"] = {}
        self.bytes = model.footprint(layout)

    def has_room(self, size) -> bool:
        return bool(self.free.get(size))

    def n_free(self) -> int:
        return sum(len(v) for v in self.free.values())


class _Flight:
    def __init__(self, req: Request, admitted: float):
        self.req = req
        self.left = int(req.steps)
        self.admitted = admitted
        self.records: List[dict] = []
        self.status = 0


class Server:
    """Continuous-batching server over the engine.  ``models`` maps a model
    key to a CompiledModel (default: the quickstart, ZGB and RuO2 models)."""

    def __init__(self, models: Optional[Dict[str, CompiledModel]] = None, hbm_budget: int = 64 << 30,
                 sizes=None, slots_per_size: int = 64, device: int = 0, intervals_per_launch: int = 4):
        if models is None:
            models = {"quickstart": compile_model(W.quickstart_project(), "local_smart"),
                      "zgb": compile_model(W.zgb_project(), "local_smart"),
                      "ruo2": compile_model(W.ruo2_project(), "local_smart")}
        self.compiled = models
        self.models = {k: engine.Model(cm) for k, cm in models.items()}
        self.hbm_budget = int(hbm_budget)
        # the lattice sizes a page holds slots for (others get pages of their own)
        self.sizes = [tuple(int(v) for v in (list(sz) + [1, 1])[:3]) for sz in
                      (sizes if sizes is not None else [(n, n, 1) for n in SIZES])]
        self.per_size = int(slots_per_size)
        # sampling intervals one launch covers (records made in the kernel):
        # more means fewer host round trips, fewer means quicker admission
        self.intervals_per_launch = int(intervals_per_launch)
        self.device = device
        self.pages: Dict[str, List[_Page]] = {}
        self.bytes_in_use = 0
        self.stats = {"ticks": 0, "pages": 0, "peak_active": 0, "peak_bytes": 0}

    # ------------------------------------------------------------------
    def _try_admit(self, req: Request, now: float, busy=()) -> bool:
        size = tuple(int(v) for v in (list(req.size) + [1, 1])[:3])
        pages = self.pages.setdefault(req.model, [])
        page = next((p for p in pages if p.has_room(size) and id(p) not in busy), None)
        if page is None and any(p.has_room(size) for p in pages):
            return False        # a page with room is mid-interval: join it next
        if page is None:
            sizes = self.sizes if size in self.sizes else [size]
            per = self.per_size
            m = self.models[req.model]
            while per >= 1 and self.bytes_in_use + m.footprint([list(z) for z in sizes] * per) > self.hbm_budget:
                per //= 2
            if per < 1:
                return False
            page = _Page(m, sizes, per, self.device)
            pages.append(page)
            self.bytes_in_use += page.bytes
            self.stats["pages"] += 1
            self.stats["peak_bytes"] = max(self.stats["peak_bytes"], self.bytes_in_use)
        slot = page.free[size].pop()
        b = page.batch
        P = self.compiled[req.model].nr_of_proc
        rates = np.asarray(req.rates, np.float64).reshape(1, P)
        b.init([req.seed & 0xFFFFFFFF], [0], r0=slot,
               configs=None if req.config is None else [np.asarray(req.config)])
        b.set_rates(rates, r0=slot)
        page.active[slot] = _Flight(req, now)
        return True

    def _release_idle_pages(self):
        for key, pages in self.pages.items():
            keep = []
            for p in pages:
                if not p.active and p.n_free() == p.slots:
                    p.batch.close()
                    self.bytes_in_use -= p.bytes
                else:
                    keep.append(p)
            self.pages[key] = keep

    # ------------------------------------------------------------------
    def serve(self, requests: Sequence[Request], realtime: bool = True) -> List[Result]:
        """Serve ``requests`` (sorted by arrival).  With realtime=True the
        arrival times are honoured against the wall clock; otherwise all
        requests are present at the start.  Returns the results in request
        order."""
        pending = sorted(requests, key=lambda r: r.arrival)
        waiting: List[Request] = []
        results: Dict[int, Result] = {}
        inflight: Dict[int, tuple] = {}      # id(page) -> (page, ids, wait)
        t0 = time.perf_counter()
        now = lambda: time.perf_counter() - t0  # noqa: E731
        i = 0
        while i < len(pending) or waiting or inflight or \
                any(p.active for ps in self.pages.values() for p in ps):
            t = now()
            while i < len(pending) and (not realtime or pending[i].arrival <= t):
                waiting.append(pending[i])
                i += 1
            # admission: first come first served, skipping requests whose
            # pool has no room while others still fit
            still = []
            for req in waiting:
                if not self._try_admit(req, now() if realtime else 0.0, inflight):
                    still.append(req)
            if still and len(still) == len(waiting) and not inflight:
                self._release_idle_pages()
                still = [r for r in still if not self._try_admit(r, now() if realtime else 0.0)]
            waiting = still
            # every page with active slots and no run in flight starts its
            # next sampling interval; pages progress independently and
            # overlap on the device (one stream each)
            for ps in self.pages.values():
                for p in ps:
                    if p.active and id(p) not in inflight:
                        ids = np.array(sorted(p.active), np.int32)
                        every = p.active[int(ids[0])].req.sample_every
                        if all(p.active[int(s)].req.sample_every == every for s in ids):
                            # several intervals per launch, sampled inside the kernel
                            nint = self.intervals_per_launch
                            budget = np.array([min(nint * every, p.active[s].left) for s in ids], np.int64)
                            wait = p.batch.run_recorded(int(budget.max()), ids, budget, every, nint)
                            inflight[id(p)] = (p, ids, wait, True)
                        else:
                            budget = np.array([min(p.active[s].req.sample_every, p.active[s].left) for s in ids],
                                              np.int64)
                            wait = p.batch.run_sampled(int(budget.max()), ids, budget)
                            inflight[id(p)] = (p, ids, wait, False)
                        self.stats["ticks"] += 1
            active = sum(len(p.active) for ps in self.pages.values() for p in ps)
            self.stats["peak_active"] = max(self.stats["peak_active"], active)
            if not inflight:
                if i < len(pending) and realtime:
                    time.sleep(max(0.0, min(pending[i].arrival - now(), 0.01)))
                elif waiting:
                    raise RuntimeError("request %d does not fit the HBM budget" % waiting[0].rid)
                continue
            finished = [k for k, v in inflight.items() if v[0].batch.idle()]
            if not finished:
                time.sleep(50e-6)
                continue
            for k in finished:
                p, ids, wait, recorded = inflight.pop(k)
                _, rec = wait()
                tdone = now()
                for j, s in enumerate(ids):
                    f = p.active[s]
                    rj = rec[j] if recorded else {key: v[j:j + 1] for key, v in rec.items()}
                    nrec = len(rj["kmc_step"])
                    ran = int(rj["kmc_step"][-1]) - (f.req.steps - f.left)
                    f.left -= ran
                    f.status = int(rj["status"][-1])
                    if ran > 0 or not f.records:
                        for q in range(nrec):
                            f.records.append({key: v[q] for key, v in rj.items()})
                    if f.left <= 0 or f.status != 0:
                        results[f.req.rid] = _finish(f, tdone)
                        del p.active[s]
                        p.free[tuple(int(v) for v in p.batch.sizes[s])].append(int(s))
        return [results[r.rid] for r in requests]

    def run_alone(self, req: Request) -> Result:
        """The same request served on its own (a fresh page, nothing else in
        flight)."""
        solo = Server(self.compiled, hbm_budget=self.hbm_budget, sizes=[], slots_per_size=1,
                      device=self.device, intervals_per_launch=self.intervals_per_launch)
        try:
            r = dataclasses.replace(req, arrival=0.0)
            return solo.serve([r], realtime=False)[0]
        finally:
            solo.close()

    def close(self):
        for ps in self.pages.values():
            for p in ps:
                p.batch.close()
        self.pages = {}
        self.bytes_in_use = 0


def _finish(f: _Flight, t: float) -> Result:
    keys = f.records[0].keys()
    samples = {k: np.stack([r[k] for r in f.records]) for k in keys}
    return Result(f.req.rid, f.status, int(f.req.steps - f.left), samples,
                  f.req.arrival, f.admitted, t)


# ---------------------------------------------------------------------------
# request streams
# ---------------------------------------------------------------------------

SIZES = (20, 64, 128, 256)


def _rates_for(model: str, pt, g) -> np.ndarray:
    P = len(pt.processes)
    if model == "quickstart":
        return np.array([100.0, 100.0])
    if model == "zgb":
        y = g.uniform(0.35, 0.60)
        return np.array([y if p.name == "CO_adsorption" else 1e10 if p.name.startswith("CO_oxidation")
                         else (1.0 - y) / 2.0 for p in pt.processes])
    return 10.0 ** g.uniform(-2.0, 9.0, P)


def poisson_stream(rate: float, n: int, seed: int = 0, models=("quickstart", "zgb", "ruo2"),
                   sizes=SIZES, steps_range=(1e4, 1e6), sample_every: int = 1000,
                   coverage=(0.3, 0.3)) -> List[Request]:
    """n requests with exponential inter-arrival times (mean 1/rate s), each
    drawn from the quickstart / ZGB / RuO2 workloads: a lattice side from
    ``sizes``, a step budget log-uniform in ``steps_range``, rate constants as
    in those workloads, a per-request seed, RuO2 starting from a random
    coverage, the others empty."""
    g = np.random.Generator(np.random.Philox(key=seed))
    projects = {"quickstart": W.quickstart_project(), "zgb": W.zgb_project(), "ruo2": W.ruo2_project()}
    t = 0.0
    out = []
    for rid in range(n):
        t += g.exponential(1.0 / rate)
        model = models[int(g.integers(len(models)))]
        side = int(sizes[int(g.integers(len(sizes)))])
        steps = int(10.0 ** g.uniform(np.log10(steps_range[0]), np.log10(steps_range[1])))
        pt = projects[model]
        config = None
        if model == "ruo2":
            ids = pt.species_ids()
            u = g.random(side * side * 2)
            config = np.full(u.size, ids["empty"], np.int8)
            config[u < coverage[0]] = ids["co"]
            config[(u >= coverage[0]) & (u < coverage[0] + coverage[1])] = ids["oxygen"]
        out.append(Request(rid, model, (side, side, 1), _rates_for(model, pt, g),
                           int(g.integers(1, 2 ** 31)), steps, sample_every, config, t))
    return out


def latency_summary(results: Sequence[Result], wall: Optional[float] = None) -> dict:
    lat = np.array([r.latency for r in results])
    steps = int(sum(r.steps_done for r in results))
    span = wall if wall is not None else max(r.finished for r in results) - min(r.arrival for r in results)
    return {"requests": len(results), "p50_latency_s": float(np.percentile(lat, 50)),
            "p99_latency_s": float(np.percentile(lat, 99)), "steps": steps,
            "steps_per_s": steps / span if span > 0 else 0.0,
            "absorbed": int(sum(r.status == 1 for r in results))}
