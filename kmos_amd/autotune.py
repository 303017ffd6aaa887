"""Autotuner for how the engine is launched.

Two choices change the speed of a run but never its results (every replica
keeps its own uniform stream and its own state, whatever shares a
workgroup with it):

* the launch shape of the step kernel for a batch: the lattice in LDS or
  in HBM, and the number of replica waves per workgroup (they share one LDS
  copy of the op program and one compute unit), together with the order of
  the replica list, which decides which replicas share a workgroup and the
  order in which workgroups are dispatched (``tune_launch``);
* the number of sampling intervals the server advances per launch, which
  trades host round trips (throughput) against how soon a freed slot can
  take a new request (latency) (``tune_serving``).

``tune_launch`` times every candidate on a probe batch built by the
caller's factory (the batch that will run is never advanced by tuning),
interleaving the candidates over the repeats so that drift of the probe's
state affects them alike, and returns the fastest by median device time:

    best = autotune.tune_launch(lambda: make_batch(...), steps=2000)
    ids = autotune.apply(batch, best)
    batch.run(nsteps, ids)
"""
from __future__ import annotations

import json
import os
from typing import Callable, Dict, Iterable, List, Optional, Sequence, Tuple

import numpy as np

# (lattice_in_lds, waves_per_group): -1 = lattice in LDS when it fits
SHAPES: Tuple[Tuple[int, int], ...] = ((-1, 1), (-1, 2), (-1, 4), (0, 1), (0, 4))
ORDERS: Tuple[str, ...] = ("given", "size_desc", "interleave")


def replica_order(volumes: Sequence[int], how: str) -> np.ndarray:
    """Replica indices in launch order: as given, largest lattice first
    (stable), or round robin over the lattice sizes from the largest (so
    that the replicas of one workgroup have different sizes)."""
    v = np.asarray(volumes, np.int64)
    if how == "given":
        return np.arange(v.size, dtype=np.int32)
    order = np.argsort(-v, kind="stable").astype(np.int32)
    if how == "size_desc":
        return order
    if how == "interleave":
        classes: Dict[int, List[int]] = {}
        for i in order:
            classes.setdefault(int(v[i]), []).append(int(i))
        lists = [classes[k] for k in sorted(classes, reverse=True)]
        out: List[int] = []
        pos = [0] * len(lists)
        while len(out) < v.size:
            for c, lst in enumerate(lists):
                if pos[c] < len(lst):
                    out.append(lst[pos[c]])
                    pos[c] += 1
        return np.asarray(out, np.int32)
    raise ValueError("unknown replica order %r" % how)


def _pick(times: Dict[tuple, List[float]]) -> Tuple[tuple, List[dict]]:
    table = []
    for c, ts in times.items():
        table.append({"lattice_in_lds": c[0], "waves_per_group": c[1], "order": c[2],
                      "ms": float(np.median(ts)), "runs": [float(t) for t in ts]})
    # fastest median; ties go to the earlier candidate
    best = min(range(len(table)), key=lambda i: (table[i]["ms"], i))
    return tuple(table[best][k] for k in ("lattice_in_lds", "waves_per_group", "order")), table


def _load_cache(path: Optional[str]) -> dict:
    if not path or not os.path.exists(path):
        return {}
    with open(path) as f:
        return json.load(f)


def _store_cache(path: Optional[str], key: str, entry: dict):
    if not path:
        return
    data = _load_cache(path)
    data[key] = entry
    d = os.path.dirname(os.path.abspath(path))
    os.makedirs(d, exist_ok=True)
    tmp = path + ".tmp"
    with open(tmp, "w") as f:
        json.dump(data, f, indent=1, sort_keys=True)
    os.replace(tmp, path)


def tune_launch(make_batch: Callable, steps: int = 2000, warmup: int = 1000, repeats: int = 3,
                shapes: Iterable[Tuple[int, int]] = SHAPES, orders: Iterable[str] = ORDERS,
                cache_path: Optional[str] = None, key: Optional[str] = None) -> dict:
    """Time every (shape, order) candidate for `steps` kMC steps of every
    replica of a probe batch from ``make_batch()`` (closed afterwards) and
    return {"lattice_in_lds", "waves_per_group", "order", "ms", "table"}.
    With ``cache_path`` and ``key``, a stored choice for the key is returned
    without timing and a new one is stored."""
    if cache_path and key:
        hit = _load_cache(cache_path).get(key)
        if hit is not None:
            return dict(hit, cached=True)
    cands = [(int(l), int(w), str(o)) for l, w in shapes for o in orders]
    if not cands or repeats < 1 or steps < 1:
        raise ValueError("nothing to tune")
    b = make_batch()
    try:
        vol = [int(v) for v in b.volumes]
        ids = {o: replica_order(vol, o) for o in set(c[2] for c in cands)}
        if warmup > 0:
            b.run(int(warmup))
        times: Dict[tuple, List[float]] = {c: [] for c in cands}
        for _ in range(repeats):
            for c in cands:
                b.set_launch(c[0], c[1])
                b.run(int(steps), ids[c[2]])
                times[c].append(b.last_run_ms())
    finally:
        b.close()
    best, table = _pick(times)
    ms = min(t["ms"] for t in table)
    entry = {"lattice_in_lds": best[0], "waves_per_group": best[1], "order": best[2], "ms": ms,
             "steps": int(steps), "table": table}
    if cache_path and key:
        _store_cache(cache_path, key, entry)
    return dict(entry, cached=False)


def apply(batch, choice: dict) -> np.ndarray:
    """Set the chosen launch shape on ``batch`` and return its replica
    order (pass it as the replica list of ``run``)."""
    batch.set_launch(int(choice["lattice_in_lds"]), int(choice["waves_per_group"]))
    return replica_order([int(v) for v in batch.volumes], choice["order"])


def tune_serving(make_server: Callable[[int], object], make_requests: Callable[[], list],
                 candidates: Iterable[int] = (1, 2, 4, 8), realtime: bool = True) -> dict:
    """Serve the same request stream with each candidate number of sampling
    intervals per launch (``make_server(k)`` builds the server) and return
    the candidate with the lowest p99 latency (ties: the higher throughput)
    with the summary of every candidate."""
    from .serve import latency_summary
    rows = []
    for k in candidates:
        srv = make_server(int(k))
        try:
            res = srv.serve(make_requests(), realtime=realtime)
        finally:
            srv.close()
        s = latency_summary(res)
        rows.append(dict(s, intervals_per_launch=int(k)))
    if not rows:
        raise ValueError("no candidates")
    best = min(rows, key=lambda r: (r["p99_latency_s"], -r["steps_per_s"]))
    return {"intervals_per_launch": best["intervals_per_launch"], "table": rows}
