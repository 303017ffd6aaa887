#!/usr/bin/env python3
"""Benchmark: aggregate kMC steps/s of the batched engine on one MI355X.

Workload (default): the batched ZGB model, 4096 replicas of 128x128, each
running 1e5 kMC steps from an empty lattice with y ~ U[0.35, 0.60].  One
bench "step" is one launch that advances every replica by
steps_per_replica / K kMC steps, so the K timed launches run the whole
workload (4.1e8 kMC steps); W warm-up launches run on a separate batch.

Printed (one JSON line): whole-job kMC steps/s (value), the plain path (one
replica per launch) on the same model, their ratio, the HBM roofline of
the step kernel from the algorithmic bytes per kMC step (int32 avail_sites
entries read or written, counted by the instrumented oracle) with the LDS
bandwidth beside it (LDS bytes per launch from the committed rocprofv3
counter summary over this run's launch time), the launch shape the
autotuner chose (kmos_amd/autotune.py, trials on a probe batch), and a CPU
baseline (the C oracle on the host cores, one process per core, on a
bounded sample of the same workload).  Two more legs ride along:

  ruo2     the batched RuO2 workload (2048 replicas, 20^2:64^2:128^2 =
           1:2:1, 2e5 steps) on the GPU, next to the reference's own
           generated Fortran (built by oracle/build_ref.py with the injected
           uniform stream) run one replica per host core
  pairwise the lateral-interaction workload (512 replicas of 256^2,
           pairwise lat_int, 5e5 steps) with its roofline and the C oracle
           on the host cores
  serving  a Poisson stream of requests drawn from the quickstart / ZGB /
           RuO2 workloads through the continuous-batching server: p50/p99
           request latency and aggregate steps/s at 100 requests/s, and the
           same at 10 and 1000 requests/s (serving_sweep)
"""
import argparse
import json
import multiprocessing as mp
import os
import sys
import time

import numpy as np

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, REPO)

from kmos_amd import workloads as W  # noqa: E402
from kmos_amd.compiler import compile_model  # noqa: E402

PEAK_HBM_GBS = 8000.0
METRIC = "aggregate kMC steps/s, batched vs plain GPU; p50/p99 request latency"


def measured_traffic(workload):
    """HBM bytes per launch of the step kernel from the rocprofv3 PMC summary
    committed for this workload (FETCH_SIZE + WRITE_SIZE, calibrated), or
    None."""
    path = os.path.join(REPO, "profiles", "kmc_steps_kernel_%s_pmc.md" % workload)
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("traffic_bytes_per_launch:"):
            return float(line.split(":")[1])
    return None


PEAK_LDS_GBS = 256 * 128 * 2.4   # CUs x bytes per clock per CU x GHz


def measured_lds(workload):
    """LDS bytes per launch of the step kernel from the rocprofv3 PMC summary
    committed for this workload (the SQ_INSTS_LDS_*_BANDWIDTH counters), or
    None."""
    path = os.path.join(REPO, "profiles", "kmc_steps_kernel_%s_lds.md" % workload)
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("lds_bytes_per_launch:"):
            return float(line.split(":")[1])
    return None


def lds_figures(workload, avg_launch_s):
    """Achieved LDS bandwidth of the step kernel: the profiled LDS bytes per
    launch over this run's average launch time, against the LDS peak."""
    b = measured_lds(workload)
    if b is None:
        return None
    achieved = b / avg_launch_s / 1e9
    return {"bytes_per_launch": b, "achieved": achieved, "peak": PEAK_LDS_GBS, "unit": "GB/s",
            "frac": achieved / PEAK_LDS_GBS}


def _oracle_worker(args):
    blob, size, seed, stream, rates, steps, count_bytes = args
    from oracle import pyoracle
    o = pyoracle.Oracle(blob, size)
    o.init(seed, stream)
    o.set_rates(rates)
    o.reset_touched()
    t0 = time.perf_counter()
    n = o.run(steps)
    dt = time.perf_counter() - t0
    return n, dt, o.touched


def algorithmic_bytes_per_step(wl, cm, sample=8, steps=20000):
    """Mean int32 avail_sites entries touched per kMC step x 4 B (oracle
    instrumentation), over `sample` replicas of the workload."""
    from oracle import pyoracle
    pyoracle.build()
    blob = cm.to_blob()
    tot_steps, tot_touch = 0, 0
    for r in range(min(sample, wl.n)):
        n, _, touched = _oracle_worker((blob, tuple(wl.sizes[r][:2]), int(wl.seeds[r]),
                                        int(wl.streams[r]), wl.rates[r], steps, True))
        tot_steps += n
        tot_touch += touched
    return 4.0 * tot_touch / max(tot_steps, 1)


def cpu_baseline(wl, cm, seconds=15.0):
    """The C oracle on all host cores (one process per core), each core
    running whole replicas of the workload for its full step budget."""
    from oracle import pyoracle
    pyoracle.build()
    cores = len(os.sched_getaffinity(0))
    cores = min(cores, 16)
    blob = cm.to_blob()
    # size the sample so that it is about `seconds` of CPU work: time one
    # replica first
    n1, dt1, _ = _oracle_worker((blob, tuple(wl.sizes[0][:2]), int(wl.seeds[0]),
                                 int(wl.streams[0]), wl.rates[0], min(wl.steps, 20000), False))
    rate1 = n1 / max(dt1, 1e-9)
    per_replica = wl.steps / rate1
    nrep = int(max(cores, min(wl.n, round(seconds * cores / per_replica))))
    nrep = max(cores, (nrep // cores) * cores)
    jobs = [(blob, tuple(wl.sizes[r % wl.n][:2]), int(wl.seeds[r % wl.n]), int(wl.streams[r % wl.n]),
             wl.rates[r % wl.n], wl.steps, False) for r in range(nrep)]
    ctx = mp.get_context("fork")
    t0 = time.perf_counter()
    with ctx.Pool(cores) as pool:
        res = pool.map(_oracle_worker, jobs, chunksize=1)
    wall = time.perf_counter() - t0
    steps = sum(r[0] for r in res)
    return {"value": steps / wall, "unit": "kMC steps/s", "cores": cores, "kind": "port",
            "sample": "%d replicas x %d steps of %s (C oracle, one process per core, %.1f s wall)"
                      % (nrep, wl.steps, wl.name, wall)}


def _ref_worker(args):
    exe, size, seed, stream, rates, config, steps = args
    import subprocess
    import tempfile
    with tempfile.TemporaryDirectory() as td:
        inp = os.path.join(td, "in.txt")
        with open(inp, "w") as fh:
            fh.write("%d %d 1\n%d %d\n%d 1\n%d\n" % (size[0], size[1], seed, stream, steps,
                                                       0 if config is None else 1))
            fh.write(" ".join(repr(float(r)) for r in rates) + "\n")
            if config is not None:
                fh.write(" ".join(str(int(c)) for c in config) + "\n")
        out = subprocess.run([exe, inp, os.path.join(td, "out.bin")], capture_output=True, text=True,
                             check=True).stdout
    secs = float(out.split("do_kmc_steps_seconds")[1].split()[0])
    return steps, secs


def reference_cpu_baseline(wl, per_core=12):
    """The reference's generated Fortran (oracle/_ref, built from the
    reference's checked-in export), one process per host core, each running
    whole replicas of the workload through the reference's do_kmc_steps."""
    exe = os.path.join(REPO, "oracle", "_ref", "ruo2_%s" % wl.backend, "ref_driver")
    if not os.path.exists(exe):
        return None
    cores = min(len(os.sched_getaffinity(0)), 16)
    nrep = min(wl.n, cores * per_core)
    # an even spread over the size mix
    idx = np.linspace(0, wl.n - 1, nrep).astype(int)
    jobs = [(exe, tuple(int(v) for v in wl.sizes[r][:2]), int(wl.seeds[r]), int(wl.streams[r]),
             wl.rates[r], None if wl.configs is None else wl.configs[r], wl.steps) for r in idx]
    ctx = mp.get_context("fork")
    t0 = time.perf_counter()
    with ctx.Pool(cores) as pool:
        res = pool.map(_ref_worker, jobs, chunksize=1)
    wall = time.perf_counter() - t0
    steps = sum(r[0] for r in res)
    loop = sum(r[1] for r in res)
    # all cores ran concurrently; the rate counts the reference's step loop
    # only (do_kmc_steps wall time), not process start-up or set-up
    return {"value": cores * steps / loop, "unit": "kMC steps/s", "cores": cores, "kind": "reference",
            "sample": "%d replicas (sizes spread over the mix) x %d steps of %s through the reference "
                      "Fortran's do_kmc_steps (flang -O3, injected uniforms), one process per core "
                      "running concurrently; value = cores x steps / summed loop seconds; %.1f s wall"
                      % (nrep, wl.steps, wl.name, wall)}


def run_batched(engine, model, wl, K, warmup, tune=True):
    """Time K launches that together run the whole workload; returns the
    batch, wall seconds, executed steps, per-launch device ms, the kMC steps
    per launch and the launch shape the autotuner chose (None untuned).
    The warm-up launches and the autotuner's trials run on a probe batch of
    the same replicas, so the timed batch starts from the workload's
    initial state."""
    from kmos_amd import autotune
    chunk = max(1, wl.steps // K)
    probe = min(chunk, 1000)
    choice = None
    if tune:
        choice = autotune.tune_launch(lambda: make_batch(engine, model, wl), steps=probe,
                                      warmup=probe * max(warmup, 0), repeats=2)
    elif warmup > 0:
        wb = make_batch(engine, model, wl, np.arange(min(wl.n, 256)))
        for _ in range(warmup):
            wb.run(min(chunk, 2000))
        wb.close()
    b = make_batch(engine, model, wl)
    ids = autotune.apply(b, choice) if choice else None
    kernel_ms = []
    done_before = b.scalars_all()[1].sum()
    b.sync()
    t0 = time.perf_counter()
    for _ in range(K):
        b.run(chunk, ids, sync=False)
        kernel_ms.append(b.last_run_ms())   # waits for the launch, reads its HIP events
    b.sync()
    wall = time.perf_counter() - t0
    executed = int(b.scalars_all()[1].sum() - done_before)
    return b, wall, executed, kernel_ms, chunk, choice


def graph_leg(engine, model, wl, steps_per_run=50, runs=300):
    """Host cost of short runs: one replica advanced `runs` times by
    `steps_per_run` kMC steps, by direct launches and by replays of a
    captured HIP graph (kmcg_set_graph), wall microseconds per run."""
    out = {"steps_per_run": steps_per_run, "runs": runs}
    for mode in ("direct", "graph"):
        b = make_batch(engine, model, wl, [0])
        b.set_graph(mode == "graph")
        ids = np.zeros(1, np.int32)
        for _ in range(5):
            b.run(steps_per_run, ids)
        t0 = time.perf_counter()
        for _ in range(runs):
            b.run(steps_per_run, ids)
        out[mode + "_us_per_run"] = (time.perf_counter() - t0) / runs * 1e6
        out[mode + "_kernel_us"] = b.last_run_ms() * 1e3
        b.close()
    return out


def launch_summary(choice):
    """The autotuner's choice and how it compares with the default shape
    (lattice in LDS when it fits, four replica waves per workgroup, replicas
    in the given order)."""
    if not choice:
        return None
    default = [t["ms"] for t in choice["table"]
               if (t["lattice_in_lds"], t["waves_per_group"], t["order"]) == (-1, 4, "given")]
    return {"lattice_in_lds": choice["lattice_in_lds"], "waves_per_group": choice["waves_per_group"],
            "replica_order": choice["order"], "candidates": len(choice["table"]),
            "trial_kmc_steps_per_replica": choice["steps"], "best_trial_ms": choice["ms"],
            "default_trial_ms": default[0] if default else None}


def ruo2_leg(engine, K, no_cpu, tune=True):
    wl = W.ruo2()
    cm = compile_model(wl.project, wl.backend)
    model = engine.Model(cm)
    b, wall, executed, kernel_ms, chunk, choice = run_batched(engine, model, wl, K, 1, tune)
    st = b.scalars_all()[2]
    b.close()
    bps = algorithmic_bytes_per_step(wl, cm)
    avg_s = float(np.mean(kernel_ms)) / 1e3
    achieved = executed / K * bps / avg_s / 1e9
    out = {"workload": wl.name, "replicas": wl.n, "steps_per_replica": wl.steps,
           "gpu_steps_per_s": executed / wall, "avg_launch_ms": avg_s * 1e3,
           "launch": launch_summary(choice),
           "replica_status_counts": {str(k): int(v) for k, v in zip(*np.unique(st, return_counts=True))},
           "roofline": {"bound": "hbm", "achieved": achieved, "peak": PEAK_HBM_GBS, "unit": "GB/s",
                        "frac": achieved / PEAK_HBM_GBS, "traffic": measured_traffic(wl.name),
                        "algorithmic_bytes_per_kmc_step": bps, "lds": lds_figures(wl.name, avg_s)}}
    if not no_cpu:
        ref = reference_cpu_baseline(wl)
        if ref is not None:
            out["cpu_baseline"] = ref
            out["gpu_over_cpu_reference"] = out["gpu_steps_per_s"] / ref["value"]
    return out


def pairwise_leg(engine, K, no_cpu, tune=True):
    """The lateral-interaction workload (512 replicas of 256^2, pairwise
    lat_int, 5e5 steps each) with its own roofline figures and the C
    oracle on the host cores beside it (no reference build exists for it)."""
    wl = W.pairwise()
    cm = compile_model(wl.project, wl.backend)
    model = engine.Model(cm)
    b, wall, executed, kernel_ms, chunk, choice = run_batched(engine, model, wl, K, 1, tune)
    st = b.scalars_all()[2]
    b.close()
    bps = algorithmic_bytes_per_step(wl, cm)
    avg_s = float(np.mean(kernel_ms)) / 1e3
    achieved = executed / K * bps / avg_s / 1e9
    out = {"workload": wl.name, "replicas": wl.n, "steps_per_replica": wl.steps,
           "gpu_steps_per_s": executed / wall, "avg_launch_ms": avg_s * 1e3,
           "launch": launch_summary(choice),
           "replica_status_counts": {str(k): int(v) for k, v in zip(*np.unique(st, return_counts=True))},
           "roofline": {"bound": "hbm", "achieved": achieved, "peak": PEAK_HBM_GBS, "unit": "GB/s",
                        "frac": achieved / PEAK_HBM_GBS, "traffic": measured_traffic(wl.name),
                        "algorithmic_bytes_per_kmc_step": bps, "lds": lds_figures(wl.name, avg_s)}}
    if not no_cpu:
        out["cpu_baseline"] = cpu_baseline(wl, cm, 5.0)
        out["gpu_over_cpu"] = out["gpu_steps_per_s"] / out["cpu_baseline"]["value"]
    return out


def serving_leg(rate, n, seed=1, tune=True):
    """The serving stream; with `tune`, the sampling intervals per launch
    are first chosen by the autotuner on a shorter stream of the same
    distribution (another seed)."""
    from kmos_amd import autotune, serve
    reqs = serve.poisson_stream(rate=rate, n=n, seed=seed)
    ipl, tuning = 4, None
    if tune:
        res = autotune.tune_serving(lambda k: serve.Server(hbm_budget=200 << 30, intervals_per_launch=k),
                                    lambda: serve.poisson_stream(rate=rate, n=max(20, n // 4), seed=seed + 1000),
                                    candidates=(1, 2, 4, 8))
        ipl = res["intervals_per_launch"]
        tuning = [{k: r[k] for k in ("intervals_per_launch", "p50_latency_s", "p99_latency_s", "steps_per_s")}
                  for r in res["table"]]
    srv = serve.Server(hbm_budget=200 << 30, intervals_per_launch=ipl)
    t0 = time.perf_counter()
    res = srv.serve(reqs, realtime=True)
    wall = time.perf_counter() - t0
    out = serve.latency_summary(res, wall)
    out.update({"arrival_rate_per_s": rate, "ticks": srv.stats["ticks"], "pages": srv.stats["pages"],
                "peak_active": srv.stats["peak_active"], "peak_hbm_bytes": srv.stats["peak_bytes"],
                "steps_range": [1e4, 1e6], "sample_every": 1000,
                "sizes": list(serve.SIZES), "models": ["quickstart", "zgb", "ruo2"],
                "intervals_per_launch": ipl, "autotune": tuning})
    srv.close()
    return out


def serving_sweep(rates=(10.0, 1000.0), seed=2, ipl=4):
    """The serving stream at other arrival rates (the stream's sweep from 10
    to 1000 requests/s), each with about five seconds of arrivals (at least
    40 and at most 400 requests): p50/p99 latency and steps/s per rate."""
    from kmos_amd import serve
    out = []
    for rate in rates:
        n = int(min(400, max(40, rate * 5)))
        srv = serve.Server(hbm_budget=200 << 30, intervals_per_launch=ipl)
        t0 = time.perf_counter()
        res = srv.serve(serve.poisson_stream(rate=rate, n=n, seed=seed), realtime=True)
        s = serve.latency_summary(res, time.perf_counter() - t0)
        srv.close()
        out.append({"arrival_rate_per_s": rate, "requests": n, "p50_latency_s": s["p50_latency_s"],
                    "p99_latency_s": s["p99_latency_s"], "steps_per_s": s["steps_per_s"]})
    return out


def calibrate_copy(nbytes=1 << 30, reps=5):
    """A plain device-to-device copy of known size: what FETCH_SIZE and
    WRITE_SIZE read for it calibrates the step kernel's counters."""
    import torch
    src = torch.ones(nbytes // 4, dtype=torch.int32, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(reps):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes_per_dispatch": nbytes, "dispatches": reps}))
    return 0


def _progress(msg):
    """One line per finished leg on stderr (stdout carries only the JSON line)."""
    print("bench.py: %s" % msg, file=sys.stderr, flush=True)


def make_batch(engine, model, wl, idx=None):
    idx = np.arange(wl.n) if idx is None else np.asarray(idx)
    b = engine.Batch(model, [list(s) for s in wl.sizes[idx]])
    b.init(wl.seeds[idx], wl.streams[idx])
    b.set_rates(wl.rates[idx])
    if wl.configs is not None:
        b.set_config([wl.configs[i] for i in idx])
    return b


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="zgb", choices=sorted(W.WORKLOADS))
    ap.add_argument("--replicas", type=int, default=None)
    ap.add_argument("--kmc-steps", type=int, default=None, help="kMC steps per replica")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--plain-steps", type=int, default=20000)
    ap.add_argument("--no-ruo2", action="store_true", help="skip the RuO2 leg")
    ap.add_argument("--no-serve", action="store_true", help="skip the serving leg")
    ap.add_argument("--no-pairwise", action="store_true", help="skip the lateral-interaction leg")
    ap.add_argument("--serve-rate", type=float, default=100.0, help="request arrivals per second")
    ap.add_argument("--serve-requests", type=int, default=200)
    ap.add_argument("--no-tune", action="store_true",
                    help="default launch shape and serving chunk length instead of the autotuner's")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only run a device copy of 1 GiB five times (calibrates HBM byte counters)")
    a = ap.parse_args()
    if a.calibrate_copy:
        return calibrate_copy()
    if a.gpus != 1:
        print("bench.py measures one GPU: replicas are independent and the engine runs "
              "them all on a single device; run with --gpus 1", file=sys.stderr)
        return 2

    kw = {}
    if a.replicas is not None:
        kw["n"] = a.replicas
    if a.kmc_steps is not None:
        kw["steps"] = a.kmc_steps
    wl = W.WORKLOADS[a.workload](**kw)
    cm = compile_model(wl.project, wl.backend)

    from kmos_amd import _lib, engine
    _lib.load()
    if _lib.device_count() < 1:
        print("no GPU visible", file=sys.stderr)
        return 1
    model = engine.Model(cm)
    K = max(1, a.steps)
    b, wall, executed, kernel_ms, chunk, choice = run_batched(engine, model, wl, K, a.warmup, not a.no_tune)
    _progress("batched %s: %d launches timed" % (wl.name, K))
    t_after, k_after, st_after = b.scalars_all()
    value = executed / wall
    avg_kernel_s = float(np.mean(kernel_ms)) / 1e3
    steps_per_launch = executed / K

    # spot check: replica 0 against the oracle, bit-exact
    spot = None
    try:
        from oracle import pyoracle
        pyoracle.build()
        from tests.paritylib import compare_state, oracle_replica
        o = oracle_replica(pyoracle, cm, tuple(wl.sizes[0][:2]), int(wl.seeds[0]), int(wl.streams[0]),
                           wl.rates[0], None if wl.configs is None else wl.configs[0])
        o.run(int(k_after[0]))
        compare_state(b, 0, o, where="bench spot check")
        spot = True
    except AssertionError as e:
        print("spot check FAILED: %s" % e, file=sys.stderr)
        spot = False

    # plain path: one replica per launch, same kernel
    pb = make_batch(engine, model, wl, [0])
    pb.run(min(a.plain_steps, 2000))
    tp0 = time.perf_counter()
    pb.run(a.plain_steps)
    plain_wall = time.perf_counter() - tp0
    plain = a.plain_steps / plain_wall
    pb.close()

    b.close()
    bytes_per_step = algorithmic_bytes_per_step(wl, cm)
    achieved = steps_per_launch * bytes_per_step / avg_kernel_s / 1e9
    traffic = measured_traffic(wl.name)
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "kMC steps/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": a.warmup,
        "ms_per_step": wall * 1e3 / K,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (workload parameters from a seeded Philox generator)",
        "config": dict(wl.describe(), kmc_steps_per_launch=chunk),
        "launch": launch_summary(choice),
        "plain_gpu_steps_per_s": plain,
        "batched_over_plain": value / plain,
        "replica_status_counts": {str(k): int(v) for k, v in zip(*np.unique(st_after, return_counts=True))},
        "parity_spot_check": spot,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": PEAK_HBM_GBS, "unit": "GB/s",
                     "frac": achieved / PEAK_HBM_GBS, "traffic": traffic,
                     "algorithmic_bytes_per_kmc_step": bytes_per_step,
                     "kernel": "kmc_steps_kernel", "avg_launch_ms": avg_kernel_s * 1e3,
                     "lds": lds_figures(wl.name, avg_kernel_s)},
    }
    _progress("plain path and spot check done")
    out["graph_replay"] = graph_leg(engine, model, wl)
    if not a.no_cpu:
        out["cpu_baseline"] = cpu_baseline(wl, cm, a.cpu_seconds)
        out["gpu_over_cpu"] = value / out["cpu_baseline"]["value"]
        _progress("cpu baseline done")
    if not a.no_ruo2 and a.workload == "zgb":
        out["ruo2"] = ruo2_leg(engine, K, a.no_cpu, not a.no_tune)
        _progress("ruo2 leg done")
    if not a.no_pairwise and a.workload == "zgb":
        out["pairwise"] = pairwise_leg(engine, K, a.no_cpu, not a.no_tune)
        _progress("pairwise leg done")
    if not a.no_serve:
        out["serving"] = serving_leg(a.serve_rate, a.serve_requests, tune=not a.no_tune)
        _progress("serving leg done")
        out["serving_sweep"] = serving_sweep(ipl=out["serving"]["intervals_per_launch"])
        _progress("serving sweep done")
        out["p50_latency_s"] = out["serving"]["p50_latency_s"]
        out["p99_latency_s"] = out["serving"]["p99_latency_s"]
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
