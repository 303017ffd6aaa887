#!/usr/bin/env python3
"""Benchmark: aggregate kMC steps/s of the batched engine on one MI355X.

Contract line (default workload): the batched RuO2(110) CO-oxidation model,
2048 replicas of 20^2 : 64^2 : 128^2 cells (1 : 2 : 1), 36 processes with
rates log-uniform in [1e-2, 1e9] per replica, random starting coverage
0.3 CO / 0.3 O, 2e5 kMC steps each.  One bench "step" is one launch that
advances every replica by steps_per_replica / K kMC steps, so the K timed
launches run the whole workload (4.1e8 kMC steps); the autotuner's trials
and the warm-up run on a separate probe batch.

Printed (one JSON line): whole-job kMC steps/s (value), the plain path (one
replica per launch) on the same model, their ratio, the HBM roofline of
the step kernel from the algorithmic bytes per kMC step (int32 avail_sites
entries read or written, counted by the instrumented oracle) with the
measured HBM traffic and LDS bandwidth beside it (bytes per kMC step from
the committed rocprofv3 counter summaries, times this run's kMC steps per
launch and over this run's launch time), the launch shape the autotuner
chose, and the CPU baseline: the reference's own generated Fortran (built
from its checked-in export by oracle/build_ref.py, injected uniforms), one
process per host core, with its per-step CPU_TIME call (faithful) and
without it (sensitivity).  Legs:

  zgb        4096 replicas of the ZGB model on 128^2, 1e5 steps, with the
             C oracle on the host cores beside it
  pairwise   512 replicas of 256^2, pairwise lat_int, 5e5 steps, with the
             C oracle on the host cores
  quickstart one replica of the quickstart model (20^2, CO ads/des at
             100/s), 1e6 steps in one launch, beside the reference's
             published single-core figure and the C oracle on one core
  serving    Poisson requests from the quickstart / ZGB / RuO2 workloads
             through the continuous-batching server: p50/p99 latency and
             steps/s at 100 requests/s, the same at 10 and 1000 requests/s
             (serving_sweep), and a burst that fills thousands of slots at
             once (serving_saturated)
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
PEAK_LDS_GBS = 256 * 128 * 2.4   # CUs x bytes per clock per CU x GHz
METRIC = "aggregate kMC steps/s, batched vs plain GPU; p50/p99 request latency"
# the reference's only published rate: `kmos benchmark` on the quickstart
# model, 1e6 steps in 1.51 s on one core (README.md:85-96)
PUBLISHED_QUICKSTART = 1e6 / 1.51
# worker processes of the CPU baselines: the host cores this process may
# use, at most 16 (the CPU share one GPU's host gives a job)
CORE_CAP = 16


def host_cores():
    avail = len(os.sched_getaffinity(0))
    return avail, min(avail, CORE_CAP)


def profile_figure(workload, kind, key, kernel="kmc_steps_kernel"):
    """A per-kMC-step figure ('traffic' or 'lds' summary) from the rocprofv3
    counter summary committed for this workload and kernel, or None."""
    path = os.path.join(REPO, "profiles", "%s_%s_%s.md" % (kernel, workload, kind))
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith(key + ":"):
            return float(line.split(":")[1])
    return None


def roofline(workload, bytes_per_step, steps_per_launch, avg_launch_s, kernel="kmc_steps_kernel"):
    """HBM roofline of the step kernel: algorithmic bytes per launch over the
    mean launch time; the measured HBM traffic and the LDS bandwidth from
    the profiles' bytes per kMC step, scaled to this run's launches."""
    achieved = bytes_per_step * steps_per_launch / avg_launch_s / 1e9
    t_step = profile_figure(workload, "pmc", "traffic_bytes_per_kmc_step", kernel)
    l_step = profile_figure(workload, "lds", "lds_bytes_per_kmc_step", kernel)
    lds = None
    if l_step is not None:
        a = l_step * steps_per_launch / avg_launch_s / 1e9
        lds = {"bytes_per_kmc_step": l_step, "achieved": a, "peak": PEAK_LDS_GBS, "unit": "GB/s",
               "frac": a / PEAK_LDS_GBS}
    return {"bound": "hbm", "achieved": achieved, "peak": PEAK_HBM_GBS, "unit": "GB/s",
            "frac": achieved / PEAK_HBM_GBS,
            "traffic": None if t_step is None else t_step * steps_per_launch,
            "traffic_bytes_per_kmc_step": t_step,
            "algorithmic_bytes_per_kmc_step": bytes_per_step,
            "kernel": kernel, "avg_launch_ms": avg_launch_s * 1e3,
            "kmc_steps_per_launch": steps_per_launch, "lds": lds}


# ---------------------------------------------------------------------------
# CPU baselines
# ---------------------------------------------------------------------------

def _oracle_worker(args):
    blob, size, seed, stream, rates, config, steps, fast = args
    from oracle import pyoracle
    o = pyoracle.Oracle(blob, size, fast=fast)
    o.init(seed, stream)
    o.set_rates(rates)
    if config is not None:
        o.set_config(config)
    o.reset_touched()
    t0 = time.perf_counter()
    n = o.run(steps)
    dt = time.perf_counter() - t0
    return n, dt, o.touched


def _job(wl, r, steps, fast):
    return (wl.project_blob, tuple(int(v) for v in wl.sizes[r][:2]), int(wl.seeds[r]), int(wl.streams[r]),
            wl.rates[r], None if wl.configs is None else wl.configs[r], steps, fast)


def algorithmic_bytes_per_step(wl, sample=8, steps=20000):
    """Mean int32 avail_sites entries touched per kMC step x 4 B (the
    oracle's counter), over `sample` replicas of the workload spread over
    its size mix."""
    from oracle import pyoracle
    pyoracle.build()
    tot_steps, tot_touch = 0, 0
    for r in np.linspace(0, wl.n - 1, min(sample, wl.n)).astype(int):
        n, _, touched = _oracle_worker(_job(wl, r, steps, False))
        tot_steps += n
        tot_touch += touched
    return 4.0 * tot_touch / max(tot_steps, 1)


def oracle_cpu_baseline(wl, seconds=15.0):
    """The C oracle (timing build: -O3, no entry counter) on the host cores,
    one process per core, each running whole replicas of the workload for
    its full step budget; the sample is sized to about `seconds` of wall."""
    from oracle import pyoracle
    pyoracle.build()
    avail, cores = host_cores()
    n1, dt1, _ = _oracle_worker(_job(wl, 0, min(wl.steps, 20000), True))
    per_replica = wl.steps / (n1 / max(dt1, 1e-9))
    nrep = int(max(cores, min(wl.n, round(seconds * cores / per_replica))))
    nrep = max(cores, (nrep // cores) * cores)
    idx = np.linspace(0, wl.n - 1, nrep).astype(int)
    ctx = mp.get_context("fork")
    t0 = time.perf_counter()
    with ctx.Pool(cores) as pool:
        res = pool.map(_oracle_worker, [_job(wl, r, wl.steps, True) for r in idx], chunksize=1)
    wall = time.perf_counter() - t0
    steps = sum(r[0] for r in res)
    return {"value": steps / wall, "unit": "kMC steps/s", "cores": cores, "cores_available": avail,
            "cores_used": cores, "kind": "port",
            "sample": "%d replicas x %d steps of %s (C oracle -O3 without its entry counter, one "
                      "process per core, %.1f s wall)" % (nrep, wl.steps, wl.name, wall)}


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


def reference_cpu_baseline(wl, per_core=12, variant=""):
    """The reference's generated Fortran (oracle/_ref, built from the
    reference's checked-in export), one process per host core, each running
    whole replicas of the workload through the reference's do_kmc_steps;
    variant "_noclock" is the build without the per-step CPU_TIME call."""
    exe = os.path.join(REPO, "oracle", "_ref", "ruo2_%s%s" % (wl.backend, variant), "ref_driver")
    if not os.path.exists(exe):
        return None
    avail, cores = host_cores()
    nrep = min(wl.n, cores * per_core)
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
    return {"value": cores * steps / loop, "unit": "kMC steps/s", "cores": cores, "cores_available": avail,
            "cores_used": cores, "kind": "reference",
            "sample": "%d replicas (sizes spread over the mix) x %d steps of %s through the reference "
                      "Fortran's do_kmc_steps (flang -O3, injected uniforms%s), one process per core "
                      "running concurrently; value = cores x steps / summed loop seconds; %.1f s wall"
                      % (nrep, wl.steps, wl.name,
                         ", per-step CPU_TIME call removed" if variant else ", per-step CPU_TIME call kept",
                         wall)}


# ---------------------------------------------------------------------------
# GPU legs
# ---------------------------------------------------------------------------

def make_batch(engine, model, wl, idx=None):
    idx = np.arange(wl.n) if idx is None else np.asarray(idx)
    b = engine.Batch(model, [list(s) for s in wl.sizes[idx]])
    b.init(wl.seeds[idx], wl.streams[idx])
    b.set_rates(wl.rates[idx])
    if wl.configs is not None:
        b.set_config([wl.configs[i] for i in idx])
    return b


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


def spot_check(b, wl, cm, k_after, replicas=(0,)):
    """The listed replicas of the timed batch against the oracle, bit-exact."""
    from oracle import pyoracle
    pyoracle.build()
    from tests.paritylib import compare_state, oracle_replica
    try:
        for r in replicas:
            o = oracle_replica(pyoracle, cm, tuple(wl.sizes[r][:2]), int(wl.seeds[r]), int(wl.streams[r]),
                               wl.rates[r], None if wl.configs is None else wl.configs[r])
            o.run(int(k_after[r]))
            compare_state(b, r, o, where="bench spot check replica %d" % r)
    except AssertionError as e:
        print("spot check FAILED: %s" % e, file=sys.stderr)
        return False
    return True


def plain_rate(engine, model, wl, steps, r=0):
    """The plain path: one replica per launch, same kernel; kMC steps/s of
    one launch of `steps` steps after a short warm-up launch."""
    pb = make_batch(engine, model, wl, [r])
    pb.run(min(steps, 2000))
    t0 = time.perf_counter()
    pb.run(steps)
    wall = time.perf_counter() - t0
    pb.close()
    return steps / wall


def batched_leg(engine, wl, K, warmup, tune, cpu, spot=(0,)):
    """One batched workload: throughput, launch shape, replica status,
    parity spot check, roofline and (with `cpu`) the C oracle baseline."""
    cm = compile_model(wl.project, wl.backend)
    wl.project_blob = cm.to_blob()
    model = engine.Model(cm)
    b, wall, executed, kernel_ms, chunk, choice = run_batched(engine, model, wl, K, warmup, tune)
    _, k_after, st = b.scalars_all()
    ok = spot_check(b, wl, cm, k_after, spot)
    b.close()
    avg_s = float(np.mean(kernel_ms)) / 1e3
    out = {"workload": wl.name, "replicas": wl.n, "steps_per_replica": wl.steps,
           "gpu_steps_per_s": executed / wall, "wall_s": wall, "avg_launch_ms": avg_s * 1e3,
           "launch": launch_summary(choice), "parity_spot_check": ok,
           "replica_status_counts": {str(k): int(v) for k, v in zip(*np.unique(st, return_counts=True))},
           "roofline": roofline(wl.name, algorithmic_bytes_per_step(wl), executed / K, avg_s, model.step_kernel)}
    if cpu:
        out["cpu_baseline"] = oracle_cpu_baseline(wl, 10.0)
        out["gpu_over_cpu"] = out["gpu_steps_per_s"] / out["cpu_baseline"]["value"]
    return out, model, cm, wall, executed, chunk, choice


def quickstart_leg(engine, cpu):
    """Config 1: one replica of the quickstart model, 1e6 steps in one
    launch (the plain path at its own workload), beside the reference's
    published single-core rate and the C oracle on one core."""
    wl = W.quickstart()
    cm = compile_model(wl.project, wl.backend)
    wl.project_blob = cm.to_blob()
    model = engine.Model(cm)
    b = make_batch(engine, model, wl)
    b.run(2000)                     # warm-up on a second batch below
    b.close()
    b = make_batch(engine, model, wl)
    t0 = time.perf_counter()
    b.run(wl.steps)
    wall = time.perf_counter() - t0
    _, k_after, _ = b.scalars_all()
    ok = spot_check(b, wl, cm, k_after)
    b.close()
    out = {"workload": wl.name, "replicas": 1, "steps_per_replica": wl.steps,
           "gpu_steps_per_s": wl.steps / wall, "kernel_ms": None, "parity_spot_check": ok,
           "published_single_core_steps_per_s": PUBLISHED_QUICKSTART,
           "published_source": "kmos README.md:85-96 (kmos benchmark, 1e6 steps in 1.51 s, hardware not stated)"}
    if cpu:
        n, dt, _ = _oracle_worker(_job(wl, 0, wl.steps, True))
        out["cpu_one_core_steps_per_s"] = n / dt
    return out


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


def tracking_leg(engine, K, steps=20000, warm=20000):
    """Displacement tracking (the MSD/ACF side channel: the tracking form of
    the step kernel) against the same batch without it.  The pairwise
    model compiled for local_smart (49 processes, hops in four directions),
    512 replicas of 256^2; after `warm` steps the A particles are numbered
    (msd_init) and K launches of `steps` steps are timed, through
    do_kmc_steps_displacement and do_kmc_steps respectively, the plain form
    (no overlapped select) in both runs."""
    from kmos_amd import acf
    wl = W.pairwise()
    cm = compile_model(wl.project, "local_smart")
    model = engine.Model(cm)
    hops, disp = acf.displacement_table(cm)
    trace = cm.project.species_ids()["A"]
    out = {"workload": "pairwise_local_smart_256x256_x%d" % wl.n, "replicas": wl.n, "tracked_species": "A",
           "kmc_steps_per_launch": wl.n * steps, "launches": K}
    for mode in ("plain", "tracking"):
        b = make_batch(engine, model, wl)
        b.set_overlap(0)
        b.run(warm)
        if mode == "tracking":
            b.msd_init(trace, hops, disp)
        # do_kmc_steps or do_kmc_steps_displacement (synchronous)
        step = (lambda k: b.run_displacement(k)) if mode == "tracking" else (lambda k: b.run(k))
        step(2000)
        ms = []
        t0 = time.perf_counter()
        for _ in range(K):
            step(steps)
            ms.append(b.last_run_ms())
        wall = time.perf_counter() - t0
        out[mode + "_gpu_steps_per_s"] = wl.n * steps * K / wall
        out[mode + "_avg_launch_ms"] = float(np.mean(ms))
        if mode == "tracking":
            st = b.msd_state(0)
            ni = int(st["nr_of_ions"])
            d = st["displacement"][1:ni + 1]
            out["tracked_particles_replica0"] = ni
            # calc_msd: mean of |displacement|^2 over the numbered particles
            out["msd_replica0"] = float(((d[:, 0] ** 2 + d[:, 1] ** 2) + d[:, 2] ** 2).mean()) if ni else 0.0
        b.close()
    out["tracking_over_plain"] = out["tracking_gpu_steps_per_s"] / out["plain_gpu_steps_per_s"]
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
    to 1000 requests/s), each with five seconds of arrivals (at least 40
    requests): p50/p99 latency and steps/s per rate."""
    from kmos_amd import serve
    out = []
    for rate in rates:
        n = int(max(40, rate * 5))
        srv = serve.Server(hbm_budget=200 << 30, intervals_per_launch=ipl)
        t0 = time.perf_counter()
        res = srv.serve(serve.poisson_stream(rate=rate, n=n, seed=seed), realtime=True)
        s = serve.latency_summary(res, time.perf_counter() - t0)
        out.append({"arrival_rate_per_s": rate, "requests": n, "p50_latency_s": s["p50_latency_s"],
                    "p99_latency_s": s["p99_latency_s"], "steps_per_s": s["steps_per_s"],
                    "peak_active": srv.stats["peak_active"]})
        srv.close()
    return out


def serving_saturated(batched_rate, n=4096, rate=1e5, seed=3, ipl=8):
    """A burst that fills the server: `n` requests arriving within
    n / rate seconds, so that thousands of trajectories are resident at
    once.  Reports p50/p99, aggregate steps/s over the whole stream, the
    steps/s while at least half the requests were still in flight, and
    both as a fraction of the contract line's batched rate."""
    from kmos_amd import serve
    reqs = serve.poisson_stream(rate=rate, n=n, seed=seed)
    srv = serve.Server(hbm_budget=200 << 30, intervals_per_launch=ipl, slots_per_size=256)
    t0 = time.perf_counter()
    res = srv.serve(reqs, realtime=True)
    wall = time.perf_counter() - t0
    s = serve.latency_summary(res, wall)
    # the saturated window: from the last arrival until half the requests
    # have finished; steps done in it from each request's sample records
    fin = np.sort([r.finished for r in res])
    t_a, t_b = max(r.arrival for r in res), fin[len(fin) // 2]
    steps_in = 0.0
    for r in res:
        if r.finished <= t_a or r.steps_done == 0:
            continue
        # steps spread evenly over the request's residence (admission to finish)
        span = max(r.finished - r.admitted, 1e-9)
        lo, hi = max(r.admitted, t_a), min(r.finished, t_b)
        if hi > lo:
            steps_in += r.steps_done * (hi - lo) / span
    sat = steps_in / (t_b - t_a) if t_b > t_a else None
    out = {"requests": n, "arrival_rate_per_s": rate, "p50_latency_s": s["p50_latency_s"],
           "p99_latency_s": s["p99_latency_s"], "steps_per_s": s["steps_per_s"],
           "steps_per_s_saturated_window": sat, "saturated_window_s": t_b - t_a,
           "peak_active": srv.stats["peak_active"], "pages": srv.stats["pages"],
           "peak_hbm_bytes": srv.stats["peak_bytes"], "intervals_per_launch": ipl,
           "steps_per_s_over_batched": s["steps_per_s"] / batched_rate,
           "saturated_over_batched": None if sat is None else sat / batched_rate}
    srv.close()
    return out


PHASES = ("select", "site_read", "post", "replay", "rest")


def phase_profile(engine, wl, K, warmup):
    """The contract workload run by the phase-counting build of the engine
    (libkmcgpu_phases.so): shader-clock ticks per kMC step in each phase of
    the step loop, over all replicas and per lattice size, and the ticks
    per microsecond implied by the kernel time."""
    cm = compile_model(wl.project, wl.backend)
    model = engine.Model(cm)
    b, wall, executed, kernel_ms, chunk, _ = run_batched(engine, model, wl, K, warmup, tune=False)
    ph = b.phase_profile().astype(np.float64)
    b.close()
    steps = ph[:, 5]
    out = {"workload": wl.name, "replicas": wl.n, "kmc_steps_per_launch": executed / K,
           "avg_launch_ms": float(np.mean(kernel_ms)), "gpu_steps_per_s": executed / wall}
    ticks = ph[:, :5].sum(axis=1)
    # a wave runs for about the whole launch: its ticks over the launch time
    out["ticks_per_us"] = float(np.median(ticks / K / (np.mean(kernel_ms) * 1e3)))

    def table(sel):
        tot = ph[sel, :5].sum(axis=0) / max(steps[sel].sum(), 1.0)
        return {name: {"ticks_per_step": float(v), "frac": float(v / tot.sum())} for name, v in zip(PHASES, tot)}

    out["all"] = table(np.ones(wl.n, bool))
    # replay blocks: batches of list reads (one HBM round trip each)
    out["replay_blocks_per_step"] = float(ph[:, 6].sum() / max(steps.sum(), 1.0))
    out["replay_blocks_per_replay"] = float(ph[:, 6].sum() / max(ph[:, 7].sum(), 1.0))
    vol = wl.sizes.prod(axis=1)
    out["by_size"] = {"%dx%d" % tuple(wl.sizes[np.nonzero(vol == v)[0][0]][:2]): table(vol == v)
                      for v in np.unique(vol)}
    return out


# ---------------------------------------------------------------------------
# otf
# ---------------------------------------------------------------------------

def _otf_worker(args):
    blob, size, seed, stream, rates, userpar, steps, fast = args
    from oracle import pyoracle
    o = pyoracle.OtfOracle(blob, size, fast=fast)
    o.init(seed, stream)
    o.set_params(rates, userpar)
    t0 = time.perf_counter()
    n = o.run(steps)
    dt = time.perf_counter() - t0
    return n, dt, o.work()


def _otf_job(wl, blob, r, steps, fast):
    return (blob, tuple(int(v) for v in wl.sizes[r][:2]), int(wl.seeds[r]), int(wl.streams[r]),
            wl.rates[r], wl.userpar[r], steps, fast)


def otf_algorithmic_bytes(wl, blob, sample=4, steps=5000):
    """Per kMC step, from the oracle's work counters over `sample`
    replicas: 16 B per list operation (the entry's cell, inverse index and
    rate), 16 B per running-sum entry that must be re-added (its rate read,
    its sum written: every entry from the first changed position of each
    list on) and 8 B per probe of a binary site search."""
    from oracle import pyoracle
    pyoracle.build_otf()
    tot = np.zeros(3)
    n_steps = 0
    for r in np.linspace(0, wl.n - 1, min(sample, wl.n)).astype(int):
        n, _, w = _otf_worker(_otf_job(wl, blob, r, steps, False))
        tot += w
        n_steps += n
    per = tot / max(n_steps, 1)
    return 16.0 * per[0] + 16.0 * per[1] + 8.0 * per[2], {
        "list_ops_per_step": per[0], "readded_entries_per_step": per[1], "search_probes_per_step": per[2]}


def otf_cpu_baseline(wl, blob, seconds=15.0):
    """The otf oracle (-O3; the reference's intZGB otf fixture does not
    build) on the host cores, one process per core, whole replicas."""
    from oracle import pyoracle
    pyoracle.build_otf()
    avail, cores = host_cores()
    n1, dt1, _ = _otf_worker(_otf_job(wl, blob, 0, min(wl.steps, 5000), True))
    per_replica = wl.steps / (n1 / max(dt1, 1e-9))
    nrep = int(max(cores, min(wl.n, round(seconds * cores / per_replica))))
    nrep = max(cores, (nrep // cores) * cores)
    idx = np.linspace(0, wl.n - 1, nrep).astype(int)
    ctx = mp.get_context("fork")
    t0 = time.perf_counter()
    with ctx.Pool(cores) as pool:
        res = pool.map(_otf_worker, [_otf_job(wl, blob, r, wl.steps, True) for r in idx], chunksize=1)
    wall = time.perf_counter() - t0
    steps = sum(r[0] for r in res)
    return {"value": steps / wall, "unit": "kMC steps/s", "cores": cores, "cores_available": avail,
            "cores_used": cores, "kind": "port",
            "sample": "%d replicas x %d steps of %s (otf oracle -O3, one process per core, %.1f s wall)"
                      % (nrep, wl.steps, wl.name, wall)}


def otf_leg(K, warmup, cpu, wl=None):
    """The otf engine on the intZGB workload: K timed launches after
    `warmup`, every replica its share of the step budget per launch; replica
    0 checked against the otf oracle afterwards."""
    from kmos_amd import otf_engine
    from oracle import pyoracle
    wl = wl or W.intzgb_otf()
    cm = compile_model(wl.project, "otf")
    blob = cm.to_blob()
    m = otf_engine.OtfModel(cm)
    b = otf_engine.OtfBatch(m, wl.sizes)
    b.set_params(wl.rates, wl.userpar)
    b.init(wl.seeds, wl.streams)
    chunk = max(1, wl.steps // (K + warmup))
    for _ in range(warmup):
        b.run(chunk)
    kernel_ms = []
    t0 = time.perf_counter()
    for _ in range(K):
        b.run(chunk)
        kernel_ms.append(b.last_run_ms())
    wall = time.perf_counter() - t0
    status = [b.scalars(r)[2] for r in range(wl.n)]
    executed = sum(b.scalars(r)[1] for r in range(wl.n)) - warmup * chunk * wl.n
    pyoracle.build_otf()
    o = pyoracle.OtfOracle(blob, tuple(wl.sizes[0][:2]))
    o.init(int(wl.seeds[0]), int(wl.streams[0]))
    o.set_params(wl.rates[0], wl.userpar[0])
    o.run((K + warmup) * chunk)
    s, bk, _ = b.avail(0)
    s2, bk2, _ = o.avail()
    spot = bool(np.array_equal(b.lattice(0), o.lattice()) and np.array_equal(s, s2) and
                np.array_equal(b.counters(0)["procstat"], o.procstat()))
    bps, work = otf_algorithmic_bytes(wl, blob)
    avg_s = float(np.mean(kernel_ms)) / 1e3
    steps_per_launch = chunk * wl.n
    rl = roofline(wl.name, bps, steps_per_launch, avg_s, kernel="otf_steps_kernel")
    out = {"workload": wl.name, "replicas": int(wl.n), "steps_per_replica": int((K + warmup) * chunk),
           "gpu_steps_per_s": executed / wall, "kmc_steps_per_launch": steps_per_launch,
           "avg_launch_ms": avg_s * 1e3, "replica_status_counts": {k: status.count(k) for k in set(status)},
           "parity_spot_check": spot, "roofline": rl, "work_per_step": work}
    if cpu:
        out["cpu_baseline"] = otf_cpu_baseline(wl, blob)
        out["gpu_over_cpu"] = out["gpu_steps_per_s"] / out["cpu_baseline"]["value"]
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


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="ruo2", choices=sorted(W.WORKLOADS))
    ap.add_argument("--replicas", type=int, default=None)
    ap.add_argument("--kmc-steps", type=int, default=None, help="kMC steps per replica")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--plain-steps", type=int, default=20000)
    ap.add_argument("--legs", default="zgb,pairwise,otf,tracking,quickstart,serving",
                    help="comma-separated legs beside the contract line (empty: none)")
    ap.add_argument("--serve-rate", type=float, default=100.0, help="request arrivals per second")
    ap.add_argument("--serve-requests", type=int, default=200)
    ap.add_argument("--no-tune", action="store_true",
                    help="default launch shape and serving chunk length instead of the autotuner's")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only run a device copy of 1 GiB five times (calibrates HBM byte counters)")
    ap.add_argument("--phases", action="store_true",
                    help="only the phase profile of the step loop on the workload (loads the "
                         "phase-counting build libkmcgpu_phases.so), one JSON line")
    a = ap.parse_args()
    if a.calibrate_copy:
        return calibrate_copy()
    if a.phases:
        os.environ["KMCG_PHASES"] = "1"
    if a.gpus != 1:
        print("bench.py measures one GPU: replicas are independent and the engine runs "
              "them all on a single device; run with --gpus 1", file=sys.stderr)
        return 2

    if a.workload == "intzgb_otf":
        # the otf backend alone: one JSON line of its leg
        from kmos_amd import _lib
        _lib.load()
        kw = {}
        if a.replicas is not None:
            kw["n"] = a.replicas
        if a.kmc_steps is not None:
            kw["steps"] = a.kmc_steps
        print(json.dumps(otf_leg(max(1, a.steps), a.warmup, not a.no_cpu, W.intzgb_otf(**kw))))
        return 0
    kw = {}
    if a.replicas is not None:
        kw["n"] = a.replicas
    if a.kmc_steps is not None:
        kw["steps"] = a.kmc_steps
    wl = W.WORKLOADS[a.workload](**kw)
    legs = [x for x in a.legs.split(",") if x]

    from kmos_amd import _lib, engine
    _lib.load()
    if _lib.device_count() < 1:
        print("no GPU visible", file=sys.stderr)
        return 1
    K = max(1, a.steps)
    if a.phases:
        print(json.dumps(phase_profile(engine, wl, K, a.warmup)))
        return 0
    head, model, cm, wall, executed, chunk, choice = batched_leg(engine, wl, K, a.warmup, not a.no_tune,
                                                                 cpu=False)
    _progress("batched %s: %d launches timed" % (wl.name, K))
    value = executed / wall
    plain = plain_rate(engine, model, wl, a.plain_steps)
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
        "launch": head["launch"],
        "plain_gpu_steps_per_s": plain,
        "batched_over_plain": value / plain,
        "replica_status_counts": head["replica_status_counts"],
        "parity_spot_check": head["parity_spot_check"],
        "roofline": head["roofline"],
    }
    _progress("plain path and spot check done")
    out["graph_replay"] = graph_leg(engine, model, wl)
    if not a.no_cpu:
        ref = reference_cpu_baseline(wl)
        if ref is not None:
            out["cpu_baseline"] = ref
            out["gpu_over_cpu"] = value / ref["value"]
            nc = reference_cpu_baseline(wl, variant="_noclock")
            if nc is not None:
                out["cpu_baseline_no_cpu_time"] = nc
                out["gpu_over_cpu_no_cpu_time"] = value / nc["value"]
        else:
            wl.project_blob = cm.to_blob()
            out["cpu_baseline"] = oracle_cpu_baseline(wl)
            out["gpu_over_cpu"] = value / out["cpu_baseline"]["value"]
        _progress("cpu baseline done")
    if "zgb" in legs and a.workload != "zgb":
        out["zgb"] = batched_leg(engine, W.zgb(), K, 1, not a.no_tune, not a.no_cpu)[0]
        _progress("zgb leg done")
    if "pairwise" in legs and a.workload != "pairwise":
        out["pairwise"] = batched_leg(engine, W.pairwise(), K, 1, not a.no_tune, not a.no_cpu)[0]
        _progress("pairwise leg done")
    if "otf" in legs:
        out["otf"] = otf_leg(K, 1, not a.no_cpu)
        _progress("otf leg done")
    if "tracking" in legs:
        out["tracking"] = tracking_leg(engine, K)
        _progress("tracking leg done")
    if "quickstart" in legs and a.workload != "quickstart":
        out["quickstart"] = quickstart_leg(engine, not a.no_cpu)
        _progress("quickstart leg done")
    if "serving" in legs:
        out["serving"] = serving_leg(a.serve_rate, a.serve_requests, tune=not a.no_tune)
        _progress("serving leg done")
        out["serving_sweep"] = serving_sweep(ipl=out["serving"]["intervals_per_launch"])
        _progress("serving sweep done")
        out["serving_saturated"] = serving_saturated(value)
        _progress("saturated serving done")
        out["p50_latency_s"] = out["serving"]["p50_latency_s"]
        out["p99_latency_s"] = out["serving"]["p99_latency_s"]
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
