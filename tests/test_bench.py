"""The benchmark's host-side contract, without a GPU: more than one GPU is
refused with a message, the committed rocprofv3 summaries parse into the
per-kMC-step traffic and LDS figures, the roofline scales them by the run's
own kMC steps per launch and launch time (so the LDS bandwidth is the
profile's bytes per step times the kernel's steps/s), the autotuner summary
names the default shape's trial, the CPU baselines state the cores they
used, and the algorithmic bytes per kMC step of the ZGB workload are the
figure its profile quotes."""
import os
import subprocess
import sys

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, REPO)
import bench  # noqa: E402

WORKLOADS = ("zgb_128x128_x4096", "ruo2_local_smart_mix_x2048", "pairwise_lat_int_256x256_x512")


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(REPO, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 2
    assert "one GPU" in r.stderr and r.stdout == ""


@pytest.mark.parametrize("workload", WORKLOADS)
def test_profile_figures_scale_with_the_run(workload):
    traffic = bench.profile_figure(workload, "pmc", "traffic_bytes_per_kmc_step")
    lds = bench.profile_figure(workload, "lds", "lds_bytes_per_kmc_step")
    assert traffic is not None and traffic > 0
    assert lds is not None and lds > traffic        # the LDS moves far more bytes than HBM
    for spl, avg_s in ((4.096e7, 0.1), (2.048e7, 0.06)):
        f = bench.roofline(workload, 200.0, spl, avg_s)
        kernel_rate = spl / avg_s                     # kMC steps/s inside the kernel
        assert f["achieved"] == pytest.approx(200.0 * kernel_rate / 1e9)
        assert f["frac"] == pytest.approx(f["achieved"] / 8000.0)
        assert f["traffic"] == pytest.approx(traffic * spl)
        assert f["lds"]["achieved"] == pytest.approx(lds * kernel_rate / 1e9)
        assert f["lds"]["peak"] == pytest.approx(256 * 128 * 2.4)
        assert f["lds"]["frac"] == pytest.approx(f["lds"]["achieved"] / f["lds"]["peak"])
    assert bench.profile_figure("no_such_workload", "lds", "lds_bytes_per_kmc_step") is None
    none = bench.roofline("no_such_workload", 100.0, 1e6, 1.0)
    assert none["lds"] is None and none["traffic"] is None


def test_host_cores_are_stated():
    avail, used = bench.host_cores()
    assert avail == len(os.sched_getaffinity(0)) and used == min(avail, 16)


def test_launch_summary():
    table = [{"lattice_in_lds": -1, "waves_per_group": 4, "order": "given", "ms": 5.0, "runs": [5.0]},
             {"lattice_in_lds": -1, "waves_per_group": 1, "order": "given", "ms": 4.0, "runs": [4.0]}]
    s = bench.launch_summary({"lattice_in_lds": -1, "waves_per_group": 1, "order": "given", "ms": 4.0,
                              "steps": 1000, "table": table})
    assert s["waves_per_group"] == 1 and s["default_trial_ms"] == 5.0 and s["best_trial_ms"] == 4.0
    assert s["candidates"] == 2 and bench.launch_summary(None) is None


def test_zgb_algorithmic_bytes_match_the_profile():
    from kmos_amd import workloads as W
    from kmos_amd.compiler import compile_model
    wl = W.zgb()
    wl.project_blob = compile_model(wl.project, wl.backend).to_blob()
    bps = bench.algorithmic_bytes_per_step(wl)
    text = open(os.path.join(REPO, "profiles", "kmc_steps_kernel_zgb_128x128_x4096_pmc.md")).read()
    assert "| %.1f |" % bps in text


def test_oracle_cpu_baseline_states_its_cores():
    from kmos_amd import workloads as W
    from kmos_amd.compiler import compile_model
    wl = W.quickstart(steps=20000)
    wl.project_blob = compile_model(wl.project, wl.backend).to_blob()
    out = bench.oracle_cpu_baseline(wl, seconds=0.5)
    assert out["kind"] == "port" and out["value"] > 0
    assert out["cores_used"] == min(out["cores_available"], 16) == out["cores"]
