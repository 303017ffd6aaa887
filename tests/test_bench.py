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

# workload -> the step kernel the library launches for it (kmcg_model_forms)
WORKLOADS = {"zgb_128x128_x4096": "kmc_steps_kernel", "ruo2_local_smart_mix_x2048": "kmc_steps_kernel",
             "pairwise_lat_int_256x256_x512": "kmc_steps_kernel_li"}


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(REPO, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 2
    assert "one GPU" in r.stderr and r.stdout == ""


@pytest.mark.parametrize("workload", sorted(WORKLOADS))
def test_profile_figures_scale_with_the_run(workload):
    kernel = WORKLOADS[workload]
    traffic = bench.profile_figure(workload, "pmc", "traffic_bytes_per_kmc_step", kernel)
    lds = bench.profile_figure(workload, "lds", "lds_bytes_per_kmc_step", kernel)
    assert traffic is not None and traffic > 0
    assert lds is not None and lds > traffic        # the LDS moves far more bytes than HBM
    for spl, avg_s in ((4.096e7, 0.1), (2.048e7, 0.06)):
        f = bench.roofline(workload, 200.0, spl, avg_s, kernel)
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


def test_profiles_name_the_kernel_the_library_launches():
    """The kernel each workload's profiles are filed under is the form the
    library reports for its model (kmcg_model_forms)."""
    from kmos_amd import engine
    from kmos_amd import workloads as W
    from kmos_amd.compiler import compile_model
    for key, wl in (("zgb_128x128_x4096", W.zgb(n=2)), ("ruo2_local_smart_mix_x2048", W.ruo2(n=4)),
                    ("pairwise_lat_int_256x256_x512", W.pairwise(n=2))):
        assert wl.name.rsplit("_x", 1)[0] == key.rsplit("_x", 1)[0]
        assert engine.Model(compile_model(wl.project, wl.backend)).step_kernel == WORKLOADS[key]
