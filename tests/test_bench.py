"""The benchmark's host-side contract, without a GPU: more than one GPU is
refused with a message, the committed rocprofv3 summaries parse into the
traffic and LDS figures the JSON line carries, the autotuner summary names
the default shape's trial, and the algorithmic bytes per kMC step of the ZGB
workload are the figure its profile and roofline quote."""
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
def test_profile_figures_parse(workload):
    traffic = bench.measured_traffic(workload)
    lds = bench.measured_lds(workload)
    assert traffic is not None and traffic > 0
    assert lds is not None and lds > traffic        # the LDS moves far more bytes than HBM
    f = bench.lds_figures(workload, 0.1)
    assert f["unit"] == "GB/s" and f["peak"] == pytest.approx(256 * 128 * 2.4)
    assert f["frac"] == pytest.approx(lds / 0.1 / 1e9 / f["peak"])
    assert bench.measured_lds("no_such_workload") is None and bench.lds_figures("no_such_workload", 1.0) is None


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
    bps = bench.algorithmic_bytes_per_step(wl, compile_model(wl.project, wl.backend))
    text = open(os.path.join(REPO, "profiles", "kmc_steps_kernel_zgb_128x128_x4096_pmc.md")).read()
    assert "| %.1f |" % bps in text
