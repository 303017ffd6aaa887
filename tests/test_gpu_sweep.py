"""Sweep drivers on the engine: every grid point of a ModelRunner batch gives
the line KMC_Model gives for that point run alone (same seed, same
parameters, same init/sample steps), for get_std_sampled_data and for
sample_steady_state; the lock file stops a second run from redoing points.
Two grid points' lines are also rebuilt from the oracle: the same replica
run on the CPU, its state read at the same sample points and averaged by
the same host code, gives the same line.
"""
import os

import numpy as np
import pytest

from kmos_amd.export import export_source
from kmos_amd.run import KMC_Model, adjustable_values, interval_tof, rate_constants, std_sampled_output, tof_matrix
from kmos_amd.steady_state import sample_steady_state
from kmos_amd.sweep import LinearParameter, ModelRunner

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _zgb_with_tof(tmp_path, backend="local_smart"):
    """The ZGB model with a CO2 turnover counted on the oxidation steps."""
    text = open(os.path.join(GOLDEN, "zgb_model.ini")).read()
    text = text.replace("rate_constant = 10**10\n", "rate_constant = 10**10\ntof_count = {'CO2': 1}\n")
    ini = tmp_path / "zgb_tof.ini"
    ini.write_text(text)
    return export_source(str(ini), str(tmp_path / ("exp_" + backend)), backend)


def _oracle_line(runner, pt, seed, init_steps, samples, interval):
    """The get_std_sampled_data line of one grid point from the oracle:
    the point's replica on the CPU, its state after init_steps and after
    every sampling interval, averaged by the host code KMC_Model uses."""
    from oracle import pyoracle
    pyoracle.build()
    cm = runner.compiled
    params = runner.point_parameters(pt)
    o = pyoracle.Oracle(cm.to_blob(), tuple(runner.size))
    o.init(seed, 0)
    o.set_rates(rate_constants(runner.settings, params, cm))
    o.run(init_steps)
    tofm = tof_matrix(runner.settings, cm)
    ncells = int(np.prod(runner.size))
    t0, k0 = o.kmc_time, o.kmc_step
    t_prev, ps_prev, ig_prev = t0, o.procstat().astype(np.float64), o.integ_rates()
    tofs, occs, dts, step_ts = [], [], [], []
    for _ in range(samples):
        o.run(interval)
        t, ps, ig = o.kmc_time, o.procstat().astype(np.float64), o.integ_rates()
        tofs.append(interval_tof(tofm, ps - ps_prev, ig - ig_prev, t - t_prev, ncells, "integ").flatten())
        occs.append(list(o.occupation().flatten()))
        dts.append(t - t_prev)
        step_ts.append(o.get_state(None)["kmc_time_step"])
        t_prev, ps_prev, ig_prev = t, ps, ig
    return std_sampled_output(adjustable_values(params), tofs, occs, dts, step_ts, t_prev - t0, t_prev,
                              o.kmc_step - k0, runner.std_header())


class ScanZGB(ModelRunner):
    yCO = LinearParameter(0.40, 0.55, steps=4)


def _alone(d, size, seed, yco):
    m = KMC_Model(export_dir=d, size=size, random_seed=seed)
    m.parameters.yCO = yco
    return m


@pytest.mark.parametrize("backend", ["local_smart", "lat_int"])
def test_model_runner_equals_kmc_model_per_point(engine, tmp_path, backend):
    print("case", backend, flush=True)
    d = _zgb_with_tof(tmp_path, backend)
    runner = ScanZGB(export_dir=d, size=[16, 16], directory=str(tmp_path))
    res = runner.run(init_steps=2000, sample_steps=3001, samples=3, random_seed=7)
    assert len(res) == 4
    for pt in (runner.points()[0], runner.points()[3]):
        assert res[runner.input_line(pt)] == _oracle_line(runner, pt, 7, 2000, 3, 1000), pt
    for pt in runner.points():
        m = _alone(d, [16, 16], 7, pt[0])
        m.do_steps(2000)
        m.get_atoms(geometry=False)
        line = m.get_std_sampled_data(samples=3, sample_size=3001, tof_method="integ")
        m.deallocate()
        assert res[runner.input_line(pt)] == line, pt
    dat = open(tmp_path / "ScanZGB.dat").read().splitlines()
    assert dat[0] == runner.std_header().strip() == "#yCO CO2 CO_sc_site O_sc_site empty_sc_site kmc_time simulated_time kmc_steps"
    assert sum(1 for ln in dat if not ln.startswith("#")) == 4
    assert runner.run(init_steps=2000, sample_steps=3001, samples=3, random_seed=7) == {}


def test_steady_state_batch_equals_sample_steady_state(engine, tmp_path):
    d = _zgb_with_tof(tmp_path)

    class SteadyZGB(ModelRunner):
        yCO = LinearParameter(0.42, 0.50, steps=2)
    runner = SteadyZGB(export_dir=d, size=[12, 12], directory=str(tmp_path))
    kw = dict(batch_size=4000, warm_up=4, check_frequency=2, tof_method="integ")
    res = runner.run_steady_state(random_seed=3, **kw)
    for pt in runner.points():
        m = _alone(d, [12, 12], 3, pt[0])
        line = sample_steady_state(m, show_progress=False, **kw)
        m.deallocate()
        assert res[runner.input_line(pt)] == line, pt
