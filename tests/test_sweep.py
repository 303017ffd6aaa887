"""Sweep drivers on the CPU: grids, the runner's bookkeeping, the sampled-data
averaging and the steady-state test.

* T_grid / p_grid against what they promise (1/T and log p evenly spaced,
  end points kept) and LogParameter / LinearParameter against numpy;
* the runner collects its parameters in definition order, walks the grid
  with the first parameter slowest, names points like kmos
  (``<runner>_<v1>_<v2>``) and honours the lock file;
* the parameters block of the .dat header in the reference's layout
  (kmos/run/__init__.py:1670-1692);
* std_sampled_output against a direct weighted mean;
* the EWMA against pandas' EWM (adjusted and plain), the control band and
  the per-cut inside fractions against a hand-worked 4-point series, the
  scrap fraction on constant and on step signals, and SteadyStateMonitor
  against sample_steady_state on the same data stream.
"""
import numpy as np
import pytest

from kmos_amd import steady_state as SS
from kmos_amd.run import parameters_repr, std_sampled_output
from kmos_amd.sweep import (LinearParameter, LogParameter, ModelRunner, PressureParameter,
                            TemperatureParameter, T_grid, p_grid)


def test_grids():
    T = np.array(T_grid(400.0, 800.0, 5))
    assert T[0] == pytest.approx(400.0) and T[-1] == pytest.approx(800.0)
    np.testing.assert_allclose(np.diff(1.0 / T), np.full(4, (1 / 800 - 1 / 400) / 4), rtol=1e-12)
    p = p_grid(1e-3, 10.0, 5)
    assert p[0] == pytest.approx(1e-3) and p[-1] == pytest.approx(10.0)
    np.testing.assert_allclose(np.diff(np.log(p)), np.full(4, np.log(1e4) / 4), rtol=1e-12)
    np.testing.assert_array_equal(LogParameter(-1, 2, steps=4).get_grid(), np.logspace(-1, 2, 4))
    np.testing.assert_array_equal(LinearParameter(0.3, 0.6, steps=4).get_grid(), np.linspace(0.3, 0.6, 4))
    fixed = PressureParameter(2.0)
    assert fixed.max == 2.0 and list(fixed.get_grid()) == [2.0]
    assert TemperatureParameter(600).unit == "K" and fixed.unit == "bar"


class _Scan(ModelRunner):
    b_par = LinearParameter(1.0, 2.0, steps=2)
    a_par = LogParameter(0, 1, steps=3)


def _bare_runner(tmp_path):
    r = object.__new__(_Scan)           # no model needed for the bookkeeping
    r.directory = str(tmp_path)
    return r


def test_runner_bookkeeping(tmp_path):
    assert _Scan.runner_name == "_Scan"
    assert list(_Scan.parameters) == ["b_par", "a_par"]
    r = _bare_runner(tmp_path)
    pts = r.points()
    assert len(pts) == 6
    assert [p[0] for p in pts] == [1.0, 1.0, 1.0, 2.0, 2.0, 2.0]       # first parameter slowest
    np.testing.assert_allclose([p[1] for p in pts[:3]], [1.0, 10 ** 0.5, 10.0])
    assert r.input_line(pts[0]) == "_Scan_1.0_1.0"
    first = r._claim(pts[:4])
    assert first == pts[:4]
    assert r._claim(pts) == pts[4:]                      # claimed points are skipped
    assert r._claim(pts) == []
    with open(tmp_path / "_Scan.lock") as f:
        assert [line.strip() for line in f] == [r.input_line(p) for p in pts]


def test_parameters_block_layout():
    params = {"yCO": {"value": 0.45, "adjustable": True}, "T": {"value": "600", "adjustable": False},
              "A": {"value": 1.5}}
    assert parameters_repr(params) == (
        "# kMC model parameters (3, fixed 2)\n"
        "# --------------------\n"
        "# A = 1.5\n"
        "# T = 600\n"
        "# yCO = 0.45  # *\n"
        "# --------------------\n"
        "# * adjustable parameters\n")


def test_std_sampled_output_weights():
    g = np.random.default_rng(1)
    tofs = g.uniform(0, 5, (4, 2))
    occs = g.uniform(0, 1, (4, 3))
    dts = g.uniform(0.1, 1, 4)
    steps_t = g.uniform(1e-3, 1e-2, 4)
    line = std_sampled_output([0.5], list(tofs), list(occs), list(dts), list(steps_t), 2.0, 7.0, 400,
                              "#y t1 t2 o1 o2 o3 kmc_time simulated_time kmc_steps\n")
    vals = [float(v) for v in line.split()]
    tof = (tofs * dts[:, None]).sum(0) / dts.sum()
    occ = (occs * steps_t[:, None]).sum(0) / steps_t.sum()
    np.testing.assert_allclose(vals, [0.5, *tof, *occ, 2.0, 7.0, 400], rtol=6e-6)
    d = std_sampled_output([0.5], list(tofs), list(occs), list(dts), list(steps_t), 2.0, 7.0, 400,
                           "#y t1 t2 o1 o2 o3 kmc_time simulated_time kmc_steps\n", output="dict")
    assert list(d) == ["y", "t1", "t2", "o1", "o2", "o3", "kmc_time", "simulated_time", "kmc_steps"]


def test_ewma_matches_pandas():
    pd = pytest.importorskip("pandas")
    y = np.random.default_rng(3).normal(size=200).cumsum()
    for alpha in (0.05, 0.3):
        for adjust in (True, False):
            ref = pd.Series(y).ewm(alpha=alpha, adjust=adjust).mean().to_numpy()
            np.testing.assert_allclose(SS.ewma(y, alpha, adjust), ref, rtol=1e-12, atol=1e-12)
        np.testing.assert_allclose(SS.ewma_alpha(y, alpha), SS.ewma(y, alpha), rtol=0, atol=0)


def test_band_and_inside_fractions_by_hand():
    # half-width L sqrt(a/(2-a) (1-(1-a)^t)) at a = 0.5, L = 1: 0, sqrt(1/6), sqrt(1/4), sqrt(7/24)
    np.testing.assert_allclose(SS.band_profile(4, 1.0, 0.5), [0.0, 6 ** -0.5, 0.5, (7 / 24) ** 0.5],
                               rtol=1e-15)
    # y = 0,1,0,1 at a = 0.5: z = 0, 2/3, 2/7, 2/3 (weights 1, 1/2, 1/4, 1/8 normalised)
    y = np.array([0.0, 1.0, 0.0, 1.0])
    np.testing.assert_allclose(SS.ewma(y, 0.5), [0.0, 2 / 3, 2 / 7, 2 / 3], rtol=1e-15)
    # cut 0: m = s = 1/2, bands 0, .204, .25, .270 -> z1, z2, z3 inside, z0 not (strict)
    # cut 1: m = 2/3, s = .471 -> z1 and z3 inside, z2 (.381 away) outside the .236 band
    # cut 2: m = s = 1/2 -> both inside; cut 3: s = 0 -> nothing inside
    np.testing.assert_allclose(SS.inside_fractions(y, 1.0, 0.5, [0, 1, 2, 3]), [0.75, 2 / 3, 1.0, 0.0])
    assert SS.scrap_fraction(y, 1.0, 0.5, 0) == 0.5          # best cut 2 of 4
    assert SS.scrap_fraction(y, 1.0, 0.5, 3) == 0.75         # only cut 3 allowed
    with pytest.raises(ValueError):
        SS.scrap_fraction(y, 1.0, 0.5, 4)


def test_scrap_fraction():
    assert SS.get_scrap_fraction(np.full(50, 3.0), 4, 0.05, 5) == 0.0
    g = np.random.default_rng(4)
    y = np.concatenate([np.linspace(10, 1, 40), 1 + 0.01 * g.normal(size=160)])   # decay, then noise
    f = SS.get_scrap_fraction(y, 4, 0.05, 5)
    assert 0.2 <= f < 1.0                       # at least the decay is scrapped
    assert SS.get_scrap_fraction(1 + 0.01 * g.normal(size=200), 4, 0.05, 5) < 0.15


class _FakeModel:
    """Deterministic stand-in: a relaxing signal plus noise per batch."""

    def __init__(self):
        self.g = np.random.default_rng(7)
        self.n = 0

    def get_std_header(self):
        return "#x kmc_time simulated_time kmc_steps\n"

    def get_std_sampled_data(self, samples, sample_size, tof_method="integ", output="dict"):
        self.n += 1
        return {"x": 1 + 5 * np.exp(-self.n / 4.0) + 0.05 * self.g.normal(), "kmc_time": 1.0 + 0.1 * self.n,
                "simulated_time": 2.0 * self.n, "kmc_steps": sample_size}


def test_monitor_is_sample_steady_state():
    a = SS.sample_steady_state(_FakeModel(), batch_size=100, warm_up=5, check_frequency=3, show_progress=False)
    m = _FakeModel()
    det = SS.SteadyStateMonitor(warm_up=5, check_frequency=3)
    while not det.add(m.get_std_sampled_data(100, 100)):
        pass
    assert det.summary(m.get_std_header()) == a
    assert det.batch >= 5 and det.batch % 3 == 0
    d = det.summary(m.get_std_header(), "dict")
    start = int(det.batch * 0.15)
    assert d["kmc_steps"] == 100 * (det.batch + 1 - start)
