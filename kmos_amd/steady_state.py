"""Steady-state detection for kMC sampling (the role of
kmos/run/steady_state.py:135-267, ``sample_steady_state``).

Method: the initial-bias test of Rossetti, Li and Qu, "Exploring
exponentially weighted moving average control charts to determine the
warm-up period" (Winter Simulation Conference 2005,
doi:10.1109/WSC.2005.1574321).  Each sampled quantity y_0..y_{n-1} is
smoothed by the adjusted EWMA (pandas ``Series.ewm(alpha, adjust=True)``)

    z_t = sum_{i<=t} (1-a)^(t-i) y_i  /  sum_{i<=t} (1-a)^(t-i),

and for every candidate end of the warm-up c the points z_c..z_{n-1} are
tested against a control band around the mean m_c of y_c..y_{n-1} with
standard deviation s_c (population form):

    |band half-width at t| = L s_c sqrt(a/(2-a) (1 - (1-a)^t)).

(kmos uses the exponent t rather than the textbook 2t of Montgomery's
EWMA chart; it is kept so that a run is judged steady after the same batch
as under kmos.)  The cut c >= warm_up that keeps the largest fraction of
its points strictly inside the band is the warm-up; c/n is the fraction of
the history to scrap.  A run is steady when no quantity (keys naming a time
or a step count are not tested) asks to scrap more than ``bias_threshold``.

``sample_steady_state`` drives one model batch by batch until then; the
batched form over a parameter grid is ``sweep.ModelRunner.run_steady_state``,
which feeds one ``SteadyStateMonitor`` per replica.
"""
from __future__ import annotations

import numpy as np
from scipy.signal import lfilter


def ewma(y, alpha, adjust=True):
    """EWMA of the series y: the adjusted (normalised-weight) form by
    default, the plain recursion z_t = a y_t + (1-a) z_{t-1}, z_0 = y_0,
    with adjust=False.  Both equal pandas' ``ewm(alpha=...).mean()``."""
    y = np.asarray(y, dtype=np.float64)
    if y.size == 0:
        return y.copy()
    decay = 1.0 - float(alpha)
    if adjust:
        # numerator and normaliser obey the same first-order recursion
        return lfilter([1.0], [1.0, -decay], y) / lfilter([1.0], [1.0, -decay], np.ones_like(y))
    # plain form: z_t - decay z_{t-1} = a y_t, started from z_0 = y_0
    z, _ = lfilter([float(alpha)], [1.0, -decay], y[1:], zi=[decay * y[0]])
    return np.concatenate([y[:1], z])


def band_profile(n, L, alpha):
    """Control-band half-width per point t = 0..n-1 for unit standard
    deviation (see the module docstring)."""
    t = np.arange(n, dtype=np.float64)
    return L * np.sqrt(alpha / (2.0 - alpha) * (1.0 - (1.0 - alpha) ** t))


def inside_fractions(y, L, alpha, cuts):
    """For every cut c in `cuts`: the fraction of EWMA points z_c..z_{n-1}
    strictly inside the band around the tail's mean."""
    y = np.asarray(y, dtype=np.float64)
    z = ewma(y, alpha)
    w = band_profile(len(y), L, alpha)
    out = np.empty(len(cuts))
    for j, c in enumerate(cuts):
        tail = y[c:]
        m, s = tail.mean(), tail.std()
        half = s * w[c:]
        zt = z[c:]
        out[j] = np.count_nonzero((m - half < zt) & (zt < m + half)) / float(tail.size)
    return out


def scrap_fraction(y, L, alpha, warm_up):
    """Fraction of the history y to discard as warm-up: the first best cut
    at or after `warm_up`, over len(y).  A constant series has none."""
    y = np.asarray(y, dtype=np.float64)
    if np.all(y == y[0]):
        return 0.0
    if warm_up >= len(y):
        raise ValueError("the history (%d batches) must be longer than warm_up (%d)" % (len(y), warm_up))
    cuts = np.arange(warm_up, len(y))
    best = cuts[int(np.argmax(inside_fractions(y, L, alpha, cuts)))]
    return best / float(len(y))


# kmos' names for the two pieces users call directly
def ewma_alpha(y, alpha, prev_ewma=None, adjust=True):
    """kmos' entry point: the EWMA of y, or with `prev_ewma` the one-step
    update a y[-1] + (1-a) prev_ewma."""
    if prev_ewma is not None:
        return alpha * y[-1] + (1.0 - alpha) * prev_ewma
    return ewma(y, alpha, adjust)


def get_scrap_fraction(y, L, alpha, warm_up):
    """kmos' entry point for ``scrap_fraction``."""
    return scrap_fraction(y, L, alpha, warm_up)


def _tested(key):
    return "time" not in key and "step" not in key


class SteadyStateMonitor:
    """Batch-by-batch state of one steady-state run.

    ``add(data)`` takes one ``get_std_sampled_data(..., output='dict')``
    result and returns whether the run is now steady; the test runs at
    batch indices b >= warm_up with b % check_frequency == 0.  ``summary``
    averages the history after the first ``bias_threshold`` of it: keys
    naming a time keep their last value, keys naming a step count are
    summed, everything else is averaged with the batches' ``kmc_time`` as
    weights."""

    def __init__(self, L=4, alpha=0.05, bias_threshold=0.15, warm_up=20, check_frequency=10):
        self.L, self.alpha = L, alpha
        self.bias_threshold = bias_threshold
        self.warm_up, self.check_frequency = warm_up, check_frequency
        self.keys = None
        self.rows = []
        self.done = False
        self.completed_percent = 0.0
        self.critical_key = ""

    @property
    def batch(self):
        """0-based index of the last batch added (-1 before the first)."""
        return len(self.rows) - 1

    def add(self, data: dict) -> bool:
        if self.keys is None:
            self.keys = list(data)
        self.rows.append([data[k] for k in self.keys])
        b = self.batch
        if b < self.warm_up or b % self.check_frequency:
            return self.done
        table = np.asarray(self.rows, dtype=np.float64)
        worst, worst_key = 0.0, ""
        for j, key in enumerate(self.keys):
            if not _tested(key):
                continue
            f = scrap_fraction(table[:, j], self.L, self.alpha, self.warm_up)
            if f > worst:
                worst, worst_key = f, key
        self.critical_key = worst_key
        self.completed_percent = (1.0 - worst) / (1.0 - self.bias_threshold) * 100.0
        self.done = self.completed_percent >= 100.0
        return self.done

    def summary(self, header: str, output="str"):
        keep = self.rows[int(self.batch * self.bias_threshold):]
        cols = {k: np.asarray([row[j] for row in keep]) for j, k in enumerate(self.keys)}
        weights = cols["kmc_time"]
        data = {}
        for k, v in cols.items():
            if "time" in k:
                data[k] = v[-1]
            elif "step" in k:
                data[k] = sum(v)
            else:
                data[k] = np.average(v, weights=weights)
        if output == "dict":
            return data
        if output == "str":
            return " ".join(format(data[name.lstrip("#")], ".5e") for name in header.split()) + "\n"
        raise UserWarning("Output format %s not defined. I only know 'str' and 'dict'" % output)


def sample_steady_state(model, batch_size=1000000, L=4, alpha=0.05, bias_threshold=0.15,
                        tof_method="integ", warm_up=20, check_frequency=10, show_progress=True,
                        make_plots=False, output="str", seed="EWMA"):
    """Sample `model` in batches of `batch_size` kMC steps (100 samples per
    batch) until the run is steady, then return the averaged sampled data
    ('str': one line in the order of the model's std header; 'dict').
    Works with any model offering get_std_sampled_data and get_std_header
    (kmos_amd.run.KMC_Model).  Same parameters as kmos'
    sample_steady_state (kmos/run/steady_state.py:135-148)."""
    if make_plots:
        raise NotImplementedError("EWMA plots need matplotlib, which this engine does not use")
    mon = SteadyStateMonitor(L, alpha, bias_threshold, warm_up, check_frequency)
    while True:
        steady = mon.add(model.get_std_sampled_data(100, batch_size, tof_method=tof_method, output="dict"))
        b = mon.batch
        if show_progress and b >= warm_up and b % check_frequency == 0:
            print("%3d%% limited by %-40s (%d)" % (int(mon.completed_percent), mon.critical_key, b))
        if steady:
            if show_progress:
                print("Done after %d batches!" % b)
            return mon.summary(model.get_std_header(), output)
