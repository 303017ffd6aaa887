"""Steady-state detection for kMC sampling (kmos/run/steady_state.py).

The initial-bias test of Rossetti, Li and Qu (Winter Simulation Conference
2005, doi:10.1109/WSC.2005.1574321): an exponentially weighted moving
average (EWMA) of each sampled quantity is checked against lower and upper
control limits (Montgomery, Introduction to Statistical Quality Control,
7th ed.); the cutoff that keeps the most EWMA points inside the limits marks
the end of the warm-up.  ``sample_steady_state`` drives a model until no
quantity asks to scrap more than ``bias_threshold`` of its history; the
batched form over many replicas is ``sweep.ModelRunner.run_steady_state``.
"""
from __future__ import annotations

import itertools

import numpy as np


def ewma_alpha(y, alpha, prev_ewma=None, adjust=True):
    """EWMA with pandas' `adjust` weighting (steady_state.py:28-49)."""
    old_wt_factor = 1.0 - alpha
    old_wt = 1.0
    new_wt = 1.0 if adjust else alpha
    alpha = np.double(alpha)
    if prev_ewma is not None:
        return alpha * y[-1] + (1 - alpha) * prev_ewma
    ewma = np.zeros_like(y, dtype=np.double)
    ewma[0] = y[0]
    for i in range(1, len(y)):
        old_wt *= old_wt_factor
        ewma[i] = (new_wt * y[i] + old_wt * ewma[i - 1]) / (old_wt + new_wt)
        if adjust:
            old_wt += new_wt
        else:
            old_wt = 1.0
    return ewma


def lcl_ucl(y, cutoff, L, lambda_factor):
    """Mean and the lower/upper control limits of the EWMA for data after
    `cutoff` (steady_state.py:52-73)."""
    mu0 = np.mean(y[cutoff:]) * np.ones_like(y)
    sigma = np.std(y[cutoff:])
    N = len(y)
    delta = L * sigma * np.sqrt(lambda_factor / (2 - lambda_factor)
                                * (1 - (1 - lambda_factor) ** np.arange(0.0, N, 1)))
    return mu0, mu0 - delta, mu0 + delta


def p2d(y, cutoff, L, alpha):
    """Fraction of EWMA points after `cutoff` inside the control limits
    (steady_state.py:76-86)."""
    ewma = ewma_alpha(y, alpha)
    _, lcl, ucl = lcl_ucl(y, cutoff, L, alpha)
    return ((lcl[cutoff:] < ewma[cutoff:]) & (ewma[cutoff:] < ucl[cutoff:])).sum() / float(len(y) - cutoff)


def get_scrap_fraction(y, L, alpha, warm_up):
    """Fraction of the history to discard as warm-up (steady_state.py:89-98)."""
    if (y[0] == y).all():
        return 0.0
    D = np.array([p2d(y, cutoff, L, alpha) for cutoff in range(len(y))])
    return (np.argmax(D[warm_up:]) + warm_up) / float(len(y))


class Detector:
    """The per-model state of sample_steady_state (steady_state.py:186-267):
    feed one get_std_sampled_data dict per batch; `done` turns true when the
    history has passed the test, `result(output)` gives the averages."""

    def __init__(self, L=4, alpha=0.05, bias_threshold=0.15, warm_up=20, check_frequency=10):
        self.L, self.alpha, self.bias_threshold = L, alpha, bias_threshold
        self.warm_up, self.check_frequency = warm_up, check_frequency
        self.hist = {}
        self.batch = -1
        self.done = False
        self.completed_percent = 0.0
        self.critical_key = ""

    def add(self, data: dict) -> bool:
        self.batch += 1
        batch = self.batch
        for key, value in data.items():
            self.hist.setdefault(key, []).append(value)
        if batch >= self.warm_up and batch % self.check_frequency == 0:
            max_scrap, critical = 0.0, ""
            for key, y in self.hist.items():
                if "time" in key or "step" in key:
                    continue
                scrap = get_scrap_fraction(np.array(y), self.L, self.alpha, self.warm_up)
                if scrap > max_scrap:
                    max_scrap, critical = scrap, key
            self.critical_key = critical
            self.completed_percent = float(1 - max_scrap) / (1 - self.bias_threshold) * 100.0
            if self.completed_percent >= 100 and batch >= self.warm_up:
                self.done = True
        return self.done

    def result(self, header: str, output="str"):
        start = int(self.batch * self.bias_threshold)
        hist = {key: np.array(value[start:]) for key, value in self.hist.items()}
        data = {}
        for key, values in hist.items():
            if "time" in key:
                data[key] = values[-1]
            elif "step" in key:
                data[key] = sum(values)
            else:
                data[key] = np.average(values, weights=hist["kmc_time"])
        if output == "dict":
            return data
        if output == "str":
            return " ".join(format(data[key.replace("#", "")], ".5e") for key in header.split()) + "\n"
        raise UserWarning("Output format %s not defined. I only know 'str' and 'dict'" % output)


def sample_steady_state(model, batch_size=1000000, L=4, alpha=0.05, bias_threshold=0.15,
                        tof_method="integ", warm_up=20, check_frequency=10, show_progress=True,
                        make_plots=False, output="str", seed="EWMA"):
    """Sample `model` in batches of `batch_size` steps (100 samples each)
    until the EWMA test finds the history free of initial bias
    (steady_state.py:135-267).  Works with any model that has
    get_std_sampled_data and get_std_header (kmos_amd.run.KMC_Model)."""
    if make_plots:
        raise NotImplementedError("EWMA plots need matplotlib, which this engine does not use")
    det = Detector(L, alpha, bias_threshold, warm_up, check_frequency)
    for _ in itertools.count():
        det.add(model.get_std_sampled_data(100, batch_size, tof_method=tof_method, output="dict"))
        if show_progress and det.batch >= warm_up and det.batch % check_frequency == 0:
            print("%3d%% limited by %-40s (%d)" % (int(det.completed_percent), det.critical_key, det.batch))
        if det.done:
            if show_progress:
                print("Done after %d batches!" % det.batch)
            break
    return det.result(model.get_std_header(), output)
