#!/usr/bin/env python3
"""Write golden trajectories of the reference Fortran engines.

Runs the reference builds of oracle/_ref (oracle/build_ref.py: the
reference's own generated Fortran, compiled with the injected Philox
uniform stream) and stores, per case, the per-step (process, site,
kmc_time) and the final state (lattice, nr_of_sites, procstat, integ_rates,
both avail_sites tables) as compressed numpy arrays under tests/golden/.

Usage: python tests/golden/make_ref_trajectories.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, REPO)
sys.path.insert(0, os.path.join(REPO, "tests"))

from kmos_amd.compiler import compile_model  # noqa: E402
from paritylib import pdopd_project, random_config  # noqa: E402
from refbin import run_reference, run_reference_otf  # noqa: E402

# (name, reference build, size, seed, stream, steps, rate seed, random config)
CASES = [
    ("ruo2_local_smart_10x10", "ruo2_local_smart", (10, 10), 3, 0, 5000, 11, False),
    ("ruo2_local_smart_8x6_config", "ruo2_local_smart", (8, 6), 4, 9, 5000, 12, True),
    ("ruo2_lat_int_10x10", "ruo2_lat_int", (10, 10), 5, 1, 5000, 13, False),
    ("ruo2_lat_int_8x6_config", "ruo2_lat_int", (8, 6), 6, 2, 5000, 14, True),
    # two layers, create/annihilate: from the initial layer and from a
    # lattice whose cells hold either layer
    ("pdopd_local_smart_5x4", "pdopd_local_smart", (5, 4), 7, 3, 4000, 15, False),
    ("pdopd_local_smart_6x5_config", "pdopd_local_smart", (6, 5), 8, 4, 4000, 16, True),
    ("pdopd_lat_int_5x4", "pdopd_lat_int", (5, 4), 9, 5, 4000, 17, False),
    ("pdopd_lat_int_6x5_config", "pdopd_lat_int", (6, 5), 10, 6, 4000, 18, True),
]
# otf (oracle/ref_driver_otf.f90): the RuO2 fixture, whose names the
# exporter shortened to 35 characters; also stores rates_matrix
OTF_CASES = [
    ("ruo2_otf_10x10", "ruo2_otf", (10, 10), 21, 7, 5000, 31, False),
    ("ruo2_otf_8x6_config", "ruo2_otf", (8, 6), 22, 8, 5000, 32, True),
]


def case_inputs(model, size, rate_seed, use_config):
    """Rates log-uniform per process; the RuO2 configurations weight the
    three species 0.3 / 0.4 / 0.3, the pdopd ones hold one layer per cell."""
    g = np.random.default_rng(rate_seed)
    if model == "ruo2":
        rates = 10.0 ** g.uniform(-2.0, 9.0, 36)
        config = None
        if use_config:
            config = g.choice([0, 1, 2], size=size[0] * size[1] * 2, p=[0.3, 0.4, 0.3]).astype(np.int8)
        return rates, config
    pt = pdopd_project()
    cm = compile_model(pt, "local_smart")
    rates = 10.0 ** g.uniform(-2.0, 6.0, cm.nr_of_proc)
    config = random_config(pt, size[0] * size[1] * cm.spuck, g) if use_config else None
    return rates, config


def main():
    for name, build, size, seed, stream, steps, rseed, use_cfg in CASES:
        model = build.split("_")[0]
        rates, config = case_inputs(model, size, rseed, use_cfg)
        res = run_reference(build, size, seed, stream, steps, rates, len(rates), config=config)
        out = os.path.join(HERE, "ref_traj_%s.npz" % name)
        np.savez_compressed(
            out, model=model, size=np.array(size), seed=seed, stream=stream, steps=steps, rates=rates,
            config=config if config is not None else np.zeros(0, np.int8),
            proc=res["proc"], site=res["site"], time=res["time"], lattice=res["lattice"],
            nr=res["nr"], procstat=res["procstat"], integ=res["integ"],
            sites=res["sites"], back=res["back"], kmc_step=res["kmc_step"])
        print(out, os.path.getsize(out))
    for name, build, size, seed, stream, steps, rseed, use_cfg in OTF_CASES:
        rates, config = case_inputs("ruo2", size, rseed, use_cfg)
        userpar = np.random.default_rng(rseed + 1000).uniform(0.5, 1.5, 14)   # 14 parameters
        res = run_reference_otf(build, size, seed, stream, steps, rates, len(rates), userpar, config=config)
        out = os.path.join(HERE, "ref_traj_%s.npz" % name)
        np.savez_compressed(
            out, model="ruo2", size=np.array(size), seed=seed, stream=stream, steps=steps, rates=rates,
            userpar=userpar, config=config if config is not None else np.zeros(0, np.int8),
            proc=res["proc"], site=res["site"], time=res["time"], lattice=res["lattice"],
            nr=res["nr"], procstat=res["procstat"], integ=res["integ"],
            sites=res["sites"], back=res["back"], kmc_step=res["kmc_step"],
            rates_matrix=res["rates_matrix"])
        print(out, os.path.getsize(out))


if __name__ == "__main__":
    main()
