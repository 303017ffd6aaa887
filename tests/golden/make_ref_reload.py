#!/usr/bin/env python3
"""Write golden .reload snapshots of the reference Fortran engines.

For each case the reference build of oracle/_ref runs `steps` traced steps
and calls its own save_system (kmos/fortran_src/base.mpy:517-578); the file
it wrote is stored as ``ref_reload_<case>.reload``.  The reference then
loads that file with its own reload_system (base.mpy:363-514) and runs
`cont` more steps with the uniform stream continued at step `steps`; the
state before and after the continuation goes to ``ref_reload_<case>.npz``.

Usage: python tests/golden/make_ref_reload.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, REPO)
sys.path.insert(0, os.path.join(REPO, "tests"))

from refbin import run_reference  # noqa: E402

# (case, reference build, size, seed, stream, steps, continuation steps, rate seed)
CASES = [
    ("ruo2_local_smart_6x6", "ruo2_local_smart", (6, 6), 7, 3, 2000, 1000, 21),
    ("ruo2_lat_int_6x6", "ruo2_lat_int", (6, 6), 8, 4, 2000, 1000, 22),
]
P = 36
FIELDS = ("kmc_time", "kmc_step", "lattice", "nr", "procstat", "integ", "sites", "back")


def main():
    for case, build, size, seed, stream, steps, cont, rseed in CASES:
        rates = 10.0 ** np.random.default_rng(rseed).uniform(-2.0, 9.0, P)
        saved = run_reference(build, size, seed, stream, steps, rates, P, mode=2)
        text = saved["reload"]
        path = os.path.join(HERE, "ref_reload_%s.reload" % case)
        with open(path, "w") as fh:
            fh.write(text)
        # the continuation: any rates given here are replaced by the file's
        again = run_reference(build, size, seed, stream, cont, rates, P, mode=3,
                              reload_text=text, rng_step=steps)
        out = os.path.join(HERE, "ref_reload_%s.npz" % case)
        arrays = {"saved_" + k: np.asarray(saved[k]) for k in FIELDS}
        arrays.update({"cont_" + k: np.asarray(again[k]) for k in FIELDS})
        np.savez_compressed(out, size=np.array(size), seed=seed, stream=stream, steps=steps,
                            cont=cont, rates=rates, cont_proc=again["proc"],
                            cont_site=again["site"], cont_time=again["time"], **arrays)
        print(path, os.path.getsize(path))
        print(out, os.path.getsize(out))


if __name__ == "__main__":
    main()
