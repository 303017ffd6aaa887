"""Run a reference Fortran driver (oracle/_ref/<name>/ref_driver) and parse
its output (format documented in oracle/ref_driver.f90)."""
import os
import subprocess
import tempfile

import numpy as np

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF_DIR = os.path.join(REPO, "oracle", "_ref")


def driver_path(name):
    p = os.path.join(REF_DIR, name, "ref_driver")
    return p if os.path.exists(p) else None


def run_reference(name, size, seed, stream, nsteps, rates, P, config=None, mode=0,
                  reload_text=None, rng_step=0):
    """mode 0 trace, 1 do_kmc_steps, 2 trace then save_system (the file's
    text comes back as res["reload"]), 3 reload_system(reload_text), then
    trace from step rng_step of the uniform stream."""
    exe = driver_path(name)
    if exe is None:
        raise FileNotFoundError("reference build %s missing (run oracle/build_ref.py)" % name)
    X, Y = size[0], size[1]
    Z = size[2] if len(size) > 2 else 1
    res = {}
    with tempfile.TemporaryDirectory() as td:
        inp = os.path.join(td, "in.txt")
        out = os.path.join(td, "out.bin")
        with open(inp, "w") as fh:
            fh.write("%d %d %d\n%d %d\n%d %d\n%d\n" % (X, Y, Z, seed, stream, nsteps, mode,
                                                      0 if config is None else 1))
            fh.write(" ".join(repr(float(r)) for r in rates) + "\n")
            if config is not None:
                fh.write(" ".join(str(int(c)) for c in config) + "\n")
            if mode == 3:
                fh.write("%d\n" % rng_step)
        if mode == 3:
            with open(os.path.join(td, "in.reload"), "w") as fh:
                fh.write(reload_text)
        subprocess.check_call([exe, inp, out], cwd=td)
        buf = open(out, "rb").read()
        if mode == 2:
            res["reload"] = open(os.path.join(td, "ref.reload")).read()
    off = 0

    def take(dtype, count):
        nonlocal off
        a = np.frombuffer(buf, dtype=dtype, count=count, offset=off)
        off += a.nbytes
        return a.copy()

    n = int(take("<i4", 1)[0])
    if mode != 1:
        res["proc"] = take("<i4", n)
        res["site"] = take("<i4", n)
        res["time"] = take("<f8", n)
    res["kmc_time"] = float(take("<f8", 1)[0])
    res["kmc_step"] = int(take("<i8", 1)[0])
    V = (len(buf) - off - 20 * P) // (4 + 8 * P)
    res["lattice"] = take("<i4", V)
    res["nr"] = take("<i4", P)
    res["procstat"] = take("<i8", P)
    res["integ"] = take("<f8", P)
    avail = take("<i4", P * V * 2).reshape(2, V, P)   # Fortran column-major (P, V, 2)
    res["sites"] = avail[0].T.copy()
    res["back"] = avail[1].T.copy()
    assert off == len(buf)
    return res


def run_reference_otf(name, size, seed, stream, nsteps, rates, P, userpar=(), config=None, mode=0):
    """The otf driver (oracle/ref_driver_otf.f90): mode 0 trace, 1
    do_kmc_steps; returns the trace, the final state and rates_matrix."""
    exe = driver_path(name)
    if exe is None:
        raise FileNotFoundError("reference build %s missing (run oracle/build_ref.py)" % name)
    X, Y = size[0], size[1]
    Z = size[2] if len(size) > 2 else 1
    res = {}
    with tempfile.TemporaryDirectory() as td:
        inp = os.path.join(td, "in.txt")
        out = os.path.join(td, "out.bin")
        with open(inp, "w") as fh:
            fh.write("%d %d %d\n%d %d\n%d %d\n%d\n" % (X, Y, Z, seed, stream, nsteps, mode,
                                                      0 if config is None else 1))
            fh.write(" ".join(repr(float(r)) for r in rates) + "\n")
            fh.write("%d\n" % len(userpar))
            if len(userpar):
                fh.write(" ".join(repr(float(u)) for u in userpar) + "\n")
            if config is not None:
                fh.write(" ".join(str(int(c)) for c in config) + "\n")
        r = subprocess.run([exe, inp, out], cwd=td, stdout=subprocess.PIPE, text=True, check=True)
        res["stdout"] = r.stdout
        buf = open(out, "rb").read()
    off = 0

    def take(dtype, count):
        nonlocal off
        a = np.frombuffer(buf, dtype=dtype, count=count, offset=off)
        off += a.nbytes
        return a.copy()

    n = int(take("<i4", 1)[0])
    if mode != 1:
        res["proc"] = take("<i4", n)
        res["site"] = take("<i4", n)
        res["time"] = take("<f8", n)
    res["kmc_time"] = float(take("<f8", 1)[0])
    res["kmc_step"] = int(take("<i8", 1)[0])
    # lattice V*4 + 20 P + avail P*V*8 + rates P*(V+1)*8
    V = (len(buf) - off - 20 * P - 8 * P) // (4 + 16 * P)
    res["lattice"] = take("<i4", V)
    res["nr"] = take("<i4", P)
    res["procstat"] = take("<i8", P)
    res["integ"] = take("<f8", P)
    avail = take("<i4", P * V * 2).reshape(2, V, P)
    res["sites"] = avail[0].T.copy()
    res["back"] = avail[1].T.copy()
    res["rates_matrix"] = take("<f8", P * (V + 1)).reshape(P, V + 1)
    assert off == len(buf)
    return res
