"""The ``.reload`` state snapshot of one replica.

kmos writes the whole state of a system to ``<system_name>.reload`` with
``base.save_system`` and reads it back with ``base.reload_system``
(kmos/fortran_src/base.mpy:363-578).  The file is ASCII, one record per
line, ``label value...``; lines starting with ``#`` are comments and
labels the reader does not know are skipped (base.mpy:414-431,455-500):

    kmc_time, walltime, kmc_step, nr_of_proc, volume     scalars
    procstat[P], nr_of_sites[P], rates[P], lattice[V]    one line each
    avail_sites p sites[V]                               one line per process
    avail_sites_back p back[V]                           one line per process

``write_reload`` keeps that layout and those labels, so the reference's
reader loads its files, and writes reals with 17 significant digits, so a
state survives the round trip bit for bit.  It adds records that the
reference's reader skips and this reader restores when present:
``kmc_time_step``, ``integ_rates``, ``rng_seed``, ``rng_stream``,
``rng_counter`` (the engine's counter-based random stream: seed, stream and
counter fix every later draw) and ``system_size``.  A file the reference
wrote has none of them: the state then continues from zero integ_rates, as
after the reference's reload.
"""
from __future__ import annotations

import numpy as np

HEADER = "#Reload file written by kmos_amd in the layout of kmos save_system"

_SCALAR_INT = ("kmc_step", "nr_of_proc", "volume", "rng_seed", "rng_stream", "rng_counter")
_SCALAR_REAL = ("kmc_time", "walltime", "kmc_time_step")
_VECTOR = {"procstat": np.int64, "nr_of_sites": np.int32, "rates": np.float64,
           "lattice": np.int8, "integ_rates": np.float64, "system_size": np.int32}


def _reals(a) -> str:
    return " ".join("%24.16E" % v for v in np.asarray(a, np.float64).ravel())


def _ints(a, width: int) -> str:
    return "".join("%*d" % (width, v) for v in np.asarray(a).ravel())


def write_reload(path: str, state: dict) -> None:
    """Write ``state`` (the dict ``engine.Batch.get_state`` returns, plus
    ``walltime``, ``rng_seed``, ``rng_stream``, ``system_size`` when known)
    in the layout of save_system (base.mpy:557-575)."""
    sites = np.asarray(state["avail_sites"], np.int64)
    back = np.asarray(state["avail_sites_back"], np.int64)
    P, V = sites.shape
    lines = [HEADER, "#Scalar variables",
             " kmc_time  %24.16E" % state["kmc_time"],
             " walltime   %14.7E" % float(state.get("walltime", 0.0)),
             " kmc_step %22d" % int(state["kmc_step"]),
             " nr_of_proc %12d" % P,
             " volume %12d" % V,
             " kmc_time_step %24.16E" % float(state.get("kmc_time_step", 0.0))]
    for key in ("rng_seed", "rng_stream", "rng_counter"):
        if state.get(key) is not None:
            lines.append(" %s %22d" % (key, int(state[key])))
    lines.append("#Vector variables")
    lines.append("procstat " + _ints(state["procstat"], 21))
    lines.append("nr_of_sites " + _ints(state["nr_of_sites"], 9))
    lines.append("rates " + _reals(state["rates"]))
    if state.get("integ_rates") is not None:
        lines.append("integ_rates " + _reals(state["integ_rates"]))
    if state.get("system_size") is not None:
        lines.append("system_size " + _ints(state["system_size"], 9))
    lines.append("lattice " + _ints(state["lattice"], 9))
    for p in range(P):
        lines.append("avail_sites " + _ints([p + 1], 9) + _ints(sites[p], 9))
    for p in range(P):
        lines.append("avail_sites_back " + _ints([p + 1], 9) + _ints(back[p], 9))
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")


class ReloadError(ValueError):
    """A .reload file that cannot be read."""


def _fortran_float(tok: str) -> float:
    return float(tok.replace("D", "E").replace("d", "e"))


def read_reload(path: str) -> dict:
    """Parse a .reload file in two passes like reload_system
    (base.mpy:398-502): the first fixes nr_of_proc and volume, the second
    fills every array.  Unknown labels are skipped; a known record with the
    wrong number of values raises ReloadError (the reference's reader reports
    read errors on the scalar records and leaves arrays partly filled)."""
    with open(path) as f:
        records = []
        for lineno, raw in enumerate(f, 1):
            text = raw.strip()
            if not text or text.startswith("#"):
                continue
            label, _, rest = text.partition(" ")
            records.append((lineno, label, rest.split()))

    def fail(lineno, msg):
        raise ReloadError("%s:%d: %s" % (path, lineno, msg))

    out: dict = {}
    for lineno, label, toks in records:                 # first pass
        if label in ("nr_of_proc", "volume"):
            if not toks:
                fail(lineno, "%s has no value" % label)
            try:
                out[label] = int(toks[0])
            except ValueError:
                fail(lineno, "cannot read %s" % label)
    if "nr_of_proc" not in out or "volume" not in out:
        raise ReloadError("%s: nr_of_proc and volume are required" % path)
    P, V = out["nr_of_proc"], out["volume"]
    if P < 1 or V < 1:
        raise ReloadError("%s: nr_of_proc and volume must be positive" % path)
    sizes = {"procstat": P, "nr_of_sites": P, "rates": P, "integ_rates": P, "lattice": V,
             "system_size": 3}
    sites = np.zeros((P, V), np.int32)
    back = np.zeros((P, V), np.int32)
    for lineno, label, toks in records:                 # second pass
        try:
            if label in _SCALAR_INT:
                out[label] = int(toks[0])
            elif label in _SCALAR_REAL:
                out[label] = _fortran_float(toks[0])
            elif label in _VECTOR:
                if len(toks) < sizes[label]:
                    fail(lineno, "%s needs %d values, found %d" % (label, sizes[label], len(toks)))
                conv = _fortran_float if _VECTOR[label] == np.float64 else int
                out[label] = np.array([conv(t) for t in toks[:sizes[label]]], _VECTOR[label])
            elif label in ("avail_sites", "avail_sites_back"):
                p = int(toks[0])
                if not 1 <= p <= P:
                    fail(lineno, "%s row %d outside 1..%d" % (label, p, P))
                if len(toks) < V + 1:
                    fail(lineno, "%s row needs %d values, found %d" % (label, V, len(toks) - 1))
                (sites if label == "avail_sites" else back)[p - 1] = [int(t) for t in toks[1:V + 1]]
        except ReloadError:
            raise
        except (IndexError, ValueError, OverflowError) as e:
            fail(lineno, "cannot read %s: %s" % (label, e))
    for key in ("kmc_time", "kmc_step", "procstat", "nr_of_sites", "rates", "lattice"):
        if key not in out:
            raise ReloadError("%s: record %s is missing" % (path, key))
    out["avail_sites"] = sites
    out["avail_sites_back"] = back
    return out
