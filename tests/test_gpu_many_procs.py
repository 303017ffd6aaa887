"""Models with more processes than lanes: 72 processes (two per lane) and
136 processes (four per lane) against the oracle bit for bit, ragged sizes
and random starting lattices in one batch.  The per-lane process slots,
the accumulate over several slots and the queue replay are shared by both
backends; these models run local_smart (as lat_int their op programs
exceed the 64 KiB the engine keeps in LDS, which model loading rejects)."""
import numpy as np
import pytest

from kmos_amd.compiler import compile_model
from kmos_amd.model import parse_ini
from paritylib import compare_state, oracle_replica

pytestmark = pytest.mark.gpu

SPECIES = ["A", "B", "C", "D"]
DIRS4 = [(1, 0), (0, 1), (-1, 0), (0, -1)]
DIRS8 = DIRS4 + [(1, 1), (-1, 1), (-1, -1), (1, -1)]


def many_process_project(dirs):
    """Adsorption and desorption of four species, their hops and the
    pairwise reactions X + Y -> empty + empty along every direction."""
    lines = ["[Meta]", "author = test", "email = test@example.org", "model_name = many",
             "model_dimension = 2", "debug = 0", "", "[SpeciesList]", "default_species = empty", "",
             "[Species empty]", "representation =", "color = #ffffff", "tags =", ""]
    for s in SPECIES:
        lines += ["[Species %s]" % s, "representation =", "color = #000000", "tags =", ""]
    lines += ["[Lattice]", "cell_size = 1.0 0.0 0.0 0.0 1.0 0.0 0.0 0.0 1.0", "default_layer = sq",
              "substrate_layer = sq", "representation =", "", "[Layer sq]", "color = #ffffff",
              "site s = (0.5, 0.5, 0.5); default_species;", ""]

    def proc(name, conds, acts):
        return ["[Process %s]" % name, "rate_constant = 1.", "enabled = True",
                "conditions = " + " + ".join(conds), "actions = " + " + ".join(acts), ""]
    for s in SPECIES:
        lines += proc("%s_ads" % s, ["empty@s"], ["%s@s" % s])
        lines += proc("%s_des" % s, ["%s@s" % s], ["empty@s"])
        for k, (dx, dy) in enumerate(dirs):
            nb = "s.(%d, %d, 0)" % (dx, dy)
            lines += proc("%s_hop%d" % (s, k), ["%s@s" % s, "empty@%s" % nb], ["empty@s", "%s@%s" % (s, nb)])
    for x in SPECIES:
        for y in SPECIES:
            if x == y:
                continue
            for k, (dx, dy) in enumerate(dirs):
                nb = "s.(%d, %d, 0)" % (dx, dy)
                lines += proc("%s_%s_react%d" % (x, y, k), ["%s@s" % x, "%s@%s" % (y, nb)],
                              ["empty@s", "empty@%s" % nb])
    return parse_ini("\n".join(lines))


@pytest.mark.parametrize("dirs,nproc", [(DIRS4, 72), (DIRS8, 136)], ids=["P72", "P136"])
@pytest.mark.parametrize("backend", ["local_smart"])
def test_many_processes_match_oracle(engine, oracle, dirs, nproc, backend):
    print("case", nproc, backend, flush=True)
    pt = many_process_project(dirs)
    cm = compile_model(pt, backend)
    assert cm.nr_of_proc == nproc
    sizes = [(9, 7), (16, 16), (30, 20)]
    g = np.random.default_rng(nproc)
    rates = 10.0 ** g.uniform(-2, 2, (len(sizes), nproc))
    m = engine.Model(cm)
    b = engine.Batch(m, [list(s) + [1] for s in sizes])
    seeds = np.arange(len(sizes), dtype=np.uint32) + 21
    b.init(seeds, np.zeros(len(sizes), np.uint32))
    b.set_rates(rates)
    configs = [g.integers(0, len(SPECIES) + 1, int(v)).astype(np.int8) for v in b.volumes]
    b.set_config(configs)
    orcs = [oracle_replica(oracle, cm, s, int(seeds[i]), 0, rates[i], configs[i]) for i, s in enumerate(sizes)]
    for r, o in enumerate(orcs):
        compare_state(b, r, o, where="P%d %s r%d after init" % (nproc, backend, r))
    done = 0
    for chunk in (1, 499, 2500):
        b.run(chunk)
        for o in orcs:
            o.run(chunk)
        done += chunk
        for r, o in enumerate(orcs):
            compare_state(b, r, o, where="P%d %s r%d after %d" % (nproc, backend, r, done))
