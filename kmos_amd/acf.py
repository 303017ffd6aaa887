"""Particle tracking of a traced species (kmos/run/acf.py): mean squared
displacements and the autocorrelation of a site property.

kmos records, per kMC step, which particle of a traced species a diffusion
process moves and by how much (do_kmc_steps_displacement,
kmos/fortran_src/proclist_generic_subroutines_acf.mpy:143-219, over the
base_acf module, base_acf.mpy:217-361 and 825-891).  The site pair and the
displacement of every process come from a generated routine,
get_diff_sites_displacement (kmos/io/acf.py:443-671); here the same
information is a table computed once on the host (``displacement_table``)
and the engine applies it inside the step kernel (kmcg_run_displacement):

    from kmos_amd import acf
    acf.initialize_msd(model, "ion")
    acf.allocate_trajectory(model, 1000)
    acf.do_kmc_steps_displacement(model, 1000, traj_on=True)
    acf.calc_msd(model), acf.get_displacement(model), acf.get_trajectory(model)

and the autocorrelation function (do_kmc_steps_acf, same file 1-98): every
step adds, to the time bins since the last wrap, the step's duration times
the mean over tracked particles of product_property(property at the
particle's first site, property at its site now):

    acf.allocate_acf(model, 2, t_bin=1e-3, t_f=1e-1)
    acf.set_property_acf(model, "default_a_1", 1); acf.set_types_acf(model, 1.0)
    acf.calc_product_property(model); acf.initialize_acf(model, "ion")
    acf.do_kmc_steps_acf(model, 100000); acf.get_acf(model, normalization=True)

``model`` is a kmos_amd.run.KMC_Model (replica 0 of its batch).  The
local_smart and lat_int backends' routines are restated (the otf backend is
not built).  A process is a *hop* when, in
the generated routine, one action takes the particle away (the site becomes
the default species) and one puts a particle down on a site that held the
default species, with no action that leaves its species unchanged (kmos's
"source"/"drain" processes) and no other species change; for such models
(pure diffusion) the trajectory is the reference's bit for bit.  For any
other process the reference routine leaves the site pair undefined
(uninitialised variables); here it moves no particle.
"""
from __future__ import annotations

import numpy as np

from .compiler import BACKEND_LAT_INT, BACKEND_LOCAL_SMART, CompiledModel

HOP_W = 9


def _f32(v) -> float:
    """A default-REAL literal of the generated lattice module (single
    precision, widened to double on assignment)."""
    return float(np.float32(v))


def displacement_table(cm: CompiledModel):
    """(hops int32[P][9], disp float64[P][3]) for kmcg_msd_init: per process
    the hop flag, the left site (dx, dy, dz, n) and the reached site
    relative to the chosen site's cell, and the displacement
    matmul(unit_cell_size, cell difference + site position difference).

    local_smart chooses a process at its executing coordinate, so the sites
    are taken relative to it (the smart writer, kmos/io/acf.py:443-671);
    lat_int chooses it at site 1 of its own cell, so they are the actions'
    coordinates themselves (the otf/lat_int writer, kmos/io/acf.py:674-1000,
    `lsite = nr2lattice(nr_site) + (/0,0,0,-1/)`).  In both the particle
    leaves the site of the action that takes it and reaches the site of the
    action that puts it."""
    if cm.backend not in (BACKEND_LOCAL_SMART, BACKEND_LAT_INT):
        raise NotImplementedError("displacement tracking covers the local_smart and lat_int backends")
    pt = cm.project
    default = pt.default_species
    cell = [[_f32(v) for v in row] for row in pt.cell]
    pos = [[_f32(v) for v in s.pos] for layer in pt.layers for s in layer.sites]
    P = cm.nr_of_proc
    hops = np.zeros((P, HOP_W), np.int32)
    disp = np.zeros((P, 3), np.float64)
    for p, proc in enumerate(pt.processes):
        ex_off = proc.executing_coord().offset if cm.backend == BACKEND_LOCAL_SMART else (0, 0, 0)

        def previous(action):
            return [c.species for c in proc.conditions if c.coord == action.coord][0]

        source = any(a.species == previous(a) for a in proc.actions)
        old = new = None
        ok = True
        for a in proc.actions:
            prev = previous(a)
            rel = (a.coord.offset[0] - ex_off[0], a.coord.offset[1] - ex_off[1],
                   a.coord.offset[2] - ex_off[2], pt.site_index(a.coord.layer, a.coord.name))
            if a.species[:1] in ("^", "$"):
                ok = False
            elif a.species == default and a.species != prev:
                if source:
                    ok = False                      # drain_process
                else:
                    old = rel
            elif prev != a.species:
                if prev != default:
                    ok = False                      # the routine would call take_<prev> here
                if source:
                    ok = False                      # source_process
                else:
                    new = rel
        if not (ok and old is not None and new is not None):
            continue
        hops[p] = (1,) + old + new
        v = [float(new[c] - old[c]) + (pos[new[3] - 1][c] - pos[old[3] - 1][c]) for c in range(3)]
        for i in range(3):
            acc = 0.0
            for j in range(3):
                acc = acc + cell[i][j] * v[j]
            disp[p, i] = acc
    return hops, disp


# ---- kmos.run.acf functions (MSD part), on replica 0 of a KMC_Model -------

def _batch(model):
    return model._kmc._b()


def initialize_msd(model, species):
    """Number the sites holding `species` 1.. and zero the displacements."""
    trace = getattr(model.proclist, species.lower())
    hops, disp = displacement_table(model.compiled)
    _batch(model).msd_init(int(trace), hops, disp, 0, 1)
    model._msd_traj0 = None


def allocate_trajectory(model, nr_of_steps):
    """Record the positions of every tracked particle for the next
    nr_of_steps steps run with traj_on."""
    b = _batch(model)
    b.msd_trajectory(int(nr_of_steps))
    st = b.msd_state(0)
    model._msd_traj0 = (int(nr_of_steps), st["site_arr"][: st["nr_of_ions"] + 1].copy())


def do_kmc_steps_displacement(model, n, traj_on=False):
    """proclist_acf.do_kmc_steps_displacement(n, traj_on); a float n is
    truncated like the f2py call."""
    _batch(model).run_displacement(int(n), [0], traj_on)
    model._kmc._check_status()


def _msd_state(model):
    return _batch(model).msd_state(0)


def get_displacement(model):
    """[1, nr_of_ions, 3] displacement of each tracked particle."""
    st = _msd_state(model)
    return st["displacement"][1: st["nr_of_ions"] + 1][None].copy()


def calc_msd(model):
    """calc_mean_squared_disp (base_acf.mpy:892-915): the particles' summed
    squared displacements, in particle order, over nr_of_ions."""
    st = _msd_state(model)
    n = st["nr_of_ions"]
    msd = 0.0
    for k in range(1, n + 1):
        d = st["displacement"][k]
        msd = msd + ((d[0] ** 2 + d[1] ** 2) + d[2] ** 2)
    return msd / n


def set_displacement_to_zero(model):
    _batch(model).msd_zero(0, 1)


def get_trajectory(model):
    """[1, nr_of_ions, nr_of_steps + 1] site of every tracked particle after
    each recorded step (allocate_trajectory's positions in column 0;
    update_trajectory, base_acf.mpy:825-830); columns of steps not run yet
    stay 0, as in the reference's zero-filled array."""
    if getattr(model, "_msd_traj0", None) is None:
        raise RuntimeError("allocate_trajectory first")
    nsteps, first = model._msd_traj0
    st = _msd_state(model)
    n = st["nr_of_ions"]
    traj = np.zeros((n + 1, nsteps + 1), np.int64)
    traj[: len(first), 0] = first
    cur = traj[:, 0].copy()
    ev = st["events"]
    for j in range(min(nsteps, len(ev))):
        p, s = int(ev[j, 0]), int(ev[j, 1])
        cur[p] = s
        cur[0] = 0
        traj[:, j + 1] = cur
    return traj[1:][None]


# ---- autocorrelation (kmos/run/acf.py:159-283) ------------------------------

def allocate_acf(model, nr_of_types, t_bin, t_f, safety_factor=None, extending_factor=None):
    """allocate_tracing_arr + allocate_config_bin_acf (base_acf.mpy:382-524)."""
    b = _batch(model)
    b.acf_alloc(int(nr_of_types), t_bin, t_f, int(safety_factor or 0), int(extending_factor or 0))
    V = int(b.volumes[0])
    model._acf = {"types": np.zeros(int(nr_of_types)), "property": np.zeros(V + 1, np.int32),
                  "product": np.zeros((int(nr_of_types) + 1, int(nr_of_types) + 1))}


def set_property_acf(model, layer_site_name, property_type):
    """Every site of type `layer_site_name` gets the property type index."""
    a = model._acf
    n = getattr(model.lattice, layer_site_name.lower())
    spuck = int(model.lattice.spuck)
    V = len(a["property"]) - 1
    for i in range(V):
        if (i + 1) % spuck + spuck == n or (i + 1) % spuck == n:
            a["property"][i + 1] = int(property_type)
    _batch(model).acf_set_properties([a["property"]], None)


def get_types_acf(model):
    return model._acf["types"][None].copy()


def set_types_acf(model, site_property):
    """The value of the first type index not set yet."""
    types = model._acf["types"]
    for i in range(len(types)):
        if types[i] == 0:
            types[i] = site_property
            break


def calc_product_property(model):
    """product_property(i, j) = types(i) * types(j) for every pair of types."""
    a = model._acf
    T = len(a["types"])
    for i in range(T):
        for j in range(T):
            a["product"][i + 1, j + 1] = a["types"][i] * a["types"][j]
    _batch(model).acf_set_properties(None, a["product"])


def get_product_property(model):
    return model._acf["product"][1:, 1:][None].copy()


def initialize_acf(model, species):
    """Number the traced particles and set their starting properties."""
    trace = getattr(model.proclist, species.lower())
    hops, _ = displacement_table(model.compiled)
    _batch(model).acf_init(int(trace), hops, 0, 1)
    model._msd_traj0 = None


def do_kmc_steps_acf(model, n, traj_on=False):
    _batch(model).run_acf(int(n), [0], traj_on)
    model._kmc._check_status()


def _acf_state(model):
    return _batch(model).acf_state(0)


def get_config_bin_acf(model):
    return _acf_state(model)["config_bin"][None].copy()


def get_counter_write_in_bin_acf(model):
    return _acf_state(model)["counter_write_in_bin"][None].copy()


def get_property_acf(model):
    return model._acf["property"][1:][None].copy()


def get_property_o(model):
    st = _acf_state(model)
    return st["property_o"][1: st["nr_of_ions"] + 1][None].copy()


def get_buffer_acf(model):
    st = _acf_state(model)
    return st["buffer_acf"][1: st["nr_of_ions"] + 1][None].copy()


def get_id_arr(model):
    return _msd_state(model)["id_arr"][1:][None].copy()


def get_site_arr(model):
    st = _msd_state(model)
    return st["site_arr"][1: st["nr_of_ions"] + 1][None].copy()


def get_acf(model, normalization=False):
    """calc_acf per bin (base_acf.mpy:917-935): config_bin over
    wrap_count * t_bin; normalised to the first bin when asked."""
    st = _acf_state(model)
    acf = np.zeros(st["nr_of_bins"])
    for i in range(st["nr_of_bins"]):
        acf[i] = st["config_bin"][i] / (st["wrap_count"] * st["t_bin"])
    if normalization:
        acf = acf / acf[0]
    return acf


def set_acf_to_zero(model):
    _batch(model).acf_zero(0, 1)
