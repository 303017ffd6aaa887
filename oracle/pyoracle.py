"""ctypes wrapper of the C oracle (oracle/kmc_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
the cpu_baseline leg of bench.py, never by the engine.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libkmc_oracle.so")
# timing build: -O3, no touched-entry counter (the CPU baseline)
FAST_PATH = os.path.join(HERE, "libkmc_oracle_fast.so")
SRC_PATH = os.path.join(HERE, "kmc_oracle.c")

STATUS = {0: "ok", 1: "absorbing", 2: "species_mismatch", 3: "del_absent", 4: "bad_arg"}


def build(force=False):
    """Compile the oracle with gcc (no FMA contraction, like the reference's
    x86-64 build), and its timing variant (-O3, no entry counter)."""
    for path, flags in ((LIB_PATH, ["-O2"]), (FAST_PATH, ["-O3", "-DORC_NO_TOUCH"])):
        if force or not os.path.exists(path) or os.path.getmtime(path) < os.path.getmtime(SRC_PATH):
            subprocess.check_call(["gcc"] + flags + ["-ffp-contract=off", "-fPIC", "-shared",
                                                     SRC_PATH, "-o", path, "-lm"])
    return LIB_PATH


_libs = {}


def lib(fast=False):
    """The oracle library; `fast` selects the timing build, whose
    touched-entry counter always reads 0."""
    if fast not in _libs:
        path = FAST_PATH if fast else LIB_PATH
        if not os.path.exists(path):
            build()
        L = ctypes.CDLL(path)
        vp, i32, i64, u32, u64, dbl = (ctypes.c_void_p, ctypes.c_int32, ctypes.c_int64,
                                       ctypes.c_uint32, ctypes.c_uint64, ctypes.c_double)
        L.orc_create.restype = vp
        L.orc_create.argtypes = [vp, i64, i32, i32, i32]
        L.orc_free.argtypes = [vp]
        L.orc_init.argtypes = [vp, u32, u32]
        L.orc_init.restype = i32
        L.orc_init_config.argtypes = [vp, u32, u32, vp]
        L.orc_init_config.restype = i32
        L.orc_set_config.argtypes = [vp, vp]
        L.orc_set_config.restype = i32
        L.orc_set_rates.argtypes = [vp, vp]
        L.orc_run.argtypes = [vp, i64, vp, vp, vp]
        L.orc_run.restype = i64
        L.orc_get_next_step.argtypes = [vp, vp, vp]
        L.orc_get_next_step.restype = i32
        L.orc_run_proc_nr.argtypes = [vp, i32, i32]
        L.orc_run_proc_nr.restype = i32
        L.orc_touchup.argtypes = [vp, i32, i32, i32, i32]
        L.orc_touchup.restype = i32
        for name, rt in (("orc_status", i32), ("orc_volume", i32), ("orc_nr_of_proc", i32),
                         ("orc_kmc_time", dbl), ("orc_kmc_time_step", dbl),
                         ("orc_kmc_step", i64), ("orc_rng_ctr", u64), ("orc_touched", u64)):
            getattr(L, name).argtypes = [vp]
            getattr(L, name).restype = rt
        L.orc_reset_touched.argtypes = [vp]
        L.orc_set_kmc_time.argtypes = [vp, dbl]
        for name in ("orc_get_lattice", "orc_get_nr", "orc_get_procstat", "orc_get_integ",
                     "orc_get_accum", "orc_get_occupation"):
            getattr(L, name).argtypes = [vp, vp]
        L.orc_get_avail.argtypes = [vp, vp, vp]
        L.orc_set_state.argtypes = [vp, dbl, dbl, i64, u64, vp, vp, vp, vp, vp, vp, vp]
        L.orc_msd_init.argtypes = [vp, i32, vp, vp]
        L.orc_msd_trajectory.argtypes = [vp, i64]
        L.orc_msd_zero.argtypes = [vp]
        L.orc_msd_get.argtypes = [vp, vp, vp, vp, vp, vp, vp]
        L.orc_run_displacement.argtypes = [vp, i64, i32]
        L.orc_run_displacement.restype = i64
        L.orc_acf_alloc.argtypes = [vp, i32, dbl, dbl, i32, i32]
        L.orc_acf_set.argtypes = [vp, vp, vp]
        L.orc_acf_init.argtypes = [vp, i32, vp]
        L.orc_acf_zero.argtypes = [vp]
        L.orc_run_acf.argtypes = [vp, i64, i32]
        L.orc_run_acf.restype = i64
        L.orc_acf_get.argtypes = [vp, vp, vp, vp, vp, vp, vp]
        L.orc_get_species.argtypes = [vp, i32]
        L.orc_get_species.restype = i32
        L.orc_replace_species.argtypes = [vp, i32, i32, i32]
        L.orc_replace_species.restype = i32
        L.orc_step_uniforms.argtypes = [u32, u32, u64, vp]
        L.orc_philox.argtypes = [vp, vp, vp]
        L.orc_nr2lattice.argtypes = [vp, i32, vp]
        L.orc_lattice2nr.argtypes = [vp, i32, i32, i32, i32]
        L.orc_lattice2nr.restype = i32
        _libs[fast] = L
    return _libs[fast]


def _ptr(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def philox4x32_10(ctr, key):
    c = np.asarray(ctr, dtype=np.uint32)
    k = np.asarray(key, dtype=np.uint32)
    out = np.zeros(4, dtype=np.uint32)
    lib().orc_philox(_ptr(c), _ptr(k), _ptr(out))
    return out


def step_uniforms(seed, stream, step):
    u = np.zeros(3, dtype=np.float64)
    lib().orc_step_uniforms(seed, stream, step, _ptr(u))
    return u


class Oracle:
    """One replica of the reference step loop on the CPU."""

    def __init__(self, blob, size, fast=False):
        self.L = lib(fast)
        blob = np.ascontiguousarray(blob, dtype=np.int32)
        X, Y = int(size[0]), int(size[1]) if len(size) > 1 else 1
        Z = int(size[2]) if len(size) > 2 else 1
        self._blob = blob
        self.h = self.L.orc_create(_ptr(blob), blob.size, X, Y, Z)
        if not self.h:
            raise ValueError("bad model blob")
        self.P = self.L.orc_nr_of_proc(self.h)
        self.V = self.L.orc_volume(self.h)
        self.S = int(blob[4])
        self.spuck = int(blob[6])

    def __del__(self):
        if getattr(self, "h", None):
            self.L.orc_free(self.h)
            self.h = None

    def init(self, seed, stream=0, config=None):
        if config is None:
            return self.L.orc_init(self.h, seed & 0xFFFFFFFF, stream & 0xFFFFFFFF)
        c = np.ascontiguousarray(config, dtype=np.int8)
        assert c.size == self.V
        return self.L.orc_init_config(self.h, seed & 0xFFFFFFFF, stream & 0xFFFFFFFF, _ptr(c))

    def set_rates(self, rates):
        r = np.ascontiguousarray(rates, dtype=np.float64)
        assert r.size == self.P
        self.L.orc_set_rates(self.h, _ptr(r))

    def set_config(self, config):
        c = np.ascontiguousarray(config, dtype=np.int8)
        assert c.size == self.V
        return self.L.orc_set_config(self.h, _ptr(c))

    def run(self, n, trace=False):
        if trace:
            p = np.zeros(n, np.int32)
            s = np.zeros(n, np.int32)
            t = np.zeros(n, np.float64)
            k = self.L.orc_run(self.h, n, _ptr(p), _ptr(s), _ptr(t))
            return k, p[:k], s[:k], t[:k]
        return self.L.orc_run(self.h, n, None, None, None)

    def get_next_step(self):
        p = np.zeros(1, np.int32)
        s = np.zeros(1, np.int32)
        self.L.orc_get_next_step(self.h, _ptr(p), _ptr(s))
        return int(p[0]), int(s[0])

    def run_proc_nr(self, proc, site):
        return self.L.orc_run_proc_nr(self.h, proc, site)

    def touchup(self, x, y, z, n):
        return self.L.orc_touchup(self.h, x, y, z, n)

    @property
    def status(self):
        return self.L.orc_status(self.h)

    @property
    def kmc_time(self):
        return self.L.orc_kmc_time(self.h)

    @property
    def kmc_step(self):
        return self.L.orc_kmc_step(self.h)

    @property
    def rng_ctr(self):
        return self.L.orc_rng_ctr(self.h)

    @property
    def touched(self):
        return self.L.orc_touched(self.h)

    def reset_touched(self):
        self.L.orc_reset_touched(self.h)

    def lattice(self):
        out = np.zeros(self.V, np.int8)
        self.L.orc_get_lattice(self.h, _ptr(out))
        return out

    def nr_of_sites(self):
        out = np.zeros(self.P, np.int32)
        self.L.orc_get_nr(self.h, _ptr(out))
        return out

    def procstat(self):
        out = np.zeros(self.P, np.int64)
        self.L.orc_get_procstat(self.h, _ptr(out))
        return out

    def integ_rates(self):
        out = np.zeros(self.P, np.float64)
        self.L.orc_get_integ(self.h, _ptr(out))
        return out

    def accum_rates(self):
        out = np.zeros(self.P, np.float64)
        self.L.orc_get_accum(self.h, _ptr(out))
        return out

    def occupation(self):
        out = np.zeros(self.S * self.spuck, np.float64)
        self.L.orc_get_occupation(self.h, _ptr(out))
        return out.reshape(self.S, self.spuck)

    def avail(self):
        sites = np.zeros((self.P, self.V), np.int32)
        back = np.zeros((self.P, self.V), np.int32)
        self.L.orc_get_avail(self.h, _ptr(sites), _ptr(back))
        return sites, back

    def get_state(self, rates):
        """The fields of a .reload snapshot (rates: the ones last set)."""
        sites, back = self.avail()
        return {"kmc_time": self.kmc_time, "kmc_time_step": self.L.orc_kmc_time_step(self.h),
                "kmc_step": self.kmc_step, "rng_counter": self.rng_ctr, "lattice": self.lattice(),
                "nr_of_sites": self.nr_of_sites(), "procstat": self.procstat(),
                "integ_rates": self.integ_rates(), "rates": np.asarray(rates, np.float64),
                "avail_sites": sites, "avail_sites_back": back}

    # displacement tracking ------------------------------------------------
    def msd_init(self, trace_species, hops, disp):
        self._hops = np.ascontiguousarray(hops, np.int32)
        self._disp = np.ascontiguousarray(disp, np.float64)
        self.L.orc_msd_init(self.h, int(trace_species), _ptr(self._hops), _ptr(self._disp))

    def msd_trajectory(self, n):
        self.L.orc_msd_trajectory(self.h, int(n))

    def msd_zero(self):
        self.L.orc_msd_zero(self.h)

    def run_displacement(self, n, traj_on=False):
        return self.L.orc_run_displacement(self.h, int(n), 1 if traj_on else 0)

    def msd_state(self):
        ni = np.zeros(1, np.int32)
        ne = np.zeros(1, np.int64)
        sa = np.zeros(self.V + 1, np.int32)
        ia = np.zeros(self.V + 1, np.int32)
        dp = np.zeros((self.V + 1, 3), np.float64)
        self.L.orc_msd_get(self.h, _ptr(ni), _ptr(sa), _ptr(ia), _ptr(dp), _ptr(ne), None)
        ev = np.zeros((max(int(ne[0]), 1), 2), np.int32)
        self.L.orc_msd_get(self.h, _ptr(ni), _ptr(sa), _ptr(ia), _ptr(dp), _ptr(ne), _ptr(ev))
        return {"nr_of_ions": int(ni[0]), "site_arr": sa, "id_arr": ia, "displacement": dp,
                "events": ev[: int(ne[0])]}

    # autocorrelation bins ---------------------------------------------------
    def acf_alloc(self, nr_types, t_bin, t_f, safety=0, extending=0):
        self.L.orc_acf_alloc(self.h, int(nr_types), float(t_bin), float(t_f), int(safety), int(extending))
        self._nt1 = int(nr_types) + 1
        self._ext = int(float(t_f) * (safety or 2) / float(t_bin) + 1) * (extending or 2)

    def acf_set(self, property_acf=None, product_property=None):
        self._pa = None if property_acf is None else np.ascontiguousarray(property_acf, np.int32)
        self._pp = None if product_property is None else np.ascontiguousarray(product_property, np.float64)
        assert self._pa is None or self._pa.size == self.V + 1
        self.L.orc_acf_set(self.h, _ptr(self._pa), _ptr(self._pp))

    def acf_init(self, trace_species, hops):
        self._hops = np.ascontiguousarray(hops, np.int32)
        self.L.orc_acf_init(self.h, int(trace_species), _ptr(self._hops))

    def acf_zero(self):
        self.L.orc_acf_zero(self.h)

    def run_acf(self, n, traj_on=False):
        return self.L.orc_run_acf(self.h, int(n), 1 if traj_on else 0)

    def acf_state(self):
        i5 = np.zeros(5, np.int32)
        d4 = np.zeros(4, np.float64)
        cb = np.zeros(self._ext, np.float64)
        cc = np.zeros(self._ext, np.int32)
        buf = np.zeros(self.V + 1, np.float64)
        po = np.zeros(self.V + 1, np.int32)
        self.L.orc_acf_get(self.h, _ptr(cb), _ptr(cc), _ptr(d4), _ptr(i5), _ptr(buf), _ptr(po))
        ext = int(i5[2])
        assert ext == self._ext
        return {"config_bin": cb, "counter_write_in_bin": cc,
                "sum_buffer": float(d4[0]), "t_o": float(d4[1]), "wrap_count": float(d4[2]),
                "t_bin": float(d4[3]), "bin_index": int(i5[0]), "nr_of_bins": int(i5[1]),
                "extended_nr_of_bins": ext, "nr_of_ions_before": int(i5[3]), "nr_of_ions": int(i5[4]),
                "buffer_acf": buf, "property_o": po}

    def set_state(self, st):
        """reload_system: overwrite the state (rng_counter defaults to kmc_step)."""
        c = lambda k, t: np.ascontiguousarray(st[k], dtype=t)
        ig = st.get("integ_rates")
        ig = None if ig is None else np.ascontiguousarray(ig, np.float64)
        keep = [c("lattice", np.int8), c("nr_of_sites", np.int32), c("procstat", np.int64),
                c("rates", np.float64), ig, c("avail_sites", np.int32), c("avail_sites_back", np.int32)]
        assert keep[0].size == self.V and keep[5].size == self.P * self.V
        ctr = st.get("rng_counter")
        self.L.orc_set_state(self.h, float(st["kmc_time"]), float(st.get("kmc_time_step", 0.0)),
                            int(st["kmc_step"]), int(st["kmc_step"] if ctr is None else ctr),
                            *[_ptr(a) for a in keep])


# ---------------------------------------------------------------------------
# otf (oracle/otf_oracle.c)
# ---------------------------------------------------------------------------

OTF_SRC = os.path.join(HERE, "otf_oracle.c")
OTF_LIB = os.path.join(HERE, "libotf_oracle.so")
OTF_FAST = os.path.join(HERE, "libotf_oracle_fast.so")
OTF_STATUS = {0: "ok", 1: "absorbing", 2: "species_mismatch", 3: "del_absent", 4: "bad_arg",
              5: "search"}


def build_otf(force=False):
    for path, opt in ((OTF_LIB, "-O2"), (OTF_FAST, "-O3")):
        if force or not os.path.exists(path) or os.path.getmtime(path) < os.path.getmtime(OTF_SRC):
            subprocess.check_call(["gcc", opt, "-ffp-contract=off", "-fPIC", "-shared",
                                   OTF_SRC, "-o", path, "-lm"])
    return OTF_LIB


_otf_libs = {}


def otf_lib(fast=False):
    if fast not in _otf_libs:
        path = OTF_FAST if fast else OTF_LIB
        if not os.path.exists(path):
            build_otf()
        L = ctypes.CDLL(path)
        vp, i32, i64, u32, u64, dbl = (ctypes.c_void_p, ctypes.c_int32, ctypes.c_int64,
                                       ctypes.c_uint32, ctypes.c_uint64, ctypes.c_double)
        L.otf_create.restype = vp
        L.otf_create.argtypes = [vp, i64, i32, i32, i32]
        L.otf_free.argtypes = [vp]
        L.otf_init.argtypes = [vp, u32, u32]
        L.otf_set_params.argtypes = [vp, vp, vp]
        L.otf_set_config.argtypes = [vp, vp]
        L.otf_set_config.restype = i32
        L.otf_update_accum_rate.argtypes = [vp]
        L.otf_run.argtypes = [vp, i64, vp, vp, vp]
        L.otf_run.restype = i64
        for name, rt in (("otf_status", i32), ("otf_volume", i32), ("otf_kmc_time", dbl),
                         ("otf_kmc_step", i64), ("otf_rng_ctr", u64)):
            getattr(L, name).argtypes = [vp]
            getattr(L, name).restype = rt
        for name in ("otf_get_lattice", "otf_get_nr", "otf_get_procstat", "otf_get_integ",
                     "otf_get_accum", "otf_get_table"):
            getattr(L, name).argtypes = [vp, vp]
        L.otf_get_avail.argtypes = [vp, vp, vp]
        L.otf_set_clock.argtypes = [vp, dbl, i64, u64]
        L.otf_work.argtypes = [vp, vp]
        _otf_libs[fast] = L
    return _otf_libs[fast]


class OtfOracle:
    """One replica of an otf model on the CPU (oracle/otf_oracle.c)."""

    def __init__(self, blob, size, fast=False):
        self.L = otf_lib(fast)
        self.blob = np.ascontiguousarray(blob, dtype=np.int32)
        X, Y = int(size[0]), int(size[1])
        Z = int(size[2]) if len(size) > 2 else 1
        self.h = self.L.otf_create(self.blob.ctypes.data_as(ctypes.c_void_p), self.blob.size, X, Y, Z)
        if not self.h:
            raise ValueError("not an otf model blob")
        self.P = int(self.blob[3])
        self.K = int(self.blob[10])
        self.V = self.L.otf_volume(self.h)

    def __del__(self):
        if getattr(self, "h", None):
            self.L.otf_free(self.h)
            self.h = None

    def init(self, seed, stream=0):
        self.L.otf_init(self.h, seed, stream)

    def set_params(self, rates, userpar=()):
        r = np.ascontiguousarray(rates, dtype=np.float64)
        u = np.ascontiguousarray(userpar if len(userpar) else [0.0], dtype=np.float64)
        assert r.size == self.P and (len(userpar) == self.K)
        self.L.otf_set_params(self.h, r.ctypes.data_as(ctypes.c_void_p), u.ctypes.data_as(ctypes.c_void_p))

    def set_config(self, config):
        c = np.ascontiguousarray(config, dtype=np.int8)
        assert c.size == self.V
        return self.L.otf_set_config(self.h, c.ctypes.data_as(ctypes.c_void_p))

    def run(self, n, trace=False):
        if trace:
            tp = np.zeros(n, np.int32)
            ts = np.zeros(n, np.int32)
            tt = np.zeros(n, np.float64)
            k = self.L.otf_run(self.h, n, tp.ctypes.data_as(ctypes.c_void_p),
                               ts.ctypes.data_as(ctypes.c_void_p), tt.ctypes.data_as(ctypes.c_void_p))
            return k, tp[:k], ts[:k], tt[:k]
        return self.L.otf_run(self.h, n, None, None, None)

    def update_accum_rate(self):
        self.L.otf_update_accum_rate(self.h)

    @property
    def status(self):
        return OTF_STATUS.get(self.L.otf_status(self.h), "?")

    @property
    def kmc_time(self):
        return self.L.otf_kmc_time(self.h)

    @property
    def kmc_step(self):
        return self.L.otf_kmc_step(self.h)

    @property
    def rng_ctr(self):
        return self.L.otf_rng_ctr(self.h)

    def _get(self, name, dtype, n):
        a = np.zeros(n, dtype)
        getattr(self.L, name)(self.h, a.ctypes.data_as(ctypes.c_void_p))
        return a

    def lattice(self):
        return self._get("otf_get_lattice", np.int8, self.V)

    def nr_of_sites(self):
        return self._get("otf_get_nr", np.int32, self.P)

    def procstat(self):
        return self._get("otf_get_procstat", np.int64, self.P)

    def integ_rates(self):
        return self._get("otf_get_integ", np.float64, self.P)

    def accum_rates(self):
        return self._get("otf_get_accum", np.float64, self.P)

    def table(self):
        return self._get("otf_get_table", np.float64, int(self.blob[13]))

    def work(self):
        """(list operations, re-added running-sum entries, binary-search
        probes) since creation."""
        return tuple(int(v) for v in self._get("otf_work", np.uint64, 3))

    def avail(self):
        """(sites [P][V], back [P][V], rates_matrix [P][V+1])"""
        av = np.zeros((self.P, self.V, 2), np.int32)
        rm = np.zeros((self.P, self.V + 1), np.float64)
        self.L.otf_get_avail(self.h, av.ctypes.data_as(ctypes.c_void_p), rm.ctypes.data_as(ctypes.c_void_p))
        return av[:, :, 0].copy(), av[:, :, 1].copy(), rm
