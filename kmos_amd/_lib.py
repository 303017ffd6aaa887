"""ctypes binding of libkmcgpu.so (the C ABI declared in include/kmcg.h and
include/kmcg_otf.h).

There is no CPU fallback: if the library is missing or no GPU is visible,
every engine call raises.
"""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libkmcgpu.so")
# the same engine built with its step-loop phase counters (KMCG_PHASES=1);
# loaded instead when the environment sets KMCG_PHASES=1
PHASES_PATH = os.path.join(HERE, "libkmcgpu_phases.so")

KMCG_OK = 0
ERRORS = {-1: "EINVAL", -2: "ENOMEM", -3: "EHIP", -4: "EUNSUPPORTED"}
ERRORS_BY_NAME = {v: k for k, v in ERRORS.items()}
REPLICA_STATUS = {0: "ok", 1: "absorbing", 2: "species_mismatch", 3: "del_absent", 4: "corrupt"}

# (name, restype, argtypes) of every entry point of include/*.h
_vp, _i32, _i64, _u32, _d = (ctypes.c_void_p, ctypes.c_int32, ctypes.c_int64,
                             ctypes.c_uint32, ctypes.c_double)
SIGNATURES = [
    ("kmcg_last_error", ctypes.c_char_p, []),
    ("kmcg_device_count", _i32, [_vp]),
    ("kmcg_model_load", _i32, [_vp, _i64, _vp]),
    ("kmcg_model_free", None, [_vp]),
    ("kmcg_model_info", _i32, [_vp, _vp]),
    ("kmcg_model_forms", _i32, [_vp, _vp]),
    ("kmcg_batch_footprint", _i32, [_vp, _i32, _vp, _vp]),
    ("kmcg_batch_create", _i32, [_vp, _i32, _vp, _i32, _vp]),
    ("kmcg_batch_free", None, [_vp]),
    ("kmcg_init", _i32, [_vp, _i32, _i32, _vp, _vp, _vp]),
    ("kmcg_set_rates", _i32, [_vp, _i32, _i32, _vp]),
    ("kmcg_get_rates", _i32, [_vp, _i32, _vp]),
    ("kmcg_set_config", _i32, [_vp, _i32, _i32, _vp]),
    ("kmcg_run", _i32, [_vp, _vp, _i32, _i64, _vp]),
    ("kmcg_sync", _i32, [_vp]),
    ("kmcg_set_launch", _i32, [_vp, _i32, _i32]),
    ("kmcg_set_overlap", _i32, [_vp, _i32]),
    ("kmcg_set_graph", _i32, [_vp, _i32]),
    ("kmcg_check_lists", _i32, [_vp, _i32, _i32, _vp]),
    ("kmcg_last_run_ms", _i32, [_vp, _vp]),
    ("kmcg_phase_profile", _i32, [_vp, _i32, _i32, _vp]),
    ("kmcg_get_next_step", _i32, [_vp, _i32, _vp, _vp]),
    ("kmcg_run_proc_nr", _i32, [_vp, _i32, _i32, _i32]),
    ("kmcg_touchup", _i32, [_vp, _i32, _i32, _i32, _i32, _i32]),
    ("kmcg_replace_species", _i32, [_vp, _i32, _i32, _i32, _i32]),
    ("kmcg_get_scalars", _i32, [_vp, _i32, _vp, _vp, _vp]),
    ("kmcg_get_scalars_all", _i32, [_vp, _vp, _vp, _vp]),
    ("kmcg_set_kmc_time", _i32, [_vp, _i32, _d]),
    ("kmcg_get_nr_of_sites", _i32, [_vp, _i32, _vp]),
    ("kmcg_get_procstat", _i32, [_vp, _i32, _vp]),
    ("kmcg_get_integ_rates", _i32, [_vp, _i32, _vp]),
    ("kmcg_update_integ_rate", _i32, [_vp, _i32]),
    ("kmcg_get_accum_rates", _i32, [_vp, _i32, _vp]),
    ("kmcg_get_lattice", _i32, [_vp, _i32, _vp]),
    ("kmcg_get_avail", _i32, [_vp, _i32, _vp, _vp]),
    ("kmcg_set_state", _i32, [_vp, _i32, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    ("kmcg_msd_init", _i32, [_vp, _i32, _i32, _i32, _vp, _vp]),
    ("kmcg_msd_trajectory", _i32, [_vp, _i64]),
    ("kmcg_run_displacement", _i32, [_vp, _vp, _i32, _i64, _vp, _i32]),
    ("kmcg_msd_zero", _i32, [_vp, _i32, _i32]),
    ("kmcg_msd_get", _i32, [_vp, _i32, _vp, _vp, _vp, _vp, _vp, _vp]),
    ("kmcg_acf_alloc", _i32, [_vp, _i32, _d, _d, _i32, _i32]),
    ("kmcg_acf_set_properties", _i32, [_vp, _i32, _i32, _vp, _vp]),
    ("kmcg_acf_init", _i32, [_vp, _i32, _i32, _i32, _vp]),
    ("kmcg_run_acf", _i32, [_vp, _vp, _i32, _i64, _vp, _i32]),
    ("kmcg_acf_zero", _i32, [_vp, _i32, _i32]),
    ("kmcg_acf_get", _i32, [_vp, _i32, _vp, _vp, _vp, _vp, _vp, _vp]),
    ("kmcg_get_occupation", _i32, [_vp, _i32, _i32, _vp]),
    ("kmcg_get_counters", _i32, [_vp, _i32, _i32, _vp, _vp]),
    ("kmcg_sample", _i32, [_vp, _vp, _i32, _vp]),
    ("kmcg_run_sampled", _i32, [_vp, _vp, _i32, _i64, _vp, _vp]),
    ("kmcg_query", _i32, [_vp, _vp]),
    ("kmcg_run_recorded", _i32, [_vp, _vp, _i32, _i64, _vp, _i32, _i32, _vp]),
    # include/kmcg_otf.h
    ("kmcg_otf_model_load", _i32, [_vp, _i64, _vp]),
    ("kmcg_otf_model_free", None, [_vp]),
    ("kmcg_otf_model_info", _i32, [_vp, _vp]),
    ("kmcg_otf_batch_create", _i32, [_vp, _i32, _vp, _i32, _vp]),
    ("kmcg_otf_batch_free", None, [_vp]),
    ("kmcg_otf_init", _i32, [_vp, _i32, _i32, _vp, _vp]),
    ("kmcg_otf_set_params", _i32, [_vp, _i32, _i32, _vp, _vp]),
    ("kmcg_otf_set_config", _i32, [_vp, _i32, _i32, _vp]),
    ("kmcg_otf_run", _i32, [_vp, _i32, _i32, _i64]),
    ("kmcg_otf_sync", _i32, [_vp]),
    ("kmcg_otf_last_run_ms", _i32, [_vp, _vp]),
    ("kmcg_otf_get_scalars", _i32, [_vp, _i32, _vp, _vp, _vp, _vp]),
    ("kmcg_otf_get_lattice", _i32, [_vp, _i32, _vp]),
    ("kmcg_otf_get_counters", _i32, [_vp, _i32, _vp, _vp, _vp, _vp]),
    ("kmcg_otf_get_avail", _i32, [_vp, _i32, _vp, _vp, _vp]),
    ("kmcg_otf_get_table", _i32, [_vp, _i32, _vp]),
]

_lib = None


class KmcgError(RuntimeError):
    pass


def load():
    """Load libkmcgpu.so; raise if it has not been built."""
    global _lib
    if _lib is None:
        path = PHASES_PATH if os.environ.get("KMCG_PHASES") == "1" else LIB_PATH
        if not os.path.exists(path):
            raise KmcgError("%s is not built (run __graft_entry__.build() "
                            "or python -m kmos_amd._build)" % os.path.basename(path))
        lib = ctypes.CDLL(path)
        for name, rt, args in SIGNATURES:
            f = getattr(lib, name)
            f.restype = rt
            f.argtypes = args
        _lib = lib
    return _lib


def check(rc):
    if rc != KMCG_OK:
        msg = load().kmcg_last_error().decode(errors="replace")
        raise KmcgError("%s: %s" % (ERRORS.get(rc, rc), msg))
    return rc


def device_count():
    c = ctypes.c_int32(0)
    load().kmcg_device_count(ctypes.byref(c))
    return c.value
