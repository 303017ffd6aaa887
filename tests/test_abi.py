"""The C ABI library loads and exports every entry point of include/kmcg.h,
and its host-side validation works without a GPU (no kernel is launched)."""
import ctypes
import os
import re

import numpy as np
import pytest

from kmos_amd import _lib, workloads as W
from kmos_amd.compiler import compile_model
from paritylib import pdopd_project

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    import glob
    text = "".join(open(f).read() for f in sorted(glob.glob(os.path.join(REPO, "include", "*.h"))))
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(kmcg_\w+)\s*\(", text)))


def test_header_matches_binding_table():
    assert _declared() == sorted(name for name, _, _ in _lib.SIGNATURES)


def test_library_exports_every_symbol():
    lib = _lib.load()
    for name in _declared():
        assert hasattr(lib, name), name


def _load(cm):
    lib = _lib.load()
    blob = np.ascontiguousarray(cm.to_blob(), dtype=np.int32)
    h = ctypes.c_void_p()
    rc = lib.kmcg_model_load(blob.ctypes.data_as(ctypes.c_void_p), blob.size, ctypes.byref(h))
    return rc, h, blob


@pytest.mark.parametrize("factory,backend", [
    (W.ruo2_project, "local_smart"), (W.ruo2_project, "lat_int"),
    (W.zgb_project, "local_smart"), (W.pairwise_project, "lat_int"),
    (W.pairwise_project, "local_smart"), (W.quickstart_project, "local_smart"),
    (pdopd_project, "local_smart"), (pdopd_project, "lat_int"),
])
def test_model_load_on_host(factory, backend):
    cm = compile_model(factory(), backend)
    rc, h, _ = _load(cm)
    assert rc == 0, _lib.load().kmcg_last_error()
    info = np.zeros(5, np.int32)
    assert _lib.load().kmcg_model_info(h, info.ctypes.data_as(ctypes.c_void_p)) == 0
    assert list(info[:2]) == [cm.nr_of_proc, cm.nr_of_species]
    fp = ctypes.c_int64(0)
    sizes = np.array([[20, 20, 1], [64, 64, 1]], np.int32)
    assert _lib.load().kmcg_batch_footprint(h, 2, sizes.ctypes.data_as(ctypes.c_void_p), ctypes.byref(fp)) == 0
    V = 20 * 20 * cm.spuck + 64 * 64 * cm.spuck
    assert fp.value >= 2 * cm.nr_of_proc * V * 4
    _lib.load().kmcg_model_free(h)


def test_malformed_blobs_are_rejected():
    cm = compile_model(W.ruo2_project(), "lat_int")
    blob = cm.to_blob().copy()
    lib = _lib.load()
    h = ctypes.c_void_p()
    bad = blob.copy()
    bad[0] = 0
    assert lib.kmcg_model_load(bad.ctypes.data_as(ctypes.c_void_p), bad.size, ctypes.byref(h)) == -1
    bad = blob.copy()
    bad[14] = bad.size + 100          # group table out of range
    assert lib.kmcg_model_load(bad.ctypes.data_as(ctypes.c_void_p), bad.size, ctypes.byref(h)) == -1
    assert b"range" in lib.kmcg_last_error() or b"bad" in lib.kmcg_last_error()


def test_batch_create_without_gpu_fails_cleanly():
    if _lib.device_count() > 0:
        pytest.skip("a GPU is visible")
    cm = compile_model(W.zgb_project(), "local_smart")
    rc, h, _ = _load(cm)
    assert rc == 0
    b = ctypes.c_void_p()
    sizes = np.array([[16, 16, 1]], np.int32)
    rc = _lib.load().kmcg_batch_create(h, 1, sizes.ctypes.data_as(ctypes.c_void_p), 0, ctypes.byref(b))
    assert rc == -1 and b"device" in _lib.load().kmcg_last_error()


def test_many_process_models_load_or_are_rejected_cleanly():
    """72 and 136 processes load as local_smart (two and four processes per
    lane); as lat_int their op programs exceed 64 KiB and loading says so."""
    from test_gpu_many_procs import DIRS4, DIRS8, many_process_project
    for dirs, nproc in ((DIRS4, 72), (DIRS8, 136)):
        pt = many_process_project(dirs)
        cm = compile_model(pt, "local_smart")
        rc, h, _ = _load(cm)
        assert rc == 0 and cm.nr_of_proc == nproc, _lib.load().kmcg_last_error()
        _lib.load().kmcg_model_free(h)
        rc, _, _ = _load(compile_model(pt, "lat_int"))
        assert rc == -4 and b"64 KiB" in _lib.load().kmcg_last_error()


def test_lattice_species_checked_before_the_int8_cast():
    """Species ids outside 0..S-1 (0..S with a multi-layer model's null
    species) are refused before the cast to int8, so that 300 or 256
    cannot wrap into a valid id (44, 0)."""
    import numpy as np
    from kmos_amd.engine import Batch
    b = Batch.__new__(Batch)
    b.n, b.S, b.max_species, b.volumes = 2, 3, 2, np.array([4, 4], np.int64)
    ok = b._species_array([np.array([0, 1, 2, 0]), np.array([2, 2, 1, 0])], 0, 2)
    assert ok.dtype == np.int8 and ok.tolist() == [0, 1, 2, 0, 2, 2, 1, 0]
    for bad in ([300, 0, 0, 0], [256, 0, 0, 0], [-1, 0, 0, 0], [3, 0, 0, 0]):
        with pytest.raises(ValueError):
            b._species_array([np.array(bad)], 0, 1)
    with pytest.raises(ValueError):
        b._species_array([np.array([0.5, 0, 0, 0])], 0, 1)
    with pytest.raises(ValueError):
        b._species_array([np.array([0, 0, 0])], 0, 1)          # wrong size
    with pytest.raises(ValueError):
        b._species_array([np.zeros(4, int)], 1, 2)             # beyond the batch
    b.max_species = 3                                          # two layers: null = S = 3
    assert b._species_array([np.array([3, 0, 1, 2])], 0, 1).tolist() == [3, 0, 1, 2]
    with pytest.raises(ValueError):
        b._species_array([np.array([4, 0, 0, 0])], 0, 1)
