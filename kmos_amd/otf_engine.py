"""Python handles over the otf C ABI (include/kmcg_otf.h).

``OtfBatch`` holds replicas of one otf model on one GPU; every method maps
to one entry point.  Arrays use the reference's numbering (processes 1..P,
sites 1..V, species 0..S-1, parameters in name order).
"""
from __future__ import annotations

import ctypes
from typing import Optional, Sequence

import numpy as np

from . import _lib
from .compiler import BACKEND_OTF, CompiledModel

_vp = ctypes.c_void_p


def _ptr(a: Optional[np.ndarray]):
    return None if a is None else a.ctypes.data_as(_vp)


class OtfModel:
    """A compiled otf model loaded into the engine."""

    def __init__(self, compiled: CompiledModel):
        if compiled.backend != BACKEND_OTF:
            raise ValueError("OtfModel needs a model compiled for the otf backend")
        self.compiled = compiled
        self.blob = np.ascontiguousarray(compiled.to_blob(), dtype=np.int32)
        lib = _lib.load()
        h = _vp()
        _lib.check(lib.kmcg_otf_model_load(_ptr(self.blob), self.blob.size, ctypes.byref(h)))
        self.h = h
        info = np.zeros(4, np.int32)
        _lib.check(lib.kmcg_otf_model_info(self.h, _ptr(info)))
        self.nr_of_proc, self.nr_of_species, self.nr_of_params, self.table_size = (int(v) for v in info)
        self.spuck = compiled.spuck
        self.dimension = compiled.dimension
        self.param_names = list(compiled.param_names)

    def __del__(self):
        if getattr(self, "h", None) is not None and _lib._lib is not None:
            _lib._lib.kmcg_otf_model_free(self.h)
            self.h = None


class OtfBatch:
    """Replicas of one otf model resident on one GPU."""

    def __init__(self, model: OtfModel, sizes: Sequence[Sequence[int]], device: int = 0):
        self.model = model
        sz = np.ones((len(sizes), 3), np.int32)
        for i, s in enumerate(sizes):
            s = list(s)
            sz[i, :len(s)] = s[:3]
        if model.dimension < 3:
            sz[:, 2] = 1
        if model.dimension < 2:
            sz[:, 1] = 1
        self.sizes = sz
        self.n = len(sz)
        self.P = model.nr_of_proc
        self.S = model.nr_of_species
        self.K = model.nr_of_params
        self.volumes = sz.prod(axis=1).astype(np.int64) * model.spuck
        h = _vp()
        _lib.check(_lib.load().kmcg_otf_batch_create(model.h, self.n, _ptr(sz), device, ctypes.byref(h)))
        self.h = h

    def close(self):
        if getattr(self, "h", None) is not None and _lib._lib is not None:
            _lib._lib.kmcg_otf_batch_free(self.h)
            self.h = None

    __del__ = close

    def _range(self, r0, n):
        n = self.n - r0 if n is None else n
        if r0 < 0 or n < 1 or r0 + n > self.n:
            raise ValueError("replica range out of the batch")
        return n

    def init(self, seeds, streams=None, r0: int = 0, n: Optional[int] = None):
        n = self._range(r0, n)
        s = np.ascontiguousarray(np.broadcast_to(np.asarray(seeds, np.uint32), (n,)))
        t = np.ascontiguousarray(np.broadcast_to(np.asarray(0 if streams is None else streams, np.uint32), (n,)))
        _lib.check(_lib.load().kmcg_otf_init(self.h, r0, n, _ptr(s), _ptr(t)))

    def set_params(self, rates, userpar=None, r0: int = 0, n: Optional[int] = None):
        """rates [n][P] and userpar [n][K] (parameters in name order)."""
        n = self._range(r0, n)
        r = np.ascontiguousarray(np.broadcast_to(np.asarray(rates, np.float64), (n, self.P)))
        if self.K:
            if userpar is None:
                raise ValueError("this model has %d parameters" % self.K)
            u = np.ascontiguousarray(np.broadcast_to(np.asarray(userpar, np.float64), (n, self.K)))
        else:
            u = np.zeros(1, np.float64)
        _lib.check(_lib.load().kmcg_otf_set_params(self.h, r0, n, _ptr(r), _ptr(u)))

    def set_config(self, configs, r0: int = 0):
        """One species array per replica from r0 on."""
        n = self._range(r0, len(configs))
        parts = []
        for i, c in enumerate(configs):
            a = np.asarray(c)
            if a.size != self.volumes[r0 + i]:
                raise ValueError("configuration %d has %d sites, the replica %d" % (i, a.size, self.volumes[r0 + i]))
            if a.dtype.kind not in "iu" or a.min(initial=0) < 0 or a.max(initial=0) >= self.S:
                raise ValueError("species ids must be integers 0..%d" % (self.S - 1))
            parts.append(a.astype(np.int8).ravel())
        cfg = np.ascontiguousarray(np.concatenate(parts))
        _lib.check(_lib.load().kmcg_otf_set_config(self.h, r0, n, _ptr(cfg)))

    def run(self, steps: int, r0: int = 0, n: Optional[int] = None, sync: bool = True):
        n = self._range(r0, n)
        _lib.check(_lib.load().kmcg_otf_run(self.h, r0, n, int(steps)))
        if sync:
            self.sync()

    def sync(self):
        _lib.check(_lib.load().kmcg_otf_sync(self.h))

    def last_run_ms(self) -> float:
        ms = ctypes.c_float(0.0)
        _lib.check(_lib.load().kmcg_otf_last_run_ms(self.h, ctypes.byref(ms)))
        return ms.value

    # ---- state ---------------------------------------------------------------
    def scalars(self, r: int):
        """(kmc_time, kmc_step, status name)"""
        t, dt, k, s = self.scalars_full(r)
        return t, k, _lib.REPLICA_STATUS.get(s, str(s))

    def scalars_full(self, r: int):
        """(kmc_time, kmc_time_step, kmc_step, status code)"""
        t = ctypes.c_double(0.0)
        dt = ctypes.c_double(0.0)
        k = ctypes.c_int64(0)
        s = ctypes.c_int32(0)
        _lib.check(_lib.load().kmcg_otf_get_scalars(self.h, r, ctypes.byref(t), ctypes.byref(dt), ctypes.byref(k),
                                                    ctypes.byref(s)))
        return t.value, dt.value, k.value, s.value

    def lattice(self, r: int) -> np.ndarray:
        out = np.zeros(int(self.volumes[r]), np.int8)
        _lib.check(_lib.load().kmcg_otf_get_lattice(self.h, r, _ptr(out)))
        return out

    def counters(self, r: int):
        nr = np.zeros(self.P, np.int32)
        ps = np.zeros(self.P, np.int64)
        integ = np.zeros(self.P, np.float64)
        acc = np.zeros(self.P, np.float64)
        _lib.check(_lib.load().kmcg_otf_get_counters(self.h, r, _ptr(nr), _ptr(ps), _ptr(integ), _ptr(acc)))
        return {"nr_of_sites": nr, "procstat": ps, "integ_rates": integ, "accum_rates": acc}

    def avail(self, r: int):
        """(sites [P][V], back [P][V], rates_matrix [P][V+1])"""
        V = int(self.volumes[r])
        s = np.zeros((self.P, V), np.int32)
        b = np.zeros((self.P, V), np.int32)
        rm = np.zeros((self.P, V + 1), np.float64)
        _lib.check(_lib.load().kmcg_otf_get_avail(self.h, r, _ptr(s), _ptr(b), _ptr(rm)))
        return s, b, rm

    def table(self, r: int) -> np.ndarray:
        out = np.zeros(max(self.model.table_size, 1), np.float64)
        _lib.check(_lib.load().kmcg_otf_get_table(self.h, r, _ptr(out)))
        return out[:self.model.table_size]
