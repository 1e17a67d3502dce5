"""Ensemble serving over libmhit: many independent periodic boxes resident on
one GPU, admitted by an HBM budget and stepped concurrently."""
import ctypes

import numpy as np

from . import _lib
from ._lib import Params, check
from .solver import default_params


class EnsConfig(ctypes.Structure):
    _fields_ = [("hbm_budget_bytes", ctypes.c_size_t), ("max_active", ctypes.c_int),
                ("inflight", ctypes.c_int), ("use_graphs", ctypes.c_int), ("autotune", ctypes.c_int),
                ("gang_size", ctypes.c_int), ("gang_max_nx", ctypes.c_int)]


class EnsResult(ctypes.Structure):
    _fields_ = [("id", ctypes.c_longlong), ("status", ctypes.c_int), ("steps_done", ctypes.c_int),
                ("nx", ctypes.c_int), ("reserved", ctypes.c_int), ("submit_ms", ctypes.c_double),
                ("admit_ms", ctypes.c_double), ("finish_ms", ctypes.c_double), ("gumax", ctypes.c_double),
                ("cou", ctypes.c_double)]


def _bind(lib):
    if getattr(lib, "_ens_bound", False):
        return
    P = ctypes.POINTER
    vp = ctypes.c_void_p
    dp = ctypes.c_void_p
    sig = {
        "mhit_ens_create": ([P(EnsConfig), P(vp)], ctypes.c_int),
        "mhit_ens_destroy": ([vp], ctypes.c_int),
        "mhit_ens_submit": ([vp, P(Params), dp, dp, dp, dp, ctypes.c_int, P(ctypes.c_longlong)], ctypes.c_int),
        "mhit_ens_progress": ([vp, P(ctypes.c_int), P(ctypes.c_int)], ctypes.c_int),
        "mhit_ens_wait_idle": ([vp], ctypes.c_int),
        "mhit_ens_results": ([vp, P(EnsResult), ctypes.c_int, P(ctypes.c_int)], ctypes.c_int),
        "mhit_ens_get_field": ([vp, ctypes.c_longlong, ctypes.c_int, dp], ctypes.c_int),
        "mhit_ens_release": ([vp, ctypes.c_longlong], ctypes.c_int),
        "mhit_ens_latency": ([vp, P(ctypes.c_double), P(ctypes.c_double), P(ctypes.c_double),
                              P(ctypes.c_longlong)], ctypes.c_int),
    }
    for name, (args, res) in sig.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = res
    lib._ens_bound = True


class Ensemble:
    def __init__(self, hbm_budget_bytes=0, max_active=64, inflight=2, use_graphs=True, autotune=True,
                 gang_size=8, gang_max_nx=128):
        """gang_size: boxes of one (nx, precision, phiflag) with nx <= gang_max_nx
        stepped together by one launch per pass (1: every box on its own)."""
        self._lib = _lib.load()
        _bind(self._lib)
        cfg = EnsConfig(hbm_budget_bytes, max_active, inflight, 1 if use_graphs else 0, 1 if autotune else 0,
                        gang_size, gang_max_nx)
        h = ctypes.c_void_p()
        check(self._lib.mhit_ens_create(ctypes.byref(cfg), ctypes.byref(h)), "mhit_ens_create")
        self._h = h
        self._n = {}

    def close(self):
        if getattr(self, "_h", None):
            self._lib.mhit_ens_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def submit(self, params=None, nsteps=1, u=None, v=None, w=None, phi=None, **overrides):
        if params is None or isinstance(params, dict):
            params = default_params(**{**(params or {}), **overrides})
        keep = []

        def ptr(a):
            if a is None:
                return None
            a = np.asfortranarray(a, dtype=np.float64)
            keep.append(a)
            return a.ctypes.data_as(ctypes.c_void_p)

        rid = ctypes.c_longlong()
        check(self._lib.mhit_ens_submit(self._h, ctypes.byref(params), ptr(u), ptr(v), ptr(w), ptr(phi),
                                        int(nsteps), ctypes.byref(rid)), "mhit_ens_submit")
        self._n[rid.value] = params.nx
        return rid.value

    def progress(self):
        a, q = ctypes.c_int(), ctypes.c_int()
        check(self._lib.mhit_ens_progress(self._h, ctypes.byref(a), ctypes.byref(q)), "mhit_ens_progress")
        return a.value, q.value

    def wait_idle(self):
        check(self._lib.mhit_ens_wait_idle(self._h), "mhit_ens_wait_idle")

    def results(self):
        arr = (EnsResult * 256)()
        out = []
        while True:
            k = ctypes.c_int()
            check(self._lib.mhit_ens_results(self._h, arr, 256, ctypes.byref(k)), "mhit_ens_results")
            out += [{name: getattr(r, name) for name, _ in EnsResult._fields_} for r in arr[:k.value]]
            if k.value < 256:
                return out

    def get_field(self, rid, fieldn):
        n = self._n[rid]
        out = np.empty((n, n, n), dtype=np.float64, order="F")
        check(self._lib.mhit_ens_get_field(self._h, int(rid), int(fieldn), out.ctypes.data_as(ctypes.c_void_p)),
              "mhit_ens_get_field")
        return out

    def release(self, rid):
        check(self._lib.mhit_ens_release(self._h, int(rid)), "mhit_ens_release")
        self._n.pop(rid, None)

    def latency(self):
        p50, p99, mean = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
        cnt = ctypes.c_longlong()
        check(self._lib.mhit_ens_latency(self._h, ctypes.byref(p50), ctypes.byref(p99), ctypes.byref(mean),
                                         ctypes.byref(cnt)), "mhit_ens_latency")
        return {"p50_ms": p50.value, "p99_ms": p99.value, "mean_ms": mean.value, "steps": cnt.value}
