"""Host-side mirror of the reference driver (multi/main.f90) over libmhit.

Arrays follow readwrite.f90's layout: an interior field is N x N x N FP64 in
Fortran order, exposed here as a numpy array of shape (N, N, N) with
order='F', so that a[i, j, k] is field(i+1, j+1, k+1) of the reference.
"""
import ctypes
import os

import numpy as np

from . import _lib
from ._lib import Params, Consts, Diag, KStat, Stats, Tuning, MHIT_TUNE_FORCE, MHIT_BATCH_MAX, check, MhitError

FIELD_NAMES = {1: "u", 2: "v", 3: "w", 4: "p", 5: "phi"}   # readwrite.f90:21-25


def default_params(**overrides):
    """mhit_params with multi/input.inp defaults, then `overrides`."""
    lib = _lib.load()
    p = Params()
    check(lib.mhit_default_params(ctypes.byref(p)), "mhit_default_params")
    for k, v in overrides.items():
        if not hasattr(p, k):
            raise KeyError(k)
        setattr(p, k, v)
    return p


def read_input(path, nx, **overrides):
    """readinput (readinput.f90:11-31) of `path`, with nx given at run time."""
    lib = _lib.load()
    p = Params()
    check(lib.mhit_read_input(os.fsencode(path), int(nx), ctypes.byref(p)), f"read_input({path})")
    for k, v in overrides.items():
        setattr(p, k, v)
    return p


def derive(params):
    """Derived constants of readinput.f90:35-43."""
    lib = _lib.load()
    c = Consts()
    check(lib.mhit_derive(ctypes.byref(params), ctypes.byref(c)), "mhit_derive")
    return {name: getattr(c, name) for name, _ in Consts._fields_}


def debug_math(op, x):
    """Device evaluation of log (0), tanh (1), sqrt (2) or 1/x (3)."""
    lib = _lib.load()
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.empty_like(x)
    check(lib.mhit_debug_math(int(op), x.size, x.ctypes.data_as(ctypes.c_void_p),
                              y.ctypes.data_as(ctypes.c_void_p)), "debug_math")
    return y


class Solver:
    """One periodic box resident on the GPU: fields, FFT plan and state."""

    def __init__(self, params=None, **overrides):
        lib = _lib.load()
        if params is None:
            params = default_params(**overrides)
        elif isinstance(params, dict):
            params = default_params(**{**params, **overrides})
        else:
            for k, v in overrides.items():
                setattr(params, k, v)
        self.params = params
        self.n = params.nx
        self._lib = lib
        h = ctypes.c_void_p()
        check(lib.mhit_create(ctypes.byref(params), ctypes.byref(h)), "mhit_create")
        self._h = h
        self.consts = derive(params)
        self.last = None

    # -- lifetime -----------------------------------------------------------
    def close(self):
        if getattr(self, "_h", None):
            self._lib.mhit_destroy(self._h)
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

    # -- fields ---------------------------------------------------------------
    def init_fields(self):
        """Taylor-Green / spherical drop initial condition (main.f90:218-278)."""
        check(self._lib.mhit_init_fields(self._h), "mhit_init_fields")

    def set_field(self, fieldn, a):
        a = np.asarray(a, dtype=np.float64)
        n = self.n
        if a.size != n ** 3:
            raise ValueError(f"field must hold {n}^3 values, got {a.size}")
        buf = np.asfortranarray(a.reshape((n, n, n), order="F") if a.ndim != 3 else a)
        check(self._lib.mhit_set_field(self._h, int(fieldn), buf.ctypes.data_as(ctypes.c_void_p)),
              f"set_field({fieldn})")

    def get_field(self, fieldn):
        n = self.n
        out = np.empty((n, n, n), dtype=np.float64, order="F")
        check(self._lib.mhit_get_field(self._h, int(fieldn), out.ctypes.data_as(ctypes.c_void_p)),
              f"get_field({fieldn})")
        return out

    # -- time stepping ----------------------------------------------------------
    def step(self, nsteps=1, raise_on_stop=True):
        """Advance `nsteps` timesteps (main.f90:329-946 each)."""
        d = Diag()
        rc = self._lib.mhit_step(self._h, int(nsteps), ctypes.byref(d))
        self.last = d
        if rc != _lib.MHIT_OK and (raise_on_stop or rc not in (_lib.MHIT_ERR_CFL, _lib.MHIT_ERR_NAN)):
            raise MhitError(rc, "mhit_step")
        return d

    def init_random_velocity(self, seed=1234, k_peak=4.0, urms=1.0):
        """Random-phase, discretely divergence-free turbulent start made on the
        device (mhit_init_random_velocity): E(k) ~ k^4 exp(-2 (k/k_peak)^2),
        scaled to urms."""
        check(self._lib.mhit_init_random_velocity(self._h, int(seed), float(k_peak), float(urms)),
              "init_random_velocity")

    def reset_time(self, tstart):
        check(self._lib.mhit_reset_time(self._h, int(tstart)), "reset_time")

    # -- files (readwrite.f90) ------------------------------------------------
    def write_field(self, directory, t, fieldn):
        check(self._lib.mhit_write_field(self._h, os.fsencode(directory), int(t), int(fieldn)),
              f"write_field({fieldn})")

    def read_field(self, path, fieldn):
        check(self._lib.mhit_read_field(self._h, os.fsencode(path), int(fieldn)), f"read_field({path})")

    # -- diagnostics ------------------------------------------------------------
    def diagnostics(self, nu=None, dxi=0.0):
        """Kinetic energy, dissipation 2 nu <S:S> (compute_diss formula) and the
        phase sum of the current fields, reduced on the device."""
        st = Stats()
        nu = self.params.mu / self.params.rho if nu is None else nu
        check(self._lib.mhit_diagnostics(self._h, float(dxi), float(nu), ctypes.byref(st)), "diagnostics")
        return {name: getattr(st, name) for name, _ in Stats._fields_}

    def save_state(self, path):
        """Checkpoint the exact device state (mhit_save_state)."""
        check(self._lib.mhit_save_state(self._h, str(path).encode()), "save_state")

    def load_state(self, path):
        """Resume from mhit_save_state's file: continues bit for bit."""
        check(self._lib.mhit_load_state(self._h, str(path).encode()), "load_state")

    def drop_count(self, max_drops=1 << 16):
        """post/DropCountSize on the current phi: number of drops, and their
        volumes (cells) and equivalent diameters in raster order of each
        drop's first cell."""
        cnt = ctypes.c_int(0)
        vol = np.zeros(max_drops)
        deq = np.zeros(max_drops)
        check(self._lib.mhit_drop_count(self._h, int(max_drops), ctypes.byref(cnt),
                                        vol.ctypes.data_as(ctypes.c_void_p), deq.ctypes.data_as(ctypes.c_void_p)),
              "drop_count")
        k = min(cnt.value, max_drops)
        return cnt.value, vol[:k].copy(), deq[:k].copy()

    # -- launch shapes -----------------------------------------------------------
    def autotune(self, force=False):
        """Time the candidate launch shapes of the tunable passes on this fresh
        box (before any field is set) and keep the fastest; cached per
        (nx, precision, phiflag) for the process (see mhit_autotune)."""
        t = Tuning()
        check(self._lib.mhit_autotune(self._h, MHIT_TUNE_FORCE if force else 0, ctypes.byref(t)), "autotune")
        return t.to_dict()

    def tuning(self):
        t = Tuning()
        check(self._lib.mhit_get_tuning(self._h, ctypes.byref(t)), "get_tuning")
        return t.to_dict()

    def set_tuning(self, **shapes):
        """Set launch shapes by knob name (0 = heuristic); see _lib.KNOBS."""
        t = Tuning()
        for k, v in shapes.items():
            t.value[_lib.KNOBS.index(k)] = int(v)
        check(self._lib.mhit_set_tuning(self._h, ctypes.byref(t)), "set_tuning")

    # -- measurement ------------------------------------------------------------
    def kernel_timing(self, enable=True):
        check(self._lib.mhit_kernel_timing(self._h, 1 if enable else 0), "kernel_timing")

    def kernel_stats(self):
        cnt = ctypes.c_int(0)
        check(self._lib.mhit_kernel_stats(self._h, None, 0, ctypes.byref(cnt)), "kernel_stats")
        arr = (KStat * max(cnt.value, 1))()
        check(self._lib.mhit_kernel_stats(self._h, arr, cnt.value, ctypes.byref(cnt)), "kernel_stats")
        out = []
        for s in arr[:cnt.value]:
            out.append({"name": s.name.decode(), "launches": s.launches, "total_ms": s.total_ms,
                        "avg_ms": s.total_ms / max(s.launches, 1), "bytes": s.bytes_per_launch})
        return out

    def debug_buffer(self, which):
        """Intermediate device buffer of the last step (see mhit_debug_buffer)."""
        n = self.n
        out = np.empty((n, n, n), dtype=np.float64, order="F")
        check(self._lib.mhit_debug_buffer(self._h, int(which), out.ctypes.data_as(ctypes.c_void_p)),
              f"debug_buffer({which})")
        return out

    def device_bytes(self):
        b = ctypes.c_size_t(0)
        check(self._lib.mhit_device_bytes(self._h, ctypes.byref(b)), "device_bytes")
        return b.value


def step_batch(solvers, nsteps=1, raise_on_stop=True):
    """Advance several boxes of one precision and phiflag, of one or several
    sizes, together: shared launches for all of them, the FFT passes one per
    size (mhit_step_batch).  Each box ends bit-identical to Solver.step on it
    alone.  Returns one Diag per box."""
    solvers = list(solvers)
    if not 1 <= len(solvers) <= MHIT_BATCH_MAX:
        raise ValueError(f"1..{MHIT_BATCH_MAX} boxes per batch, got {len(solvers)}")
    lib = _lib.load()
    hs = (ctypes.c_void_p * len(solvers))(*[s._h.value for s in solvers])
    ds = (Diag * len(solvers))()
    rc = lib.mhit_step_batch(hs, len(solvers), int(nsteps), ds)
    out = list(ds)
    for s, d in zip(solvers, out):
        s.last = d
    if rc != _lib.MHIT_OK and (raise_on_stop or rc not in (_lib.MHIT_ERR_CFL, _lib.MHIT_ERR_NAN)):
        raise MhitError(rc, "mhit_step_batch")
    return out
