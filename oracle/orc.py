"""ctypes binding of oracle/liboracle.so -- TEST INFRASTRUCTURE.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use this
module, always as the checker or the timed CPU baseline, never as the product.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liboracle.so")


class OrcParams(ctypes.Structure):
    """orc_params; same layout as mhit_params in include/mhit.h."""
    _fields_ = [
        ("nx", ctypes.c_int), ("restart", ctypes.c_int), ("tstart", ctypes.c_int),
        ("tfin", ctypes.c_int), ("dump", ctypes.c_int), ("inflow", ctypes.c_int),
        ("inphi", ctypes.c_int), ("phiflag", ctypes.c_int),
        ("dt", ctypes.c_double), ("mu", ctypes.c_double), ("rho", ctypes.c_double),
        ("f1", ctypes.c_double), ("f2", ctypes.c_double), ("f3", ctypes.c_double),
        ("k0", ctypes.c_double), ("radius", ctypes.c_double), ("sigma", ctypes.c_double),
        ("epsr", ctypes.c_double), ("gamma_factor", ctypes.c_double),
        ("wavenumbers", ctypes.c_int), ("poisson_rhs_fp32", ctypes.c_int),
        ("precision", ctypes.c_int), ("reserved", ctypes.c_int),
    ]


class OrcConsts(ctypes.Structure):
    _fields_ = [(name, ctypes.c_double) for name in
                ("twopi", "lx", "dx", "dxi", "ddxi", "rhoi", "eps", "epsi", "enum_")]


class OrcDiag(ctypes.Structure):
    _fields_ = [
        ("steps_done", ctypes.c_int), ("t", ctypes.c_int), ("status", ctypes.c_int),
        ("reserved", ctypes.c_int), ("gumax", ctypes.c_double), ("cou", ctypes.c_double),
        ("umean", ctypes.c_double), ("vmean", ctypes.c_double), ("wmean", ctypes.c_double),
    ]


DEFAULTS = dict(nx=16, restart=0, tstart=0, tfin=20000, dump=1000, inflow=0, inphi=0, phiflag=0,
                dt=0.0002, mu=0.006, rho=1.0, f1=0.0, f2=0.0, f3=0.0, k0=2.0, radius=0.5,
                sigma=1.0, epsr=1.0, gamma_factor=1.0, wavenumbers=0, poisson_rhs_fp32=1,
                precision=0, reserved=0)

_libs = {}


def load(path=None):
    """The oracle library at `path` (default: oracle/liboracle.so), loaded
    once per path.  Other paths are builds of the same source with other
    flags (the CPU baseline's -O3 -march=native build) or with a planted
    mistake (the mutation tests of tests/test_oracle_mutants.py)."""
    path = os.path.abspath(path or LIB_PATH)
    if path in _libs:
        return _libs[path]
    if not os.path.exists(path):
        raise ImportError(f"{path} missing: run make -C oracle")
    lib = ctypes.CDLL(path)
    P = ctypes.POINTER
    vp = ctypes.c_void_p
    dp = P(ctypes.c_double)
    sig = {
        "orc_derive": ([P(OrcParams), P(OrcConsts)], None),
        "orc_create": ([P(OrcParams)], vp),
        "orc_destroy": ([vp], None),
        "orc_nthreads": ([], ctypes.c_int),
        "orc_set_threads": ([ctypes.c_int], None),
        "orc_set_field": ([vp, ctypes.c_int, dp], None),
        "orc_get_field": ([vp, ctypes.c_int, dp], None),
        "orc_init_ic": ([vp], None),
        "orc_step": ([vp, ctypes.c_int, P(OrcDiag)], ctypes.c_int),
        "orc_get_consts": ([vp, P(OrcConsts)], None),
        "orc_set_sharpening": ([vp, ctypes.c_int], None),
        "orc_get_tables": ([vp, dp, dp], None),
        "orc_debug_buffer": ([vp, ctypes.c_int, dp], ctypes.c_int),
        "orc_fft_plan_radices": ([ctypes.c_int, P(ctypes.c_int)], ctypes.c_int),
        "orc_fft_1d": ([ctypes.c_int, ctypes.c_int, dp, ctypes.c_int], None),
        "orc_poisson": ([ctypes.c_int, P(OrcParams), dp, dp], None),
        "orc_sum_lane64": ([dp, ctypes.c_int, ctypes.c_ssize_t], ctypes.c_double),
        "orc_sum3": ([dp, ctypes.c_int], ctypes.c_double),
        "orc_dissipation": ([dp, dp, dp, ctypes.c_int, ctypes.c_double, ctypes.c_double], ctypes.c_double),
        "orc_kinetic_energy": ([dp, dp, dp, ctypes.c_int], ctypes.c_double),
        "orc_log": ([ctypes.c_double], ctypes.c_double),
        "orc_tanh": ([ctypes.c_double], ctypes.c_double),
        "orc_math_table": ([dp], None),
        "orc_drop_count": ([dp, ctypes.c_int, ctypes.c_int, dp, dp], ctypes.c_int),
    }
    for name, (args, res) in sig.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = res
    _libs[path] = lib
    return lib


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def params(**kw):
    p = OrcParams()
    for k, v in {**DEFAULTS, **kw}.items():
        setattr(p, k, v)
    return p


def consts(**kw):
    c = OrcConsts()
    load().orc_derive(ctypes.byref(params(**kw)), ctypes.byref(c))
    return {name: getattr(c, name) for name, _ in OrcConsts._fields_}


SHARPENING = {"literal": 0, "identity": 1}


class Oracle:
    """One box on the CPU, restating multi/main.f90 at a 1x1 decomposition.

    sharpening: how the ACDI sharpening factor 1/4 (1 - tanh^2(psi/(2 eps)))
    is evaluated -- "identity" (default: the form the HIP path evaluates,
    A B / (A + B)^2 with A = val + enum, B = 1 - val + enum) or "literal"
    (tanh of psi as main.f90:463-465 writes it).  They differ by rounding
    only."""

    def __init__(self, lib_path=None, sharpening="identity", **kw):
        self.lib = load(lib_path)
        self.p = params(**kw)
        self.n = self.p.nx
        self.h = self.lib.orc_create(ctypes.byref(self.p))
        if not self.h:
            raise ValueError(f"oracle rejects nx={self.n}")
        self.lib.orc_set_sharpening(self.h, SHARPENING[sharpening])

    def close(self):
        if self.h:
            self.lib.orc_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def init_ic(self):
        self.lib.orc_init_ic(self.h)

    def set_field(self, fieldn, a):
        a = np.asfortranarray(np.asarray(a, dtype=np.float64).reshape((self.n,) * 3, order="F"))
        self.lib.orc_set_field(self.h, int(fieldn), _dp(a))

    def get_field(self, fieldn):
        out = np.empty((self.n,) * 3, dtype=np.float64, order="F")
        self.lib.orc_get_field(self.h, int(fieldn), _dp(out))
        return out

    def step(self, nsteps=1):
        d = OrcDiag()
        self.lib.orc_step(self.h, int(nsteps), ctypes.byref(d))
        return {name: getattr(d, name) for name, _ in OrcDiag._fields_}

    def debug_buffer(self, which):
        out = np.empty((self.n,) * 3, dtype=np.float64, order="F")
        if self.lib.orc_debug_buffer(self.h, int(which), _dp(out)) != 0:
            raise ValueError(which)
        return out

    def tables(self):
        x = np.empty(self.n)
        k = np.empty(self.n)
        self.lib.orc_get_tables(self.h, _dp(x), _dp(k))
        return x, k

    def consts(self):
        c = OrcConsts()
        self.lib.orc_get_consts(self.h, ctypes.byref(c))
        return {name: getattr(c, name) for name, _ in OrcConsts._fields_}


def fft_1d(z, sign):
    """Unnormalised FFT along the last axis of a complex array (copy)."""
    z = np.ascontiguousarray(z, dtype=np.complex128)
    n = z.shape[-1]
    out = z.copy()
    buf = out.view(np.float64)
    load().orc_fft_1d(n, int(sign), _dp(buf), int(out.size // n))
    return out


def plan_radices(n):
    r = (ctypes.c_int * 16)()
    ns = load().orc_fft_plan_radices(int(n), r)
    return list(r[:ns])


def poisson(rhsp, **kw):
    n = rhsp.shape[0]
    p = params(nx=n, **kw)
    a = np.asfortranarray(rhsp, dtype=np.float64)
    out = np.empty((n, n, n), dtype=np.float64, order="F")
    load().orc_poisson(n, ctypes.byref(p), _dp(a), _dp(out))
    return out


def sum_lane64(v):
    v = np.ascontiguousarray(v, dtype=np.float64)
    return load().orc_sum_lane64(_dp(v), v.size, 1)


def sum3(f):
    f = np.asfortranarray(f, dtype=np.float64)
    return load().orc_sum3(_dp(f), f.shape[0])


def dissipation(u, v, w, dxi, nu):
    u, v, w = (np.asfortranarray(a, dtype=np.float64) for a in (u, v, w))
    return load().orc_dissipation(_dp(u), _dp(v), _dp(w), u.shape[0], float(dxi), float(nu))


def kinetic_energy(u, v, w):
    u, v, w = (np.asfortranarray(a, dtype=np.float64) for a in (u, v, w))
    return load().orc_kinetic_energy(_dp(u), _dp(v), _dp(w), u.shape[0])


def log(x):
    return load().orc_log(float(x))


def tanh(x):
    return load().orc_tanh(float(x))


def drop_count(phi, max_drops=1 << 16):
    """post/DropCountSize restated: (count, volumes in cells, equivalent
    diameters) of the phi >= 0.5 sets, 26-connected, periodic, in raster order."""
    phi = np.asfortranarray(phi, dtype=np.float64)
    vol = np.zeros(max_drops)
    deq = np.zeros(max_drops)
    cnt = load().orc_drop_count(_dp(phi), phi.shape[0], int(max_drops), _dp(vol), _dp(deq))
    if cnt < 0:
        raise MemoryError("orc_drop_count")
    k = min(cnt, max_drops)
    return cnt, vol[:k].copy(), deq[:k].copy()


def math_table():
    t = np.zeros(451)
    load().orc_math_table(_dp(t))
    return t


def set_threads(n):
    load().orc_set_threads(int(n))


def nthreads():
    return load().orc_nthreads()
