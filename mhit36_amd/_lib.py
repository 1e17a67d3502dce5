"""ctypes binding of libmhit.so, the C ABI declared in include/mhit.h.

The library is the product path: it is loaded from the package directory and
nothing here falls back to another implementation when it is missing.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libmhit.so")

MHIT_OK = 0
MHIT_ERR_HIP = 1
MHIT_ERR_OOM = 2
MHIT_ERR_BAD_N = 3
MHIT_ERR_ARG = 4
MHIT_ERR_CFL = 5
MHIT_ERR_NAN = 6
MHIT_ERR_IO = 7
MHIT_ERR_NODEV = 8
MHIT_ERR_UNSUPPORTED = 9

MHIT_FP64 = 0
MHIT_MIXED = 1
MHIT_K_REFERENCE = 0
MHIT_K_MODIFIED = 1
MHIT_BATCH_MAX = 16
MHIT_NOALLOC_MAX_N = 128

# Names of the C-ABI entry points declared in include/mhit.h.
EXPORTS = (
    "mhit_default_params", "mhit_read_input", "mhit_derive", "mhit_create",
    "mhit_destroy", "mhit_init_fields", "mhit_set_field", "mhit_get_field",
    "mhit_step", "mhit_step_batch", "mhit_reset_time", "mhit_init_random_velocity", "mhit_write_field", "mhit_read_field",
    "mhit_diagnostics", "mhit_kernel_timing", "mhit_kernel_stats", "mhit_device_bytes",
    "mhit_debug_math", "mhit_debug_buffer", "mhit_autotune", "mhit_get_tuning", "mhit_set_tuning", "mhit_drop_count", "mhit_drop_count_noalloc", "mhit_save_state", "mhit_load_state",
    "mhit_ens_create", "mhit_ens_destroy", "mhit_ens_submit", "mhit_ens_progress", "mhit_ens_wait_idle",
    "mhit_ens_results", "mhit_ens_get_field", "mhit_ens_release", "mhit_ens_latency",
    "mhit_strerror", "mhit_abi_version",
)


class Params(ctypes.Structure):
    """mhit_params: input.inp values (readinput.f90:13-31) plus switches."""
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

    def to_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_}


class Consts(ctypes.Structure):
    """mhit_consts: readinput.f90:35-43."""
    _fields_ = [(name, ctypes.c_double) for name in
                ("twopi", "lx", "dx", "dxi", "ddxi", "rhoi", "eps", "epsi", "enum_")]


class Diag(ctypes.Structure):
    _fields_ = [
        ("steps_done", ctypes.c_int), ("t", ctypes.c_int), ("status", ctypes.c_int),
        ("reserved", ctypes.c_int), ("gumax", ctypes.c_double), ("cou", ctypes.c_double),
        ("umean", ctypes.c_double), ("vmean", ctypes.c_double), ("wmean", ctypes.c_double),
        ("ms", ctypes.c_double),
    ]


class Stats(ctypes.Structure):
    _fields_ = [(name, ctypes.c_double) for name in ("ke", "diss", "phi_sum", "phi_volume")]


class KStat(ctypes.Structure):
    _fields_ = [
        ("name", ctypes.c_char * 32), ("launches", ctypes.c_int), ("reserved", ctypes.c_int),
        ("total_ms", ctypes.c_double), ("bytes_per_launch", ctypes.c_double),
    ]


KNOBS = ("kz_phase", "kz_momentum", "x_pairs", "xinv_pairs", "y_tile", "z_tile")
MHIT_TUNE_FORCE = 1


class Tuning(ctypes.Structure):
    """mhit_tuning: launch shape of each tunable pass."""
    _fields_ = [("value", ctypes.c_int * len(KNOBS)), ("heuristic", ctypes.c_int * len(KNOBS)),
                ("ms_heuristic", ctypes.c_double), ("ms_tuned", ctypes.c_double)]

    def to_dict(self):
        return {"value": dict(zip(KNOBS, self.value)), "heuristic": dict(zip(KNOBS, self.heuristic)),
                "ms_heuristic": self.ms_heuristic, "ms_tuned": self.ms_tuned}


class MhitError(RuntimeError):
    def __init__(self, code, what=""):
        self.code = code
        msg = strerror(code) if _lib is not None else str(code)
        super().__init__(f"{what}: {msg} (code {code})" if what else f"{msg} (code {code})")


_lib = None


def load():
    """Load libmhit.so from the package directory; raise if it is not built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"{LIB_PATH} is missing: build it first (python -c 'import __graft_entry__ as g; g.build()')")
    lib = ctypes.CDLL(LIB_PATH)
    P = ctypes.POINTER
    vp = ctypes.c_void_p
    sig = {
        "mhit_default_params": ([P(Params)], ctypes.c_int),
        "mhit_read_input": ([ctypes.c_char_p, ctypes.c_int, P(Params)], ctypes.c_int),
        "mhit_derive": ([P(Params), P(Consts)], ctypes.c_int),
        "mhit_create": ([P(Params), P(vp)], ctypes.c_int),
        "mhit_destroy": ([vp], ctypes.c_int),
        "mhit_init_fields": ([vp], ctypes.c_int),
        "mhit_set_field": ([vp, ctypes.c_int, vp], ctypes.c_int),
        "mhit_get_field": ([vp, ctypes.c_int, vp], ctypes.c_int),
        "mhit_step": ([vp, ctypes.c_int, P(Diag)], ctypes.c_int),
        "mhit_step_batch": ([P(vp), ctypes.c_int, ctypes.c_int, P(Diag)], ctypes.c_int),
        "mhit_reset_time": ([vp, ctypes.c_int], ctypes.c_int),
        "mhit_init_random_velocity": ([vp, ctypes.c_ulonglong, ctypes.c_double, ctypes.c_double], ctypes.c_int),
        "mhit_write_field": ([vp, ctypes.c_char_p, ctypes.c_int, ctypes.c_int], ctypes.c_int),
        "mhit_read_field": ([vp, ctypes.c_char_p, ctypes.c_int], ctypes.c_int),
        "mhit_diagnostics": ([vp, ctypes.c_double, ctypes.c_double, P(Stats)], ctypes.c_int),
        "mhit_kernel_timing": ([vp, ctypes.c_int], ctypes.c_int),
        "mhit_kernel_stats": ([vp, P(KStat), ctypes.c_int, P(ctypes.c_int)], ctypes.c_int),
        "mhit_device_bytes": ([vp, P(ctypes.c_size_t)], ctypes.c_int),
        "mhit_debug_math": ([ctypes.c_int, ctypes.c_size_t, vp, vp], ctypes.c_int),
        "mhit_debug_buffer": ([vp, ctypes.c_int, vp], ctypes.c_int),
        "mhit_autotune": ([vp, ctypes.c_int, P(Tuning)], ctypes.c_int),
        "mhit_get_tuning": ([vp, P(Tuning)], ctypes.c_int),
        "mhit_drop_count": ([vp, ctypes.c_int, P(ctypes.c_int), vp, vp], ctypes.c_int),
        "mhit_drop_count_noalloc": ([vp, ctypes.c_int, P(ctypes.c_int), vp, vp], ctypes.c_int),
        "mhit_save_state": ([vp, ctypes.c_char_p], ctypes.c_int),
        "mhit_load_state": ([vp, ctypes.c_char_p], ctypes.c_int),
        "mhit_set_tuning": ([vp, P(Tuning)], ctypes.c_int),
        "mhit_strerror": ([ctypes.c_int], ctypes.c_char_p),
        "mhit_abi_version": ([], ctypes.c_int),
    }
    for name, (args, res) in sig.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = res
    _lib = lib
    return lib


def strerror(code):
    return load().mhit_strerror(int(code)).decode()


def check(code, what=""):
    if code != MHIT_OK:
        raise MhitError(code, what)
    return code
