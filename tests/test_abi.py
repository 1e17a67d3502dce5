"""The C ABI of libmhit.so without a GPU: it loads, exports every function
include/mhit.h declares, keeps the struct layouts the header fixes, parses
input.inp like readinput.f90 and derives the same constants as the oracle."""
import ctypes
import os
import re

import pytest

from mhit36_amd import _lib
from mhit36_amd.solver import default_params, read_input, derive
from oracle import orc

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "mhit.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(mhit_[a-z_]+)\s*\(", text)))


def test_library_exports_every_declared_function():
    lib = _lib.load()
    names = declared_functions()
    assert len(names) >= 19
    for name in names:
        assert hasattr(lib, name), name
    assert set(names) == set(_lib.EXPORTS)


def test_abi_version_and_strerror():
    lib = _lib.load()
    assert lib.mhit_abi_version() == 3
    for code in range(10):
        assert _lib.strerror(code)
    assert _lib.strerror(99) == "unknown error"


def test_struct_layouts():
    # 8 ints, 11 doubles, 4 ints
    assert ctypes.sizeof(_lib.Params) == 8 * 4 + 11 * 8 + 4 * 4
    assert ctypes.sizeof(_lib.Params) == ctypes.sizeof(orc.OrcParams)
    assert [f[0] for f in _lib.Params._fields_] == [f[0] for f in orc.OrcParams._fields_]
    assert ctypes.sizeof(_lib.Diag) == 4 * 4 + 6 * 8
    assert ctypes.sizeof(_lib.KStat) == 32 + 8 + 16
    from mhit36_amd.ensemble import EnsConfig
    # size_t budget, then max_active, inflight, use_graphs, autotune, gang_size, gang_max_nx
    assert ctypes.sizeof(EnsConfig) == 8 + 6 * 4
    assert [f[0] for f in EnsConfig._fields_][-2:] == ["gang_size", "gang_max_nx"]


def test_step_batch_rejects_bad_arguments_before_any_device_call():
    lib = _lib.load()
    hs = (ctypes.c_void_p * 2)(None, None)
    assert lib.mhit_step_batch(None, 1, 1, None) == _lib.MHIT_ERR_ARG
    assert lib.mhit_step_batch(hs, 0, 1, None) == _lib.MHIT_ERR_ARG
    assert lib.mhit_step_batch(hs, _lib.MHIT_BATCH_MAX + 1, 1, None) == _lib.MHIT_ERR_ARG
    assert lib.mhit_step_batch(hs, 2, 1, None) == _lib.MHIT_ERR_ARG      # null contexts
    assert lib.mhit_step_batch(hs, 1, -1, None) == _lib.MHIT_ERR_ARG


def test_defaults_follow_input_inp():
    p = default_params()
    assert (p.restart, p.tstart, p.tfin, p.dump, p.inflow, p.inphi) == (0, 0, 20000, 1000, 0, 0)
    assert (p.dt, p.mu, p.rho) == (0.0002, 0.006, 1.0)
    assert (p.f1, p.f2, p.f3, p.k0) == (0.0, 0.0, 0.0, 2.0)
    assert (p.radius, p.sigma, p.epsr) == (0.5, 1.0, 1.0)
    assert p.phiflag == 1 and p.gamma_factor == 1.0 and p.poisson_rhs_fp32 == 1


def test_read_input_reference_file():
    """tests/golden/input.inp is the reference's multi/input.inp."""
    p = read_input(os.path.join(ROOT, "tests", "golden", "input.inp"), 64)
    assert p.nx == 64
    assert (p.restart, p.tstart, p.tfin, p.dump, p.inflow, p.inphi) == (0, 0, 20000, 1000, 0, 0)
    assert (p.dt, p.mu, p.rho, p.f1, p.f2, p.f3, p.k0) == (0.0002, 0.006, 1.0, 0.0, 0.0, 0.0, 2.0)
    assert (p.radius, p.sigma, p.epsr) == (0.5, 1.0, 1.0)     # "1.0d0" is a Fortran double literal


def test_read_input_errors(tmp_path):
    with pytest.raises(_lib.MhitError) as e:
        read_input(str(tmp_path / "missing.inp"), 32)
    assert e.value.code == _lib.MHIT_ERR_IO
    short = tmp_path / "short.inp"
    short.write_text("0\n0\n10\n")
    with pytest.raises(_lib.MhitError) as e:
        read_input(str(short), 32)
    assert e.value.code == _lib.MHIT_ERR_IO


@pytest.mark.parametrize("n", [16, 64, 128, 512])
def test_derive_matches_oracle_bitwise(n):
    p = default_params(nx=n, rho=1.3, epsr=0.75)
    got = derive(p)
    ref = orc.consts(nx=n, rho=1.3, epsr=0.75)
    assert got == ref
    assert got["dx"] == got["lx"] / (n - 1)                 # dx = lx/(nx-1)
    assert got["enum_"] == 1.0000000168623835e-16            # REAL(4) 1.e-16


def test_create_rejects_bad_grid_before_touching_the_device():
    lib = _lib.load()
    h = ctypes.c_void_p()
    for n in (0, 2, 7, 10, 50, 4096):
        p = default_params(nx=n)
        assert lib.mhit_create(ctypes.byref(p), ctypes.byref(h)) == _lib.MHIT_ERR_BAD_N
    p = default_params(nx=32, precision=2)
    assert lib.mhit_create(ctypes.byref(p), ctypes.byref(h)) == _lib.MHIT_ERR_UNSUPPORTED
    p = default_params(nx=32, dt=0.0)
    assert lib.mhit_create(ctypes.byref(p), ctypes.byref(h)) == _lib.MHIT_ERR_ARG
    assert lib.mhit_create(None, ctypes.byref(h)) == _lib.MHIT_ERR_ARG


def test_null_context_arguments():
    lib = _lib.load()
    assert lib.mhit_step(None, 1, None) == _lib.MHIT_ERR_ARG
    assert lib.mhit_get_field(None, 1, None) == _lib.MHIT_ERR_ARG
    assert lib.mhit_destroy(None) == _lib.MHIT_ERR_ARG
