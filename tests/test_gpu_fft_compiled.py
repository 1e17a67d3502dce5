"""The FFT passes compiled for n = 2^L (64 <= n <= 1024) against the kernels
compiled for any n.

The compiled kernels take the stage plan (radices 8,8 at 64; 8,4,4 at 128;
8,8,4 at 256; 8,8,8 at 512; 8,8,4,4 at 1024) as constants, the y/z passes
also their default column tile.  A 2-column y/z tile runs the kernels that
read the plan at run time; the x passes stay compiled for any number of line
pairs per block, and run here with one pair instead of the default.  Same
inputs, same arithmetic: the fields must agree bit for bit.  128^3 (plan
8,4,4) is also checked against the oracle; 64^3, 256^3 and 512^3 are in the
parity and configuration tests.
"""
import numpy as np
import pytest

from mhit36_amd import Solver, ic
from oracle import orc

pytestmark = pytest.mark.gpu

OTHER_SHAPES = dict(x_pairs=1, xinv_pairs=1, y_tile=2, z_tile=2)


def _drops_case(n, seed):
    u, v, w = ic.random_velocity(n, seed=seed, k_peak=4.0, urms=1.0)
    _, dx = ic.grid(n)
    umax = max(u.max(), v.max(), w.max())
    phi = ic.drops(n, ic.drop_centres(8, 0.4, dx, seed=seed + 1), 0.4, dx)
    kw = dict(nx=n, dt=0.05 * dx / umax, mu=0.004, f1=1.0, f2=1.0, f3=1.0, phiflag=1, inflow=1, inphi=1,
              radius=0.4, sigma=0.1, gamma_factor=1.2)
    return kw, {1: u, 2: v, 3: w, 5: phi}


def test_128_compiled_plan_bit_identical_to_oracle():
    kw, fields = _drops_case(128, 31)
    with Solver(**kw) as s:
        for f, a in fields.items():
            s.set_field(f, a)
        o = orc.Oracle(**kw)
        for f, a in fields.items():
            o.set_field(f, a)
        for _ in range(2):
            s.step(1)
            o.step(1)
            for f in (1, 2, 3, 4, 5):
                assert np.array_equal(s.get_field(f), o.get_field(f)), f
        o.close()


@pytest.mark.parametrize("precision", [0, 1])
def test_128_compiled_equals_generic(precision):
    kw, fields = _drops_case(128, 41)
    out = []
    for shapes in ({}, OTHER_SHAPES):
        with Solver(**kw, precision=precision) as s:
            if shapes:
                s.set_tuning(**shapes)
            for f, a in fields.items():
                s.set_field(f, a)
            s.step(3)
            out.append([s.get_field(f) for f in (1, 2, 3, 4)])
    for a, b in zip(*out):
        assert np.array_equal(a, b)


def test_1024_compiled_equals_generic():
    """1024^3 (plan 8,8,4,4, one x line spanning two wavefronts, 4-column
    tiles): two steps of forced HIT from the device-made turbulent start."""
    n = 1024
    dx = 2 * np.pi / (n - 1)
    kw = dict(nx=n, dt=0.06 * dx, mu=0.001, f1=1.0, f2=1.0, f3=1.0, k0=2.0, phiflag=0, inflow=1)
    res = []
    for shapes in ({}, OTHER_SHAPES):
        print(f"1024^3 shapes={shapes or 'default'}", flush=True)
        with Solver(**kw) as s:
            s.init_random_velocity(seed=99, k_peak=4.0, urms=1.0)
            if shapes:
                s.set_tuning(**shapes)
            d = s.step(2)
            assert d.steps_done == 2
            res.append((d.gumax, s.diagnostics(), s.get_field(4)))
    (g0, d0, p0), (g1, d1, p1) = res
    assert g0 == g1
    assert d0 == d1
    assert np.array_equal(p0, p1)
