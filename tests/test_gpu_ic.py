"""The device's random-phase turbulent start (mhit_init_random_velocity)
against its host restatement (oracle/ic_random.py: the same SplitMix64 draws,
numpy's inverse real DFT, the same staggered curl): equal within 1e-12 of the
field's maximum (the two differ in libm rounding and the FFT algorithm only),
on power-of-two (register FFT) and radix-3 (LDS FFT) sizes.  The device field
is discretely divergence-free (main.f90:723-725) to 1e-13 and has u_rms = 1;
in mixed precision it is the FP64 field rounded to FP32."""
import numpy as np
import pytest

from mhit36_amd import Solver
from oracle import ic_random

pytestmark = pytest.mark.gpu


def device_field(n, seed, k_peak=4.0, precision=0):
    with Solver(nx=n, dt=1e-3, mu=0.01, phiflag=0, inflow=1, precision=precision) as s:
        s.init_random_velocity(seed=seed, k_peak=k_peak, urms=1.0)
        return tuple(s.get_field(f) for f in (1, 2, 3))


@pytest.mark.parametrize("n,seed,k_peak", [(32, 1234, 4.0), (48, 7, 3.0), (64, 99, 4.0), (96, 5, 6.0)])
def test_device_generator_matches_host_restatement(n, seed, k_peak):
    got = device_field(n, seed, k_peak)
    ref = ic_random.random_velocity(n, seed=seed, k_peak=k_peak)
    for a, b in zip(got, ref):
        assert np.abs(a - b).max() <= 1e-12 * np.abs(b).max()
    u, v, w = got
    div = (np.roll(u, -1, 0) - u) + (np.roll(v, -1, 1) - v) + (np.roll(w, -1, 2) - w)
    assert np.abs(div).max() <= 1e-13 * max(np.abs(u).max(), np.abs(v).max(), np.abs(w).max())
    urms = np.sqrt((np.mean(u * u) + np.mean(v * v) + np.mean(w * w)) / 3.0)
    assert abs(urms - 1.0) < 1e-12


def test_mixed_precision_field_is_the_fp64_field_rounded():
    n = 64
    a = device_field(n, 3, precision=0)
    b = device_field(n, 3, precision=1)
    for x, y in zip(a, b):
        assert np.array_equal(y, x.astype(np.float32).astype(np.float64))
