"""The measured configurations at their full sizes.

- 256^3 forced random HIT (NS + ABC forcing): bit-identical to the oracle
  over the Euler step and two AB2 steps; then 1000 steps on the device with
  the mean flow removed to round-off and the CFL stop never tripped.  dt is
  fixed at CFL 0.3 of the initial umax: at CFL 0.5 the forcing raises umax
  and the explicit scheme (AB2, central differences) blows up near step 140,
  on the device and in the oracle alike (the first steps agree bit for bit).
- 256^3 HIT with 64 drops of radius 0.25 (NS + ACDI, CSF): bit-identical to
  the oracle over three steps; then FP64 against mixed precision over 1000
  steps: dissipation within 1e-4 and phase volume within 1e-5 of FP64 at every
  100th step, the same drop count in both.  CFL 0.05: the reference's explicit
  phase-field update goes unstable near CFL 0.14 once forcing raises umax.
- 512^3 NS+ACDI with 512 drops of radius 0.1 (the benchmark box): bit-identical
  to the oracle over two steps.
- 192^3 with 16 drops (a radix-3 size of the serving stream, LDS-staged FFT):
  bit-identical over three steps; mixed within 1e-5 of FP64 over 50 steps.
- 1024^3 NS-only forced HIT (the random-phase spectrum made on the device,
  ABC forcing, nu = 0.001) resident on one GPU: FP64 against mixed precision
  over 100 steps, kinetic energy within 1e-5 and dissipation within 1e-4.
- 512^3 NS+ACDI with 512 drops (the benchmark box): mixed against FP64 over
  100 steps, dissipation within 1e-4, phase volume within 1e-5, same drops.
"""
import numpy as np
import pytest

from mhit36_amd import Solver, ic
from oracle import orc

pytestmark = pytest.mark.gpu


def hit(n, seed=1234):
    u, v, w = ic.random_velocity(n, seed=seed, k_peak=4.0, urms=1.0)
    _, dx = ic.grid(n)
    return u, v, w, dx, max(u.max(), v.max(), w.max())


def load(box, fields):
    for f, a in fields.items():
        box.set_field(f, a)


def assert_same(s, o, fields):
    for f in fields:
        assert np.array_equal(s.get_field(f), o.get_field(f)), f


def test_256_forced_hit():
    n = 256
    u, v, w, dx, umax = hit(n)
    kw = dict(nx=n, dt=0.3 * dx / umax, mu=0.005, f1=1.0, f2=1.0, f3=1.0, phiflag=0, inflow=1)
    fields = {1: u, 2: v, 3: w}
    with Solver(**kw) as s:
        load(s, fields)
        o = orc.Oracle(**kw)
        load(o, fields)
        for _ in range(3):
            s.step(1)
            o.step(1)
            assert_same(s, o, (1, 2, 3, 4))
        o.close()
        for _ in range(10):
            d = s.step(100)
            assert np.isfinite(d.gumax) and d.cou <= 7.0
        for f in (1, 2, 3):
            a = s.get_field(f)
            assert abs(a.mean()) < 1e-13 * np.abs(a).max()      # mean flow removed (main.f90:889-920)


def test_256_acdi_fp64_vs_mixed():
    n = 256
    u, v, w, dx, umax = hit(n)
    centres = ic.drop_centres(64, 0.25, dx, seed=4321)
    phi = ic.drops(n, centres, 0.25, dx)
    kw = dict(nx=n, dt=0.05 * dx / umax, mu=0.005, f1=1.0, f2=1.0, f3=1.0, phiflag=1, inflow=1, inphi=1,
              radius=0.25, sigma=0.1, gamma_factor=1.2)
    fields = {1: u, 2: v, 3: w, 5: phi}
    with Solver(**kw) as s:
        load(s, fields)
        o = orc.Oracle(**kw)
        load(o, fields)
        for _ in range(3):
            s.step(1)
            o.step(1)
            assert_same(s, o, (1, 2, 3, 4, 5))
        o.close()
    series, drops = {}, {}
    for prec in (0, 1):
        with Solver(**kw, precision=prec) as s:
            load(s, fields)
            assert s.drop_count()[0] == 64
            series[prec] = []
            for _ in range(10):
                s.step(100)
                series[prec].append(s.diagnostics())
            drops[prec] = s.drop_count()[0]
    assert drops[1] == drops[0] > 0, drops
    for a, b in zip(series[1], series[0]):
        assert abs(a["diss"] - b["diss"]) <= 1e-4 * b["diss"]
        assert abs(a["phi_sum"] - b["phi_sum"]) <= 1e-5 * b["phi_sum"]


def test_512_acdi_benchmark_box():
    import bench
    n = 512
    u, v, w, phi, dt = bench.make_inputs(n, 1)
    kw = bench.params_for(n, 1, dt)
    fields = {1: u, 2: v, 3: w, 5: phi}
    with Solver(**kw) as s:
        load(s, fields)
        s.step(2)
        o = orc.Oracle(**kw)
        load(o, fields)
        o.step(2)
        assert_same(s, o, (1, 2, 3, 4, 5))
        o.close()


def test_1024_ns_resident():
    """1024^3 forced HIT: the turbulent start generated on the device (u_rms 1,
    E(k) ~ k^4 exp(-2 (k/4)^2)), dt = 0.06 dx (CFL about 0.3 at the field's
    maximum), 100 steps in FP64 and in mixed precision, each context inside
    the GPU's 288 GB."""
    n = 1024
    dx = 2 * np.pi / (n - 1)
    kw = dict(nx=n, dt=0.06 * dx, mu=0.001, f1=1.0, f2=1.0, f3=1.0, k0=2.0, phiflag=0, inflow=1)
    res = {}
    for prec in (0, 1):
        print(f"1024^3 NS-only precision={prec}", flush=True)
        with Solver(**kw, precision=prec) as s:
            assert s.device_bytes() < 288 * 10 ** 9
            s.init_random_velocity(seed=1234, k_peak=4.0, urms=1.0)
            d0 = s.diagnostics()
            assert abs(np.sqrt(2.0 * d0["ke"] / 3.0) - 1.0) < 1e-6
            d = s.step(100)
            assert d.steps_done == 100 and np.isfinite(d.gumax) and d.cou < 0.6
            res[prec] = (d.gumax, s.diagnostics())
    assert abs(res[1][0] - res[0][0]) <= 1e-4 * res[0][0]
    assert abs(res[1][1]["diss"] - res[0][1]["diss"]) <= 1e-4 * res[0][1]["diss"]
    assert abs(res[1][1]["ke"] - res[0][1]["ke"]) <= 1e-5 * res[0][1]["ke"]


def test_512_mixed_against_fp64():
    """The benchmark box in mixed precision against FP64 over 100 steps."""
    import bench
    n = 512
    u, v, w, phi, dt = bench.make_inputs(n, 1)
    kw = bench.params_for(n, 1, dt)
    res = {}
    for prec in (0, 1):
        print(f"512^3 100 steps precision={prec}", flush=True)
        with Solver(**kw, precision=prec) as s:
            load(s, {1: u, 2: v, 3: w, 5: phi})
            d = s.step(100)
            assert d.steps_done == 100
            res[prec] = (s.diagnostics(), s.drop_count()[0])
    (a, na), (b, nb) = res[0], res[1]
    assert abs(b["diss"] - a["diss"]) <= 1e-4 * a["diss"]
    assert abs(b["phi_sum"] - a["phi_sum"]) <= 1e-5 * a["phi_sum"]
    assert na == nb


@pytest.mark.parametrize("precision", [0, 1])
def test_192_radix3_acdi(precision):
    """A radix-3 size of the serving stream (192 = 2^6 3: the LDS-staged FFT
    kernels): FP64 bit-identical to the oracle over three steps; mixed within
    1e-5 of FP64 over 50 steps."""
    n = 192
    u, v, w, dx, umax = hit(n, seed=77)
    phi = ic.drops(n, ic.drop_centres(16, 0.3, dx, seed=5), 0.3, dx)
    kw = dict(nx=n, dt=0.05 * dx / umax, mu=0.004, f1=1.0, f2=1.0, f3=1.0, phiflag=1, inflow=1, inphi=1,
              radius=0.3, sigma=0.1, gamma_factor=1.2)
    fields = {1: u, 2: v, 3: w, 5: phi}
    if precision == 0:
        with Solver(**kw) as s:
            load(s, fields)
            o = orc.Oracle(**kw)
            load(o, fields)
            for _ in range(3):
                s.step(1)
                o.step(1)
                assert_same(s, o, (1, 2, 3, 4, 5))
            o.close()
    else:
        res = {}
        for prec in (0, 1):
            with Solver(**kw, precision=prec) as s:
                load(s, fields)
                s.step(50)
                res[prec] = s.diagnostics()
        assert abs(res[1]["diss"] - res[0]["diss"]) <= 1e-5 * res[0]["diss"]
        assert abs(res[1]["phi_sum"] - res[0]["phi_sum"]) <= 1e-6 * res[0]["phi_sum"]
