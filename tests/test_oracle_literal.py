"""The C oracle (periodic index wrap, own FFT, R2C) against the literal numpy
restatement (padded X-pencils with cudecompUpdateHalosX plane copies, full
complex DFT via pocketfft), oracle/literal.py."""
import numpy as np
import pytest

from oracle import orc
from oracle.literal import LiteralBox

from gpu_util import random_velocity


def compare(kw, fields, steps):
    o = orc.Oracle(**kw)
    lit = LiteralBox(**{k: v for k, v in kw.items() if k not in ("inflow", "inphi", "wavenumbers")})
    if fields is None:
        o.init_ic()
        fields = {f: o.get_field(f) for f in ((1, 2, 3, 5) if kw.get("phiflag") else (1, 2, 3))}
    for f, a in fields.items():
        o.set_field(f, a)
        lit.set_interior({1: lit.u, 2: lit.v, 3: lit.w, 5: lit.phi}[f], a)
    for _ in range(steps):
        d = o.step(1)
        dl = lit.step()
        assert d["gumax"] == pytest.approx(dl["gumax"], rel=1e-13, abs=1e-15)
    for f, a in ((1, lit.u), (2, lit.v), (3, lit.w), (4, lit.p)) + (((5, lit.phi),) if kw.get("phiflag") else ()):
        ref = lit.interior(a)
        got = o.get_field(f)
        scale = max(np.abs(ref).max(), 1e-12)
        assert np.abs(got - ref).max() <= 1e-12 * scale, (f, np.abs(got - ref).max() / scale)


def test_literal_ns_tgv():
    compare(dict(nx=16, dt=1e-3, mu=1 / 1600), None, 4)


def test_literal_ns_forced_random():
    n = 12
    u, v, w = random_velocity(n, seed=11)
    compare(dict(nx=n, dt=2e-3, mu=0.01, f1=1.0, f2=0.7, f3=0.4, k0=2.0, inflow=1), {1: u, 2: v, 3: w}, 5)


def test_literal_acdi_drop():
    compare(dict(nx=16, dt=1e-3, mu=0.01, f1=0.5, f2=0.3, f3=0.2, radius=1.0, sigma=0.1, phiflag=1), None, 5)


def test_literal_acdi_gamma_factor_no_fp32():
    compare(dict(nx=16, dt=1e-3, mu=0.01, radius=1.2, sigma=0.2, phiflag=1, gamma_factor=1.2,
                 poisson_rhs_fp32=0), None, 3)
