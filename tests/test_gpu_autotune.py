"""Launch shapes split the work over blocks differently and never change the
arithmetic: every valid shape of every tunable pass gives bit-identical
fields, the autotuner picks one of them, and its choice is cached per
(nx, precision, phiflag)."""
import numpy as np
import pytest

from mhit36_amd import Solver, MhitError
from mhit36_amd._lib import KNOBS

pytestmark = pytest.mark.gpu

CASES = [
    dict(nx=32, dt=5e-4, mu=0.01, phiflag=1, radius=0.8, sigma=0.1, gamma_factor=1.2, precision=0),
    dict(nx=48, dt=5e-4, mu=0.01, phiflag=1, radius=0.9, sigma=0.1, precision=1),   # radix-3 plan (LDS FFT)
    dict(nx=64, dt=1e-3, mu=0.005, phiflag=0, f1=1.0, f2=1.0, f3=1.0, precision=0),
]
STEPS = 4


def fields(p, steps=STEPS, **shapes):
    with Solver(**p) as s:
        if shapes:
            s.set_tuning(**shapes)
        s.init_fields()
        s.step(steps)
        return [s.get_field(f) for f in ((1, 2, 3, 4, 5) if p["phiflag"] else (1, 2, 3, 4))]


def valid_values(p, knob):
    out = []
    with Solver(**p) as s:
        for v in (1, 2, 4, 8, 16, 32, 64):
            try:
                s.set_tuning(**{knob: v})
                out.append(v)
            except MhitError:
                pass
    return out


@pytest.mark.parametrize("q", range(len(CASES)))
def test_every_shape_is_bit_identical(q):
    p = CASES[q]
    ref = fields(p, **{k: 0 for k in KNOBS})
    with Solver(**p) as s:
        heur = s.tuning()["heuristic"]
    for knob in KNOBS:
        if knob == "kz_phase" and not p["phiflag"]:
            continue
        vals = valid_values(p, knob)
        assert heur[knob] in vals, (knob, heur[knob], vals)
        for v in vals:
            if v == heur[knob]:
                continue
            got = fields(p, **{knob: v})
            for a, b in zip(ref, got):
                assert np.array_equal(a, b), (q, knob, v)


def test_invalid_shape_is_rejected():
    with Solver(**CASES[0]) as s:
        with pytest.raises(MhitError):
            s.set_tuning(kz_momentum=3)          # does not divide nx
        with pytest.raises(MhitError):
            s.set_tuning(y_tile=1024)


@pytest.mark.parametrize("q", range(len(CASES)))
def test_autotune_choice_is_cached_and_exact(q):
    p = CASES[q]
    ref = fields(p, **{k: 0 for k in KNOBS})
    with Solver(**p) as s:
        t = s.autotune(force=True)
        assert t["ms_tuned"] <= t["ms_heuristic"] + 1e-12
        for knob in KNOBS:
            if knob == "kz_phase" and not p["phiflag"]:
                continue
            assert t["value"][knob] in valid_values(p, knob)
        s.init_fields()
        s.step(STEPS)
        got = [s.get_field(f) for f in ((1, 2, 3, 4, 5) if p["phiflag"] else (1, 2, 3, 4))]
    for a, b in zip(ref, got):
        assert np.array_equal(a, b)
    with Solver(**p) as s2:                      # a new box of the same kind starts tuned
        assert s2.tuning()["value"] == t["value"]


def test_autotune_needs_a_fresh_box():
    p = CASES[0]
    with Solver(**p) as s:
        s.init_fields()
        with pytest.raises(MhitError):
            s.autotune(force=True)
