"""Gang stepping (mhit_step_batch): several boxes of one precision and
phiflag, of one or several sizes, advanced by shared launches (one per pass,
the FFT passes one per size).  Every box must end bit-identical to the same
box stepped alone with mhit_step: all fields, the step count, the stop status
and the scalars of the last step."""
import ctypes

import numpy as np
import pytest

from mhit36_amd import Solver, MhitError, ic, step_batch
from mhit36_amd import _lib

pytestmark = pytest.mark.gpu


def box_params(n, q, phiflag, precision, forced=False):
    """Box q of a gang: its own dt, viscosity, forcing and drops."""
    _, dx = ic.grid(n)
    p = dict(nx=n, dt=(0.04 + 0.01 * q) * dx, mu=0.004 + 0.003 * q, phiflag=phiflag, inflow=0,
             inphi=1 if phiflag else 0, precision=precision, sigma=0.05 + 0.01 * q, gamma_factor=1.0 + 0.1 * q)
    if forced:
        p.update(f1=1.0, f2=0.5 + 0.25 * q, f3=1.0)
    if phiflag:
        p["radius"] = 0.3 + 0.02 * q
    return p


def start(p, q, pre_steps=0):
    s = Solver(**p)
    s.init_fields()
    n = p["nx"]
    if p["phiflag"]:
        _, dx = ic.grid(n)
        R = p["radius"]
        s.set_field(5, ic.drops(n, ic.drop_centres(1 + q % 3, R, dx, seed=100 + q), R, dx))
    if pre_steps:
        s.step(pre_steps)
    return s


def fields(s):
    return [s.get_field(f) for f in ((1, 2, 3, 4, 5) if s.params.phiflag else (1, 2, 3, 4))]


def check_gang(n, phiflag, precision, nbox, nsteps, pre=(), forced=False, calls=1):
    """n: one size for all nbox boxes, or a list of the boxes' sizes."""
    ns = list(n) if isinstance(n, (list, tuple)) else [n] * nbox
    ps = [box_params(ns[q], q, phiflag, precision, forced) for q in range(nbox)]
    pre = list(pre) + [0] * (nbox - len(pre))
    alone, alone_d = [], []
    for q, p in enumerate(ps):
        with start(p, q, pre[q]) as s:
            for _ in range(calls):
                d = s.step(nsteps, raise_on_stop=False)
            alone.append(fields(s))
            alone_d.append(d)
    gang = [start(p, q, pre[q]) for q, p in enumerate(ps)]
    try:
        for _ in range(calls):
            ds = step_batch(gang, nsteps, raise_on_stop=False)
        for q, s in enumerate(gang):
            for f, (a, b) in enumerate(zip(alone[q], fields(s))):
                assert np.array_equal(a, b), (n, phiflag, precision, q, f)
            a, b = alone_d[q], ds[q]
            assert (a.steps_done, a.t, a.status) == (b.steps_done, b.t, b.status), q
            assert (a.gumax, a.cou, a.umean, a.vmean, a.wmean) == (b.gumax, b.cou, b.umean, b.vmean, b.wmean), q
    finally:
        for s in gang:
            s.close()


@pytest.mark.parametrize("n,phiflag,precision", [
    (32, 1, 0),      # register FFT, phase field, FP64
    (32, 1, 1),
    (48, 1, 1),      # radix-3 plan: LDS-staged FFT kernels
    (48, 0, 0),
    (64, 0, 0),
    (64, 0, 1),
])
def test_gang_equals_alone(n, phiflag, precision):
    check_gang(n, phiflag, precision, nbox=4, nsteps=5, forced=True)


def test_gang_mixes_euler_and_ab2_steps():
    """Boxes that already ran 0, 1, 2 and 3 steps alone: the first gang step
    is Euler for one box and AB2 for the others, and the phi buffers start in
    both orientations."""
    check_gang(32, 1, 0, nbox=4, nsteps=3, pre=(0, 1, 2, 3))


def test_gang_repeated_calls():
    check_gang(32, 1, 1, nbox=3, nsteps=2, calls=3)


def test_gang_of_sixteen():
    check_gang(32, 1, 1, nbox=16, nsteps=2)


@pytest.mark.parametrize("phiflag,precision", [(1, 0), (1, 1), (0, 0), (0, 1)])
def test_gang_of_uneven_sizes(phiflag, precision):
    """Boxes of 64, 32, 128, 48, 64 and 96 in one gang, in that order: the
    shared passes in one launch, the FFT passes per size (compiled plans at 64
    and 128, the runtime plan at 32, LDS-staged radix-3 kernels at 48 and 96)."""
    check_gang([64, 32, 128, 48, 64, 96], phiflag, precision, nbox=6, nsteps=4, forced=True)


@pytest.mark.parametrize("phiflag,precision", [(1, 0), (0, 1)])
def test_gang_with_large_boxes(phiflag, precision):
    """A 256^3 box (register FFT with the plan read at run time, 8-column
    tiles) and a 192^3 box (radix-3, LDS-staged) in one gang with a 64^3 box."""
    check_gang([256, 64, 192], phiflag, precision, nbox=3, nsteps=2, forced=True)


def test_gang_of_uneven_sizes_mixes_euler_and_ab2_steps():
    check_gang([48, 32, 64, 32], 1, 0, nbox=4, nsteps=3, pre=(2, 0, 1, 3))


def test_gang_box_that_stops():
    """A box whose CFL number passes 7 stops after its first step (halted,
    MHIT_ERR_CFL) exactly as alone; the other boxes go on."""
    n = 32
    _, dx = ic.grid(n)
    ps = [box_params(n, q, 1, 0) for q in range(3)]
    ps[1]["dt"] = 20.0 * dx                      # umax about 1: cou about 20 after step 1
    alone = []
    for q, p in enumerate(ps):
        with start(p, q) as s:
            d = s.step(4, raise_on_stop=False)
            alone.append((fields(s), d.steps_done, d.status))
    assert alone[1][1] == 1 and alone[1][2] == _lib.MHIT_ERR_CFL
    gang = [start(p, q) for q, p in enumerate(ps)]
    try:
        with pytest.raises(MhitError):
            step_batch(gang, 4)
        for q, s in enumerate(gang):
            d = s.last
            assert (d.steps_done, d.status) == alone[q][1:], q
            for a, b in zip(alone[q][0], fields(s)):
                assert np.array_equal(a, b), q
        ds = step_batch(gang, 2, raise_on_stop=False)   # the stopped box stays stopped
        assert ds[1].steps_done == 0 and ds[1].status == _lib.MHIT_ERR_CFL
        assert ds[0].steps_done == 2 and ds[2].steps_done == 2
    finally:
        for s in gang:
            s.close()


def test_gang_rejects_mixed_classes():
    """Precision and phiflag must agree (sizes may differ)."""
    with Solver(**box_params(32, 0, 1, 0)) as a, Solver(**box_params(32, 1, 1, 1)) as b, \
            Solver(**box_params(32, 1, 0, 0)) as d:
        for other in (b, d):
            with pytest.raises(MhitError):
                step_batch([a, other], 1)
        with pytest.raises(MhitError):
            step_batch([a, a], 1)                 # the same box twice
        lib = _lib.load()
        hs = (ctypes.c_void_p * 17)(*([a._h.value] * 17))
        assert lib.mhit_step_batch(hs, 17, 1, None) == _lib.MHIT_ERR_ARG
        assert lib.mhit_step_batch(hs, 0, 1, None) == _lib.MHIT_ERR_ARG
