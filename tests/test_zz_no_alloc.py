"""The allocation-free drop count (mhit_drop_count_noalloc) against the
allocating one (mhit_drop_count), both called through the C ABI: the same
count, volumes and diameters on two calls in a row, the second on a smaller
box with fewer drops (so stale scratch from the first call would show), and
the limits of the scratch enforced."""
import ctypes

import numpy as np
import pytest

from mhit36_amd import Solver, _lib, ic

pytestmark = pytest.mark.gpu


def _count(s, entry, max_drops=1024):
    lib = _lib.load()
    cnt = ctypes.c_int(-1)
    vol = np.full(max_drops, -1.0)
    deq = np.full(max_drops, -1.0)
    rc = getattr(lib, entry)(s._h, max_drops, ctypes.byref(cnt), vol.ctypes.data_as(ctypes.c_void_p),
                             deq.ctypes.data_as(ctypes.c_void_p))
    k = max(0, min(cnt.value, max_drops))
    return rc, cnt.value, vol[:k].copy(), deq[:k].copy()


def _box(n, ndrops, radius, seed, precision=0):
    _, dx = ic.grid(n)
    phi = ic.drops(n, ic.drop_centres(ndrops, radius, dx, seed=seed), radius, dx)
    s = Solver(nx=n, dt=1e-3, mu=0.01, phiflag=1, inphi=1, radius=radius, precision=precision)
    s.init_fields()
    s.set_field(5, phi)
    return s


def test_noalloc_equals_allocating_count_on_two_calls():
    cases = ((128, 40, 0.3, 11, 0), (64, 6, 0.4, 12, 1))
    boxes = [_box(*c) for c in cases]
    try:
        got = [_count(s, "mhit_drop_count_noalloc") for s in boxes]     # two calls in a row, big then small
        ref = [_count(s, "mhit_drop_count") for s in boxes]
        for (n, nd, _, _, _), g, r in zip(cases, got, ref):
            print(f"noalloc n={n}: {g[1]} drops (allocating entry: {r[1]})", flush=True)
            assert g[0] == _lib.MHIT_OK and r[0] == _lib.MHIT_OK
            assert g[1] == r[1] == nd
            assert np.array_equal(g[2], r[2]) and np.array_equal(g[3], r[3])
            assert np.all(g[2] > 0)
        again = _count(boxes[0], "mhit_drop_count_noalloc")            # the big box once more
        assert again[1] == ref[0][1] and np.array_equal(again[2], ref[0][2])
    finally:
        for s in boxes:
            s.close()


def test_noalloc_rejects_what_the_scratch_cannot_hold():
    lib = _lib.load()
    cnt = ctypes.c_int(0)
    with Solver(nx=_lib.MHIT_NOALLOC_MAX_N + 64, dt=1e-3, mu=0.01, phiflag=1) as s:   # 192^3 > 128^3
        s.init_fields()
        assert lib.mhit_drop_count_noalloc(s._h, 0, ctypes.byref(cnt), None, None) == _lib.MHIT_ERR_BAD_N
    with Solver(nx=32, dt=1e-3, mu=0.01, phiflag=0) as s:                            # no phase field
        assert lib.mhit_drop_count_noalloc(s._h, 0, ctypes.byref(cnt), None, None) == _lib.MHIT_ERR_ARG
    assert lib.mhit_drop_count_noalloc(None, 0, ctypes.byref(cnt), None, None) == _lib.MHIT_ERR_ARG
