"""Drop counting on the device (mhit_drop_count, post/DropCountSize) equals
the oracle's flood fill: count, order, volumes and diameters."""
import numpy as np
import pytest

from mhit36_amd import Solver, ic
from oracle import orc

pytestmark = pytest.mark.gpu


def device_count(n, phi, precision=0):
    with Solver(nx=n, dt=1e-3, mu=0.01, phiflag=1, inphi=1, precision=precision) as s:
        s.set_field(5, phi)
        cnt, vol, deq = s.drop_count()
        used = s.get_field(5)          # phi as stored (FP32 in mixed precision)
    return cnt, vol, deq, used


@pytest.mark.parametrize("seed,power", [(0, 3), (1, 6), (2, 12)])
@pytest.mark.parametrize("precision", [0, 1])
def test_random_fields(seed, power, precision):
    n = 48
    phi = np.random.default_rng(seed).uniform(0, 1, (n, n, n)) ** power
    cnt, vol, deq, used = device_count(n, phi, precision)
    rc, rvol, rdeq = orc.drop_count(used)
    assert cnt == rc
    assert np.array_equal(vol, rvol)
    assert np.array_equal(deq, rdeq)


def test_drops_initial_condition():
    n = 96                                         # radix-3 size
    _, dx = ic.grid(n)
    centres = ic.drop_centres(24, 0.3, dx, seed=7)
    phi = ic.drops(n, centres, 0.3, dx)
    cnt, vol, deq, used = device_count(n, phi)
    rc, rvol, rdeq = orc.drop_count(used)
    assert cnt == rc == 24
    assert np.array_equal(vol, rvol) and np.array_equal(deq, rdeq)
    assert np.allclose(deq, 0.6, rtol=0.1)          # diameter of a drop of radius 0.3


def test_wrapping_and_single_cells():
    n = 32
    phi = np.zeros((n, n, n))
    phi[0, 5, 5] = phi[n - 1, 6, 6] = 1.0          # across the x boundary, diagonal
    phi[::4, ::4, 20] = 1.0                         # 64 isolated cells
    cnt, vol, deq, used = device_count(n, phi)
    rc, rvol, rdeq = orc.drop_count(used)
    assert cnt == rc == 65
    assert np.array_equal(vol, rvol) and np.array_equal(deq, rdeq)


def test_as_many_drops_as_the_box_holds():
    """Isolated cells two apart in every direction: n^3/8 drops, the most a
    box can hold, spread over every chunk of the device's root listing."""
    n = 32
    phi = np.zeros((n, n, n))
    phi[::2, ::2, ::2] = 1.0
    cnt, vol, deq, used = device_count(n, phi)
    rc, rvol, rdeq = orc.drop_count(used)
    assert cnt == rc == n ** 3 // 8
    assert np.array_equal(vol, rvol) and np.array_equal(deq, rdeq)


def test_workspace_reused_across_calls():
    """A context keeps its labelling workspace: a second count on another
    phi sees nothing of the first."""
    n = 48
    a = np.random.default_rng(3).uniform(0, 1, (n, n, n)) ** 4
    b = np.random.default_rng(4).uniform(0, 1, (n, n, n)) ** 9
    with Solver(nx=n, dt=1e-3, mu=0.01, phiflag=1, inphi=1) as s:
        for phi in (a, b, a):
            s.set_field(5, phi)
            cnt, vol, deq = s.drop_count()
            rc, rvol, rdeq = orc.drop_count(s.get_field(5))
            assert cnt == rc and np.array_equal(vol, rvol) and np.array_equal(deq, rdeq)
