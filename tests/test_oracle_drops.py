"""The oracle's restatement of post/DropCountSize (get_interface.f90:11-47,
flood_fill.f90, calculate_deq.f90): constructed cases with known answers
(face, edge and corner contacts, periodic wrap on each axis, the >= 0.5
threshold) and an independent labelling of random fields (scipy.ndimage.label
with the 26-neighbourhood on a 3x3x3 periodic tiling, copies merged by
union-find)."""
import math

import numpy as np
import pytest

from oracle import orc

ndimage = pytest.importorskip("scipy.ndimage")


def deq(vol, n):
    return (2 * math.pi / (n - 1)) * (6.0 / math.pi * vol) ** (1.0 / 3.0)


def test_contacts_and_threshold():
    n = 12
    phi = np.zeros((n, n, n))
    phi[1, 1, 1] = phi[2, 2, 2] = 1.0          # corner contact: one drop of 2
    phi[5, 5, 5] = phi[6, 6, 5] = 0.5           # edge contact, threshold is >= 0.5: one drop of 2
    phi[9, 9, 9] = 0.4999999                    # below the threshold: nothing
    phi[9, 2, 9] = phi[10, 2, 9] = 0.8          # face contact: one drop of 2
    cnt, vol, d = orc.drop_count(phi)
    assert cnt == 3
    assert list(vol) == [2.0, 2.0, 2.0]
    assert np.allclose(d, [deq(2.0, n)] * 3, rtol=0, atol=1e-15)


@pytest.mark.parametrize("axis", [0, 1, 2])
def test_periodic_wrap(axis):
    n = 10
    phi = np.zeros((n, n, n))
    idx = [4, 4, 4]
    a = list(idx); a[axis] = 0
    b = list(idx); b[axis] = n - 1
    phi[tuple(a)] = phi[tuple(b)] = 1.0
    b2 = list(b); b2[(axis + 1) % 3] += 1       # diagonal across the boundary
    phi[tuple(b2)] = 1.0
    cnt, vol, _ = orc.drop_count(phi)
    assert cnt == 1 and vol[0] == 3.0


def test_order_is_raster_of_first_cell():
    n = 8
    phi = np.zeros((n, n, n))
    phi[5, 0, 0] = 1.0                          # linear index 5
    phi[1, 1, 0] = phi[1, 2, 0] = phi[1, 3, 0] = 1.0   # first cell index 9, three cells
    phi[0, 0, 3] = 1.0                          # index 192
    cnt, vol, _ = orc.drop_count(phi)
    assert cnt == 3 and list(vol) == [1.0, 3.0, 1.0]


def periodic_labels(top):
    """Independent labelling: scipy 26-connected labels of a 3x3x3 tiling,
    merged across the 27 copies of each cell."""
    n = top.shape[0]
    tiled = np.tile(top, (3, 3, 3))
    lab, nl = ndimage.label(tiled, structure=np.ones((3, 3, 3)))
    parent = list(range(nl + 1))

    def find(a):
        while parent[a] != a:
            parent[a] = parent[parent[a]]
            a = parent[a]
        return a

    centre = lab[n:2 * n, n:2 * n, n:2 * n]
    for ox in range(3):
        for oy in range(3):
            for oz in range(3):
                copy = lab[ox * n:(ox + 1) * n, oy * n:(oy + 1) * n, oz * n:(oz + 1) * n]
                m = centre > 0
                for a, b in zip(centre[m], copy[m]):
                    ra, rb = find(a), find(b)
                    if ra != rb:
                        parent[max(ra, rb)] = min(ra, rb)
    roots = np.zeros_like(centre)
    m = centre > 0
    roots[m] = [find(a) for a in centre[m]]
    return roots


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_random_fields_against_independent_labelling(seed):
    n = 14
    rng = np.random.default_rng(seed)
    phi = rng.uniform(0, 1, (n, n, n)) ** 3          # sparse enough to make many sets
    top = phi >= 0.5
    roots = periodic_labels(top)
    flat = roots.ravel(order="F")                    # x fastest: raster order of the reference
    first = {}
    sizes = {}
    for x, r in enumerate(flat):
        if r:
            first.setdefault(r, x)
            sizes[r] = sizes.get(r, 0) + 1
    order = sorted(first, key=first.get)
    cnt, vol, d = orc.drop_count(phi)
    assert cnt == len(order)
    assert list(vol) == [float(sizes[r]) for r in order]
    assert np.allclose(d, [deq(sizes[r], n) for r in order], rtol=1e-15, atol=0)
