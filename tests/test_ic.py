"""The seeded initial-condition generators of mhit36_amd/ic.py on the CPU.

random_velocity takes u = curl A on the staggered grid (forward differences
over dx = 2 pi / (N - 1), the spacing of main.f90:149-153) of a random vector
potential whose modes have E|A_k|^2 = k^4 exp(-2 (k/k_p)^2) / (4 pi k^4).
What that construction fixes, independent of how the code is written:

* the staggered divergence of the solver (main.f90:723-725) is zero to
  rounding, because forward differences commute;
* the mean of each component is zero (a difference of periodic values);
* u_rms = sqrt((<u^2> + <v^2> + <w^2>) / 3) is the requested value;
* the expected energy of mode k of u + v + w is 2 |D(k)|^2 E|A_k|^2, D the
  difference symbol (e^{2 pi i k / N} - 1) / dx per direction, so the shell
  spectrum follows k^4 exp(-2 (k/k_p)^2) up to |D(k)|^2 / k^2 and peaks
  near k_p.

drop_centres places centres at periodic separation >= 2R + 8 eps; drops is
max over drops of 0.5 (1 - tanh((r - R) / (2 eps))) (main.f90:273-274).
"""
import numpy as np
import pytest

from mhit36_amd import ic


def _div(u, v, w, dx):
    return ((np.roll(u, -1, 0) - u) + (np.roll(v, -1, 1) - v) + (np.roll(w, -1, 2) - w)) / dx


@pytest.mark.parametrize("n,seed", [(32, 1), (48, 7)])
def test_divergence_free_zero_mean_unit_rms(n, seed):
    u, v, w = ic.random_velocity(n, seed=seed, k_peak=4.0, urms=1.0)
    _, dx = ic.grid(n)
    umax = max(np.abs(u).max(), np.abs(v).max(), np.abs(w).max())
    assert np.abs(_div(u, v, w, dx)).max() * dx / umax < 1e-13
    for a in (u, v, w):
        assert abs(a.mean()) < 1e-14 * umax
    rms = np.sqrt((np.mean(u * u) + np.mean(v * v) + np.mean(w * w)) / 3.0)
    assert abs(rms - 1.0) < 1e-12
    # a central difference (not what the solver uses) does not vanish: the test sees the stencil
    cd = ((np.roll(u, -1, 0) - np.roll(u, 1, 0)) + (np.roll(v, -1, 1) - np.roll(v, 1, 1))
          + (np.roll(w, -1, 2) - np.roll(w, 1, 2))) / (2 * dx)
    assert np.abs(cd).max() * dx / umax > 1e-3


def _shells(n, e3):
    k = np.fft.fftfreq(n, 1.0 / n)
    kk = np.sqrt(k[:, None, None] ** 2 + k[None, :, None] ** 2 + k[None, None, :] ** 2)
    sh = np.rint(kk).astype(int)
    return np.bincount(sh.ravel(), weights=e3.ravel())


def test_shell_spectrum_follows_k4_exp_with_the_difference_symbol():
    n, kp = 64, 4.0
    _, dx = ic.grid(n)
    k = np.fft.fftfreq(n, 1.0 / n)
    KX, KY, KZ = np.meshgrid(k, k, k, indexing="ij")
    kk = np.sqrt(KX ** 2 + KY ** 2 + KZ ** 2)
    kmax = min(n // 2 - 1, int(6 * kp))
    with np.errstate(divide="ignore", invalid="ignore"):
        a2 = np.where((kk > 0) & (kk <= kmax), np.exp(-2.0 * (kk / kp) ** 2) / (4 * np.pi), 0.0)
    d2 = sum(4.0 * np.sin(np.pi * K / n) ** 2 for K in (KX, KY, KZ)) / dx ** 2
    expect = _shells(n, 2.0 * d2 * a2)
    got = np.zeros_like(expect)
    for seed in (11, 12, 13, 14):
        u, v, w = ic.random_velocity(n, seed=seed, k_peak=kp, urms=1.0)
        got += _shells(n, sum(np.abs(np.fft.fftn(a)) ** 2 for a in (u, v, w)))
    ks = np.arange(2, 13)
    ratio = (got[ks] / got[1:kmax + 1].sum()) / (expect[ks] / expect[1:kmax + 1].sum())
    assert np.abs(ratio - 1.0).max() < 0.3, ratio
    slope = np.polyfit(np.log(ks), np.log(ratio), 1)[0]
    assert abs(slope) < 0.25, slope
    # the model itself: a k^4 exp(-2 (k/kp)^2) shell law peaks at k = kp
    assert abs(int(np.argmax(got[1:kmax + 1])) + 1 - kp) <= 1
    # and a k^2 law (a plausible slip in the amplitude) would be caught by the slope
    alt = expect[ks] / ks ** 2
    alt_ratio = (got[ks] / got[ks].sum()) / (alt / alt.sum())
    assert abs(np.polyfit(np.log(ks), np.log(alt_ratio), 1)[0]) > 1.0


def test_random_velocity_is_seeded():
    a = ic.random_velocity(24, seed=5)
    b = ic.random_velocity(24, seed=5)
    c = ic.random_velocity(24, seed=6)
    assert all(np.array_equal(x, y) for x, y in zip(a, b))
    assert not np.array_equal(a[0], c[0])


def _periodic_dist(a, b, lx):
    d = np.abs(a - b)
    d = np.minimum(d, lx - d)
    return np.sqrt((d * d).sum(-1))


@pytest.mark.parametrize("ndrops,radius,n", [(64, 0.25, 64), (512, 0.1, 128)])
def test_drop_centres_separation(ndrops, radius, n):
    _, dx = ic.grid(n)
    eps = dx
    c = ic.drop_centres(ndrops, radius, eps, seed=4321)
    assert c.shape == (ndrops, 3)
    lx = 8.0 * np.arctan(1.0)
    assert (c >= 0).all() and (c < lx).all()
    d = _periodic_dist(c[:, None, :], c[None, :, :], lx)
    d[np.arange(ndrops), np.arange(ndrops)] = np.inf
    assert d.min() >= 2 * radius + 8 * eps


def test_drops_profile_and_volume():
    n, radius = 96, 0.5
    _, dx = ic.grid(n)
    eps = dx
    x, _ = ic.grid(n)
    c = np.array([[x[30], x[40], x[50]], [x[80], x[10], x[70]]])
    phi = ic.drops(n, c, radius, eps)
    assert phi.min() >= 0.0 and phi.max() <= 1.0
    # centre value 0.5 (1 - tanh(-R / 2 eps)) and the half level at r = R
    assert abs(phi[30, 40, 50] - 0.5 * (1 - np.tanh(-radius / 2 / eps))) < 1e-15
    # volume of phi >= 1/2 against two spheres of radius R (cell-counted)
    vol = (phi >= 0.5).sum() * dx ** 3
    assert abs(vol / (2 * 4.0 / 3.0 * np.pi * radius ** 3) - 1.0) < 0.05
    # a drop across the periodic seam is whole: centred on node 0 it is the drop centred on
    # node 48, rolled by 48 cells (the period is n dx, node n-1 is not node 0's image)
    ps = ic.drops(n, np.array([[x[0], x[40], x[50]]]), radius, eps)
    p0 = ic.drops(n, np.array([[x[48], x[40], x[50]]]), radius, eps)
    assert np.abs(np.roll(p0, -48, axis=0) - ps).max() < 1e-12
    assert (ps >= 0.5).sum() == (p0 >= 0.5).sum()
