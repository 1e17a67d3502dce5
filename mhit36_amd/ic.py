"""Seeded synthetic initial conditions for the forced-HIT and drop configurations.

The reference reads such fields from ./input/{u,v,w,phi}.dat (inflow=1,
inphi=1; readwrite.f90:141-194).  These generators write the same layout:
interior N^3 FP64 arrays in Fortran order.

* random_velocity: random-phase field with E(k) ~ k^4 exp(-2 (k/k_p)^2),
  made discretely divergence-free on the staggered grid by taking it as the
  staggered curl of a random vector potential (the divergence used by the
  solver, main.f90:723-725, then vanishes up to rounding), scaled to u_rms.
* drops: non-overlapping spherical drops, phi = max over drops of
  0.5 (1 - tanh((r - R) / (2 eps))) with periodic distances (the profile of
  main.f90:273-274).
"""
import numpy as np

try:
    import scipy.fft as _fft
    _FFT_KW = {"workers": -1}
except ImportError:  # numpy fallback for the transform only
    _fft = np.fft
    _FFT_KW = {}


def grid(n):
    """Node coordinates of main.f90:149-153 (x(1)=0, x(i)=x(i-1)+dx, dx=2pi/(n-1))."""
    dx = (8.0 * np.arctan(1.0)) / (n - 1)
    x = np.zeros(n)
    for i in range(1, n):
        x[i] = x[i - 1] + dx
    return x, dx


def random_velocity(n, seed=1234, k_peak=4.0, urms=1.0, kmax=None):
    rng = np.random.default_rng(seed)
    kmax = kmax if kmax is not None else min(n // 2 - 1, int(6 * k_peak))
    k1 = np.fft.fftfreq(n, 1.0 / n)
    kr = np.fft.rfftfreq(n, 1.0 / n)
    KX = k1[:, None, None]
    KY = k1[None, :, None]
    KZ = kr[None, None, :]
    kk = np.sqrt(KX ** 2 + KY ** 2 + KZ ** 2)
    with np.errstate(divide="ignore", invalid="ignore"):
        amp = np.where((kk > 0) & (kk <= kmax),
                       np.sqrt(kk ** 4 * np.exp(-2.0 * (kk / k_peak) ** 2) / (4 * np.pi * kk ** 4)), 0.0)
    del KX, KY, KZ, kk
    _, dx = grid(n)
    A = []
    for _c in range(3):
        ph = rng.uniform(0, 2 * np.pi, size=amp.shape)
        mag = rng.standard_normal(size=amp.shape)
        spec = (amp * mag * np.exp(1j * ph)).astype(np.complex128)
        del ph, mag
        # axes ordered (x, y, z) in memory as Fortran order: transform over all three
        a = _fft.irfftn(spec, s=(n, n, n), axes=(0, 1, 2), **_FFT_KW)
        A.append(a)
        del spec
    Ax, Ay, Az = A
    d = lambda a, ax: (np.roll(a, -1, axis=ax) - a) / dx
    # staggered curl: u = dy Az - dz Ay, v = dz Ax - dx Az, w = dx Ay - dy Ax
    u = d(Az, 1) - d(Ay, 2)
    v = d(Ax, 2) - d(Az, 0)
    w = d(Ay, 0) - d(Ax, 1)
    del A, Ax, Ay, Az
    rms = np.sqrt((np.mean(u * u) + np.mean(v * v) + np.mean(w * w)) / 3.0)
    s = urms / rms
    return tuple(np.asfortranarray(a * s) for a in (u, v, w))


def drop_centres(ndrops, radius, eps, seed=4321, lx=None, max_tries=200000):
    """Random centres in [0, lx) with periodic separation >= 2R + 8 eps over
    the period lx (default 2 pi; separation over 2 pi implies it over the
    slightly longer period n dx that drops() uses)."""
    rng = np.random.default_rng(seed)
    lx = lx if lx is not None else 8.0 * np.arctan(1.0)
    sep = 2 * radius + 8 * eps
    centres = []
    tries = 0
    while len(centres) < ndrops:
        tries += 1
        if tries > max_tries:
            raise RuntimeError(f"could only place {len(centres)} of {ndrops} drops")
        c = rng.uniform(0, lx, size=3)
        ok = True
        for q in centres:
            dd = np.abs(c - q)
            dd = np.minimum(dd, lx - dd)
            if np.sqrt(np.sum(dd * dd)) < sep:
                ok = False
                break
        if ok:
            centres.append(c)
    return np.array(centres)


def drops(n, centres, radius, eps):
    """Phase field of spheres at `centres` on the node grid of main.f90:149-153.

    Distances are periodic over the n cells of the box, whose period is
    n dx = 2 pi n / (n - 1) on this grid (node n-1 sits at 2 pi, one cell short
    of node 0's periodic image), so a drop that crosses the seam is as whole
    as one in the middle of the box."""
    x, dx = grid(n)
    lx = n * dx
    phi = np.zeros((n, n, n), order="F")
    half = radius + 40 * eps      # profile below 1e-17 beyond: no jump at the box edge
    w = min(int(np.ceil(half / dx)) + 1, n // 2)     # beyond n/2 the box only repeats cells
    for c in centres:
        idx = []
        dist = []
        for ax in range(3):
            i0 = int(np.round(c[ax] / dx))
            ii = np.arange(i0 - w, i0 + w + 1) % n
            d = np.abs(x[ii] - c[ax])
            d = np.minimum(d, lx - d)
            idx.append(ii)
            dist.append(d)
        r = np.sqrt(dist[0][:, None, None] ** 2 + dist[1][None, :, None] ** 2 + dist[2][None, None, :] ** 2)
        val = 0.5 * (1.0 - np.tanh((r - radius) / 2 / eps))
        sub = phi[np.ix_(idx[0], idx[1], idx[2])]
        phi[np.ix_(idx[0], idx[1], idx[2])] = np.maximum(sub, val)
    return phi
