"""Host restatement of the device random-phase velocity generator
(mhit_init_random_velocity) -- TEST INFRASTRUCTURE: the checker the GPU tests
compare the device field with, never the product path.

The reference reads its turbulent initial field from input/{u,v,w}.dat
(inflow = 1; main.f90:234-239, readwrite.f90:107-194) and ships no generator.
This one, and the device's, draw a random vector potential A on the half
spectrum (x the halved axis, as the device's FFT stores it):

  A_c(q, ky, kz) = amp(k) g e^{i theta},   amp(k) = sqrt(exp(-2 (k/k_p)^2) / (4 pi))
                                           for 0 < k <= k_max, else 0,

with g a standard normal (Box-Muller) and theta uniform on [0, 2 pi), both
from a counter-based hash of (seed, component, kz, ky, q) -- SplitMix64
(Steele, Lea, Flood, "Fast splittable pseudorandom number generators",
OOPSLA 2014) -- so the field does not depend on the order in which it is
made.  A is brought to physical space by an unnormalised inverse real DFT,
and u = curl A is taken with the forward differences of the staggered grid
(u on x-faces, v on y-faces, w on z-faces), which makes the discrete
divergence of main.f90:723-725 vanish up to rounding.  |u_k| ~ k |A_k| gives
E(k) ~ k^4 exp(-2 (k/k_p)^2).  Finally u, v, w are scaled to the requested
u_rms = sqrt(<u^2 + v^2 + w^2> / 3).
"""
import numpy as np

M64 = np.uint64(0xFFFFFFFFFFFFFFFF)


def splitmix64(x):
    """SplitMix64 finaliser on uint64 arrays (wrapping arithmetic)."""
    x = np.asarray(x, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = x + np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
    return z ^ (z >> np.uint64(31))


def uniform(h):
    """[0, 1) from the top 53 bits."""
    return (h >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


def kmax_default(n, k_peak):
    return float(min(n // 2 - 1, int(6 * k_peak)))


def wavenumbers(n):
    k = np.arange(n, dtype=np.float64)
    k[k >= n // 2 + (n % 2)] -= n            # 0..n/2-1, -n/2..-1 (fftfreq order)
    return k


def potential_spectrum(n, comp, seed, k_peak=4.0, kmax=None):
    """Half spectrum of A_comp, shape (n/2+1, n, n) indexed [q, ky, kz]."""
    kmax = kmax_default(n, k_peak) if kmax is None else kmax
    nh = n // 2 + 1
    kq = np.arange(nh, dtype=np.float64)[:, None, None]
    kf = wavenumbers(n)
    ky = kf[None, :, None]
    kz = kf[None, None, :]
    kk = np.sqrt((kq * kq + ky * ky) + kz * kz)
    q = np.arange(nh, dtype=np.uint64)[:, None, None]
    j = np.arange(n, dtype=np.uint64)[None, :, None]
    k = np.arange(n, dtype=np.uint64)[None, None, :]
    un = np.uint64(n)
    with np.errstate(over="ignore"):
        idx = ((np.uint64(comp) * un + k) * un + j) * np.uint64(nh) + q
        key = splitmix64(np.uint64(seed)) ^ (idx * np.uint64(4))
    u1 = uniform(splitmix64(key))
    u2 = uniform(splitmix64(key + np.uint64(1)))
    u3 = uniform(splitmix64(key + np.uint64(2)))
    twopi = 2.0 * np.pi
    g = np.sqrt(-2.0 * np.log(1.0 - u1)) * np.cos(twopi * u2)
    th = twopi * u3
    amp = np.where((kk > 0.0) & (kk <= kmax), np.sqrt(np.exp(-2.0 * (kk / k_peak) ** 2) / (4.0 * np.pi)), 0.0)
    a = amp * g
    return a * np.cos(th) + 1j * (a * np.sin(th))


def potential(n, comp, seed, k_peak=4.0, kmax=None):
    """A_comp on the grid, (n, n, n) indexed [i, j, k]: unnormalised inverse real DFT."""
    spec = potential_spectrum(n, comp, seed, k_peak, kmax)
    return np.fft.irfftn(spec, s=(n, n, n), axes=(1, 2, 0), norm="forward")


def random_velocity(n, seed=1234, k_peak=4.0, urms=1.0, kmax=None):
    dx = (8.0 * np.arctan(1.0)) / (n - 1)
    Ax, Ay, Az = (potential(n, c, seed, k_peak, kmax) for c in range(3))
    d = lambda a, ax: (np.roll(a, -1, axis=ax) - a) / dx
    u = d(Az, 1) - d(Ay, 2)
    v = d(Ax, 2) - d(Az, 0)
    w = d(Ay, 0) - d(Ax, 1)
    rms = np.sqrt((np.sum(u * u) + np.sum(v * v) + np.sum(w * w)) / (3.0 * n ** 3))
    s = urms / rms
    return tuple(np.asfortranarray(a * s) for a in (u, v, w))
