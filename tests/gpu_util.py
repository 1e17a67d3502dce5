"""Helpers shared by the GPU parity tests."""
import numpy as np

from oracle import orc


def random_velocity(n, seed, amp=1.0):
    """Smooth random periodic velocity (a few low Fourier modes), seeded."""
    rng = np.random.default_rng(seed)
    x = 2 * np.pi * np.arange(n) / n
    X, Y, Z = np.meshgrid(x, x, x, indexing="ij")
    out = []
    for _ in range(3):
        f = np.zeros((n, n, n))
        for _m in range(4):
            kx, ky, kz = rng.integers(-3, 4, size=3)
            ph = rng.uniform(0, 2 * np.pi)
            f += rng.normal() * np.cos(kx * X + ky * Y + kz * Z + ph)
        out.append(np.asfortranarray(amp * f))
    return out


def drops_phi(n, centers, radius, eps):
    x = 2 * np.pi * np.arange(n) / (n - 1)
    X, Y, Z = np.meshgrid(x, x, x, indexing="ij")
    phi = np.zeros((n, n, n))
    L = x[-1] + x[1]
    for c in centers:
        d2 = np.zeros_like(X)
        for A, cc in ((X, c[0]), (Y, c[1]), (Z, c[2])):
            d = np.abs(A - cc)
            d = np.minimum(d, L - d)
            d2 = d2 + d * d
        phi = np.maximum(phi, 0.5 * (1.0 - np.tanh((np.sqrt(d2) - radius) / 2 / eps)))
    return np.asfortranarray(phi)


def run_pair(kw, steps, fields=None, chunks=None):
    """Run the GPU solver and the oracle from the same state; return both."""
    from mhit36_amd import Solver
    s = Solver(**kw)
    o = orc.Oracle(**kw)
    if fields is None:
        s.init_fields()
        o.init_ic()
    else:
        for f, a in fields.items():
            s.set_field(f, a)
            o.set_field(f, a)
    chunks = chunks or [steps]
    ds, do = None, None
    for c in chunks:
        ds = s.step(c)
        do = o.step(c)
    return s, o, ds, do


def max_rel(a, b):
    den = max(np.abs(b).max(), 1e-300)
    return float(np.abs(a - b).max() / den)
