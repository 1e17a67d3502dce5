"""The oracle's FFT and Poisson solve (main.f90:732-850) against numpy's
pocketfft, a brute-force DFT and discrete-mode closed forms."""
import numpy as np
import pytest

from oracle import orc

SIZES = [2, 4, 6, 8, 12, 16, 24, 32, 48, 64, 96, 128, 192, 256, 512]


def test_plan_radices():
    assert orc.plan_radices(512) == [8, 8, 8]
    assert orc.plan_radices(256) == [8, 8, 4]
    assert orc.plan_radices(128) == [8, 4, 4]
    assert orc.plan_radices(1024) == [8, 8, 4, 4]
    assert orc.plan_radices(96) == [8, 4, 3]
    assert orc.plan_radices(192) == [8, 8, 3]
    assert orc.plan_radices(2) == [2]
    assert orc.plan_radices(10) == []          # 5 is not a supported factor
    assert orc.plan_radices(27) == []          # n must be even
    for n in SIZES:
        assert int(np.prod(orc.plan_radices(n))) == n


@pytest.mark.parametrize("n", SIZES)
def test_fft_matches_numpy(n):
    rng = np.random.default_rng(n)
    z = rng.standard_normal((5, n)) + 1j * rng.standard_normal((5, n))
    f = orc.fft_1d(z, -1)
    b = orc.fft_1d(z, +1)
    assert np.abs(f - np.fft.fft(z, axis=-1)).max() <= 1e-14 * np.abs(f).max()
    assert np.abs(b - np.fft.ifft(z, axis=-1) * n).max() <= 1e-14 * np.abs(b).max()


@pytest.mark.parametrize("n", [6, 16, 24, 48])
def test_fft_matches_bruteforce_dft(n):
    rng = np.random.default_rng(100 + n)
    z = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    m = np.arange(n)
    W = np.exp(-2j * np.pi * np.outer(m, m) / n)
    ref = W @ z
    assert np.abs(orc.fft_1d(z[None, :], -1)[0] - ref).max() <= 1e-13 * np.abs(ref).max()
    assert np.abs(orc.fft_1d(ref[None, :], +1)[0] / n - z).max() <= 1e-13 * np.abs(z).max()


def _mode(n, mx, my, mz):
    i = np.arange(n)
    return np.cos(2 * np.pi * (mx * i[:, None, None] + my * i[None, :, None] + mz * i[None, None, :]) / n)


@pytest.mark.parametrize("n,m", [(16, (1, 0, 0)), (16, (2, 3, 1)), (24, (5, -2, 7)), (32, (16, 16, 16)),
                                 (32, (0, 0, 3))])
def test_poisson_discrete_mode_kat(n, m):
    """RHS = a discrete Fourier mode with integer wavenumber m: under the
    reference's integer-k table p = -rhs/|m|^2 exactly.  This is the
    commented-out KAT of main.f90:29, 739-741, redone with a true DFT mode."""
    rhs = np.asfortranarray(_mode(n, *m))
    p = orc.poisson(rhs, poisson_rhs_fp32=0)
    k2 = float(m[0] ** 2 + m[1] ** 2 + m[2] ** 2)
    assert np.abs(p - (-rhs / k2)).max() <= 1e-13


def test_poisson_mean_mode_is_dropped():
    n = 16
    rhs = np.asfortranarray(_mode(n, 1, 2, 0) + 3.0)
    p = orc.poisson(rhs, poisson_rhs_fp32=0)
    assert np.abs(p - (-_mode(n, 1, 2, 0) / 5.0)).max() <= 1e-13


@pytest.mark.parametrize("fp32", [0, 1])
def test_poisson_matches_numpy_z2z(fp32):
    """Against the reference's formulation: Z2Z 3-D DFT of cmplx(rhsp, 0.0),
    divide by -k^2/N^3, zero mean, unnormalised inverse, real part."""
    n = 24
    rng = np.random.default_rng(3)
    rhs = np.asfortranarray(rng.standard_normal((n, n, n)))
    p = orc.poisson(rhs, poisson_rhs_fp32=fp32)
    r = rhs.astype(np.float32).astype(np.float64) if fp32 else rhs
    k = np.array([i if i < n // 2 else i - n for i in range(n)], dtype=float)
    k2 = k[:, None, None] ** 2 + k[None, :, None] ** 2 + k[None, None, :] ** 2
    psi = np.fft.fftn(r.astype(complex))
    with np.errstate(divide="ignore", invalid="ignore"):
        psi = -psi / k2 / n ** 3
    psi[0, 0, 0] = 0
    ref = np.real(np.fft.ifftn(psi, norm="forward"))
    assert np.abs(p - ref).max() <= 1e-13 * np.abs(ref).max()


def test_poisson_modified_wavenumbers_invert_fd_laplacian():
    """With the FD2 modified wavenumbers the solve inverts the staggered
    div-grad Laplacian exactly (up to the mean)."""
    n = 16
    rng = np.random.default_rng(5)
    rhs = rng.standard_normal((n, n, n))
    rhs -= rhs.mean()
    rhs = np.asfortranarray(rhs)
    p = orc.poisson(rhs, poisson_rhs_fp32=0, wavenumbers=1)
    dxi = orc.consts(nx=n)["dxi"]
    lap = sum(np.roll(p, 1, a) - 2 * p + np.roll(p, -1, a) for a in range(3)) * dxi * dxi
    assert np.abs(lap - rhs).max() <= 1e-11 * np.abs(rhs).max()
