"""Checks of the oracle's phase field, surface tension, convection of phi and
ABC forcing against what the mathematics fixes, independently of how the
oracle writes the reference's expressions.

Each function runs the oracle built at `lib` (None: oracle/liboracle.so),
with the sharpening factor in `form` ("identity", the form the HIP path
evaluates, or "literal", tanh of psi as written), and raises AssertionError
when its check fails.  tests/test_oracle_pins.py runs
them on the oracle; tests/test_oracle_mutants.py runs them on copies of the
oracle with one planted mistake each and requires every mistake to be caught.

Reference anchors (multi/main.f90):
  341-359  convection of phi (face velocities times face averages of phi)
  362-381  diffusion eps*lap(phi), times gamma
  384-443  psi = eps log(phi/(1-phi)), undivided normals, normalisation
  446-486  sharpening flux 1/4 (1 - tanh^2(psi/(2 eps))) n, Euler update
  584-600  ABC forcing at node coordinates
  604-658  CSF force 6 sigma mu grad(phi), mu = phi(1-phi)(1-2phi)/eps - eps lap(phi),
           interpolated to the faces
"""
import numpy as np

from oracle import orc


def _box(lib, n, form="identity", **kw):
    o = orc.Oracle(lib_path=lib, sharpening=form, nx=n, inflow=1, inphi=1, **kw)
    c = o.consts()
    x = np.arange(n) * c["dx"]
    return o, c, x


def _set(o, n, u=None, v=None, w=None, phi=None):
    z = np.zeros((n, n, n))
    for f, a in ((1, u), (2, v), (3, w), (5, phi)):
        if f == 5 and a is None:
            continue
        o.set_field(f, z if a is None else np.asfortranarray(a))


def _lap(f, dx):
    return sum(np.roll(f, 1, a) - 2 * f + np.roll(f, -1, a) for a in range(3)) / dx ** 2


# ---------------------------------------------------------------------------
# the hyperbolic-tangent profile is the ACDI equilibrium.
# For phi = 1/2 (1 - tanh((r - R)/(2 eps))), psi = eps log(phi/(1-phi)) is the
# signed distance (R - r), 1/4 (1 - tanh^2(psi/(2 eps))) = phi (1 - phi), the
# unit normal of psi is -r_hat and eps grad(phi) = -phi(1-phi) r_hat: the
# diffusive flux eps grad(phi) and the sharpening flux cancel pointwise, for a
# sphere as for a plane.  With u = 0 the discrete phi update is therefore pure
# truncation error, second order in dx/eps: at fixed R the residual, measured
# against eps |lap phi| on the interface band |r - R| < 3 eps, must fall about
# 4x each time eps/dx doubles.  A wrong sign, a factor 2 in the tanh factor or
# its argument, a one-sided normal or a missing eps leaves an O(1) or
# first-order residual.
# ---------------------------------------------------------------------------
def sphere_residuals(lib=None, n=128, R=2.0, epsrs=(1.0, 2.0, 4.0), form="identity"):
    out = []
    for er in epsrs:
        o, c, x = _box(lib, n, form, dt=1.0, mu=0.0, phiflag=1, sigma=0.0, epsr=er)
        eps, dx = c["eps"], c["dx"]
        X, Y, Z = np.meshgrid(x, x, x, indexing="ij")
        cx, cy, cz = x[n // 2] + 0.37 * dx, x[n // 2] - 0.21 * dx, x[n // 2] + 0.11 * dx
        d = np.sqrt((X - cx) ** 2 + (Y - cy) ** 2 + (Z - cz) ** 2)
        phi = 0.5 * (1.0 - np.tanh((d - R) / (2.0 * eps)))
        _set(o, n, phi=phi)
        o.step(1)                                     # dt = 1: phi^{n+1} - phi = rhs_phi
        rhs = o.get_field(5) - phi
        band = np.abs(d - R) < 3.0 * eps
        out.append(np.abs(rhs[band]).max() / np.abs(eps * _lap(phi, dx)[band]).max())
    return out


def check_sphere_equilibrium(lib=None, form="identity"):
    r1, r2, r4 = sphere_residuals(lib, form=form)
    assert r1 < 0.12, (r1, r2, r4)
    assert 3.3 < r1 / r2 < 4.8, (r1, r2, r4)
    assert 3.3 < r2 / r4 < 5.0, (r1, r2, r4)


# ---------------------------------------------------------------------------
# convection of phi.  With a uniform velocity U along axis a and a
# discrete Fourier mode phi_i = c0 + c1 cos(theta i + s) along the same axis
# (gamma = 0, so only the convection acts), the central face averages give exactly
#   rhs_i = U c1 sin(theta) sin(theta i + s) / dx,
# the eigenvalue of the centred difference on that mode.
# ---------------------------------------------------------------------------
def check_convection_mode(lib=None, form="identity", n=32, m=3, U=0.7, dt=1e-2):
    for axis in range(3):
        o, c, x = _box(lib, n, form, dt=dt, mu=0.0, phiflag=1, sigma=0.0, gamma_factor=0.0)
        dx = c["dx"]
        th = 2 * np.pi * m / n
        i = np.arange(n)
        sh = [1, 1, 1]
        sh[axis] = n
        phi = np.broadcast_to((0.5 + 0.25 * np.cos(th * i + 0.4)).reshape(sh), (n,) * 3)
        vel = [None, None, None]
        vel[axis] = np.full((n,) * 3, U)
        _set(o, n, *vel, phi=phi)
        o.step(1)
        rhs = (o.get_field(5) - phi) / dt
        want = np.broadcast_to((U * 0.25 * np.sin(th) * np.sin(th * i + 0.4) / dx).reshape(sh), (n,) * 3)
        err = np.abs(rhs - want).max() / np.abs(want).max()
        assert err < 1e-12, (axis, err)


# ---------------------------------------------------------------------------
# staggering: u(i) sits on the face i - 1/2 (main.f90:723-725, 881-883)
# and phi at the centre i.  Sampling smooth periodic u and phi there, the
# discrete convective term must approach -(u phi)' at x_i at second order
# (error ratio ~4 per doubling of N); velocities read from the wrong faces
# shift u by dx and leave a first-order error.
# ---------------------------------------------------------------------------
def convection_errors(lib=None, axis=0, ns=(32, 64, 128), dt=1e-2, form="identity"):
    errs = []
    for n in ns:
        o, c, x = _box(lib, n, form, dt=dt, mu=0.0, phiflag=1, sigma=0.0, gamma_factor=0.0)
        dx = c["dx"]
        k = 2 * np.pi / (n * dx)                      # one period over the n-point grid
        xc = np.arange(n) * dx
        xf = xc - 0.5 * dx
        ua = lambda s: 0.6 + 0.3 * np.sin(k * s + 0.2)
        pa = lambda s: 0.5 + 0.2 * np.cos(k * s - 0.5)
        dflux = (0.3 * k * np.cos(k * xc + 0.2)) * pa(xc) + ua(xc) * (-0.2 * k * np.sin(k * xc - 0.5))
        sh = [1, 1, 1]
        sh[axis] = n
        vel = [None, None, None]
        vel[axis] = np.broadcast_to(ua(xf).reshape(sh), (n,) * 3)
        phi = np.broadcast_to(pa(xc).reshape(sh), (n,) * 3)
        _set(o, n, *vel, phi=phi)
        o.step(1)
        rhs = (o.get_field(5) - phi) / dt
        want = np.broadcast_to((-dflux).reshape(sh), (n,) * 3)
        errs.append(np.abs(rhs - want).max() / np.abs(want).max())
    return errs


def check_convection_staggering(lib=None, form="identity"):
    for axis in range(3):
        e = convection_errors(lib, axis, form=form)
        assert e[0] < 0.02, (axis, e)
        assert 3.5 < e[0] / e[1] < 4.5 and 3.5 < e[1] / e[2] < 4.5, (axis, e)


# ---------------------------------------------------------------------------
# the chemical potential mu = phi(1-phi)(1-2phi)/eps - eps lap(phi)
# vanishes on the planar tanh profile (eps phi'' = phi(1-phi)(1-2phi)/eps
# there), so the CSF cell force 6 sigma mu grad(phi) of a plane slab is pure
# truncation error: against its natural size 6 sigma |phi(1-phi)(1-2phi)/eps|
# |grad phi| it must be small and fall ~4x per doubling of eps/dx.  A sign
# error in either term, or a wrong factor, leaves it O(1).  The force is
# checked in each axis, so a wrong neighbour in one direction shows too.
# ---------------------------------------------------------------------------
def csf_plane_residuals(lib=None, axis=0, n=128, epsrs=(1.0, 2.0, 4.0), R=1.5, form="identity"):
    out = []
    for er in epsrs:
        o, c, x = _box(lib, n, form, dt=1e-3, mu=0.0, phiflag=1, sigma=0.3, epsr=er, gamma_factor=0.0)
        eps, dx = c["eps"], c["dx"]
        p1 = 0.5 * (1.0 - np.tanh((np.abs(x - x[n // 2] - 0.3 * dx) - R) / (2.0 * eps)))
        sh = [1, 1, 1]
        sh[axis] = n
        phi = np.broadcast_to(p1.reshape(sh), (n,) * 3)
        _set(o, n, phi=phi)
        o.step(1)
        f = o.debug_buffer(15 + axis)                 # CSF cell force, this axis
        others = [o.debug_buffer(15 + a) for a in range(3) if a != axis]
        t1 = phi * (1 - phi) * (1 - 2 * phi) / eps
        g = (np.roll(phi, -1, axis) - np.roll(phi, 1, axis)) / (2 * dx)
        scale = np.abs(6 * 0.3 * t1 * g).max()
        out.append(np.abs(f).max() / scale)
        assert max(np.abs(a).max() for a in others) == 0.0, "tangential CSF force of a plane"
    return out


def check_csf_plane(lib=None, form="identity"):
    for axis in range(3):
        r1, r2, r4 = csf_plane_residuals(lib, axis, form=form)
        assert r1 < 0.3, (axis, r1, r2, r4)
        assert 3.3 < r1 / r2 < 4.8 and 3.3 < r2 / r4 < 4.8, (axis, r1, r2, r4)


# ---------------------------------------------------------------------------
# Laplace's law.  A drop of radius R at rest, no sharpening
# (gamma = 0), exact projection (FD2 modified wavenumbers, no REAL(4)
# rounding): after one step the pressure holds the CSF force, and its jump
# across the interface is 2 sigma / R up to O(eps/R).  The drop is centred on
# a cell centre, so the pressure must also be symmetric under each axis
# reflection about the centre and under axis permutations: forces
# interpolated to the wrong faces break that.
# ---------------------------------------------------------------------------
def laplace_jump(lib=None, n=96, R=1.0, epsr=1.0, sigma=0.1, form="identity"):
    o, c, x = _box(lib, n, form, dt=1e-3, mu=0.0, phiflag=1, sigma=sigma, epsr=epsr, gamma_factor=0.0,
                   wavenumbers=1, poisson_rhs_fp32=0)
    eps = c["eps"]
    X, Y, Z = np.meshgrid(x, x, x, indexing="ij")
    cc = x[n // 2]
    d = np.sqrt((X - cc) ** 2 + (Y - cc) ** 2 + (Z - cc) ** 2)
    phi = 0.5 * (1.0 - np.tanh((d - R) / (2.0 * eps)))
    _set(o, n, phi=phi)
    o.step(1)
    p = o.get_field(4)
    jump = (p[d < R - 4 * eps].mean() - p[d > R + 6 * eps].mean()) / (2 * sigma / R)
    idx = (2 * (n // 2) - np.arange(n)) % n            # reflection about the centre cell
    pm = np.abs(p).max()
    asym = max(np.abs(p - p[idx, :, :]).max(), np.abs(p - p[:, idx, :]).max(),
               np.abs(p - p[:, :, idx]).max(), np.abs(p - p.transpose(1, 0, 2)).max(),
               np.abs(p - p.transpose(2, 1, 0)).max()) / pm
    return jump, asym


def check_laplace(lib=None, form="identity"):
    j1, a1 = laplace_jump(lib, epsr=1.0, form=form)
    j2, a2 = laplace_jump(lib, epsr=2.0, form=form)
    assert abs(j1 - 1.0) < 0.05, j1
    assert abs(j2 - 1.0) < 0.02, j2
    assert a1 < 1e-11 and a2 < 1e-11, (a1, a2)


# ---------------------------------------------------------------------------
# ABC forcing, at the node coordinates x_i = (i-1) dx of each index
# (main.f90:591-593).  From rest (no convection, no viscous term, no CSF) the
# first, Euler, step gives pointwise
#   u* = dt (f3 sin(k0 z_k) + f2 cos(k0 y_j))
#   v* = dt (f1 sin(k0 x_i) + f3 cos(k0 z_k))
#   w* = dt (f2 sin(k0 y_j) + f1 cos(k0 x_i)).
# ---------------------------------------------------------------------------
def check_forcing(lib=None, form="identity", n=16, dt=1e-2, k0=2.0):
    f1, f2, f3 = 0.3, 0.5, 0.7
    o, c, _ = _box(lib, n, form, dt=dt, mu=0.01, f1=f1, f2=f2, f3=f3, k0=k0)
    x = np.arange(n) * c["dx"]
    _set(o, n)
    o.step(1)
    X, Y, Z = np.meshgrid(x, x, x, indexing="ij")
    want = (dt * (f3 * np.sin(k0 * Z) + f2 * np.cos(k0 * Y)),
            dt * (f1 * np.sin(k0 * X) + f3 * np.cos(k0 * Z)),
            dt * (f2 * np.sin(k0 * Y) + f1 * np.cos(k0 * X)))
    for b, wnt in zip((18, 19, 20), want):
        got = o.debug_buffer(b)
        err = np.abs(got - wnt).max()
        assert err < 1e-15, (b, err)


ALL = {
    "sphere_equilibrium": check_sphere_equilibrium,
    "convection_mode": check_convection_mode,
    "convection_staggering": check_convection_staggering,
    "csf_plane": check_csf_plane,
    "laplace": check_laplace,
    "forcing": check_forcing,
}
