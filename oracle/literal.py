"""Literal numpy restatement of one MHIT36 timestep -- TEST INFRASTRUCTURE.

A second, independent restatement of multi/main.f90:323-946 used to pin the C
oracle (oracle/mhit_oracle.c).  Where the C oracle replaces every halo plane
by periodic index wrap and uses its own FFT, this one keeps the reference's
X-pencil storage literally:

  * arrays are padded, shape (N, N+2, N+2), Fortran order; the interior is
    j, k in 1..N (0-based here), x is periodic by index wrap (main.f90:351-352);
  * `halo()` restates cudecompUpdateHalosX at a 1x1 decomposition: periodic
    plane copies along y (dim 2) then along z (dim 3), the z planes spanning
    the y halo rows so that the corner cells are filled;
  * the Poisson RHS is cmplx(rhsp, 0.0) -> COMPLEX(4) (main.f90:738), the 3-D
    transform is a full complex-to-complex DFT (numpy.fft, pocketfft), divided
    by -k^2 and N^3 with integer wavenumbers (main.f90:154-159, 791-798), and p
    is the real part of the unnormalised inverse (main.f90:842-850);
  * the velocity mean is a plain numpy sum (the reference's OpenACC reduction
    order is unspecified), removed from the padded arrays (main.f90:917-919).

Expressions keep the Fortran operand order; numpy evaluates elementwise
arithmetic without FMA contraction.  Small grids only (pure numpy).
"""
import numpy as np


class LiteralBox:
    def __init__(self, nx, dt, mu, rho=1.0, f1=0.0, f2=0.0, f3=0.0, k0=2.0, radius=0.5,
                 sigma=1.0, epsr=1.0, phiflag=0, gamma_factor=1.0, poisson_rhs_fp32=1):
        n = nx
        self.n = n
        self.phiflag = phiflag
        self.dt, self.mu, self.rho = dt, mu, rho
        self.f1, self.f2, self.f3, self.k0 = f1, f2, f3, k0
        self.radius, self.sigma = radius, sigma
        self.gamma_factor = gamma_factor
        self.fp32 = poisson_rhs_fp32
        # readinput.f90:35-43
        self.twopi = 8.0 * np.arctan(1.0)
        self.lx = self.twopi
        self.dx = self.lx / (n - 1)
        self.dxi = 1.0 / self.dx
        self.ddxi = 1.0 / self.dx / self.dx
        self.rhoi = 1.0 / rho
        self.eps = epsr * self.dx
        self.epsi = 1.0 / self.eps
        self.enum = float(np.float32(1.e-16))
        # main.f90:149-159
        x = np.zeros(n)
        for i in range(1, n):
            x[i] = x[i - 1] + self.dx
        self.x = x
        kx = np.array([(i - 1) * (self.twopi / self.lx) if i <= n // 2 else (i - 1 - n) * (self.twopi / self.lx)
                       for i in range(1, n + 1)])
        self.kx = kx
        shp = (n, n + 2, n + 2)
        z = lambda: np.zeros(shp, order="F")
        self.u, self.v, self.w = z(), z(), z()
        self.rhsu_o, self.rhsv_o, self.rhsw_o = z(), z(), z()
        self.phi = z()
        self.alpha, self.beta = 1.0, 0.0          # main.f90:323-324
        self.gumax = 1.0                          # main.f90:325

    # -- helpers ----------------------------------------------------------------
    def halo(self, a):
        n = self.n
        a[:, 0, :] = a[:, n, :]
        a[:, n + 1, :] = a[:, 1, :]
        a[:, :, 0] = a[:, :, n]
        a[:, :, n + 1] = a[:, :, 1]

    def set_interior(self, a, data):
        n = self.n
        a[:, 1:n + 1, 1:n + 1] = np.asarray(data).reshape((n, n, n), order="F")
        self.halo(a)

    def interior(self, a):
        n = self.n
        return np.asfortranarray(a[:, 1:n + 1, 1:n + 1])

    # -- one step (main.f90:329-946) -------------------------------------------
    def step(self):
        n = self.n
        dt, dxi, ddxi, rhoi = self.dt, self.dxi, self.ddxi, self.rhoi
        J = slice(1, n + 1)
        Jp = slice(2, n + 2)
        Jm = slice(0, n)
        c = (slice(None), J, J)
        xp = lambda a: np.roll(a, -1, axis=0)
        xm = lambda a: np.roll(a, 1, axis=0)
        u, v, w = self.u, self.v, self.w

        if self.phiflag:
            phi = self.phi
            eps, epsi = self.eps, self.epsi
            rhsphi = np.zeros_like(phi)
            P, Pxp, Pxm = phi[c], xp(phi)[c], xm(phi)[c]
            Pyp, Pym = phi[:, Jp, J], phi[:, Jm, J]
            Pzp, Pzm = phi[:, J, Jp], phi[:, J, Jm]
            # main.f90:353-355
            rhsphi[c] = (- (xp(u)[c] * 0.5 * (Pxp + P) - u[c] * 0.5 * (P + Pxm)) * dxi
                         - (v[:, Jp, J] * 0.5 * (Pyp + P) - v[c] * 0.5 * (P + Pym)) * dxi
                         - (w[:, J, Jp] * 0.5 * (Pzp + P) - w[c] * 0.5 * (P + Pzm)) * dxi)
            # main.f90:362-377
            gamma = self.gamma_factor * self.gumax
            rhsphi[c] = rhsphi[c] + gamma * (eps * (Pxp - 2.0 * P + Pxm) * ddxi
                                             + eps * (Pyp - 2.0 * P + Pym) * ddxi
                                             + eps * (Pzp - 2.0 * P + Pzm) * ddxi)
            # main.f90:384-393 (whole padded array)
            val = np.minimum(phi, 1.0)
            psidi = eps * np.log((val + self.enum) / (1.0 - val + self.enum))
            # main.f90:397-414 + halos 418-429
            normx, normy, normz = np.zeros_like(phi), np.zeros_like(phi), np.zeros_like(phi)
            normx[c] = xp(psidi)[c] - xm(psidi)[c]
            normy[c] = psidi[:, Jp, J] - psidi[:, Jm, J]
            normz[c] = psidi[:, J, Jp] - psidi[:, J, Jm]
            for a in (normx, normy, normz):
                self.halo(a)
            # main.f90:433-442
            normod = 1.0 / (np.sqrt(normx ** 2 + normy ** 2 + normz ** 2) + float(np.float32(1.0E-16)))
            normx, normy, normz = normx * normod, normy * normod, normz * normod
            # main.f90:463-465
            th = lambda q: np.tanh(0.5 * q * epsi)
            fxp = 0.25 * (1.0 - th(xp(psidi)[c]) ** 2) * xp(normx)[c]
            fxm = 0.25 * (1.0 - th(xm(psidi)[c]) ** 2) * xm(normx)[c]
            fyp = 0.25 * (1.0 - th(psidi[:, Jp, J]) ** 2) * normy[:, Jp, J]
            fym = 0.25 * (1.0 - th(psidi[:, Jm, J]) ** 2) * normy[:, Jm, J]
            fzp = 0.25 * (1.0 - th(psidi[:, J, Jp]) ** 2) * normz[:, J, Jp]
            fzm = 0.25 * (1.0 - th(psidi[:, J, Jm]) ** 2) * normz[:, J, Jm]
            rhsphi[c] = rhsphi[c] - gamma * ((fxp - fxm) * 0.5 * dxi + (fyp - fym) * 0.5 * dxi
                                             + (fzp - fzm) * 0.5 * dxi)
            # main.f90:476 + halos 484-485
            phi[c] = phi[c] + dt * rhsphi[c]
            self.halo(phi)

        # convection main.f90:527-549
        U, V, W = u, v, w
        u0, v0, w0 = U[c], V[c], W[c]
        h11 = (xp(U)[c] + u0) * (xp(U)[c] + u0) - (u0 + xm(U)[c]) * (u0 + xm(U)[c])
        h12 = (U[:, Jp, J] + u0) * (V[:, Jp, J] + xm(V)[:, Jp, J]) - (u0 + U[:, Jm, J]) * (v0 + xm(V)[c])
        h13 = (U[:, J, Jp] + u0) * (W[:, J, Jp] + xm(W)[:, J, Jp]) - (u0 + U[:, J, Jm]) * (w0 + xm(W)[c])
        h21 = (xp(U)[c] + xp(U)[:, Jm, J]) * (xp(V)[c] + v0) - (u0 + U[:, Jm, J]) * (v0 + xm(V)[c])
        h22 = (V[:, Jp, J] + v0) * (V[:, Jp, J] + v0) - (v0 + V[:, Jm, J]) * (v0 + V[:, Jm, J])
        h23 = (W[:, J, Jp] + W[:, Jm, Jp]) * (V[:, J, Jp] + v0) - (w0 + W[:, Jm, J]) * (v0 + V[:, J, Jm])
        h31 = (xp(W)[c] + w0) * (xp(U)[c] + xp(U)[:, J, Jm]) - (w0 + xm(W)[c]) * (u0 + U[:, J, Jm])
        h32 = (V[:, Jp, J] + V[:, Jp, Jm]) * (W[:, Jp, J] + w0) - (v0 + V[:, J, Jm]) * (w0 + W[:, Jm, J])
        h33 = (W[:, J, Jp] + w0) * (W[:, J, Jp] + w0) - (w0 + W[:, J, Jm]) * (w0 + W[:, J, Jm])
        hs = [h * 0.25 * dxi for h in (h11, h12, h13, h21, h22, h23, h31, h32, h33)]
        rhsu = -(hs[0] + hs[1] + hs[2])
        rhsv = -(hs[3] + hs[4] + hs[5])
        rhsw = -(hs[6] + hs[7] + hs[8])
        # viscous main.f90:567-578
        mu = self.mu
        lap = lambda A: (mu * (xp(A)[c] - 2.0 * A[c] + xm(A)[c]) * ddxi,
                         mu * (A[:, Jp, J] - 2.0 * A[c] + A[:, Jm, J]) * ddxi,
                         mu * (A[:, J, Jp] - 2.0 * A[c] + A[:, J, Jm]) * ddxi)
        a1, a2, a3 = lap(U)
        rhsu = rhsu + (a1 + a2 + a3) * rhoi
        a1, a2, a3 = lap(V)
        rhsv = rhsv + (a1 + a2 + a3) * rhoi
        a1, a2, a3 = lap(W)
        rhsw = rhsw + (a1 + a2 + a3) * rhoi
        # ABC forcing main.f90:585-593 (interior plane j at y = x(j - halo_ext))
        x = self.x
        k0 = self.k0
        sx = np.sin(k0 * x)[:, None, None]
        cx = np.cos(k0 * x)[:, None, None]
        sy = np.sin(k0 * x)[None, :, None]
        cy = np.cos(k0 * x)[None, :, None]
        sz = np.sin(k0 * x)[None, None, :]
        cz = np.cos(k0 * x)[None, None, :]
        rhsu = rhsu + self.f3 * sz + self.f2 * cy
        rhsv = rhsv + self.f1 * sx + self.f3 * cz
        rhsw = rhsw + self.f2 * sy + self.f1 * cx
        # CSF main.f90:606-657
        if self.phiflag:
            phi = self.phi
            eps, epsi = self.eps, self.epsi
            P, Pxp, Pxm = phi[c], xp(phi)[c], xm(phi)[c]
            Pyp, Pym = phi[:, Jp, J], phi[:, Jm, J]
            Pzp, Pzm = phi[:, J, Jp], phi[:, J, Jm]
            chempot = P * (1.0 - P) * (1.0 - 2.0 * P) * epsi - eps * (Pxp + Pxm + Pyp + Pym + Pzp + Pzm - 6.0 * P) * ddxi
            fxst, fyst, fzst = np.zeros_like(phi), np.zeros_like(phi), np.zeros_like(phi)
            fxst[c] = 6.0 * self.sigma * chempot * (0.5 * (Pxp - Pxm) * dxi)
            fyst[c] = 6.0 * self.sigma * chempot * (0.5 * (Pyp - Pym) * dxi)
            fzst[c] = 6.0 * self.sigma * chempot * (0.5 * (Pzp - Pzm) * dxi)
            for a in (fxst, fyst, fzst):
                self.halo(a)
            rhsu = rhsu + 0.5 * (xm(fxst)[c] + fxst[c]) * rhoi
            rhsv = rhsv + 0.5 * (fyst[:, Jm, J] + fyst[c]) * rhoi
            rhsw = rhsw + 0.5 * (fzst[:, J, Jm] + fzst[c]) * rhoi
        # main.f90:666-681, halos 686-695
        ustar, vstar, wstar = np.zeros_like(u), np.zeros_like(u), np.zeros_like(u)
        ustar[c] = u0 + dt * (self.alpha * rhsu - self.beta * self.rhsu_o[c])
        vstar[c] = v0 + dt * (self.alpha * rhsv - self.beta * self.rhsv_o[c])
        wstar[c] = w0 + dt * (self.alpha * rhsw - self.beta * self.rhsw_o[c])
        self.alpha, self.beta = 1.5, 0.5
        self.rhsu_o[c], self.rhsv_o[c], self.rhsw_o[c] = rhsu, rhsv, rhsw
        for a in (ustar, vstar, wstar):
            self.halo(a)
        # divergence main.f90:723-725
        cd = self.rho * dxi / dt
        rhsp = cd * (xp(ustar)[c] - ustar[c])
        rhsp = rhsp + cd * (vstar[:, Jp, J] - vstar[c])
        rhsp = rhsp + cd * (wstar[:, J, Jp] - wstar[c])
        # main.f90:738
        if self.fp32:
            rhsp = rhsp.astype(np.float32).astype(np.float64)
        # Z2Z forward, divide, inverse, real part
        psi = np.fft.fftn(rhsp.astype(np.complex128), norm="backward")
        kk = self.kx
        k2 = kk[:, None, None] ** 2 + kk[None, :, None] ** 2 + kk[None, None, :] ** 2
        with np.errstate(divide="ignore", invalid="ignore"):
            psi = -psi / k2 / float(n * n * n)
        psi[0, 0, 0] = 0.0
        pint = np.real(np.fft.ifftn(psi, norm="forward"))
        p = np.zeros_like(u)
        p[c] = pint
        self.halo(p)
        # projection main.f90:881-883
        u[c] = ustar[c] - dt / self.rho * (p[c] - xm(p)[c]) * dxi
        v[c] = vstar[c] - dt / self.rho * (p[c] - p[:, Jm, J]) * dxi
        w[c] = wstar[c] - dt / self.rho * (p[c] - p[:, J, Jm]) * dxi
        for a in (u, v, w):
            self.halo(a)
        # mean removal main.f90:890-919
        gum = np.sum(u[c]) / n / n / n
        gvm = np.sum(v[c]) / n / n / n
        gwm = np.sum(w[c]) / n / n / n
        u -= gum
        v -= gvm
        w -= gwm
        for a in (u, v, w):
            self.halo(a)
        # main.f90:937-942 signed max over the padded arrays
        self.gumax = max(w.max(), max(u.max(), v.max()))
        self.cou = self.gumax * dt * dxi
        self.p = p
        return dict(gumax=self.gumax, cou=self.cou, umean=gum, vmean=gvm, wmean=gwm)
