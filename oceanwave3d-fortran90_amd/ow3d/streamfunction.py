"""Steady nonlinear (stream-function) waves by Fenton's Fourier method, for initial conditions and as
an analytic check of the time stepping: a wave of permanent form travels as eta(x - c t).

The reference computes the same family of solutions with Fenton's program
(utilities/stream_func_set_up.f90:46-51, utilities/stream_func_coeffs.f:1-153) and evaluates
eta and the surface potential with utilities/stream_func_wave_finite.f:3-71.  This is an independent
implementation of the published method (J. D. Fenton, "The numerical solution of steady water wave
problems", Computers & Geosciences 14 (1988) 357-368): Newton iteration on the nodal surface
elevations and the Fourier coefficients of the stream function, with the wave height reached in a
few continuation steps.

Frame moving with the wave, X = x - c t, z from the mean level, bed at z = -d:
    psi(X, z) = -ubar (z + d) + sum_j B_j sinh(j k (z + d)) / cosh(j k d) cos(j k X)
Free-surface conditions at z = eta(X): psi = -Q (kinematic), (u^2 + w^2)/2 + g eta = R (dynamic);
mean of eta zero; eta(0) - eta(pi/k) = H; zero Eulerian mean current (c = ubar).
"""
import math

import numpy as np

G = 9.81


class StreamFunctionWave:
    """Solution for wavenumber k, depth d, height H (zero mean Eulerian current)."""

    def __init__(self, k, d, H, N=32, g=G, steps=None, tol=1e-13, maxit=60):
        self.k, self.d, self.H, self.N, self.g = float(k), float(d), float(H), int(N), float(g)
        self._solve(steps or max(2, int(math.ceil(8 * H * k / (2 * math.pi * 0.142) + 1))), tol, maxit)

    # unknowns: eta_0..eta_N, B_1..B_N, ubar, c, Q, R  (2N + 5)
    def _residual(self, u, H):
        N, k, d, g = self.N, self.k, self.d, self.g
        eta = u[:N + 1]
        B = u[N + 1:2 * N + 1]
        ubar, c, Q, R = u[2 * N + 1:]
        X = np.arange(N + 1) * math.pi / (k * N)
        j = np.arange(1, N + 1)[None, :]
        zk = k * (eta[:, None] + d)
        ch = np.cosh(j * zk) / np.cosh(j * k * d)
        sh = np.sinh(j * zk) / np.cosh(j * k * d)
        cs = np.cos(j * k * X[:, None])
        sn = np.sin(j * k * X[:, None])
        psi = -ubar * (eta + d) + (sh * cs) @ B
        uu = -ubar + (j * k * ch * cs) @ B
        ww = (j * k * sh * sn) @ B
        f = np.empty(2 * N + 5)
        f[:N + 1] = psi + Q
        f[N + 1:2 * N + 2] = 0.5 * (uu * uu + ww * ww) + g * eta - R
        wts = np.ones(N + 1); wts[0] = wts[-1] = 0.5
        f[2 * N + 2] = (wts @ eta) / N
        f[2 * N + 3] = eta[0] - eta[N] - H
        f[2 * N + 4] = c - ubar
        return f

    def _solve(self, steps, tol, maxit):
        N, k, d, g = self.N, self.k, self.d, self.g
        c0 = math.sqrt(g * math.tanh(k * d) / k)
        X = np.arange(N + 1) * math.pi / (k * N)
        u = np.zeros(2 * N + 5)
        for s in range(1, steps + 1):
            H = self.H * s / steps
            if s == 1:
                a = H / 2
                u[:N + 1] = a * np.cos(k * X)
                u[N + 1] = c0 * a / math.tanh(k * d)
                u[2 * N + 1] = c0; u[2 * N + 2] = c0
                u[2 * N + 3] = c0 * d
                u[2 * N + 4] = 0.5 * c0 * c0
            for _ in range(maxit):
                f = self._residual(u, H)
                J = np.empty((u.size, u.size))
                for q in range(u.size):
                    h = 1e-7 * max(1.0, abs(u[q]))
                    up = u.copy(); up[q] += h
                    um = u.copy(); um[q] -= h
                    J[:, q] = (self._residual(up, H) - self._residual(um, H)) / (2 * h)
                du = np.linalg.solve(J, -f)
                u += du
                if np.max(np.abs(du)) < tol * max(1.0, np.max(np.abs(u))):
                    break
            else:
                raise RuntimeError("stream-function iteration did not converge")
        self.u = u
        self.eta_nodes = u[:N + 1]
        self.B = u[N + 1:2 * N + 1]
        self.ubar, self.c, self.Q, self.R = u[2 * N + 1:]
        # cosine series of eta from the nodal values (trapezoidal rule over the half wavelength)
        jj = np.arange(N + 1)
        wts = np.ones(N + 1); wts[0] = wts[-1] = 0.5
        C = np.cos(np.outer(jj, jj) * math.pi / N)
        self.E = (2.0 / N) * (C * wts[None, :]) @ self.eta_nodes
        self.E[0] *= 0.5
        self.E[N] *= 0.5

    @property
    def wavelength(self):
        return 2 * math.pi / self.k

    @property
    def period(self):
        return self.wavelength / self.c

    def eta(self, x, t=0.0):
        X = np.asarray(x, dtype=float) - self.c * t
        j = np.arange(self.N + 1)
        return np.cos(self.k * np.multiply.outer(X, j)) @ self.E

    def potential(self, x, z, t=0.0):
        """Velocity potential in the fixed frame (zero Eulerian mean current)."""
        X = np.asarray(x, dtype=float) - self.c * t
        z = np.asarray(z, dtype=float)
        j = np.arange(1, self.N + 1)
        k, d = self.k, self.d
        ch = np.cosh(k * np.multiply.outer(z + d, j)) / np.cosh(j * k * d)
        return np.sum(self.B * ch * np.sin(k * np.multiply.outer(X, j)), axis=-1)

    def surface(self, x, t=0.0):
        """eta and the surface potential Phi~ = phi(x, eta(x)) on the nodes x."""
        e = self.eta(x, t)
        return e, self.potential(x, e, t)


def sf_tank(nx, wavelengths, kh=1.0, steepness=0.05, h=1.0, N=32, courant=0.5):
    """x-z tank with walls, Nx points over `wavelengths` wavelengths, stream-function wave of height
    steepness * L and depth h (kh given); returns (wave, x, E, P, dt)."""
    from .problems import coords
    k = kh / h
    lam = 2 * math.pi / k
    wave = StreamFunctionWave(k, h, steepness * lam, N=N)
    lx = wavelengths * lam
    x, _ = coords(nx, 1, lx, 1.0)
    E, P = wave.surface(x)
    dx = lx / (nx - 1)
    dt = courant * dx / wave.c
    return wave, x, E[None, :], P[None, :], dt
