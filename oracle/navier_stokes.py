"""Oracle (test infrastructure only): CPU restatement of the reference's
pseudo-spectral Navier-Stokes example on one global grid (the per-rank
distribution does not change the arithmetic).

Follows ``examples/navier-stokes-MPI/main.cpp``:
* Grid (:31-206): domain [0, lx*pi)^3, wavenumbers fftfreq(n, lx/(2n)),
  rfftfreq along z, ksq, inv_ksq (0 at k=0), 2/3 dealias mask with cutoff
  int((n/2)*2/3) on the integer indices;
* Variables (:270-430): Taylor-Green initial condition u = v0 sin x cos y cos z,
  v = -v0 cos x sin y cos z, w = 0, forward transform, dealias, projection;
* rhs (:572-666): dealias, 9 spectral derivatives, 12 backward transforms,
  advective term u.grad(u) in physical space, 3 forward transforms, dealias,
  combine (-adv - nu k^2 u), projection;
* RK4th (examples/utils/math_utils.hpp:47-141): the four stage updates;
* diag_energy (:855-910): E = sum 0.5 f (|u|^2+|v|^2+|w|^2) (lx pi)^3 / n^6
  with the Hermitian weight f = 1 on the kz = 0 plane and 2 elsewhere,
  evaluated here with the global kz index.

FFTs are ``oracle.fft`` (FP64); backward transforms use the plan's default
backward normalisation (1/n^3).
"""

from __future__ import annotations

import numpy as np

from . import fft as F


def fftfreq(n, d):
    """numpy-compatible fftfreq (kokkos-fft KokkosFFT::fftfreq)."""
    return np.fft.fftfreq(n, d)


class Grid:
    def __init__(self, n, lx=2.0):
        self.n = n
        self.lx = lx
        self.x = np.arange(n) * (lx * np.pi / n)
        d = lx / (2.0 * n)
        self.kx = fftfreq(n, d)
        self.ky = fftfreq(n, d)
        self.kz = np.fft.rfftfreq(n, d)
        KX, KY, KZ = np.meshgrid(self.kx, self.ky, self.kz, indexing="ij")
        self.KX, self.KY, self.KZ = KX, KY, KZ
        self.ksq = KX ** 2 + KY ** 2 + KZ ** 2
        self.inv_ksq = np.zeros_like(self.ksq)
        nz = self.ksq != 0
        self.inv_ksq[nz] = 1.0 / self.ksq[nz]
        self.inv_ksq[0, 0, 0] = 0.0
        kmax = n // 2
        cutoff = (kmax * 2) // 3  # integer arithmetic of main.cpp:166-169
        ix = np.abs(np.fft.fftfreq(n, 1.0) * n)
        iz = np.abs(np.fft.rfftfreq(n, 1.0) * n)
        IX, IY, IZ = np.meshgrid(ix, ix, iz, indexing="ij")
        self.mask = (~((IX > cutoff) | (IY > cutoff) | (IZ > cutoff))).astype(np.float64)
        self.cutoff = cutoff


def forward(u):
    return F.fftn(u, [0, 1, 2], real_input=True)


def backward(uk, n):
    return F.ifftn(uk, [0, 1, 2], n_last=n, real_output=True)


def projection(g, uk, vk, wk):
    div = (g.KX * uk + g.KY * vk + g.KZ * wk) * g.inv_ksq
    return uk - g.KX * div, vk - g.KY * div, wk - g.KZ * div


def taylor_green(g, v0=1.0):
    X, Y, Z = np.meshgrid(g.x, g.x, g.x, indexing="ij")
    u = v0 * np.sin(X) * np.cos(Y) * np.cos(Z)
    v = -v0 * np.cos(X) * np.sin(Y) * np.cos(Z)
    w = np.zeros_like(u)
    return u, v, w


def initial_spectrum(g, u, v, w):
    uk, vk, wk = forward(u) * g.mask, forward(v) * g.mask, forward(w) * g.mask
    return projection(g, uk, vk, wk)


def rhs(g, nu, uk, vk, wk):
    """Returns (dukdt, dvkdt, dwkdt, uk_dealiased, vk_dealiased, wk_dealiased)."""
    n = g.n
    uk, vk, wk = uk * g.mask, vk * g.mask, wk * g.mask
    comps = (uk, vk, wk)
    real = [backward(c, n) for c in comps]
    grads = [[backward(1j * K * c, n) for K in (g.KX, g.KY, g.KZ)] for c in comps]
    u, v, w = real
    adv = [u * gr[0] + v * gr[1] + w * gr[2] for gr in grads]
    advk = [forward(a) * g.mask for a in adv]
    visc = -nu * g.ksq
    out = [-a + visc * c for a, c in zip(advk, comps)]
    out = projection(g, *out)
    return out[0], out[1], out[2], uk, vk, wk


def rk4_step(g, nu, dt, state):
    """One RK4th step (math_utils.hpp:89-141) of the three spectral fields.
    The reference dealiases the state in place inside rhs (main.cpp:576-578),
    so the stage value it advances from is the dealiased one."""
    y = list(state)
    y_copy = None
    ks = []
    for step in range(4):
        d0, d1, d2, a0, a1, a2 = rhs(g, nu, *y)
        y = [a0, a1, a2]  # in-place dealias of the stage value
        dy = [d0, d1, d2]
        if step == 0:
            y_copy = list(y)
            k1 = [d * dt for d in dy]
            ks.append(k1)
            y = [yc + k / 2.0 for yc, k in zip(y_copy, k1)]
        elif step == 1:
            k2 = [d * dt for d in dy]
            ks.append(k2)
            y = [yc + k / 2.0 for yc, k in zip(y_copy, k2)]
        elif step == 2:
            k3 = [d * dt for d in dy]
            ks.append(k3)
            y = [yc + k for yc, k in zip(y_copy, k3)]
        else:
            y = [yc + (k1_ + 2.0 * k2_ + 2.0 * k3_ + d * dt) / 6.0
                 for yc, k1_, k2_, k3_, d in zip(y_copy, ks[0], ks[1], ks[2], dy)]
    return y


def kinetic_energy(g, uk, vk, wk):
    coef = (g.lx * np.pi) ** 3 / float(g.n) ** 6
    f = np.full(g.ksq.shape, 2.0)
    f[:, :, 0] = 1.0
    e = 0.5 * f * (np.abs(uk) ** 2 + np.abs(vk) ** 2 + np.abs(wk) ** 2)
    return float(e.sum() * coef)


# ---------------------------------------------------------------------------
# Forward-mode (tangent-linear) restatement, for checking gradients.  No
# counterpart in the reference; pinned against central differences of
# rk4_step / kinetic_energy above.
# ---------------------------------------------------------------------------

def rhs_tangent(g, nu, y, dy):
    """d/de rhs(y + e dy) at e = 0 for the advective form (main.cpp:572-666).
    Returns (tangent of the three rhs fields, tangent of the dealiased state)."""
    n = g.n
    ym = [c * g.mask for c in y]
    dm = [c * g.mask for c in dy]
    u = [backward(c, n) for c in ym]
    du = [backward(c, n) for c in dm]
    grads = [[backward(1j * K * c, n) for K in (g.KX, g.KY, g.KZ)] for c in ym]
    dgrads = [[backward(1j * K * c, n) for K in (g.KX, g.KY, g.KZ)] for c in dm]
    dadv = [sum(du[j] * grads[i][j] + u[j] * dgrads[i][j] for j in range(3)) for i in range(3)]
    dadvk = [forward(a) * g.mask for a in dadv]
    visc = -nu * g.ksq
    out = [-a + visc * c for a, c in zip(dadvk, dm)]
    out = projection(g, *out)
    return list(out), dm


def rk4_step_tangent(g, nu, dt, state, dstate):
    """One RK4 step and its tangent: returns (next state, next tangent)."""
    y, dy = list(state), list(dstate)
    ks, dks = [], []
    y_copy = dy_copy = None
    for step in range(4):
        d0, d1, d2, a0, a1, a2 = rhs(g, nu, *y)
        dd, dm = rhs_tangent(g, nu, y, dy)
        y, dy = [a0, a1, a2], dm
        k, dk = [v * dt for v in (d0, d1, d2)], [v * dt for v in dd]
        if step == 0:
            y_copy, dy_copy = list(y), list(dy)
        if step < 3:
            ks.append(k)
            dks.append(dk)
            h = 0.5 if step < 2 else 1.0
            y = [a + h * b for a, b in zip(y_copy, k)]
            dy = [a + h * b for a, b in zip(dy_copy, dk)]
        else:
            y = [a + (k1 + 2.0 * k2 + 2.0 * k3 + k4) / 6.0
                 for a, k1, k2, k3, k4 in zip(y_copy, ks[0], ks[1], ks[2], k)]
            dy = [a + (k1 + 2.0 * k2 + 2.0 * k3 + k4) / 6.0
                  for a, k1, k2, k3, k4 in zip(dy_copy, dks[0], dks[1], dks[2], dk)]
    return y, dy


def kinetic_energy_tangent(g, y, dy):
    coef = (g.lx * np.pi) ** 3 / float(g.n) ** 6
    f = np.full(g.ksq.shape, 2.0)
    f[:, :, 0] = 1.0
    return float(coef * sum((f * (np.conj(a) * b).real).sum() for a, b in zip(y, dy)))


def final_energy_directional(g, nu, dt, nsteps, u0, delta):
    """(J, dJ/du0 . delta) for J = kinetic energy after nsteps RK4 steps from the
    physical velocity u0 (three arrays), the state starting as D P FFT(u0)."""
    y = list(initial_spectrum(g, *u0))
    dy = list(initial_spectrum(g, *delta))
    for _ in range(nsteps):
        y, dy = rk4_step_tangent(g, nu, dt, y, dy)
    return kinetic_energy(g, *y), kinetic_energy_tangent(g, y, dy)
