"""Pins the Navier-Stokes oracle (examples/navier-stokes-MPI/main.cpp) with
closed forms and invariants, independent of the oracle's own formulas."""

import numpy as np
import pytest

from oracle import navier_stokes as NS


def _tg_rhs_closed_form(g, nu):
    """d/dt of the Taylor-Green field at t = 0: -P(u.grad u) + nu lap u.
    u.grad u = (sin2x cos^2 z, sin2y cos^2 z, 0)/2, pressure (cos2x + cos2y)(cos2z + 2)/16."""
    X, Y, Z = np.meshgrid(g.x, g.x, g.x, indexing="ij")
    u, v, w = NS.taylor_green(g)
    du = -np.sin(2 * X) * np.cos(2 * Z) / 8.0 - 3.0 * nu * u
    dv = -np.sin(2 * Y) * np.cos(2 * Z) / 8.0 - 3.0 * nu * v
    dw = (np.cos(2 * X) + np.cos(2 * Y)) * np.sin(2 * Z) / 8.0
    return du, dv, dw


@pytest.mark.parametrize("n", [8, 16])
def test_taylor_green_energy_is_pi_cubed(n):
    g = NS.Grid(n)
    y = NS.initial_spectrum(g, *NS.taylor_green(g))
    # E = 1/2 * (2 pi)^3 * <u^2 + v^2> = 1/2 * 8 pi^3 * (1/8 + 1/8)
    assert abs(NS.kinetic_energy(g, *y) - np.pi ** 3) < 1e-12 * np.pi ** 3


@pytest.mark.parametrize("n", [8, 16])
def test_rhs_of_taylor_green_matches_closed_form(n):
    g = NS.Grid(n)
    nu = 0.01
    y = NS.initial_spectrum(g, *NS.taylor_green(g))
    d = NS.rhs(g, nu, *y)[:3]
    want = [NS.forward(c) for c in _tg_rhs_closed_form(g, nu)]
    for a, b in zip(d, want):
        assert np.abs(a - b).max() < 1e-12 * n ** 3


def _random_solenoidal(g, seed, kmax=3):
    rng = np.random.default_rng(seed)
    u = [rng.standard_normal((g.n,) * 3) for _ in range(3)]
    y = NS.initial_spectrum(g, *u)
    ix = np.abs(np.fft.fftfreq(g.n, 1.0) * g.n)
    iz = np.abs(np.fft.rfftfreq(g.n, 1.0) * g.n)
    IX, IY, IZ = np.meshgrid(ix, ix, iz, indexing="ij")
    low = (IX <= kmax) & (IY <= kmax) & (IZ <= kmax)
    return [c * low for c in y]


def test_inviscid_steps_conserve_energy_and_divergence():
    """The dealiased Galerkin system conserves energy at nu = 0; RK4 adds O(dt^5) per step."""
    g = NS.Grid(16)
    y = _random_solenoidal(g, 3)
    e0 = NS.kinetic_energy(g, *y)
    for _ in range(4):
        y = NS.rk4_step(g, 0.0, 1e-3, y)
    assert abs(NS.kinetic_energy(g, *y) - e0) < 1e-10 * e0
    div = g.KX * y[0] + g.KY * y[1] + g.KZ * y[2]
    assert np.abs(div).max() < 1e-10 * np.abs(y[0]).max()


def test_viscous_decay_rate_of_taylor_green():
    """Energy of the TG vortex starts decaying at dE/dt = -2 nu |k|^2 E = -6 nu E."""
    g = NS.Grid(16)
    nu, dt = 0.05, 1e-4
    y = NS.initial_spectrum(g, *NS.taylor_green(g))
    e0 = NS.kinetic_energy(g, *y)
    e1 = NS.kinetic_energy(g, *NS.rk4_step(g, nu, dt, y))
    rate = (e1 - e0) / dt
    assert abs(rate + 6 * nu * e0) < 1e-3 * 6 * nu * e0


def test_tangent_matches_central_differences():
    g = NS.Grid(8)
    nu, dt, nsteps = 0.05, 0.05, 3
    rng = np.random.default_rng(1)
    u0 = [a + rng.normal(0, 0.1, a.shape) for a in NS.taylor_green(g)]
    de = [rng.normal(0, 1, a.shape) for a in u0]
    j, dj = NS.final_energy_directional(g, nu, dt, nsteps, u0, de)

    def jf(e):
        y = NS.initial_spectrum(g, *[a + e * b for a, b in zip(u0, de)])
        for _ in range(nsteps):
            y = NS.rk4_step(g, nu, dt, y)
        return NS.kinetic_energy(g, *y)

    assert abs(jf(0.0) - j) < 1e-14 * abs(j)
    fd = (jf(1e-4) - jf(-1e-4)) / 2e-4
    assert abs(fd - dj) < 1e-8 * abs(dj)
