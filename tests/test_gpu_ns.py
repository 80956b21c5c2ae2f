"""GPU Navier-Stokes solver (examples/navier-stokes-MPI) against the oracle,
and its tangent / adjoint / checkpointed gradient."""

import numpy as np
import pytest

from oracle import navier_stokes as NS

pytestmark = pytest.mark.gpu


def _solver(n, px, py, form, nu=1.0 / 1600.0, dt=1e-3):
    import dfft_ns
    return dfft_ns.NavierStokes(n, px, py, lx=2.0, nu=nu, dt=dt, form=form)


def _scatter(s, arrs, which, dev):
    """Global (3, ...) numpy field set -> per-rank device blocks."""
    import torch
    out = []
    for r in range(s.nranks):
        e, st = s.local_extents(r, which)
        sl = tuple(slice(a, a + b) for a, b in zip(st, e))
        blk = np.stack([a[sl] for a in arrs])
        dt = torch.float64 if which == 0 else torch.complex128
        out.append(torch.from_numpy(np.ascontiguousarray(blk)).to(dev, dt))
    return out


def _gather(s, blocks, which):
    n = s.n
    shape = (n, n, n) if which == 0 else (n, n, n // 2 + 1)
    out = np.zeros((3,) + shape, dtype=np.float64 if which == 0 else np.complex128)
    for r, b in enumerate(blocks):
        e, st = s.local_extents(r, which)
        sl = (slice(None),) + tuple(slice(a, a + b) for a, b in zip(st, e))
        out[sl] = b.cpu().numpy()
    return out


def _hdot(n, a, b):
    """<a, b>_h = sum h Re(conj a b), h = 1 on the kz = 0, n/2 planes, 2 elsewhere."""
    h = np.full(a.shape[-1], 2.0)
    h[0] = 1.0
    if n % 2 == 0:
        h[n // 2] = 1.0
    return float((h * (np.conj(a) * b).real).sum())


def _rel(a, b):
    return float(np.linalg.norm((a - b).ravel()) / np.linalg.norm(b.ravel()))


@pytest.mark.parametrize("n,px,py,form", [
    (16, 1, 1, 0), (16, 2, 2, 0), (16, 1, 3, 0), (32, 2, 1, 0),
    (16, 1, 1, 1), (24, 2, 2, 1), (16, 3, 2, 1),
])
def test_taylor_green_steps_match_oracle(dev, n, px, py, form):
    nu, dt, nsteps = 0.01, 1e-2, 3
    s = _solver(n, px, py, form, nu, dt)
    s.init_taylor_green(1.0)
    g = NS.Grid(n)
    y = list(NS.initial_spectrum(g, *NS.taylor_green(g)))
    assert abs(s.energy() - np.pi ** 3) < 1e-12 * np.pi ** 3
    assert _rel(_gather(s, s.get_state(), 1), np.stack(y)) < 1e-12
    s.step(nsteps)
    for _ in range(nsteps):
        y = NS.rk4_step(g, nu, dt, y)
    got = _gather(s, s.get_state(), 1)
    assert _rel(got, np.stack(y)) < 1e-12
    assert abs(s.energy() - NS.kinetic_energy(g, *y)) < 1e-12 * np.pi ** 3


@pytest.mark.parametrize("form", [0, 1])
def test_random_field_steps_match_oracle(dev, form):
    n, nu, dt = 16, 0.02, 5e-3
    rng = np.random.default_rng(11)
    u = [rng.standard_normal((n, n, n)) for _ in range(3)]
    s = _solver(n, 2, 2, form, nu, dt)
    s.set_physical(_scatter(s, u, 0, dev))
    g = NS.Grid(n)
    y = list(NS.initial_spectrum(g, *u))
    s.step(2)
    for _ in range(2):
        y = NS.rk4_step(g, nu, dt, y)
    assert _rel(_gather(s, s.get_state(), 1), np.stack(y)) < 1e-12


@pytest.mark.parametrize("n,px,py", [(16, 1, 1), (16, 2, 2), (32, 2, 1), (16, 1, 3), (64, 2, 2)])
def test_products_formed_in_the_forward_pass_match_stored_products(dev, n, px, py):
    """The divergence form's 6 products formed at the loads of the forward
    R2C pass (no product fields in HBM) against the stored-products path."""
    nu, dt = 0.01, 1e-2
    rng = np.random.default_rng(n + 7 * px + 13 * py)
    u = [rng.standard_normal((n, n, n)) for _ in range(3)]
    a, b = _solver(n, px, py, 0, nu, dt), _solver(n, px, py, 0, nu, dt)
    b.set_option("fuse_products", 0)
    assert a.get_option("products_fused") == 1
    assert b.get_option("products_fused") == 0 and b.get_option("fuse_products") == 0
    for s in (a, b):
        s.set_physical(_scatter(s, u, 0, dev))
    ya = a.get_state()
    assert _rel(_gather(a, a.rhs(ya), 1), _gather(b, b.rhs(ya), 1)) < 1e-13
    a.step(2, use_graph=True)
    b.step(2, use_graph=True)
    assert _rel(_gather(a, a.get_state(), 1), _gather(b, b.get_state(), 1)) < 1e-13
    # switching back re-captures the step with the fused launches
    b.set_option("fuse_products", 1)
    assert b.get_option("products_fused") == 1
    a.step(1, use_graph=True)
    b.step(1, use_graph=True)
    assert _rel(_gather(a, a.get_state(), 1), _gather(b, b.get_state(), 1)) < 1e-13


def test_products_option_where_the_pass_cannot_form_them(dev):
    import dfft
    s = _solver(24, 2, 2, 0)         # 24-point lines: no products kernel
    assert s.get_option("fuse_products") == 1 and s.get_option("products_fused") == 0
    t = _solver(16, 1, 1, 1)         # advective form: no products at all
    assert t.get_option("products_fused") == 0
    with pytest.raises(dfft.DFFTError):
        s.set_option("fuse_products", 2)
    with pytest.raises(dfft.DFFTError):
        s.get_option("no_such_option")


def test_graph_replay_matches_eager_bitwise(dev):
    a = _solver(32, 2, 2, 0)
    b = _solver(32, 2, 2, 0)
    a.init_taylor_green(1.0)
    b.init_taylor_green(1.0)
    a.step(5, use_graph=False)
    b.step(2, use_graph=True)   # eager first step, then a captured step replayed
    b.step(3, use_graph=True)   # replays only
    ya, yb = a.get_state(), b.get_state()
    for ta, tb in zip(ya, yb):
        assert bool((ta == tb).all())
    assert a.energy() == b.energy()


def test_autotune_keeps_state_and_results(dev):
    a = _solver(32, 2, 2, 0)
    b = _solver(32, 2, 2, 0)
    for s in (a, b):
        s.init_taylor_green(1.0)
    b.autotune()
    ya0, yb0 = a.get_state(), b.get_state()
    for ta, tb in zip(ya0, yb0):
        assert bool((ta == tb).all())
    a.step(3, use_graph=True)
    b.step(3, use_graph=True)
    ga, gb = _gather(a, a.get_state(), 1), _gather(b, b.get_state(), 1)
    assert _rel(gb, ga) < 1e-13


def test_rhs_of_taylor_green_matches_closed_form(dev):
    n, nu = 32, 0.01
    s = _solver(n, 2, 2, 0, nu)
    s.init_taylor_green(1.0)
    d = _gather(s, s.rhs(s.get_state()), 1)
    g = NS.Grid(n)
    X, Y, Z = np.meshgrid(g.x, g.x, g.x, indexing="ij")
    u, v, _ = NS.taylor_green(g)
    want = [-np.sin(2 * X) * np.cos(2 * Z) / 8.0 - 3 * nu * u,
            -np.sin(2 * Y) * np.cos(2 * Z) / 8.0 - 3 * nu * v,
            (np.cos(2 * X) + np.cos(2 * Y)) * np.sin(2 * Z) / 8.0]
    assert _rel(d, np.stack([NS.forward(c) for c in want])) < 1e-12


def _random_states(s, dev, seeds):
    out = []
    for sd in seeds:
        rng = np.random.default_rng(sd)
        u = [rng.standard_normal((s.n,) * 3) for _ in range(3)]
        s.set_physical(_scatter(s, u, 0, dev))
        out.append(s.get_state())
    return out


def test_tangent_is_exact_central_difference(dev):
    """R is quadratic, so (R(y + e d) - R(y - e d)) / 2e equals R'(y) d up to rounding."""
    import torch
    s = _solver(16, 2, 2, 0, nu=0.01)
    y, d = _random_states(s, dev, [1, 2])
    eps = 1e-3
    rp = s.rhs([a + eps * b for a, b in zip(y, d)])
    rm = s.rhs([a - eps * b for a, b in zip(y, d)])
    fd = [(p - m) / (2 * eps) for p, m in zip(rp, rm)]
    tl = s.rhs_tangent(y, d)
    num = sum(float(torch.linalg.vector_norm(a - b)) ** 2 for a, b in zip(tl, fd))
    den = sum(float(torch.linalg.vector_norm(b)) ** 2 for b in fd)
    assert (num / den) ** 0.5 < 1e-9


@pytest.mark.parametrize("n,px,py", [(16, 1, 1), (16, 2, 2), (24, 2, 3)])
def test_adjoint_dot_product(dev, n, px, py):
    """<R'(y) d, l>_h == <d, R'(y)^* l>_h."""
    s = _solver(n, px, py, 0, nu=0.01)
    y, d, lam = _random_states(s, dev, [3, 4, 5])
    lhs = _hdot(n, _gather(s, s.rhs_tangent(y, d), 1), _gather(s, lam, 1))
    rhs = _hdot(n, _gather(s, d, 1), _gather(s, s.rhs_adjoint(y, lam), 1))
    assert abs(lhs - rhs) < 1e-12 * max(abs(lhs), abs(rhs))


def _tg_noise(n, seed, sigma=1e-2):
    g = NS.Grid(n)
    rng = np.random.default_rng(seed)
    return g, [a + rng.normal(0.0, sigma, a.shape) for a in NS.taylor_green(g)]


@pytest.mark.parametrize("nsteps,ck,px,py", [(3, 1, 1, 1), (3, 2, 2, 2), (4, 5, 1, 2)])
def test_gradient_matches_oracle_directional_derivative(dev, nsteps, ck, px, py):
    n, nu, dt = 16, 0.01, 1e-2
    g, u0 = _tg_noise(n, 5)
    rng = np.random.default_rng(6)
    delta = [rng.standard_normal((n, n, n)) for _ in range(3)]
    s = _solver(n, px, py, 0, nu, dt)
    j, grad = s.gradient(nsteps, _scatter(s, u0, 0, dev), checkpoint_every=ck)
    j_ref, dj_ref = NS.final_energy_directional(g, nu, dt, nsteps, u0, delta)
    assert abs(j - j_ref) < 1e-12 * abs(j_ref)
    dj = float((_gather(s, grad, 0) * np.stack(delta)).sum())
    assert abs(dj - dj_ref) < 1e-10 * abs(dj_ref)


def test_gradient_128_cubed_20_steps_against_central_differences(dev):
    """Taylor-Green at 128^3, Re = 1600, dt = 1e-3, N(0, 1e-2) noise (seed 5),
    20 RK4 steps, checkpoint every 5: dJ/du0 . delta vs central differences, eps = 1e-6."""
    import torch
    n, nsteps = 128, 20
    _, u0 = _tg_noise(n, 5)
    s = _solver(n, 1, 1, 0)
    u0d = _scatter(s, u0, 0, dev)
    j, grad = s.gradient(nsteps, u0d, checkpoint_every=5)
    rng = np.random.default_rng(7)
    directions = {"u0": u0d, "noise": _scatter(s, [rng.standard_normal((n, n, n)) for _ in range(3)], 0, dev)}
    eps = 1e-6
    for name, dd in directions.items():
        jp = s.final_energy(nsteps, [a + eps * b for a, b in zip(u0d, dd)])
        jm = s.final_energy(nsteps, [a - eps * b for a, b in zip(u0d, dd)])
        fd = (jp - jm) / (2 * eps)
        ad = sum(float((g_ * d_).sum()) for g_, d_ in zip(grad, dd))
        print(f"direction {name}: adjoint {ad:.15e} fd {fd:.15e} rel {abs(ad - fd) / abs(fd):.2e}")
        assert abs(ad - fd) < 1e-6 * abs(fd)
    assert abs(j - s.final_energy(nsteps, u0d)) < 1e-13 * abs(j)
    del torch


def test_autograd_harness(dev):
    import torch
    import dfft_ns
    n, nsteps = 16, 3
    _, u0 = _tg_noise(n, 5)
    s = _solver(n, 2, 1, 0, 0.01, 1e-2)
    u0d = [t.requires_grad_(True) for t in _scatter(s, u0, 0, dev)]
    e = dfft_ns.final_energy(s, nsteps, u0d, checkpoint_every=2)
    gr = torch.autograd.grad(3.0 * e, u0d)
    j, grad = s.gradient(nsteps, [t.detach() for t in u0d], checkpoint_every=2)
    assert abs(float(e) - j) == 0.0
    for a, b in zip(gr, grad):
        assert torch.allclose(a, 3.0 * b, rtol=0, atol=0)


def test_diagnostics_of_taylor_green(dev, tmp_path):
    """diag_fields / diag_energy (main.cpp:812-910): per-rank u, v, w and vorticity files
    named as the reference names them, against the Taylor-Green closed forms; the
    spectrum holds all energy in the |k| = sqrt(3) bin."""
    n, px, py = 16, 2, 2
    s = _solver(n, px, py, 0)
    s.init_taylor_green(1.0)
    e = s.diag(str(tmp_path / "data"), 7)
    assert abs(e - np.pi ** 3) < 1e-12 * np.pi ** 3
    g = NS.Grid(n)
    X, Y, Z = np.meshgrid(g.x, g.x, g.x, indexing="ij")
    want = {"u": np.sin(X) * np.cos(Y) * np.cos(Z), "v": -np.cos(X) * np.sin(Y) * np.cos(Z), "w": 0 * X,
            "vor_x": -np.cos(X) * np.sin(Y) * np.sin(Z), "vor_y": -np.sin(X) * np.cos(Y) * np.sin(Z),
            "vor_z": 2 * np.sin(X) * np.sin(Y) * np.cos(Z)}
    d = tmp_path / f"data_px{px}_py{py}"
    for r in range(px * py):
        ext, st = s.local_extents(r, 0)
        rx, ry = r // py, r % py
        sl = tuple(slice(a, a + b) for a, b in zip(st, ext))
        for label, ref in want.items():
            f = d / f"{label}_rx{rx:03d}_ry{ry:03d}_{7:010d}.dat"
            got = np.fromfile(f, dtype=np.float64).reshape(ext)
            assert np.abs(got - ref[sl]).max() < 1e-13, (label, r)
    e_spec = np.fromfile(d / f"E_spec_rx000_ry000_{7:010d}.dat", dtype=np.float64)
    k_vals = np.fromfile(d / f"k_vals_rx000_ry000_{7:010d}.dat", dtype=np.float64)
    assert e_spec.shape == (50,) and k_vals.shape == (50,)
    k_max = np.sqrt(3.0) * (n // 2)
    bins = np.linspace(0.0, k_max, 51)
    assert np.allclose(k_vals, 0.5 * (bins[:-1] + bins[1:]), rtol=0, atol=1e-14)
    hit = int(np.sqrt(3.0) // (k_max / 50))
    coef = (2 * np.pi) ** 3 / n ** 6
    assert abs(e_spec[hit] * coef - np.pi ** 3) < 1e-12 * np.pi ** 3
    # other bins hold only the squares of transform round-off
    assert np.abs(np.delete(e_spec, hit)).max() < 1e-20 * e_spec[hit]


def test_energy_spectrum_sums_to_energy(dev):
    n = 24
    s = _solver(n, 2, 3, 0)
    y, = _random_states(s, dev, [9])
    s.set_state(y)
    e_spec, k_vals = s.energy_spectrum(50)
    coef = (2 * np.pi) ** 3 / n ** 6
    assert abs(sum(e_spec) * coef - s.energy()) < 1e-12 * s.energy()
    # velocity and vorticity agree with the oracle's backward transforms of the state
    ys = _gather(s, s.get_state(), 1)
    g = NS.Grid(n)
    u_ref = np.stack([NS.backward(c, n) for c in ys])
    assert _rel(_gather(s, s.velocity(), 0), u_ref) < 1e-12
    w_hat = [1j * (g.KY * ys[2] - g.KZ * ys[1]), 1j * (g.KZ * ys[0] - g.KX * ys[2]),
             1j * (g.KX * ys[1] - g.KY * ys[0])]
    assert _rel(_gather(s, s.vorticity(), 0), np.stack([NS.backward(c, n) for c in w_hat])) < 1e-12


# ---------------------------------------------------------------------------
# Taylor-Green 256^3 float64 on a 2 x 2 grid of virtual ranks (Re = 1600,
# dt = 1e-3, divergence form, autotuned plans, hipGraph replay): the
# production-size run of examples/navier-stokes-MPI/main.cpp:349-353, 572-666.
#
# The oracle cannot step 256^3 in test time, but the Taylor-Green state is
# band-limited: the four stages of one RK4 step double the largest wavenumber
# component each (1 -> 2 -> 4 -> 8 -> 16), and every product stays below
# n/2 for n = 64 (no aliasing onto kept modes) and inside both dealias
# cutoffs (21 at n = 64, 85 at n = 256).  One step at n = 64 therefore holds
# the same coefficients as one step at n = 256, up to the factor (256/64)^3
# of the unscaled forward transform, and every other coefficient is zero.
# ---------------------------------------------------------------------------

def _embed(small, n):
    """Spectral field set of a small grid, scaled to the unscaled transform of
    grid n and placed at the same wavenumbers of the (n, n, n/2+1) half spectrum."""
    m = small.shape[1]
    big = np.zeros(small.shape[:1] + (n, n, n // 2 + 1), dtype=np.complex128)
    k = np.arange(-(m // 2) + 1, m // 2)  # the small grid's Nyquist mode stays out
    kz = np.arange(0, m // 2)
    ix, iz = k % m, kz
    jx, jz = k % n, kz
    scale = (n / m) ** 3
    big[np.ix_(np.arange(small.shape[0]), jx, jx, jz)] = small[np.ix_(np.arange(small.shape[0]), ix, ix, iz)] * scale
    return big


def _max_divergence(n, y):
    g = NS.Grid(n)
    div = g.KX * y[0] + g.KY * y[1] + g.KZ * y[2]
    return float(np.abs(div).max() / np.abs(y).max())


def test_taylor_green_256_cubed_2x2_production_run(dev):
    import dfft_ns
    n, nu, dt = 256, 1.0 / 1600.0, 1e-3
    s = _solver(n, 2, 2, dfft_ns.DIVERGENCE, nu, dt)
    s.autotune()
    assert s.get_option("products_fused") == 1
    s.init_taylor_green(1.0)
    # E(0) = pi^3 and the initial state
    assert abs(s.energy() - np.pi ** 3) < 1e-12 * np.pi ** 3
    g16 = NS.Grid(16)
    y16 = np.stack(NS.initial_spectrum(g16, *NS.taylor_green(g16)))
    y0 = _gather(s, s.get_state(), 1)
    assert _rel(y0, _embed(y16, n)) < 1e-12
    # closed-form right-hand side at t = 0 (every mode has |k_i| <= 2)
    d = _gather(s, s.rhs(s.get_state()), 1)
    X, Y, Z = np.meshgrid(g16.x, g16.x, g16.x, indexing="ij")
    u, v, _ = NS.taylor_green(g16)
    want = [-np.sin(2 * X) * np.cos(2 * Z) / 8.0 - 3 * nu * u,
            -np.sin(2 * Y) * np.cos(2 * Z) / 8.0 - 3 * nu * v,
            (np.cos(2 * X) + np.cos(2 * Y)) * np.sin(2 * Z) / 8.0]
    assert _rel(d, _embed(np.stack([NS.forward(c) for c in want]), n)) < 1e-12
    del d
    # one eager RK4 step against the oracle's step at n = 64
    g64 = NS.Grid(64)
    y64 = NS.rk4_step(g64, nu, dt, list(NS.initial_spectrum(g64, *NS.taylor_green(g64))))
    s.step(1, use_graph=False)
    y1 = _gather(s, s.get_state(), 1)
    assert _rel(y1, _embed(np.stack(y64), n)) < 1e-12
    assert abs(s.energy() - NS.kinetic_energy(g64, *y64)) < 1e-12 * np.pi ** 3
    # two graph-replayed steps equal two eager steps bitwise and stay solenoidal
    s.init_taylor_green(1.0)
    s.step(2, use_graph=False)
    eager = s.get_state()
    e_eager = s.energy()
    s.init_taylor_green(1.0)
    s.step(1, use_graph=True)  # eager first step, then capture
    s.step(1, use_graph=True)  # replay
    replay = s.get_state()
    for a, b in zip(eager, replay):
        assert bool((a == b).all())
    assert s.energy() == e_eager
    y2 = _gather(s, replay, 1)
    assert _max_divergence(n, y2) < 1e-12
    # viscous decay: dE/dt(0) = -6 nu E(0) for Taylor-Green; over two steps the
    # O(dt^2) remainder is far below the tolerance
    rate = (e_eager - np.pi ** 3) / (2 * dt)
    assert abs(rate + 6 * nu * np.pi ** 3) < 1e-3 * 6 * nu * np.pi ** 3
    s.destroy()


def test_taylor_green_256_cubed_2x2_inviscid_conserves_energy(dev):
    """nu = 0: the dealiased Galerkin system conserves kinetic energy; RK4 adds
    an O(dt^5) error per step (far below 1e-12 at dt = 1e-3)."""
    import dfft_ns
    n = 256
    s = _solver(n, 2, 2, dfft_ns.DIVERGENCE, 0.0, 1e-3)
    s.init_taylor_green(1.0)
    s.step(3, use_graph=True)
    assert abs(s.energy() - np.pi ** 3) < 1e-12 * np.pi ** 3
    assert _max_divergence(n, _gather(s, s.get_state(), 1)) < 1e-12
    s.destroy()
