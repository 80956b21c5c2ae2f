"""Adjoint of the distributed plan (dfft_execute_adjoint) and the
torch.autograd harness over it: dot-product identities, the FP64 gradient of
torch.fft on the CPU, and central differences."""

import numpy as np
import pytest

from helpers import out_extents, seeded_input
from oracle import planner as PL
from oracle import plan_exec as PE

pytestmark = pytest.mark.gpu


def _plan(dev, gin, axes, tin, tout, real, fp64, norm, exec_mode=0):
    import torch
    import dfft
    gout = out_extents(gin, axes, real)
    rdt = torch.float64 if fp64 else torch.float32
    cdt = torch.complex128 if fp64 else torch.complex64
    ins, outs = [], []
    for r in range(int(np.prod(tin))):
        e, _ = PL.compute_local_extents_and_starts(gin, list(tin), r, PL.LAYOUT_RIGHT)
        ins.append(torch.zeros(e, dtype=rdt if real else cdt, device=dev))
    for r in range(int(np.prod(tout))):
        e, _ = PL.compute_local_extents_and_starts(gout, list(tout), r, PL.LAYOUT_RIGHT)
        outs.append(torch.zeros(e, dtype=cdt, device=dev))
    plan = dfft.Plan(ins, outs, axes, tin, tout, norm=norm, exec_mode=exec_mode,
                     global_in_extents=gin, global_out_extents=gout)
    return plan, gout, rdt if real else cdt, cdt


def _blocks(dev, arr, gext, topo, dt):
    import torch
    return [torch.from_numpy(np.ascontiguousarray(a)).to(dev, dt) for a in PE.scatter(arr, gext, list(topo))]


def _rdot(a, b):
    """Re sum conj(a) b over all blocks, in FP64."""
    s = 0.0
    for x, y in zip(a, b):
        x = x.cpu().numpy().astype(np.complex128)
        y = y.cpu().numpy().astype(np.complex128)
        s += float((np.conj(x) * y).real.sum())
    return s


def _norm(a):
    return float(np.sqrt(sum(float(np.linalg.norm(x.cpu().numpy().astype(np.complex128).ravel())) ** 2
                             for x in a)))


SHAPES = [
    ([16, 12, 10], [0, 1, 2], [2, 1, 1], [1, 2, 1]),
    ([15, 12, 9], [0, 1, 2], [1, 3, 1], [3, 1, 1]),
    ([8, 7, 6], [2, 0, 1], [1, 2, 2], [2, 2, 1]),
    ([8, 7, 6], [1, 2, 0], [2, 1, 3], [2, 3, 1]),
    ([12, 10, 8], [0, 2], [1, 2, 1], [2, 1, 1]),
    ([12, 10, 9], [1], [3, 1, 1], [1, 1, 3]),
    ([64, 32, 64], [0, 1, 2], [4, 1, 1], [1, 4, 1]),
]


@pytest.mark.parametrize("gin,axes,tin,tout", SHAPES)
@pytest.mark.parametrize("real", [False, True])
@pytest.mark.parametrize("fp64", [True, False])
def test_dot_product_identity(dev, gin, axes, tin, tout, real, fp64):
    """<F x, y> = <x, F^* y> and <B y, x> = <y, B^* x> (real inner products)."""
    import torch
    import dfft
    tol = 1e-12 if fp64 else 2e-6
    for norm in ["backward", "ortho", "forward", "none"]:
        plan, gout, idt, cdt = _plan(dev, gin, axes, tin, tout, real, fp64, norm)
        x = _blocks(dev, seeded_input(gin, real, 1), gin, tin, idt)
        y = _blocks(dev, seeded_input(gout, False, 2), gout, tout, cdt)
        fx = [torch.empty_like(t) for t in y]
        fty = [torch.empty_like(t) for t in x]
        plan._execute(x, fx, dfft.Direction.forward)
        plan.execute_adjoint(y, fty, dfft.Direction.forward)
        lhs, rhs = _rdot(fx, y), _rdot(x, fty)
        assert abs(lhs - rhs) <= tol * _norm(x) * _norm(fty) + tol * _norm(fx) * _norm(y), (norm, lhs, rhs)
        by = [torch.empty_like(t) for t in x]
        btx = [torch.empty_like(t) for t in y]
        plan._execute(y, by, dfft.Direction.backward)
        plan.execute_adjoint(x, btx, dfft.Direction.backward)
        lhs, rhs = _rdot(by, x), _rdot(y, btx)
        assert abs(lhs - rhs) <= tol * _norm(by) * _norm(x) + tol * _norm(y) * _norm(btx), (norm, lhs, rhs)
        plan.destroy()


@pytest.mark.parametrize("real", [False, True])
@pytest.mark.parametrize("exec_mode", [0, 1])
def test_autograd_matches_torch_fft_gradient(dev, real, exec_mode):
    """Gradient of L = Re <a, F x> + Re <b, B F x> through the harness vs torch.fft's own
    autograd on the CPU in FP64 (the FP64 reference gradient)."""
    import torch
    import dfft
    import dfft_autograd
    gin, axes, tin, tout = [12, 10, 8], [0, 1, 2], [1, 2, 1], [2, 1, 1]
    for norm in ["backward", "ortho"]:
        plan, gout, idt, cdt = _plan(dev, gin, axes, tin, tout, real, True, norm, exec_mode)
        xg = seeded_input(gin, real, 3)
        ag = seeded_input(gout, False, 4)
        bg = seeded_input(gin, real, 5)
        x = [t.requires_grad_(True) for t in _blocks(dev, xg, gin, tin, idt)]
        a = _blocks(dev, ag, gout, tout, cdt)
        b = _blocks(dev, bg, gin, tin, idt)
        fx = dfft_autograd.execute(plan, x, dfft.Direction.forward)
        bfx = dfft_autograd.execute(plan, fx, dfft.Direction.backward)
        loss = sum((torch.conj(p) * q).real.sum() for p, q in zip(a, fx))
        loss = loss + sum((torch.conj(p) * q).real.sum() for p, q in zip(b, bfx))
        gx = torch.autograd.grad(loss, x)
        got = PE.gather([t.cpu().numpy() for t in gx], gin, list(tin), np.float64 if real else np.complex128)

        xt = torch.from_numpy(xg).requires_grad_(True)
        at, bt = torch.from_numpy(ag), torch.from_numpy(bg)
        f = torch.fft.rfftn(xt, dim=axes, norm=norm if norm != "none" else "backward") if real else \
            torch.fft.fftn(xt, dim=axes, norm=norm)
        bf = torch.fft.irfftn(f, s=[gin[k] for k in axes], dim=axes, norm=norm) if real else \
            torch.fft.ifftn(f, dim=axes, norm=norm)
        lt = (torch.conj(at) * f).real.sum() + (torch.conj(bt) * bf).real.sum()
        (want,) = torch.autograd.grad(lt, xt)
        want = want.numpy()
        assert np.linalg.norm((got - want).ravel()) <= 1e-12 * np.linalg.norm(want.ravel())
        plan.destroy()


def test_autograd_central_differences(dev):
    """R2C forward + C2R backward round trip inside a nonlinear loss: harness gradient vs
    central differences along a random direction (eps = 1e-6)."""
    import torch
    import dfft
    import dfft_autograd
    gin, axes, tin, tout = [16, 12, 10], [0, 1, 2], [2, 1, 1], [1, 2, 1]
    plan, gout, idt, cdt = _plan(dev, gin, axes, tin, tout, True, True, "backward")
    x0 = _blocks(dev, seeded_input(gin, True, 6), gin, tin, idt)
    d = _blocks(dev, seeded_input(gin, True, 7), gin, tin, idt)

    def loss(x):
        f = dfft_autograd.execute(plan, x, dfft.Direction.forward)
        k = sum((t.abs() ** 3).sum() for t in f)  # non-quadratic in x
        back = dfft_autograd.execute(plan, [t * t.abs() for t in f], dfft.Direction.backward)
        return k + sum((v ** 2).sum() for v in back)

    x = [t.clone().requires_grad_(True) for t in x0]
    gx = torch.autograd.grad(loss(x), x)
    ad = sum(float((g * v).sum()) for g, v in zip(gx, d))
    eps = 1e-6
    with torch.no_grad():
        lp = float(loss([a + eps * b for a, b in zip(x0, d)]))
        lm = float(loss([a - eps * b for a, b in zip(x0, d)]))
    fd = (lp - lm) / (2 * eps)
    assert abs(ad - fd) <= 1e-6 * abs(fd), (ad, fd)
