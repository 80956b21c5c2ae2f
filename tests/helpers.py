"""Shared helpers of the test-suite (seeded inputs, GPU plan runner)."""

import numpy as np

from oracle import fft as F
from oracle import planner as PL
from oracle import plan_exec as PE


def seeded_input(shape, real, seed, dist="uniform"):
    rng = np.random.default_rng(seed)
    if dist == "uniform":
        re = rng.uniform(-1.0, 1.0, shape)
        im = rng.uniform(-1.0, 1.0, shape)
    else:
        re = rng.standard_normal(shape)
        im = rng.standard_normal(shape)
    return re if real else re + 1j * im


def out_extents(gin, axes, real):
    gout = list(gin)
    if real:
        gout[axes[-1]] = gin[axes[-1]] // 2 + 1
    return gout


def rel_l2(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    nb = np.linalg.norm(b.ravel())
    return float(np.linalg.norm((a - b).ravel()) / (nb if nb > 0 else 1.0))


def gather_rel_l2(gpu_locals, ref_locals):
    num = sum(np.linalg.norm((g - r).ravel()) ** 2 for g, r in zip(gpu_locals, ref_locals))
    den = sum(np.linalg.norm(r.ravel()) ** 2 for r in ref_locals)
    return float(np.sqrt(num / den)) if den > 0 else float(np.sqrt(num))


def device_array(a, dev, dtype, left=False):
    """numpy array -> device tensor of the same logical shape, row-major
    (LayoutRight) or column-major (LayoutLeft: strides grow along the axes)."""
    import torch
    if not left:
        return torch.from_numpy(np.ascontiguousarray(a)).to(dev, dtype)
    rev = list(range(a.ndim))[::-1]
    return torch.from_numpy(np.ascontiguousarray(a.transpose(rev))).to(dev, dtype).permute(rev)


def device_full(shape, value, dtype, dev, left=False):
    import torch
    if not left:
        return torch.full(list(shape), value, dtype=dtype, device=dev)
    rev = list(range(len(shape)))[::-1]
    return torch.full(list(shape)[::-1], value, dtype=dtype, device=dev).permute(rev)


def run_gpu_plan(dev, x, gin, axes, tin, tout, real, fp64=True, norm="backward", exec_mode=0,
                 tin_layout=None, tout_layout=None, configure=None, view_layout=None):
    """Forward + backward through the C ABI; returns (outs, backs, plan) as numpy.
    view_layout="left": every view is a LayoutLeft (column-major) array."""
    import torch
    import dfft
    left = view_layout == "left"

    tin_t = dfft.Topology(tin, dfft.LayoutLeft if tin_layout == "left" else dfft.LayoutRight)
    tout_t = dfft.Topology(tout, dfft.LayoutLeft if tout_layout == "left" else dfft.LayoutRight)
    lin = PL.LAYOUT_LEFT if tin_layout == "left" else PL.LAYOUT_RIGHT
    lout = PL.LAYOUT_LEFT if tout_layout == "left" else PL.LAYOUT_RIGHT
    gout = out_extents(gin, axes, real)
    rdt = torch.float64 if fp64 else torch.float32
    cdt = torch.complex128 if fp64 else torch.complex64
    ins = [device_array(a, dev, rdt if real else cdt, left) for a in PE.scatter(x, gin, tin, lin)]
    outs = []
    for r in range(len(ins)):
        e, _ = PL.compute_local_extents_and_starts(gout, list(tout), r, lout)
        outs.append(device_full(e, float("nan"), cdt, dev, left))
    plan = dfft.Plan(ins, outs, axes, tin_t, tout_t, norm=norm, exec_mode=exec_mode,
                     global_in_extents=gin, global_out_extents=gout,
                     layout=dfft.LayoutLeft if left else dfft.LayoutRight)
    if configure is not None:
        configure(plan, ins, outs)
        ins = [device_array(a, dev, rdt if real else cdt, left) for a in PE.scatter(x, gin, tin, lin)]
    dfft.execute(plan, ins, outs, dfft.Direction.forward)
    backs = [device_full(list(a.shape), float("nan"), a.dtype, dev, left) for a in ins]
    dfft.execute(plan, outs, backs, dfft.Direction.backward)
    torch.cuda.synchronize()
    res = ([o.cpu().numpy() for o in outs], [b.cpu().numpy() for b in backs],
           [a.cpu().numpy() for a in ins], plan)
    return res


def reference_spectrum(x, gin, axes, tout, real, norm="backward", tout_layout=None):
    lout = PL.LAYOUT_LEFT if tout_layout == "left" else PL.LAYOUT_RIGHT
    gout = out_extents(gin, axes, real)
    return PE.scatter(F.fftn(x, axes, real, norm), gout, list(tout), lout)
