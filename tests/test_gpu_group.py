"""Groups of independent transforms of mixed extents (dfft_group_*): every
member against the oracle, graph replay against eager, and the 64-member
mixed batch N in {96, 128, 160}^3 complex64."""

import numpy as np
import pytest

from helpers import out_extents, rel_l2, seeded_input
from oracle import fft as F

pytestmark = pytest.mark.gpu


def _members(dev, shapes, real, fp64, axes_list, seed):
    import torch
    rdt = torch.float64 if fp64 else torch.float32
    cdt = torch.complex128 if fp64 else torch.complex64
    xs, ins, outs = [], [], []
    for i, (shape, axes) in enumerate(zip(shapes, axes_list)):
        x = seeded_input(shape, real, seed + i)
        xs.append(x)
        ins.append(torch.from_numpy(x).to(dev, rdt if real else cdt))
        outs.append(torch.full(out_extents(shape, axes, real), float("nan"), dtype=cdt, device=dev))
    return xs, ins, outs


@pytest.mark.parametrize("real", [False, True])
@pytest.mark.parametrize("fp64", [True, False])
def test_mixed_group_matches_oracle(dev, real, fp64):
    """Lengths with a grouped kernel (16, 32, 64, 96, 160) mixed with lengths
    launched one by one (24, 12, 7), batched 4-D members, 1- and 2-axis members."""
    import torch
    import dfft
    shapes = [[16, 16, 16], [32, 24, 16], [12, 96, 8], [7, 16, 32], [3, 32, 32, 32], [160, 8, 4], [16, 64],
              [20, 96]]
    axes_list = [[0, 1, 2], [0, 1, 2], [0, 1, 2], [0, 1, 2], [1, 2, 3], [0, 2, 1], [0, 1], [1]]
    xs, ins, outs = _members(dev, shapes, real, fp64, axes_list, 10)
    g = dfft.Group(ins, outs, axes_list)
    g.execute(ins, outs, dfft.Direction.forward)
    backs = [torch.full(list(t.shape), float("nan"), dtype=t.dtype, device=dev) for t in ins]
    g.execute(outs, backs, dfft.Direction.backward)
    torch.cuda.synchronize()
    tol = 1e-12 if fp64 else 2e-6
    for x, o, b, axes in zip(xs, outs, backs, axes_list):
        ref = F.fftn(x, axes, real)
        assert rel_l2(o.cpu().numpy(), ref) <= tol
        assert rel_l2(b.cpu().numpy(), x) <= tol
    # the nine members with a grouped length share one launch per pass; 7, 12, 20, 24 get their own
    assert g.num_launches(dfft.Direction.forward) >= 3


def test_group_unaligned_complex64_members(dev):
    """complex64 members viewed at an odd element offset (rows start 8 bytes
    past a 16-byte boundary) are moved element by element, beside aligned
    members of the same lengths whose contiguous lines go 16 bytes per lane
    through LDS and whose strided lines go two per lane; inputs offset on even
    members, backward outputs offset on odd ones."""
    import torch
    import dfft
    shapes = [[8, 16, 96], [8, 16, 96], [4, 8, 160], [4, 8, 160], [8, 8, 128], [8, 8, 128]]
    axes_list = [[0, 1, 2]] * len(shapes)
    xs, ins, outs = _members(dev, shapes, False, False, axes_list, 40)

    def shifted(t):
        buf = torch.full((t.numel() + 1,), float("nan"), dtype=t.dtype, device=dev)
        v = buf[1:].view(t.shape)
        v.copy_(t)
        return v

    ins = [shifted(t) if i % 2 == 0 else t for i, t in enumerate(ins)]
    backs = [torch.full(list(t.shape), float("nan"), dtype=t.dtype, device=dev) for t in ins]
    backs = [shifted(t) if i % 2 == 1 else t for i, t in enumerate(backs)]
    assert ins[0].data_ptr() % 16 == 8 and backs[1].data_ptr() % 16 == 8
    g = dfft.Group(ins, outs, axes_list)
    g.execute(ins, outs, dfft.Direction.forward)
    g.execute(outs, backs, dfft.Direction.backward)
    torch.cuda.synchronize()
    for x, o, b, axes in zip(xs, outs, backs, axes_list):
        assert rel_l2(o.cpu().numpy(), F.fftn(x, axes, False)) <= 2e-6
        assert rel_l2(b.cpu().numpy(), x) <= 2e-6
    g.destroy()


@pytest.mark.parametrize("lanes,stage", [(256, 1), (64, 0), (64, 1)])
def test_group_graph_replay_matches_eager(dev, lanes, stage):
    """Graph replays equal the eager launches bit for bit, every replay, also
    with the contiguous launches in 64-lane blocks, staged or not."""
    import torch
    import dfft
    shapes = [[32, 32, 32], [16, 96, 16], [160, 16, 8], [24, 16, 16]]
    axes_list = [[0, 1, 2]] * 4
    xs, ins, outs = _members(dev, shapes, False, False, axes_list, 20)
    eager = [torch.empty_like(t) for t in outs]
    g = dfft.Group(ins, outs, axes_list, norm="ortho")
    g.set_option("lanes", lanes)
    g.set_option("stage", stage)
    g.execute(ins, eager, dfft.Direction.forward)
    for rep in range(4):
        for t in outs:
            t.fill_(float("nan"))
        g.execute(ins, outs, dfft.Direction.forward, use_graph=True)
        torch.cuda.synchronize()
        for a, b in zip(outs, eager):
            assert bool((a == b).all()), rep
    # new pointers rebuild the tables and the graph
    outs2 = [torch.empty_like(t) for t in outs]
    g.execute(ins, outs2, dfft.Direction.forward, use_graph=True)
    g.execute(ins, outs2, dfft.Direction.forward, use_graph=True)
    torch.cuda.synchronize()
    for a, b in zip(outs2, eager):
        assert bool((a == b).all())


def test_group_rejects_bad_members(dev):
    import torch
    import dfft
    a = torch.zeros(8, 8, 8, dtype=torch.complex64, device=dev)
    b = torch.zeros(8, 8, 8, dtype=torch.complex128, device=dev)
    with pytest.raises(dfft.DFFTError) as e:
        dfft.Group([a, b], [a.clone(), b.clone()])
    assert e.value.code == dfft.ERROR_MISMATCH
    r = torch.zeros(8, 8, 8, dtype=torch.float32, device=dev)
    c = torch.zeros(8, 8, 5, dtype=torch.complex64, device=dev)
    g = dfft.Group([r], [c])
    # arrays of the wrong type are refused before any pointer reaches the library
    with pytest.raises(dfft.DFFTError) as e:
        g.execute([c], [c], dfft.Direction.forward)
    assert e.value.code == dfft.ERROR_MISMATCH
    # an R2C member whose real input and half spectrum share their memory
    base = torch.zeros(8 * 8 * 10, dtype=torch.float32, device=dev)
    r_alias = base[:8 * 8 * 8].view(8, 8, 8)
    c_alias = base.view(torch.complex64).view(8, 8, 5)
    with pytest.raises(dfft.DFFTError) as e:
        g.execute([r_alias], [c_alias], dfft.Direction.forward)
    assert e.value.code == dfft.ERROR_ALIAS
    with pytest.raises(dfft.DFFTError) as e:
        dfft.Group([r], [torch.zeros(8, 8, 6, dtype=torch.complex64, device=dev)])
    assert e.value.code == dfft.ERROR_INVALID_EXTENTS


def test_mixed_batch_64_transforms_96_128_160(dev):
    """64 C2C transforms, N cycling through {96, 128, 160}^3, complex64, uniform [-1, 1]:
    one launch per pass; spot members against the FP64 oracle, all round trips."""
    import torch
    import dfft
    sizes = [(96, 128, 160)[i % 3] for i in range(64)]
    gen = torch.Generator(device=dev).manual_seed(1234)
    ins, outs = [], []
    for n in sizes:
        re = torch.rand((n, n, n), generator=gen, device=dev, dtype=torch.float32) * 2 - 1
        im = torch.rand((n, n, n), generator=gen, device=dev, dtype=torch.float32) * 2 - 1
        ins.append(torch.complex(re, im))
        outs.append(torch.empty((n, n, n), dtype=torch.complex64, device=dev))
    g = dfft.Group(ins, outs)
    assert g.num_launches(dfft.Direction.forward) == 3
    assert g.num_launches(dfft.Direction.backward) == 3
    g.execute(ins, outs, dfft.Direction.forward, use_graph=True)
    backs = [torch.empty_like(t) for t in ins]
    g.execute(outs, backs, dfft.Direction.backward, use_graph=True)
    torch.cuda.synchronize()
    worst = 0.0
    for i, (x, b) in enumerate(zip(ins, backs)):
        num = float(torch.linalg.vector_norm(b - x))
        den = float(torch.linalg.vector_norm(x))
        worst = max(worst, num / den)
    assert worst <= 1e-6
    for i in (0, 1, 2, 63):
        x = ins[i].cpu().numpy().astype(np.complex128)
        ref = np.fft.fftn(x)
        err = rel_l2(outs[i].cpu().numpy().astype(np.complex128), ref)
        print(f"member {i} N={sizes[i]} rel-L2 vs FP64 {err:.2e}")
        assert err <= 1e-6


def test_group_dual_lines(dev):
    """FP32 groups with two adjacent lines per lane on the strided passes (dual
    on) and without (off) match the oracle; the member whose rows are odd (its
    lines are no adjacent pairs) is grouped with the one-line kernel, so with
    dual on each of its two strided passes adds a second grouped launch."""
    import torch
    import dfft
    shapes = [[96, 128, 16], [128, 96, 16], [160, 64, 16], [128, 64, 7]]
    axes_list = [[0, 1, 2]] * len(shapes)
    launches = {}
    for dual in (1, 0):
        xs, ins, outs = _members(dev, shapes, False, False, axes_list, 30)
        g = dfft.Group(ins, outs, axes_list)
        g.set_option("dual", dual)
        assert g.get_option("dual") == dual
        g.execute(ins, outs, dfft.Direction.forward)
        backs = [torch.full(list(t.shape), float("nan"), dtype=t.dtype, device=dev) for t in ins]
        g.execute(outs, backs, dfft.Direction.backward)
        torch.cuda.synchronize()
        for x, o, b, axes in zip(xs, outs, backs, axes_list):
            assert rel_l2(o.cpu().numpy(), F.fftn(x, axes, False)) <= 2e-6, f"dual={dual}"
            assert rel_l2(b.cpu().numpy(), x) <= 2e-6, f"dual={dual}"
        launches[dual] = g.num_launches(dfft.Direction.forward)
        g.destroy()
    assert launches[1] == launches[0] + 2


@pytest.mark.parametrize("prec", ["fp64", "fp32-staged", "fp32"])
@pytest.mark.parametrize("real", [False, True])
@pytest.mark.parametrize("lanes", [64, 128, 256])
def test_group_contiguous_lanes(dev, lanes, real, prec):
    """Launches of contiguous lines in blocks of 64, 128 or 256 lanes (every
    grouped length: 2 to 32 lanes per line, so 2 to 32 lines per block), with
    persistent and one-block-per-virtual-block grids, FP32 with and without
    the LDS staging of contiguous lines, cached and nontemporal accesses of
    those launches, against the oracle.  The
    first pass of these members walks contiguous lines only; with the
    (16, 64, 5) member over axes [0, 1] added, its strided axis-1 lines join
    that launch, which then keeps 256-lane blocks."""
    import torch
    import dfft
    shapes = [[16, 16, 16], [8, 32, 32], [4, 6, 64], [12, 96, 8], [4, 8, 128], [160, 8, 4], [2, 4, 256],
              [6, 5, 160], [16, 64], [20, 96]]
    axes_list = [[0, 1, 2]] * 8 + [[0, 1], [1]]
    fp64 = prec == "fp64"
    tol = 1e-12 if fp64 else 2e-6
    for mult, extra in ((0, False), (1, False), (0, True)):
        sh = shapes + ([[16, 64, 5]] if extra else [])
        ax = axes_list + ([[0, 1]] if extra else [])
        xs, ins, outs = _members(dev, sh, real, fp64, ax, 50 + lanes)
        g = dfft.Group(ins, outs, ax)
        g.set_option("lanes", lanes)
        g.set_option("grid_mult", mult)
        g.set_option("stage", 0 if prec == "fp32" else 1)
        g.set_option("contig_nt", mult)
        assert g.get_option("lanes") == lanes and g.get_option("contig_nt") == mult
        g.execute(ins, outs, dfft.Direction.forward)
        backs = [torch.full(list(t.shape), float("nan"), dtype=t.dtype, device=dev) for t in ins]
        g.execute(outs, backs, dfft.Direction.backward)
        torch.cuda.synchronize()
        for x, o, b, axes in zip(xs, outs, backs, ax):
            what = f"lanes={lanes} mult={mult} extra={extra} {x.shape}"
            assert rel_l2(o.cpu().numpy(), F.fftn(x, axes, real)) <= tol, what
            assert rel_l2(b.cpu().numpy(), x) <= tol, what
        if extra:
            with pytest.raises(dfft.DFFTError):
                g.set_option("lanes", 96)
            with pytest.raises(dfft.DFFTError):
                g.set_option("stage", 2)
            with pytest.raises(dfft.DFFTError):
                g.set_option("contig_nt", 2)
        g.destroy()
