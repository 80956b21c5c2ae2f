"""Bit-exact parity of the exchange layer on the GPU against the oracle.

Known-answer construction of the reference's Test_Pack.cpp:254-473 /
Test_Unpack.cpp:197-...: the global array holds i2 + i1*gn2 + i0*gn2*gn1 on
(8,7,5); every rank's local slab and every axis order is packed, and the
buffer must equal the oracle's bit for bit (pure copies -- no tolerance).
"""

import itertools

import numpy as np
import pytest

from oracle import packunpack as PU
from oracle import planner as PL
from oracle import plan_exec as PE

pytestmark = pytest.mark.gpu

GN = (8, 7, 5)


def _global(dtype):
    i0, i1, i2 = np.meshgrid(*[np.arange(n) for n in GN], indexing="ij")
    g = (i2 + i1 * GN[2] + i0 * GN[2] * GN[1]).astype(np.float64)
    if np.issubdtype(dtype, np.complexfloating):
        return (g + 1j * (1000 + g)).astype(dtype)
    return g.astype(dtype)


def _device(a, dev, layout):
    """numpy array -> device tensor of the same logical shape, stored
    row-major (LayoutRight) or column-major (LayoutLeft)."""
    import torch
    if layout == PL.LAYOUT_RIGHT:
        return torch.from_numpy(np.ascontiguousarray(a)).to(dev)
    rev = list(range(a.ndim))[::-1]
    return torch.from_numpy(np.ascontiguousarray(a.transpose(rev))).to(dev).permute(rev)


def _filled(shape, value, dtype, dev, layout):
    import torch
    if layout == PL.LAYOUT_RIGHT:
        return torch.full(list(shape), value, dtype=dtype, device=dev)
    rev = list(range(len(shape)))[::-1]
    return torch.full(list(shape)[::-1], value, dtype=dtype, device=dev).permute(rev)


@pytest.mark.parametrize("layout", [PL.LAYOUT_RIGHT, PL.LAYOUT_LEFT])
@pytest.mark.parametrize("dtype", [np.float32, np.float64, np.complex64, np.complex128])
def test_pack_unpack_known_answer(dev, dtype, layout):
    """LayoutLeft: column-major views and (e0, e1, e2, P) column-major buffers
    (the float/double x LayoutLeft/LayoutRight types of Test_Pack.cpp:11-14)."""
    import torch
    import dfft

    lay = dfft.LayoutLeft if layout == PL.LAYOUT_LEFT else dfft.LayoutRight
    g = _global(dtype)
    for nprocs in range(2, 5):
        topos = [[1, 1, nprocs], [1, nprocs, 1], [nprocs, 1, 1]]
        for tsrc, tdst in itertools.permutations(topos, 2):
            in_axis, out_axis = PL.slab_in_out_axes(tsrc, tdst)
            bext = PL.compute_buffer_extents(layout, list(GN), tsrc, tdst)
            for rank in range(nprocs):
                loc = PE.scatter(g, list(GN), tsrc)[rank]
                for m in itertools.permutations(range(3)):
                    src = np.ascontiguousarray(loc.transpose(m))
                    ref = PU.pack(src, bext, list(m), in_axis, layout)
                    ts = _device(src, dev, layout)
                    tb = _filled(bext, 7, ts.dtype, dev, layout)
                    dfft.pack(ts, tb, list(m), in_axis, layout=lay)
                    got = np.ascontiguousarray(tb.cpu().numpy())
                    assert np.array_equal(got.view(np.uint8), np.ascontiguousarray(ref).view(np.uint8)), \
                        (nprocs, tsrc, tdst, rank, m)
                    # unpack the same buffer back along in_axis into a zeroed array
                    td = _filled(src.shape, 0, ts.dtype, dev, layout)
                    dfft.unpack(tb, td, list(m), in_axis, layout=lay)
                    refu = PU.unpack(ref, src.shape, list(m), in_axis, layout=layout)
                    assert np.array_equal(np.ascontiguousarray(td.cpu().numpy()).view(np.uint8),
                                          np.ascontiguousarray(refu).view(np.uint8))


def test_pack_rejects_a_buffer_of_the_other_layout(dev):
    import torch
    import dfft
    src = torch.zeros(8, 7, 5, dtype=torch.float64, device=dev)
    right = torch.zeros(2, 8, 4, 5, dtype=torch.float64, device=dev)
    with pytest.raises(dfft.DFFTError):
        dfft.pack(src, right, [0, 1, 2], 1, layout=dfft.LayoutLeft)  # row-major arrays, LayoutLeft requested


def test_all2all_closed_form(dev):
    """Test_All2All.cpp:136-170: values encode (rank, index)."""
    import torch
    import dfft

    for n in [1, 2, 3, 4, 7]:
        sends = [torch.arange(n * 6, dtype=torch.float64, device=dev).reshape(n, 6) + 1000 * r for r in range(n)]
        recvs = [torch.zeros(n, 6, dtype=torch.float64, device=dev) for _ in range(n)]
        dfft.all2all(sends, recvs)
        for q in range(n):
            for p in range(n):
                expect = torch.arange(q * 6, q * 6 + 6, dtype=torch.float64) + 1000 * p
                assert torch.equal(recvs[q][p].cpu(), expect)
    with pytest.raises(dfft.DFFTError):
        bad = [torch.zeros(3, 2, device=dev) for _ in range(2)]
        dfft.all2all(bad, [torch.zeros(3, 2, device=dev) for _ in range(2)])


def _oracle_transblock(locals_, plan, bi, forward=True):
    """pack -> all2all -> unpack of block bi for every (sub-)communicator."""
    return plan._exchange(locals_, bi, forward)


@pytest.mark.parametrize("case", [
    ([8, 7, 5], [2, 1, 1], [1, 2, 1], [0, 1, 2], False),
    ([8, 7, 5], [1, 1, 3], [3, 1, 1], [2, 0, 1], False),
    ([9, 7, 6], [1, 4, 1], [1, 1, 4], [0, 1, 2], True),
    ([8, 7, 6], [2, 2, 1], [1, 2, 2], [0, 1, 2], True),
    ([8, 7, 6], [1, 2, 3], [2, 3, 1], [0, 2, 1], False),
    ([8, 7, 6], [3, 2, 1], [1, 3, 2], [1, 2, 0], False),
])
def test_exchange_matches_pack_all2all_unpack(dev, case):
    """dfft_exchange (one fused remap) == oracle pack->all2all->unpack, bit for bit,
    for every transpose block of the plan, forward and backward."""
    import torch
    import dfft

    gin, tin, tout, axes, real = case
    plan = PE.OraclePlan(gin, [n // 2 + 1 if (real and a == axes[-1]) else n for a, n in enumerate(gin)],
                         axes, tin, tout, real)
    rng = np.random.default_rng(3)
    topo_idx = 0
    for bi, b in enumerate(plan.blocks):
        if b.block_type != PL.TRANSPOSE:
            continue
        src_t = plan.analyses[0].topologies[topo_idx]
        dst_t = plan.analyses[0].topologies[topo_idx + 1]
        lays = plan.analyses[0].layouts
        src_l = PL.LAYOUT_RIGHT if (not lays or lays[topo_idx]) else PL.LAYOUT_LEFT
        dst_l = PL.LAYOUT_RIGHT if (not lays or lays[topo_idx + 1]) else PL.LAYOUT_LEFT
        topo_idx += 1
        # global extents of the data at this block: real input before the first FFT block
        nfft_before = sum(1 for k in range(bi) if plan.blocks[k].block_type == PL.FFT)
        gext = plan.gin if (real and nfft_before == 0) else plan.gout
        dtype = np.float64 if (real and nfft_before == 0) else np.complex128
        locals_ = [np.ascontiguousarray(
            (rng.standard_normal(plan.analyses[r].blocks[bi].in_extents)
             + (0 if dtype == np.float64 else 1j * rng.standard_normal(plan.analyses[r].blocks[bi].in_extents))
             ).astype(dtype)) for r in range(plan.P)]
        ref = _oracle_transblock(locals_, plan, bi, True)
        ins = [torch.from_numpy(a).to(dev) for a in locals_]
        outs = [torch.zeros(plan.analyses[r].blocks[bi].out_extents, dtype=ins[0].dtype, device=dev)
                for r in range(plan.P)]
        dfft.exchange(ins, outs, gext, dfft.Topology(src_t, dfft.LayoutRight if src_l == PL.LAYOUT_RIGHT
                                                     else dfft.LayoutLeft), b.in_map,
                      dfft.Topology(dst_t, dfft.LayoutRight if dst_l == PL.LAYOUT_RIGHT else dfft.LayoutLeft),
                      b.out_map)
        torch.cuda.synchronize()
        for r in range(plan.P):
            assert np.array_equal(outs[r].cpu().numpy(), ref[r]), (case, bi, r)


@pytest.mark.parametrize("shape", [[8, 7, 5], [4, 6, 3, 5]])
def test_local_transpose_is_numpy_transpose(dev, shape):
    """kokkos-fft Impl::transpose(in, out, map) on one rank (SlabPlan.hpp final
    local transposes): out.extent(k) = in.extent(map[k]), out = numpy.transpose(in, map),
    bit for bit, through dfft_exchange with a single rank."""
    import itertools
    import torch
    import dfft
    rng = np.random.default_rng(11)
    x = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
    ones = [1] * len(shape)
    ident = list(range(len(shape)))
    for m in itertools.permutations(ident):
        want = np.transpose(x, m)
        out = torch.zeros(list(want.shape), dtype=torch.complex128, device=dev)
        dfft.exchange([torch.from_numpy(x).to(dev)], [out], shape, ones, ident, ones, list(m))
        torch.cuda.synchronize()
        assert np.array_equal(out.cpu().numpy(), want), m
