"""CPU tests that pin the oracle itself (no GPU).

* planner: the reference's own golden tables (Test_Topologies.cpp), the
  BlockInfo known answers of Test_BlockAnalyses.cpp and the mapping table
  of Test_Mapping.cpp;
* pack/unpack: the expected buffers of Test_Pack.cpp / Test_Unpack.cpp,
  rebuilt from the global array as those tests do, plus bin_mapping cases;
* FFT: direct long-double DFT, numpy's pocketfft, closed forms;
* executor: the plan tests' strategy (global FFT sliced per rank).
"""

import itertools

import numpy as np
import pytest

from oracle import fft as F
from oracle import packunpack as PU
from oracle import plan_exec as PE
from oracle import planner as PL
from topology_cases import pencil_cases, slab_cases


# ----------------------------------------------------------------- planner
@pytest.mark.parametrize("nprocs", [2, 3, 4, 5, 6])
def test_slab_topologies_golden(nprocs):
    n = 0
    for fn, tin, tout, axes, ref in slab_cases(nprocs):
        assert PL.get_all_slab_topologies(tin, tout, axes) == ref, (fn, tin, tout, axes)
        n += 1
    assert n == 257


@pytest.mark.parametrize("nprocs", [2, 3, 4, 5, 6])
def test_pencil_topologies_golden(nprocs):
    n = 0
    for fn, tin, lin, tout, lout, axes, rax, rlay, ref in pencil_cases(nprocs):
        got = PL.get_all_pencil_topologies(tin, tout, axes, lin, lout)
        assert got == (ref, rax, rlay), (fn, tin, lin, tout, lout, axes)
        n += 1
    assert n == 410


def test_slab_topologies_throw_for_shared():
    """Test_Topologies.cpp:2022-2033: nprocs == 1 topologies are not slabs."""
    for axes in itertools.permutations([0, 1, 2]):
        with pytest.raises(PL.PlannerError):
            PL.get_all_slab_topologies([1, 1, 1], [1, 1, 1], list(axes))


def test_permute_map_by_axes_golden():
    """Test_Mapping.cpp:60-200 (3-D, LayoutRight and LayoutLeft)."""
    R, L = PL.LAYOUT_RIGHT, PL.LAYOUT_LEFT
    src = [0, 1, 2]
    assert PL.permute_map_by_axes(R, src, 0) == [1, 2, 0]
    assert PL.permute_map_by_axes(R, src, 1) == [0, 2, 1]
    assert PL.permute_map_by_axes(R, src, 2) == [0, 1, 2]
    assert PL.permute_map_by_axes(L, src, 0) == [0, 1, 2]
    assert PL.permute_map_by_axes(L, src, 1) == [1, 0, 2]
    assert PL.permute_map_by_axes(L, src, 2) == [2, 0, 1]
    # the header's worked example: src map (0, 2, 1)
    assert PL.permute_map_by_axes(L, [0, 2, 1], 0) == [0, 2, 1]
    assert PL.permute_map_by_axes(L, [0, 2, 1], 1) == [1, 0, 2]
    assert PL.permute_map_by_axes(L, [0, 2, 1], 2) == [2, 0, 1]
    assert PL.permute_map_by_axes(R, [0, 2, 1], 0) == [2, 1, 0]
    assert PL.permute_map_by_axes(R, [0, 2, 1], 1) == [0, 2, 1]
    assert PL.permute_map_by_axes(R, [0, 2, 1], 2) == [0, 1, 2]
    assert PL.permute_map_by_axes(R, src, [1, 2]) == [0, 1, 2]
    assert PL.permute_map_by_axes(R, src, [2, 1]) == [0, 2, 1]
    assert PL.permute_map_by_axes(L, src, [1, 2]) == [2, 1, 0]


def test_buffer_extents_golden():
    """Test_Extents.cpp:20-70: (n0,n1,n2,n3) with x/y pencils over (p0, p1)."""
    n0, n1, n2, n3 = 13, 8, 17, 5
    p0, p1 = 2, 3
    tin, tout = [1, p0, p1, 1], [p0, 1, p1, 1]
    exp_r = [p0, (n0 - 1) // p0 + 1, (n1 - 1) // p0 + 1, (n2 - 1) // p1 + 1, n3]
    assert PL.compute_buffer_extents(PL.LAYOUT_RIGHT, [n0, n1, n2, n3], tin, tout) == exp_r
    assert PL.compute_buffer_extents(PL.LAYOUT_LEFT, [n0, n1, n2, n3], tin, tout) == exp_r[1:] + [p0]
    with pytest.raises(PL.PlannerError):
        PL.compute_buffer_extents(PL.LAYOUT_RIGHT, [n0, n1, n2, n3], [1, p0, p1, 1], [p1, 1, p0, 1])


def test_fft_extents_golden():
    """Extents.hpp:101-115 worked example: in (8,7,8), out (8,7,5), axes (0,1,2)."""
    assert PL.compute_fft_extents([8, 7, 8], [8, 7, 5], [0, 1, 2]) == [8, 7, 8]


def test_rank_to_coord_golden():
    """MPI_Extents.hpp:474-491 examples (LayoutRight / LayoutLeft on (2,2))."""
    assert [PL.rank_to_coord([2, 2], r) for r in range(4)] == [[0, 0], [0, 1], [1, 0], [1, 1]]
    assert [PL.rank_to_coord([2, 2], r, PL.LAYOUT_LEFT) for r in range(4)] == [[0, 0], [1, 0], [0, 1], [1, 1]]


def test_local_extents_and_starts_golden():
    """MPI_Extents.hpp:20-50 and :531-545 worked examples."""
    # global (5, 11, 19) on topology (1, 2, 4), LayoutRight
    ext = [PL.compute_local_extents_and_starts([5, 11, 19], [1, 2, 4], r) for r in range(8)]
    assert [e for e, _ in ext] == [[5, 6, 5], [5, 6, 5], [5, 6, 5], [5, 6, 4],
                                   [5, 5, 5], [5, 5, 5], [5, 5, 5], [5, 5, 4]]
    assert [s for _, s in ext] == [[0, 0, 0], [0, 0, 5], [0, 0, 10], [0, 0, 15],
                                   [0, 6, 0], [0, 6, 5], [0, 6, 10], [0, 6, 15]]
    extl = [PL.compute_local_extents_and_starts([5, 11, 19], [1, 2, 4], r, PL.LAYOUT_LEFT) for r in range(8)]
    assert [s for _, s in extl] == [[0, 0, 0], [0, 6, 0], [0, 0, 5], [0, 6, 5],
                                    [0, 0, 10], [0, 6, 10], [0, 0, 15], [0, 6, 15]]
    # global (10, Y, Z) on (3, 1, 1)
    e = [PL.compute_local_extents_and_starts([10, 4, 4], [3, 1, 1], r) for r in range(3)]
    assert [x[0][0] for x in e] == [4, 3, 3] and [x[1][0] for x in e] == [0, 4, 7]


def test_block_analyses_known_answers():
    """Config A / config B block lists (hand-derived from BlockAnalyses.hpp:158-395)."""
    # 64^3 C2C slab (2,1,1) -> (1,2,1)
    ba = PL.block_analyses(PL.SLAB, False, PL.LAYOUT_RIGHT, [32, 64, 64], [64] * 3, [64] * 3,
                           [2, 1, 1], [1, 2, 1], [0, 1, 2], 0)
    assert [b.block_type for b in ba.blocks] == [PL.FFT, PL.TRANSPOSE, PL.FFT]
    assert ba.blocks[0].fft_dim == 2 and ba.blocks[0].in_extents == [32, 64, 64]
    t = ba.blocks[1]
    assert (t.in_axis, t.out_axis, t.buffer_extents, t.out_map) == (1, 0, [2, 32, 32, 64], [1, 2, 0])
    assert t.out_extents == [32, 64, 64]
    assert ba.blocks[2].fft_dim == 1 and ba.blocks[2].in_map == [1, 2, 0]
    # 256^3 R2C pencil (2,2,1) -> (1,2,2): topology chain through the dummy substitution
    ba = PL.block_analyses(PL.PENCIL, True, PL.LAYOUT_RIGHT, [128, 128, 256], [256] * 3, [256, 256, 129],
                           [2, 2, 1], [1, 2, 2], [0, 1, 2], 0)
    assert ba.topologies == [[2, 2, 1], [2, 1, 2], [1, 2, 2]]
    assert ba.trans_axes == [1, 0] and ba.layouts == [1, 1, 1]
    assert ba.all_axes == [[2], [1], [0]]
    types = [b.block_type for b in ba.blocks]
    assert types == [PL.FFT, PL.TRANSPOSE, PL.FFT, PL.TRANSPOSE, PL.FFT]
    assert ba.blocks[0].out_extents == [128, 128, 129]
    assert ba.blocks[1].buffer_extents == [2, 128, 128, 65] and ba.blocks[1].out_map == [0, 2, 1]
    assert ba.blocks[3].buffer_extents == [2, 128, 128, 65] and ba.blocks[3].out_map == [2, 1, 0]
    assert (ba.blocks[3].in_axis, ba.blocks[3].out_axis, ba.blocks[3].comm_axis) == (1, 0, 0)


# ------------------------------------------------------------- pack/unpack
@pytest.mark.parametrize("n,nbins,starts,lengths", [
    (10, 3, [0, 4, 7], [4, 3, 3]),
    (12, 3, [0, 4, 8], [4, 4, 4]),
    (128, 6, [0, 22, 44, 65, 86, 107], [22, 22, 21, 21, 21, 21]),
])
def test_bin_mapping_golden(n, nbins, starts, lengths):
    """Test_Pack.cpp:590-608."""
    got = [PU.bin_mapping(n, nbins, p) for p in range(nbins)]
    assert [s for s, _ in got] == starts and [l for _, l in got] == lengths


def test_merge_indices():
    """Test_Pack.cpp:57-84."""
    for start, length in [(0, 4), (4, 3), (7, 3)]:
        for idx in range(6):
            m = PU.merge_indices(idx, start, length, 0, 0)
            assert m == (PU.OUT_OF_BOUNDS if idx >= length else idx + start)
            assert PU.merge_indices(idx, start, length, 0, 1) == idx


def _test_pack_expected(gu, tsrc_axis_len, nprocs, rank, split, src_local, buffer_shape):
    """Expected send buffer as Test_Pack.cpp:386-473 builds it: for each buffer
    element (p, i0, i1, i2) take the global-index value when the bin of the split
    axis and the local source extents allow it."""
    ref = np.zeros(buffer_shape, dtype=gu.dtype)
    P = buffer_shape[0]
    for p in range(P):
        for idx in itertools.product(*[range(n) for n in buffer_shape[1:]]):
            start, length = PU.bin_mapping(src_local.shape[split], P, p)
            gi = list(idx)
            gi[split] = idx[split] + start
            if idx[split] < length and all(gi[a] < src_local.shape[a] for a in range(3)):
                ref[(p,) + idx] = src_local[tuple(gi)]
    return ref


@pytest.mark.parametrize("layout", [PL.LAYOUT_RIGHT, PL.LAYOUT_LEFT])
@pytest.mark.parametrize("nprocs", [1, 2, 3, 4])
def test_pack_matches_reference_construction(nprocs, layout):
    """Test_Pack.cpp:386-473 for both buffer layouts: LayoutLeft buffers are
    (e0, e1, e2, P) with the process index last (p = i3)."""
    gn = (8, 7, 5)
    i0, i1, i2 = np.meshgrid(*[np.arange(n) for n in gn], indexing="ij")
    gu = (i2 + i1 * gn[2] + i0 * gn[2] * gn[1]).astype(np.float64)
    topos = [[1, 1, nprocs], [1, nprocs, 1], [nprocs, 1, 1]]
    for tsrc, tdst in itertools.permutations(topos, 2):
        if nprocs == 1:
            continue
        split, _ = PL.slab_in_out_axes(tsrc, tdst)
        bext = PL.compute_buffer_extents(layout, list(gn), tsrc, tdst)
        bright = PL.compute_buffer_extents(PL.LAYOUT_RIGHT, list(gn), tsrc, tdst)
        assert bext == (bright if layout == PL.LAYOUT_RIGHT else bright[1:] + bright[:1])
        for rank in range(nprocs):
            loc = PE.scatter(gu, list(gn), tsrc)[rank]
            exp = _test_pack_expected(gu, None, nprocs, rank, split, loc, bright)
            if layout == PL.LAYOUT_LEFT:
                exp = np.moveaxis(exp, 0, -1)
            for m in itertools.permutations(range(3)):
                src = np.ascontiguousarray(loc.transpose(m))
                got = PU.pack(src, bext, list(m), split, layout)
                assert np.array_equal(got, exp), (tsrc, tdst, rank, m)
                assert np.array_equal(PU.pack_literal(src, bext, list(m), split, layout), exp)
                back = PU.unpack(got, src.shape, list(m), split, layout=layout)
                assert np.array_equal(back, src)
                assert np.array_equal(PU.unpack_literal(got, src.shape, list(m), split, layout), src)


def test_all2all_left_moves_the_last_axis():
    """LayoutLeft buffers exchange blocks of their last (process) axis."""
    sends = [np.arange(12.0).reshape(2, 2, 3) + 100 * r for r in range(3)]
    recvs = PU.all2all(sends, PL.LAYOUT_LEFT)
    right = PU.all2all([np.moveaxis(a, -1, 0) for a in sends])
    for a, b in zip(recvs, right):
        assert np.array_equal(a, np.moveaxis(b, 0, -1))
    for q in range(3):
        for p in range(3):
            assert np.array_equal(recvs[q][..., p], sends[p][..., q])
    with pytest.raises(RuntimeError):
        PU.all2all([np.zeros((2, 3)), np.zeros((2, 3))], PL.LAYOUT_LEFT)


def test_all2all_rejects_wrong_proc_extent():
    """Test_All2All.cpp:196-230."""
    with pytest.raises(RuntimeError):
        PU.all2all([np.zeros((3, 2)), np.zeros((3, 2))])


def test_exchange_is_a_redistribution():
    """pack -> all2all -> unpack of every transpose block (including pencil
    sub-communicators) moves each rank's block of topology t_i to its block of
    t_{i+1} -- the property the fused one-GPU executor relies on."""
    cases = [([8, 7, 5], [2, 1, 1], [1, 2, 1], [0, 1, 2], False),
             ([8, 7, 6], [2, 2, 1], [1, 2, 2], [0, 1, 2], True),
             ([8, 7, 6], [1, 2, 3], [2, 3, 1], [0, 2, 1], False),
             ([8, 7, 6], [3, 2, 1], [1, 3, 2], [1, 2, 0], False),
             ([9, 7, 6], [1, 4, 1], [1, 1, 4], [0, 1, 2], True)]
    rng = np.random.default_rng(0)
    for gin, tin, tout, axes, real in cases:
        gout = list(gin)
        if real:
            gout[axes[-1]] = gin[axes[-1]] // 2 + 1
        plan = PE.OraclePlan(gin, gout, axes, tin, tout, real)
        ti = 0
        for bi, b in enumerate(plan.blocks):
            if b.block_type != PL.TRANSPOSE:
                continue
            lays = plan.analyses[0].layouts
            lsrc = PL.LAYOUT_RIGHT if (not lays or lays[ti]) else PL.LAYOUT_LEFT
            ldst = PL.LAYOUT_RIGHT if (not lays or lays[ti + 1]) else PL.LAYOUT_LEFT
            tsrc, tdst = plan.analyses[0].topologies[ti], plan.analyses[0].topologies[ti + 1]
            ti += 1
            nfft = sum(1 for k in range(bi) if plan.blocks[k].block_type == PL.FFT)
            gext = gin if (real and nfft == 0) else gout
            G = rng.standard_normal(gext)
            src = [np.ascontiguousarray(a.transpose(b.in_map)) for a in PE.scatter(G, gext, tsrc, lsrc)]
            for r in range(plan.P):
                assert list(src[r].shape) == plan.analyses[r].blocks[bi].in_extents
            got = plan._exchange(src, bi, True)
            exp = [np.ascontiguousarray(a.transpose(b.out_map)) for a in PE.scatter(G, gext, tdst, ldst)]
            for r in range(plan.P):
                assert np.array_equal(got[r], exp[r]), (gin, tin, tout, bi, r)


# --------------------------------------------------------------------- FFT
@pytest.mark.parametrize("n", [1, 2, 3, 4, 5, 6, 7, 8, 9, 12, 16, 30, 64, 96, 97, 160])
def test_fft_against_direct_dft_and_numpy(n):
    rng = np.random.default_rng(n)
    x = rng.standard_normal((3, n)) + 1j * rng.standard_normal((3, n))
    fast = F.c2c(x, 1, -1)
    direct = F.c2c(x, 1, -1, direct=True)
    assert np.abs(fast - direct).max() <= 1e-13 * max(1.0, np.abs(direct).max())
    assert np.abs(fast - np.fft.fft(x, axis=1)).max() <= 1e-12 * max(1.0, np.abs(direct).max())
    inv = F.c2c(x, 1, +1)
    assert np.abs(inv - np.fft.ifft(x, axis=1) * n).max() <= 1e-12 * max(1.0, np.abs(inv).max())
    r = rng.standard_normal((2, n))
    spec = F.r2c(r, 1)
    assert np.abs(spec - np.fft.rfft(r, axis=1)).max() <= 1e-12 * max(1.0, np.abs(spec).max())
    assert np.abs(F.c2r(spec, n, 1) / n - r).max() <= 1e-13


def test_fft_closed_form_single_mode():
    n = 16
    t = 2 * np.pi * np.arange(n) / n
    X, Y, Z = np.meshgrid(t, t, t, indexing="ij")
    s = F.fftn(np.sin(3 * X + 5 * Y + 7 * Z), [0, 1, 2])
    exp = np.zeros([n] * 3, complex)
    exp[3, 5, 7] = -0.5j * n ** 3
    exp[-3, -5, -7] = 0.5j * n ** 3
    assert np.abs(s - exp).max() <= 1e-12 * n ** 3


def test_fft_parseval_and_linearity():
    rng = np.random.default_rng(4)
    x = rng.standard_normal((6, 10)) + 1j * rng.standard_normal((6, 10))
    y = rng.standard_normal((6, 10)) + 1j * rng.standard_normal((6, 10))
    fx, fy = F.fftn(x, [0, 1]), F.fftn(y, [0, 1])
    assert abs(np.vdot(fx, fx).real / x.size - np.vdot(x, x).real) <= 1e-12 * np.vdot(x, x).real
    assert np.abs(F.fftn(2 * x - 3j * y, [0, 1]) - (2 * fx - 3j * fy)).max() <= 1e-12 * np.abs(fx).max()


# ---------------------------------------------------------------- executor
@pytest.mark.parametrize("P", [2, 3, 4])
@pytest.mark.parametrize("real", [False, True])
def test_oracle_slab_plan_vs_global(P, real):
    rng = np.random.default_rng(P)
    topos = [[1, 1, P], [1, P, 1], [P, 1, 1]]
    gin = [8, 7, 5]
    for tin, tout in itertools.product(topos, topos):
        for axes in itertools.permutations([0, 1, 2]):
            gout = list(gin)
            if real:
                gout[axes[-1]] = gin[axes[-1]] // 2 + 1
            p = PE.OraclePlan(gin, gout, list(axes), tin, tout, real)
            x = rng.standard_normal(gin) if real else rng.standard_normal(gin) + 1j * rng.standard_normal(gin)
            outs = p.forward(PE.scatter(x, gin, tin))
            ref = PE.scatter(F.fftn(x, list(axes), real), gout, tout)
            for o, r in zip(outs, ref):
                assert o.shape == r.shape
                if r.size:
                    assert np.abs(o - r).max() <= 1e-12 * np.abs(F.fftn(x, list(axes), real)).max()
            back = p.backward(outs)
            for b, a in zip(back, PE.scatter(x, gin, tin)):
                if a.size:
                    assert np.abs(b - a).max() <= 1e-12


@pytest.mark.parametrize("px,py", [(2, 2), (2, 3)])
def test_oracle_pencil_plan_vs_global(px, py):
    rng = np.random.default_rng(px * py)
    topos = [[1, px, py], [px, 1, py], [px, py, 1]]
    gin = [8, 7, 6]
    for tin, tout in itertools.product(topos, topos):
        for axes in itertools.permutations([0, 1, 2]):
            for real in (False, True):
                gout = list(gin)
                if real:
                    gout[axes[-1]] = gin[axes[-1]] // 2 + 1
                p = PE.OraclePlan(gin, gout, list(axes), tin, tout, real)
                assert p.name == "PencilPlan"
                x = rng.standard_normal(gin) if real else rng.standard_normal(gin) + 1j * rng.standard_normal(gin)
                outs = p.forward(PE.scatter(x, gin, tin))
                ref = PE.scatter(F.fftn(x, list(axes), real), gout, tout)
                scale = np.abs(F.fftn(x, list(axes), real)).max()
                for o, r in zip(outs, ref):
                    assert np.abs(o - r).max() <= 1e-12 * scale
                back = p.backward(outs)
                for b, a in zip(back, PE.scatter(x, gin, tin)):
                    assert np.abs(b - a).max() <= 1e-12


def test_oracle_plan_errors():
    """Plan.hpp:62-68 and :85-112 throw sites."""
    with pytest.raises(PL.PlannerError):
        PE.OraclePlan([8, 8, 8], [8, 8, 8], [0, 1, 2], [2, 1, 1], [1, 1, 4])
    with pytest.raises(PL.PlannerError):  # brick
        PE.OraclePlan([8, 8, 8], [8, 8, 8], [0, 1, 2], [2, 2, 2], [2, 2, 2])
    with pytest.raises(PL.PlannerError):  # R2C extents
        PE.OraclePlan([8, 8, 8], [8, 8, 8], [0, 1, 2], [2, 1, 1], [1, 2, 1], real_input=True)
