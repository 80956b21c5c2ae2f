"""Planner known answers of the reference's own unit tests, against the
oracle planner and the product (C ABI through dfft.py):

* Test_MPI_Extents.cpp:163-317 -- the rank_to_coord table (LayoutRight and
  LayoutLeft topologies) and its out-of-range throws;
* Test_MPI_Extents.cpp:35-160, 337-900 -- global extents, local starts,
  local extents under every map, and compute_local_extents_and_starts, each
  expected value built the way the reference test builds it (the
  ``distribute_extents`` rule per rank coordinate) for the process grids the
  test is run with;
* Test_BlockAnalyses.cpp:81-208 -- propose_fft_block and
  propose_transpose_block (both view layouts x both topology layouts, FFT
  axes {2} or none, last block or not).
"""

import itertools

import pytest

import dfft
from oracle import planner as PL

LAYOUTS = [PL.LAYOUT_RIGHT, PL.LAYOUT_LEFT]


def _topo(values, layout):
    return dfft.Topology(values, dfft.LayoutLeft if layout == PL.LAYOUT_LEFT else dfft.LayoutRight)


# --------------------------------------------------------------------------
# rank_to_coord (Test_MPI_Extents.cpp:163-317)
# --------------------------------------------------------------------------
RANK_TO_COORD = {
    # topology: {rank: coord} for LayoutRight; LayoutLeft differs where listed
    (2,): {0: (0,), 1: (1,)},
    (1, 2): {0: (0, 0), 1: (0, 1)},
    (4, 2): {0: (0, 0), 1: (0, 1), 2: (1, 0), 3: (1, 1), 4: (2, 0), 5: (2, 1), 6: (3, 0), 7: (3, 1)},
    (1, 2, 1): {0: (0, 0, 0), 1: (0, 1, 0)},
    (4, 2, 1): {0: (0, 0, 0), 1: (0, 1, 0), 2: (1, 0, 0), 3: (1, 1, 0), 4: (2, 0, 0), 5: (2, 1, 0),
                6: (3, 0, 0), 7: (3, 1, 0)},
    (1, 4, 1, 1): {0: (0, 0, 0, 0), 1: (0, 1, 0, 0), 2: (0, 2, 0, 0), 3: (0, 3, 0, 0)},
    (1, 1, 4, 2): {0: (0, 0, 0, 0), 1: (0, 0, 0, 1), 2: (0, 0, 1, 0), 3: (0, 0, 1, 1), 4: (0, 0, 2, 0),
                   5: (0, 0, 2, 1), 6: (0, 0, 3, 0), 7: (0, 0, 3, 1)},
}
RANK_TO_COORD_LEFT = {
    (4, 2): {0: (0, 0), 1: (1, 0), 2: (2, 0), 3: (3, 0), 4: (0, 1), 5: (1, 1), 6: (2, 1), 7: (3, 1)},
    (4, 2, 1): {0: (0, 0, 0), 1: (1, 0, 0), 2: (2, 0, 0), 3: (3, 0, 0), 4: (0, 1, 0), 5: (1, 1, 0),
                6: (2, 1, 0), 7: (3, 1, 0)},
    (1, 1, 4, 2): {0: (0, 0, 0, 0), 1: (0, 0, 1, 0), 2: (0, 0, 2, 0), 3: (0, 0, 3, 0), 4: (0, 0, 0, 1),
                   5: (0, 0, 1, 1), 6: (0, 0, 2, 1), 7: (0, 0, 3, 1)},
}


@pytest.mark.parametrize("layout", LAYOUTS)
def test_rank_to_coord_table(layout):
    for topo, table in RANK_TO_COORD.items():
        if layout == PL.LAYOUT_LEFT and topo in RANK_TO_COORD_LEFT:
            table = RANK_TO_COORD_LEFT[topo]
        for r, c in table.items():
            assert tuple(PL.rank_to_coord(list(topo), r, layout)) == c, (topo, r)
            assert tuple(dfft.rank_to_coord(_topo(topo, layout), r)) == c, (topo, r)
    for topo, bad in (((2,), 2), ((1, 2), 2), ((1, 2, 1), 2), ((1, 4, 1, 1), 4)):
        with pytest.raises(PL.PlannerError):
            PL.rank_to_coord(list(topo), bad, layout)
        with pytest.raises(dfft.DFFTError):
            dfft.rank_to_coord(_topo(topo, layout), bad)


# --------------------------------------------------------------------------
# extents (Test_MPI_Extents.cpp:35-160, 337-900)
# --------------------------------------------------------------------------
def _distribute(n, r, t):
    q, rem = divmod(n, t)
    return q + 1 if r < rem else q


GRIDS_2D = [1, 2, 3, 4]
GRIDS_3D = [(1, 1), (1, 2), (2, 1), (2, 2), (1, 3), (3, 1), (2, 3), (3, 2)]


@pytest.mark.parametrize("layout", LAYOUTS)
@pytest.mark.parametrize("nprocs", GRIDS_2D)
def test_extents_2d(layout, nprocs):
    gn0, gn1 = 19, 32
    t0, t1 = [1, nprocs], [nprocs, 1]
    for rank in range(nprocs):
        e_t0 = [gn0, _distribute(gn1, rank, nprocs)]
        e_t1 = [_distribute(gn0, rank, nprocs), gn1]
        st_t0 = [0, sum(_distribute(gn1, r, nprocs) for r in range(rank))]
        st_t1 = [sum(_distribute(gn0, r, nprocs) for r in range(rank)), 0]
        for t, e, st in ((t0, e_t0, st_t0), (t1, e_t1, st_t1)):
            for m in ([0, 1], [1, 0]):
                assert PL.compute_local_extents([gn0, gn1], t, m, rank) == [e[a] for a in m]
            assert PL.compute_local_extents_and_starts([gn0, gn1], t, rank, layout) == (e, st)
            assert dfft.compute_local_extents_and_starts([gn0, gn1], _topo(t, layout), rank) == (e, st)
        # the reference's "random" per-rank pattern: starts are prefix sums
        rnd_t0 = [[7, 10 + r] for r in range(nprocs)]
        rnd_t1 = [[5 + 3 * r, 12] for r in range(nprocs)]
        assert PL.compute_local_starts(rnd_t0, t0, rank, layout) == [0, sum(10 + r for r in range(rank))]
        assert PL.compute_local_starts(rnd_t1, t1, rank, layout) == [sum(5 + 3 * r for r in range(rank)), 0]
    # global extents from every rank's local extents (Allgather + sums)
    for t, fn in ((t0, lambda r: [gn0, _distribute(gn1, r, nprocs)]),
                  (t1, lambda r: [_distribute(gn0, r, nprocs), gn1])):
        locs = [fn(r) for r in range(nprocs)]
        for rank in range(nprocs):
            assert PL.compute_global_extents(locs, t, rank, layout) == [gn0, gn1]


@pytest.mark.parametrize("layout", LAYOUTS)
@pytest.mark.parametrize("npx,npy", GRIDS_3D)
def test_extents_3d(layout, npx, npy):
    gn = [8, 7, 5]
    topos = {"t0": [1, npx, npy], "t1": [npx, 1, npy], "t2": [npx, npy, 1]}
    for rank in range(npx * npy):
        rx, ry = (rank % npx, rank // npx) if layout == PL.LAYOUT_LEFT else (rank // npy, rank % npy)
        exp = {
            "t0": [gn[0], _distribute(gn[1], rx, npx), _distribute(gn[2], ry, npy)],
            "t1": [_distribute(gn[0], rx, npx), gn[1], _distribute(gn[2], ry, npy)],
            "t2": [_distribute(gn[0], rx, npx), _distribute(gn[1], ry, npy), gn[2]],
        }
        sx = lambda n: sum(_distribute(n, r, npx) for r in range(rx))  # noqa: E731
        sy = lambda n: sum(_distribute(n, r, npy) for r in range(ry))  # noqa: E731
        starts = {"t0": [0, sx(gn[1]), sy(gn[2])], "t1": [sx(gn[0]), 0, sy(gn[2])], "t2": [sx(gn[0]), sy(gn[1]), 0]}
        for k, t in topos.items():
            assert PL.compute_local_extents_and_starts(gn, t, rank, layout) == (exp[k], starts[k]), (k, rank)
            assert dfft.compute_local_extents_and_starts(gn, _topo(t, layout), rank) == (exp[k], starts[k])
        # local extents under every map (LayoutRight topologies, plus the
        # LayoutLeft topology {npy, npx, 1} of the reference test)
        rxr, ryr = rank // npy, rank % npy
        ref_r = {
            "t0": [gn[0], _distribute(gn[1], rxr, npx), _distribute(gn[2], ryr, npy)],
            "t1": [_distribute(gn[0], rxr, npx), gn[1], _distribute(gn[2], ryr, npy)],
            "t2": [_distribute(gn[0], rxr, npx), _distribute(gn[1], ryr, npy), gn[2]],
        }
        t3 = [npy, npx, 1]
        e3 = [_distribute(gn[0], ryr, npy), _distribute(gn[1], rxr, npx), gn[2]]
        for m in itertools.permutations(range(3)):
            for k, t in topos.items():
                assert PL.compute_local_extents(gn, t, list(m), rank) == [ref_r[k][a] for a in m]
            assert PL.compute_local_extents(gn, t3, list(m), rank, PL.LAYOUT_LEFT) == [e3[a] for a in m]
    # the reference's "random" per-rank pattern (extents grow with the rank's
    # coordinates rx, ry): starts are prefix sums along each coordinate
    def coords(rank):
        return (rank % npx, rank // npx) if layout == PL.LAYOUT_LEFT else (rank // npy, rank % npy)

    pattern = {"t0": lambda x, y: [4, 3 + x, 2 + y], "t1": lambda x, y: [4 + x, 4, 2 + y],
               "t2": lambda x, y: [4 + x, 3 + y, 2]}
    split = {"t0": (1, 2), "t1": (0, 2), "t2": (0, 1)}  # axes distributed over npx, npy
    for k, t in topos.items():
        locs = [pattern[k](*coords(r)) for r in range(npx * npy)]
        ax, ay = split[k]
        for rank in range(npx * npy):
            x, y = coords(rank)
            want = [0, 0, 0]
            want[ax] = sum(pattern[k](r, 0)[ax] for r in range(x))
            want[ay] = sum(pattern[k](0, r)[ay] for r in range(y))
            assert PL.compute_local_starts(locs, t, rank, layout) == want, (k, rank)
            glob = [locs[0][a] for a in range(3)]
            glob[ax] = sum(pattern[k](r, 0)[ax] for r in range(npx))
            glob[ay] = sum(pattern[k](0, r)[ay] for r in range(npy))
            assert PL.compute_global_extents(locs, t, rank, layout) == glob, (k, rank)


# --------------------------------------------------------------------------
# block proposals (Test_BlockAnalyses.cpp:81-208)
# --------------------------------------------------------------------------
def test_propose_fft_block():
    b, size = PL.propose_fft_block([0, 1, 2], [16, 16, 16], [9, 16, 16], 3, 0)
    assert b.block_type == PL.FFT and b.fft_dim == 3 and b.block_idx == 0
    assert b.in_extents == [16, 16, 16] and b.out_extents == [9, 16, 16]
    assert b.in_map == [0, 1, 2] and b.out_map == [0, 1, 2]
    assert size == 9 * 16 * 16 * 2


@pytest.mark.parametrize("layout", LAYOUTS)
@pytest.mark.parametrize("topo_layout", LAYOUTS)
@pytest.mark.parametrize("axes", [[2], []])
@pytest.mark.parametrize("is_last", [True, False])
def test_propose_transpose_block(layout, topo_layout, axes, is_last):
    m, in_ext, tin, tout, gext = [2, 0, 1], [16, 16, 16], [4, 4, 1], [4, 1, 4], [64, 64, 16]
    is_lr = topo_layout == PL.LAYOUT_RIGHT
    for rank in range(16):
        b, size = PL.propose_transpose_block(layout, m, tin, tout, in_ext, gext, axes, rank, is_lr, is_last,
                                             1, 2, 1, 3, 2)
        assert b.block_type == PL.TRANSPOSE
        assert (b.in_map, b.in_axis, b.out_axis, b.comm_axis) == (m, 1, 2, 1)
        assert (b.in_topology, b.out_topology, b.in_extents, b.block_idx) == (tin, tout, in_ext, 3)
        if is_last:
            want_map = [0, 1, 2]
        else:
            want_map = PL.permute_map_by_axes(layout, m, axes if axes else [2])
        assert b.out_map == want_map
        assert b.out_extents == PL.compute_local_extents(gext, tout, want_map, rank, topo_layout)
        assert b.buffer_extents == PL.compute_buffer_extents(layout, gext, tin, tout)
        vol = lambda e: e[0] * e[1] * e[2] * (e[3] if len(e) > 3 else 1)  # noqa: E731
        assert size == 2 * max(vol(in_ext), vol(b.buffer_extents), vol(b.out_extents))
    # literal values for rank 0 (hand-evaluated from the rules above)
    b, size = PL.propose_transpose_block(layout, m, tin, tout, in_ext, gext, axes, 0, is_lr, is_last, 1, 2, 1, 3, 2)
    assert size == 8192
    if layout == PL.LAYOUT_RIGHT:
        assert b.buffer_extents == [4, 16, 16, 4]
        assert b.out_map == [0, 1, 2]  # LayoutRight: axis 2 last; also the last block's identity
        assert b.out_extents == [16, 64, 4]
    else:
        assert b.buffer_extents == [16, 16, 4, 4]
        assert b.out_map == ([0, 1, 2] if is_last else [2, 0, 1])
        assert b.out_extents == ([16, 64, 4] if is_last else [4, 16, 64])
