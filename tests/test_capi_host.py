"""CPU tests of libdfft.so: it loads, exports every symbol of include/dfft.h,
and its host planner matches the reference's golden tables and the oracle
planner (no device calls)."""

import ctypes
import itertools
import re
from pathlib import Path

import pytest

import dfft
from oracle import planner as PL
from topology_cases import pencil_cases, slab_cases

ROOT = Path(__file__).resolve().parents[1]


def test_library_exports_every_declared_symbol(lib):
    header = (ROOT / "include" / "dfft.h").read_text()
    declared = set(re.findall(r"\b(dfft_[a-z0-9_]+)\s*\(", header))
    assert declared == set(dfft.EXPORTED)
    for name in declared:
        assert hasattr(lib, name), name


def test_version_and_error_string(lib):
    major, minor = ctypes.c_int(), ctypes.c_int()
    assert lib.dfft_version(ctypes.byref(major), ctypes.byref(minor)) == 0
    assert (major.value, minor.value) >= (0, 1)
    assert lib.dfft_version(None, None) == dfft.ERROR_INVALID_ARGUMENT
    assert b"null" in lib.dfft_last_error()


@pytest.mark.parametrize("nprocs", [2, 3, 4, 5, 6])
def test_product_slab_topologies_golden(nprocs):
    for fn, tin, tout, axes, ref in slab_cases(nprocs):
        assert dfft.get_all_slab_topologies(tin, tout, axes) == ref, (fn, tin, tout, axes)


@pytest.mark.parametrize("nprocs", [2, 3, 4, 5, 6])
def test_product_pencil_topologies_golden(nprocs):
    for fn, tin, lin, tout, lout, axes, rax, rlay, ref in pencil_cases(nprocs):
        t_in = dfft.Topology(tin, dfft.LayoutLeft if lin == "left" else dfft.LayoutRight)
        t_out = dfft.Topology(tout, dfft.LayoutLeft if lout == "left" else dfft.LayoutRight)
        assert dfft.get_all_pencil_topologies(t_in, t_out, axes) == (ref, rax, rlay), (fn, tin, tout, axes)


def test_product_slab_topologies_throw():
    with pytest.raises(dfft.DFFTError) as e:
        dfft.get_all_slab_topologies([1, 1, 1], [1, 1, 1], [0, 1, 2])
    assert e.value.code == dfft.ERROR_INVALID_TOPOLOGY


def _oracle_blocks(gin, gout, axes, tin, tout, rank, real, layout=PL.LAYOUT_RIGHT):
    kind = PL.get_common_topology_type(tin, tout)
    local = PL.compute_local_extents(gin, tin, list(range(len(gin))), rank)
    ba = PL.block_analyses(kind, real, layout, local, gin, gout, tin, tout, axes, rank)
    out = []
    for b in ba.blocks:
        out.append({"in_topology": b.in_topology, "out_topology": b.out_topology,
                    "in_extents": b.in_extents, "out_extents": b.out_extents,
                    "buffer_extents": b.buffer_extents, "in_map": b.in_map, "out_map": b.out_map,
                    "in_axis": b.in_axis, "out_axis": b.out_axis, "fft_dim": b.fft_dim,
                    "comm_axis": b.comm_axis, "block_type": b.block_type, "block_idx": b.block_idx})
    return out, ba.max_buffer_size


@pytest.mark.parametrize("layout", [PL.LAYOUT_RIGHT, PL.LAYOUT_LEFT])
def test_product_block_analyses_match_oracle(layout):
    """Every rank's BlockInfo list of the C++ planner equals the oracle's
    (slab P=2..4 and pencil 2x2, 2x3 with every pair and axis order, R2C/C2C)
    for LayoutRight and LayoutLeft views."""
    lay = dfft.LayoutLeft if layout == PL.LAYOUT_LEFT else dfft.LayoutRight
    n = 0
    for gin in ([8, 7, 5], [9, 6, 7]):
        for P in (2, 3, 4):
            topos = [[1, 1, P], [1, P, 1], [P, 1, 1]]
            for tin, tout in itertools.product(topos, topos):
                for axes in itertools.permutations([0, 1, 2]):
                    for real in (False, True):
                        gout = list(gin)
                        if real:
                            gout[axes[-1]] = gin[axes[-1]] // 2 + 1
                        for r in range(P):
                            exp = _oracle_blocks(gin, gout, list(axes), tin, tout, r, real, layout)
                            got = dfft.block_analyses(gin, gout, list(axes), tin, tout, r, real, lay)
                            assert got == exp, (gin, tin, tout, axes, real, r)
                            n += 1
    for px, py in ((2, 2), (2, 3), (3, 2)):
        topos = [[1, px, py], [px, 1, py], [px, py, 1]]
        gin = [8, 7, 6]
        for tin, tout in itertools.product(topos, topos):
            for axes in itertools.permutations([0, 1, 2]):
                for real in (False, True):
                    gout = list(gin)
                    if real:
                        gout[axes[-1]] = gin[axes[-1]] // 2 + 1
                    for r in range(px * py):
                        exp = _oracle_blocks(gin, gout, list(axes), tin, tout, r, real, layout)
                        got = dfft.block_analyses(gin, gout, list(axes), tin, tout, r, real, lay)
                        assert got == exp, (gin, tin, tout, axes, real, r)
                        n += 1
    assert n > 3000


def test_python_extent_helpers():
    assert dfft.compute_local_extents_and_starts([10, 4, 4], [3, 1, 1], 1) == ([3, 4, 4], [4, 0, 0])
    assert dfft.rank_to_coord(dfft.Topology([2, 2], dfft.LayoutLeft), 1) == [1, 0]
    assert dfft.rank_to_coord([2, 2], 1) == [0, 1]
