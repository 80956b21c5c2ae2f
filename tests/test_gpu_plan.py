"""GPU parity of the distributed plans against the oracle (through the C ABI).

Mirrors the reference's plan tests (Test_SlabPlan.cpp, Test_PencilPlan.cpp):
every virtual rank's output must equal the global transform sliced to that
rank's block of the output topology, and backward(forward(x)) == x.  The
reference checks float at rtol 1e-5; here FP64 is held to rel-L2 <= 1e-12 and
FP32 to rel-L2 <= 1e-5 against the FP64 oracle.
"""

import itertools

import numpy as np
import pytest

from helpers import gather_rel_l2, reference_spectrum, run_gpu_plan, seeded_input

pytestmark = pytest.mark.gpu

TOL64 = 1e-12
TOL32 = 1e-5


def _check(dev, gin, axes, tin, tout, real, fp64=True, norm="backward", exec_mode=0, seed=0, **kw):
    x = seeded_input(gin, real, seed)
    outs, backs, ins, plan = run_gpu_plan(dev, x, gin, axes, tin, tout, real, fp64, norm, exec_mode, **kw)
    ref = reference_spectrum(x, gin, axes, tout, real, norm, kw.get("tout_layout"))
    tol = TOL64 if fp64 else TOL32
    e_fwd = gather_rel_l2(outs, ref)
    e_rt = gather_rel_l2(backs, ins)
    assert e_fwd <= tol, f"forward rel-L2 {e_fwd:.3e} {gin} {axes} {tin}->{tout} real={real}"
    assert e_rt <= tol, f"round trip rel-L2 {e_rt:.3e} {gin} {axes} {tin}->{tout} real={real}"
    return plan


@pytest.mark.parametrize("P", [1, 2, 3, 4])
@pytest.mark.parametrize("real", [False, True])
def test_slab_3d_all_pairs_all_orders(dev, P, real):
    """Test_SlabPlan.cpp:2084-2290 -- (8,7,5), every slab pair, every axis order."""
    topos = [[1, 1, P], [1, P, 1], [P, 1, 1]]
    for tin, tout in itertools.product(topos, topos):
        for axes in itertools.permutations([0, 1, 2]):
            plan = _check(dev, [8, 7, 5], list(axes), tin, tout, real, seed=P)
            assert plan.name() == ("SharedPlan" if P == 1 else "SlabPlan")


@pytest.mark.parametrize("real", [False, True])
@pytest.mark.parametrize("fp64", [True, False])
def test_pencil_2x2_and_2x3(dev, real, fp64):
    """Test_PencilPlan.cpp:3653-... (8,7,6) style, pencil pairs, every axis order."""
    for px, py in [(2, 2), (2, 3)]:
        topos = [[1, px, py], [px, 1, py], [px, py, 1]]
        for tin, tout in itertools.product(topos, topos):
            for axes in itertools.permutations([0, 1, 2]):
                plan = _check(dev, [8, 7, 6], list(axes), tin, tout, real, fp64, seed=px + py)
                assert plan.name() == "PencilPlan"


@pytest.mark.parametrize("real", [False, True])
def test_slab_2d_and_1d_axes(dev, real):
    """FFT over 1 or 2 axes of a 3-D view, and 2-D views (Test_SlabPlan.cpp:91, :333)."""
    for P in [2, 3]:
        topos = [[1, 1, P], [1, P, 1], [P, 1, 1]]
        for tin, tout in itertools.product(topos, topos):
            for axes in [[0], [1], [2], [0, 1], [1, 2], [2, 0]]:
                _check(dev, [8, 7, 9], axes, tin, tout, real, seed=7)
        for tin, tout in itertools.product([[1, P], [P, 1]], repeat=2):
            for axes in [[0, 1], [1, 0], [0], [1]]:
                _check(dev, [8, 7], axes, tin, tout, real, seed=8)


@pytest.mark.parametrize("norm", ["backward", "forward", "ortho", "none"])
def test_normalizations(dev, norm):
    x = seeded_input([12, 10, 8], False, 3)
    outs, backs, ins, _ = run_gpu_plan(dev, x, [12, 10, 8], [0, 1, 2], [2, 1, 1], [1, 2, 1], False, True, norm)
    ref = reference_spectrum(x, [12, 10, 8], [0, 1, 2], [1, 2, 1], False, norm)
    assert gather_rel_l2(outs, ref) <= TOL64
    scale = {"backward": 1.0, "forward": 1.0, "ortho": 1.0, "none": 12 * 10 * 8}[norm]
    assert gather_rel_l2([b / scale for b in backs], ins) <= TOL64


def test_4d_batched_views(dev):
    """4-D views with a leading batch axis (Test_SlabPlan.cpp:2738, NS batched example)."""
    for real in [False, True]:
        _check(dev, [3, 8, 7, 6], [1, 2, 3], [1, 2, 1, 1], [1, 1, 2, 1], real, seed=11)
        _check(dev, [3, 8, 7, 6], [1, 2, 3], [1, 2, 2, 1], [1, 1, 2, 2], real, seed=12)


def test_config_a_64cubed_c2c_slab(dev):
    """64^3 complex128, slab (2,1,1) -> (1,2,1), U[-1,1] seed 1, fwd then bwd."""
    _check(dev, [64, 64, 64], [0, 1, 2], [2, 1, 1], [1, 2, 1], False, seed=1)


def test_config_b_256cubed_r2c_pencil(dev):
    """256^3 float64 N(0,1) seed 2, pencil (2,2,1) -> (1,2,2): rel-L2 <= 1e-12."""
    gin = [256, 256, 256]
    rng = np.random.default_rng(2)
    x = rng.standard_normal(gin)
    outs, backs, ins, plan = run_gpu_plan(dev, x, gin, [0, 1, 2], [2, 2, 1], [1, 2, 2], True)
    ref = reference_spectrum(x, gin, [0, 1, 2], [1, 2, 2], True)
    assert plan.name() == "PencilPlan"
    assert gather_rel_l2(outs, ref) <= TOL64
    assert gather_rel_l2(backs, ins) <= TOL64


def test_closed_form_single_mode(dev):
    """sin(3x+5y+7z) on [0,2pi)^3: forward spectrum -i N^3/2 at (3,5,7), +i N^3/2 at -k."""
    n = 32
    t = 2 * np.pi * np.arange(n) / n
    X, Y, Z = np.meshgrid(t, t, t, indexing="ij")
    f = np.sin(3 * X + 5 * Y + 7 * Z)
    for fp64 in [True, False]:
        outs, _, _, plan = run_gpu_plan(dev, f, [n] * 3, [0, 1, 2], [4, 1, 1], [1, 4, 1], False, fp64)
        from oracle import plan_exec as PE
        spec = PE.gather(outs, [n] * 3, [1, 4, 1], np.complex128)
        expect = np.zeros([n] * 3, complex)
        expect[3, 5, 7] = -0.5j * n ** 3
        expect[-3, -5, -7] = 0.5j * n ** 3
        tol = 1e-12 if fp64 else 1e-5
        assert np.abs(spec - expect).max() / (0.5 * n ** 3) <= tol


def test_blockwalk_matches_fused(dev):
    """The reference block sequence (FFT blocks + exchanges) gives the same result."""
    for gin, tin, tout, axes, real in [
        ([8, 7, 5], [2, 1, 1], [1, 2, 1], [0, 1, 2], False),
        ([8, 7, 5], [1, 1, 3], [3, 1, 1], [2, 0, 1], True),
        ([8, 7, 6], [1, 2, 2], [2, 1, 2], [0, 1, 2], True),
        ([8, 7, 6], [2, 2, 1], [1, 2, 2], [1, 0, 2], False),
        ([8, 7, 6], [1, 2, 3], [2, 3, 1], [0, 1, 2], True),
    ]:
        _check(dev, gin, axes, tin, tout, real, exec_mode=1, seed=5)


def test_layout_left_topologies(dev):
    """Topology<..., LayoutLeft> (Test_PencilPlan.cpp:102): rank ordering per layout."""
    _check(dev, [8, 7, 6], [0, 1, 2], [1, 2, 3], [3, 2, 1], False, tin_layout="right", tout_layout="left")
    _check(dev, [8, 7, 6], [0, 1, 2], [2, 3, 1], [1, 2, 3], True, tin_layout="left", tout_layout="right")


@pytest.mark.parametrize("exec_mode", [0, 1])
@pytest.mark.parametrize("fp64", [True, False])
def test_layout_left_views(dev, exec_mode, fp64):
    """LayoutLeft (column-major) views, the {float, double} x LayoutLeft cases
    of Test_SlabPlan.cpp:13-16 / Test_PencilPlan.cpp: fused execution and the
    block walk (LayoutLeft block analyses: maps with the FFT axes first, (e..., P)
    buffers), slab and pencil, C2C and R2C, against the oracle."""
    cases = [
        ([8, 7, 5], [2, 1, 1], [1, 2, 1], [0, 1, 2]),
        ([8, 7, 5], [1, 1, 3], [3, 1, 1], [2, 0, 1]),
        ([9, 8, 6], [1, 2, 1], [2, 1, 1], [1, 2]),
        ([8, 7, 6], [2, 2, 1], [1, 2, 2], [0, 1, 2]),
        ([8, 7, 6], [1, 2, 3], [2, 3, 1], [0, 2, 1]),
        ([3, 8, 7, 6], [1, 2, 1, 1], [1, 1, 2, 1], [1, 2, 3]),
    ]
    for gin, tin, tout, axes in cases:
        for real in (False, True):
            plan = _check(dev, gin, axes, tin, tout, real, fp64, exec_mode=exec_mode, seed=13, view_layout="left")
            assert plan.name() in ("SlabPlan", "PencilPlan")


def test_layout_left_views_with_left_topologies(dev):
    _check(dev, [8, 7, 6], [0, 1, 2], [1, 2, 3], [3, 2, 1], True, tin_layout="right", tout_layout="left",
           view_layout="left")
    _check(dev, [8, 7, 6], [0, 1, 2], [2, 3, 1], [1, 2, 3], False, tin_layout="left", tout_layout="right",
           view_layout="left", exec_mode=1)


def test_layout_left_views_refuse_row_major_arrays(dev):
    import torch
    import dfft
    ins = [torch.zeros(4, 8, 5, dtype=torch.complex128, device=dev) for _ in range(2)]
    outs = [torch.zeros(8, 4, 5, dtype=torch.complex128, device=dev).permute(2, 1, 0).contiguous().permute(2, 1, 0)
            for _ in range(2)]
    plan = dfft.Plan(ins, outs, [0, 1, 2], [2, 1, 1], [1, 2, 1], layout=dfft.LayoutLeft)
    with pytest.raises(dfft.DFFTError):
        dfft.execute(plan, ins, outs, dfft.Direction.forward)  # inputs are row-major


@pytest.mark.parametrize("exec_mode", [0, 1, 2])
def test_ranks_with_empty_blocks(dev, exec_mode):
    """Extents smaller than a topology factor leave some ranks with empty
    local views (compute_local_extents gives 0; MPI_Extents.hpp:195-235): the
    other ranks' results are unchanged, empty views are never touched."""
    for gin, tin, tout, real in [
        ([3, 8, 8], [4, 1, 1], [1, 4, 1], False),
        ([8, 3, 6], [4, 1, 1], [1, 4, 1], True),
        ([2, 2, 9], [1, 1, 4], [4, 1, 1], False),
        ([5, 6, 3], [2, 3, 1], [1, 2, 3], False),
    ]:
        for fp64 in (True, False):
            _check(dev, gin, [0, 1, 2], tin, tout, real, fp64, exec_mode=exec_mode, seed=19)


def test_fp32_two_lines_per_lane_from_odd_rows(dev):
    """FP32 C2C passes take two adjacent lines per lane also when their input
    has odd rows (the R2C spectrum's 257 elements per row, read under the
    scratch's padded enumeration): the backward x pass here.  Same results
    with two lines per lane on and off, against the oracle."""
    gin, tin, tout = [256, 8, 512], [4, 1, 1], [1, 4, 1]

    def off(plan, ins, outs):
        for k in range(3):
            plan.set_option(f"fwd.{k}.dual", 0)
            plan.set_option(f"bwd.{k}.dual", 0)

    for configure in (None, off):
        _check(dev, gin, [0, 1, 2], tin, tout, True, fp64=False, seed=17, configure=configure)


def test_chunked_and_tuned_launch_shapes(dev):
    """Launch-shape options, chunked first passes and the autotuner change the
    schedule, never the result."""
    def chunk(c):
        return lambda plan, ins, outs: plan.set_option("chunk", c)

    def shapes(plan, ins, outs):
        for d in ("fwd", "bwd"):
            for i in range(3):
                plan.set_option(f"{d}.{i}.vpt", 16)
                plan.set_option(f"{d}.{i}.lpb", 2)
                plan.set_option(f"{d}.{i}.lf", 1)

    def tune(plan, ins, outs):
        plan.autotune(ins, outs)
        assert plan.get_option("fwd.chunk") in (0, 4, 8, 16, 32, 64)

    for cfg in (chunk(4), chunk(7), shapes, tune):
        _check(dev, [32, 24, 64], [0, 1, 2], [2, 1, 1], [1, 2, 1], False, seed=21, configure=cfg)
        _check(dev, [32, 32, 64], [0, 1, 2], [2, 2, 1], [1, 2, 2], True, seed=22, configure=cfg)
        _check(dev, [16, 64, 32], [2, 0, 1], [1, 1, 2], [2, 1, 1], True, fp64=False, seed=23, configure=cfg)


def test_nonsmooth_lengths_generic_kernel(dev):
    """Prime and mixed lengths go through the generic kernel."""
    for gin in ([11, 13, 17], [9, 25, 49], [96, 10, 160]):
        for real in [False, True]:
            _check(dev, gin, [0, 1, 2], [2, 1, 1], [1, 2, 1], real, seed=9)


@pytest.mark.parametrize("real", [False, True])
@pytest.mark.parametrize("fp64", [True, False])
def test_mixed_radix_fast_lengths(dev, real, fp64):
    """Compile-time kernels for 96 = 4*4*2*3 and 160 = 4*4*2*5 (several butterflies per
    thread in the first stage) along the contiguous and the strided axes."""
    for gin, axes in (([6, 5, 96], [2]), ([4, 3, 160], [2]), ([96, 6, 5], [0]), ([160, 4, 3], [0]),
                      ([8, 6, 160], [0, 1, 2]), ([96, 4, 6], [1, 2, 0])):
        _check(dev, gin, axes, [1, 1, 1], [1, 1, 1], real, fp64, seed=5)
        _check(dev, gin, axes, [1, 2, 1], [1, 2, 1], real, fp64, seed=6)


@pytest.mark.parametrize("pad", [0, 1])
@pytest.mark.parametrize("real", [False, True])
@pytest.mark.parametrize("fp64", [True, False])
def test_padded_scratch_rows(dev, pad, real, fp64):
    """Scratch rows padded to 128-byte multiples (forced on) or not at all, for odd row
    lengths: slab pairs in every axis order, a pencil pair, and chunked first passes."""
    cfg = lambda plan, ins, outs: plan.set_option("pad", pad)
    for tin, tout in itertools.product([[1, 1, 2], [1, 2, 1], [2, 1, 1]], repeat=2):
        for axes in itertools.permutations([0, 1, 2]):
            _check(dev, [8, 7, 5], list(axes), tin, tout, real, fp64, seed=3, configure=cfg)
    _check(dev, [9, 8, 13], [0, 1, 2], [2, 2, 1], [1, 2, 2], real, fp64, seed=4, configure=cfg)

    def chunked(plan, ins, outs):
        plan.set_option("pad", pad)
        plan.set_option("chunk", 3)
    _check(dev, [12, 10, 9], [0, 1, 2], [2, 1, 1], [1, 2, 1], real, fp64, seed=5, configure=chunked)


@pytest.mark.parametrize("real", [False, True])
@pytest.mark.parametrize("vpt", [4, 8, 16])
@pytest.mark.parametrize("fp64", [True, False])
def test_split_lds_exchanges(dev, real, vpt, fp64):
    """Radix-stage exchanges through LDS in real / imaginary halves (the default
    for FP64 lines and FP32 lines of 1024 points or more) against complex
    exchanges: both match the oracle, and since only
    data movement differs they agree to rounding (the compiler may contract
    multiply-adds differently in the two kernels, so not bit for bit).  Every
    fast length along contiguous and strided axes, paired R2C / C2R lines
    included."""
    shapes = [([16, 96, 160], 2), ([128, 64, 256], 1), ([512, 8, 32], 4), ([32, 12, 512], 1),
              ([1024, 4, 6], 1), ([4, 6, 1024], 2)]
    if not fp64:  # FP32 has the split form for lines of 1024 points and more only
        shapes = shapes[-2:]
    tol = TOL64 if fp64 else TOL32
    for gin, lpb in shapes:
        results = []
        for split in (1, 0):
            def cfg(plan, ins, outs, split=split):
                for d in ("fwd", "bwd"):
                    for i in range(plan.num_passes(0 if d == "fwd" else 1)):
                        plan.set_option(f"{d}.{i}.split", split)
                        plan.set_option(f"{d}.{i}.vpt", vpt)
                        plan.set_option(f"{d}.{i}.lpb", lpb)
            x = seeded_input(gin, real, 11)
            outs, backs, ins, _ = run_gpu_plan(dev, x, gin, [0, 1, 2], [2, 1, 1], [1, 2, 1], real, fp64,
                                               configure=cfg)
            ref = reference_spectrum(x, gin, [0, 1, 2], [1, 2, 1], real)
            assert gather_rel_l2(outs, ref) <= tol, f"split={split} {gin} vpt={vpt}"
            assert gather_rel_l2(backs, ins) <= tol, f"split={split} {gin} vpt={vpt}"
            results.append((outs, backs))
        for k in range(2):
            e = gather_rel_l2(results[0][k], results[1][k])
            assert e <= (1e-14 if fp64 else 1e-6), f"split and complex exchanges differ by {e:.2e} for {gin} vpt={vpt}"


@pytest.mark.parametrize("fp64", [True, False])
@pytest.mark.parametrize("pad", [0, 1])
def test_lanes_across_lines_blocks(dev, fp64, pad):
    """Lanes-across-lines blocks of 1..16 lines on every pass (blocks that end
    inside a row of the fastest batch axis, cross rank blocks, or hold padding
    lines of padded scratch rows) against the oracle, in slab and pencil plans."""
    for lpb in (1, 2, 4, 8, 16):
        def cfg(plan, ins, outs, lpb=lpb):
            plan.set_option("pad", pad)
            for d in ("fwd", "bwd"):
                for i in range(plan.num_passes(0 if d == "fwd" else 1)):
                    plan.set_option(f"{d}.{i}.lf", 1)
                    plan.set_option(f"{d}.{i}.lpb", lpb)
        _check(dev, [16, 24, 40], [0, 1, 2], [2, 1, 1], [1, 2, 1], False, fp64, seed=21, configure=cfg)
        _check(dev, [32, 12, 20], [0, 1, 2], [1, 1, 2], [2, 1, 1], True, fp64, seed=22, configure=cfg)
        _check(dev, [16, 32, 9], [0, 1, 2], [2, 2, 1], [1, 2, 2], False, fp64, seed=23, configure=cfg)


@pytest.mark.parametrize("vpt", [4, 8, 16])
def test_dual_line_lanes_fp32(dev, vpt):
    """FP32 C2C lanes-across-lines passes with two adjacent lines per lane
    (forced on, 2..32 lines per block) against the oracle; shapes whose lines
    are not adjacent pairs (odd rows) run the one-line kernels instead."""
    for lpb in (2, 4, 8, 16, 32):
        def cfg(plan, ins, outs, lpb=lpb):
            for d in ("fwd", "bwd"):
                for i in range(plan.num_passes(0 if d == "fwd" else 1)):
                    plan.set_option(f"{d}.{i}.dual", 1)
                    plan.set_option(f"{d}.{i}.lf", 1)
                    plan.set_option(f"{d}.{i}.lpb", lpb)
                    plan.set_option(f"{d}.{i}.vpt", vpt)
        _check(dev, [128, 256, 16], [0, 1, 2], [2, 1, 1], [1, 2, 1], False, False, seed=41, configure=cfg)
        _check(dev, [512, 8, 6], [0, 1, 2], [2, 1, 1], [1, 2, 1], False, False, seed=42, configure=cfg)
        _check(dev, [128, 128, 10], [0, 1, 2], [2, 1, 1], [1, 2, 1], True, False, seed=43, configure=cfg)
        _check(dev, [256, 6, 7], [0, 1, 2], [2, 1, 1], [1, 2, 1], False, False, seed=44, configure=cfg)
        _check(dev, [256, 128, 8], [0, 1, 2], [2, 2, 1], [1, 2, 2], False, False, seed=45, configure=cfg)


_SLAB_VIEWS = {
    # Test_SlabPlan.cpp:77-91 (View2D: one or two FFT axes), :330-341 / :1040-1052 /
    # :2081-2093 (View3D: one, two or three axes), :2735-2752 (View4D: three axes)
    2: ([8, 7], [[0], [1], [0, 1], [1, 0]]),
    3: ([8, 7, 6], [[0], [1], [2], [0, 1], [0, 2], [1, 0], [1, 2], [2, 0], [2, 1]]
        + [list(p) for p in itertools.permutations([0, 1, 2])]),
    4: ([8, 7, 6, 6], [[0, 1, 2], [0, 2, 1], [1, 0, 2], [1, 2, 0], [2, 0, 1], [2, 1, 0],
                       [1, 2, 3], [1, 3, 2], [2, 1, 3]]),
}


@pytest.mark.parametrize("view_layout", ["right", "left"])
@pytest.mark.parametrize("fp64", [True, False])
@pytest.mark.parametrize("ndim", [2, 3, 4])
def test_reference_slab_matrix(dev, ndim, fp64, view_layout):
    """The slab test matrix of Test_SlabPlan.cpp:3549-3640 (TestSlab1D / 2D / 3D x
    View2D / 3D / 4D x R2C / C2C, typed over {float, double} x {LayoutLeft,
    LayoutRight}): every pair of slab topologies, every axis set the reference
    lists, on 2 and 3 virtual ranks (ragged blocks: 7 = 4 + 3, 8 = 3 + 3 + 2)."""
    gin, axis_sets = _SLAB_VIEWS[ndim]
    for P in (2, 3):
        topos = []
        for k in range(ndim):
            t = [1] * ndim
            t[k] = P
            topos.append(t)
        for tin, tout in itertools.product(topos, repeat=2):
            for axes in axis_sets:
                for real in (False, True):
                    plan = _check(dev, gin, axes, tin, tout, real, fp64, seed=31 + P,
                                  view_layout=None if view_layout == "right" else "left")
                    assert plan.name() in ("SharedPlan", "SlabPlan")


def _pencil_topologies(ndim, px, py):
    """x-, y- and z-pencils of Test_PencilPlan.cpp:99-100 / :1763-1764 / :3648-3650
    (the z-pencil is a LayoutLeft topology), with the 4-D (1, px, py, 1) pencil."""
    if ndim == 3:
        return [([1, px, py], "right"), ([px, 1, py], "right"), ([py, px, 1], "left")]
    return [([1, 1, px, py], "right"), ([1, px, 1, py], "right"), ([1, py, px, 1], "left"),
            ([1, px, py, 1], "right")]


_PENCIL_VIEWS = {
    # Test_PencilPlan.cpp:102-110, :587-596, :1766-1779, :2436-2447, :3653-3667
    3: ([8, 7, 5], [[0], [1], [2], [0, 1], [0, 2], [1, 0], [1, 2], [2, 0], [2, 1]]
        + [list(p) for p in itertools.permutations([0, 1, 2])]),
    4: ([8, 7, 5, 6], [[0, 1], [0, 2], [1, 0], [1, 2], [2, 0], [2, 1], [1, 3], [2, 3], [0, 3]]
        + [list(p) for p in itertools.permutations([0, 1, 2])] + [[1, 2, 3], [1, 3, 2], [2, 1, 3]]),
}


@pytest.mark.parametrize("view_layout", ["right", "left"])
@pytest.mark.parametrize("fp64", [True, False])
@pytest.mark.parametrize("ndim", [3, 4])
def test_reference_pencil_matrix(dev, ndim, fp64, view_layout):
    """The pencil test matrix of Test_PencilPlan.cpp:4521-4593 (TestPencil1D / 2D /
    3D x View3D / 4D x R2C / C2C over {float, double} x {LayoutLeft, LayoutRight}),
    every pair of the x / y / z pencils for every axis set, on 2 x 2 and 2 x 3
    virtual ranks.  A pair whose transposes include one between equal topologies
    of different layouts has no buffer extents in the reference
    (Extents.hpp:35-60 merges the two topologies, ContainerAnalyses.hpp:181-226
    refuses equal ones): the oracle's planner refuses it and so must the plan."""
    import dfft
    from oracle import planner as OPL
    from helpers import out_extents
    gin, axis_sets = _PENCIL_VIEWS[ndim]
    vl = None if view_layout == "right" else "left"
    refused = 0
    for px, py in ((2, 2), (2, 3)):
        topos = _pencil_topologies(ndim, px, py)
        for (tin, lin), (tout, lout) in itertools.product(topos, repeat=2):
            for axes in axis_sets:
                for real in (False, True):
                    try:
                        lay = OPL.LAYOUT_LEFT if vl else OPL.LAYOUT_RIGHT
                        gout = out_extents(gin, axes, real)
                        ext = OPL.compute_local_extents(gin, tin, list(range(ndim)), 0,
                                                        OPL.LAYOUT_LEFT if lin == "left" else OPL.LAYOUT_RIGHT)
                        OPL.block_analyses(OPL.PENCIL, real, lay, ext, gin, gout, tin, tout, axes, 0,
                                           OPL.LAYOUT_LEFT if lin == "left" else OPL.LAYOUT_RIGHT,
                                           OPL.LAYOUT_LEFT if lout == "left" else OPL.LAYOUT_RIGHT)
                    except OPL.PlannerError:
                        refused += 1
                        with pytest.raises(dfft.DFFTError, match="differ exactly two positions"):
                            _check(dev, gin, axes, tin, tout, real, fp64, tin_layout=lin, tout_layout=lout,
                                   view_layout=vl)
                        continue
                    plan = _check(dev, gin, axes, tin, tout, real, fp64, seed=41 + px + py,
                                  tin_layout=lin, tout_layout=lout, view_layout=vl)
                    assert plan.name() in ("SharedPlan", "SlabPlan", "PencilPlan")
    assert refused == (8 if ndim == 4 else 0)


@pytest.mark.parametrize("fp64", [True, False])
@pytest.mark.parametrize("lf", [0, 1])
def test_unpaired_real_lines(dev, fp64, lf):
    """R2C / C2R passes with one real line per complex FFT (`pair` off: the
    one-line kernel loads reals for R2C and the Hermitian extension of the
    half spectrum for C2R), lanes along or across the lines, every fast
    length class, against the oracle."""
    def cfg(plan, ins, outs):
        for d in ("fwd", "bwd"):
            for i in range(plan.num_passes(0 if d == "fwd" else 1)):
                plan.set_option(f"{d}.{i}.pair", 0)
                plan.set_option(f"{d}.{i}.lf", lf)
    for gin, axes, tin, tout in (([8, 6, 512], [0, 1, 2], [2, 1, 1], [1, 2, 1]),
                                 ([6, 96, 10], [2, 0, 1], [1, 1, 2], [2, 1, 1]),
                                 ([160, 4, 6], [1, 2, 0], [2, 1, 1], [1, 2, 1]),
                                 ([16, 1024, 3], [0, 2, 1], [1, 2, 1], [2, 1, 1]),
                                 ([12, 16, 32], [0, 1, 2], [2, 2, 1], [1, 2, 2])):
        _check(dev, gin, axes, tin, tout, True, fp64, seed=61, configure=cfg)


@pytest.mark.parametrize("fp64", [True, False])
@pytest.mark.parametrize("lin", [0, 1])
def test_linear_line_enumeration(dev, fp64, lin):
    """One-line lanes-across-lines C2C passes from a padded scratch into odd
    unpadded output rows (the forward x pass of an R2C slab plan, rows of
    n/2 + 1 elements), with the lines enumerated over the real elements
    (lin 1: blocks cross row ends) or over the padded rows (lin 0), every
    lines-per-block size, against the oracle."""
    for lpb in (2, 8, 16):
        def cfg(plan, ins, outs, lpb=lpb):
            plan.set_option("pad", 1)
            for d in ("fwd", "bwd"):
                for i in range(plan.num_passes(0 if d == "fwd" else 1)):
                    plan.set_option(f"{d}.{i}.lin", lin)
                    plan.set_option(f"{d}.{i}.lf", 1)
                    plan.set_option(f"{d}.{i}.dual", 0)
                    plan.set_option(f"{d}.{i}.lpb", lpb)
            assert plan.get_option("fwd.2.lin") == lin
        _check(dev, [64, 12, 64], [0, 1, 2], [2, 1, 1], [1, 2, 1], True, fp64, seed=71, configure=cfg)
        _check(dev, [32, 10, 30], [0, 1, 2], [2, 1, 1], [1, 2, 1], True, fp64, seed=72, configure=cfg)
        _check(dev, [16, 8, 64], [0, 1, 2], [2, 2, 1], [1, 2, 2], True, fp64, seed=73, configure=cfg)
