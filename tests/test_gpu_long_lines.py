"""Transform lengths longer than one line in LDS (n = n1 * n2, two launches
through a complex scratch) against the oracle, in every executor that runs
them natively, and their adjoints.  The reference's DynPlan takes any
extent (SlabPlan.hpp:1098-1113)."""

import numpy as np
import pytest

from helpers import gather_rel_l2, out_extents, reference_spectrum, run_gpu_plan, seeded_input
from oracle import plan_exec as PE
from oracle import planner as PL

pytestmark = pytest.mark.gpu

CASES = [
    # gin, axes, tin, tout, fp64
    ([8192, 6, 4], [0, 1, 2], [1, 2, 1], [1, 1, 2], True),     # 8192 = 64 x 128, C2C axis first
    ([4, 6, 8192], [0, 1, 2], [2, 1, 1], [1, 2, 1], True),     # the long axis is the R2C axis
    ([6405, 3, 4], [0, 1, 2], [1, 1, 2], [1, 2, 1], True),     # 6405 = 61 x 105 (odd radices, prime 61)
    ([6, 8, 8192], [2, 1, 0], [1, 2, 2], [2, 2, 1], True),     # pencil, long axis transformed first
    ([2, 3, 12288], [0, 1, 2], [2, 1, 1], [1, 2, 1], False),   # FP32: 12288 = 96 x 128
    ([12288, 4], [0, 1], [1, 2], [2, 1], False),               # 2-D view
]


@pytest.mark.parametrize("gin,axes,tin,tout,fp64", CASES)
@pytest.mark.parametrize("real", [False, True])
@pytest.mark.parametrize("exec_mode", [0, 1])
def test_long_lines_match_oracle(dev, gin, axes, tin, tout, fp64, real, exec_mode):
    x = seeded_input(gin, real, 3)
    outs, backs, ins, plan = run_gpu_plan(dev, x, gin, axes, tin, tout, real, fp64, "backward", exec_mode)
    ref = reference_spectrum(x, gin, axes, tout, real)
    tol = 1e-12 if fp64 else 1e-5
    e_fwd, e_rt = gather_rel_l2(outs, ref), gather_rel_l2(backs, ins)
    assert e_fwd <= tol, f"forward rel-L2 {e_fwd:.3e}"
    assert e_rt <= tol, f"round trip rel-L2 {e_rt:.3e}"


@pytest.mark.parametrize("gin,axes,tin,tout", [
    ([8192, 6, 4], [0, 1, 2], [1, 2, 1], [1, 1, 2]),   # two launches
    ([6007, 4, 3], [0, 1, 2], [1, 2, 1], [1, 1, 2]),   # Bluestein
])
@pytest.mark.parametrize("real", [False, True])
@pytest.mark.parametrize("exec_mode", [0, 1])
def test_long_lines_layout_left_views(dev, gin, axes, tin, tout, real, exec_mode):
    """The same paths on LayoutLeft (column-major) views."""
    x = seeded_input(gin, real, 9)
    outs, backs, ins, _ = run_gpu_plan(dev, x, gin, axes, tin, tout, real, True, "backward", exec_mode,
                                       view_layout="left")
    ref = reference_spectrum(x, gin, axes, tout, real)
    assert gather_rel_l2(outs, ref) <= 1e-12
    assert gather_rel_l2(backs, ins) <= 1e-12


@pytest.mark.parametrize("gin,tin,tout", [
    ([6, 8192, 4], [2, 1, 1], [1, 1, 2]),   # the long axis in the first two (chunked) passes
    ([6, 6007, 4], [2, 1, 1], [1, 1, 2]),   # Bluestein in a chunked pass
])
@pytest.mark.parametrize("configure", ["chunk", "autotune"])
def test_long_lines_chunked_and_tuned(dev, gin, tin, tout, configure):
    """Chunked first passes (lines of one chunk at a line offset) and the
    autotuner on plans whose passes run two launches or Bluestein."""
    def conf(plan, ins, outs):
        if configure == "chunk":
            plan.set_option("chunk", 2)
        else:
            plan.autotune(ins, outs)

    for real in (False, True):
        x = seeded_input(gin, real, 13)
        outs, backs, ins, _ = run_gpu_plan(dev, x, gin, [0, 1, 2], tin, tout, real, True, "backward", 0,
                                           configure=conf)
        ref = reference_spectrum(x, gin, [0, 1, 2], tout, real)
        assert gather_rel_l2(outs, ref) <= 1e-12
        assert gather_rel_l2(backs, ins) <= 1e-12


@pytest.mark.parametrize("n", [4096, 8192])
@pytest.mark.parametrize("fp64", [True, False])
@pytest.mark.parametrize("real", [False, True])
def test_fast_kernels_for_4096_and_8192(dev, n, fp64, real):
    """Compile-time kernels of 4096 and 8192 points (one line, pairs of real
    lines, two FP32 lines per lane) along the contiguous and a strided axis,
    including every launch shape the options allow."""
    for gin, tin, tout in (([4, 6, n], [2, 1, 1], [1, 2, 1]), ([n, 6, 4], [1, 2, 1], [1, 1, 2])):
        x = seeded_input(gin, real, 29)
        for vpt in (4, 8, 16):
            def conf(plan, ins, outs, vpt=vpt):
                for k in range(3):
                    plan.set_option(f"fwd.{k}.vpt", vpt)
                    plan.set_option(f"bwd.{k}.vpt", vpt)
            outs, backs, ins, _ = run_gpu_plan(dev, x, gin, [0, 1, 2], tin, tout, real, fp64, "backward", 0,
                                               configure=conf)
            ref = reference_spectrum(x, gin, [0, 1, 2], tout, real)
            tol = 1e-12 if fp64 else 1e-5
            assert gather_rel_l2(outs, ref) <= tol, (gin, vpt)
            assert gather_rel_l2(backs, ins) <= tol, (gin, vpt)


def test_long_line_closed_form(dev):
    """A single mode exp(2 pi i m j / n) along an 8192-point axis transforms
    to n at k = m and 0 elsewhere (independent of the oracle's FFT)."""
    n, m = 8192, 1237
    j = np.arange(n)
    x = np.broadcast_to(np.exp(2j * np.pi * m * j / n)[:, None], (n, 4)).copy()
    outs, _, _, _ = run_gpu_plan(dev, x, [n, 4], [0, 1], [1, 2], [2, 1], False, True)
    got = PE.gather(outs, out_extents([n, 4], [0, 1], False), [2, 1], np.complex128)
    want = np.zeros((n, 4), np.complex128)
    want[m, 0] = n * 4
    assert np.abs(got - want).max() <= 1e-9 * n * 4


BLUESTEIN_CASES = [
    ([6007, 4], [0, 1], [1, 2], [2, 1], True),         # 6007 prime: Bluestein, m = 16384 = 128 x 128
    ([3, 4, 6007], [0, 1, 2], [3, 1, 1], [1, 3, 1], True),  # ... as the R2C axis
    ([10243, 2], [0, 1], [1, 2], [2, 1], False),       # FP32 beyond 10240, prime
]


@pytest.mark.parametrize("gin,axes,tin,tout,fp64", BLUESTEIN_CASES)
@pytest.mark.parametrize("real", [False, True])
@pytest.mark.parametrize("exec_mode", [0, 1])
def test_prime_lengths_beyond_one_line_run_bluestein(dev, gin, axes, tin, tout, fp64, real, exec_mode):
    x = seeded_input(gin, real, 5)
    outs, backs, ins, _ = run_gpu_plan(dev, x, gin, axes, tin, tout, real, fp64, "ortho", exec_mode)
    ref = reference_spectrum(x, gin, axes, tout, real, "ortho")
    tol = 1e-12 if fp64 else 1e-5
    e_fwd, e_rt = gather_rel_l2(outs, ref), gather_rel_l2(backs, ins)
    assert e_fwd <= tol, f"forward rel-L2 {e_fwd:.3e}"
    assert e_rt <= tol, f"round trip rel-L2 {e_rt:.3e}"


@pytest.mark.parametrize("gin,fp64", [
    ([127, 6, 4], True), ([1009, 4, 3], True), ([2018, 3, 2], True),   # large prime factors: Bluestein
    ([61, 8, 4], False), ([1009, 4, 3], False),
    ([61, 8, 4], True), ([31, 8, 4], False),                         # small ones: the generic kernel
])
@pytest.mark.parametrize("real", [False, True])
def test_lengths_with_large_prime_factors(dev, gin, fp64, real):
    """Lengths whose largest prime factor makes the generic kernel's r-term
    sums expensive run by Bluestein (cost model); either way the results
    match the oracle."""
    tin, tout = [1, 2, 1], [1, 1, 2]
    x = seeded_input(gin, real, 21)
    outs, backs, ins, _ = run_gpu_plan(dev, x, gin, [0, 1, 2], tin, tout, real, fp64)
    ref = reference_spectrum(x, gin, [0, 1, 2], tout, real)
    tol = 1e-12 if fp64 else 1e-5
    assert gather_rel_l2(outs, ref) <= tol
    assert gather_rel_l2(backs, ins) <= tol


def test_bluestein_single_mode(dev):
    """A single mode along a 6007-point (prime) axis: n at k = m, 0 elsewhere."""
    n, m = 6007, 2999
    j = np.arange(n)
    x = np.broadcast_to(np.exp(2j * np.pi * m * j / n)[:, None], (n, 2)).copy()
    outs, _, _, _ = run_gpu_plan(dev, x, [n, 2], [0, 1], [1, 2], [2, 1], False, True)
    got = PE.gather(outs, out_extents([n, 2], [0, 1], False), [2, 1], np.complex128)
    want = np.zeros((n, 2), np.complex128)
    want[m, 0] = n * 2
    assert np.abs(got - want).max() <= 1e-9 * n * 2


@pytest.mark.parametrize("gin", [[4, 6, 8192], [4, 6, 6007]])
@pytest.mark.parametrize("real", [False, True])
def test_long_line_adjoint_dot_product(dev, gin, real):
    """<F x, y> = <x, F^* y> with a 8192-point axis (two launches) or a
    6007-point one (Bluestein), the R2C axis when real."""
    import torch
    import dfft
    axes, tin, tout = [0, 1, 2], [2, 1, 1], [1, 2, 1]
    gout = out_extents(gin, axes, real)
    cdt = torch.complex128
    idt = torch.float64 if real else cdt

    def blocks(arr, ext, topo, dt):
        return [torch.from_numpy(np.ascontiguousarray(a)).to(dev, dt) for a in PE.scatter(arr, ext, list(topo))]

    def rdot(a, b):
        return sum(float((np.conj(p.cpu().numpy()) * q.cpu().numpy()).real.sum()) for p, q in zip(a, b))

    def norm(a):
        return float(np.sqrt(sum(float(np.linalg.norm(t.cpu().numpy().ravel())) ** 2 for t in a)))

    ins = [torch.zeros(PL.compute_local_extents_and_starts(gin, tin, r, PL.LAYOUT_RIGHT)[0], dtype=idt, device=dev)
           for r in range(2)]
    outs = [torch.zeros(PL.compute_local_extents_and_starts(gout, tout, r, PL.LAYOUT_RIGHT)[0], dtype=cdt,
                        device=dev) for r in range(2)]
    plan = dfft.Plan(ins, outs, axes, tin, tout, global_in_extents=gin, global_out_extents=gout)
    x = blocks(seeded_input(gin, real, 1), gin, tin, idt)
    y = blocks(seeded_input(gout, False, 2), gout, tout, cdt)
    fx = [torch.empty_like(t) for t in y]
    fty = [torch.empty_like(t) for t in x]
    plan._execute(x, fx, dfft.Direction.forward)
    plan.execute_adjoint(y, fty, dfft.Direction.forward)
    lhs, rhs = rdot(fx, y), rdot(x, fty)
    assert abs(lhs - rhs) <= 1e-12 * (norm(x) * norm(fty) + norm(fx) * norm(y)), (lhs, rhs)
