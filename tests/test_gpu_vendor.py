"""rocFFT behind the plan boundary (DFFT_EXEC_VENDOR, the reference's TplPlan
role, KokkosFFT_Distributed_TplPlan.hpp:25-181): same views, same results as
the fused executor and the oracle."""

import pytest

from helpers import gather_rel_l2, reference_spectrum, run_gpu_plan, seeded_input

pytestmark = pytest.mark.gpu

VENDOR = 2


@pytest.mark.parametrize("fp64", [True, False])
@pytest.mark.parametrize("real", [False, True])
def test_vendor_plan_matches_oracle(dev, fp64, real):
    tol = 1e-12 if fp64 else 1e-5
    for gin, tin, tout, axes in [
        ([8, 7, 5], [2, 1, 1], [1, 2, 1], [0, 1, 2]),
        ([8, 7, 5], [1, 1, 3], [3, 1, 1], [2, 0, 1]),
        ([8, 7, 6], [2, 2, 1], [1, 2, 2], [0, 1, 2]),
        ([8, 7, 6], [1, 2, 3], [2, 3, 1], [1, 2]),
        ([3, 8, 7, 6], [1, 2, 1, 1], [1, 1, 2, 1], [1, 2, 3]),
        ([64, 64, 64], [4, 1, 1], [1, 4, 1], [0, 1, 2]),
    ]:
        x = seeded_input(gin, real, 21)
        outs, backs, ins, plan = run_gpu_plan(dev, x, gin, axes, tin, tout, real, fp64, exec_mode=VENDOR)
        assert plan.name() == "rocFFTPlan"
        assert gather_rel_l2(outs, reference_spectrum(x, gin, axes, tout, real)) <= tol
        assert gather_rel_l2(backs, ins) <= tol


@pytest.mark.parametrize("norm", ["backward", "forward", "ortho", "none"])
def test_vendor_plan_normalizations(dev, norm):
    gin = [12, 10, 8]
    x = seeded_input(gin, False, 3)
    outs, backs, ins, _ = run_gpu_plan(dev, x, gin, [0, 1, 2], [2, 1, 1], [1, 2, 1], False, True, norm,
                                       exec_mode=VENDOR)
    assert gather_rel_l2(outs, reference_spectrum(x, gin, [0, 1, 2], [1, 2, 1], False, norm)) <= 1e-12
    scale = 12 * 10 * 8 if norm == "none" else 1.0
    assert gather_rel_l2([b / scale for b in backs], ins) <= 1e-12


def test_vendor_plan_layout_left_and_any_length(dev):
    """LayoutLeft views, and a length (8191, prime) no line kernel of this
    library runs in FP64 -- rocFFT takes any length."""
    x = seeded_input([8, 7, 6], True, 4)
    outs, backs, ins, _ = run_gpu_plan(dev, x, [8, 7, 6], [0, 1, 2], [2, 1, 1], [1, 2, 1], True,
                                       exec_mode=VENDOR, view_layout="left")
    assert gather_rel_l2(outs, reference_spectrum(x, [8, 7, 6], [0, 1, 2], [1, 2, 1], True)) <= 1e-12
    gin = [2, 8191]
    x = seeded_input(gin, False, 5)
    outs, backs, ins, _ = run_gpu_plan(dev, x, gin, [1], [1, 1], [1, 1], False, exec_mode=VENDOR)
    assert gather_rel_l2(outs, reference_spectrum(x, gin, [1], [1, 1], False)) <= 1e-12
    assert gather_rel_l2(backs, ins) <= 1e-12


def test_vendor_plan_has_no_adjoint(dev):
    import torch
    import dfft
    ins = [torch.zeros(4, 8, 8, dtype=torch.complex128, device=dev) for _ in range(2)]
    outs = [torch.zeros(8, 4, 8, dtype=torch.complex128, device=dev) for _ in range(2)]
    plan = dfft.Plan(ins, outs, [0, 1, 2], [2, 1, 1], [1, 2, 1], exec_mode=VENDOR)
    with pytest.raises(dfft.DFFTError) as e:
        plan.execute_adjoint(outs, ins, dfft.Direction.forward)
    assert e.value.code == dfft.ERROR_UNSUPPORTED
