"""The oracle's OpenMP CPU executor (oracle/cpu_plan.c, the timed CPU
baseline) against the Python oracle plan and numpy on the same inputs."""

import numpy as np
import pytest

from oracle import cpu_plan as CP
from oracle import plan_exec as PE


@pytest.fixture(scope="module", autouse=True)
def _threads():
    CP.set_threads(4)


@pytest.mark.parametrize("shape", [(8, 16, 32), (6, 12, 10), (3, 64, 5), (2, 128, 9)])
def test_fft_axis_against_numpy(shape):
    """Power-of-two lengths take the radix-4/2 Stockham path (contiguous lines
    and interleaved chunks of 8 strided lines, with a ragged last chunk),
    other lengths the recursive transform."""
    rng = np.random.default_rng(0)
    x = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
    for ax in range(3):
        for sign in (-1, 1):
            y = CP.fft_axis([x.copy()], ax, sign)[0]
            ref = np.fft.fft(x, axis=ax) if sign < 0 else np.fft.ifft(x, axis=ax) * shape[ax]
            assert np.abs(y - ref).max() < 1e-13 * np.abs(ref).max(), (ax, sign)
    x0 = x.copy()
    out = np.empty_like(x)
    CP.fft_axis([x], 1, -1, [out])  # out of place leaves the input alone
    assert np.array_equal(x, x0)
    ref = np.fft.fft(x, axis=1)
    assert np.abs(out - ref).max() < 1e-13 * np.abs(ref).max()


def test_fft_axis_rejects_uneven_lengths():
    a = np.zeros((4, 8, 8), np.complex128)
    b = np.zeros((4, 6, 8), np.complex128)
    with pytest.raises(ValueError):
        CP.fft_axis([a, b], 1, -1)


@pytest.mark.parametrize("g,tin,tout", [
    ([16, 12, 10], [2, 1, 1], [1, 2, 1]),
    ([16, 16, 16], [4, 1, 1], [1, 4, 1]),
    ([9, 8, 7], [3, 1, 1], [1, 3, 1]),
    ([16, 8, 8], [1, 1, 2], [2, 1, 1]),
    ([8, 12, 10], [1, 2, 1], [2, 1, 1]),
    ([16, 12, 8], [2, 2, 1], [1, 2, 2]),
    ([12, 10, 9], [2, 3, 1], [1, 2, 3]),
])
@pytest.mark.parametrize("norm", ["backward", "ortho"])
def test_block_walk_matches_oracle_plan(g, tin, tout, norm):
    rng = np.random.default_rng(1)
    x = rng.uniform(-1, 1, g) + 1j * rng.uniform(-1, 1, g)
    op = PE.OraclePlan(g, g, [0, 1, 2], tin, tout, norm=norm)
    cp = CP.CpuPlan(g, [0, 1, 2], tin, tout, norm=norm)
    ins = PE.scatter(x, g, tin)
    ref = op.forward(ins)
    outs = [np.full_like(r, np.nan) for r in ref]
    cp.forward(ins, outs)
    scale = max(np.abs(r).max() for r in ref)
    assert max(np.abs(a - b).max() for a, b in zip(outs, ref)) < 1e-13 * scale
    back = [np.full_like(a, np.nan) for a in ins]
    cp.backward(outs, back)
    rb = op.backward(ref)
    assert max(np.abs(a - b).max() for a, b in zip(back, rb)) < 1e-13
    assert max(np.abs(a - b).max() for a, b in zip(back, ins)) < 1e-13
