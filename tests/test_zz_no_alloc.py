"""dfft_group_execute_noalloc (pass tables in the library's static table
area) against dfft_group_execute (allocated tables): the same launches, so
bitwise equal outputs; two calls in a row, the second with a smaller group;
tables beyond the static area are refused before anything is launched."""

import pytest

pytestmark = pytest.mark.gpu


def _members(dev, sizes, seed):
    """sizes: (n, real) pairs -> member inputs and outputs (complex64 / float32)."""
    import torch
    gen = torch.Generator(device=dev).manual_seed(seed)
    ins, outs = [], []
    for n, real in sizes:
        shape = (n, n, n)
        re = torch.rand(shape, generator=gen, device=dev, dtype=torch.float32) * 2 - 1
        if real:
            ins.append(re.contiguous())
            outs.append(torch.empty((n, n, n // 2 + 1), dtype=torch.complex64, device=dev))
        else:
            im = torch.rand(shape, generator=gen, device=dev, dtype=torch.float32) * 2 - 1
            ins.append(torch.complex(re, im).contiguous())
            outs.append(torch.empty(shape, dtype=torch.complex64, device=dev))
    return ins, outs


def _same_bits(a, b):
    import torch

    def bits(t):
        return (torch.view_as_real(t) if t.is_complex() else t).contiguous().view(torch.uint8)
    return bool(torch.equal(bits(a), bits(b)))


def _compare(dev, sizes, seed):
    import torch
    import dfft
    ins, outs = _members(dev, sizes, seed)
    g = dfft.Group(ins, outs)
    ref = [torch.full_like(o, float("nan")) for o in outs]
    got = [torch.full_like(o, float("nan")) for o in outs]
    g.execute(ins, ref, dfft.Direction.forward)
    g.execute_noalloc(ins, got, dfft.Direction.forward)
    torch.cuda.synchronize()
    for i, (a, b) in enumerate(zip(ref, got)):
        assert _same_bits(a, b), f"forward member {i} {sizes[i]}"
    ref_b = [torch.full_like(a, float("nan")) for a in ins]
    got_b = [torch.full_like(a, float("nan")) for a in ins]
    g.execute(ref, ref_b, dfft.Direction.backward)
    g.execute_noalloc(got, got_b, dfft.Direction.backward)
    torch.cuda.synchronize()
    for i, (a, b, x) in enumerate(zip(ref_b, got_b, ins)):
        assert _same_bits(a, b), f"backward member {i} {sizes[i]}"
        assert float(torch.linalg.vector_norm(b - x) / torch.linalg.vector_norm(x)) < 1e-5
    g.destroy()


def test_noalloc_matches_allocating_execute_twice(dev):
    # grouped lengths (64, 96, 128, 160, 32) and one without a grouped kernel (48)
    _compare(dev, [(64, False), (96, False), (128, True), (48, False), (32, True), (160, False)], seed=7)
    # a second call with a smaller group overwrites the shared tables
    _compare(dev, [(32, False), (16, True)], seed=8)


def test_noalloc_refuses_tables_beyond_the_static_area(dev):
    import torch
    import dfft
    n = 3000  # 3 passes x 3000 member records > 4 MiB of tables
    ins = [torch.zeros(16, 16, 16, dtype=torch.complex64, device=dev) for _ in range(n)]
    outs = [torch.full((16, 16, 16), 5.0, dtype=torch.complex64, device=dev) for _ in range(n)]
    g = dfft.Group(ins, outs)
    with pytest.raises(dfft.DFFTError) as e:
        g.execute_noalloc(ins, outs, dfft.Direction.forward)
    assert e.value.code == dfft.ERROR_UNSUPPORTED
    torch.cuda.synchronize()
    assert all(bool((o == 5.0).all()) for o in outs[:8])  # nothing was launched
    g.destroy()
