"""FP32 path against FP64 at the precision-report sizes: 512^3 C2C and
1024^3 R2C, slab (4,1,1) -> (1,4,1), on U[-1,1] inputs and on the single mode
sin(2 pi (3x + 5y + 7z)/n), whose spectrum is known exactly."""

import math

import pytest

pytestmark = pytest.mark.gpu

P = 4


def _forward(dev, n, real, fp64, field):
    import torch
    import dfft
    gin = [n] * 3
    gout = [n, n, n // 2 + 1] if real else gin
    rdt = torch.float64 if fp64 else torch.float32
    cdt = torch.complex128 if fp64 else torch.complex64
    gen = torch.Generator(device=dev).manual_seed(3)
    ins, outs = [], []
    for r in range(P):
        ext, st = dfft.compute_local_extents_and_starts(gin, [P, 1, 1], r)
        if field == "uniform":
            x = torch.rand(ext, generator=gen, device=dev, dtype=torch.float64) * 2 - 1
            if not real:
                x = torch.complex(x, torch.rand(ext, generator=gen, device=dev, dtype=torch.float64) * 2 - 1)
        else:
            ii = [torch.arange(st[a], st[a] + ext[a], device=dev, dtype=torch.float64) for a in range(3)]
            X, Y, Z = torch.meshgrid(*ii, indexing="ij")
            x = torch.sin(2 * math.pi * (3 * X + 5 * Y + 7 * Z) / n)
            del X, Y, Z
            if not real:
                x = torch.complex(x, torch.zeros_like(x))
        ins.append(x.to(rdt if real else cdt).contiguous())
        del x
        oe, _ = dfft.compute_local_extents_and_starts(gout, [1, P, 1], r)
        outs.append(torch.empty(oe, dtype=cdt, device=dev))
    plan = dfft.Plan(ins, outs, [0, 1, 2], [P, 1, 1], [1, P, 1], global_in_extents=gin, global_out_extents=gout)
    dfft.execute(plan, ins, outs, dfft.Direction.forward)
    torch.cuda.synchronize()
    plan.destroy()
    del ins
    return outs, gout


def _exact_error(blocks, gout, n, real):
    import torch
    import dfft
    num = den = 0.0
    for r, blk in enumerate(blocks):
        ext, st = dfft.compute_local_extents_and_starts(gout, [1, P, 1], r)
        want = torch.zeros(blk.shape, dtype=torch.complex128, device=blk.device)
        for sgn in (1, -1):
            if real and sgn < 0:
                continue
            k = [(sgn * 3) % n, (sgn * 5) % n, (sgn * 7) % n]
            loc = [k[a] - st[a] for a in range(3)]
            if all(0 <= loc[a] < ext[a] for a in range(3)):
                want[loc[0], loc[1], loc[2]] = complex(0.0, -sgn * n ** 3 / 2)
        num += float(torch.linalg.vector_norm(blk.to(torch.complex128) - want)) ** 2
        den += float(torch.linalg.vector_norm(want)) ** 2
    return math.sqrt(num / den)


@pytest.mark.parametrize("n,real", [(512, False), (1024, True)])
def test_fp32_against_fp64_and_exact(dev, n, real):
    import torch
    for field in ("uniform", "sin"):
        o64, gout = _forward(dev, n, real, True, field)
        o32, _ = _forward(dev, n, real, False, field)
        num = sum(float(torch.linalg.vector_norm(a.to(torch.complex128) - b)) ** 2 for a, b in zip(o32, o64))
        den = sum(float(torch.linalg.vector_norm(b)) ** 2 for b in o64)
        err = math.sqrt(num / den)
        print(f"{n}^3 {'R2C' if real else 'C2C'} {field}: fp32 vs fp64 rel-L2 {err:.3e}")
        assert err <= 1e-5
        if field == "sin":
            e64, e32 = _exact_error(o64, gout, n, real), _exact_error(o32, gout, n, real)
            print(f"  vs exact: fp64 {e64:.3e} fp32 {e32:.3e}")
            assert e64 <= 1e-12
            assert e32 <= 1e-5
        del o64, o32
        torch.cuda.empty_cache()
