"""FP32 path against FP64 at the precision-report sizes: 512^3 C2C and
1024^3 R2C, slab (4,1,1) -> (1,4,1), on U[-1,1] inputs and on the single mode
sin(2 pi (3x + 5y + 7z)/n), whose spectrum is known exactly.  For the
uniform inputs both spectra are also checked against the exact DFT of the
FP64 input, evaluated on the host at a grid of sampled wavenumbers."""

import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

P = 4


def _forward(dev, n, real, fp64, field, keep_input=None):
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
        if keep_input is not None:
            keep_input.append(x.cpu().numpy())
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


def _sample_ks(n, real, seed):
    """Wavenumbers sampled per axis: DC, the first, a middle, the Nyquist and
    the last index, plus random ones (half spectrum along z for R2C)."""
    rng = np.random.default_rng(seed)
    out = []
    for a in range(3):
        top = n // 2 + 1 if (real and a == 2) else n
        ks = {0, 1, n // 2, top - 1, n // 3}
        ks.update(int(v) for v in rng.integers(0, top, 3))
        out.append(sorted(ks))
    return out


def _phase_matrix(n, ks):
    """E[x, j] = exp(-2 pi i (x k_j mod n) / n), the phase reduced exactly."""
    x = np.arange(n, dtype=np.int64)[:, None]
    m = (x * np.asarray(ks, dtype=np.int64)[None, :]) % n
    return np.exp(-2j * np.pi * m / n)


def _sampled_dft(x, ks):
    """Exact (FP64) DFT of the global array x at every k in ks[0] x ks[1] x ks[2],
    by three separable contractions (z first, over the whole array)."""
    n0, n1, n2 = x.shape
    e0, e1, e2 = (_phase_matrix(nn, k) for nn, k in zip(x.shape, ks))
    flat = x.reshape(n0 * n1, n2)
    if np.iscomplexobj(x):
        a = flat @ e2
    else:  # real input: two real products instead of a complex copy of the array
        a = (flat @ np.ascontiguousarray(e2.real)) + 1j * (flat @ np.ascontiguousarray(e2.imag))
    a = a.reshape(n0, n1, len(ks[2]))
    b = np.einsum("xyk,yj->xjk", a, e1)
    return np.einsum("xjk,xi->ijk", b, e0)


def _gpu_at(blocks, gout, ks):
    import torch
    import dfft
    out = np.zeros([len(k) for k in ks], dtype=np.complex128)
    for r, blk in enumerate(blocks):
        ext, st = dfft.compute_local_extents_and_starts(gout, [1, P, 1], r)
        sel = [[(j, k - st[a]) for j, k in enumerate(ks[a]) if st[a] <= k < st[a] + ext[a]] for a in range(3)]
        if not all(sel):
            continue
        idx = [torch.tensor([v for _, v in s_], device=blk.device) for s_ in sel]
        vals = blk[idx[0][:, None, None], idx[1][None, :, None], idx[2][None, None, :]]
        out[np.ix_(*[[j for j, _ in s_] for s_ in sel])] = vals.to(torch.complex128).cpu().numpy()
    return out


@pytest.mark.parametrize("n,real", [(512, False), (1024, True)])
def test_fp32_against_fp64_and_exact(dev, n, real):
    import torch
    for field in ("uniform", "sin"):
        x64 = [] if field == "uniform" else None
        o64, gout = _forward(dev, n, real, True, field, keep_input=x64)
        o32, _ = _forward(dev, n, real, False, field)
        if x64 is not None:
            ks = _sample_ks(n, real, seed=n)
            exact = _sampled_dft(np.concatenate(x64, axis=0), ks)
            del x64
            e64 = np.linalg.norm(_gpu_at(o64, gout, ks) - exact) / np.linalg.norm(exact)
            e32 = np.linalg.norm(_gpu_at(o32, gout, ks) - exact) / np.linalg.norm(exact)
            print(f"{n}^3 {'R2C' if real else 'C2C'} uniform at {exact.size} sampled k vs exact DFT: "
                  f"fp64 {e64:.3e} fp32 {e32:.3e}")
            assert e64 <= 1e-12
            assert e32 <= 1e-5
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
