"""Benchmark: 3-D C2C FFT forward+backward round trip on one MI355X.

Workload (one "step"): 512^3 complex128, slab decomposition over 4 virtual
ranks, in-topology (4,1,1) -> out-topology (1,4,1), axes (0,1,2), forward
then backward with the default backward normalisation; the same round trip in
complex64 is timed beside it.  Inputs: U[-1,1] real and imaginary parts,
seeded, resident in HBM before the timed region.

Prints one JSON line.  `value` is the FP64 round-trip time in ms (lower is
better).  `roofline` is the line-pass kernel's algorithmic bytes per launch
(the compulsory bytes of a transform, input + output array, shared by its
three axis-pass launches) over the launches' average duration, timed with
HIP events on the launch stream: frac = 2B / t over the passes' own time;
`pass_efficiency` beside it counts each launch's own read + write of the
array.  `cpu_baseline` times the oracle's OpenMP CPU executor of the
reference's block sequence on the same 512^3 round trip (median of 10 after
3 warm-ups, every host core).

Further sections of the same line:
* r2c_1024: 1024^3 R2C forward + C2R backward, slab (4,1,1) -> (1,4,1),
  FP64 and FP32, autotuned, with per-pass algorithmic bytes and HBM fraction;
* ns: Taylor-Green vortex 256^3 float64, Re = 1600, dt = 1e-3, 2x2 virtual
  ranks, divergence form (3 velocity + 6 product fields per RK stage, each set
  in one batched transform), steps replayed from a captured hipGraph;
* ns_adjoint: 20 RK4 steps at 128^3 from the Taylor-Green field plus
  N(0, 1e-2) noise, checkpoint every 5 steps: gradient of the final kinetic
  energy; steps/s = 20 / (forward + recomputation + adjoint time);
* mixed_batch: 64 C2C transforms, N cycling through {96, 128, 160}^3
  complex64, forward + backward, one grouped launch per pass, hipGraph;
* precision: FP32 against FP64 (rel-L2) for 512^3 C2C and 1024^3 R2C on
  U[-1,1] inputs, and against the exact spectrum of sin(k.x), k = (3, 5, 7);
* vendor_plan: rocFFT behind the plan boundary (exec mode EXEC_VENDOR, the
  reference's TplPlan role) on the same views and round trip;
* rocfft_comparator: the vendor single-GPU FFT (torch.fft -> hipFFT/rocFFT)
  on one global 512^3 array, for reference;
* cpu_library_baseline: scipy.fft on all host cores beside the oracle.
"""

import argparse
import json
import math
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT / "distributed-fft-for-kokkos_amd"))
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0
N = 512
P = 4
# measured with rocprofv3 --pmc on this workload (profiles/fft_line_fast_512cubed_c2c_slab4.md)
TRAFFIC_BYTES_PER_LAUNCH = 4.2956e9
TRAFFIC_SOURCE = "profiles/fft_line_fast_512cubed_c2c_slab4.md: FETCH_SIZE x2.000 + WRITE_SIZE x1.000 per launch"


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def make_inputs(torch, dfft, dev, cdt, seed):
    g = torch.Generator(device=dev)
    g.manual_seed(seed)
    ins, outs = [], []
    for r in range(P):
        ext, _ = dfft.compute_local_extents_and_starts([N] * 3, [P, 1, 1], r)
        rdt = torch.float64 if cdt == torch.complex128 else torch.float32
        re = torch.rand(ext, generator=g, device=dev, dtype=rdt) * 2 - 1
        im = torch.rand(ext, generator=g, device=dev, dtype=rdt) * 2 - 1
        ins.append(torch.complex(re, im).contiguous())
        oext, _ = dfft.compute_local_extents_and_starts([N] * 3, [1, P, 1], r)
        outs.append(torch.empty(oext, dtype=cdt, device=dev))
    return ins, outs


def time_roundtrip(torch, dfft, plan, ins, outs, backs, steps, warmup):
    for _ in range(warmup):
        dfft.execute(plan, ins, outs, dfft.Direction.forward)
        dfft.execute(plan, outs, backs, dfft.Direction.backward)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        dfft.execute(plan, ins, outs, dfft.Direction.forward)
        dfft.execute(plan, outs, backs, dfft.Direction.backward)
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) * 1e3 / steps


def pass_times(dfft, plan, ins, outs, backs, reps):
    fw, bw = [], []
    for _ in range(reps):
        fw.append(plan.execute_timed(ins, outs, dfft.Direction.forward))
        bw.append(plan.execute_timed(outs, backs, dfft.Direction.backward))
    avg = lambda rows: [sum(r[i] for r in rows) / len(rows) for i in range(len(rows[0]))]
    return avg(fw), avg(bw)


def _median(xs):
    xs = sorted(xs)
    k = len(xs) // 2
    return xs[k] if len(xs) % 2 else 0.5 * (xs[k - 1] + xs[k])


def cpu_baseline(warmup=3, runs=10):
    """The oracle's OpenMP CPU executor of the reference's block sequence
    (oracle/cpu_plan.c: FFT2 along (y, z) per x-slab, pack -> all-to-all ->
    unpack into the (y, z, x) mapped layout, FFT along x, local transpose back,
    normalisation sweep) on the same 512^3 complex128 slab (4,1,1) -> (1,4,1)
    round trip, P = 4 virtual ranks on every host core (cores / P threads per
    rank).  Timed as distributed/perf_test/PerfTest_FFT3D.cpp:50-60 times the
    reference: the call returns after every rank is done (fence), wall clock,
    `warmup` untimed round trips, then the median of `runs`."""
    import numpy as np
    from oracle import cpu_plan as CP

    cores = CP.cores_available()
    CP.set_threads(cores)
    g = [N] * 3
    plan = CP.CpuPlan(g, [0, 1, 2], [P, 1, 1], [1, P, 1])
    rng = np.random.default_rng(3)
    ins = []
    for e in plan.meta.in_local:
        a = np.empty(e, np.complex128)
        a.real = rng.uniform(-1, 1, e)
        a.imag = rng.uniform(-1, 1, e)
        ins.append(a)
    outs = [np.empty(e, np.complex128) for e in plan.meta.out_local]
    backs = [np.empty_like(a) for a in ins]
    times = []
    for i in range(warmup + runs):
        t0 = time.perf_counter()
        plan.forward(ins, outs)
        plan.backward(outs, backs)
        dt = time.perf_counter() - t0
        if i >= warmup:
            times.append(dt * 1e3)
    err = max(float(np.abs(a - b).max()) for a, b in zip(ins, backs))
    return {"value": round(_median(times), 2), "unit": "ms per 512^3 complex128 fwd+bwd round trip",
            "cores": cores, "kind": "port", "threads_per_rank": cores / P, "runs": runs, "warmup": warmup,
            "min_ms": round(min(times), 2), "max_ms": round(max(times), 2), "roundtrip_max_abs_err": err,
            "sample": f"the full 512^3 complex128 slab P={P} workload: OpenMP CPU block walk of the oracle "
                      f"(own radix-4 FFT, reference's pack/all2all/unpack and transposes), {cores} threads, "
                      f"median of {runs} after {warmup} warm-ups"}


def cpu_library_baseline(warmup=3, runs=10):
    """scipy.fft (pocketfft) fftn + ifftn of one global 512^3 complex128 array
    on all host cores: a library CPU bound beside the block walk (no
    distribution, no exchanges)."""
    import numpy as np
    import scipy.fft as sf
    cores = len(os.sched_getaffinity(0))
    rng = np.random.default_rng(3)
    x = np.empty((N,) * 3, np.complex128)
    x.real = rng.uniform(-1, 1, x.shape)
    x.imag = rng.uniform(-1, 1, x.shape)
    times = []
    for i in range(warmup + runs):
        t0 = time.perf_counter()
        sf.ifftn(sf.fftn(x, workers=cores), workers=cores)
        dt = time.perf_counter() - t0
        if i >= warmup:
            times.append(dt * 1e3)
    return {"value": round(_median(times), 2), "unit": "ms per 512^3 complex128 fwd+bwd", "cores": cores,
            "kind": "library", "runs": runs, "warmup": warmup,
            "sample": f"scipy.fft fftn+ifftn of one 512^3 complex128 array, workers={cores}, "
                      f"median of {runs} after {warmup} warm-ups"}


def vendor_plan(torch, dfft, dev, steps, warmup):
    """rocFFT behind the plan boundary (dfft.EXEC_VENDOR, the reference's
    TplPlan role): the same 512^3 slab round trip through the same views, the
    ranks' blocks gathered into one array for one rocFFT call and scattered
    back."""
    out = {}
    for tag, cdt in (("fp64", torch.complex128), ("fp32", torch.complex64)):
        ins, outs = make_inputs(torch, dfft, dev, cdt, seed=3)
        backs = [torch.empty_like(a) for a in ins]
        plan = dfft.Plan(ins, outs, [0, 1, 2], [P, 1, 1], [1, P, 1], exec_mode=dfft.EXEC_VENDOR)
        ms = time_roundtrip(torch, dfft, plan, ins, outs, backs, steps, warmup)
        fw, bw = pass_times(dfft, plan, ins, outs, backs, 2)
        err = math.sqrt(sum(float(torch.linalg.vector_norm(b - a)) ** 2 for a, b in zip(ins, backs)) /
                        sum(float(torch.linalg.vector_norm(a)) ** 2 for a in ins))
        out[tag] = {"ms": round(ms, 4), "fwd_step_ms": [round(x, 4) for x in fw],
                    "bwd_step_ms": [round(x, 4) for x in bw], "roundtrip_rel_l2": err}
        del plan, ins, outs, backs
        torch.cuda.empty_cache()
    out["what"] = ("dfft plan with exec_mode=EXEC_VENDOR (plan name rocFFTPlan): gather remap, one rocFFT "
                   "call, scatter remap per direction; steps = [gather, rocFFT, scatter]")
    return out


def rocfft_comparator(torch, dev, steps):
    """The vendor single-GPU 3-D FFT (torch.fft -> hipFFT/rocFFT) on the same
    512^3 round trip, no distribution: a reference point, not this library."""
    out = {}
    for tag, dt in (("fp64", torch.complex128), ("fp32", torch.complex64)):
        g = torch.Generator(device=dev).manual_seed(3)
        rdt = torch.float64 if dt == torch.complex128 else torch.float32
        x = torch.complex(torch.rand((N,) * 3, generator=g, device=dev, dtype=rdt),
                          torch.rand((N,) * 3, generator=g, device=dev, dtype=rdt))
        for _ in range(2):
            y = torch.fft.ifftn(torch.fft.fftn(x))
        del y
        ms = _sync_ms(torch, lambda: torch.fft.ifftn(torch.fft.fftn(x)), steps)
        out[tag + "_ms"] = round(ms, 4)
        del x
        torch.cuda.empty_cache()
    out["what"] = "torch.fft.fftn + ifftn (hipFFT/rocFFT), 512^3, one GPU array, incl. its temporaries"
    return out


def _sync_ms(torch, fn, reps):
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        fn()
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) * 1e3 / reps


def bench_ns(torch, steps):
    import dfft_ns
    n = 256
    s = dfft_ns.NavierStokes(n, 2, 2, lx=2.0, nu=1.0 / 1600.0, dt=1e-3, form=dfft_ns.DIVERGENCE)
    s.autotune()
    fused = s.get_option("products_fused")
    s.init_taylor_green(1.0)
    e0 = s.energy()
    s.step(2, use_graph=True)  # eager first step + capture
    ms = _sync_ms(torch, lambda: s.step(1, use_graph=True), steps)
    e1 = s.energy()
    # the same step with the 6 product fields stored before the forward transform
    s.set_option("fuse_products", 0)
    s.init_taylor_green(1.0)
    s.step(2, use_graph=True)
    ms_stored = _sync_ms(torch, lambda: s.step(1, use_graph=True), steps)
    s.destroy()
    fields = 9 * 4  # transforms per step: (3 backward + 6 forward) x 4 stages
    return {"workload": "Taylor-Green 256^3 float64, 2x2 virtual ranks, RK4, divergence form, hipGraph",
            "ms_per_step": round(ms, 4), "steps_per_s": round(1e3 / ms, 3),
            "products_in_forward_pass": fused, "stored_products_ms_per_step": round(ms_stored, 4),
            "transforms_per_step": fields, "energy_t0": e0, "energy_after": e1}


def bench_ns_adjoint(torch, dev):
    import dfft_ns
    n, nsteps, ck = 128, 20, 5
    # Taylor-Green on [0, 2 pi)^3 (u = sin x cos y cos z, v = -cos x sin y cos z, w = 0) + N(0, 1e-2)
    x = torch.arange(n, dtype=torch.float64, device=dev) * (2 * math.pi / n)
    X, Y, Z = torch.meshgrid(x, x, x, indexing="ij")
    gen = torch.Generator(device=dev).manual_seed(5)
    u0 = torch.stack([torch.sin(X) * torch.cos(Y) * torch.cos(Z), -torch.cos(X) * torch.sin(Y) * torch.cos(Z),
                      torch.zeros_like(X)])
    u0 = (u0 + 1e-2 * torch.randn(u0.shape, generator=gen, device=dev, dtype=torch.float64)).contiguous()
    s = dfft_ns.NavierStokes(n, 1, 1, lx=2.0, nu=1.0 / 1600.0, dt=1e-3, form=dfft_ns.DIVERGENCE)
    u0d = [u0]
    s.autotune()
    s.gradient(2, u0d, checkpoint_every=1)  # warm-up (allocations, kernel setup)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    j, grad = s.gradient(nsteps, u0d, checkpoint_every=ck)
    torch.cuda.synchronize()
    t_adj = time.perf_counter() - t0
    fwd_ms = _sync_ms(torch, lambda: s.final_energy(nsteps, u0d), 2)
    s.destroy()
    return {"workload": "adjoint of 20 RK4 steps, Taylor-Green 128^3 float64 + N(0,1e-2) noise, "
                        "checkpoint every 5", "steps_per_s": round(nsteps / t_adj, 3),
            "gradient_s": round(t_adj, 4), "forward_only_steps_per_s": round(nsteps / (fwd_ms * 1e-3), 3),
            "adjoint_over_forward": round(t_adj / (fwd_ms * 1e-3), 3), "J": j}


def bench_mixed_batch(torch, dfft, dev, steps, warmup):
    sizes = [(96, 128, 160)[i % 3] for i in range(64)]
    gen = torch.Generator(device=dev).manual_seed(1234)
    ins, outs = [], []
    for n in sizes:
        re = torch.rand((n, n, n), generator=gen, device=dev, dtype=torch.float32) * 2 - 1
        im = torch.rand((n, n, n), generator=gen, device=dev, dtype=torch.float32) * 2 - 1
        ins.append(torch.complex(re, im))
        outs.append(torch.empty((n, n, n), dtype=torch.complex64, device=dev))
    backs = [torch.empty_like(t) for t in ins]
    g = dfft.Group(ins, outs)
    snapshot = [t.clone() for t in ins]
    g.autotune(ins, outs)  # overwrites ins/outs: restored below
    for a, b in zip(ins, snapshot):
        a.copy_(b)
    del snapshot

    def rt():
        g.execute(ins, outs, dfft.Direction.forward, use_graph=True)
        g.execute(outs, backs, dfft.Direction.backward, use_graph=True)

    for _ in range(max(1, warmup)):
        rt()
    ms = _sync_ms(torch, rt, steps)
    err = max(float(torch.linalg.vector_norm(b - a) / torch.linalg.vector_norm(a)) for a, b in zip(ins, backs))
    vol = sum(n ** 3 for n in sizes)
    launches = g.num_launches(dfft.Direction.forward) + g.num_launches(dfft.Direction.backward)
    pass_bytes = 2 * vol * 8  # one axis pass of every member reads and writes all of them
    passes = 6  # 3 axis passes per 3-D transform and direction
    opts = {k: g.get_option(k) for k in ("lds_cap", "grid_mult", "nt", "dual", "lanes", "stage", "contig_nt")}
    g.destroy()
    return {"workload": "64 C2C transforms, N in {96,128,160}^3 (22/21/21) complex64, fwd+bwd, grouped, hipGraph",
            "ms": round(ms, 4), "launches": launches, "options": opts,
            "avg_pass_GBs": round(pass_bytes * passes / (ms * 1e-3) / 1e9, 1),
            "roundtrip_rel_l2_max": err}


def bench_r2c_1024(torch, dfft, dev, steps, warmup, tune):
    """1024^3 R2C forward + C2R backward, slab (4,1,1) -> (1,4,1), FP64 and FP32,
    U[-1,1] real inputs; per-pass algorithmic bytes: R2C / C2R passes move the
    real array and the half spectrum, C2C passes read and write the half
    spectrum."""
    n, P_ = 1024, 4
    gin, gout = [n, n, n], [n, n, n // 2 + 1]
    res = {"workload": "3D R2C/C2R FFT 1024^3, slab (4,1,1)->(1,4,1), 4 virtual ranks, fwd+bwd"}
    for tag, rdt, cdt in (("fp64", torch.float64, torch.complex128), ("fp32", torch.float32, torch.complex64)):
        log(f"[bench] r2c 1024^3 {tag} ...")
        gen = torch.Generator(device=dev).manual_seed(7)

        def inputs():
            return [torch.rand(dfft.compute_local_extents_and_starts(gin, [P_, 1, 1], r)[0], generator=gen,
                               device=dev, dtype=rdt) * 2 - 1 for r in range(P_)]
        ins = inputs()
        outs = [torch.empty(dfft.compute_local_extents_and_starts(gout, [1, P_, 1], r)[0], dtype=cdt, device=dev)
                for r in range(P_)]
        plan = dfft.Plan(ins, outs, [0, 1, 2], [P_, 1, 1], [1, P_, 1], global_in_extents=gin,
                         global_out_extents=gout)
        if tune:
            t0 = time.perf_counter()
            plan.autotune(ins, outs)
            torch.cuda.synchronize()
            log(f"[bench] r2c {tag}: autotune {time.perf_counter() - t0:.1f} s")
            del ins
            ins = inputs()
        backs = [torch.empty_like(a) for a in ins]
        ms = time_roundtrip(torch, dfft, plan, ins, outs, backs, steps, warmup)
        fw, bw = pass_times(dfft, plan, ins, outs, backs, max(2, steps // 2))
        err = math.sqrt(sum(float(torch.linalg.vector_norm(b - a)) ** 2 for a, b in zip(ins, backs)) /
                        sum(float(torch.linalg.vector_norm(a)) ** 2 for a in ins))
        rs = 8 if tag == "fp64" else 4
        real_b = n ** 3 * rs
        half_b = n * n * (n // 2 + 1) * 2 * rs
        # forward: R2C, C2C, C2C; backward: C2C, C2C, C2R
        pass_bytes = [real_b + half_b, 2 * half_b, 2 * half_b, 2 * half_b, 2 * half_b, half_b + real_b]
        allp = fw + bw
        achieved = sum(pass_bytes) / (sum(allp) * 1e-3) / 1e9
        res[tag] = {"ms": round(ms, 3), "fwd_pass_ms": [round(x, 4) for x in fw],
                    "bwd_pass_ms": [round(x, 4) for x in bw],
                    "avg_pass_GBs": round(achieved, 1),
                    "pass_efficiency": round(achieved / HBM_PEAK_GBS, 4),
                    "roofline_2B": round(2 * (real_b + half_b) / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                    "roundtrip_rel_l2": err}
        log(f"[bench] r2c {tag}: {ms:.3f} ms, passes {fw} {bw}, {achieved:.0f} GB/s per pass")
        del plan, ins, outs, backs
        torch.cuda.empty_cache()
    res["fp32_speedup"] = round(res["fp64"]["ms"] / res["fp32"]["ms"], 3)
    return res


def bench_precision(torch, dfft, dev):
    """FP32 vs FP64 (both on the GPU; the FP64 path agrees with the CPU oracle to 1e-12
    wherever the oracle runs) and vs the exact spectrum of a single mode."""
    out = {}
    for tag, n, real in (("c2c_512", 512, False), ("r2c_1024", 1024, True)):
        gin = [n] * 3
        res = {}
        for field in ("uniform", "sin_k357"):
            spectra = {}
            for fp64 in (True, False):
                rdt = torch.float64 if fp64 else torch.float32
                cdt = torch.complex128 if fp64 else torch.complex64
                gen = torch.Generator(device=dev).manual_seed(11)
                ins, outs = [], []
                for r in range(P):
                    ext, st = dfft.compute_local_extents_and_starts(gin, [P, 1, 1], r)
                    if field == "uniform":
                        re = torch.rand(ext, generator=gen, device=dev, dtype=torch.float64) * 2 - 1
                        x = re if real else torch.complex(re, torch.rand(ext, generator=gen, device=dev,
                                                                        dtype=torch.float64) * 2 - 1)
                    else:
                        ii = [torch.arange(st[a], st[a] + ext[a], device=dev, dtype=torch.float64)
                              for a in range(3)]
                        X, Y, Z = torch.meshgrid(*ii, indexing="ij")
                        x = torch.sin(2 * math.pi * (3 * X + 5 * Y + 7 * Z) / n)
                        if not real:
                            x = torch.complex(x, torch.zeros_like(x))
                    ins.append(x.to(rdt if real else cdt).contiguous())
                    oe, _ = dfft.compute_local_extents_and_starts([n, n, n // 2 + 1] if real else gin, [1, P, 1], r)
                    outs.append(torch.empty(oe, dtype=cdt, device=dev))
                gout = [n, n, n // 2 + 1] if real else gin
                plan = dfft.Plan(ins, outs, [0, 1, 2], [P, 1, 1], [1, P, 1], global_in_extents=gin,
                                 global_out_extents=gout)
                dfft.execute(plan, ins, outs, dfft.Direction.forward)
                torch.cuda.synchronize()
                spectra[fp64] = [o.to(torch.complex128) for o in outs]
                del plan, ins, outs
                torch.cuda.empty_cache()
            num = sum(float(torch.linalg.vector_norm(a - b)) ** 2 for a, b in zip(spectra[False], spectra[True]))
            den = sum(float(torch.linalg.vector_norm(b)) ** 2 for b in spectra[True])
            res[field] = {"fp32_vs_fp64": math.sqrt(num / den)}
            if field == "sin_k357":
                # exact: n^3/(2i) at k and -n^3/(2i) at -k (the R2C half holds only k = (3,5,7))
                for fp64 in (True, False):
                    num = den = 0.0
                    for r, blk in enumerate(spectra[fp64]):
                        ext, st = dfft.compute_local_extents_and_starts(gout, [1, P, 1], r)
                        want = torch.zeros_like(blk)
                        for sgn in (1, -1):
                            k = [(sgn * 3) % n, (sgn * 5) % n, (sgn * 7) % n]
                            if real and sgn < 0:
                                continue
                            loc = [k[a] - st[a] for a in range(3)]
                            if all(0 <= loc[a] < ext[a] for a in range(3)):
                                want[loc[0], loc[1], loc[2]] = complex(0.0, -sgn * n ** 3 / 2)
                        num += float(torch.linalg.vector_norm(blk - want)) ** 2
                        den += float(torch.linalg.vector_norm(want)) ** 2
                    res[field]["fp64_vs_exact" if fp64 else "fp32_vs_exact"] = math.sqrt(num / den)
            del spectra
            torch.cuda.empty_cache()
        out[tag] = res
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-tune", action="store_true", help="skip the plan autotuner")
    ap.add_argument("--only", choices=["fp64", "fp32"], default=None,
                    help="run one precision only (profiling runs)")
    ap.add_argument("--sections", default="fft,r2c,ns,adjoint,batch,precision,rocfft",
                    help="comma list of: fft (headline), r2c, ns, adjoint, batch, precision, rocfft")
    ap.add_argument("--copy-probe", action="store_true",
                    help="only run a plain 1 GiB device-to-device copy (counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: the P virtual ranks share one MI355X; runs on more "
              "than one GPU are not measured by this benchmark", file=sys.stderr)
        sys.exit(2)

    import torch
    import dfft

    dev = torch.device("cuda:0")
    if args.copy_probe:
        a = torch.ones(1 << 27, dtype=torch.float64, device=dev)  # 1 GiB
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes": a.numel() * 8, "launches": 3}))
        return
    sections = set(args.sections.split(","))
    if "fft" not in sections:
        # profiling runs of single sections: no headline line
        line = {"sections": sorted(sections)}
        run_sections(torch, dfft, dev, args, sections, line)
        print(json.dumps(line), flush=True)
        return
    results = {}
    precisions = (("fp64", torch.complex128), ("fp32", torch.complex64))
    if args.only:
        precisions = tuple(p for p in precisions if p[0] == args.only)
    for tag, cdt in precisions:
        log(f"[bench] {tag}: allocating 512^3 slab P=4")
        ins, outs = make_inputs(torch, dfft, dev, cdt, seed=3)
        backs = [torch.empty_like(a) for a in ins]
        plan = dfft.Plan(ins, outs, [0, 1, 2], [P, 1, 1], [1, P, 1])
        if not args.no_tune:
            t0 = time.perf_counter()
            plan.autotune(ins, outs)  # overwrites ins/outs: inputs are regenerated below
            torch.cuda.synchronize()
            log(f"[bench] {tag}: autotune {time.perf_counter() - t0:.2f} s -> {plan.options()}")
            del ins, outs
            ins, outs = make_inputs(torch, dfft, dev, cdt, seed=3)
        ms = time_roundtrip(torch, dfft, plan, ins, outs, backs, args.steps, args.warmup)
        fw, bw = pass_times(dfft, plan, ins, outs, backs, max(2, args.steps // 2))
        log(f"[bench] {tag}: {ms:.3f} ms per fwd+bwd; passes fwd {fw} bwd {bw}")
        # round-trip accuracy on the resident data
        err = math.sqrt(sum(float(torch.linalg.vector_norm(b - a)) ** 2 for a, b in zip(ins, backs)) /
                        sum(float(torch.linalg.vector_norm(a)) ** 2 for a in ins))
        results[tag] = {"ms": ms, "fwd_pass_ms": fw, "bwd_pass_ms": bw, "roundtrip_rel_l2": err,
                        "workspace_bytes": plan.workspace_bytes(), "options": plan.options()}
        del plan, ins, outs, backs
        torch.cuda.empty_cache()

    if args.only:
        print(json.dumps({"only": args.only, **{k: v for k, v in results[args.only].items()}}), flush=True)
        return
    esz = 16
    B = 2 * N ** 3 * esz  # input + output array of one transform (bytes)
    r64 = results["fp64"]
    allp = r64["fwd_pass_ms"] + r64["bwd_pass_ms"]
    avg_launch = sum(allp) / len(allp)
    launches = len(r64["fwd_pass_ms"])  # line-pass launches per transform (one per axis)
    # algorithmic bytes per launch: the transform's compulsory bytes B (input +
    # output array) shared by its launches, so achieved / peak is the
    # round-trip figure 2B / t over the passes' own time
    alg_per_launch = B / launches
    achieved = alg_per_launch / avg_launch / 1e6  # GB/s
    pass_bytes = 2 * N ** 3 * esz  # each launch reads and writes the whole array
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4),
                # HBM bytes per launch: FETCH_SIZE x2 + WRITE_SIZE x1 (copy-calibrated), rocprofv3 --pmc
                "traffic": TRAFFIC_BYTES_PER_LAUNCH, "traffic_source": TRAFFIC_SOURCE,
                "algorithmic_bytes_per_launch": alg_per_launch,
                "algorithmic_bytes_per_transform": B, "launches_per_transform": launches,
                "kernel": "fft_line_fast<double,512,R,NT> (the six axis-pass launches of the round trip)",
                "avg_launch_ms": round(avg_launch, 4), "slowest_launch_ms": round(max(allp), 4),
                # each launch's own read + write of the whole array over its duration
                "pass_efficiency": round(pass_bytes / avg_launch / 1e6 / HBM_PEAK_GBS, 4),
                "sweep_cap": round(1.0 / launches, 4),
                # measured HBM bytes of a launch over its compulsory share: one
                # full sweep per axis pass (3.00 for 3 launches per transform)
                "traffic_over_algorithmic": round(TRAFFIC_BYTES_PER_LAUNCH / alg_per_launch, 4)}
    line = {
        "metric": "3D FFT fwd+bwd ms & %HBM roofline @512^3 FP64 vs FP32; NS adjoint steps/s",
        "value": round(r64["ms"], 4),
        "unit": "ms per 512^3 complex128 fwd+bwd round trip",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(r64["ms"], 4),
        "higher_is_better": False, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64", "data": "synthetic U[-1,1] seeded",
        "config": {"workload": "3D C2C FFT 512^3 complex128, slab (4,1,1)->(1,4,1), 4 virtual ranks, fwd+bwd",
                   "global_extents": [N] * 3, "virtual_ranks": P},
        "roofline_2B": round(2 * B / (r64["ms"] * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
        "fp32_ms": round(results["fp32"]["ms"], 4),
        "fp32_speedup": round(r64["ms"] / results["fp32"]["ms"], 3),
        "roundtrip_rel_l2": {"fp64": r64["roundtrip_rel_l2"], "fp32": results["fp32"]["roundtrip_rel_l2"]},
        "pass_ms": {"fp64": {"fwd": r64["fwd_pass_ms"], "bwd": r64["bwd_pass_ms"]},
                    "fp32": {"fwd": results["fp32"]["fwd_pass_ms"], "bwd": results["fp32"]["bwd_pass_ms"]}},
        "roofline": roofline,
    }
    run_sections(torch, dfft, dev, args, sections, line)
    if "rocfft" in sections:
        log("[bench] rocFFT comparators ...")
        line["vendor_plan"] = vendor_plan(torch, dfft, dev, max(3, args.steps // 2), 1)
        line["rocfft_comparator"] = rocfft_comparator(torch, dev, max(3, args.steps // 2))
        torch.cuda.empty_cache()
    if not args.no_cpu_baseline:
        log("[bench] cpu baseline (oracle) ...")
        line["cpu_baseline"] = cpu_baseline()
        line["cpu_library_baseline"] = cpu_library_baseline()
    print(json.dumps(line), flush=True)


def run_sections(torch, dfft, dev, args, sections, line):
    if "ns" in sections:
        log("[bench] NS Taylor-Green 256^3 ...")
        line["ns"] = bench_ns(torch, max(3, args.steps))
        torch.cuda.empty_cache()
    if "adjoint" in sections:
        log("[bench] NS adjoint 128^3, 20 steps ...")
        line["ns_adjoint"] = bench_ns_adjoint(torch, dev)
        torch.cuda.empty_cache()
    if "batch" in sections:
        log("[bench] mixed batch of 64 transforms ...")
        line["mixed_batch"] = bench_mixed_batch(torch, dfft, dev, args.steps, args.warmup)
        torch.cuda.empty_cache()
    if "r2c" in sections:
        line["r2c_1024"] = bench_r2c_1024(torch, dfft, dev, max(3, args.steps // 2), max(1, args.warmup),
                                          not args.no_tune)
        torch.cuda.empty_cache()
    if "precision" in sections:
        log("[bench] precision FP32 vs FP64 ...")
        line["precision"] = bench_precision(torch, dfft, dev)
        torch.cuda.empty_cache()


if __name__ == "__main__":
    main()
