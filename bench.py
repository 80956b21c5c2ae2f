"""Benchmark: 3-D C2C FFT forward+backward round trip on one MI355X.

Workload (one "step"): 512^3 complex128, slab decomposition over 4 virtual
ranks, in-topology (4,1,1) -> out-topology (1,4,1), axes (0,1,2), forward
then backward with the default backward normalisation; the same round trip in
complex64 is timed beside it.  Inputs: U[-1,1] real and imaginary parts,
seeded, resident in HBM before the timed region.

Prints one JSON line.  `value` is the FP64 round-trip time in ms (lower is
better).  `roofline` is the dominant kernel's algorithmic bytes (input +
output array of the pass) over its average duration, timed with HIP events
on the launch stream; `cpu_baseline` times the CPU restatement of the plan
(oracle, single thread for the FFT lines) on a bounded sample and scales it.
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


def cpu_baseline():
    """Oracle (CPU restatement of the slab plan, P=4 virtual ranks) on a
    bounded sample: one 128^3 complex128 forward+backward, scaled to 512^3 by
    the N log N operation count."""
    import numpy as np
    from oracle import plan_exec as PE

    n = 256
    g = [n] * 3
    rng = np.random.default_rng(3)
    x = rng.uniform(-1, 1, g) + 1j * rng.uniform(-1, 1, g)
    plan = PE.OraclePlan(g, g, [0, 1, 2], [P, 1, 1], [1, P, 1])
    ins = PE.scatter(x, g, [P, 1, 1])
    t0 = time.perf_counter()
    outs = plan.forward(ins)
    plan.backward(outs)
    dt = time.perf_counter() - t0
    scale = (N ** 3 * math.log2(N ** 3)) / (n ** 3 * math.log2(n ** 3))
    return {"value": dt * 1e3 * scale, "unit": "ms per 512^3 complex128 fwd+bwd (extrapolated)",
            "cores": 1, "kind": "port",
            "sample": f"oracle slab plan P=4, one {n}^3 complex128 fwd+bwd in {dt:.2f} s, "
                      f"scaled x{scale:.1f} by N log N"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-tune", action="store_true", help="skip the plan autotuner")
    ap.add_argument("--only", choices=["fp64", "fp32"], default=None,
                    help="run one precision only (profiling runs)")
    ap.add_argument("--copy-probe", action="store_true",
                    help="only run a plain 1 GiB device-to-device copy (counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: the P virtual ranks share one MI355X; runs on more "
              "than one GPU are not part of this build", file=sys.stderr)
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
    dom = max(allp)
    pass_bytes = 2 * N ** 3 * esz  # each pass reads and writes the whole array
    achieved = pass_bytes / (sum(allp) / len(allp)) / 1e6  # GB/s, average pass
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                "kernel": "fft_line_fast<double,512> (all three axis passes averaged)",
                "dominant_pass_ms": round(dom, 4)}
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
    if not args.no_cpu_baseline:
        log("[bench] cpu baseline (oracle) ...")
        line["cpu_baseline"] = cpu_baseline()
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
