#!/usr/bin/env python3
"""Benchmark of the batched mixed-radix FFT path on one MI355X.

Workload (one "step"): forward transforms of
    n = 2520 (2^3 3^2 5 7) x batch 4096  and  n = 2^20 x batch 64,
each in ComplexF32 and ComplexF64, contiguous rows, re/im i.i.d. N(0,1) (synthetic, seeded),
inputs resident in HBM before the timed region.  Throughput = sum of 5 n log2 n over the step
/ time (GFLOP/s); algorithmic HBM bytes = 2 n b per transform (b = 8 / 16 bytes).

Prints one JSON line.  --gpus must be 1: every transform of the workload runs on one GPU.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "catabrafft.jl_amd"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))

HBM_PEAK_GBS = 8000.0
CASES = [  # (n, batch, dtype)
    (2520, 4096, np.complex64),
    (2520, 4096, np.complex128),
    (1 << 20, 64, np.complex64),
    (1 << 20, 64, np.complex128),
]


def flops(n, batch):
    return 5.0 * n * math.log2(n) * batch


def algo_bytes(n, batch, dt):
    return 2.0 * n * batch * np.dtype(dt).itemsize


def cpu_baseline(budget_s: float = 12.0):
    """The oracle (C restatement of CatabraFFT's algorithm, `port`) on the host cores, over a
    bounded sample of the same workload: rows of each case until the time budget is spent."""
    import oracle as O

    threads = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    threads = max(1, min(threads, 16))
    rng = np.random.Generator(np.random.PCG64(2))
    per_case = budget_s / len(CASES)
    tot_flops, tot_time, rows = 0.0, 0.0, []
    for n, batch, dt in CASES:
        prec = "f" if dt == np.complex64 else "d"
        plan = O.Plan(n, prec)
        b = max(threads, min(batch, 64 if n < 4096 else threads))
        x = (rng.standard_normal((b, n)) + 1j * rng.standard_normal((b, n))).astype(dt)
        plan.exec(x[:threads], threads=threads)  # warm-up
        done, t0 = 0, time.perf_counter()
        while time.perf_counter() - t0 < per_case:
            plan.exec(x, threads=threads)
            done += b
        el = time.perf_counter() - t0
        tot_flops += flops(n, done)
        tot_time += el
        rows.append(f"{n}x{done} {np.dtype(dt).name}")
    return {
        "value": round(tot_flops / tot_time / 1e9, 3),
        "unit": "GFLOP/s",
        "cores": threads,
        "kind": "port",
        "sample": "oracle (C restatement of CatabraFFT's algorithm, OpenMP over rows), "
                  + ", ".join(rows) + f" in {tot_time:.1f} s",
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cases", default="all", help="comma list of case indices")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this FFT engine runs on one GPU; the batched workload is measured at "
              "--gpus 1 only (transforms shard by rows across GPUs with no communication).",
              file=sys.stderr)
        sys.exit(2)

    import torch

    import catabrafft as C
    import oracle as O

    cases = CASES if args.cases == "all" else [CASES[int(i)] for i in args.cases.split(",")]
    dev = torch.device("cuda:0")
    stream = torch.cuda.current_stream(dev)
    plans, xs, ys = [], [], []
    g = torch.Generator(device=dev)
    g.manual_seed(2)
    for n, batch, dt in cases:
        tdt = torch.complex64 if dt == np.complex64 else torch.complex128
        rdt = torch.float32 if dt == np.complex64 else torch.float64
        x = torch.complex(torch.randn(batch, n, generator=g, device=dev, dtype=rdt),
                          torch.randn(batch, n, generator=g, device=dev, dtype=rdt))
        assert x.dtype == tdt
        xs.append(x)
        ys.append(torch.empty_like(x))
        plans.append(C.Plan(n, batch, dt))
    h = stream.cuda_stream

    def step(evs=None, k=0):
        for i, p in enumerate(plans):
            if evs is not None:
                evs[i][0][k].record(stream)
            p.exec_ptr(xs[i].data_ptr(), ys[i].data_ptr(), -1, False, h)
            if evs is not None:
                evs[i][1][k].record(stream)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    evs = [[[torch.cuda.Event(enable_timing=True) for _ in range(args.steps)] for _ in range(2)]
           for _ in plans]
    t_start = torch.cuda.Event(enable_timing=True)
    t_end = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    t_start.record(stream)
    for k in range(args.steps):
        step(evs, k)
    t_end.record(stream)
    torch.cuda.synchronize()
    total_ms = t_start.elapsed_time(t_end)
    ms_per_step = total_ms / args.steps

    # per-case times (HIP events on the stream the kernels run on)
    per_case = []
    for i, (n, batch, dt) in enumerate(cases):
        ms = [evs[i][0][k].elapsed_time(evs[i][1][k]) for k in range(args.steps)]
        per_case.append(float(np.mean(ms)))

    # accuracy per case on sampled rows vs the oracle in long double (outside the timed region)
    errs = []
    for i, (n, batch, dt) in enumerate(cases):
        rows = [0, batch - 1]
        xr = xs[i][rows].cpu().numpy().astype(np.complex128)
        yr = ys[i][rows].cpu().numpy().astype(np.complex128)
        ref = O.fft_l(xr) if n <= 4096 else O.fft(xr, "d", threads=8)
        errs.append(float(np.linalg.norm(yr - ref) / np.linalg.norm(ref)))

    tot_flops = sum(flops(n, b) for n, b, _ in cases)
    value = tot_flops / (ms_per_step * 1e-3) / 1e9
    breakdown = []
    for (n, batch, dt), ms, e, p in zip(cases, per_case, errs, plans):
        gbs = algo_bytes(n, batch, dt) / (ms * 1e-3) / 1e9
        breakdown.append({
            "n": n, "batch": batch, "dtype": np.dtype(dt).name, "ms": round(ms, 4),
            "gflops": round(flops(n, batch) / (ms * 1e-3) / 1e9, 1),
            "algo_GBs": round(gbs, 1), "hbm_frac": round(gbs / HBM_PEAK_GBS, 4),
            "rel_l2_vs_oracle": e, "plan": p.describe(),
        })
    dom = int(np.argmax(per_case))
    n_d, b_d, dt_d = cases[dom]
    achieved = algo_bytes(n_d, b_d, dt_d) / (per_case[dom] * 1e-3) / 1e9
    out = {
        "metric": "GFLOP/s (5N·log2N) & % HBM peak, batched FP32 vs FP64; serving p50/p99 latency",
        "value": round(value, 1),
        "unit": "GFLOP/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64",
        "data": "synthetic: re/im i.i.d. N(0,1), torch PCG seed 2, resident in HBM",
        "config": {"workload": "batched mixed-radix C2C FFT: n=2520 x 4096 and n=2^20 x 64, "
                               "ComplexF32 and ComplexF64, contiguous rows, forward",
                   "global_batch": sum(b for _, b, _ in cases), "parallelism": "single GPU"},
        "roofline": {
            "bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
            "kernel": f"n={n_d} x {b_d} {np.dtype(dt_d).name}: {plans[dom].describe()}",
        },
        "breakdown": breakdown,
    }
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline()
    print(json.dumps(out))


if __name__ == "__main__":
    main()
