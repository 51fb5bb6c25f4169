#!/usr/bin/env python3
"""Benchmark of the batched mixed-radix FFT path on one MI355X.

Workload (one "step"): forward transforms of
    n = 2520 (2^3 3^2 5 7) x batch 4096  and  n = 2^20 x batch 64,
each in ComplexF32 and ComplexF64, contiguous rows, re/im i.i.d. N(0,1) (synthetic, seeded),
inputs resident in HBM before the timed region; plans made with the MEASURE flag (the radix
sequence and kernel shapes timed on the device at plan time, outside the timed region; --flags
none gives the default heuristics).  Throughput = sum of 5 n log2 n over the step / time
(GFLOP/s); algorithmic HBM bytes = 2 n b per transform (b = 8 / 16 bytes).

The same line carries the rest of the metric, each measured in this run: `serving` (a Poisson
trace of mixed requests through the continuous-batching server: p50 / p99 latency, throughput,
and the speed-up over one ComplexF64 launch per transform issued from C), `rader` (the five
Rader primes at 2^24 elements each) and `uneven` (10,000 transforms of uneven lengths in one
plan), plus `cpu_baseline` (the oracle on the host, one thread and all of this job's threads).

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


# planner flags of the batched plans (spells.jl:4-8): MEASURE times candidate radix sequences
# and kernel shapes on the device when the plan is made, outside the timed region
FLAGS = {"none": 0, "measure": 1, "enchant": 2}


def flops(n, batch):
    return 5.0 * n * math.log2(n) * batch


def algo_bytes(n, batch, dt):
    return 2.0 * n * batch * np.dtype(dt).itemsize


def _host_threads():
    """Threads the CPU leg may use: the job's CPU affinity, capped by OMP_NUM_THREADS when the
    environment sets it (a GPU job is given a share of the host's cores)."""
    aff = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    omp = os.environ.get("OMP_NUM_THREADS")
    if omp and omp.isdigit() and int(omp) > 0:
        return max(1, min(aff, int(omp)))
    return max(1, aff)


def _cpu_leg(threads, budget_s, rng):
    """The oracle over rows of each batched case until budget_s is spent: (GFLOP/s, sample)."""
    import oracle as O

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
    return tot_flops / tot_time / 1e9, ", ".join(rows) + f" in {tot_time:.1f} s"


def cpu_baseline(budget_s: float = 20.0):
    """The oracle (C restatement of CatabraFFT's algorithm, `port`) on the host, over a bounded
    sample of the same workload, in the two modes of the CPU plan: one thread (the reference's
    own benchmark method, julia --threads=1) and every thread this job may use (OpenMP over
    rows).  `value` is the multi-thread leg."""
    threads = _host_threads()
    rng = np.random.Generator(np.random.PCG64(2))
    single, s_sample = _cpu_leg(1, budget_s / 2, rng)
    multi, m_sample = _cpu_leg(threads, budget_s / 2, rng)
    return {
        "value": round(multi, 3),
        "unit": "GFLOP/s",
        "cores": threads,
        "kind": "port",
        "sample": "oracle (C restatement of CatabraFFT's algorithm, OpenMP over rows), "
                  + m_sample,
        "single_thread": {"value": round(single, 3), "cores": 1, "sample": s_sample},
        "host_logical_cpus": os.cpu_count(),
        "cores_note": "threads = this job's CPU affinity capped by OMP_NUM_THREADS (the host's "
                      "share for one GPU); host_logical_cpus is the whole machine",
    }


def measured_traffic(plan_desc: str):
    """HBM bytes per step of a case as measured by rocprofv3 FETCH_SIZE / WRITE_SIZE (scaled by a
    plain copy) for this exact plan, from profiles/batched_traffic.json; None when the plan was
    not profiled."""
    path = os.path.join(ROOT, "profiles", "batched_traffic.json")
    try:
        with open(path) as f:
            table = json.load(f)["cases"]
    except (OSError, ValueError, KeyError):
        return None
    key = plan_desc.split(" tuned(")[0].split(" (wisdom)")[0]  # the kernel configuration
    entry = table.get(key)
    return entry["bytes_per_step"] if entry else None


def bench_batched(args):
    import torch

    import catabrafft as C
    import oracle as O

    cases = CASES if args.cases == "all" else [CASES[int(i)] for i in args.cases.split(",")]
    dev = torch.device("cuda:0")
    stream = torch.cuda.current_stream(dev)
    plans, xs, ys = [], [], []
    g = torch.Generator(device=dev)
    g.manual_seed(2)
    if args.wisdom:  # plans of an earlier run (counter runs profile the same kernels)
        C.import_wisdom(args.wisdom)
    for n, batch, dt in cases:
        tdt = torch.complex64 if dt == np.complex64 else torch.complex128
        rdt = torch.float32 if dt == np.complex64 else torch.float64
        x = torch.complex(torch.randn(batch, n, generator=g, device=dev, dtype=rdt),
                          torch.randn(batch, n, generator=g, device=dev, dtype=rdt))
        assert x.dtype == tdt
        xs.append(x)
        ys.append(torch.empty_like(x))
        if args.fields:  # one configuration forced (counter runs of alternative plans)
            plans.append(_forced_plan(C, n, batch, dt, args.fields))
        else:
            plans.append(C.Plan(n, batch, dt, flags=FLAGS[args.flags]))
    if args.save_wisdom:
        C.export_wisdom(args.save_wisdom)
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
    traffic = measured_traffic(plans[dom].describe())
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
                   "plan_flags": args.flags,
                   "global_batch": sum(b for _, b, _ in cases), "parallelism": "single GPU"},
        "roofline": {
            "bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": traffic,
            "traffic_vs_algorithmic": (round(traffic / algo_bytes(n_d, b_d, dt_d), 3)
                                       if traffic else None),
            "kernel": f"n={n_d} x {b_d} {np.dtype(dt_d).name}: {plans[dom].describe()}",
            "per": "one step of the case: every launch of its plan (achieved, traffic in bytes)",
        },
        "breakdown": breakdown,
    }
    # the rest of the metric, each in its own key: serving p50 / p99 and the speed-up over one
    # FP64 launch per transform, batched Rader primes, the uneven batch
    del xs, ys, plans
    torch.cuda.empty_cache()
    sections = {"serving": serving_section, "rader": rader_section, "uneven": uneven_section}
    for name in [x for x in args.sections.split(",") if x and x != "none"]:
        t0 = time.perf_counter()
        print(f"bench: {name} section", file=sys.stderr, flush=True)
        out[name] = sections[name](args)
        out[name]["section_seconds"] = round(time.perf_counter() - t0, 1)
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline()
    print(json.dumps(out))


# ----------------------------------------------------------------------------- other configs
def _time_ms(fn, steps, warmup):
    import torch

    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    a.record()
    for _ in range(steps):
        fn()
    b.record()
    torch.cuda.synchronize()
    return a.elapsed_time(b) / steps


def bench_single(args):
    """n = 1024 ComplexF64, one transform: forward + inverse latency on the device and through
    the host-pointer ABI (the Julia drop-in's path, PCIe included)."""
    import torch

    import catabrafft as C
    import oracle as O

    n = 1024
    rng = np.random.Generator(np.random.PCG64(1))
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    xd = torch.from_numpy(x).cuda()
    yd, zd = torch.empty_like(xd), torch.empty_like(xd)
    p = C.Plan(n, 1, np.complex128)
    s = torch.cuda.current_stream().cuda_stream

    def fwd_inv():
        p.exec_ptr(xd.data_ptr(), yd.data_ptr(), -1, False, s)
        p.exec_ptr(yd.data_ptr(), zd.data_ptr(), 1, True, s)

    dev_us = _time_ms(fwd_inv, args.steps * 50, args.warmup) * 1e3
    t0 = time.perf_counter()
    reps = 200
    for _ in range(reps):
        y = C.fft(x)
        z = C.ifft(y)
    host_us = (time.perf_counter() - t0) / reps * 1e6
    err = float(np.linalg.norm(y - O.fft_l(x)) / np.linalg.norm(O.fft_l(x)))
    rt = float(np.linalg.norm(z - x) / np.linalg.norm(x))
    plan = O.Plan(n, "d")
    t0 = time.perf_counter()
    for _ in range(2000):
        plan.exec(x, -1)
        plan.exec(x, 1, True)
    cpu_us = (time.perf_counter() - t0) / 2000 * 1e6
    print(json.dumps({"workload": "n=1024 ComplexF64 forward+inverse, single transform",
                      "device_us_per_fwd_inv": round(dev_us, 2),
                      "host_abi_us_per_fwd_inv": round(host_us, 1),
                      "cpu_oracle_us_per_fwd_inv_1thread": round(cpu_us, 2),
                      "rel_l2_vs_oracle": err, "round_trip": rt, "plan": p.describe()}))


RADER_PRIMES = (17, 257, 4099, 65537, 1048573)


def _sync_ms(torch, fn, steps, warmup, stream):
    """Mean time of fn() over `steps` launches after `warmup`, by HIP events on `stream`."""
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    a.record(stream)
    for _ in range(steps):
        fn()
    b.record(stream)
    torch.cuda.synchronize()
    return a.elapsed_time(b) / steps


def rader_section(args):
    """Batched Rader primes p in {17, 257, 4099, 65537, 1048573}, ComplexF32 (FP64 convolution for
    p > 2^16), batch floor(2^24 / p), re/im U(-1, 1); two rows per prime against the oracle."""
    import torch

    import catabrafft as C
    import oracle as O

    rows, tot_t, tot_b, tot_f = [], 0.0, 0.0, 0.0
    stream = torch.cuda.current_stream()
    for p_ in RADER_PRIMES:
        batch = (1 << 24) // p_
        g = torch.Generator(device="cuda")
        g.manual_seed(3)
        x = torch.complex(torch.rand(batch, p_, generator=g, device="cuda") * 2 - 1,
                          torch.rand(batch, p_, generator=g, device="cuda") * 2 - 1)
        y = torch.empty_like(x)
        plan = C.Plan(p_, batch, np.complex64)
        h = stream.cuda_stream
        ms = _sync_ms(torch, lambda: plan.exec_ptr(x.data_ptr(), y.data_ptr(), -1, False, h),
                      args.steps, args.warmup, stream)
        sample = [0, batch // 2, batch - 1] if batch > 2 else list(range(batch))
        xr = x[sample].cpu().numpy().astype(np.complex128)
        ref = O.fft(xr, "d", threads=8) if p_ > 4096 else O.fft_l(xr)
        err = float(np.linalg.norm(y[sample].cpu().numpy() - ref) / np.linalg.norm(ref))
        gbs = 2 * p_ * batch * 8 / (ms * 1e-3) / 1e9
        rows.append({"p": p_, "batch": batch, "ms": round(ms, 4),
                     "gflops": round(flops(p_, batch) / (ms * 1e-3) / 1e9, 1),
                     "algo_GBs": round(gbs, 1), "hbm_frac": round(gbs / HBM_PEAK_GBS, 4),
                     "rel_l2_vs_oracle": err, "plan": plan.describe()})
        tot_t += ms
        tot_b += 2 * p_ * batch * 8
        tot_f += flops(p_, batch)
        del x, y, plan
        torch.cuda.empty_cache()
    return {"workload": "batched Rader primes, ComplexF32 (FP64 convolution for p > 2^16), "
                        "2^24 elements per prime, re/im U(-1,1) seed 3",
            "total_ms": round(tot_t, 4), "gflops": round(tot_f / (tot_t * 1e-3) / 1e9, 1),
            "hbm_frac": round(tot_b / (tot_t * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
            "max_rel_l2_vs_oracle": max(r["rel_l2_vs_oracle"] for r in rows),
            "rel_l2_bound": 1e-6, "cases": rows}


def bench_rader(args):
    print(json.dumps(rader_section(args)))


def cfg4_lengths(count=10000, seed=4, lo=64, hi=1 << 18):
    """L = round(exp(U(ln lo, ln hi))), snapped to the nearest 7-smooth (70%), non-7-smooth
    composite (20%) or prime (10%) length."""
    import oracle as O

    rng = np.random.Generator(np.random.PCG64(seed))

    def smooth7(m):
        for q in (2, 3, 5, 7):
            while m % q == 0:
                m //= q
        return m == 1

    def nearest(L, pred):
        d = 0
        while True:
            for c in (L - d, L + d):
                if 2 <= c and pred(c):
                    return c
            d += 1

    out = []
    for _ in range(count):
        L = int(round(math.exp(rng.uniform(math.log(lo), math.log(hi)))))
        u = rng.uniform()
        if u < 0.7:
            out.append(nearest(L, smooth7))
        elif u < 0.9:
            out.append(nearest(L, lambda m: not smooth7(m) and not O.isprime(m)))
        else:
            out.append(nearest(L, O.isprime))
    return out


def uneven_section(args):
    """10,000 transforms of uneven lengths in one plan (work-list launches, hipGraph replay),
    ComplexF32; 12 sampled transforms against the oracle."""
    import torch

    import catabrafft as C
    import oracle as O

    lengths = cfg4_lengths()
    offs = np.concatenate([[0], np.cumsum(lengths)[:-1]]).astype(np.int64)
    total = int(sum(lengths))
    g = torch.Generator(device="cuda")
    g.manual_seed(4)
    x = torch.complex(torch.randn(total, generator=g, device="cuda"),
                      torch.randn(total, generator=g, device="cuda"))
    y = torch.empty_like(x)
    t0 = time.perf_counter()
    plan = C.UnevenPlan(lengths, offs, np.complex64)
    plan_s = time.perf_counter() - t0
    stream = torch.cuda.current_stream()
    h = stream.cuda_stream
    ms = _sync_ms(torch, lambda: plan.exec_ptr(x.data_ptr(), y.data_ptr(), -1, False, h),
                  args.steps, args.warmup, stream)
    fl = sum(5.0 * n * math.log2(n) for n in lengths)
    gbs = 2 * total * 8 / (ms * 1e-3) / 1e9
    rng = np.random.Generator(np.random.PCG64(7))
    errs = []
    xs, ys = x.cpu().numpy(), y.cpu().numpy()
    for i in rng.choice(len(lengths), 12, replace=False):
        n, o = lengths[i], offs[i]
        ref = O.fft(xs[o:o + n].astype(np.complex128), "d")
        errs.append(float(np.linalg.norm(ys[o:o + n] - ref) / np.linalg.norm(ref)))
    desc = plan.describe()
    del x, y, plan
    torch.cuda.empty_cache()
    return {"workload": "uneven batch: 10,000 lengths log-uniform in [64, 2^18] (70% 7-smooth, "
                        "20% other composite, 10% prime), ComplexF32, N(0,1) seed 4",
            "elements": total, "ms": round(ms, 4), "gflops": round(fl / (ms * 1e-3) / 1e9, 1),
            "algo_GBs": round(gbs, 1), "hbm_frac": round(gbs / HBM_PEAK_GBS, 4),
            "plan_seconds": round(plan_s, 2), "max_rel_l2_sampled": max(errs),
            "rel_l2_bound": 1e-6, "plan": desc}


def bench_uneven(args):
    print(json.dumps(uneven_section(args)))


def _one_per_launch_lib():
    """lib/libcfft_one_per_launch.so: the C loop of one cfft_exec per transform."""
    import ctypes

    import catabrafft._lib as L

    L.lib()  # the library it links against, loaded (and bound to torch's runtime) first
    path = os.path.join(os.path.dirname(L.LIB_PATH), "libcfft_one_per_launch.so")
    if not os.path.exists(path):
        raise OSError(f"{path} is missing: build with catabrafft.jl_amd/build.py")
    lib = ctypes.CDLL(path)
    P, vp, i64, ci = ctypes.POINTER, ctypes.c_void_p, ctypes.c_int64, ctypes.c_int
    lib.cfft_one_per_launch.argtypes = [P(vp), P(i64), P(i64), P(i64), i64, vp, vp, i64, ci, ci, vp]
    lib.cfft_one_per_launch.restype = ci
    return lib


def serving_section(args, nreq=None):
    """cfg5: Poisson arrivals at 20k req/s, lengths as the uneven batch, batch per request
    Geometric(p = 1/16) truncated to 1..256, 10% ComplexF64, 200 GB admission cap; requests read
    device-resident inputs.  Runs:
      offered  -- the trace replayed at its arrival times: p50 / p99 latency (submit to
                  completion) and throughput;
      capacity -- the same requests submitted back to back;
      fp64     -- the same requests all in ComplexF64, back to back (the precision of the
                  baseline);
      baseline -- every transform of the trace as its own FP64 launch (one batch-1 plan per
                  length), issued from C (lib/libcfft_one_per_launch.so), one after another on
                  one stream.
    Accuracy: two rows of sampled requests of each dtype against the oracle."""
    import ctypes

    import torch

    import catabrafft as C
    import oracle as O

    nreq = nreq or args.requests
    rng = np.random.Generator(np.random.PCG64(5))
    lengths = cfg4_lengths(count=nreq, seed=5)
    batches = np.minimum(np.maximum(rng.geometric(1 / 16, nreq), 1), 256).astype(np.int64)
    f64 = rng.uniform(size=nreq) < 0.1
    gaps = rng.exponential(1 / 20000.0, nreq)
    sizes = [int(n) * int(b) for n, b in zip(lengths, batches)]
    pool_elems = max(sizes)
    g = torch.Generator(device="cuda")
    g.manual_seed(5)
    pool32 = torch.complex(torch.randn(pool_elems, generator=g, device="cuda"),
                           torch.randn(pool_elems, generator=g, device="cuda"))
    pool64 = pool32.to(torch.complex128)
    ring = 4 * pool_elems
    out32 = torch.empty(ring, dtype=torch.complex64, device="cuda")
    out64 = torch.empty(ring, dtype=torch.complex128, device="cuda")
    # sampled requests (accuracy) get outputs of their own
    samp_rng = np.random.Generator(np.random.PCG64(55))
    idx32 = [i for i in range(nreq) if not f64[i]]
    idx64 = [i for i in range(nreq) if f64[i]]
    sampled = set(samp_rng.choice(idx32, min(8, len(idx32)), replace=False).tolist())
    sampled |= set(samp_rng.choice(idx64, min(8, len(idx64)), replace=False).tolist())
    kept = {}
    srv = C.Server(hbm_cap_bytes=200 * 10 ** 9, nstreams=4)

    def src_of(i, dbl):
        n, b = lengths[i], int(batches[i])
        return (pool64 if dbl else pool32)[:n * b].view(b, n)

    def run(paced, all64=False, keep=False):
        srv.reset_stats()
        tickets, ooff32, ooff64 = [], 0, 0
        t_next = time.perf_counter()
        for i in range(nreq):
            if paced:
                t_next += gaps[i]
                while time.perf_counter() < t_next:
                    pass
            n, b = lengths[i], int(batches[i])
            dbl = all64 or bool(f64[i])
            src = src_of(i, dbl)
            if keep and i in sampled:
                dst = torch.empty_like(src)
                kept[i] = (src, dst)
            elif dbl:
                if ooff64 + n * b > ring:
                    ooff64 = 0
                dst = out64[ooff64:ooff64 + n * b].view(b, n)
                ooff64 += n * b
            else:
                if ooff32 + n * b > ring:
                    ooff32 = 0
                dst = out32[ooff32:ooff32 + n * b].view(b, n)
                ooff32 += n * b
            tickets.append(srv.submit(src, out=dst))
        for t in tickets:
            srv.wait(t)
        return srv.stats()

    # the server's resident plans, made once per (length, dtype) before any timed run
    warm = []
    for n, d in sorted(set(zip(lengths, f64.tolist()))):
        warm.append(srv.submit((pool64 if d else pool32)[:n], out=(out64 if d else out32)[:n]))
    for n in sorted(set(lengths)):
        warm.append(srv.submit(pool64[:n], out=out64[:n]))
    for t in warm:
        srv.wait(t)
    offered = run(paced=True, keep=True)
    capacity = run(paced=False)
    cap64 = run(paced=False, all64=True)
    srv.close()
    transforms = int(batches.sum())
    # accuracy of the sampled requests
    worst = {"complex64": 0.0, "complex128": 0.0}
    for i, (src, dst) in kept.items():
        b = src.shape[0]
        rows = sorted({0, b - 1})
        xs = src[rows].cpu().numpy()
        k = xs.dtype.name
        ref = O.fft(xs.astype(np.complex128), "d", threads=8)
        e = float(np.linalg.norm(dst[rows].cpu().numpy() - ref) / np.linalg.norm(ref))
        worst[k] = max(worst[k], e)
    if args.no_baseline:  # (profiling the serving loop alone)
        return _serving_result(nreq, transforms, gaps, offered, capacity, cap64, worst, kept, None)
    # baseline: one FP64 launch per transform, from C
    loop = _one_per_launch_lib()
    plans = {n: C.Plan(n, 1, np.complex128) for n in sorted(set(lengths))}
    order = sorted(plans)
    handles = (ctypes.c_void_p * len(order))(*[plans[n].h.value for n in order])
    pos = {n: k for k, n in enumerate(order)}
    which, ioff, ooff = [], [], []
    o = 0
    for i in range(nreq):
        n, b = lengths[i], int(batches[i])
        if o + n * b > ring:
            o = 0
        for r in range(b):
            which.append(pos[n])
            ioff.append(r * n)
            ooff.append(o + r * n)
        o += n * b
    cnt = len(which)
    arr = lambda v: (ctypes.c_int64 * cnt)(*v)  # noqa: E731
    w_, i_, o_ = arr(which), arr(ioff), arr(ooff)
    stream = torch.cuda.current_stream()

    def base_run():
        rc = loop.cfft_one_per_launch(handles, w_, i_, o_, cnt, ctypes.c_void_p(pool64.data_ptr()),
                                      ctypes.c_void_p(out64.data_ptr()), 16, -1, 0,
                                      ctypes.c_void_p(stream.cuda_stream))
        assert rc == 0, rc

    base_run()  # warm-up
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    base_run()
    torch.cuda.synchronize()
    base_s = time.perf_counter() - t0
    base_tps = cnt / base_s
    del plans
    torch.cuda.empty_cache()
    base = {"transforms": cnt, "seconds": base_s, "transforms_per_s": base_tps}
    return _serving_result(nreq, transforms, gaps, offered, capacity, cap64, worst, kept, base)


def _serving_result(nreq, transforms, gaps, offered, capacity, cap64, worst, kept, base):
    """The serving section's JSON object (base: the one-per-launch baseline, or None)."""
    base_tps = base["transforms_per_s"] if base else None
    ratio = (lambda v: round(v / base_tps, 2)) if base else (lambda v: None)
    return {"workload": f"serving: {nreq} Poisson requests at 20k req/s (seed 5), lengths "
                        "log-uniform [64, 2^18], batch Geometric(1/16) in 1..256, 10% ComplexF64, "
                        "200 GB cap, device-resident I/O, 4 streams",
            "requests": nreq, "transforms": transforms,
            "p50_ms": round(offered["p50_s"] * 1e3, 4), "p99_ms": round(offered["p99_s"] * 1e3, 4),
            "transforms_per_s": round(offered["transforms_per_s"], 1),
            "offered_transforms_per_s": round(transforms / float(gaps.sum()), 1),
            "gflops": round(offered["gflops"], 1),
            "capacity_transforms_per_s": round(capacity["transforms_per_s"], 1),
            "capacity_gflops": round(capacity["gflops"], 1),
            "capacity_p50_ms": round(capacity["p50_s"] * 1e3, 4),
            "capacity_p99_ms": round(capacity["p99_s"] * 1e3, 4),
            "fp64_capacity_transforms_per_s": round(cap64["transforms_per_s"], 1),
            "fp64_capacity_gflops": round(cap64["gflops"], 1),
            "baseline": ({"what": "each transform of the trace as its own ComplexF64 cfft_exec "
                                  "(batch-1 plan per length), looped in C on one stream",
                          "transforms": base["transforms"], "seconds": round(base["seconds"], 4),
                          "transforms_per_s": round(base_tps, 1)} if base else None),
            "speedup_vs_baseline_offered": ratio(offered["transforms_per_s"]),
            "speedup_vs_baseline_capacity": ratio(capacity["transforms_per_s"]),
            "speedup_vs_baseline_fp64_equal_precision": ratio(cap64["transforms_per_s"]),
            "max_rel_l2_sampled": {k: v for k, v in worst.items()},
            "rel_l2_bounds": {"complex64": 1e-6, "complex128": "4 eps log2 n (<= 8e-15)"},
            "sampled_requests": len(kept)}


def bench_serve(args):
    print(json.dumps(serving_section(args)))


def bench_f16(args):
    """The batched cases in ComplexF16 storage (float arithmetic): half the bytes of F32.  The
    check is bitwise: the F16 output equals the F32 plan's output rounded to float16."""
    import torch

    import catabrafft as C

    s = torch.cuda.current_stream().cuda_stream
    out = []
    for n, batch in ((2520, 4096), (1 << 20, 64)):
        g = torch.Generator(device="cuda").manual_seed(5)
        x32 = torch.randn((batch, n, 2), generator=g, device="cuda", dtype=torch.float32)
        xh = x32.to(torch.float16)
        x32 = xh.to(torch.float32)  # the same values in both precisions
        yh, y32 = torch.empty_like(xh), torch.empty_like(x32)
        p16, p32 = C.Plan(n, batch, C.C16), C.Plan(n, batch, np.complex64)
        ms = _time_ms(lambda: p16.exec_ptr(xh.data_ptr(), yh.data_ptr(), -1, False, s),
                      args.steps, args.warmup)
        p32.exec_ptr(x32.data_ptr(), y32.data_ptr(), -1, False, s)
        torch.cuda.synchronize()
        same = bool(torch.equal(yh.view(torch.int16), y32.to(torch.float16).view(torch.int16)))
        gbs = 2 * n * batch * 4 / (ms * 1e-3) / 1e9
        out.append({"n": n, "batch": batch, "ms": round(ms, 4),
                    "gflops": round(flops(n, batch) / (ms * 1e-3) / 1e9, 1),
                    "algo_GBs": round(gbs, 1), "hbm_frac": round(gbs / 8000, 4),
                    "bits_equal_rounded_f32": same, "plan": p16.describe()})
    print(json.dumps({"workload": "batched ComplexF16 storage (float arithmetic), forward",
                      "cases": out}))


def _forced_plan(C, n, batch, dt, fields):
    """A plan of the given shape built from one wisdom line (the knobs `fields`), MEASURE flag."""
    import tempfile

    code = 0 if np.dtype(dt) == np.complex64 else 1
    fd, path = tempfile.mkstemp(suffix=".txt")
    os.close(fd)
    with open(path, "w") as f:
        f.write(f"{n} {batch} {n} {n} {code} 1 : {fields}\n")
    C.import_wisdom(path)
    os.unlink(path)
    return C.Plan(n, batch, dt, flags=C.MEASURE)


WAVE_CASES = [  # n, dtype, E, threads per workgroup, radix sequence ("" = default)
    (256, np.complex64, 16, 64, ""), (256, np.complex64, 16, 256, ""),
    (512, np.complex64, 8, 64, ""), (512, np.complex64, 8, 256, ""),
    (1024, np.complex64, 16, 64, ""), (1024, np.complex64, 16, 256, ""),
    (1024, np.complex64, 8, 128, ""), (1024, np.complex64, 8, 256, ""),
    (1024, np.complex128, 8, 128, ""), (1024, np.complex128, 8, 256, ""),
    (1024, np.complex128, 16, 64, ""), (1024, np.complex128, 16, 256, ""),
    (1024, np.complex64, 16, 64, "ra=4,16,16"), (257, np.complex64, 16, 64, ""),
]


def bench_wave(args):
    """Lines whose last pass runs across wavefront lanes (wv=1) against the same plans with the
    LDS last pass (wv=0): 128 MiB per buffer, forward.  --case i times one case and one variant
    (--wave 0 / 1) for three launches only (for counter runs)."""
    import torch

    import catabrafft as C

    s = torch.cuda.current_stream()
    cases = list(enumerate(WAVE_CASES)) if args.case < 0 else [(args.case, WAVE_CASES[args.case])]
    out = []
    for i, (n, dt, E, tt, extra) in cases:
        batch = (1 << 27) // (n * np.dtype(dt).itemsize)
        x = torch.randn(batch, n, dtype=torch.complex64 if dt == np.complex64 else torch.complex128,
                        device="cuda")
        y = torch.empty_like(x)
        row = {"case": i, "n": n, "dtype": np.dtype(dt).name, "E": E, "threads": tt,
               "radices": extra or "default", "batch": batch}
        for wv in ((0, 1) if args.case < 0 else (args.wave,)):
            p = _forced_plan(C, n, batch, dt, f"{E} {tt} 0 0 0 -1 0 0 {extra} wv={wv}")
            fn = lambda: p.exec_ptr(x.data_ptr(), y.data_ptr(), -1, False, s.cuda_stream)  # noqa
            if args.case >= 0:
                for _ in range(3):
                    fn()
                torch.cuda.synchronize()
                ms = None
            else:
                ms = _sync_ms(torch, fn, 50, 5, s)
            row[f"wv{wv}"] = {"ms": None if ms is None else round(ms, 4),
                              "algo_GBs": None if ms is None else
                              round(2 * n * batch * np.dtype(dt).itemsize / (ms * 1e-3) / 1e9, 1),
                              "plan": p.describe()}
        out.append(row)
        del x, y
    print(json.dumps({"workload": "last pass across wavefront lanes (wv=1) vs through LDS (wv=0)",
                      "cases": out}))


def bench_chunk(args):
    """The 2^20 x 64 four-step run whole (chunk 0) or in rounds of `chunk` transforms whose
    middle could stay in the Infinity Cache.  --chunks restricts the list (e.g. "f32:8")."""
    import torch

    import catabrafft as C

    s = torch.cuda.current_stream()
    n, batch = 1 << 20, 64
    want = set(args.chunks.split(",")) if args.chunks else None
    out = []
    for dt in (np.complex64, np.complex128):
        tag0 = "f32" if dt == np.complex64 else "f64"
        x = torch.randn(batch, n, dtype=torch.complex64 if dt == np.complex64 else torch.complex128,
                        device="cuda")
        y = torch.empty_like(x)
        for chunk in (0, 4, 8, 16, 32):
            tag = f"{tag0}:{chunk}"
            if want and tag not in want:
                continue
            E = 16 if dt == np.complex64 else 8
            p = _forced_plan(C, n, batch, dt, f"{E} 64 0 0 0 {chunk} 0 0")
            ms = _sync_ms(torch, lambda: p.exec_ptr(x.data_ptr(), y.data_ptr(), -1, False,
                                                    s.cuda_stream), 5 if want else 20, 3, s)
            out.append({"config": tag, "ms": round(ms, 4),
                        "algo_GBs": round(2 * n * batch * np.dtype(dt).itemsize / (ms * 1e-3) / 1e9, 1),
                        "plan": p.describe()})
        del x, y
    print(json.dumps({"workload": "2^20 x 64 four-step, whole batch vs chunked rounds",
                      "cases": out}))


TUNE_SIZES = [1 << 16, 1 << 18, 1 << 20, 1 << 22, 1 << 24, 10 ** 6, 3 ** 12, 4 * 5 ** 8, 2520,
              1024, 4099, 65537]


def bench_tune(args):
    """What MEASURE buys: for each length (512 MiB of input per precision), the NO_FLAG plan and
    the MEASURE plan timed alike, each launch from cold caches (a 512 MiB write first), median of
    9 launches; planning time of MEASURE beside it."""
    import torch

    import catabrafft as C

    s = torch.cuda.current_stream()
    flush = torch.empty(512 << 20, dtype=torch.uint8, device="cuda")

    def cold_ms(p, x, y, reps=9):
        for _ in range(2):
            p.exec_ptr(x.data_ptr(), y.data_ptr(), -1, False, s.cuda_stream)
        ms = []
        for r in range(reps):
            flush.fill_(r & 255)
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record(s)
            p.exec_ptr(x.data_ptr(), y.data_ptr(), -1, False, s.cuda_stream)
            e1.record(s)
            e1.synchronize()
            ms.append(e0.elapsed_time(e1))
        return float(np.median(ms))

    out = []
    for n in TUNE_SIZES:
        for dt in (np.complex64, np.complex128):
            es = np.dtype(dt).itemsize
            batch = max(1, (512 << 20) // (n * es))
            tdt = torch.complex64 if dt == np.complex64 else torch.complex128
            x = torch.randn(batch, n, dtype=tdt, device="cuda")
            y = torch.empty_like(x)
            p0 = C.Plan(n, batch, dt)
            t0 = time.perf_counter()
            p1 = C.Plan(n, batch, dt, flags=C.MEASURE)
            plan_s = time.perf_counter() - t0
            m0, m1 = cold_ms(p0, x, y), cold_ms(p1, x, y)
            gbs = lambda m: round(2 * n * batch * es / (m * 1e-3) / 1e9, 1)  # noqa
            out.append({"n": n, "batch": batch, "dtype": np.dtype(dt).name,
                        "default_ms": round(m0, 4), "measure_ms": round(m1, 4),
                        "default_GBs": gbs(m0), "measure_GBs": gbs(m1),
                        "speedup": round(m0 / m1, 3), "measure_plan_s": round(plan_s, 2),
                        "default_plan": p0.describe(), "measure_plan": p1.describe()})
            print(f"tune n={n} {np.dtype(dt).name}: {m0:.4f} -> {m1:.4f} ms", file=sys.stderr,
                  flush=True)
            del x, y
    print(json.dumps({"workload": "NO_FLAG against MEASURE plans, 512 MiB per case, cold launches",
                      "cases": out}))


def bench_copy(args):
    """A plain device-to-device copy of 1 GiB (torch): the reference kernel whose FETCH_SIZE /
    WRITE_SIZE counters calibrate the HBM-traffic figures of the FFT kernels."""
    import torch

    nbytes = 1 << 30
    x = torch.empty(nbytes // 4, dtype=torch.float32, device="cuda").uniform_()
    y = torch.empty_like(x)
    ms = _time_ms(lambda: y.copy_(x), args.steps, args.warmup)
    print(json.dumps({"workload": "device copy, 1 GiB read + 1 GiB written", "ms": round(ms, 4),
                      "GBs_read_plus_write": round(2 * nbytes / (ms * 1e-3) / 1e9, 1)}))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cases", default="all", help="comma list of case indices (batched)")
    ap.add_argument("--flags", default="measure", choices=sorted(FLAGS),
                    help="batched: planner flag of the plans")
    ap.add_argument("--workload", default="batched",
                    choices=["batched", "single", "rader", "uneven", "serve", "copy", "f16",
                             "wave", "chunk", "tune"])
    ap.add_argument("--case", type=int, default=-1, help="wave: one case index (counter runs)")
    ap.add_argument("--wave", type=int, default=1, help="wave: the variant of --case")
    ap.add_argument("--chunks", default="", help="chunk: comma list such as f32:0,f32:8")
    ap.add_argument("--requests", type=int, default=2000, help="serve: trace length")
    ap.add_argument("--no-baseline", action="store_true",
                    help="serve: skip the one-launch-per-transform baseline (profiling)")
    ap.add_argument("--wisdom", default="", help="batched: import this wisdom file first")
    ap.add_argument("--save-wisdom", default="", help="batched: export the plans' wisdom here")
    ap.add_argument("--fields", default="",
                    help="batched: plan the selected cases with these wisdom knobs "
                         "(E threads n1 ca cb chunk pipe large [ra=..] [rb=..] [pr=..])")
    ap.add_argument("--sections", default="serving,rader,uneven",
                    help="batched: comma list of the extra sections to run (or none)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this FFT engine runs on one GPU; the batched workload is measured at "
              "--gpus 1 only (transforms shard by rows across GPUs with no communication).",
              file=sys.stderr)
        sys.exit(2)
    {"batched": bench_batched, "single": bench_single, "rader": bench_rader,
     "uneven": bench_uneven, "serve": bench_serve, "copy": bench_copy,
     "f16": bench_f16, "wave": bench_wave, "chunk": bench_chunk,
     "tune": bench_tune}[args.workload](args)


if __name__ == "__main__":
    main()
