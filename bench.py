"""Benchmark: batched real FFT round trip on one MI355X (drop-in C ABI, HIP kernels).

Workload (one step): CkFftRealForwardBatch then CkFftRealInverseBatch over 256
independent signals of 65536 real FP32 samples (iid N(0,1), seeded), inputs resident
in HBM before timing starts.  The value is whole-step throughput in GFLOP/s with the
2.5*N*log2(N) convention for real transforms; each launch's duration is taken with HIP
events on the launch stream, which gives the roofline fraction against 8 TB/s.

Also reported on the same line:
  * batched_uneven: 20,000 transforms of mixed sizes (2^6..2^16), kinds and directions,
    each with its own buffers, in two persistent launches (autotuned, graph-replayed);
  * plain: the same signals through the one-signal-per-call drop-in entry points
    (CkFftRealForward / CkFftRealInverse with device pointers, synchronous);
  * cpu_baseline: the reference's CPU path (oracle/_ref, built from the reference's
    own sources; the CPU restatement if that is absent) timed on this host's cores on a
    bounded sample of the same workload;
  * parity: a sample of the benchmarked outputs against the CPU oracle and FP64;
  * serving: Poisson arrivals at 200k requests/s of the uneven mix through the
    continuous-batching server under a 4 GiB HBM budget: p50/p99 latency, throughput,
    and a bitwise alone-vs-mix check on a sample of the outputs;
  * split_bf16x3: the opt-in matrix-core precision mode at N = 4096..65536 (256 MiB
    batches): throughput, matrix-core FLOP/s and error against FP64, FP32 mode beside it.

python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
from __future__ import annotations

import argparse
import json
import math
import os
import platform
import sys
import threading
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "FFT GFLOP/s (5N·log2N) & % HBM roofline, batched vs plain; p50/p99 latency"
HBM_PEAK_GBS = 8000.0
N_REAL = 65536
BATCH = 256


def real_flops(n: int) -> float:
    return 2.5 * n * math.log2(n)


def real_bytes(n: int) -> int:
    # one read of the input and one write of the output, per transform
    return 4 * n + 8 * (n // 2 + 1)


def load_traffic(workload: str):
    """HBM bytes per launch from the committed rocprofv3 FETCH_SIZE/WRITE_SIZE runs."""
    p = os.path.join(ROOT, "profiles", f"traffic_{workload}.json")
    if not os.path.exists(p):
        return None
    with open(p) as f:
        return json.load(f)


def cpu_info():
    """This host: CPU model, the cores this process may run on (sched_getaffinity), the
    cgroup CPU quota if one is set, and the thread count the CPU baseline uses: every
    affinity core, but no more than the CPU share (the cgroup quota, else 16 host cores, the
    share of one GPU on a shared MI355X node) -- threads beyond the share would only
    time-slice, or take cores from other tenants of the node."""
    model = platform.processor() or platform.machine()
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    affinity = len(os.sched_getaffinity(0))
    quota = None
    try:
        with open("/sys/fs/cgroup/cpu.max") as f:
            q, period = f.read().split()[:2]
            if q != "max":
                quota = float(q) / float(period)
    except (OSError, ValueError):
        pass
    share = math.ceil(quota) if quota else 16
    return {"cpu": model, "affinity_cores": affinity, "cgroup_cpu_quota": quota, "share_cores": share,
            "threads": max(1, min(affinity, share)), "online_cpus": os.cpu_count()}


def _timed_passes(one_pass, seconds: float):
    one_pass()  # warm
    passes, t0 = 0, time.perf_counter()
    while True:
        one_pass()
        passes += 1
        el = time.perf_counter() - t0
        if el >= seconds:
            return passes, el


def _median_us(fn, reps: int) -> float:
    t = []
    for k in range(reps):
        t1 = time.perf_counter()
        fn(k)
        t.append(time.perf_counter() - t1)
    return float(np.median(t)) * 1e6


def cpu_baseline(n: int, seconds: float = 10.0, info=None):
    """Reference CPU path on this host for the benchmarked workload (real forward +
    inverse, n = 65536): `threads` threads sharing one context on a bounded sample of
    signals, and one core timing one signal's forward + inverse, median of 500 repetitions
    (the reference's own convention, src/test/test.cpp:611).  GFLOP/s at 2.5 N log2 N per
    real transform, with the literal 5 N log2 N figure beside it."""
    import oracle as O

    info = info or cpu_info()
    threads = info["threads"]
    sample = 64
    x = O.normal01(0x5EED, sample * n).reshape(sample, n)
    use_ref = O.Reference.available()
    single = None
    if use_ref:
        R = O.Reference()
        ctx = R.init(n, 3)
        lib = R.lib
        yb = np.empty((sample, 2 * (n // 2 + 1)), np.float32)
        zb = np.empty((sample, n), np.float32)
        tmp = np.empty(2 * (n // 2 + 1), np.float32)

        def one_pass():
            R.run_threads(ctx, [2] * sample, [n] * sample, list(x), list(yb), threads)
            R.run_threads(ctx, [3] * sample, [n] * sample, list(yb), list(zb), threads)

        passes, el = _timed_passes(one_pass, seconds)

        def one(k):
            lib.CkFftRealForward(ctx, n, x[k % sample].ctypes.data, yb[0].ctypes.data)
            lib.CkFftRealInverse(ctx, n, yb[0].ctypes.data, zb[0].ctypes.data, tmp.ctypes.data)

        one(0)
        med = _median_us(one, 500) * 1e-6
        single = {"value": round(2 * real_flops(n) / med / 1e9, 3), "unit": "GFLOP/s", "cores": 1,
                  "value_literal_5N": round(4 * real_flops(n) / med / 1e9, 3),
                  "us_per_forward_plus_inverse": round(med * 1e6, 1), "sample": "median of 500 repetitions"}
        R.shutdown(ctx)
    else:
        o = O.Restatement(n, 3)

        def one_pass():
            y = o.real_forward(x, threads)
            o.real_inverse(y, n, threads)

        passes, el = _timed_passes(one_pass, seconds)
    gflops = passes * sample * 2 * real_flops(n) / el / 1e9
    return {"value": round(gflops, 3), "unit": "GFLOP/s", "cores": threads,
            "value_literal_5N": round(2 * gflops, 3),
            "kind": "reference" if use_ref else "port",
            "sample": f"{passes} passes x {sample} signals x (real fwd + real inv), n={n}, "
                      f"{el:.1f} s wall, {threads} POSIX threads sharing one context",
            **{k: info[k] for k in ("cpu", "affinity_cores", "cgroup_cpu_quota", "share_cores", "online_cpus")},
            "single_core": single}


def cpu_other_configs(info, seconds: float = 4.0):
    """The reference CPU path (oracle/_ref) beside the other workloads, on this host:
      * complex 4096 x 1024 forward + inverse: `threads` threads over the 1024 signals
        for `seconds`, and one core's median of 500 forward + inverse pairs;
      * complex 2^24 forward: one transform on one core (the reference computes one
        transform on one thread, so more cores do not shorten it), two repetitions;
      * the uneven mix: a 2,000-transform sample of the same size / kind / direction
        distribution, split over `threads` threads for `seconds`, and on one core.
    GFLOP/s at 5 N log2 N complex and 2.5 N log2 N real."""
    import oracle as O

    if not O.Reference.available():
        return None
    R = O.Reference()
    lib = R.lib
    threads = info["threads"]
    out = {}
    # complex 4096 x 1024
    n, count = 4096, 1024
    ctx = R.init(n, 3)
    x = O.uniform_pm1(0xC1, count * 2 * n).reshape(count, 2 * n)
    y = np.empty_like(x)
    z = np.empty_like(x)

    def one_pass():
        R.run_threads(ctx, [0] * count, [n] * count, list(x), list(y), threads)
        R.run_threads(ctx, [1] * count, [n] * count, list(y), list(z), threads)

    passes, el = _timed_passes(one_pass, seconds)

    def one(k):
        lib.CkFftComplexForward(ctx, n, x[k % count].ctypes.data, y[0].ctypes.data)
        lib.CkFftComplexInverse(ctx, n, y[0].ctypes.data, z[0].ctypes.data)

    med = _median_us(one, 500) * 1e-6
    fl = 2 * 5.0 * n * 12
    out["complex_4096_x1024"] = {
        "value": round(passes * count * fl / el / 1e9, 3), "unit": "GFLOP/s", "cores": threads,
        "sample": f"{passes} passes x {count} signals x (fwd + inv), {el:.1f} s wall",
        "single_core": {"value": round(fl / med / 1e9, 3), "us_per_forward_plus_inverse": round(med * 1e6, 1),
                        "sample": "median of 500 repetitions"}}
    R.shutdown(ctx)
    del x, y, z
    # complex 2^24, one transform
    n = 1 << 24
    ctx = R.init(n, 1)
    x = O.uniform_pm1(0xC3, 2 * n)
    y = np.empty_like(x)
    lib.CkFftComplexForward(ctx, n, x.ctypes.data, y.ctypes.data)
    reps = []
    for _ in range(2):
        t0 = time.perf_counter()
        lib.CkFftComplexForward(ctx, n, x.ctypes.data, y.ctypes.data)
        reps.append(time.perf_counter() - t0)
    t = min(reps)
    out["complex_2p24"] = {"value": round(5.0 * n * 24 / t / 1e9, 3), "unit": "GFLOP/s", "cores": 1,
                           "ms": round(t * 1e3, 1), "sample": "one transform, best of 2 after one warm-up"}
    R.shutdown(ctx)
    del x, y
    # uneven mix sample
    rng = np.random.default_rng(0xC4)
    m = 2000
    ks = rng.integers(6, 17, size=m)
    real = rng.random(m) < 0.5
    fwd = rng.random(m) < 0.5
    ctx = R.init(1 << 16, 3)
    bufs = []
    for i in range(m):
        nn = 1 << int(ks[i])
        if real[i] and fwd[i]:
            a, b = O.normal01(i, nn), np.empty(2 * (nn // 2 + 1), np.float32)
        elif real[i]:
            a, b = O.normal01(i, 2 * (nn // 2 + 1)), np.empty(nn, np.float32)
        else:
            a, b = O.normal01(i, 2 * nn), np.empty(2 * nn, np.float32)
        bufs.append((a, b))
    flops = float(np.sum(np.where(real, 2.5, 5.0) * (1 << ks) * ks))

    kinds = np.where(real, np.where(fwd, 2, 3), np.where(fwd, 0, 1))
    ns = 1 << ks
    ins = [b[0] for b in bufs]
    outs = [b[1] for b in bufs]
    passes, el = _timed_passes(lambda: R.run_threads(ctx, kinds, ns, ins, outs, threads), seconds)
    t0 = time.perf_counter()
    R.run_threads(ctx, kinds, ns, ins, outs, 1)
    el1 = time.perf_counter() - t0
    out["uneven_mix"] = {"value": round(passes * flops / el / 1e9, 3), "unit": "GFLOP/s", "cores": threads,
                         "sample": f"{passes} passes x {m} transforms of the uneven distribution, {el:.1f} s wall",
                         "single_core": {"value": round(flops / el1 / 1e9, 3), "sample": f"one pass of {m} transforms"}}
    R.shutdown(ctx)
    return out


def uneven_mix(torch, dev, count: int, seed: int):
    """`count` transforms, n = 2^k with k uniform in 6..16, kinds 50% complex / 50% real and
    directions 50/50, iid N(0,1) inputs, each with its own 256-byte aligned buffers."""
    rng = np.random.default_rng(seed)
    ks = rng.integers(6, 17, size=count)
    real = rng.random(count) < 0.5
    fwd = rng.random(count) < 0.5
    kinds = np.where(real, np.where(fwd, 2, 3), np.where(fwd, 0, 1))
    ns = (1 << ks).astype(np.int64)
    in_len = np.where(kinds == 2, ns, np.where(kinds == 3, 2 * (ns // 2 + 1), 2 * ns))
    out_len = np.where(kinds == 2, 2 * (ns // 2 + 1), np.where(kinds == 3, ns, 2 * ns))
    align = 64                                             # 256-byte aligned buffers
    in_off = np.concatenate([[0], np.cumsum((in_len + align - 1) // align * align)])
    out_off = np.concatenate([[0], np.cumsum((out_len + align - 1) // align * align)])
    gen = torch.Generator(device=dev)
    gen.manual_seed(seed)
    x = torch.randn(int(in_off[-1]), device=dev, generator=gen)
    y = torch.empty(int(out_off[-1]), device=dev)
    xb, yb = x.data_ptr(), y.data_ptr()
    descs = [(int(kinds[i]), int(ns[i]), xb + 4 * int(in_off[i]), yb + 4 * int(out_off[i])) for i in range(count)]
    flops = np.where(real, 2.5, 5.0) * ns * ks
    return dict(ks=ks, ns=ns, kinds=kinds, in_off=in_off, out_off=out_off, in_len=in_len, out_len=out_len,
                x=x, y=y, descs=descs, flops=flops)


def uneven_leg(torch, C, O, dev, steps: int, warmup: int, plain_sample: int = 2000):
    """20,000 transforms of the uneven mix run in two persistent launches (autotuned,
    replayed as a hipGraph).  FLOPs: 5 n log2 n complex, 2.5 n log2 n real."""
    count = 20000
    m = uneven_mix(torch, dev, count, 0xC4)
    ks, ns, kinds, x, y, descs = m["ks"], m["ns"], m["kinds"], m["x"], m["y"], m["descs"]
    in_off, out_off, in_len, out_len = m["in_off"], m["out_off"], m["in_len"], m["out_len"]
    xb, yb = x.data_ptr(), y.data_ptr()
    flops = float(np.sum(m["flops"]))
    stream = torch.cuda.current_stream(dev)

    def replay_ms(c, reps=10):
        p = C.BatchPlan(c, descs, graph=True)
        for _ in range(20):
            p.execute(stream)
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record(stream)
        for _ in range(reps):
            p.execute(stream)
        b.record(stream)
        torch.cuda.synchronize()
        p.close()
        return a.elapsed_time(b) / reps

    # the default layouts (no tuning), timed as the tuners time plans: graph replays
    with C.Context(1 << 16, C.kCkFftDirection_Both) as c0:
        default_ms = replay_ms(c0)
    # layouts chosen in this mix (CkFftAutotuneMix from the defaults; choices made on each
    # class alone lost in the mix on this GPU)
    ctx = C.Context(1 << 16, C.kCkFftDirection_Both)
    t0 = time.perf_counter()
    mix_ok, mix_before, mix_after = ctx.autotune_mix(descs)
    mix_s = time.perf_counter() - t0
    tuned, tune_s = mix_ok == 1, mix_s
    plan = C.BatchPlan(ctx, descs, graph=True)
    info = plan.info()
    # untimed replays first: a new plan's replays get faster over the first few tens
    for _ in range(max(20, warmup)):
        assert plan.execute(stream) == 1
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(steps):
        plan.execute(stream)
    e1.record(stream)
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / steps
    gbs = info["algorithmic_bytes"] / (ms * 1e-3) / 1e9
    # parity sample: a few transforms against the CPU oracle
    worst = 0.0
    for i in range(0, count, count // 12):
        kind, n = int(kinds[i]), int(ns[i])
        xi = x[int(in_off[i]): int(in_off[i]) + int(in_len[i])].cpu().numpy()[None]
        yi = y[int(out_off[i]): int(out_off[i]) + int(out_len[i])].cpu().numpy()[None]
        o = O.Restatement(n, 3)
        ref = (o.complex(xi, kind == 1) if kind < 2 else o.real_forward(xi) if kind == 2 else o.real_inverse(xi, n))
        worst = max(worst, O.split_norm_err(yi, ref))
    # plain: the first `plain_sample` transforms, one synchronous drop-in call each
    tmp = torch.empty(2 * ((1 << 16) // 2 + 1), device=dev)
    calls = {0: ctx.complex_forward, 1: ctx.complex_inverse, 2: ctx.real_forward}
    t0 = time.perf_counter()
    pf = 0.0
    for i in range(plain_sample):
        kind, n = int(kinds[i]), int(ns[i])
        xi, yi = xb + 4 * int(in_off[i]), yb + 4 * int(out_off[i])
        ok = calls[kind](n, xi, yi) if kind < 3 else ctx.real_inverse(n, xi, yi, tmp)
        assert ok == 1
        pf += (2.5 if kind >= 2 else 5.0) * n * int(ks[i])
    plain_s = time.perf_counter() - t0
    plan.close()
    ctx.close()
    del x, y
    return {
        "workload": "20000 uneven transforms, n=2^k k~U{6..16}, 50% complex/real, 50% fwd/inv, own buffers",
        "ms_per_batch": round(ms, 4), "value": round(flops / (ms * 1e-3) / 1e9, 1), "unit": "GFLOP/s",
        "achieved_GBs": round(gbs, 1), "frac_of_hbm_peak": round(gbs / HBM_PEAK_GBS, 4),
        "algorithmic_bytes": info["algorithmic_bytes"], "launch": "table-driven passes for complex 2^16, narrow and wide persistent kernels, as three concurrent branches of one hipGraph replay",
        "pass_transforms": info["pass_transforms"],
        "work_items": info["work_items"], "workgroups": info["workgroups"],
        "extra_launches": info["extra_launches"], "autotuned": tuned, "autotune_s": round(tune_s, 2),
        "autotune_mix": {"ok": mix_ok == 1, "s": round(mix_s, 2),
                         "ms_per_batch_default_layouts": round(default_ms, 4),
                         "ms_per_batch_after_mix_tuning": round(mix_after * 1e-3, 4),
                         "timing": "graph replays of the plan, 10 timed after 20 untimed"},
        "parity_worst_vs_ref": float(f"{worst:.3e}"),
        "plain": {"value": round(pf / plain_s / 1e9, 2), "unit": "GFLOP/s",
                  "sample": f"first {plain_sample} transforms, one synchronous call each"},
        "batched_over_plain": round(flops / (ms * 1e-3) / (pf / plain_s), 1),
    }


def serving_leg(torch, C, dev, rate: float = 200e3, seconds: float = 2.0, budget: int = 4 << 30,
                slots: int = 4, pool: int = 20000):
    """Poisson arrivals at `rate` requests/s for `seconds`, sizes and kinds drawn as in the
    uneven mix, served by the continuous-batching server under an HBM admission budget.
    Requests cycle over a pool of `pool` distinct buffer pairs; two requests sharing a pair
    arrive pool/rate seconds apart (100 ms at the defaults), far beyond any latency here.
    Latency is arrival (scheduled time) to completion seen by the dispatcher."""
    m = uneven_mix(torch, dev, pool, 0x5E)
    kinds, ns, x, y = m["kinds"], m["ns"], m["x"], m["y"]
    torch.cuda.synchronize()
    count = int(rate * seconds)
    arrivals = np.cumsum(np.random.default_rng(0x5E).exponential(1e6 / rate, count))
    descs = [m["descs"][i % pool] for i in range(count)]
    offered_bytes = float(np.mean([(4 * n + 8 * (n // 2 + 1)) if k >= 2 else 16 * n
                                   for k, n in zip(kinds, ns)]))
    ctx = C.Context(1 << 16, C.kCkFftDirection_Both)
    ctx.autotune(None)
    out = {}
    with C.Server(ctx, budget, max_in_flight=slots) as srv:
        # warm-up (first launches of each kernel), not measured
        srv.replay(descs[:200], np.arange(200) * 100.0)
        assert srv.drain() == 1
        srv.reset_stats()
        # unloaded: one request at a time (2,000 req/s for 1 s), the latency floor
        srv.replay(descs[:2000], np.arange(2000) * 500.0)
        assert srv.drain() == 1
        idle = srv.stats()
        srv.reset_stats()
        srv.replay(descs, arrivals)
        assert srv.drain() == 1
        st = srv.stats()
    # alone vs in the mix: the last request on each sampled buffer pair against a single call
    tmp = torch.empty(2 * ((1 << 16) // 2 + 1), device=dev)
    calls = {0: ctx.complex_forward, 1: ctx.complex_inverse, 2: ctx.real_forward}
    mismatches = 0
    checked = 0
    for j in range(0, pool, pool // 200):
        kind, n = int(kinds[j]), int(ns[j])
        xi = x[int(m["in_off"][j]): int(m["in_off"][j]) + int(m["in_len"][j])]
        yi = y[int(m["out_off"][j]): int(m["out_off"][j]) + int(m["out_len"][j])]
        want = torch.empty_like(yi)
        ok = calls[kind](n, xi, want) if kind < 3 else ctx.real_inverse(n, xi, want, tmp)
        assert ok == 1
        mismatches += int(not torch.equal(want.view(torch.int32), yi.view(torch.int32)))
        checked += 1
    ctx.close()
    del x, y
    out = {
        "workload": f"Poisson {rate:.0f} req/s for {seconds:g} s, uneven mix (n=2^k k~U{{6..16}}, "
                    f"50% complex/real, 50% fwd/inv), HBM budget {budget / 2**30:g} GiB, {slots} slots",
        "requests": count, "completed": st["completed"], "failed": st["failed"],
        "p50_us": round(st["p50_us"], 1), "p99_us": round(st["p99_us"], 1), "max_us": round(st["max_us"], 1),
        "max_request": {"client_us": round(st["max_client_us"], 1), "wait_us": round(st["max_wait_us"], 1),
                        "device_us": round(st["max_device_us"], 1),
                        "meaning": "the slowest request: scheduled arrival to submission (the replaying client), "
                                   "submission to its batch's launch, launch to completion seen"},
        "mean_us": round(st["mean_us"], 1),
        "throughput_req_s": round(st["requests_per_s"], 0), "offered_req_s": rate,
        "achieved_GBs": round(st["GB_per_s"], 2), "offered_GBs": round(rate * offered_bytes / 1e9, 2),
        "batches": st["batches"], "mean_batch": round(st["mean_batch"], 2),
        "peak_inflight_bytes": st["peak_inflight_bytes"], "budget_bytes": st["budget_bytes"],
        "unloaded": {"rate_req_s": 2000, "requests": 2000, "after_warmup": 200,
                     "p50_us": round(idle["p50_us"], 1), "p99_us": round(idle["p99_us"], 1)},
        "alone_vs_mix": {"checked": checked, "bitwise_mismatches": mismatches},
    }
    return out


def event_ms(torch, stream, fn, reps: int, warmup: int = 2):
    """Median duration in ms of `fn()` (one enqueue on `stream`) over `reps` event pairs."""
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(reps)]
    for a, b in ev:
        a.record(stream)
        fn()
        b.record(stream)
    torch.cuda.synchronize()
    return float(np.median([a.elapsed_time(b) for a, b in ev]))


def back_to_back_ms(torch, stream, fn, reps: int, warmup: int = 2):
    """Mean duration in ms of `fn()` over `reps` enqueues back to back between one pair of
    events (dispatch gaps included; no event between the launches)."""
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    a = torch.cuda.Event(enable_timing=True)
    b = torch.cuda.Event(enable_timing=True)
    a.record(stream)
    for _ in range(reps):
        fn()
    b.record(stream)
    torch.cuda.synchronize()
    return a.elapsed_time(b) / reps


def complex_4096_leg(torch, C, O, dev, reps: int = 50, rotate_mib: int = 1024):
    """Complex forward then inverse round trip, N=4096, 1024 signals, re/im iid U[-1,1]
    (SplitMix64 seed 0xC1): per-launch times back to back, cycling through buffer sets of
    `rotate_mib` MiB in total (beyond the Infinity Cache), with the cache-resident times
    (one buffer set) and the median of individually event-timed launches beside them;
    GB/s against 8 TB/s; errors of a sample."""
    n, count = 4096, 1024
    x_h = O.uniform_pm1(0xC1, count * 2 * n).reshape(count, 2 * n)
    set_bytes = 3 * 8 * n * count
    nsets = max(1, math.ceil((rotate_mib << 20) / set_bytes))
    xs = [torch.from_numpy(x_h).to(dev)]
    xs += [xs[0].clone() for _ in range(nsets - 1)]
    ys = [torch.empty_like(xs[0]) for _ in range(nsets)]
    zs = [torch.empty_like(xs[0]) for _ in range(nsets)]
    stream = torch.cuda.current_stream(dev)
    reps = max(reps, 2 * nsets)
    ctr = {"i": 0}

    def nxt(rotate):
        ctr["i"] += 1
        return ctr["i"] % nsets if rotate else 0

    with C.Context(n, C.kCkFftDirection_Both) as ctx:
        def fwd(rotate=True):
            k = nxt(rotate)
            return ctx.complex_forward_batch(n, count, xs[k], ys[k], stream)

        def inv(rotate=True):
            k = nxt(rotate)
            return ctx.complex_inverse_batch(n, count, ys[k], zs[k], stream)

        # untimed: one pass over every buffer set in each direction, so that the writes that
        # created the sets have left the caches before anything is timed
        fwd_ms = back_to_back_ms(torch, stream, fwd, reps, warmup=nsets)
        inv_ms = back_to_back_ms(torch, stream, inv, reps, warmup=nsets)
        fwd_c = back_to_back_ms(torch, stream, lambda: fwd(False), reps)
        inv_c = back_to_back_ms(torch, stream, lambda: inv(False), reps)
        fwd_ev = event_ms(torch, stream, fwd, reps)
        # the round trip on set 0 for the error sample
        assert ctx.complex_forward_batch(n, count, xs[0], ys[0], stream) == 1
        assert ctx.complex_inverse_batch(n, count, ys[0], zs[0], stream) == 1
    torch.cuda.synchronize()
    idx = [0, 511, 1023]
    yh, zh = ys[0][idx].cpu().numpy(), zs[0][idx].cpu().numpy()
    o = O.Restatement(n, 3)
    nbytes = 16.0 * n * count
    frac = lambda ms: round(nbytes / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)
    del xs, ys, zs
    return {"workload": "complex fwd + inv, N=4096 x 1024, U[-1,1]",
            "forward_ms": round(fwd_ms, 5), "inverse_ms": round(inv_ms, 5),
            "timing": f"{reps} launches back to back between two events, elapsed / {reps}, cycling through "
                      f"{nsets} buffer sets ({nsets * set_bytes / 2**30:.2f} GiB)",
            "working_set_bytes": nsets * set_bytes,
            "forward_ms_event_median": round(fwd_ev, 5),
            "GFLOP_s": round(2 * 5.0 * n * 12 * count / ((fwd_ms + inv_ms) * 1e-3) / 1e9, 1),
            "forward_GBs": round(nbytes / (fwd_ms * 1e-3) / 1e9, 1),
            "frac_of_hbm_peak": frac(fwd_ms), "inverse_frac_of_hbm_peak": frac(inv_ms),
            "in_cache": {"forward_ms": round(fwd_c, 5), "inverse_ms": round(inv_c, 5),
                         "frac_of_hbm_peak": frac(fwd_c), "inverse_frac_of_hbm_peak": frac(inv_c),
                         "working_set_bytes": set_bytes},
            "forward_vs_ref": float(f"{O.split_norm_err(yh, o.complex(x_h[idx], False, 8)):.3e}"),
            "forward_relL2_fp64": float(f"{O.rel_l2(yh, O.interleave(O.fp64_complex(x_h[idx], False))):.3e}"),
            "round_trip_relL2": float(f"{O.rel_l2(zh / n, x_h[idx]):.3e}")}


def huge_leg(torch, C, O, dev, reps: int = 10):
    """One complex forward transform of N=2^24 (128 MiB in, 128 MiB out), U[-1,1]: time
    (calls back to back, each with its own intermediate allocation on the stream),
    GB/s against 8 TB/s (three global-memory passes move 3x the algorithmic bytes), and
    the relative L2 error against the FP64 FFT."""
    n = 1 << 24
    gen = torch.Generator(device=dev)
    gen.manual_seed(0xC3)
    x = torch.rand(2 * n, device=dev, generator=gen) * 2 - 1
    y = torch.empty_like(x)
    stream = torch.cuda.current_stream(dev)
    with C.Context(n, C.kCkFftDirection_Forward) as ctx:
        ms = back_to_back_ms(torch, stream, lambda: ctx.complex_forward_batch(n, 1, x, y, stream), reps)
    torch.cuda.synchronize()
    err = O.rel_l2(y.cpu().numpy(), O.interleave(O.fp64_complex(x.cpu().numpy(), False)))
    gbs = 16.0 * n / (ms * 1e-3) / 1e9
    return {"workload": "single complex forward, N=2^24, U[-1,1]", "ms": round(ms, 4),
            "GFLOP_s": round(5.0 * n * 24 / (ms * 1e-3) / 1e9, 1), "achieved_GBs": round(gbs, 1),
            "frac_of_hbm_peak": round(gbs / HBM_PEAK_GBS, 4), "passes": 3,
            "working_set_bytes": 3 * 8 * n,   # input, output and the intermediate
            "relL2_fp64": float(f"{err:.3e}")}


def host_pointer_leg(C, O, signals: int = 64):
    """PCIe-inclusive rate of the unmodified drop-in call: CkFftRealForward/Inverse on
    pageable host (numpy) buffers, n=65536, one synchronous call per signal (staged through
    pinned bounce buffers inside the call), from one thread and from 4 threads sharing the
    context (the reference allows concurrent use of a context); median of 3 timed runs."""
    n = N_REAL
    x = O.normal01(0xB0, signals * n).reshape(signals, n)
    y = np.zeros((signals, 2 * (n // 2 + 1)), np.float32)   # touched: no first-write page faults timed
    z = np.zeros((signals, n), np.float32)
    out = {"workload": f"real fwd + inv, n=65536, {signals} signals, pageable host buffers, one call each"}
    with C.Context(n, C.kCkFftDirection_Both) as ctx:
        def work(lo, hi):
            tmp = np.empty(2 * (n // 2 + 1), np.float32)
            for s in range(lo, hi):
                assert ctx.real_forward(n, x[s], y[s]) == 1
                assert ctx.real_inverse(n, y[s], z[s], tmp) == 1
        def run(threads):
            per = signals // threads
            ts = [threading.Thread(target=work, args=(i * per, (i + 1) * per)) for i in range(threads)]
            t0 = time.perf_counter()
            for t in ts:
                t.start()
            for t in ts:
                t.join()
            return time.perf_counter() - t0

        for threads in (1, 4):
            run(threads)           # untimed: creates the staging slots the threads will use
            el = float(np.median([run(threads) for _ in range(3)]))
            out[f"threads_{threads}"] = {"value": round(signals * 2 * real_flops(n) / el / 1e9, 2),
                                         "unit": "GFLOP/s", "us_per_call": round(el / (2 * signals) * 1e6, 1)}
    return out


def split_bf16_leg(torch, C, O, dev, reps: int = 20, sizes=(4096, 8192, 16384, 32768, 65536),
                   in_bytes: int = 256 << 20):
    """Opt-in split-BF16x3 mode: complex forward transforms of N = 4096..65536, batched to
    256 MiB of input, radix-16 stages on the matrix cores.  Per N: launch time (median of
    `reps` event-timed launches), GFLOP/s (5 N log2 N), algorithmic GB/s (16 N per
    transform) against 8 TB/s, matrix-core TFLOP/s actually issued (12 MFMA of 32x32x16
    per 32 radix-16 butterflies: 768 N flops per radix-16 stage, log2(N)/4 stages), the
    relative L2 error against the FP64 DFT on a sample of signals, and the same for the
    FP32 mode beside it."""
    stream = torch.cuda.current_stream(dev)
    rows = []
    for n in sizes:
        count = in_bytes // (8 * n)
        logn = int(math.log2(n))
        gen = torch.Generator(device=dev)
        gen.manual_seed(0x5B + logn)
        x = torch.rand(count, 2 * n, device=dev, generator=gen) * 2 - 1
        y = torch.empty_like(x)
        row = {"n": n, "count": count}
        for mode, name in ((C.kCkFftPrecision_SplitBf16x3, "split_bf16x3"), (C.kCkFftPrecision_Fp32, "fp32")):
            ctx = C.Context(n, C.kCkFftDirection_Both)
            assert ctx.set_precision(mode) == 1
            for _ in range(3):
                assert ctx.complex_forward_batch(n, count, x, y, stream) == 1
            torch.cuda.synchronize()
            ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(reps)]
            for a, b in ev:
                a.record(stream)
                ctx.complex_forward_batch(n, count, x, y, stream)
                b.record(stream)
            torch.cuda.synchronize()
            ms = float(np.median([a.elapsed_time(b) for a, b in ev]))
            idx = [0, count // 2, count - 1]
            xs = x[idx].cpu().numpy()
            err = O.rel_l2(y[idx].cpu().numpy(), O.interleave(O.fp64_complex(xs, False)))
            gbs = 16.0 * n * count / (ms * 1e-3) / 1e9
            r = {"ms": round(ms, 4), "GFLOP_s": round(5.0 * n * logn * count / (ms * 1e-3) / 1e9, 1),
                 "achieved_GBs": round(gbs, 1), "frac_of_hbm_peak": round(gbs / HBM_PEAK_GBS, 4),
                 "relL2_fp64": float(f"{err:.3e}")}
            if mode == C.kCkFftPrecision_SplitBf16x3:
                mfma_flops = 768.0 * n * (logn // 4) * count
                r["mfma_TFLOP_s"] = round(mfma_flops / (ms * 1e-3) / 1e12, 1)
                r["mfma_frac_of_bf16_peak"] = round(mfma_flops / (ms * 1e-3) / 2.5e15, 4)
            row[name] = r
            ctx.close()
        rows.append(row)
        del x, y
    return {"workload": "complex forward, N=4096..65536, 256 MiB of input per N, iid U[-1,1]",
            "mfma_flops_convention": "768 N per radix-16 stage (6 split products x 2 K-halves x 32x32x16)",
            "gate": "relL2_fp64 <= 1e-6", "sizes": rows}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-plain", action="store_true", help="skip the one-call-per-signal leg")
    ap.add_argument("--no-uneven", action="store_true", help="skip the uneven-batch (persistent launch) leg")
    ap.add_argument("--no-serve", action="store_true", help="skip the serving-loop leg")
    ap.add_argument("--no-split", action="store_true", help="skip the split-BF16x3 precision leg")
    ap.add_argument("--no-configs", action="store_true",
                    help="skip the complex 4096 x 1024, 2^24 and host-pointer legs")
    ap.add_argument("--rotate-mib", type=int, default=1024,
                    help="rotate the timed launches through buffer sets of at least this many MiB in "
                         "total (beyond the 256 MB Infinity Cache); 0 = one set")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="run only a device-to-device copy of 2 x 128 MiB (known bytes) to scale the "
                         "HBM counters FETCH_SIZE / WRITE_SIZE")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}): the transform path "
              "runs on one GPU and this build has no multi-GPU path.", file=sys.stderr)
        sys.exit(2)

    import torch

    import ckfft_amd as C

    assert torch.cuda.is_available(), "bench.py needs a GPU"
    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    if args.calibrate_copy:
        nbytes = 128 << 20
        a = torch.ones(nbytes // 4, dtype=torch.float32, device=dev)
        b = torch.empty_like(a)
        for _ in range(5):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"calibration": "device copy", "bytes_read_per_launch": nbytes,
                          "bytes_written_per_launch": nbytes, "launches": 5}))
        return
    n, B = N_REAL, BATCH
    half = n // 2 + 1

    import oracle as O  # seeded input generator and the parity checker
    rng_x = O.normal01(0xBE7C4, B * n).reshape(B, n)
    # Buffer sets (x, y, z) rotated from launch to launch, so that the working set of the
    # timed launches (>= args.rotate_mib) exceeds the 256 MB Infinity Cache and every
    # launch reads its input from HBM.  Every set holds the same input.
    set_bytes = B * (4 * n + 8 * half + 4 * n)
    nsets = max(1, math.ceil((args.rotate_mib << 20) / set_bytes))
    xs = [torch.from_numpy(rng_x).to(dev)]
    xs += [xs[0].clone() for _ in range(nsets - 1)]
    ys = [torch.empty((B, 2 * half), dtype=torch.float32, device=dev) for _ in range(nsets)]
    zs = [torch.empty((B, n), dtype=torch.float32, device=dev) for _ in range(nsets)]
    x, y, z = xs[0], ys[0], zs[0]
    ctx = C.Context(n, C.kCkFftDirection_Both)
    stream = torch.cuda.current_stream(dev)

    def step(i=0):
        k = i % nsets
        assert ctx.real_forward_batch(n, B, xs[k], ys[k], stream) == 1
        assert ctx.real_inverse_batch(n, B, ys[k], zs[k], stream) == 1

    for i in range(max(args.warmup, nsets)):
        step(i)
    torch.cuda.synchronize()

    # whole-step timing: K steps back to back between two events, synchronised on both
    # sides (no events between the launches: they would add a gap to every launch)
    t_start = torch.cuda.Event(enable_timing=True)
    t_end = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    t_start.record(stream)
    for k in range(args.steps):
        step(k)
    t_end.record(stream)
    torch.cuda.synchronize()
    total_ms = t_start.elapsed_time(t_end)
    ms_per_step = total_ms / args.steps
    # per-launch durations for the roofline: each kernel K times back to back between two
    # events on the launch stream, elapsed / K (the dispatch gaps between the launches are
    # included, so these are upper bounds of the kernels' own durations); `rotate` cycles
    # through the buffer sets, else every launch reuses set 0 (cache-resident)
    reps = max(args.steps, 2 * nsets)

    def back_to_back(launch, rotate=True):
        a = torch.cuda.Event(enable_timing=True)
        b = torch.cuda.Event(enable_timing=True)
        torch.cuda.synchronize()
        a.record(stream)
        for i in range(reps):
            launch(i % nsets if rotate else 0)
        b.record(stream)
        torch.cuda.synchronize()
        return a.elapsed_time(b) / reps

    # the same K steps on one buffer set (cache-resident between steps), for comparison
    t_start.record(stream)
    for k in range(args.steps):
        step(0)
    t_end.record(stream)
    torch.cuda.synchronize()
    ms_per_step_cached = t_start.elapsed_time(t_end) / args.steps

    fwd_launch = lambda k: ctx.real_forward_batch(n, B, xs[k], ys[k], stream)
    inv_launch = lambda k: ctx.real_inverse_batch(n, B, ys[k], zs[k], stream)
    fwd_ms = back_to_back(fwd_launch)
    inv_ms = back_to_back(inv_launch)
    fwd_ms_cached = back_to_back(fwd_launch, rotate=False)
    inv_ms_cached = back_to_back(inv_launch, rotate=False)

    # distribution: 100 launches of each kernel, each between its own pair of events (the
    # events add a small gap to every launch), median and p99; rotated buffer sets
    def distribution(launch, reps=100):
        ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(reps)]
        torch.cuda.synchronize()
        for i, (a, b) in enumerate(ev):
            a.record(stream)
            launch(i % nsets)
            b.record(stream)
        torch.cuda.synchronize()
        t = np.array([a.elapsed_time(b) for a, b in ev])
        return {"reps": reps, "median_ms": round(float(np.median(t)), 5), "p99_ms": round(float(np.percentile(t, 99)), 5)}

    dist = {"real_forward": distribution(fwd_launch), "real_inverse": distribution(inv_launch),
            "timing": "each launch between its own pair of HIP events on the launch stream, buffer sets rotated"}

    flops_step = B * 2 * real_flops(n)
    value = flops_step / (ms_per_step * 1e-3) / 1e9
    bytes_launch = B * real_bytes(n)
    dom_name, dom_ms = ("real_inverse", inv_ms) if inv_ms >= fwd_ms else ("real_forward", fwd_ms)
    achieved = bytes_launch / (dom_ms * 1e-3) / 1e9
    workload = f"real_fft_{n}_x{B}"
    traffic = load_traffic(workload)
    frac = lambda ms: round(bytes_launch / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)
    roofline = {
        "bound": "hbm", "kernel": f"fft_lds_kernel ({dom_name.replace('_', ' ')}, n={n}: one workgroup per signal)",
        "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
        "frac": round(achieved / HBM_PEAK_GBS, 4),
        "traffic": (traffic or {}).get(dom_name),
        "algorithmic_bytes_per_launch": bytes_launch,
        "working_set_bytes": nsets * set_bytes,
        "buffer_sets": nsets,
        "launch_ms": {"real_forward": round(fwd_ms, 5), "real_inverse": round(inv_ms, 5)},
        "launch_timing": f"{reps} launches of each kernel back to back between two events, elapsed / {reps}, "
                         f"cycling through {nsets} buffer sets ({nsets * set_bytes / 2**30:.2f} GiB)",
        "step_GBs": round(2 * bytes_launch / (ms_per_step * 1e-3) / 1e9, 1),
        "frac_per_launch": {"real_forward": frac(fwd_ms), "real_inverse": frac(inv_ms)},
        "in_cache": {"ms_per_step": round(ms_per_step_cached, 5),
                     "GFLOP_s": round(flops_step / (ms_per_step_cached * 1e-3) / 1e9, 2),
                     "launch_ms": {"real_forward": round(fwd_ms_cached, 5), "real_inverse": round(inv_ms_cached, 5)},
                     "frac_per_launch": {"real_forward": frac(fwd_ms_cached), "real_inverse": frac(inv_ms_cached)},
                     "working_set_bytes": set_bytes,
                     "timing": "the same launches on one buffer set (Infinity Cache resident between launches)"},
        "launch_distribution": dist,
    }

    # parity of the benchmarked outputs (sample of signals) against the CPU oracle and FP64
    idx = [0, 97, B - 1]
    o = O.Restatement(n, 3)
    yh = y.cpu().numpy()
    zh = z.cpu().numpy()
    parity = {
        "real_forward_vs_ref": O.split_norm_err(yh[idx], o.real_forward(rng_x[idx], 8)),
        "real_forward_relL2_fp64": O.rel_l2(yh[idx], O.interleave(O.fp64_real_forward(rng_x[idx]))),
        "real_inverse_vs_ref": O.split_norm_err(zh[idx], o.real_inverse(yh[idx], n, 8)),
        "real_inverse_relL2_fp64": O.rel_l2(zh[idx], O.fp64_real_inverse(yh[idx], n)),
        "gate": "vs_ref <= 1e-5 (max|d| / (||ref||/sqrt(len))), relL2_fp64 <= 1e-6",
    }
    parity = {k: (float(f"{v:.3e}") if isinstance(v, float) else v) for k, v in parity.items()}

    plain = None
    if not args.no_plain:
        tmp = torch.empty(2 * half, dtype=torch.float32, device=dev)
        for b in range(4):
            ctx.real_forward(n, x[b], y[b])
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for b in range(B):
            assert ctx.real_forward(n, x[b], y[b]) == 1
            assert ctx.real_inverse(n, y[b], z[b], tmp) == 1
        el = time.perf_counter() - t0
        plain_gflops = flops_step / el / 1e9
        plain = {"value": round(plain_gflops, 2), "unit": "GFLOP/s", "ms_per_step": round(el * 1e3, 3),
                 "calls": 2 * B, "batched_over_plain": round(value / plain_gflops, 1)}

    # host-pointer calls before the large legs below, on a small process footprint
    host = None if args.no_configs else host_pointer_leg(C, O)
    info = cpu_info()
    cpu = None if args.no_cpu else cpu_baseline(n, info=info)
    cpu_others = None if (args.no_cpu or args.no_configs) else cpu_other_configs(info)
    ctx.close()
    del x, y, z, xs, ys, zs
    uneven = None if args.no_uneven else uneven_leg(torch, C, O, dev, steps=10, warmup=10)
    serving = None if args.no_serve else serving_leg(torch, C, dev)
    split = None if args.no_split else split_bf16_leg(torch, C, O, dev)
    configs = None
    if not args.no_configs:
        configs = {"complex_4096_x1024": complex_4096_leg(torch, C, O, dev, rotate_mib=args.rotate_mib),
                   "complex_2p24": huge_leg(torch, C, O, dev),
                   "host_pointers": host,
                   "cpu_baseline": cpu_others}

    line = {
        "metric": METRIC, "value": round(value, 2), "unit": "GFLOP/s", "n_gpus": 1,
        "value_literal_5N": round(2 * value, 2),
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms_per_step, 5),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": workload, "n": n, "batch": B, "kinds": "real forward + real inverse",
                   "flops_convention": "2.5*N*log2(N) per real transform (value_literal_5N: 5*N*log2(N))",
                   "buffer_sets": nsets, "working_set_bytes": nsets * set_bytes,
                   "input": "iid N(0,1) float32, SplitMix64/Box-Muller seed 0xBE7C4"},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "plain": plain,
        "parity": parity,
        "batched_uneven": uneven,
        "serving": serving,
        "split_bf16x3": split,
        "other_configs": configs,
        "device": torch.cuda.get_device_name(dev),
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
