"""Benchmark of the batched FP32 transform path on one MI355X.

Workload (one step): complex FFT<float> N in {1920, 3000, 16384} x 8192
transforms (N(0,1) inputs), forward and inverse, plus RealFFT<float> N=2048 over
16384 real frames (8 sinusoids + 0.01 N(0,1)), forward and inverse.  Inputs
are generated in HBM from fixed seeds before timing; every launch reads and
writes HBM.

value  = algorithmic bytes of the step / step time, in GB/s (16 N bytes per
         complex transform, 8 N per real transform of N reals).
Also reported: the plain per-call path (one FFT<float>::fft call per
transform), the dominant kernel against the HBM roofline, rel L2 error
against the reference's FFT<double>, and the reference CPU path on the host's
cores.

    python bench.py [--steps K] [--warmup W] [--gpus 1] [--no-cpu]
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
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
METRIC = "batched FP32 FFT GB/s (% HBM peak) vs plain per-call; serving p50/p99 latency"
C2C_SIZES = (1920, 3000, 16384)
C2C_BATCH = 8192
REAL_N, REAL_FRAMES = 2048, 16384


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--plain-sample", type=int, default=256,
                    help="transforms per size timed on the plain per-call path (0: skip)")
    ap.add_argument("--no-extras", action="store_true",
                    help="skip the uneven-batch and serving measurements")
    ap.add_argument("--uneven-only", action="store_true",
                    help="only run the uneven batch's persistent launch (5 launches after one "
                         "warm-up; for profiling its kernel) and print its timing")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only time a 1 GiB device-to-device copy (the reference kernel for "
                         "scaling HBM byte counters) and exit")
    return ap.parse_args()


def copy_calibration(torch):
    """Three 1 GiB device copies: 1 GiB read + 1 GiB written per launch."""
    nbytes = 1 << 30
    src = torch.empty(nbytes, dtype=torch.uint8, device="cuda")
    dst = torch.empty_like(src)
    src.fill_(1)
    dst.copy_(src)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(3):
        dst.copy_(src)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / 3
    print(json.dumps({"copy_bytes_read": nbytes, "copy_bytes_written": nbytes,
                      "ms": round(ms, 4), "GB/s": round(2 * nbytes / ms / 1e6, 1)}))


class Launch:
    """One batched launch of the step: a plan, buffers, a direction.  Every
    launch reads a buffer that no launch of the step writes (the inverse legs
    read a copy of the forward output made at set-up), so with the step's
    6 GB of traffic between two reads of a buffer no input is still in the
    256 MB last-level cache: every figure is an HBM measurement."""

    def __init__(self, name, plan, x, y, batch, direction, alg_bytes, flops):
        self.name, self.plan, self.x, self.y = name, plan, x, y
        self.batch, self.direction, self.alg_bytes = batch, direction, alg_bytes
        self.flops = flops
        self.times_ms = []

    def run(self, stream_handle):
        self.plan.execute_ptr(self.x.data_ptr(), self.y.data_ptr(), self.batch, self.direction,
                              stream_handle)


def build_workload(torch, fft_amd, L):
    from fft_amd import synth
    launches = []
    bufs = {}
    for n in C2C_SIZES:
        plan = fft_amd.Plan(n, L.C2C, L.F32)
        x = torch.empty(C2C_BATCH, n, dtype=torch.complex64, device="cuda")
        fft_amd.fill(x, seed=n, dist="normal")
        y = torch.empty_like(x)
        z = torch.empty_like(x)
        plan.execute_ptr(x.data_ptr(), y.data_ptr(), C2C_BATCH, L.FORWARD,
                         torch.cuda.current_stream().cuda_stream)
        yi = y.clone()  # the inverse legs' input: spectra, in a buffer the step never writes
        bufs[n] = (x, y, z, yi)
        ab = 16 * n * C2C_BATCH
        fl = 5 * n * math.log2(n) * C2C_BATCH
        launches.append(Launch(f"c2c_{n}_fwd", plan, x, y, C2C_BATCH, L.FORWARD, ab, fl))
        launches.append(Launch(f"c2c_{n}_inv", plan, yi, z, C2C_BATCH, L.INVERSE, ab, fl))
    # RealFFT frames: tones from the seeded stream, noise generated in HBM
    n, frames = REAL_N, REAL_FRAMES
    fr = torch.from_numpy(synth.uniform01(n, frames * 8).reshape(frames, 8) * (n / 2)).cuda()
    ph = torch.from_numpy(synth.uniform01(n ^ 0x5A5A5A5A, frames * 8).reshape(frames, 8)
                          * (2 * math.pi)).cuda()
    t = torch.arange(n, dtype=torch.float64, device="cuda")
    xr = torch.zeros(frames, n, dtype=torch.float64, device="cuda")
    for k in range(8):
        xr += torch.cos(2 * math.pi * fr[:, k:k + 1] * t[None, :] / n + ph[:, k:k + 1])
    noise = torch.empty(frames, n, dtype=torch.float32, device="cuda")
    fft_amd.fill(noise, seed=n ^ 0xA5A5A5A5, dist="normal")
    xr = (xr + 0.01 * noise.double()).float().contiguous()
    del noise
    yr = torch.empty(frames, n // 2, dtype=torch.complex64, device="cuda")
    zr = torch.empty_like(xr)
    plan = fft_amd.Plan(n, L.R2C, L.F32)
    plan.execute_ptr(xr.data_ptr(), yr.data_ptr(), frames, L.FORWARD,
                     torch.cuda.current_stream().cuda_stream)
    yri = yr.clone()
    bufs["real"] = (xr, yr, zr, yri)
    ab = 8 * n * frames
    fl = 2.5 * n * math.log2(n) * frames
    launches.append(Launch(f"r2c_{n}_fwd", plan, xr, yr, frames, L.FORWARD, ab, fl))
    launches.append(Launch(f"c2r_{n}_inv", plan, yri, zr, frames, L.INVERSE, ab, fl))
    torch.cuda.synchronize()
    return launches, bufs


def settle(torch, launches, seconds=0.3):
    """Set-up: run the step's launches for a fixed time so that the GPU's
    clocks have left their idle state before the warm-up steps (the first
    step after a pause measured 3 % slower than steady state)."""
    sh = torch.cuda.current_stream().cuda_stream
    t0 = time.perf_counter()
    while time.perf_counter() - t0 < seconds:
        for ln in launches:
            ln.run(sh)
        torch.cuda.synchronize()


def time_launches(torch, launches, steps):
    """Untimed pass for the per-launch figures: `steps` steps with every
    launch bracketed by events.  (Events between launches cost about 4 % of
    the step, so the timed steps bracket only the dominant launch.)"""
    stream = torch.cuda.current_stream()
    sh = stream.cuda_stream
    ev = [[(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
           for _ in launches] for _ in range(steps)]
    for s in range(steps):
        for i, ln in enumerate(launches):
            ev[s][i][0].record(stream)
            ln.run(sh)
            ev[s][i][1].record(stream)
    torch.cuda.synchronize()
    for s in range(steps):
        for i, ln in enumerate(launches):
            ln.times_ms.append(ev[s][i][0].elapsed_time(ev[s][i][1]))


def time_steps(torch, launches, steps, warmup, dom):
    """The timed steps: `warmup` untimed steps, then `steps` steps between two
    events, with the dominant launch `dom` bracketed by events of its own in
    every step (its launch time for the roofline, measured live)."""
    stream = torch.cuda.current_stream()
    sh = stream.cuda_stream
    for _ in range(warmup):
        for ln in launches:
            ln.run(sh)
    torch.cuda.synchronize()
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
          for _ in range(steps)]
    begin, end = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    begin.record(stream)
    for s in range(steps):
        for ln in launches:
            if ln is dom:
                ev[s][0].record(stream)
                ln.run(sh)
                ev[s][1].record(stream)
            else:
                ln.run(sh)
    end.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    total_ms = begin.elapsed_time(end)
    dom_times = [a.elapsed_time(b) for a, b in ev]
    return total_ms, wall, dom_times


def plain_per_call(torch, fft_amd, L, bufs, sample):
    """One FFT<float>::fft call (one launch, batch 1) per transform, on the
    same resident data; sample transforms per size and direction."""
    stream = torch.cuda.current_stream()
    sh = stream.cuda_stream
    total_bytes, total_ms = 0, 0.0
    per = {}
    for n in C2C_SIZES:
        plan = fft_amd.Plan(n, L.C2C, L.F32)
        x = bufs[n][0]
        y = torch.empty(sample + 4, n, dtype=torch.complex64, device="cuda")  # own outputs
        for direction in (L.FORWARD, L.INVERSE):
            for i in range(4):  # warm
                plan.execute_ptr(x[i].data_ptr(), y[i].data_ptr(), 1, direction, sh)
            a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            a.record(stream)
            for i in range(sample):
                plan.execute_ptr(x[i].data_ptr(), y[i].data_ptr(), 1, direction, sh)
            b.record(stream)
            torch.cuda.synchronize()
            ms = a.elapsed_time(b)
            by = 16 * n * sample
            per[f"c2c_{n}_{'fwd' if direction == L.FORWARD else 'inv'}"] = round(by / ms / 1e6, 2)
            total_bytes += by
            total_ms += ms
    return total_bytes / total_ms / 1e6, per


def fast_sizes_batched(torch, fft_amd, lo=64, hi=65536, target_bytes=256 << 20, reps=3):
    """Every fast size 2^a 3^b 5^c in [lo, hi] as one batched FFT<float>
    forward launch of about target_bytes of input (inputs the last-level
    cache does not hold: two buffer sets alternate): the median and the
    spread of GB/s over the sizes, and which kernel family ran them."""
    from fft_amd import _lib as L
    from fft_amd import workloads
    sh = torch.cuda.current_stream().cuda_stream
    sizes = workloads.fast_sizes(lo, hi)
    cap = target_bytes // 8
    bufs = [torch.empty(cap, dtype=torch.complex64, device="cuda") for _ in range(4)]
    for t in bufs:
        fft_amd.fill(t, seed=11)
    rates, compiled = [], 0
    # every plan first: plan creation (host work) does not leave the GPU idle
    # between the timed launches
    plans = [fft_amd.Plan(n, L.C2C, L.F32) for n in sizes]
    for n, p in zip(sizes, plans):
        batch = max(1, cap // n)
        compiled += L.lib().sfft_plan_variants(n, L.C2C, L.F32) > 0
        for i in range(2):
            p.execute_ptr(bufs[i].data_ptr(), bufs[2 + i].data_ptr(), batch, 0, sh)
        # the faster of two timed rounds: a transient stall of the device in
        # one round (seen once at 83 GB/s on a length that runs at 4 TB/s)
        # does not stand for the length
        ms = float("inf")
        for _ in range(2):
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            for r in range(2 * reps):
                p.execute_ptr(bufs[r % 2].data_ptr(), bufs[2 + r % 2].data_ptr(), batch, 0, sh)
            e1.record()
            torch.cuda.synchronize()
            ms = min(ms, e0.elapsed_time(e1) / (2 * reps))
        rates.append(16 * n * batch / ms / 1e6)
    del bufs
    torch.cuda.empty_cache()
    r = np.array(rates)
    return {"sizes": len(sizes), "compiled_kernels": int(compiled),
            "median_GBs": round(float(np.median(r)), 1),
            "p10_GBs": round(float(np.percentile(r, 10)), 1),
            "p90_GBs": round(float(np.percentile(r, 90)), 1),
            "min_GBs": round(float(r.min()), 1), "max_GBs": round(float(r.max()), 1),
            "bytes_per_launch": target_bytes}


def uneven_batch(torch, fft_amd, count=20000, reps=5, tune=True):
    """The uneven batch: 20000 transforms, fast sizes log-uniform in [64, 65536],
    direction 50/50, U[-1,1) inputs; one persistent graph-captured launch with
    the planner's plans, then with autotuned plans (every distinct size tuned
    first), vs the plain per-size loop (one batched launch per item)."""
    from fft_amd import _lib as L
    from fft_amd import workloads
    sizes, dirs = workloads.uneven_batch(seed=3, count=count)
    fft_amd.clear_tuning()
    b = fft_amd.UnevenBatch([int(x) for x in sizes], [int(d) for d in dirs])
    xin = torch.empty(b.in_bytes, dtype=torch.uint8, device="cuda")
    fft_amd.fill(xin.view(torch.float32), seed=3)
    out = torch.empty(b.out_bytes, dtype=torch.uint8, device="cuda")
    alg = b.in_bytes + b.out_bytes
    res = {}
    if not tune:
        legs = (("persistent", b.execute),)
    else:
        t0 = time.perf_counter()
        uniq = sorted(set(int(x) for x in sizes))
        for n in uniq:
            fft_amd.autotune(n, L.C2C, L.F32, max(1, (128 << 20) // (16 * n)))
        tune_s = time.perf_counter() - t0
        bt = fft_amd.UnevenBatch([int(x) for x in sizes], [int(d) for d in dirs])
        legs = (("persistent", b.execute), ("persistent_tuned", bt.execute),
                ("per_size_loop", b.execute_per_size))
    for name, fn in legs:
        fn(xin, out)
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            fn(xin, out)
        e1.record()
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / reps
        res[name] = {"ms": round(ms, 3), "GB/s": round(alg / ms / 1e6, 1)}
    if not tune:
        return res
    res["speedup"] = round(res["per_size_loop"]["ms"] / res["persistent"]["ms"], 2)
    res["speedup_tuned"] = round(res["per_size_loop"]["ms"] / res["persistent_tuned"]["ms"], 2)
    res["autotune"] = {"sizes": len(uniq), "seconds": round(tune_s, 2)}
    fft_amd.clear_tuning()
    res["transforms"] = count
    res["alg_bytes"] = alg
    del xin, out, b, bt
    torch.cuda.empty_cache()
    return res


def facade_round_trip(n=4096, reps=200):
    """Host-staged fft + ifft round trip through the C++ drop-in header
    (fft_amd/sfft_facade_latency: std::vector in, std::vector out, wall clock,
    median of `reps`), run as a child process."""
    import subprocess
    exe = os.path.join(ROOT, "fft_amd", "sfft_facade_latency")
    if not os.path.exists(exe):
        return {"error": "fft_amd/sfft_facade_latency not built"}
    r = subprocess.run([exe, str(n), str(reps)], capture_output=True, text=True, timeout=120)
    if r.returncode != 0:
        return {"error": (r.stdout + r.stderr)[-300:]}
    return json.loads(r.stdout.strip().splitlines()[-1])


def single_transform(torch, fft_amd, reps=200):
    """One FFT<float> / FFT<double> of N = 4096 plus its ifft round trip, U[-1,1]
    inputs (seed 4096): device-resident latency (HIP events around fft + ifft,
    median of `reps`), the host-staged latency of the C++ drop-in header
    (host vector in, host vector out, wall clock) and of the Python mirror
    (numpy in, numpy out), with the round-trip error."""
    from fft_amd import _lib as L
    from fft_amd import synth
    n = 4096
    out = {}
    for name, prec, cdt in (("f32", L.F32, torch.complex64), ("f64", L.F64, torch.complex128)):
        f = fft_amd.FFT(n, dtype=cdt)
        xh = synth.complex_uniform(n, (n,), np.complex64 if prec == L.F32 else np.complex128)
        x = torch.from_numpy(xh).cuda()
        y = torch.empty_like(x)
        z = torch.empty_like(x)
        ts = []
        for _ in range(reps):
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            f.fft(x, y)
            f.ifft(y, z)
            e1.record()
            torch.cuda.synchronize()
            ts.append(e0.elapsed_time(e1) * 1e3)
        rt = float(torch.linalg.vector_norm((z / n - x).to(torch.complex128)) /
                   torch.linalg.vector_norm(x.to(torch.complex128)))
        hs = []
        yh = np.empty_like(xh)
        zh = np.empty_like(xh)
        for _ in range(20):
            t0 = time.perf_counter()
            f.fft(xh, yh)
            f.ifft(yh, zh)
            hs.append((time.perf_counter() - t0) * 1e6)
        out[name] = {"device_round_trip_us": round(float(np.median(ts)), 2),
                     "python_mirror_host_staged_round_trip_us": round(float(np.median(hs)), 1),
                     "round_trip_rel_l2": float(f"{rt:.3e}")}
    fac = facade_round_trip(n, reps)
    for name in ("f32", "f64"):
        if name in fac:
            out[name]["host_staged_round_trip_us"] = fac[name]["host_staged_round_trip_us"]
            out[name]["facade_round_trip_rel_l2"] = fac[name]["round_trip_rel_l2"]
    if "error" in fac:
        out["facade_error"] = fac["error"]
    out["host_staged_via"] = "C++ drop-in header (include/signalsmith-fft.h), std::vector in/out"
    return out


def precision_sweep(torch, fft_amd, reps=5):
    """Long 1D transforms N = 2^20 x 16 and 2^24 x 1, N(0,1) complex inputs
    from a seed: FP64, FP32 and the 16-bit storage variants (IEEE half,
    bfloat16; radix-16 MFMA stages).  Errors are against the FP64 path on the
    same (storage-rounded) input values: rel L2 and max |error| / max |X|.
    GB/s counts N complex values read + written in the variant's storage
    (32 N bytes FP64, 16 N FP32, 8 N 16-bit)."""
    from fft_amd import _lib as L
    sh = torch.cuda.current_stream().cuda_stream
    out = {"note": "f32x stores FP32 and computes in FP64 (twiddles, LDS, butterflies); "
                   "f16 / bf16 store 16-bit values and run float butterflies; the "
                   "*_mfma_radix16 rows run the radix-16 stages on MFMA (split hi+lo operands)"}
    for n, batch, seed in ((1 << 20, 16, 20), (1 << 24, 1, 24)):
        g = torch.Generator(device="cuda").manual_seed(seed)
        base = torch.randn(batch, n, 2, device="cuda", generator=g)
        row = {}
        for name, prec, dt, eb, mfma in (
                ("f64", L.F64, torch.float64, 8, -1), ("f32", L.F32, torch.float32, 4, -1),
                ("f32x", L.F32X, torch.float32, 4, -1),
                ("f16", L.F16, torch.float16, 2, -1), ("bf16", L.BF16, torch.bfloat16, 2, -1),
                ("f16_mfma_radix16", L.F16, torch.float16, 2, 1),
                ("bf16_mfma_radix16", L.BF16, torch.bfloat16, 2, 1)):
            xs = base.to(dt).contiguous()
            ys = torch.empty_like(xs)
            plan = (fft_amd.Plan(n, L.C2C, prec) if mfma < 0 else
                    fft_amd.Plan(n, L.C2C, prec, mfma=mfma))
            plan.execute_ptr(xs.data_ptr(), ys.data_ptr(), batch, L.FORWARD, sh)
            torch.cuda.synchronize()
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            for _ in range(reps):
                plan.execute_ptr(xs.data_ptr(), ys.data_ptr(), batch, L.FORWARD, sh)
            e1.record()
            torch.cuda.synchronize()
            ms = e0.elapsed_time(e1) / reps
            # FP64 reference on the same input values
            xd = xs.double().contiguous()
            yd = torch.empty_like(xd)
            fft_amd.Plan(n, L.C2C, L.F64).execute_ptr(xd.data_ptr(), yd.data_ptr(), batch,
                                                       L.FORWARD, sh)
            d = ys.double() - yd
            rel = float(torch.linalg.vector_norm(d) / torch.linalg.vector_norm(yd))
            mx = float(d.abs().max() / yd.abs().max())
            info = plan.info()
            row[name] = {"ms": round(ms, 4), "GB/s": round(2 * eb * 2 * n * batch / ms / 1e6, 1),
                         "rel_l2_vs_fp64": float(f"{rel:.3e}"),
                         "max_abs_err_rel_to_max": float(f"{mx:.3e}"),
                         "passes": info["passes"], "mfma_stages": info["mfma_stages"]}
            del xs, ys, xd, yd, d
        out[f"n{n}_x{batch}"] = row
        del base
    torch.cuda.empty_cache()
    return out


def serving(rates=(2000, 5000, 10000, 20000, 50000), duration=0.25):
    """Poisson arrivals: C2C fwd / inv / RealFFT fwd (1/3 each), fast sizes
    log-uniform in [256, 2^18], batch U{1..512}, per-request seeds; 64 GB
    HBM budget.  p50/p99 latency (arrival -> completion) and throughput, and
    the plain path (one request at a time, synchronous) at the same rates."""
    from fft_amd.serving import Server, latency_summary
    out = {}
    with Server(budget_bytes=64 << 30, max_inflight=64, streams=4) as srv:
        for rate in rates:
            rec = srv.run_poisson(rate=rate, duration=duration, n_min=256, n_max=1 << 18,
                                  batch_max=512, seed=5)
            out[f"rate_{rate}"] = {k: (round(v, 4) if isinstance(v, float) else v)
                                   for k, v in latency_summary(rec).items()}
    with Server(budget_bytes=64 << 30, plain=True) as srv:
        for rate in rates:
            rec = srv.run_poisson(rate=rate, duration=duration, n_min=256, n_max=1 << 18,
                                  batch_max=512, seed=5)
            out[f"plain_rate_{rate}"] = {k: (round(v, 4) if isinstance(v, float) else v)
                                         for k, v in latency_summary(rec).items()}
    return out


def cpu_baseline(torch, bufs, launches):
    """The reference's own FFT<float> / RealFFT<float> on the host cores (one
    transform per thread, each thread its own instance), over the whole step's
    transforms, repeated until about 16 thread-seconds of CPU work; plus rel L2
    of sample GPU outputs against the reference's FFT<double> / RealFFT<double>."""
    from oracle import pyoracle as po
    # every CPU this process may run on, capped by the host's CPU share per
    # GPU where the host states one (OMP_NUM_THREADS)
    avail = len(os.sched_getaffinity(0))
    share = int(os.environ.get("OMP_NUM_THREADS", avail) or avail)
    threads = max(1, min(avail, share))
    model = ""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    out = {"threads": threads, "nproc": os.cpu_count(), "affinity_cpus": avail,
           "cpu_model": model}
    errors = {}
    sample_rows = np.linspace(0, C2C_BATCH - 1, 8).astype(int)
    if po.ref_available():
        kind = "reference"
        total_bytes, total_s = 0, 0.0
        host = {n: (bufs[n][0].cpu().numpy(), bufs[n][3].cpu().numpy()) for n in C2C_SIZES}
        xr, yr, zr_, yri = bufs["real"]
        xrh, yrh = xr.cpu().numpy(), yri.cpu().numpy()
        # the step's transforms, repeated until about 16 thread-seconds of CPU work
        passes = 0
        while passes == 0 or total_s * threads < 16.0:
            for n in C2C_SIZES:
                xh, yh = host[n]
                for inverse, src in ((False, xh), (True, yh)):
                    total_s += po.ref_time_c2c(src, inverse, threads)
                    total_bytes += 16 * n * C2C_BATCH
            total_s += po.ref_time_real(xrh, threads)
            total_s += po.ref_time_real(yrh, threads, inverse=True)
            total_bytes += 2 * 8 * REAL_N * REAL_FRAMES
            passes += 1
        for n in C2C_SIZES:
            x, y, z, yi = bufs[n]
            xh, yih = host[n]
            ref64 = po.ref_fft(xh[sample_rows].astype(np.complex128))
            errors[f"c2c_{n}_fwd"] = po.rel_l2(y[sample_rows].cpu().numpy(), ref64)
            zr = z[sample_rows].cpu().numpy()
            errors[f"c2c_{n}_inv"] = po.rel_l2(
                zr, po.ref_fft(yih[sample_rows].astype(np.complex128), inverse=True))
        xh = xrh
        rows = np.linspace(0, REAL_FRAMES - 1, 8).astype(int)
        yh = yr[rows].cpu().numpy()
        errors[f"r2c_{REAL_N}_fwd"] = po.rel_l2(yh, po.ref_real_fft(xh[rows].astype(np.float64)))
        errors[f"c2r_{REAL_N}_inv"] = po.rel_l2(
            zr_[rows].cpu().numpy(), po.ref_real_ifft(yrh[rows].astype(np.complex128)))
        value = total_bytes / total_s / 1e9
        out.update({"value": round(value, 3), "unit": "GB/s", "cores": threads, "kind": kind,
                    "sample": f"the whole step {passes}x: every transform of the step, forward "
                              "and inverse (reference FFT<float>/RealFFT<float>, g++ -O3 "
                              "-march=x86-64-v3), one instance per thread, on the host's CPU "
                              "share for one GPU"})
    else:
        kind = "port"
        x = bufs[1920][0][:64].cpu().numpy()
        t0 = time.perf_counter()
        po.fft(x.astype(np.complex128))
        dt = time.perf_counter() - t0
        out.update({"value": round(16 * 1920 * 64 / dt / 1e9, 3), "unit": "GB/s", "cores": 1,
                    "kind": kind, "sample": "oracle FFT<double> restatement, 64 x 1920"})
    return out, errors


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only; --gpus must be 1",
                          "n_gpus": args.gpus}))
        sys.exit(2)
    import torch

    import fft_amd
    from fft_amd import _lib as L

    torch.cuda.set_device(0)
    if args.copy_calibration:
        copy_calibration(torch)
        return
    if args.uneven_only:
        print(json.dumps(uneven_batch(torch, fft_amd, tune=False)))
        return
    launches, bufs = build_workload(torch, fft_amd, L)
    settle(torch, launches)
    time_launches(torch, launches, args.steps)
    dom = max(launches, key=lambda ln: float(np.sum(ln.times_ms)))
    total_ms, wall, dom_times = time_steps(torch, launches, args.steps, args.warmup, dom)
    step_bytes = sum(ln.alg_bytes for ln in launches)
    step_flops = sum(ln.flops for ln in launches)
    ms_per_step = total_ms / args.steps
    value = step_bytes / (ms_per_step * 1e-3) / 1e9

    per = {}
    for ln in launches:
        avg = float(np.mean(ln.times_ms))
        per[ln.name] = {"ms": round(avg, 4), "GB/s": round(ln.alg_bytes / avg / 1e6, 1),
                        "hbm_frac": round(ln.alg_bytes / avg / 1e6 / HBM_PEAK_GBS, 4)}
    dom_ms = float(np.mean(dom_times))
    achieved = dom.alg_bytes / dom_ms / 1e6
    traffic = None
    tfile = os.path.join(ROOT, "profiles", "hbm_traffic.json")
    if os.path.exists(tfile):
        with open(tfile) as f:
            traffic = json.load(f).get(dom.name)

    plain_gbs, plain_per = (plain_per_call(torch, fft_amd, L, bufs, args.plain_sample)
                            if args.plain_sample > 0 else (None, None))

    result = {
        "metric": METRIC,
        "value": round(value, 2),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic: seeded splitmix64 N(0,1) complex64 (c2c), 8 sinusoids + 0.01 N(0,1) "
                "frames (real), generated in HBM",
        "config": {
            "workload": "batched FFT<float> N in {1920, 3000, 16384} x 8192 fwd+inv and "
                        "RealFFT<float> N=2048 x 16384 frames fwd+inv",
            "c2c_sizes": list(C2C_SIZES), "c2c_batch": C2C_BATCH,
            "real_n": REAL_N, "real_frames": REAL_FRAMES,
            "bytes_per_step": step_bytes,
        },
        "hbm_frac": round(value / HBM_PEAK_GBS, 4),
        "gflops": round(step_flops / (ms_per_step * 1e-3) / 1e9, 1),
        "gflops_convention": "5 N log2 N per complex transform, 2.5 N log2 N per real one",
        "per_launch": per,
        "per_launch_note": "an untimed pass of the same steps with events around every "
                           "launch; the timed steps bracket only the roofline's launch",
        "plain_per_call": None if plain_gbs is None else {
            "value": round(plain_gbs, 2), "unit": "GB/s",
            "sample": f"{args.plain_sample} single-transform calls per size and direction",
            "per_size": plain_per},
        "speedup_vs_plain": None if plain_gbs is None else round(value / plain_gbs, 1),
        "roofline": {"bound": "hbm", "kernel": dom.name, "achieved": round(achieved, 1),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "launch_ms": round(dom_ms, 4), "alg_bytes_per_launch": dom.alg_bytes},
        "wall_s_timed": round(wall, 3),
    }
    if not args.no_extras:
        result["single_transform_4096"] = single_transform(torch, fft_amd)
        result["fast_sizes_batched"] = fast_sizes_batched(torch, fft_amd)
        result["uneven_batch"] = uneven_batch(torch, fft_amd)
        result["precision_sweep"] = precision_sweep(torch, fft_amd)
        result["serving"] = serving()
        sv = result["serving"]
        result["serving_p50_ms"] = sv["rate_10000"]["p50_ms"]
        result["serving_p99_ms"] = sv["rate_10000"]["p99_ms"]
    if not args.no_cpu:
        cpu, errors = cpu_baseline(torch, bufs, launches)
        result["cpu_baseline"] = cpu
        result["rel_l2_vs_reference_fp64"] = {k: float(f"{v:.3e}") for k, v in errors.items()}
    print(json.dumps(result))


if __name__ == "__main__":
    main()
