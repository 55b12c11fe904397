"""Results in the reference benchmark's own format, for its chart tooling.

The reference harness (benchmark/benchmark.h:25-97, benchmark/main.cpp)
times FFT<double> forward, out of place, one transform per call, for the
sizes 2^k and 3 * 2^k, 9 * 2^k below 2^24, and writes

    results/<tag>.csv   "size,ops/sec,<name>" then size, transforms/s, scaled
    results/<tag>.js    addResults("<name>", [{size: .., rate: scaled}, ...]);

with scaled = transforms/s * max(1, N ln N) * 1e-6 (its "normalised rate").
This module measures the same quantity on the GPU for the per-call path
(one sfft_execute of one transform per call, data resident in HBM) and for
the batched path (many transforms per call), in FP64 or FP32.

    python -m fft_amd.benchmark_csv [--precision f64|f32] [--mode single|batched]
                                    [--max-size 16777216] [--out results]
"""
from __future__ import annotations

import argparse
import math
import os

import numpy as np


def reference_sizes(max_size: int = 1 << 24) -> list:
    """benchmark/benchmark.h:27-52: powers of two up to max_size, and 3x, 9x below it."""
    out = []
    size = 1
    while size <= max_size:
        out.append(size)
        for mult in (3, 9):
            if size * mult < max_size:
                out.append(size * mult)
        size *= 2
    return sorted(out)


def measure(precision: str = "f64", mode: str = "single", max_size: int = 1 << 24,
            min_ms: float = 20.0) -> list:
    """[(size, transforms per second, scaled rate)] for each reference size."""
    import torch

    import fft_amd
    from fft_amd import _lib as L
    prec = L.F64 if precision == "f64" else L.F32
    cdt = torch.complex128 if precision == "f64" else torch.complex64
    sh = torch.cuda.current_stream().cuda_stream
    rows = []
    for n in reference_sizes(max_size):
        batch = 1 if mode == "single" else max(1, (1 << 23) // n)
        plan = fft_amd.Plan(n, L.C2C, prec)
        g = torch.Generator(device="cuda").manual_seed(n)
        x = (torch.rand(batch, n, 2, device="cuda", generator=g, dtype=torch.float64) - 0.5)
        x = torch.view_as_complex(x).to(cdt).contiguous()  # U[-1/2, 1/2), as randomComplex
        y = torch.empty_like(x)
        plan.execute_ptr(x.data_ptr(), y.data_ptr(), batch, L.FORWARD, sh)
        torch.cuda.synchronize()
        reps = 1
        while True:
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            for _ in range(reps):
                plan.execute_ptr(x.data_ptr(), y.data_ptr(), batch, L.FORWARD, sh)
            e1.record()
            torch.cuda.synchronize()
            ms = e0.elapsed_time(e1)
            if ms >= min_ms or reps >= 1 << 16:
                break
            reps *= 4
        rate = reps * batch / (ms * 1e-3)
        rows.append((n, rate, rate * max(1.0, n * math.log(n)) * 1e-6))
        del x, y
    return rows


def write(rows, name: str, prefix: str) -> None:
    os.makedirs(os.path.dirname(prefix) or ".", exist_ok=True)
    with open(prefix + ".csv", "w") as f:
        f.write(f"size,ops/sec,{name}\n")
        for n, rate, scaled in rows:
            f.write(f"{n},{rate:.15g},{scaled:.15g}\n")
    with open(prefix + ".js", "w") as f:
        f.write(f'addResults("{name}", [')
        f.write(",".join(f"\n\t{{size: {n}, rate: {scaled:.15g}}}" for n, _, scaled in rows))
        f.write("\n]);")


def main(argv=None) -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("--precision", choices=("f64", "f32"), default="f64")
    ap.add_argument("--mode", choices=("single", "batched"), default="single")
    ap.add_argument("--max-size", type=int, default=1 << 24)
    ap.add_argument("--out", default="results")
    args = ap.parse_args(argv)
    rows = measure(args.precision, args.mode, args.max_size)
    tag = f"sfft-mi355x-{args.precision}-{args.mode}"
    name = f"sfft MI355X ({args.precision}, {args.mode})"
    write(rows, name, os.path.join(args.out, tag))
    for n, rate, scaled in rows:
        print(f"{n:>10} {rate:14.1f}/s  normalised {scaled:12.1f}")


if __name__ == "__main__":
    main()
