"""Writes the golden fixtures in this directory from the reference itself.

Runs the reference header (signalsmith-fft.h) compiled by `make -C oracle ref`
into oracle/_ref/libsfft_ref.so, on seeded inputs from fft_amd.synth, and saves
inputs + outputs as compressed .npz files (plain arrays, no pickles) plus a
manifest.  Run from the repository root:

    make -C oracle ref && python tests/golden/make_golden.py

Fixtures (inputs U[-1, 1), seed = n unless stated):
  c2c_f64.npz   FFT<double> fft and ifft of the same input, n in C2C_SIZES
  c2c_f32.npz   FFT<float> fft of the complex64 input, n in {1920, 3000, 4096}
  real_f64.npz  RealFFT<double> and ModifiedRealFFT<double>: fft of a real
                input, ifft of that spectrum and ifft of a random spectrum,
                n in {2, 4, ..., 98, 2048}
  sizes.npz     FFT::sizeMinimum/sizeMaximum for n in [0, 70000],
                RealFFT::sizeMinimum/sizeMaximum for n in [1, 70000],
                setSize/setSizeMinimum/setSizeMaximum return values and
                size() after construction with fastDirection -1/0/+1
"""
from __future__ import annotations

import json
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from fft_amd import synth  # noqa: E402
from oracle import pyoracle as po  # noqa: E402

HERE = os.path.dirname(os.path.abspath(__file__))

C2C_SIZES = list(range(1, 33)) + [48, 60, 64, 120, 243, 256, 625, 1024, 1920, 3000, 4096]
F32_SIZES = [1920, 3000, 4096]
REAL_SIZES = list(range(2, 100, 2)) + [2048]
SIZE_TABLE_MAX = 70000
SETTER_MAX = 5000


def main() -> None:
    if not po.ref_available():
        raise SystemExit("build the reference shim first: make -C oracle ref")
    manifest = {"generator": "tests/golden/make_golden.py",
                "reference": "signalsmith-fft.h (FFT<V>, RealFFT<V>, ModifiedRealFFT<V>)",
                "inputs": "fft_amd.synth counter-based splitmix64, U[-1,1), seed = n",
                "files": {}}

    arrs = {}
    for n in C2C_SIZES:
        x = synth.complex_uniform(n, (n,), np.complex128)
        arrs[f"in_{n}"] = x
        arrs[f"fwd_{n}"] = po.ref_fft(x, inverse=False)
        arrs[f"inv_{n}"] = po.ref_fft(x, inverse=True)
    np.savez_compressed(os.path.join(HERE, "c2c_f64.npz"), **arrs)
    manifest["files"]["c2c_f64.npz"] = {"sizes": C2C_SIZES, "keys": "in_N, fwd_N, inv_N",
                                        "precision": "FFT<double>"}

    arrs = {}
    for n in F32_SIZES:
        x = synth.complex_uniform(n, (n,), np.complex64)
        arrs[f"in_{n}"] = x
        arrs[f"fwd_{n}"] = po.ref_fft(x, inverse=False)
    np.savez_compressed(os.path.join(HERE, "c2c_f32.npz"), **arrs)
    manifest["files"]["c2c_f32.npz"] = {"sizes": F32_SIZES, "keys": "in_N, fwd_N",
                                        "precision": "FFT<float>"}

    arrs = {}
    for n in REAL_SIZES:
        x = synth.real_uniform(n, (n,), np.float64)
        spec = synth.complex_uniform(n + 1_000_000, (n // 2,), np.complex128)
        arrs[f"in_{n}"] = x
        arrs[f"spec_{n}"] = spec
        for tag, hs in (("real", False), ("mod", True)):
            fwd = po.ref_real_fft(x, half_shift=hs)
            arrs[f"{tag}_fwd_{n}"] = fwd
            arrs[f"{tag}_rt_{n}"] = po.ref_real_ifft(fwd, half_shift=hs)
            arrs[f"{tag}_inv_{n}"] = po.ref_real_ifft(spec, half_shift=hs)
    np.savez_compressed(os.path.join(HERE, "real_f64.npz"), **arrs)
    manifest["files"]["real_f64.npz"] = {
        "sizes": REAL_SIZES,
        "keys": "in_N (real), spec_N (random half spectrum, seed N+1e6), "
                "{real,mod}_fwd_N, {real,mod}_rt_N = ifft(fwd), {real,mod}_inv_N = ifft(spec)",
        "precision": "RealFFT<double>, ModifiedRealFFT<double>"}

    n_all = np.arange(0, SIZE_TABLE_MAX + 1, dtype=np.int64)
    n_pos = np.arange(1, SIZE_TABLE_MAX + 1, dtype=np.int64)
    n_set = np.arange(1, SETTER_MAX + 1, dtype=np.int64)
    tables = {
        "n_all": n_all,
        "fft_min": np.array([po.ref_sizes("size_minimum", int(n)) for n in n_all], dtype=np.int64),
        "fft_max": np.array([po.ref_sizes("size_maximum", int(n)) for n in n_all], dtype=np.int64),
        "n_pos": n_pos,
        "real_min": np.array([po.ref_sizes("real_size_minimum", int(n)) for n in n_pos], dtype=np.int64),
        "real_max": np.array([po.ref_sizes("real_size_maximum", int(n)) for n in n_pos], dtype=np.int64),
        "n_set": n_set,
    }
    for real in (0, 1):
        tag = "real" if real else "fft"
        for which, wname in enumerate(("set", "setmin", "setmax")):
            tables[f"{tag}_{wname}"] = np.array(
                [po.ref_setter_return(bool(real), which, int(n)) for n in n_set], dtype=np.int64)
        for d in (-1, 0, 1):
            tables[f"{tag}_ctor_{d + 1}"] = np.array(
                [po.ref_ctor_size(bool(real), int(n), d) for n in n_set], dtype=np.int64)
    np.savez_compressed(os.path.join(HERE, "sizes.npz"), **tables)
    manifest["files"]["sizes.npz"] = {
        "keys": "fft_min/fft_max over n_all=[0,70000]; real_min/real_max over n_pos=[1,70000]; "
                "{fft,real}_{set,setmin,setmax} and {fft,real}_ctor_{0,1,2} (fastDirection -1,0,+1) "
                "over n_set=[1,5000]"}

    with open(os.path.join(HERE, "manifest.json"), "w") as f:
        json.dump(manifest, f, indent=1)
        f.write("\n")
    total = sum(os.path.getsize(os.path.join(HERE, k)) for k in manifest["files"])
    print(f"wrote {len(manifest['files'])} fixture files, {total / 1024:.0f} KiB")


if __name__ == "__main__":
    main()
