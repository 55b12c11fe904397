"""Pins the CPU oracle (oracle/sfft_oracle.c) before anything trusts it.

Each check is chosen so that a plausible mistake in the restatement (a wrong
twiddle sign, a transposed digit order, a dropped 1/2 in the real post-process,
an off-by-one in the size helpers) fails at least one of them:
  * golden vectors written by the reference itself (tests/golden),
  * a brute-force long-double DFT,
  * closed forms and invariants the reference tests use
    (tests/00-fft.cpp, tests/01-real.cpp).
"""
import os

import numpy as np
import pytest

from oracle import pyoracle as po
from tests.conftest import GOLDEN

TEST_SIZES = [1, 2, 4, 8, 16, 32, 64, 128, 256, 3, 6, 9, 12, 18, 24, 5, 10, 15, 20, 25,
              7, 14, 21, 28, 49, 11, 13, 17, 19, 22, 23]  # tests/00-fft.cpp:8-16


def _load(name):
    return np.load(os.path.join(GOLDEN, name))


# ------------------------------------------------------------ size helpers --

def test_size_helpers_match_reference_tables():
    t = _load("sizes.npz")
    mn = np.array([po.size_minimum(int(n)) for n in t["n_all"]])
    mx = np.array([po.size_maximum(int(n)) for n in t["n_all"]])
    np.testing.assert_array_equal(mn, t["fft_min"])
    np.testing.assert_array_equal(mx, t["fft_max"])
    rmn = np.array([po.real_size_minimum(int(n)) for n in t["n_pos"]])
    rmx = np.array([po.real_size_maximum(int(n)) for n in t["n_pos"]])
    np.testing.assert_array_equal(rmn, t["real_min"])
    np.testing.assert_array_equal(rmx, t["real_max"])


def _factor235(n):
    p = {2: 0, 3: 0, 5: 0}
    for f in (2, 3, 5):
        while n % f == 0:
            n //= f
            p[f] += 1
    return p, n


def test_size_helpers_reference_properties():
    # tests/00-fft.cpp:168-187
    for i in range(1, 1000):
        above, below = po.size_minimum(i), po.size_maximum(i)
        assert above >= i and below <= i
        pa, ra = _factor235(above)
        pb, rb = _factor235(below)
        assert ra == 1 and rb == 1
        assert pa[3] + pa[5] <= 2 and pb[3] + pb[5] <= 2


def test_size_helper_quirks():
    # Values the reference's size helpers return (signalsmith-fft.h:317-348,
    # :403-441): never a multiple of 5; RealFFT helpers built on the half size.
    assert po.size_minimum(5) == 6 and po.size_minimum(10) == 12
    assert po.size_minimum(1000) == 1024
    assert po.size_minimum(3000) == 3072 and po.size_maximum(3000) == 2304
    assert po.real_size_minimum(2048) == 2050
    assert po.real_size_maximum(10) == 12


# ------------------------------------------------------------- complex FFT --

def test_c2c_matches_reference_golden():
    g = _load("c2c_f64.npz")
    sizes = sorted(int(k[3:]) for k in g.files if k.startswith("in_"))
    assert len(sizes) >= 40
    for n in sizes:
        x = g[f"in_{n}"]
        assert po.rel_l2(po.fft(x), g[f"fwd_{n}"]) <= 1e-14, n
        assert po.rel_l2(po.ifft(x), g[f"inv_{n}"]) <= 1e-14, n


@pytest.mark.parametrize("n", TEST_SIZES + [60, 120, 243, 625, 1920, 3000])
def test_c2c_matches_bruteforce_dft(n):
    rng = np.random.default_rng(n)
    x = rng.uniform(-0.5, 0.5, n) + 1j * rng.uniform(-0.5, 0.5, n)
    assert po.rel_l2(po.fft(x), po.dft_ld(x, -1)) <= 1e-13
    assert po.rel_l2(po.ifft(x), po.dft_ld(x, +1)) <= 1e-13


@pytest.mark.parametrize("n", TEST_SIZES)
def test_c2c_single_bins(n):
    # tests/00-fft.cpp:19-60: e^{+2 pi i b t/N} -> N delta[k - b]
    t = np.arange(n)
    for b in range(n):
        x = np.exp(2j * np.pi * t * b / n)
        expect = np.zeros(n, complex)
        expect[b] = n
        y = po.fft(x)
        assert np.max(np.abs(y - expect)) <= 1e-9 * n, (n, b)


def test_dft_ld_closed_forms():
    for n in (1, 2, 3, 7, 16, 30):
        d = np.zeros(n, complex)
        d[0] = 1
        np.testing.assert_allclose(po.dft_ld(d, -1), np.ones(n), atol=1e-15)
        np.testing.assert_allclose(po.dft_ld(np.ones(n, complex), -1),
                                   np.eye(1, n, 0).ravel() * n, atol=1e-12)
        t = np.arange(n)
        for b in range(n):
            x = np.exp(-2j * np.pi * t * b / n)  # conj bin: inverse maps to N delta
            y = po.dft_ld(x, +1)
            expect = np.zeros(n, complex)
            expect[b] = n
            np.testing.assert_allclose(y, expect, atol=1e-12 * n)


@pytest.mark.parametrize("n", [1, 5, 49, 1920, 3000, 4096, 65536])
def test_c2c_invariants(n):
    rng = np.random.default_rng(7 + n)
    a = rng.normal(size=n) + 1j * rng.normal(size=n)
    b = rng.normal(size=n) + 1j * rng.normal(size=n)
    fa, fb = po.fft(a), po.fft(b)
    assert po.rel_l2(po.fft(a + b), fa + fb) <= 1e-14          # linearity
    assert po.rel_l2(po.ifft(fa), n * a) <= 1e-14              # unnormalised inverse
    assert abs(np.vdot(fa, fa).real - n * np.vdot(a, a).real) <= 1e-12 * n * np.vdot(a, a).real
    batch = np.stack([a, b])
    assert po.rel_l2(po.fft(batch), np.stack([fa, fb])) == 0.0  # batch rows independent


def test_reference_float_error_level_pinned():
    # Reference FFT<float> against this FP64 oracle: the error level the GPU
    # FP32 path must meet or beat (rel L2 well under 1e-6).
    g = _load("c2c_f32.npz")
    for n in (1920, 3000, 4096):
        x = g[f"in_{n}"]
        err = po.rel_l2(g[f"fwd_{n}"], po.fft(x.astype(np.complex128)))
        assert 1e-8 < err < 1e-6, (n, err)


# ---------------------------------------------------------------- RealFFT --

def test_real_matches_reference_golden():
    g = _load("real_f64.npz")
    sizes = sorted(int(k[3:]) for k in g.files if k.startswith("in_"))
    for n in sizes:
        x, spec = g[f"in_{n}"], g[f"spec_{n}"]
        for tag, hs in (("real", False), ("mod", True)):
            fwd = po.real_fft(x, half_shift=hs)
            assert po.rel_l2(fwd, g[f"{tag}_fwd_{n}"]) <= 1e-14, (tag, n)
            assert po.rel_l2(po.real_ifft(fwd, half_shift=hs), g[f"{tag}_rt_{n}"]) <= 1e-14, (tag, n)
            assert po.rel_l2(po.real_ifft(spec, half_shift=hs), g[f"{tag}_inv_{n}"]) <= 1e-14, (tag, n)


@pytest.mark.parametrize("n", list(range(2, 100, 2)) + [2048])
def test_real_against_complex(n):
    # tests/01-real.cpp:17-100 with the brute-force DFT as the complex side.
    rng = np.random.default_rng(n)
    x = rng.uniform(-0.5, 0.5, n)
    X = po.dft_ld(x.astype(complex), -1)
    y = po.real_fft(x)
    assert abs(y[0].real - X[0].real) <= 1e-9 * n
    assert abs(y[0].imag - X[n // 2].real) <= 1e-9 * n
    np.testing.assert_allclose(y[1:], X[1:n // 2], atol=1e-9 * n)
    # round trip = n x
    np.testing.assert_allclose(po.real_ifft(y), n * x, atol=1e-9 * n)
    # modified: bins at k + 1/2 = DFT of x[t] e^{-i pi t / n}
    xr = x * np.exp(-1j * np.pi * np.arange(n) / n)
    Xm = po.dft_ld(xr, -1)
    ym = po.real_fft(x, half_shift=True)
    np.testing.assert_allclose(ym, Xm[: n // 2], atol=1e-9 * n)
    np.testing.assert_allclose(po.real_ifft(ym, half_shift=True), n * x, atol=1e-9 * n)
