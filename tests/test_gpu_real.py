"""GPU parity of RealFFT<V> / ModifiedRealFFT<V> through the C ABI.

Reference tests re-expressed: tests/01-real.cpp (bins against the complex FFT,
bin-0 packing, the second half of the output left untouched, ifft against
the complex ifft, the half-bin shift).  Golden vectors: tests/golden/real_f64.npz
written by the reference's RealFFT<double> and ModifiedRealFFT<double>.
"""
import os

import numpy as np
import pytest
import torch

import fft_amd
from fft_amd import synth
from oracle import pyoracle as po
from tests.conftest import GOLDEN

pytestmark = pytest.mark.gpu

TOL32 = 1e-6
TOL64 = 1e-13
REAL_SIZES = list(range(2, 100, 2)) + [2048]


def _cls(mod):
    return fft_amd.ModifiedRealFFT if mod else fft_amd.RealFFT


def _fwd(n, x, mod=False, dtype=torch.float64):
    f = _cls(mod)(n, dtype=dtype)
    xd = torch.from_numpy(np.ascontiguousarray(x)).cuda()
    yd = torch.empty(xd.shape[:-1] + (n // 2,), dtype=f.cdtype, device="cuda")
    f.fft(xd, yd)
    torch.cuda.synchronize()
    return yd.cpu().numpy()


def _inv(n, y, mod=False, dtype=torch.float64):
    f = _cls(mod)(n, dtype=dtype)
    yd = torch.from_numpy(np.ascontiguousarray(y)).cuda()
    xd = torch.empty(yd.shape[:-1] + (n,), dtype=dtype, device="cuda")
    f.ifft(yd, xd)
    torch.cuda.synchronize()
    return xd.cpu().numpy()


def test_golden_real_double():
    from fft_amd import _lib as L
    g = np.load(os.path.join(GOLDEN, "real_f64.npz"))
    checked = 0
    for n in REAL_SIZES:
        if not L.lib().sfft_size_supported(n, L.R2C):
            continue
        checked += 1
        x, spec = g[f"in_{n}"], g[f"spec_{n}"]
        for tag, mod in (("real", False), ("mod", True)):
            fwd = _fwd(n, x, mod)
            assert po.rel_l2(fwd, g[f"{tag}_fwd_{n}"]) <= TOL64, (tag, n)
            assert po.rel_l2(_inv(n, fwd, mod), g[f"{tag}_rt_{n}"]) <= TOL64, (tag, n)
            assert po.rel_l2(_inv(n, spec, mod), g[f"{tag}_inv_{n}"]) <= TOL64, (tag, n)
    assert checked >= 30


@pytest.mark.parametrize("mod", [False, True])
def test_reference_real_properties(mod):
    # tests/01-real.cpp:17-100 with the GPU complex FFT<double> as the partner
    from fft_amd import _lib as L
    for n in range(2, 100, 2):
        if not L.lib().sfft_size_supported(n, L.R2C):
            continue
        v = synth.real_uniform(n, (n,), np.float64) * 0.5
        cin = v.astype(np.complex128)
        if mod:
            cin = cin * np.exp(-1j * np.pi * np.arange(n) / n)
        f = fft_amd.FFT(n, dtype=torch.complex128)
        cd = torch.from_numpy(cin).cuda()
        cmid = torch.empty_like(cd)
        f.fft(cd, cmid)
        rf = _cls(mod)(n, dtype=torch.float64)
        rd = torch.from_numpy(v).cuda()
        mid = torch.full((n,), 52 + 21j, dtype=torch.complex128, device="cuda")
        rf.fft(rd, mid[: n // 2])
        torch.cuda.synchronize()
        cm, rm = cmid.cpu().numpy(), mid.cpu().numpy()
        if not mod:
            assert abs(cm[0].imag) <= 1e-6
            assert abs(cm[0].real - rm[0].real) <= 1e-6
            assert abs(cm[n // 2].real - rm[0].imag) <= 1e-6
        for i in range(0 if mod else 1, n // 2):
            assert abs(cm[i] - rm[i]) <= n * 1e-6, (n, i)
        assert np.all(rm[n // 2:] == 52 + 21j), "second half of the output was written"
        cout = torch.empty_like(cd)
        f.ifft(cmid, cout)
        rout = torch.empty(n, dtype=torch.float64, device="cuda")
        rf.ifft(mid[: n // 2], rout)
        torch.cuda.synchronize()
        co, ro = cout.cpu().numpy(), rout.cpu().numpy()
        if mod:
            co = co * np.exp(1j * np.pi * np.arange(n) / n)
        assert np.all(np.abs(co.imag) <= n * 1e-6)
        assert np.all(np.abs(co.real - ro) <= n * 1e-6)


def test_real_config_2048_frames():
    # RealFFT<float> N=2048 over 16384 frames of 8 sinusoids + 0.01 N(0,1)
    n, frames = 2048, 16384
    x = synth.sinusoid_frames(2048, frames, n)
    f = fft_amd.RealFFT(n)
    xd = torch.from_numpy(x).cuda()
    yd = torch.empty(frames, n // 2, dtype=torch.complex64, device="cuda")
    zd = torch.empty_like(xd)
    f.fft(xd, yd)
    f.ifft(yd, zd)
    torch.cuda.synchronize()
    rows = np.linspace(0, frames - 1, 16).astype(int)
    y = yd[rows].cpu().numpy()
    assert po.rel_l2(y, po.real_fft(x[rows].astype(np.float64))) <= TOL32
    zi = _inv(n, y, dtype=torch.float32)
    assert po.rel_l2(zi, po.real_ifft(y.astype(np.complex128))) <= TOL32
    rt = ((zd.double() / n - xd.double()).norm() / xd.double().norm()).item()
    assert rt <= 2 * TOL32, rt


@pytest.mark.parametrize("n", [6, 10, 30, 250, 1000, 1920, 4096, 6000, 16384, 32768,
                               65536, 96000, 2 ** 18])
@pytest.mark.parametrize("mod", [False, True])
def test_real_float_sizes(n, mod):
    x = synth.real_uniform(n + 3, (3, n), np.float32)
    y = _fwd(n, x, mod, dtype=torch.float32)
    assert po.rel_l2(y, po.real_fft(x.astype(np.float64), half_shift=mod)) <= TOL32
    z = _inv(n, y, mod, dtype=torch.float32)
    assert po.rel_l2(z, po.real_ifft(y.astype(np.complex128), half_shift=mod)) <= TOL32


@pytest.mark.parametrize("n,batch", [(1 << 24, 1), (1 << 22, 2)])
@pytest.mark.parametrize("mod", [False, True])
def test_real_long_few_frames(n, batch, mod):
    """Long two-pass real transforms in small batches (the forward pass'
    pairwise post-process then spreads one transform over many workgroups),
    forward and inverse against the oracle (signalsmith-fft.h:446-502)."""
    x = synth.real_uniform(n + 5, (batch, n), np.float32)
    y = _fwd(n, x, mod, dtype=torch.float32)
    err = po.rel_l2(y, po.real_fft(x.astype(np.float64), half_shift=mod))
    print("R2C", n, batch, mod, "forward rel L2", err)
    assert err <= TOL32
    z = _inv(n, y, mod, dtype=torch.float32)
    assert po.rel_l2(z, po.real_ifft(y.astype(np.complex128), half_shift=mod)) <= TOL32


def test_real_setters_mirror_reference():
    t = np.load(os.path.join(GOLDEN, "sizes.npz"))
    rf = fft_amd.RealFFT(2)
    for n in (4, 100, 1000, 2048, 3000):
        assert rf.setSize(n) == t["real_set"][n - 1]  # returns the half size
    for n in (5, 7, 100, 1000, 2047):
        # every length is planned (2047 -> 2050, half 1025 = 5^2 * 41, runs chirp-z)
        assert rf.setSizeMinimum(n) == t["real_setmin"][n - 1]
        assert rf.setSizeMaximum(n) == t["real_setmax"][n - 1]


@pytest.mark.parametrize("mod", [False, True])
@pytest.mark.parametrize("n", [12, 100, 4096, 65536])
def test_real_half_shift_pair_overlap_single_writer(n, mod):
    # h = n/2 even with the half shift: bins h/2-1 and h/2 are written by one
    # pair only; results are deterministic across repeated runs.
    x = synth.real_uniform(n + 11, (4, n), np.float32)
    y1 = _fwd(n, x, mod, dtype=torch.float32)
    y2 = _fwd(n, x, mod, dtype=torch.float32)
    assert np.array_equal(y1, y2)
    assert po.rel_l2(y1, po.real_fft(x.astype(np.float64), half_shift=mod)) <= TOL32
    z1 = _inv(n, y1, mod, dtype=torch.float32)
    z2 = _inv(n, y1, mod, dtype=torch.float32)
    assert np.array_equal(z1, z2)
