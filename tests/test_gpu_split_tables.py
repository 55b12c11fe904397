"""Two-pass plans whose split comes from the planner's per-precision measured
tables (16-bit storage, F32X, FP64; fft_amd/csrc/engine.cpp): the default plan
uses the listed first-pass length, gives the bits of a plan asked for that
split explicitly, and matches the FP64 oracle forward and inverse on the
same input (the unnormalised pair of signalsmith-fft.h:295-315, :374-386).

The tables are read from the library's source so every listed length is
covered as the tables change.
"""
import math
import os
import re

import numpy as np
import pytest
import torch

import fft_amd
from fft_amd import _lib as L
from fft_amd import synth
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu

ENGINE = os.path.join(os.path.dirname(__file__), "..", "fft_amd", "csrc", "engine.cpp")


def _split_table(name):
    src = open(ENGINE).read()
    array = "kMeasuredRealSplits" if name == "Real" else "kMeasuredSplits" + name
    m = re.search(r"%s\[\]\s*=\s*\{(.*?)\};" % array, src, re.S)
    assert m, name
    return [(int(a), int(b)) for a, b in re.findall(r"\{\s*(\d+)\s*,\s*(\d+)\s*\}", m.group(1))]


CASES = ([(L.F16, n, n1) for n, n1 in _split_table("16")]
         + [(L.BF16, n, n1) for n, n1 in _split_table("16")]
         + [(L.F32X, n, n1) for n, n1 in _split_table("F32X")]
         + [(L.F64, n, n1) for n, n1 in _split_table("F64")])
NAME = {L.F16: "f16", L.BF16: "bf16", L.F32X: "f32x", L.F64: "f64"}
HALF = {L.F16: torch.float16, L.BF16: torch.bfloat16}


def _device_input(prec, n, seed):
    """Input on the device in the plan's storage format, and its exact values
    as complex128 for the oracle."""
    if prec in HALF:
        g = torch.Generator(device="cpu").manual_seed(seed)
        xd = torch.randn(1, n, 2, generator=g).to(HALF[prec]).cuda()
        xv = xd.float().cpu().numpy().astype(np.float64)
        return xd, xv[..., 0] + 1j * xv[..., 1]
    dt = np.complex128 if prec == L.F64 else np.complex64
    x = synth.complex_normal(seed, (1, n), dt)
    return torch.from_numpy(x).cuda(), x.astype(np.complex128)


def _host(yd, prec):
    if prec in HALF:
        y = yd.float().cpu().numpy().astype(np.float64)
        return y[..., 0] + 1j * y[..., 1]
    return yd.cpu().numpy().astype(np.complex128)


def _round_half(z, prec):
    t = torch.from_numpy(np.stack([z.real, z.imag], -1).astype(np.float32)).to(HALF[prec]).float()
    t = t.numpy().astype(np.float64)
    return t[..., 0] + 1j * t[..., 1]


def test_tables_present():
    assert len(_split_table("16")) >= 4
    assert len(_split_table("F32X")) >= 2
    assert len(_split_table("F64")) >= 2


@pytest.mark.parametrize("prec,n,n1", CASES,
                         ids=[f"{NAME[p]}-{n}-{n1}" for p, n, n1 in CASES])
def test_measured_split(prec, n, n1):
    plan = fft_amd.Plan(n, L.C2C, prec)
    info = plan.info()
    assert info["passes"] == 2, info
    assert math.prod(info["radices"]) == n1, info  # first-pass radices
    xd, x = _device_input(prec, n, n + 5)
    sh = torch.cuda.current_stream().cuda_stream
    y1, y2, z = torch.empty_like(xd), torch.empty_like(xd), torch.empty_like(xd)
    plan.execute_ptr(xd.data_ptr(), y1.data_ptr(), 1, L.FORWARD, sh)
    fft_amd.Plan(n, L.C2C, prec, split=n1).execute_ptr(xd.data_ptr(), y2.data_ptr(), 1,
                                                     L.FORWARD, sh)
    # inverse of the same input (a round trip would overflow 16-bit storage:
    # the unnormalised pair scales by n)
    plan.execute_ptr(xd.data_ptr(), z.data_ptr(), 1, L.INVERSE, sh)
    torch.cuda.synchronize()
    iv = torch.int16 if prec in HALF else torch.int32
    assert torch.equal(y1.view(iv), y2.view(iv))

    y, zi = _host(y1, prec), _host(z, prec)
    want = po.fft(x)
    err = po.rel_l2(y, want)
    wanti = po.ifft(x)
    erri = po.rel_l2(zi, wanti)
    print(NAME[prec], n, n1, "forward rel L2", err, "inverse", erri)
    if prec in HALF:
        # storage rounding dominates: the output rounding plus the stored
        # intermediate between the passes
        ideal = po.rel_l2(_round_half(want, prec), want)
        assert err <= 2.0 * ideal + 1e-6, (err, ideal)
        assert erri <= 2.0 * po.rel_l2(_round_half(wanti, prec), wanti) + 1e-6
    else:
        tol = 1e-7 if prec == L.F32X else 1e-12
        assert err <= tol and erri <= tol, (err, erri)


def _real_cases():
    t = _split_table("Real")
    picks = {t[0], t[len(t) // 2], t[-1]} | {e for e in t if e[0] == 30000}
    return sorted(picks)


@pytest.mark.parametrize("h,n1", _real_cases(), ids=lambda v: str(v))
def test_measured_real_split(h, n1):
    """RealFFT<float> two-pass plans take their half length's table split
    (first, middle, last and one re-measured entry): forward and inverse
    against the oracle (signalsmith-fft.h:446-502), forward bitwise against
    the plan asked for that split."""
    n = 2 * h
    plan = fft_amd.Plan(n, L.R2C, L.F32)
    info = plan.info()
    assert info["passes"] == 2 and math.prod(info["radices"]) == n1, info
    x = synth.real_uniform(n + 3, (2, n), np.float32)
    xd = torch.from_numpy(x).cuda()
    sh = torch.cuda.current_stream().cuda_stream
    y1 = torch.empty(2, h, dtype=torch.complex64, device="cuda")
    y2 = torch.empty_like(y1)
    z = torch.empty_like(xd)
    plan.execute_ptr(xd.data_ptr(), y1.data_ptr(), 2, L.FORWARD, sh)
    fft_amd.Plan(n, L.R2C, L.F32, split=n1).execute_ptr(xd.data_ptr(), y2.data_ptr(), 2,
                                                       L.FORWARD, sh)
    plan.execute_ptr(y1.data_ptr(), z.data_ptr(), 2, L.INVERSE, sh)
    torch.cuda.synchronize()
    assert torch.equal(y1, y2)
    y = y1.cpu().numpy()
    err = po.rel_l2(y, po.real_fft(x.astype(np.float64)))
    erri = po.rel_l2(z.cpu().numpy(), po.real_ifft(y.astype(np.complex128)))
    print("R2C", n, n1, "forward rel L2", err, "inverse", erri)
    assert err <= 1e-6 and erri <= 1e-6, (err, erri)
