"""Every compiled schedule variant of every compile-time planned length (the
autotuner may pick any of them): forward and inverse against numpy's FFT
(exact to rounding in double, independent of this library), and variants
with the same radices as the default give the default's bits.  The lengths
of the measured kernel tables (fixed_single*, fixed_real_single*,
fixed_f64_single*) are read from those sources, and each is also checked
against the runtime-planned pass of the same length: bitwise where the
radices match."""
import glob
import os
import re

import numpy as np
import pytest
import torch

import fft_amd
from fft_amd import _lib as L
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu

TOL = {L.F32: 2e-6, L.F64: 1e-12}
# the single-pass compiled lengths (the compiled column passes are exercised
# by two-pass plans)
CASES = [(1 << k, L.C2C, L.F32) for k in range(8, 15)] + [
    (1920, L.C2C, L.F32), (3000, L.C2C, L.F32)] + [
    (n, L.R2C, L.F32) for n in (1024, 2048, 4096, 8192, 16384)] + [
    (1 << k, L.C2C, L.F64) for k in range(8, 14)] + [
    (n, L.R2C, L.F64) for n in (2048, 4096, 8192)]


CSRC = os.path.join(os.path.dirname(__file__), "..", "fft_amd", "csrc")


def _table(pattern, rx):
    out = set()
    for f in glob.glob(os.path.join(CSRC, pattern)):
        out |= {int(m) for m in re.findall(rx, open(f).read())}
    return sorted(out)


TABLE_CASES = ([(n, L.C2C, L.F32) for n in _table("fixed_single*.hip", r"FixedC2C<float, (\d+),")] +
               [(2 * h, L.R2C, L.F32)
                for h in _table("fixed_real_single*.hip", r"FixedReal<float, (\d+),")] +
               [(n, L.C2C, L.F64) for n in _table("fixed_f64_single*.hip", r"FixedC2C<double, (\d+),")])


def _cases():
    return CASES + TABLE_CASES


def test_tables_found():
    assert len(TABLE_CASES) > 250, len(TABLE_CASES)


@pytest.mark.parametrize("n,kind,prec", _cases())
def test_every_variant(n, kind, prec):
    nv = L.lib().sfft_plan_variants(n, kind, prec)
    assert nv >= 1
    batch = max(2, (1 << 16) // n)
    rng = np.random.default_rng(n + 31 * kind + 7 * prec)
    cdt = torch.complex64 if prec == L.F32 else torch.complex128
    rdt = torch.float32 if prec == L.F32 else torch.float64
    sh = torch.cuda.current_stream().cuda_stream
    default = fft_amd.Plan(n, kind, prec, variant=0)
    for direction in (0, 1):
        if kind == L.C2C:
            xa = rng.standard_normal((batch, n)) + 1j * rng.standard_normal((batch, n))
            x = torch.from_numpy(xa).to(cdt).cuda()
            xd = x.to(torch.complex128).cpu().numpy()
            want = np.fft.ifft(xd, axis=-1) * n if direction else np.fft.fft(xd, axis=-1)
            shape, odt = (batch, n), cdt
        elif direction == 0:
            x = torch.from_numpy(rng.standard_normal((batch, n))).to(rdt).cuda()
            want = po.real_fft(x.double().cpu().numpy())
            shape, odt = (batch, n // 2), cdt
        else:
            xa = rng.standard_normal((batch, n // 2)) + 1j * rng.standard_normal((batch, n // 2))
            x = torch.from_numpy(xa).to(cdt).cuda()
            want = po.real_ifft(x.to(torch.complex128).cpu().numpy())
            shape, odt = (batch, n), rdt
        y0 = torch.empty(shape, dtype=odt, device="cuda")
        default.execute_ptr(x.data_ptr(), y0.data_ptr(), batch, direction, sh)
        for v in range(nv):
            p = fft_amd.Plan(n, kind, prec, variant=v)
            y = torch.full(shape, float("nan"), dtype=odt, device="cuda")
            p.execute_ptr(x.data_ptr(), y.data_ptr(), batch, direction, sh)
            torch.cuda.synchronize()
            err = po.rel_l2(y.cpu().numpy(), want)
            assert err <= TOL[prec], (n, kind, prec, v, direction, p.info(), err)
            if p.info()["radices"] == default.info()["radices"]:
                assert torch.equal(y, y0), (n, kind, prec, v, direction)
        # the runtime-planned pass of the same length
        rt = fft_amd.Plan(n, kind, prec, variant=-2)
        y = torch.full(shape, float("nan"), dtype=odt, device="cuda")
        rt.execute_ptr(x.data_ptr(), y.data_ptr(), batch, direction, sh)
        torch.cuda.synchronize()
        assert po.rel_l2(y.cpu().numpy(), want) <= TOL[prec], (n, kind, prec, direction)
        if rt.info()["radices"] == default.info()["radices"]:
            assert torch.equal(y, y0), (n, kind, prec, "runtime", direction)
