"""Seeded random parity sweep over the plan space: random lengths from 1 to
2^18 (mostly with large prime factors: chirp-z) and random fast lengths up to
2^20 (one- and two-pass, compiled and runtime-planned kernels), every kind, both directions, both precisions, batches of
1-5.  Complex results are checked against numpy's FFT (exact to rounding in
double, independent of this library); real results against the oracle's
real restatement, which the golden vectors of the reference pin."""
import numpy as np
import pytest
import torch

import fft_amd
from fft_amd import _lib as L
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu

TOL = {L.F32: 2e-6, L.F64: 1e-12}


def _smooth(n):
    for p in (2, 3, 5, 7, 11, 13):
        while n % p == 0:
            n //= p
    return n == 1


def _cases(count=60, seed=2024):
    rng = np.random.default_rng(seed)
    out = []
    while len(out) < count:
        kind = int(rng.choice([L.C2C, L.C2C, L.R2C, L.R2C_HALFSHIFT]))
        prec = int(rng.choice([L.F32, L.F64]))
        n = int(np.exp(rng.uniform(0, np.log(1 << 18))))
        if kind != L.C2C:
            n = max(2, n + (n % 2))
            if not _smooth(n // 2) and n // 2 > 8192:
                continue  # the oracle's prime step is O(h^2)
        out.append((n, kind, prec, int(rng.integers(1, 6)), int(rng.integers(0, 2))))
    return out


def _fast_cases(count=60, seed=77):
    from fft_amd import workloads
    fast = workloads.fast_sizes(2, 1 << 20)
    rng = np.random.default_rng(seed)
    out = []
    while len(out) < count:
        kind = int(rng.choice([L.C2C, L.C2C, L.R2C, L.R2C_HALFSHIFT]))
        prec = int(rng.choice([L.F32, L.F64]))
        n = int(fast[int(rng.integers(0, len(fast)))])
        if kind != L.C2C and n % 2:
            continue
        batch = int(rng.integers(1, 6)) if n <= (1 << 16) else 1
        out.append((n, kind, prec, batch, int(rng.integers(0, 2))))
    return out


@pytest.mark.parametrize("n,kind,prec,batch,direction", _cases() + _fast_cases())
def test_random_plan(n, kind, prec, batch, direction):
    rng = np.random.default_rng(n * 7 + kind)
    p = fft_amd.Plan(n, kind, prec)
    cdt = torch.complex64 if prec == L.F32 else torch.complex128
    rdt = torch.float32 if prec == L.F32 else torch.float64
    sh = torch.cuda.current_stream().cuda_stream
    hs = kind == L.R2C_HALFSHIFT
    if kind == L.C2C:
        xa = rng.standard_normal((batch, n)) + 1j * rng.standard_normal((batch, n))
        x = torch.from_numpy(xa).to(cdt).cuda()
        y = torch.empty_like(x)
        want = np.fft.ifft(xa, axis=-1) * n if direction else np.fft.fft(xa, axis=-1)
    elif direction == 0:
        xa = rng.standard_normal((batch, n))
        x = torch.from_numpy(xa).to(rdt).cuda()
        y = torch.empty(batch, n // 2, dtype=cdt, device="cuda")
        want = po.real_fft(x.double().cpu().numpy(), half_shift=hs)
    else:
        xa = rng.standard_normal((batch, n // 2)) + 1j * rng.standard_normal((batch, n // 2))
        x = torch.from_numpy(xa).to(cdt).cuda()
        y = torch.empty(batch, n, dtype=rdt, device="cuda")
        want = po.real_ifft(x.to(torch.complex128).cpu().numpy(), half_shift=hs)
    if kind == L.C2C:
        xa = x.to(torch.complex128).cpu().numpy()  # the values actually transformed
        want = np.fft.ifft(xa, axis=-1) * n if direction else np.fft.fft(xa, axis=-1)
    p.execute_ptr(x.data_ptr(), y.data_ptr(), batch, direction, sh)
    torch.cuda.synchronize()
    err = po.rel_l2(y.cpu().numpy(), want)
    assert err <= TOL[prec], (n, kind, prec, batch, direction, p.info(), err)
