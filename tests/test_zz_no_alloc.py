"""sfft_execute_noalloc (workspace from the library's static device array)
against sfft_execute on the same plans and inputs: bitwise equal results on
two calls in a row, the second with a smaller batch, plus the FP64 oracle on
one transform; a call whose workspace exceeds the array is refused."""
import ctypes

import numpy as np
import pytest
import torch

import fft_amd
from fft_amd import _lib as L
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu


def _stream():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def _noalloc(plan, x, y, batch, d):
    return L.lib().sfft_execute_noalloc(plan.handle, ctypes.c_void_p(x.data_ptr()),
                                        ctypes.c_void_p(y.data_ptr()), batch, d, _stream())


def _input(n, kind, prec, batch, d, seed):
    cdt = torch.complex128 if prec == L.F64 else torch.complex64
    rdt = torch.float64 if prec == L.F64 else torch.float32
    if kind == L.C2C:
        x = torch.empty(batch, n, dtype=cdt, device="cuda")
    elif d == 0:
        x = torch.empty(batch, n, dtype=rdt, device="cuda")
    else:
        x = torch.empty(batch, n // 2, dtype=cdt, device="cuda")
    return fft_amd.fill(x, seed=seed)


def _out_like(n, kind, prec, batch, d):
    cdt = torch.complex128 if prec == L.F64 else torch.complex64
    rdt = torch.float64 if prec == L.F64 else torch.float32
    if kind == L.C2C:
        return torch.empty(batch, n, dtype=cdt, device="cuda")
    if d == 0:
        return torch.empty(batch, n // 2, dtype=cdt, device="cuda")
    return torch.empty(batch, n, dtype=rdt, device="cuda")


def _oracle(x, kind, d):
    if kind == L.C2C:
        return po.fft(x.astype(np.complex128), inverse=bool(d))
    if d == 0:
        return po.real_fft(x.astype(np.float64))
    return po.real_ifft(x.astype(np.complex128))


# (n, kind, precision, first batch, second batch, in place)
CASES = [
    (1009, L.C2C, L.F32, 64, 9, False),       # chirp-z rows in the scratch array
    (2 * 4099, L.R2C, L.F32, 16, 5, False),   # real kind, chirp-z half length
    (1 << 16, L.C2C, L.F32, 8, 3, True),      # two-pass in place: the intermediate
    (3 << 15, L.C2C, L.F64, 4, 1, True),
    (4096, L.C2C, L.F32, 32, 7, True),        # one pass: no workspace at all
]


@pytest.mark.parametrize("n,kind,prec,b1,b2,in_place", CASES)
def test_matches_execute_two_calls(n, kind, prec, b1, b2, in_place):
    plan = fft_amd.Plan(n, kind, prec)
    sh = torch.cuda.current_stream().cuda_stream
    tol = 1e-12 if prec == L.F64 else 1e-6
    for call, batch in enumerate((b1, b2)):
        for d in (0, 1):
            x = _input(n, kind, prec, batch, d, seed=100 * call + d + n)
            want = _out_like(n, kind, prec, batch, d)
            plan.execute_ptr(x.data_ptr(), want.data_ptr(), batch, d, sh)
            if in_place:
                got = x.clone()
                rc = _noalloc(plan, got, got, batch, d)
            else:
                got = _out_like(n, kind, prec, batch, d)
                rc = _noalloc(plan, x, got, batch, d)
            assert rc == L.SFFT_OK, L.lib().sfft_last_error()
            torch.cuda.synchronize()
            a = got.contiguous().view(torch.uint8).flatten()
            b = want.contiguous().view(torch.uint8).flatten()
            assert torch.equal(a, b), (n, kind, d, call, batch)
            err = po.rel_l2(got[0].cpu().numpy(), _oracle(x[0].cpu().numpy(), kind, d))
            assert err <= tol, (n, kind, d, err)


def test_oversized_workspace_refused():
    n, batch = 1 << 20, 16  # in place: 16 x 8 MiB of intermediate > the 64 MiB array
    plan = fft_amd.Plan(n, L.C2C, L.F32)
    assert plan.workspace_size(batch, True) > (64 << 20)
    x = torch.zeros(batch, n, dtype=torch.complex64, device="cuda")
    rc = _noalloc(plan, x, x, batch, 0)
    assert rc == L.SFFT_ERR_BUDGET
    assert b"scratch" in L.lib().sfft_last_error()
    y = torch.empty_like(x)  # out of place needs no workspace
    assert _noalloc(plan, x, y, batch, 0) == L.SFFT_OK
    torch.cuda.synchronize()
