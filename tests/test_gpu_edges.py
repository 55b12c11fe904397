"""Edge cases of the transform path: empty batches, the smallest and the
largest lengths, strided batches (padding untouched), argument errors,
empty uneven batches.  The largest lengths are checked through closed forms
evaluated on the device (a single bin transforms to N at that bin)."""
import ctypes

import numpy as np
import pytest
import torch

import fft_amd
from fft_amd import _lib as L
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu


def _sh():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def _strided(plan, x, in_dist, y, out_dist, batch, direction):
    L.check(L.lib().sfft_execute_strided(plan._h, ctypes.c_void_p(x.data_ptr()), in_dist,
                                         ctypes.c_void_p(y.data_ptr()), out_dist, batch,
                                         direction, _sh()))


def test_zero_batch_is_a_no_op():
    p = fft_amd.Plan(64, L.C2C, L.F32)
    x = torch.ones(64, dtype=torch.complex64, device="cuda")
    y = torch.full_like(x, 7)
    p.execute_ptr(x.data_ptr(), y.data_ptr(), 0, 0, torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    assert torch.all(y == 7)


@pytest.mark.parametrize("n", [1, 2, 3])
def test_smallest_lengths(n):
    x = (np.arange(n) + 1 + 1j * np.arange(n)).astype(np.complex128)
    for inv in (False, True):
        f = fft_amd.FFT(n, dtype=torch.complex128)
        xd = torch.from_numpy(x).cuda()
        yd = torch.empty_like(xd)
        (f.ifft if inv else f.fft)(xd, yd)
        torch.cuda.synchronize()
        assert po.rel_l2(yd.cpu().numpy(), po.dft_ld(x, 1 if inv else -1)) < 1e-14


def test_smallest_real_length():
    rf = fft_amd.RealFFT(2, dtype=torch.float64)
    x = torch.tensor([[3.0, 5.0]], dtype=torch.float64, device="cuda")
    y = torch.empty(1, 1, dtype=torch.complex128, device="cuda")
    rf.fft(x, y)
    torch.cuda.synchronize()
    assert complex(y[0, 0]) == complex(8.0, -2.0)  # (DC, Nyquist) packed in bin 0


@pytest.mark.parametrize("n,prec", [(4096, L.F32), (1920, L.F32), (777, L.F32), (40000, L.F32),
                                    (17, L.F32), (4096, L.F64), (8192, L.F64)])
def test_strided_batches_leave_padding_untouched(n, prec):
    """sfft_execute_strided with distances larger than the length (fixed-plan,
    runtime-planned, two-pass and chirp-z lengths)."""
    batch, din, dout = 5, n + 37, n + 11
    cdt = np.complex64 if prec == L.F32 else np.complex128
    rng = np.random.default_rng(n)
    xa = (rng.standard_normal((batch, din)) + 1j * rng.standard_normal((batch, din))).astype(cdt)
    x = torch.from_numpy(xa).cuda()
    y = torch.full((batch, dout), complex(9.0, -9.0), dtype=x.dtype, device="cuda")
    p = fft_amd.Plan(n, L.C2C, prec)
    tol = 1e-6 if prec == L.F32 else 1e-13
    for direction in (0, 1):
        _strided(p, x, din, y, dout, batch, direction)
        torch.cuda.synchronize()
        got = y.cpu().numpy()
        want = po.fft(xa[:, :n].astype(np.complex128), inverse=bool(direction))
        assert po.rel_l2(got[:, :n], want) <= tol, (n, direction)
        assert np.all(got[:, n:] == cdt(9 - 9j))


@pytest.mark.parametrize("prec", [L.F16, L.BF16], ids=["f16", "bf16"])
@pytest.mark.parametrize("n", [1 << 16, 1 << 18])
def test_strided_16bit_two_pass(prec, n):
    """16-bit two-pass plans (compiled column passes) with batch distances
    larger than the length: the transforms equal the contiguous ones bit for
    bit, and the padding between them is untouched."""
    dt = torch.float16 if prec == L.F16 else torch.bfloat16
    batch, din, dout = 3, n + 41, n + 12
    g = torch.Generator(device="cpu").manual_seed(n + prec)
    x = torch.randn(batch, din, 2, generator=g).to(dt).cuda()
    p = fft_amd.Plan(n, L.C2C, prec)
    assert p.info()["passes"] == 2
    for direction in (0, 1):
        y = torch.full((batch, dout, 2), 7.0, dtype=dt, device="cuda")
        _strided(p, x, din, y, dout, batch, direction)
        xc = x[:, :n].contiguous()
        yc = torch.empty_like(xc)
        p.execute_ptr(xc.data_ptr(), yc.data_ptr(), batch, direction,
                      torch.cuda.current_stream().cuda_stream)
        torch.cuda.synchronize()
        assert torch.equal(y[:, :n].view(torch.int16), yc.view(torch.int16)), direction
        assert torch.all(y[:, n:] == 7.0)


def test_strided_real_batches():
    n, batch = 2048, 3
    din, dout = n + 6, n // 2 + 5  # reals in, complex out
    xr = torch.randn(batch, din, dtype=torch.float32, device="cuda")
    y = torch.full((batch, dout), complex(9.0, 9.0), dtype=torch.complex64, device="cuda")
    p = fft_amd.Plan(n, L.R2C, L.F32)
    _strided(p, xr, din, y, dout, batch, 0)
    torch.cuda.synchronize()
    want = po.real_fft(xr[:, :n].double().cpu().numpy())
    got = y.cpu().numpy()
    assert po.rel_l2(got[:, :n // 2], want) <= 1e-6
    assert np.all(got[:, n // 2:] == np.complex64(9 + 9j))


def test_argument_errors():
    p = fft_amd.Plan(64, L.C2C, L.F32)
    x = torch.zeros(4, 64, dtype=torch.complex64, device="cuda")
    y = torch.zeros_like(x)
    lib = L.lib()
    rc = lib.sfft_execute(p._h, None, ctypes.c_void_p(y.data_ptr()), 1, 0, None)
    assert rc == L.SFFT_ERR_INVALID_ARG
    rc = lib.sfft_execute(p._h, ctypes.c_void_p(x.data_ptr()), ctypes.c_void_p(y.data_ptr()), 1, 7,
                          None)
    assert rc == L.SFFT_ERR_INVALID_ARG
    rc = lib.sfft_execute_strided(p._h, ctypes.c_void_p(x.data_ptr()), 32,
                                  ctypes.c_void_p(y.data_ptr()), 64, 2, 0, None)
    assert rc == L.SFFT_ERR_INVALID_ARG  # distance shorter than a transform
    # overlapping strided ranges: output starts inside the input's last transform
    rc = lib.sfft_execute_strided(p._h, ctypes.c_void_p(x.data_ptr()), 64,
                                  ctypes.c_void_p(x.data_ptr() + 8 * 100), 64, 2, 0, None)
    assert rc == L.SFFT_ERR_ALIASING
    assert lib.sfft_last_error()
    with pytest.raises(L.SfftError):
        fft_amd.Plan(0, L.C2C, L.F32)
    with pytest.raises(L.SfftError):
        fft_amd.Plan(63, L.R2C, L.F32)  # odd real length
    with pytest.raises(L.SfftError):
        fft_amd.Plan(64, 9, L.F32)
    with pytest.raises(L.SfftError):
        fft_amd.Plan(64, L.C2C, 9)
    with pytest.raises(L.SfftError) as e:  # F64 two-pass columns fit LDS up to 2^26
        fft_amd.Plan(1 << 27, L.C2C, L.F64)
    assert e.value.code == L.SFFT_ERR_UNSUPPORTED_SIZE


def test_empty_uneven_batch():
    b = fft_amd.UnevenBatch([], [])
    xin = torch.zeros(16, dtype=torch.uint8, device="cuda")
    out = torch.zeros(16, dtype=torch.uint8, device="cuda")
    b.execute(xin, out)
    b.execute_per_size(xin, out)
    torch.cuda.synchronize()


def _single_bin(n, b, dtype, cdtype):
    k = torch.arange(n, device="cuda", dtype=torch.int64)
    ph = ((k * b) % n).to(torch.float64) * (2 * np.pi / n)
    return torch.polar(torch.ones_like(ph), ph).to(cdtype)


@pytest.mark.parametrize("n,prec", [(1 << 28, L.F32), ((1 << 25) - 1, L.F64)],
                         ids=["2^28_f32_two_pass", "2^25-1_f64_chirp_z"])
def test_largest_lengths_single_bin(n, prec):
    """e^{2 pi i b k / N} -> N delta_b; inverse of N delta_b -> N e^{...}."""
    cdt = torch.complex64 if prec == L.F32 else torch.complex128
    b = 12345
    x = _single_bin(n, b, None, cdt)
    y = torch.empty_like(x)
    p = fft_amd.Plan(n, L.C2C, prec)
    p.execute_ptr(x.data_ptr(), y.data_ptr(), 1, 0, torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    peak = complex(y[b])
    y[b] = 0
    rest = float(torch.linalg.vector_norm(y.to(torch.complex128)))
    # rel L2 against the exact N delta_b (whose norm is N)
    rel = np.hypot(abs(peak - n), rest) / n
    print(n, "rel L2", rel)
    assert rel < (1e-6 if prec == L.F32 else 1e-12), rel
    del x, y
    torch.cuda.empty_cache()


@pytest.mark.parametrize("n,kind,prec", [(4096, L.C2C, L.F32), (777, L.C2C, L.F32),
                                         (17, L.C2C, L.F32), (2048, L.R2C, L.F32),
                                         (6000, L.R2C_HALFSHIFT, L.F32), (34, L.R2C, L.F32),
                                         (4096, L.C2C, L.F64), (1024, L.C2C, L.F64)])
def test_in_place_one_pass_and_chirp_z(n, kind, prec):
    """in == out with the same layout: results equal the out-of-place ones."""
    p = fft_amd.Plan(n, kind, prec)
    assert p.info()["in_place"] == 1
    batch = 6
    cdt = torch.complex64 if prec == L.F32 else torch.complex128
    rdt = torch.float32 if prec == L.F32 else torch.float64
    sh = torch.cuda.current_stream().cuda_stream
    for d in (0, 1):
        if kind == L.C2C or d == 1:
            x = torch.randn(batch, n if kind == L.C2C else n // 2, dtype=cdt, device="cuda")
        else:
            x = torch.randn(batch, n, dtype=rdt, device="cuda")
        nbytes = x.numel() * x.element_size()
        want = torch.empty(nbytes, dtype=torch.uint8, device="cuda")
        p.execute_ptr(x.data_ptr(), want.data_ptr(), batch, d, sh)
        buf = x.clone()
        p.execute_ptr(buf.data_ptr(), buf.data_ptr(), batch, d, sh)
        torch.cuda.synchronize()
        assert torch.equal(buf.view(torch.uint8).flatten(), want), (n, kind, d)


@pytest.mark.parametrize("n,kind,prec,batch", [
    (1 << 16, L.C2C, L.F32, 3), (1 << 20, L.C2C, L.F32, 2), (3 << 15, L.C2C, L.F64, 2),
    (1 << 17, L.R2C, L.F32, 2), (3 << 16, L.R2C_HALFSHIFT, L.F32, 2), (1 << 16, L.C2C, L.F32X, 2)])
def test_in_place_two_pass(n, kind, prec, batch):
    """Two-pass lengths in place (in == out) through the workspace: bitwise
    equal to out of place, with the library's workspace and with the caller's."""
    p = fft_amd.Plan(n, kind, prec)
    info = p.info()
    assert info["passes"] == 2 and info["in_place"] == 1
    cb = 16 if prec == L.F64 else 8
    assert info["workspace_bytes"] == (n if kind == L.C2C else n // 2) * cb
    assert p.workspace_size(batch, True) == batch * info["workspace_bytes"]
    assert p.workspace_size(batch, False) == 0
    cdt = torch.complex128 if prec == L.F64 else torch.complex64
    rdt = torch.float64 if prec == L.F64 else torch.float32
    sh = torch.cuda.current_stream().cuda_stream
    for d in (0, 1):
        if kind == L.C2C or d == 1:
            x = torch.randn(batch, n if kind == L.C2C else n // 2, dtype=cdt, device="cuda")
        else:
            x = torch.randn(batch, n, dtype=rdt, device="cuda")
        nbytes = x.numel() * x.element_size()
        want = torch.empty(nbytes, dtype=torch.uint8, device="cuda")
        p.execute_ptr(x.data_ptr(), want.data_ptr(), batch, d, sh)
        buf = x.clone()
        p.execute_ptr(buf.data_ptr(), buf.data_ptr(), batch, d, sh)
        buf2 = x.clone()
        ws = torch.empty(p.workspace_size(batch, True), dtype=torch.uint8, device="cuda")
        p.execute_workspace_ptr(buf2.data_ptr(), buf2.data_ptr(), batch, d, ws.data_ptr(),
                                ws.numel(), sh)
        torch.cuda.synchronize()
        assert torch.equal(buf.view(torch.uint8).flatten(), want), (n, kind, d)
        assert torch.equal(buf2.view(torch.uint8).flatten(), want), (n, kind, d)


def test_workspace_errors():
    p = fft_amd.Plan(1 << 16, L.C2C, L.F32)
    x = torch.zeros(2, 1 << 16, dtype=torch.complex64, device="cuda")
    lib = L.lib()
    ws = torch.empty(p.workspace_size(2, True), dtype=torch.uint8, device="cuda")
    xp = ctypes.c_void_p(x.data_ptr())
    # in place needs the workspace; one byte short is refused
    rc = lib.sfft_execute_workspace(p._h, xp, xp, 2, 0, None, 0, None)
    assert rc == L.SFFT_ERR_INVALID_ARG
    rc = lib.sfft_execute_workspace(p._h, xp, xp, 2, 0, ctypes.c_void_p(ws.data_ptr()),
                                    ws.numel() - 1, None)
    assert rc == L.SFFT_ERR_INVALID_ARG
    # a workspace inside the data is refused
    rc = lib.sfft_execute_workspace(p._h, xp, xp, 1, 0, ctypes.c_void_p(x.data_ptr()),
                                    ws.numel(), None)
    assert rc == L.SFFT_ERR_ALIASING
    # out of place needs none
    y = torch.empty_like(x)
    rc = lib.sfft_execute_workspace(p._h, xp, ctypes.c_void_p(y.data_ptr()), 2, 0, None, 0, None)
    assert rc == L.SFFT_OK
    torch.cuda.synchronize()
