"""Batches spanning more than 4 GiB of HBM.

The compiled kernels address HBM with 32-bit byte offsets from their launch
base, so the engine splits such batches into launches of less than 4 GiB
each.  Every transform of the batch must equal the same transform run alone,
bitwise, on both sides of each split (the reference transforms one buffer
per call: signalsmith-fft.h:374-386, :446-502).
"""
import pytest
import torch

import fft_amd
from fft_amd import _lib as L

pytestmark = pytest.mark.gpu

GIB4 = 1 << 32


def _alone(run, x, i, out_shape, dtype):
    yi = torch.empty((1,) + out_shape, dtype=dtype, device="cuda")
    run(x[i:i + 1], yi)
    torch.cuda.synchronize()
    return yi[0]


def _check_rows(run, x, y, per_in, per_out):
    """Rows on both sides of every 4 GiB launch boundary, and the last."""
    batch = x.shape[0]
    chunk = (GIB4 - 1) // max(per_in, per_out)
    rows = {0, batch - 1}
    for c0 in range(chunk, batch, chunk):
        rows |= {c0 - 1, c0}
    for i in sorted(rows):
        yi = _alone(run, x, i, tuple(y.shape[1:]), y.dtype)
        assert torch.equal(yi, y[i]), i
    return sorted(rows)


@pytest.mark.parametrize("n,batch,dtype", [(16384, 40000, torch.complex64),
                                           (1 << 20, 600, torch.complex64),
                                           (4096, 70000, torch.complex128),
                                           (1 << 18, 1100, torch.complex128)],
                         ids=["single-pass-16384", "two-pass-2^20", "f64-single-pass-4096",
                              "f64-two-pass-2^18"])
def test_c2c_batch_over_4gib(n, batch, dtype):
    f = fft_amd.FFT(n, dtype=dtype)
    eb = 16 if dtype == torch.complex128 else 8
    x = torch.empty(batch, n, dtype=dtype, device="cuda")
    assert x.numel() * eb > GIB4
    fft_amd.fill(x, seed=n, dist="normal")
    y = torch.empty_like(x)
    f.fft(x, y)
    torch.cuda.synchronize()
    rows = _check_rows(f.fft, x, y, eb * n, eb * n)
    print("C2C", n, batch, "passes", f.plan.info()["passes"], "rows checked", rows)
    del x, y
    torch.cuda.empty_cache()


def test_real_batch_over_4gib():
    """RealFFT of 65536 reals (a two-pass half length) over 17000 frames:
    4.5 GB in and out, forward then inverse."""
    n, batch = 1 << 16, 17000
    f = fft_amd.RealFFT(n, dtype=torch.float32)
    assert f.plan.info()["passes"] == 2
    x = torch.empty(batch, n, dtype=torch.float32, device="cuda")
    assert x.numel() * 4 > GIB4
    fft_amd.fill(x, seed=5, dist="normal")
    y = torch.empty(batch, n // 2, dtype=torch.complex64, device="cuda")
    f.fft(x, y)
    torch.cuda.synchronize()
    rows = _check_rows(f.fft, x, y, 4 * n, 4 * n)
    del x
    torch.cuda.empty_cache()
    z = torch.empty(batch, n, dtype=torch.float32, device="cuda")
    f.ifft(y, z)
    torch.cuda.synchronize()
    _check_rows(f.ifft, y, z, 4 * n, 4 * n)
    print("R2C", n, batch, "rows checked", rows)
    del y, z
    torch.cuda.empty_cache()
