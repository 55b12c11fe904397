"""Device synthesis of seeded inputs (sfft_fill): the counter-based splitmix64
stream of fft_amd.synth, generated in HBM.  Uniform values are exact in FP32
and FP64, so the device output equals synth.uniform_pm1 bitwise, at every
alignment of the buffer (the vector body, the head before the first 16-byte
boundary and the tail), and past one pass of the launch grid."""
import numpy as np
import pytest
import torch

import fft_amd
from fft_amd import synth

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
@pytest.mark.parametrize("count", [1, 2, 3, 5, 7, 16, 17, 1000003])
@pytest.mark.parametrize("offset", [0, 1, 2, 3])
def test_uniform_matches_host_stream(dtype, count, offset):
    base = torch.full((count + 8,), 7.0, dtype=dtype, device="cuda")
    fft_amd.fill(base[offset:offset + count], seed=1234 + count)
    got = base.cpu().numpy()
    want = synth.uniform_pm1(1234 + count, count).astype(got.dtype)
    assert np.array_equal(got[offset:offset + count], want)
    # nothing outside the buffer is written
    assert np.all(got[:offset] == 7.0) and np.all(got[offset + count:] == 7.0)


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
def test_uniform_beyond_one_grid_pass(dtype):
    """More values than one pass of the largest grid writes (16384 workgroups
    of 256 threads, 16 bytes each), so every thread loops."""
    count = (16384 * 256 * 16) // (4 if dtype == torch.float32 else 8) * 2 + 5
    x = torch.empty(count, dtype=dtype, device="cuda")
    fft_amd.fill(x, seed=99)
    got = x.cpu().numpy()
    for lo in (0, count // 3, count - 4099):
        want = synth.uniform_pm1(99, 4099, offset=lo).astype(got.dtype)
        assert np.array_equal(got[lo:lo + 4099], want), lo


def test_complex_uniform_layout():
    """A complex buffer takes words 2k (re) and 2k + 1 (im)."""
    x = torch.empty(3, 1001, dtype=torch.complex64, device="cuda")
    fft_amd.fill(x, seed=5)
    assert np.array_equal(x.cpu().numpy(), synth.complex_uniform(5, (3, 1001)))


@pytest.mark.parametrize("dtype,tol", [(torch.complex64, 1e-6), (torch.complex128, 1e-12)])
def test_normal_matches_host_stream(dtype, tol):
    """Box-Muller values: host and device libm differ by a few ulp."""
    x = torch.empty(4097, dtype=dtype, device="cuda")
    fft_amd.fill(x, seed=21, dist="normal")
    want = synth.complex_normal(21, (4097,), np.complex128)
    got = x.cpu().numpy().astype(np.complex128)
    assert np.max(np.abs(got - want)) <= tol * max(1.0, np.max(np.abs(want)))
