"""The autotuner on the GPU: every tuned plan gives oracle-level results, the
choice becomes the plan sfft_plan_create returns, and the JSON cache
round-trips on the same device."""
import numpy as np
import pytest
import torch

import fft_amd
from fft_amd import _lib as L
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu


@pytest.fixture(autouse=True)
def _clean():
    fft_amd.clear_tuning()
    yield
    fft_amd.clear_tuning()


def _check_c2c(n, batch=4, tol=1e-6):
    x = torch.randn(batch, n, dtype=torch.complex64, device="cuda")
    y = torch.empty_like(x)
    fft_amd.FFT(n).fft(x, y)
    torch.cuda.synchronize()
    want = po.fft(x.cpu().numpy().astype(np.complex128))
    err = po.rel_l2(y.cpu().numpy(), want)
    assert err <= tol, (n, err)
    return y


@pytest.mark.parametrize("n,batch", [(1920, 2048), (4096, 1024), (65536, 64), (3000, 1024)])
def test_tuned_c2c_plans(n, batch):
    r = fft_amd.autotune(n, L.C2C, L.F32, batch)
    print(n, r)
    assert r["candidates"] >= 3 and r["GB/s"] > 0
    h = r["hints"]
    info = fft_amd.Plan(n, L.C2C, L.F32).info()
    if h["radices"]:
        assert info["radices"] == h["radices"]
    if h["split"] > 1:
        assert info["passes"] == 2
    _check_c2c(n)


def test_tuned_real_plan():
    r = fft_amd.autotune(2048, L.R2C, L.F32, 4096)
    print(r)
    x = torch.randn(8, 2048, device="cuda")
    y = torch.empty(8, 1024, dtype=torch.complex64, device="cuda")
    fft_amd.RealFFT(2048).fft(x, y)
    torch.cuda.synchronize()
    want = po.real_fft(x.cpu().numpy().astype(np.float64))
    assert po.rel_l2(y.cpu().numpy(), want) <= 1e-6


def test_tuned_16bit_plan():
    r = fft_amd.autotune(4096, L.C2C, L.F16, 1024)
    print(r)
    assert r["candidates"] >= 2


def test_cache_round_trip(tmp_path):
    fft_amd.autotune(1920, L.C2C, L.F32, 1024)
    fft_amd.autotune(8192, L.C2C, L.F32, 512)
    y1 = _check_c2c(1920)
    path = str(tmp_path / "tune.json")
    fft_amd.save_tuning(path)
    fft_amd.clear_tuning()
    assert fft_amd.load_tuning(path) == 2
    info = fft_amd.Plan(1920, L.C2C, L.F32).info()
    assert info["n"] == 1920
    _check_c2c(8192)
    assert y1.shape == (4, 1920)
