"""16-bit storage variant (IEEE half / bfloat16 in HBM, float arithmetic),
with radix-16 stages as float butterflies (default) or on MFMA with split
hi+lo operands (plan hint mfma = 1).

The reference has no such variant; it is checked against the FP64 oracle on
the same (16-bit) input values.  Besides a plain tolerance, the error is
compared with the error of rounding the exact FP64 result to the storage
format once: the transform may add little beyond that, which is what the
split-precision MFMA contraction buys (a single 16-bit product per stage
would add about 2^-11 per radix-16 stage and fail these bounds).
"""
import numpy as np
import pytest
import torch

import fft_amd
from fft_amd import _lib as L
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu

DT = {L.F16: torch.float16, L.BF16: torch.bfloat16}
NAME = {L.F16: "f16", L.BF16: "bf16"}
# plain rel L2 bounds against FP64 (storage rounding dominates)
TOL = {L.F16: 1.5e-3, L.BF16: 1.2e-2}


def _run(n, batch, prec, direction, seed=1, mfma=-1):
    g = torch.Generator(device="cpu").manual_seed(seed + n)
    x32 = torch.randn(batch, n, 2, generator=g)
    xs = x32.to(DT[prec]).cuda()
    ys = torch.empty_like(xs)
    p = fft_amd.Plan(n, L.C2C, prec) if mfma < 0 else fft_amd.Plan(n, L.C2C, prec, mfma=mfma)
    p.execute_ptr(xs.data_ptr(), ys.data_ptr(), batch, direction,
                  torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    xin = xs.float().cpu().numpy().astype(np.float64)
    xin = xin[..., 0] + 1j * xin[..., 1]
    y = ys.float().cpu().numpy().astype(np.float64)
    y = y[..., 0] + 1j * y[..., 1]
    return p, xin, y


def _round_to(z, prec):
    t = torch.from_numpy(np.stack([z.real, z.imag], -1).astype(np.float32)).to(DT[prec]).float()
    t = t.numpy().astype(np.float64)
    return t[..., 0] + 1j * t[..., 1]


@pytest.mark.parametrize("mfma", [1, -1], ids=["mfma", "butterflies"])
@pytest.mark.parametrize("prec", [L.F16, L.BF16], ids=["f16", "bf16"])
@pytest.mark.parametrize("n,batch", [(256, 64), (1024, 32), (4096, 16), (16384, 4), (8192, 8)])
@pytest.mark.parametrize("direction", [0, 1], ids=["fwd", "inv"])
def test_single_pass_pow2(prec, n, batch, direction, mfma):
    p, x, y = _run(n, batch, prec, direction, mfma=mfma)
    info = p.info()
    assert info["passes"] == 1
    # MFMA stages run in the 16-element instance (<= 512 threads per
    # transform): up to 8192 in one pass
    assert (info["mfma_stages"] >= 1) == (mfma == 1 and n <= 8192), info
    want = po.fft(x, inverse=bool(direction))
    err = po.rel_l2(y, want)
    ideal = po.rel_l2(_round_to(want, prec), want)
    print(NAME[prec], n, direction, "rel L2", err, "storage rounding alone", ideal, info["radices"])
    assert err <= TOL[prec]
    assert err <= 1.3 * ideal + 1e-6, (err, ideal)


@pytest.mark.parametrize("mfma", [1, -1], ids=["mfma", "butterflies"])
@pytest.mark.parametrize("prec", [L.F16, L.BF16], ids=["f16", "bf16"])
@pytest.mark.parametrize("n,batch", [(1 << 16, 2), (1 << 20, 1)])
def test_two_pass_pow2(prec, n, batch, mfma):
    p, x, y = _run(n, batch, prec, 0, seed=3, mfma=mfma)
    info = p.info()
    assert info["passes"] == 2 and (info["mfma_stages"] >= 1) == (mfma == 1), info
    want = po.fft(x)
    err = po.rel_l2(y, want)
    ideal = po.rel_l2(_round_to(want, prec), want)
    print(NAME[prec], n, "rel L2", err, "storage rounding alone", ideal)
    assert err <= TOL[prec]
    # one extra rounding of the intermediate between the passes
    assert err <= 2.0 * ideal + 1e-6, (err, ideal)


@pytest.mark.parametrize("prec", [L.F16, L.BF16], ids=["f16", "bf16"])
@pytest.mark.parametrize("n,batch", [(256, 64), (1024, 64), (2048, 32), (4096, 16), (16384, 4),
                                     (1 << 16, 8), (1 << 17, 4), (1 << 18, 4), (1 << 20, 2)])
@pytest.mark.parametrize("direction", [0, 1], ids=["fwd", "inv"])
def test_compiled_kernels_match_runtime_pass(prec, n, batch, direction):
    """16-bit plans run compiled kernels (single pass: powers of two 256-16384;
    two passes: column lengths 128-1024); where the radices match the
    runtime planner's (all these lengths) they must give the bits of the
    runtime-planned passes (plan hint variant -2), which the tests above pin
    to the FP64 oracle."""
    g = torch.Generator(device="cpu").manual_seed(n + direction)
    xs = torch.randn(batch, n, 2, generator=g).to(DT[prec]).cuda()
    y1, y2 = torch.empty_like(xs), torch.empty_like(xs)
    sh = torch.cuda.current_stream().cuda_stream
    fft_amd.Plan(n, L.C2C, prec).execute_ptr(xs.data_ptr(), y1.data_ptr(), batch, direction, sh)
    fft_amd.Plan(n, L.C2C, prec, variant=-2).execute_ptr(xs.data_ptr(), y2.data_ptr(), batch,
                                                        direction, sh)
    torch.cuda.synchronize()
    assert torch.equal(y1.view(torch.int16), y2.view(torch.int16))


@pytest.mark.parametrize("prec", [L.F16, L.BF16], ids=["f16", "bf16"])
@pytest.mark.parametrize("n,batch", [(512, 64), (8192, 8)])
@pytest.mark.parametrize("direction", [0, 1], ids=["fwd", "inv"])
def test_compiled_own_radices(prec, n, batch, direction):
    """Compiled 16-bit kernels whose radices differ from the runtime
    planner's (512: 16 8 4, 8192: 16 16 8 4): against the FP64 oracle, within
    the storage rounding bound of the single-pass tests."""
    p, x, y = _run(n, batch, prec, direction, seed=11)
    assert p.info()["passes"] == 1
    want = po.fft(x, inverse=bool(direction))
    err = po.rel_l2(y, want)
    ideal = po.rel_l2(_round_to(want, prec), want)
    assert err <= TOL[prec]
    assert err <= 1.3 * ideal + 1e-6, (err, ideal)


@pytest.mark.parametrize("prec", [L.F16, L.BF16], ids=["f16", "bf16"])
@pytest.mark.parametrize("n,batch", [(1920, 8), (3000, 8), (48, 16), (100000, 1)])
def test_mixed_sizes(prec, n, batch):
    _, x, y = _run(n, batch, prec, 1, seed=5)
    want = po.fft(x, inverse=True)
    err = po.rel_l2(y, want)
    print(NAME[prec], n, "rel L2", err)
    assert err <= TOL[prec]


@pytest.mark.parametrize("mfma", [1, -1], ids=["mfma", "butterflies"])
@pytest.mark.parametrize("prec", [L.F16, L.BF16], ids=["f16", "bf16"])
def test_2e24_against_oracle(prec, mfma):
    """N = 2^24 against the FP64 oracle on the same (16-bit) input values."""
    n = 1 << 24
    p, x, y = _run(n, 1, prec, 0, seed=7, mfma=mfma)
    want = po.fft(x)
    err = po.rel_l2(y, want)
    ideal = po.rel_l2(_round_to(want, prec), want)
    print(NAME[prec], "2^24 rel L2", err, "storage rounding alone", ideal)
    assert err <= TOL[prec]
    assert err <= 2.0 * ideal + 1e-6


def test_real_kinds_rejected():
    with pytest.raises(L.SfftError):
        fft_amd.Plan(2048, L.R2C, L.F16)
