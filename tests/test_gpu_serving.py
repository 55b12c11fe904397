"""Serving loop: continuous batching under an HBM budget.

The same request gives the same result alone (a plain plan execution on the
same synthetic input) and in the mix, bitwise; a budget smaller than the
offered load applies backpressure without failures; a request larger than the
budget is rejected with SFFT_ERR_BUDGET."""
import numpy as np
import pytest
import torch

import fft_amd
from fft_amd import _lib as L
from fft_amd.serving import Server, latency_summary
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu


def _alone(n, kind, direction, batch, seed):
    """The request served alone: synthesise the same input, run the plan."""
    plan = fft_amd.Plan(n, kind, L.F32)
    if kind == L.C2C:
        x = torch.empty(batch, n, dtype=torch.complex64, device="cuda")
        y = torch.empty_like(x)
    elif direction == L.FORWARD:
        x = torch.empty(batch, n, dtype=torch.float32, device="cuda")
        y = torch.empty(batch, n // 2, dtype=torch.complex64, device="cuda")
    else:
        x = torch.empty(batch, n // 2, dtype=torch.complex64, device="cuda")
        y = torch.empty(batch, n, dtype=torch.float32, device="cuda")
    fft_amd.fill(x, seed=seed)
    plan.execute_ptr(x.data_ptr(), y.data_ptr(), batch, direction,
                     torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    return x.cpu().numpy(), y.cpu().numpy()


def test_same_result_alone_and_in_the_mix():
    reqs = [(4096, L.C2C, 0, 7), (1920, L.C2C, 1, 3), (2048, L.R2C, 0, 16), (65536, L.C2C, 0, 2),
            (3000, L.C2C, 0, 5), (1024, L.R2C, 1, 9), (16384, L.C2C, 1, 4), (6000, L.R2C, 0, 3),
            (256, L.C2C, 0, 500), (48000, L.C2C, 1, 1)]
    with Server(budget_bytes=1 << 30, max_inflight=6, streams=4) as srv:
        for i, (n, k, d, b) in enumerate(reqs):
            srv.submit(i, n, k, d, b, seed=1000 + i, retain=True)
        srv.drain()
        rec = srv.poll()
        assert len(rec) == len(reqs) and (rec["status"] == 0).all()
        for i, (n, k, d, b) in enumerate(reqs):
            x, y = _alone(n, k, d, b, 1000 + i)
            got = srv.fetch(i, y.nbytes).view(y.dtype).reshape(y.shape)
            srv.release(i)
            assert np.array_equal(got, y), (n, k, d, b)
            # and the oracle on the first transform of the request
            if k == L.C2C:
                want = po.fft(x[0].astype(np.complex128), inverse=bool(d))
            elif d == 0:
                want = po.real_fft(x[0].astype(np.float64))
            else:
                want = po.real_ifft(x[0].astype(np.complex128))
            assert po.rel_l2(got[0], want) <= 1e-6, (n, k, d)


def test_same_result_at_the_largest_requests():
    """Requests at the top of the serving range (N up to 2^18, batches up to
    512, all three kinds) served together equal the plain plan bitwise; the
    first transform of each is checked against the oracle."""
    reqs = [(1 << 18, L.C2C, 0, 512), (1 << 18, L.R2C, 0, 300), (3 << 16, L.C2C, 1, 200),
            (1 << 18, L.R2C, 1, 64), (256, L.C2C, 1, 512), (3 << 12, L.C2C, 0, 512),
            (5 ** 7, L.C2C, 1, 100), (2 ** 17, L.R2C, 0, 512)]
    with Server(budget_bytes=16 << 30, max_inflight=8, streams=4) as srv:
        for i, (n, k, d, b) in enumerate(reqs):
            srv.submit(i, n, k, d, b, seed=2000 + i, retain=True)
        srv.drain()
        rec = srv.poll()
        assert len(rec) == len(reqs) and (rec["status"] == 0).all()
        for i, (n, k, d, b) in enumerate(reqs):
            x, y = _alone(n, k, d, b, 2000 + i)
            got = srv.fetch(i, y.nbytes).view(y.dtype).reshape(y.shape)
            srv.release(i)
            assert np.array_equal(got, y), (n, k, d, b)
            if k == L.C2C:
                want = po.fft(x[0].astype(np.complex128), inverse=bool(d))
            elif d == 0:
                want = po.real_fft(x[0].astype(np.float64))
            else:
                want = po.real_ifft(x[0].astype(np.complex128))
            assert po.rel_l2(got[0], want) <= 1e-6, (n, k, d)
            del x, y, got


def test_poisson_run_full_range():
    """Poisson arrivals over the whole serving range: sizes log-uniform in
    [256, 2^18], batches 1..512, a 64 GB budget."""
    with Server(budget_bytes=64 << 30, max_inflight=64, streams=4) as srv:
        rec = srv.run_poisson(rate=2000, duration=0.25, n_min=256, n_max=1 << 18,
                              batch_max=512, seed=5)
    s = latency_summary(rec)
    assert s["requests"] > 300 and s["failed"] == 0, s
    assert s["p50_ms"] > 0 and s["p99_ms"] >= s["p50_ms"]


def test_burst_of_small_requests_merged():
    """A burst of small requests (admitted together, so they run in merged
    persistent launches) of every kind and both directions, retained: each
    equals the plain plan on the same input, bitwise."""
    rng = np.random.default_rng(4)
    sizes = [256, 300, 512, 1000, 1024, 1920, 2048, 3000, 4096, 6000, 8192, 16384, 20000,
             32768, 65536]
    reqs = []
    for i in range(120):
        n = int(rng.choice(sizes))
        kind = [L.C2C, L.C2C, L.R2C][i % 3]
        d = int(rng.integers(0, 2))
        b = int(rng.integers(1, 9))
        reqs.append((n, kind, d, b))
    with Server(budget_bytes=4 << 30, max_inflight=64, streams=4) as srv:
        for i, (n, k, d, b) in enumerate(reqs):
            srv.submit(i, n, k, d, b, seed=5000 + i, retain=True)
        srv.drain()
        rec = srv.poll()
        assert len(rec) == len(reqs) and (rec["status"] == 0).all()
        for i, (n, k, d, b) in enumerate(reqs):
            x, y = _alone(n, k, d, b, 5000 + i)
            got = srv.fetch(i, y.nbytes).view(y.dtype).reshape(y.shape)
            srv.release(i)
            assert np.array_equal(got, y), (i, n, k, d, b)


def test_backpressure_under_small_budget():
    # 32 requests of 32 MiB each through a 96 MiB pool: at most a few fit at once
    with Server(budget_bytes=96 << 20, max_inflight=16) as srv:
        for i in range(32):
            srv.submit(i, 16384, L.C2C, i % 2, 256, seed=i)
        srv.drain()
        rec = srv.poll()
    assert len(rec) == 32 and (rec["status"] == 0).all()
    order = np.argsort(rec["t_admit"])
    assert list(rec["id"][order]) == list(range(32)), "admission is in arrival order"
    s = latency_summary(rec)
    assert s["p99_ms"] >= s["p50_ms"] > 0


def test_oversized_request_rejected():
    with Server(budget_bytes=16 << 20) as srv:
        srv.submit(0, 1 << 20, L.C2C, 0, 4, seed=0)  # 64 MiB of buffers
        srv.submit(1, 1024, L.C2C, 0, 4, seed=1)
        srv.drain()
        rec = srv.poll()
    st = {int(r["id"]): int(r["status"]) for r in rec}
    assert st[0] == L.SFFT_ERR_BUDGET and st[1] == 0


def test_poisson_run_small():
    with Server(budget_bytes=8 << 30) as srv:
        rec = srv.run_poisson(rate=2000, duration=0.25, n_min=256, n_max=1 << 16, batch_max=64)
    s = latency_summary(rec)
    assert s["requests"] > 300 and s["failed"] == 0, s
    assert s["p50_ms"] > 0 and s["p99_ms"] >= s["p50_ms"]


def test_double_precision_request_matches_plain_plan():
    """An F64 server: a retained request equals the same transform run alone."""
    n, batch, seed = 6000, 3, 11
    with Server(budget_bytes=1 << 30, precision=L.F64) as srv:
        srv.submit(1, n, L.C2C, L.INVERSE, batch, seed, retain=True)
        srv.drain()
        got = srv.fetch(1, batch * n * 16).view(np.complex128).reshape(batch, n)
        srv.release(1)
    plan = fft_amd.Plan(n, L.C2C, L.F64)
    x = torch.empty(batch, n, dtype=torch.complex128, device="cuda")
    y = torch.empty_like(x)
    fft_amd.fill(x, seed=seed)
    plan.execute_ptr(x.data_ptr(), y.data_ptr(), batch, L.INVERSE,
                     torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    assert np.array_equal(got, y.cpu().numpy())
    assert po.rel_l2(got, po.fft(x.cpu().numpy(), inverse=True)) <= 1e-12
