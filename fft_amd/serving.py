"""Serving loop (sfft_server_*): continuous batching of transform requests.

    srv = Server(budget_bytes=64 << 30)
    srv.submit(id, n, kind, direction, batch, seed, retain=False)
    srv.drain(); srv.poll() -> completion records
    Server.run_poisson(...) -> records of a Poisson arrival run

Latency of a request = t_done - t_arrive on the server clock; the records
also carry the admission time and the request's algorithmic bytes.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib as L

RECORD_DTYPE = np.dtype([("id", np.uint64), ("status", np.int32), ("t_arrive", np.float64),
                         ("t_admit", np.float64), ("t_done", np.float64), ("bytes", np.uint64)])


def _records(buf, n):
    out = np.empty(n, dtype=RECORD_DTYPE)
    for i in range(n):
        c = buf[i]
        out[i] = (c.id, c.status, c.t_arrive, c.t_admit, c.t_done, c.bytes)
    return out


class Server:
    def __init__(self, budget_bytes: int = 64 << 30, max_inflight: int = 64, streams: int = 4,
                 precision: int = L.F32, plain: bool = False):
        cfg = L.ServerConfig(budget_bytes, max_inflight, streams, precision, int(plain))
        h = ctypes.c_void_p()
        L.check(L.lib().sfft_server_create(ctypes.byref(h), ctypes.byref(cfg)))
        self._h = h

    def close(self):
        if getattr(self, "_h", None) is not None and L._lib is not None:
            L._lib.sfft_server_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def submit(self, id: int, n: int, kind: int, direction: int, batch: int, seed: int,
               retain: bool = False):
        r = L.Request(id, n, kind, direction, batch, seed, int(retain))
        L.check(L.lib().sfft_server_submit(self._h, ctypes.byref(r)))

    def drain(self):
        L.check(L.lib().sfft_server_drain(self._h))

    def poll(self, max_records: int = 1 << 16) -> np.ndarray:
        buf = (L.Completion * max_records)()
        cnt = ctypes.c_size_t()
        L.check(L.lib().sfft_server_poll(self._h, buf, max_records, ctypes.byref(cnt)))
        return _records(buf, cnt.value)

    def fetch(self, id: int, nbytes: int) -> np.ndarray:
        out = np.empty(nbytes, dtype=np.uint8)
        L.check(L.lib().sfft_server_fetch(self._h, id, out.ctypes.data_as(ctypes.c_void_p), nbytes))
        return out

    def release(self, id: int):
        L.check(L.lib().sfft_server_release(self._h, id))

    def now(self) -> float:
        return L.lib().sfft_server_now(self._h)

    def run_poisson(self, rate: float, duration: float, n_min: int = 256, n_max: int = 1 << 18,
                    batch_max: int = 512, seed: int = 5, max_records: int = 1 << 20) -> np.ndarray:
        cfg = L.PoissonConfig(rate, duration, n_min, n_max, batch_max, seed)
        buf = (L.Completion * max_records)()
        cnt = ctypes.c_size_t()
        L.check(L.lib().sfft_server_run_poisson(self._h, ctypes.byref(cfg), buf, max_records,
                                                ctypes.byref(cnt)))
        return _records(buf, min(cnt.value, max_records))


def latency_summary(rec: np.ndarray) -> dict:
    ok = rec[rec["status"] == 0]
    lat = (ok["t_done"] - ok["t_arrive"]) * 1e3
    span = float(ok["t_done"].max() - ok["t_arrive"].min()) if len(ok) else 0.0
    return {
        "requests": int(len(rec)), "failed": int((rec["status"] != 0).sum()),
        "p50_ms": float(np.percentile(lat, 50)) if len(lat) else None,
        "p99_ms": float(np.percentile(lat, 99)) if len(lat) else None,
        "throughput_req_s": float(len(ok) / span) if span else None,
        "throughput_GBs": float(ok["bytes"].sum() / span / 1e9) if span else None,
    }
