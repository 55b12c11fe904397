"""Synthetic workload descriptions (sizes, directions, kinds) from seeds.

Pure host-side bookkeeping used by bench.py and the tests: which transforms
a workload holds, never how they are computed.
"""
from __future__ import annotations

import numpy as np

from . import synth


def fast_sizes(lo: int, hi: int):
    """All 2^a 3^b 5^c in [lo, hi], ascending."""
    out = []
    p2 = 1
    while p2 <= hi:
        p3 = p2
        while p3 <= hi:
            p5 = p3
            while p5 <= hi:
                if p5 >= lo:
                    out.append(p5)
                p5 *= 5
            p3 *= 3
        p2 *= 2
    return sorted(out)


def log_uniform_fast_sizes(seed: int, count: int, lo: int, hi: int) -> np.ndarray:
    """count sizes: log N ~ U[ln lo, ln hi], snapped to the nearest fast size
    (nearest in log)."""
    fs = np.array(fast_sizes(lo, hi), dtype=np.int64)
    u = synth.uniform01(seed, count)
    logn = np.log(lo) + u * (np.log(hi) - np.log(lo))
    idx = np.abs(np.log(fs)[None, :] - logn[:, None]).argmin(axis=1)
    return fs[idx]


def uneven_batch(seed: int = 3, count: int = 20000, lo: int = 64, hi: int = 65536):
    """The uneven-batch workload: sizes log-uniform over the fast sizes in
    [lo, hi], direction Bernoulli(1/2) (0 forward, 1 inverse)."""
    sizes = log_uniform_fast_sizes(seed, count, lo, hi)
    dirs = (synth.uniform01(seed ^ 0xD1CE, count) < 0.5).astype(np.int64)
    return sizes, dirs
