"""Seeded synthetic inputs: a counter-based splitmix64 stream.

The same generator runs on the device (sfft_fill_* in the C ABI), so a
workload can be synthesised directly in HBM; this numpy version produces the
identical uniform values on the host (Box-Muller normals agree to a few ulp,
since host and device libm differ).

Element e of stream `seed` draws the 64-bit word
    w(seed, e) = mix(mix(seed) + (e + 1) * 0x9E3779B97F4A7C15)
with mix() the splitmix64 finaliser.  Complex element k uses words 2k (re)
and 2k+1 (im).
    uniform [0, 1):  (w >> 40) * 2^-24           (exact in fp32)
    uniform [-1, 1): 2 * u - 1                   (exact in fp32)
    normal:          Box-Muller on words 2k, 2k+1 with u1 in (0, 1]
"""
from __future__ import annotations

import numpy as np

GOLDEN = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def _mix(z: np.ndarray) -> np.ndarray:
    z = np.asarray(z, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = (z ^ (z >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
    return z ^ (z >> np.uint64(31))


def words(seed: int, count: int, offset: int = 0) -> np.ndarray:
    """count 64-bit words of stream `seed`, starting at element `offset`."""
    key = _mix(np.uint64(seed & 0xFFFFFFFFFFFFFFFF))
    e = np.arange(offset + 1, offset + count + 1, dtype=np.uint64)
    with np.errstate(over="ignore"):
        return _mix(key + e * GOLDEN)


def uniform01(seed: int, count: int, offset: int = 0) -> np.ndarray:
    w = words(seed, count, offset)
    return (w >> np.uint64(40)).astype(np.float64) * (2.0 ** -24)


def uniform_pm1(seed: int, count: int, offset: int = 0) -> np.ndarray:
    return 2.0 * uniform01(seed, count, offset) - 1.0


def complex_uniform(seed: int, shape, dtype=np.complex64) -> np.ndarray:
    """Complex values with re, im i.i.d. U[-1, 1)."""
    n = int(np.prod(shape))
    u = uniform_pm1(seed, 2 * n)
    z = u[0::2] + 1j * u[1::2]
    return z.astype(dtype).reshape(shape)


def real_uniform(seed: int, shape, dtype=np.float32) -> np.ndarray:
    n = int(np.prod(shape))
    return uniform_pm1(seed, n).astype(dtype).reshape(shape)


def complex_normal(seed: int, shape, dtype=np.complex64) -> np.ndarray:
    """Complex values with re, im i.i.d. N(0, 1) (Box-Muller)."""
    n = int(np.prod(shape))
    w = words(seed, 2 * n)
    u1 = ((w[0::2] >> np.uint64(11)).astype(np.float64) + 1.0) * (2.0 ** -53)
    u2 = (w[1::2] >> np.uint64(11)).astype(np.float64) * (2.0 ** -53)
    r = np.sqrt(-2.0 * np.log(u1))
    z = r * np.cos(2 * np.pi * u2) + 1j * (r * np.sin(2 * np.pi * u2))
    return z.astype(dtype).reshape(shape)


def sinusoid_frames(seed: int, frames: int, n: int, tones: int = 8,
                    noise: float = 0.01, dtype=np.float32) -> np.ndarray:
    """Real frames: a sum of `tones` sinusoids with frequency U(0, n/2) bins,
    phase U[0, 2 pi), amplitude 1, plus noise * N(0, 1)."""
    fr = uniform01(seed, frames * tones, 0).reshape(frames, tones) * (n / 2)
    ph = uniform01(seed ^ 0x5A5A5A5A, frames * tones, 0).reshape(frames, tones) * (2 * np.pi)
    t = np.arange(n, dtype=np.float64)
    x = np.zeros((frames, n), dtype=np.float64)
    for k in range(tones):
        x += np.cos(2 * np.pi * fr[:, k:k + 1] * t[None, :] / n + ph[:, k:k + 1])
    nz = complex_normal(seed ^ 0xA5A5A5A5, (frames * n + 1) // 2, np.complex128)
    nz = np.concatenate([nz.real, nz.imag])[: frames * n].reshape(frames, n)
    x += noise * nz
    return x.astype(dtype)
