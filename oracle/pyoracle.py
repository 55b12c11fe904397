"""ctypes bindings for the CPU oracle -- TEST INFRASTRUCTURE ONLY.

Only tests/, the smoke check in __graft_entry__.py and the cpu_baseline leg of
bench.py may import this module, and only as the checker or as the timed CPU
baseline.  The product path (fft_amd/) never imports it.

Two libraries:
  libsfft_oracle.so   the C restatement of the reference algorithm
                      (sfft_oracle.c; built by `make -C oracle`)
  _ref/libsfft_ref.so the reference header itself compiled with a C shim
                      (ref_driver.cpp; built by `make -C oracle ref` where the
                      reference tree is present; the prebuilt file travels)
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(_HERE, "libsfft_oracle.so")
REF_SO = os.path.join(_HERE, "_ref", "libsfft_ref.so")

_sz = ctypes.c_size_t
_dp = ctypes.POINTER(ctypes.c_double)
_fp = ctypes.POINTER(ctypes.c_float)

_oracle = None
_ref = None


def _load_oracle():
    global _oracle
    if _oracle is None:
        if not os.path.exists(ORACLE_SO):
            raise RuntimeError(f"oracle library missing: {ORACLE_SO} (run `make -C oracle`)")
        lib = ctypes.CDLL(ORACLE_SO)
        for name in ("oracle_size_minimum", "oracle_size_maximum",
                     "oracle_real_size_minimum", "oracle_real_size_maximum"):
            f = getattr(lib, name)
            f.argtypes = [_sz]
            f.restype = _sz
        lib.oracle_fft_c2c.argtypes = [_sz, _sz, _dp, _dp, ctypes.c_int]
        lib.oracle_fft_c2c.restype = ctypes.c_int
        lib.oracle_fft_real.argtypes = [_sz, _sz, _dp, _dp, ctypes.c_int, ctypes.c_int]
        lib.oracle_fft_real.restype = ctypes.c_int
        lib.oracle_dft_ld.argtypes = [_sz, _dp, _dp, ctypes.c_int]
        lib.oracle_dft_ld.restype = ctypes.c_int
        _oracle = lib
    return _oracle


def ref_available() -> bool:
    return os.path.exists(REF_SO)


def _load_ref():
    global _ref
    if _ref is None:
        if not ref_available():
            raise RuntimeError(f"reference shim missing: {REF_SO}")
        lib = ctypes.CDLL(REF_SO)
        lib.ref_fft_c2c_f64.argtypes = [_sz, _sz, _dp, _dp, ctypes.c_int]
        lib.ref_fft_c2c_f32.argtypes = [_sz, _sz, _fp, _fp, ctypes.c_int]
        lib.ref_fft_real_f64.argtypes = [_sz, _sz, _dp, _dp, ctypes.c_int, ctypes.c_int]
        lib.ref_fft_real_f32.argtypes = [_sz, _sz, _fp, _fp, ctypes.c_int, ctypes.c_int]
        for name in ("ref_size_minimum", "ref_size_maximum",
                     "ref_real_size_minimum", "ref_real_size_maximum"):
            f = getattr(lib, name)
            f.argtypes = [_sz]
            f.restype = _sz
        lib.ref_setter_return.argtypes = [ctypes.c_int, ctypes.c_int, _sz]
        lib.ref_setter_return.restype = _sz
        lib.ref_ctor_size.argtypes = [ctypes.c_int, _sz, ctypes.c_int]
        lib.ref_ctor_size.restype = _sz
        for name in ("ref_time_c2c_f32", "ref_time_real_f32"):
            f = getattr(lib, name)
            f.argtypes = [_sz, _sz, _fp, _fp, ctypes.c_int, ctypes.c_int]
            f.restype = ctypes.c_double
        lib.ref_time_c2c_f64.argtypes = [_sz, _sz, _dp, _dp, ctypes.c_int, ctypes.c_int]
        lib.ref_time_c2c_f64.restype = ctypes.c_double
        _ref = lib
    return _ref


def _as_c128(x) -> np.ndarray:
    return np.ascontiguousarray(x, dtype=np.complex128)


# ----------------------------------------------------------------- oracle --

def size_minimum(n: int) -> int:
    return _load_oracle().oracle_size_minimum(n)


def size_maximum(n: int) -> int:
    return _load_oracle().oracle_size_maximum(n)


def real_size_minimum(n: int) -> int:
    return _load_oracle().oracle_real_size_minimum(n)


def real_size_maximum(n: int) -> int:
    return _load_oracle().oracle_real_size_maximum(n)


def fft(x, inverse: bool = False) -> np.ndarray:
    """FP64 complex FFT of the last axis (unnormalised inverse)."""
    x = _as_c128(x)
    n = x.shape[-1]
    out = np.empty_like(x)
    batch = x.size // n if n else 0
    rc = _load_oracle().oracle_fft_c2c(n, batch, x.ctypes.data_as(_dp), out.ctypes.data_as(_dp),
                                       int(bool(inverse)))
    if rc:
        raise RuntimeError("oracle_fft_c2c failed")
    return out


def ifft(x) -> np.ndarray:
    return fft(x, inverse=True)


def real_fft(x, half_shift: bool = False) -> np.ndarray:
    """FP64 RealFFT forward: (..., n) real -> (..., n/2) complex, bin-0 packed."""
    x = np.ascontiguousarray(x, dtype=np.float64)
    n = x.shape[-1]
    out = np.empty(x.shape[:-1] + (n // 2,), dtype=np.complex128)
    batch = x.size // n
    rc = _load_oracle().oracle_fft_real(n, batch, x.ctypes.data_as(_dp), out.ctypes.data_as(_dp),
                                        0, int(bool(half_shift)))
    if rc:
        raise RuntimeError("oracle_fft_real failed")
    return out


def real_ifft(y, half_shift: bool = False) -> np.ndarray:
    """FP64 RealFFT inverse: (..., n/2) complex -> (..., n) real (unnormalised)."""
    y = _as_c128(y)
    h = y.shape[-1]
    n = 2 * h
    out = np.empty(y.shape[:-1] + (n,), dtype=np.float64)
    batch = y.size // h
    rc = _load_oracle().oracle_fft_real(n, batch, y.ctypes.data_as(_dp), out.ctypes.data_as(_dp),
                                        1, int(bool(half_shift)))
    if rc:
        raise RuntimeError("oracle_fft_real failed")
    return out


def dft_ld(x, sign: int = -1) -> np.ndarray:
    """Brute-force O(n^2) DFT evaluated in long double."""
    x = _as_c128(x)
    out = np.empty_like(x)
    _load_oracle().oracle_dft_ld(x.shape[-1], x.ctypes.data_as(_dp), out.ctypes.data_as(_dp), sign)
    return out


# -------------------------------------------------------------- reference --

def ref_fft(x, inverse: bool = False) -> np.ndarray:
    """The reference FFT<double> (or FFT<float> for complex64 input)."""
    lib = _load_ref()
    x = np.ascontiguousarray(x)
    n = x.shape[-1]
    batch = x.size // n
    if x.dtype == np.complex64:
        out = np.empty_like(x)
        lib.ref_fft_c2c_f32(n, batch, x.ctypes.data_as(_fp), out.ctypes.data_as(_fp), int(inverse))
    else:
        x = _as_c128(x)
        out = np.empty_like(x)
        lib.ref_fft_c2c_f64(n, batch, x.ctypes.data_as(_dp), out.ctypes.data_as(_dp), int(inverse))
    return out


def ref_real_fft(x, half_shift: bool = False) -> np.ndarray:
    lib = _load_ref()
    x = np.ascontiguousarray(x)
    n = x.shape[-1]
    batch = x.size // n
    if x.dtype == np.float32:
        out = np.empty(x.shape[:-1] + (n // 2,), dtype=np.complex64)
        lib.ref_fft_real_f32(n, batch, x.ctypes.data_as(_fp), out.ctypes.data_as(_fp), 0, int(half_shift))
    else:
        x = np.ascontiguousarray(x, dtype=np.float64)
        out = np.empty(x.shape[:-1] + (n // 2,), dtype=np.complex128)
        lib.ref_fft_real_f64(n, batch, x.ctypes.data_as(_dp), out.ctypes.data_as(_dp), 0, int(half_shift))
    return out


def ref_real_ifft(y, half_shift: bool = False) -> np.ndarray:
    lib = _load_ref()
    y = np.ascontiguousarray(y)
    h = y.shape[-1]
    batch = y.size // h
    if y.dtype == np.complex64:
        out = np.empty(y.shape[:-1] + (2 * h,), dtype=np.float32)
        lib.ref_fft_real_f32(2 * h, batch, y.ctypes.data_as(_fp), out.ctypes.data_as(_fp), 1, int(half_shift))
    else:
        y = _as_c128(y)
        out = np.empty(y.shape[:-1] + (2 * h,), dtype=np.float64)
        lib.ref_fft_real_f64(2 * h, batch, y.ctypes.data_as(_dp), out.ctypes.data_as(_dp), 1, int(half_shift))
    return out


def ref_sizes(fn: str, n: int) -> int:
    return getattr(_load_ref(), "ref_" + fn)(n)


def ref_setter_return(real: bool, which: int, n: int) -> int:
    return _load_ref().ref_setter_return(int(real), which, n)


def ref_ctor_size(real: bool, n: int, fast_direction: int) -> int:
    return _load_ref().ref_ctor_size(int(real), n, fast_direction)


def ref_time_c2c(x: np.ndarray, inverse: bool, threads: int) -> float:
    """Wall seconds for the reference to transform every row of x, one
    transform per thread (each thread owns an FFT<V> instance)."""
    lib = _load_ref()
    x = np.ascontiguousarray(x)
    n = x.shape[-1]
    batch = x.size // n
    out = np.empty_like(x)
    if x.dtype == np.complex64:
        return lib.ref_time_c2c_f32(n, batch, x.ctypes.data_as(_fp), out.ctypes.data_as(_fp),
                                    int(inverse), threads)
    return lib.ref_time_c2c_f64(n, batch, x.ctypes.data_as(_dp), out.ctypes.data_as(_dp),
                                int(inverse), threads)


def ref_time_real(x: np.ndarray, threads: int, inverse: bool = False) -> float:
    """Wall seconds for RealFFT<float> over every row: forward takes (..., n)
    reals, inverse takes (..., n/2) complex."""
    lib = _load_ref()
    if inverse:
        x = np.ascontiguousarray(x, dtype=np.complex64)
        n = 2 * x.shape[-1]
        batch = x.size // (n // 2)
        out = np.empty(x.shape[:-1] + (n,), dtype=np.float32)
    else:
        x = np.ascontiguousarray(x, dtype=np.float32)
        n = x.shape[-1]
        batch = x.size // n
        out = np.empty(x.shape[:-1] + (n // 2,), dtype=np.complex64)
    return lib.ref_time_real_f32(n, batch, x.ctypes.data_as(_fp), out.ctypes.data_as(_fp),
                                 int(inverse), threads)


# ---------------------------------------------------------------- metrics --

def rel_l2(a, b) -> float:
    """||a - b||_2 / ||b||_2 in float64 (b is the reference)."""
    a = np.asarray(a, dtype=np.complex128).ravel()
    b = np.asarray(b, dtype=np.complex128).ravel()
    den = np.linalg.norm(b)
    num = np.linalg.norm(a - b)
    return float(num / den) if den else float(num)


def close_enough_metric(a, b) -> float:
    """The reference's own test metric (tests/tests-common.h:17-33):
    sqrt(sum |a-b|^2 / sum(|a*a|^2 + |b*b|^2)); the reference passes at < 1e-6."""
    a = np.asarray(a, dtype=np.complex128).ravel()
    b = np.asarray(b, dtype=np.complex128).ravel()
    err = np.sum(np.abs(a - b) ** 2)
    energy = np.sum(np.abs(a * a) ** 2 + np.abs(b * b) ** 2)
    if energy == 0:
        return 0.0
    return float(np.sqrt(err / energy))
