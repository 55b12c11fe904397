"""Python mirror of the reference interface, backed by the HIP engine.

    signalsmith::FFT<V>             -> FFT(size, fastDirection=0, dtype=torch.complex64)
    signalsmith::RealFFT<V>         -> RealFFT(size, fastDirection=0, dtype=torch.float32)
    signalsmith::ModifiedRealFFT<V> -> ModifiedRealFFT(...)

Same method names, argument meaning and return values as signalsmith-fft.h
(setSize / setSizeMinimum / setSizeMaximum / size / sizeMinimum / sizeMaximum /
fft / ifft), including the reference's quirks:
  * the inverse is unnormalised (ifft(fft(x)) = N x)      signalsmith-fft.h:295-315
  * RealFFT.setSize returns the HALF size                  signalsmith-fft.h:434
  * RealFFT.sizeMaximum may exceed its argument            signalsmith-fft.h:406-408
  * RealFFT output holds N/2 bins, bin 0 = (DC, Nyquist)   signalsmith-fft.h:459-462
Differences, by design: inputs and outputs are torch tensors on the GPU (a
leading batch shape runs as one batched launch) or host numpy arrays (staged
through device memory); aliasing input and output raises instead of silently
corrupting.  Every transform runs through libsfft.so; there is no CPU path.
"""
from __future__ import annotations

import ctypes

import numpy as np
import torch

from . import _lib as L

_CPLX = {torch.complex64: L.F32, torch.complex128: L.F64}
_REAL = {torch.float32: L.F32, torch.float64: L.F64}


def _stream_handle(stream=None) -> int:
    s = stream if stream is not None else torch.cuda.current_stream()
    return s.cuda_stream


class Plan:
    """An engine plan: (n, kind, precision) -> device twiddles + launch layout."""

    def __init__(self, n: int, kind: int = L.C2C, precision: int = L.F32, variant=None,
                 radices=None, transforms_per_block: int = 0, split: int = 0,
                 elements_per_thread: int = 0, mfma: int = -1):
        """variant / radices / transforms_per_block / split / elements_per_thread /
        mfma override the planner (sfft_plan_create_hinted); without them the
        cached (possibly autotuned) plan is used."""
        h = ctypes.c_void_p()
        if (variant is None and radices is None and not transforms_per_block and not split
                and not elements_per_thread and mfma == -1):
            L.check(L.lib().sfft_plan_create(ctypes.byref(h), n, kind, precision))
        else:
            hints = L.PlanHints()
            hints.variant = -1 if variant is None else int(variant)
            rs = list(radices or [])
            hints.nradices = len(rs)
            for i, r in enumerate(rs):
                hints.radices[i] = int(r)
            hints.transforms_per_block = int(transforms_per_block)
            hints.split = int(split)
            hints.elements_per_thread = int(elements_per_thread)
            hints.mfma = int(mfma)
            L.check(L.lib().sfft_plan_create_hinted(ctypes.byref(h), n, kind, precision,
                                                    ctypes.byref(hints)))
        self._h = h
        self.n, self.kind, self.precision = n, kind, precision

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and L._lib is not None:
            L._lib.sfft_plan_destroy(h)
            self._h = None

    @property
    def handle(self):
        return self._h

    def info(self) -> dict:
        inf = L.PlanInfo()
        L.check(L.lib().sfft_plan_get_info(self._h, ctypes.byref(inf)))
        return {
            "n": inf.n, "kind": inf.kind, "precision": inf.precision, "passes": inf.passes,
            "radices": [inf.radices[i] for i in range(inf.nradices)],
            "threads_per_transform": inf.threads_per_transform,
            "transforms_per_block": inf.transforms_per_block,
            "workspace_bytes": inf.workspace_bytes, "table_bytes": inf.table_bytes,
            "mfma_stages": inf.mfma_stages, "chirp_length": inf.chirp_length,
            "in_place": inf.in_place,
        }

    def execute_ptr(self, in_ptr: int, out_ptr: int, batch: int, direction: int,
                    stream_handle: int) -> None:
        L.check(L.lib().sfft_execute(self._h, ctypes.c_void_p(in_ptr), ctypes.c_void_p(out_ptr),
                                     batch, direction, ctypes.c_void_p(stream_handle)))

    def workspace_size(self, batch: int, in_place: bool) -> int:
        """Device workspace bytes an execute call needs (sfft_plan_workspace_size)."""
        b = ctypes.c_size_t(0)
        L.check(L.lib().sfft_plan_workspace_size(self._h, batch, int(in_place), ctypes.byref(b)))
        return b.value

    def execute_workspace_ptr(self, in_ptr: int, out_ptr: int, batch: int, direction: int,
                              ws_ptr: int, ws_bytes: int, stream_handle: int) -> None:
        """sfft_execute_workspace: the caller's device workspace, no allocation."""
        L.check(L.lib().sfft_execute_workspace(
            self._h, ctypes.c_void_p(in_ptr), ctypes.c_void_p(out_ptr), batch, direction,
            ctypes.c_void_p(ws_ptr), ws_bytes, ctypes.c_void_p(stream_handle)))


def _check_device_pair(inp: torch.Tensor, out: torch.Tensor):
    if not (inp.is_cuda and out.is_cuda):
        raise ValueError("device tensors expected")
    if not inp.is_contiguous() or not out.is_contiguous():
        raise ValueError("contiguous tensors expected")


class FFT:
    """signalsmith::FFT<V> (signalsmith-fft.h:69-387) on the GPU."""

    def __init__(self, size: int, fastDirection: int = 0, dtype=torch.complex64):
        if dtype not in _CPLX:
            raise TypeError("dtype must be torch.complex64 or torch.complex128")
        self.dtype = dtype
        self._prec = _CPLX[dtype]
        self._size = 0
        self._plan = None
        if fastDirection > 0:
            size = FFT.sizeMinimum(size)
        if fastDirection < 0:
            size = FFT.sizeMaximum(size)
        self.setSize(size)

    @staticmethod
    def sizeMinimum(size: int) -> int:
        return L.lib().sfft_size_minimum(size)

    @staticmethod
    def sizeMaximum(size: int) -> int:
        return L.lib().sfft_size_maximum(size)

    def setSize(self, size: int) -> int:
        if size != self._size:
            plan = Plan(size, L.C2C, self._prec) if size > 0 else None
            self._size, self._plan = size, plan
        return self._size

    def setSizeMinimum(self, size: int) -> int:
        return self.setSize(FFT.sizeMinimum(size))

    def setSizeMaximum(self, size: int) -> int:
        return self.setSize(FFT.sizeMaximum(size))

    def size(self) -> int:
        return self._size

    @property
    def plan(self) -> Plan:
        return self._plan

    def _run(self, inp, out, direction: int, stream=None):
        n = self._size
        if n == 0:
            return out
        if isinstance(inp, np.ndarray) or isinstance(out, np.ndarray):
            return self._run_host(inp, out, direction)
        _check_device_pair(inp, out)
        if inp.dtype != self.dtype or out.dtype != self.dtype:
            raise TypeError(f"expected {self.dtype}")
        if inp.shape[-1] != n or out.shape != inp.shape:
            raise ValueError(f"last dimension must be {n} and shapes must match")
        batch = inp.numel() // n
        self._plan.execute_ptr(inp.data_ptr(), out.data_ptr(), batch, direction,
                               _stream_handle(stream))
        return out

    def _run_host(self, inp, out, direction):
        npdt = np.complex64 if self.dtype == torch.complex64 else np.complex128
        x = torch.from_numpy(np.ascontiguousarray(inp, dtype=npdt)).cuda()
        y = torch.empty_like(x)
        self._run(x, y, direction)
        res = y.cpu().numpy()
        if out is None:
            return res
        out[...] = res
        return out

    def fft(self, input, output, stream=None):
        """Forward transform (e^{-2 pi i nk/N}); output has input's shape."""
        return self._run(input, output, L.FORWARD, stream)

    def ifft(self, input, output, stream=None):
        """Unnormalised inverse transform: ifft(fft(x)) = N x."""
        return self._run(input, output, L.INVERSE, stream)


class RealFFT:
    """signalsmith::RealFFT<V, optionFlags> (signalsmith-fft.h:393-503) on the GPU."""

    _half_shift = False

    def __init__(self, size: int, fastDirection: int = 0, dtype=torch.float32):
        if dtype not in _REAL:
            raise TypeError("dtype must be torch.float32 or torch.float64")
        self.dtype = dtype
        self.cdtype = torch.complex64 if dtype == torch.float32 else torch.complex128
        self._prec = _REAL[dtype]
        self._half = 0
        self._plan = None
        if fastDirection > 0:
            size = type(self).sizeMinimum(size)
        if fastDirection < 0:
            size = type(self).sizeMaximum(size)
        self.setSize(size)

    @staticmethod
    def sizeMinimum(size: int) -> int:
        return L.lib().sfft_real_size_minimum(size)

    @staticmethod
    def sizeMaximum(size: int) -> int:
        return L.lib().sfft_real_size_maximum(size)

    def setSize(self, size: int) -> int:
        """Returns the half size, as the reference does (signalsmith-fft.h:434)."""
        half = size // 2
        if half != self._half or (self._plan is None and half > 0):
            kind = L.R2C_HALFSHIFT if self._half_shift else L.R2C
            plan = Plan(2 * half, kind, self._prec) if half > 0 else None
            self._half, self._plan = half, plan
        return self._half

    def setSizeMinimum(self, size: int) -> int:
        return self.setSize(type(self).sizeMinimum(size))

    def setSizeMaximum(self, size: int) -> int:
        return self.setSize(type(self).sizeMaximum(size))

    def size(self) -> int:
        return 2 * self._half

    @property
    def plan(self) -> Plan:
        return self._plan

    def fft(self, input, output, stream=None):
        """N reals -> N/2 complex bins; bin 0 = (DC, Nyquist) unless half-shifted."""
        n = self.size()
        if n == 0:
            return output
        if isinstance(input, np.ndarray):
            npdt = np.float32 if self.dtype == torch.float32 else np.float64
            x = torch.from_numpy(np.ascontiguousarray(input, dtype=npdt)).cuda()
            y = torch.empty(x.shape[:-1] + (n // 2,), dtype=self.cdtype, device="cuda")
            self.fft(x, y)
            output[..., : n // 2] = y.cpu().numpy()
            return output
        _check_device_pair(input, output)
        if input.dtype != self.dtype or output.dtype != self.cdtype:
            raise TypeError("real input / complex output dtypes expected")
        if input.shape[-1] != n or output.shape[-1] != n // 2 or \
                output.shape[:-1] != input.shape[:-1]:
            raise ValueError(f"input (..., {n}) and output (..., {n // 2}) expected")
        batch = input.numel() // n
        self._plan.execute_ptr(input.data_ptr(), output.data_ptr(), batch, L.FORWARD,
                               _stream_handle(stream))
        return output

    def ifft(self, input, output, stream=None):
        """N/2 complex bins -> N reals, unnormalised (round trip = N x)."""
        n = self.size()
        if n == 0:
            return output
        if isinstance(input, np.ndarray):
            npdt = np.complex64 if self.dtype == torch.float32 else np.complex128
            x = torch.from_numpy(np.ascontiguousarray(input[..., : n // 2], dtype=npdt)).cuda()
            y = torch.empty(x.shape[:-1] + (n,), dtype=self.dtype, device="cuda")
            self.ifft(x, y)
            output[...] = y.cpu().numpy()
            return output
        _check_device_pair(input, output)
        if input.dtype != self.cdtype or output.dtype != self.dtype:
            raise TypeError("complex input / real output dtypes expected")
        if input.shape[-1] != n // 2 or output.shape[-1] != n or \
                output.shape[:-1] != input.shape[:-1]:
            raise ValueError(f"input (..., {n // 2}) and output (..., {n}) expected")
        batch = input.numel() // (n // 2)
        self._plan.execute_ptr(input.data_ptr(), output.data_ptr(), batch, L.INVERSE,
                               _stream_handle(stream))
        return output


class ModifiedRealFFT(RealFFT):
    """signalsmith::ModifiedRealFFT<V> = RealFFT<V, FFTOptions::halfFreqShift>
    (signalsmith-fft.h:389-391, :505-508): bins at k + 1/2, no bin-0 packing."""

    _half_shift = True


def transform_bytes(n: int, kind: int, direction: int, precision: int):
    """(input bytes, output bytes) of one transform."""
    rb = 8 if precision == L.F64 else 4
    if kind == L.C2C:
        return 2 * n * rb, 2 * n * rb
    if direction == L.FORWARD:
        return n * rb, n * rb  # n reals -> n/2 complex
    return n * rb, n * rb


class UnevenBatch:
    """Transforms of many sizes, kinds and directions in one persistent,
    graph-captured launch (sfft_batch_*).  Inputs and outputs are packed back
    to back, in the given order, in two device arenas."""

    def __init__(self, sizes, directions, kinds=None, precision: int = L.F32):
        m = len(sizes)
        kinds = [L.C2C] * m if kinds is None else list(kinds)
        items = (L.BatchItem * max(1, m))()
        ioff = ooff = 0
        self.in_offsets, self.out_offsets = [], []
        for i, (n, d, k) in enumerate(zip(sizes, directions, kinds)):
            ib, ob = transform_bytes(int(n), int(k), int(d), precision)
            items[i] = L.BatchItem(int(n), int(k), int(d), 1, ioff, ooff)
            self.in_offsets.append(ioff)
            self.out_offsets.append(ooff)
            ioff += ib
            ooff += ob
        self.in_bytes, self.out_bytes = ioff, ooff
        self.sizes, self.directions, self.kinds = list(sizes), list(directions), kinds
        self.precision = precision
        h = ctypes.c_void_p()
        L.check(L.lib().sfft_batch_create(ctypes.byref(h), items, m, precision))
        self._h = h

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and L._lib is not None:
            L._lib.sfft_batch_destroy(h)
            self._h = None

    def _check(self, a: torch.Tensor, b: torch.Tensor):
        if not (a.is_cuda and b.is_cuda):
            raise ValueError("device arenas expected")
        if a.numel() * a.element_size() < self.in_bytes or \
                b.numel() * b.element_size() < self.out_bytes:
            raise ValueError("arena too small")

    def execute(self, in_arena: torch.Tensor, out_arena: torch.Tensor, stream=None):
        self._check(in_arena, out_arena)
        L.check(L.lib().sfft_batch_execute(self._h, ctypes.c_void_p(in_arena.data_ptr()),
                                           ctypes.c_void_p(out_arena.data_ptr()),
                                           ctypes.c_void_p(_stream_handle(stream))))
        return out_arena

    def execute_per_size(self, in_arena: torch.Tensor, out_arena: torch.Tensor, stream=None):
        self._check(in_arena, out_arena)
        L.check(L.lib().sfft_batch_execute_per_size(
            self._h, ctypes.c_void_p(in_arena.data_ptr()), ctypes.c_void_p(out_arena.data_ptr()),
            ctypes.c_void_p(_stream_handle(stream))))
        return out_arena


def fill(t: torch.Tensor, seed: int, dist: str = "uniform", stream=None) -> torch.Tensor:
    """Fill a device tensor with the seeded synthetic stream of fft_amd.synth
    (uniform [-1, 1) or normal), generated on the device."""
    prec = L.F64 if t.dtype in (torch.float64, torch.complex128) else L.F32
    count = t.numel() * (2 if t.is_complex() else 1)
    L.check(L.lib().sfft_fill(ctypes.c_void_p(t.data_ptr()), count, prec,
                              0 if dist == "uniform" else 1, seed,
                              ctypes.c_void_p(_stream_handle(stream))))
    return t


def _hints_dict(h) -> dict:
    return {"variant": h.variant, "radices": [h.radices[i] for i in range(h.nradices)],
            "transforms_per_block": h.transforms_per_block, "split": h.split,
            "elements_per_thread": h.elements_per_thread, "mfma": h.mfma}


def autotune(n: int, kind: int = L.C2C, precision: int = L.F32, batch: int = 0) -> dict:
    """Time the planner's candidates for (n, kind, precision) on the GPU and
    make the fastest the plan Plan(n, kind, precision) returns from now on
    (sfft_autotune).  batch 0: about 1 GB per timed launch."""
    r = L.TuneResult()
    L.check(L.lib().sfft_autotune(n, kind, precision, batch, ctypes.byref(r)))
    return {"hints": _hints_dict(r.hints), "ms": r.ms, "GB/s": r.gbs, "batch": r.batch,
            "candidates": r.candidates}


def save_tuning(path: str) -> None:
    """Write the tuned choices as JSON (keyed by GPU architecture and HIP runtime)."""
    L.check(L.lib().sfft_tune_save(str(path).encode()))


def load_tuning(path: str, check_device: bool = True) -> int:
    """Read tuned choices; returns how many were taken (0 if the file was
    tuned on another architecture or runtime and check_device is set)."""
    k = ctypes.c_size_t(0)
    L.check(L.lib().sfft_tune_load(str(path).encode(), 1 if check_device else 0, ctypes.byref(k)))
    return k.value


def clear_tuning() -> None:
    L.check(L.lib().sfft_tune_clear())
