"""CPU checks of the C ABI: the library loads, exports every symbol that
include/*.h declares, and its host-only helpers are bit-exact against the
reference's size tables.  No compute calls (there is no GPU here)."""
import ctypes
import os
import re

import numpy as np

from fft_amd import _lib as L
from tests.conftest import GOLDEN, ROOT


def _declared_symbols():
    syms = set()
    inc = os.path.join(ROOT, "include")
    for fn in os.listdir(inc):
        if not fn.endswith(".h") or fn == "signalsmith-fft.h":
            continue
        text = open(os.path.join(inc, fn)).read()
        for m in re.finditer(r"^\s*(?:const\s+)?[A-Za-z_][A-Za-z0-9_ \*]*?\b(sfft_[a-z0-9_]+)\s*\(",
                             text, re.M):
            syms.add(m.group(1))
    return syms


def test_library_loads_and_exports_every_declared_symbol():
    lib = L.lib()
    declared = _declared_symbols()
    assert len(declared) >= 18, declared
    missing = [s for s in declared if not hasattr(lib, s)]
    assert not missing, missing
    assert set(L.EXPORTED) == declared, set(L.EXPORTED) ^ declared


def test_size_helpers_bit_exact_against_reference_tables():
    lib = L.lib()
    t = np.load(os.path.join(GOLDEN, "sizes.npz"))
    mn = np.array([lib.sfft_size_minimum(int(n)) for n in t["n_all"]])
    mx = np.array([lib.sfft_size_maximum(int(n)) for n in t["n_all"]])
    np.testing.assert_array_equal(mn, t["fft_min"])
    np.testing.assert_array_equal(mx, t["fft_max"])
    rmn = np.array([lib.sfft_real_size_minimum(int(n)) for n in t["n_pos"]])
    rmx = np.array([lib.sfft_real_size_maximum(int(n)) for n in t["n_pos"]])
    np.testing.assert_array_equal(rmn, t["real_min"])
    np.testing.assert_array_equal(rmx, t["real_max"])


def test_size_supported_flags():
    lib = L.lib()
    assert lib.sfft_size_supported(4096, L.C2C) == 1
    assert lib.sfft_size_supported(3000, L.C2C) == 1
    assert lib.sfft_size_supported(0, L.C2C) == 0
    assert lib.sfft_size_supported(2048, L.R2C) == 1
    assert lib.sfft_size_supported(2049, L.R2C) == 0
    assert lib.sfft_size_supported(17, L.C2C) == 1  # chirp-z plan
    assert lib.sfft_size_supported(34, L.R2C) == 1
    assert lib.sfft_size_supported((1 << 28) + 1, L.C2C) == 0


def test_last_error_is_empty_string_initially():
    assert L.lib().sfft_last_error() in (b"", None) or isinstance(L.lib().sfft_last_error(), bytes)


def test_python_mirror_has_reference_method_names():
    import fft_amd
    for cls in (fft_amd.FFT, fft_amd.RealFFT, fft_amd.ModifiedRealFFT):
        for name in ("setSize", "setSizeMinimum", "setSizeMaximum", "size", "sizeMinimum",
                     "sizeMaximum", "fft", "ifft"):
            assert hasattr(cls, name), (cls, name)
