"""fft_amd: MI355X-native engine behind the Signalsmith FFT API.

    FFT, RealFFT, ModifiedRealFFT   mirrors of signalsmith::FFT<V>, RealFFT<V>,
                                    ModifiedRealFFT<V> (fft_amd/api.py)
    Plan                            a cached engine plan (C ABI handle)
    UnevenBatch                     many sizes in one persistent launch
    fill                            seeded synthetic inputs generated in HBM
    autotune, save_tuning,          the autotuner and its JSON plan cache
    load_tuning, clear_tuning

All transforms run through libsfft.so (include/sfft.h); importing this package
does not load it -- the first call does, and raises if it is missing.
"""
from .api import (FFT, ModifiedRealFFT, Plan, RealFFT, UnevenBatch, autotune,  # noqa: F401
                  clear_tuning, fill, load_tuning, save_tuning)
