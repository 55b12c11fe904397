// ref_driver.cpp -- C-ABI shim around the reference header, for test
// infrastructure only (golden fixtures and the CPU baseline of bench.py).
//
// Built by oracle/Makefile against signalsmith-fft.h where it lies in the
// reference tree (-I points there); the header is never copied into this
// repository.  The product never links or loads the result
// (oracle/_ref/libsfft_ref.so).
#include "signalsmith-fft.h"

#include <chrono>
#include <complex>
#include <cstddef>
#include <thread>
#include <vector>

namespace {

template <typename V>
void run_c2c(size_t n, size_t batch, const V *in, V *out, int inverse) {
    signalsmith::FFT<V> fft(n);
    const std::complex<V> *ci = reinterpret_cast<const std::complex<V> *>(in);
    std::complex<V> *co = reinterpret_cast<std::complex<V> *>(out);
    for (size_t b = 0; b < batch; ++b) {
        if (inverse)
            fft.ifft(ci + b * n, co + b * n);
        else
            fft.fft(ci + b * n, co + b * n);
    }
}

template <typename V, int flags>
void run_real(size_t n, size_t batch, const V *in, V *out, int inverse) {
    signalsmith::RealFFT<V, flags> fft(n);
    const size_t h = n / 2;
    for (size_t b = 0; b < batch; ++b) {
        if (inverse) {
            const std::complex<V> *ci = reinterpret_cast<const std::complex<V> *>(in) + b * h;
            fft.ifft(ci, out + b * n);
        } else {
            std::complex<V> *co = reinterpret_cast<std::complex<V> *>(out) + b * h;
            fft.fft(in + b * n, co);
        }
    }
}

// One transform per thread, each thread with its own instance (instances
// are not reentrant).  Thread t takes transforms t, t+T, t+2T, ...
template <typename V>
double timed_c2c(size_t n, size_t batch, const V *in, V *out, int inverse, int threads) {
    auto t0 = std::chrono::steady_clock::now();
    std::vector<std::thread> pool;
    for (int t = 0; t < threads; ++t) {
        pool.emplace_back([=]() {
            signalsmith::FFT<V> fft(n);
            const std::complex<V> *ci = reinterpret_cast<const std::complex<V> *>(in);
            std::complex<V> *co = reinterpret_cast<std::complex<V> *>(out);
            for (size_t b = (size_t)t; b < batch; b += (size_t)threads) {
                if (inverse)
                    fft.ifft(ci + b * n, co + b * n);
                else
                    fft.fft(ci + b * n, co + b * n);
            }
        });
    }
    for (auto &th : pool) th.join();
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

template <typename V>
double timed_real(size_t n, size_t batch, const V *in, V *out, int inverse, int threads) {
    auto t0 = std::chrono::steady_clock::now();
    std::vector<std::thread> pool;
    const size_t h = n / 2;
    for (int t = 0; t < threads; ++t) {
        pool.emplace_back([=]() {
            signalsmith::RealFFT<V> fft(n);
            for (size_t b = (size_t)t; b < batch; b += (size_t)threads) {
                if (inverse)
                    fft.ifft(reinterpret_cast<const std::complex<V> *>(in) + b * h, out + b * n);
                else
                    fft.fft(in + b * n, reinterpret_cast<std::complex<V> *>(out) + b * h);
            }
        });
    }
    for (auto &th : pool) th.join();
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

}  // namespace

extern "C" {

void ref_fft_c2c_f64(size_t n, size_t batch, const double *in, double *out, int inverse) {
    run_c2c<double>(n, batch, in, out, inverse);
}
void ref_fft_c2c_f32(size_t n, size_t batch, const float *in, float *out, int inverse) {
    run_c2c<float>(n, batch, in, out, inverse);
}
void ref_fft_real_f64(size_t n, size_t batch, const double *in, double *out, int inverse, int half_shift) {
    if (half_shift)
        run_real<double, signalsmith::FFTOptions::halfFreqShift>(n, batch, in, out, inverse);
    else
        run_real<double, 0>(n, batch, in, out, inverse);
}
void ref_fft_real_f32(size_t n, size_t batch, const float *in, float *out, int inverse, int half_shift) {
    if (half_shift)
        run_real<float, signalsmith::FFTOptions::halfFreqShift>(n, batch, in, out, inverse);
    else
        run_real<float, 0>(n, batch, in, out, inverse);
}

size_t ref_size_minimum(size_t n) { return signalsmith::FFT<double>::sizeMinimum(n); }
size_t ref_size_maximum(size_t n) { return signalsmith::FFT<double>::sizeMaximum(n); }
size_t ref_real_size_minimum(size_t n) { return signalsmith::RealFFT<double>::sizeMinimum(n); }
size_t ref_real_size_maximum(size_t n) { return signalsmith::RealFFT<double>::sizeMaximum(n); }

// Return values of the setters and size() on fresh objects.
// which: 0 setSize, 1 setSizeMinimum, 2 setSizeMaximum; real: RealFFT.
size_t ref_setter_return(int real, int which, size_t n) {
    if (real) {
        signalsmith::RealFFT<double> f(2);
        return which == 0 ? f.setSize(n) : which == 1 ? f.setSizeMinimum(n) : f.setSizeMaximum(n);
    }
    signalsmith::FFT<double> f(1);
    return which == 0 ? f.setSize(n) : which == 1 ? f.setSizeMinimum(n) : f.setSizeMaximum(n);
}
// size() after construction with a fastDirection.
size_t ref_ctor_size(int real, size_t n, int fast_direction) {
    if (real) return signalsmith::RealFFT<double>(n, fast_direction).size();
    return signalsmith::FFT<double>(n, fast_direction).size();
}

double ref_time_c2c_f32(size_t n, size_t batch, const float *in, float *out, int inverse, int threads) {
    return timed_c2c<float>(n, batch, in, out, inverse, threads);
}
double ref_time_c2c_f64(size_t n, size_t batch, const double *in, double *out, int inverse, int threads) {
    return timed_c2c<double>(n, batch, in, out, inverse, threads);
}
double ref_time_real_f32(size_t n, size_t batch, const float *in, float *out, int inverse, int threads) {
    return timed_real<float>(n, batch, in, out, inverse, threads);
}

}  // extern "C"
