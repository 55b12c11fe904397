// signalsmith-fft.h -- drop-in C++ facade of the MI355X transform engine.
//
// Same class templates, method names, argument meaning and return values as
// the reference header of the same name (signalsmith-fft.h:69-508):
//
//   signalsmith::FFT<V>                  FFT(size, fastDirection = 0), setSize,
//                                        setSizeMinimum/Maximum, size(),
//                                        static sizeMinimum/Maximum, fft, ifft
//   signalsmith::FFTOptions              halfFreqShift = 1
//   signalsmith::RealFFT<V, optionFlags> same set; setSize returns the half size
//   signalsmith::ModifiedRealFFT<V>      RealFFT<V, FFTOptions::halfFreqShift>
//
// V is float or double.  fft/ifft accept any random-access iterator or
// container of std::complex<V> (V for real input/output), as the reference
// does (signalsmith-fft.h:55-67); the data is staged through pinned host
// memory to HBM, transformed by libsfft.so (include/sfft.h) and copied back.
// Additions: fftDevice / ifftDevice run on device pointers (optionally a
// batch and a hipStream_t), and every method reports engine errors by
// throwing signalsmith::FFTError (the reference has no error path).
//
// Differences the reference leaves undefined, defined here: in-place use
// (input == output) is supported by the host path (staging separates them)
// and, on device pointers, by one-pass and chirp-z lengths (others throw);
// odd RealFFT sizes round down to even.  Link with -lsfft.
#ifndef SIGNALSMITH_FFT_GPU_H
#define SIGNALSMITH_FFT_GPU_H

#ifndef SIGNALSMITH_FFT_NAMESPACE
#define SIGNALSMITH_FFT_NAMESPACE signalsmith
#endif

#include <sfft.h>

#include <complex>
#include <cstddef>
#include <iterator>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <utility>

namespace SIGNALSMITH_FFT_NAMESPACE {

struct FFTError : std::runtime_error {
    int code;
    FFTError(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

namespace gpu_detail {

inline void check(int rc) {
    if (rc != SFFT_OK) throw FFTError(rc, std::string("sfft: ") + sfft_last_error());
}

template <typename V>
struct Precision;
template <>
struct Precision<float> {
    static constexpr int value = SFFT_F32;
};
template <>
struct Precision<double> {
    static constexpr int value = SFFT_F64;
};

// Containers give std::begin(c); anything else (pointer, iterator) is used as is.
template <typename T, typename = void>
struct GetIterator {
    static T get(const T &t) { return t; }
};
template <typename T>
struct GetIterator<T, decltype((void)std::begin(std::declval<T>()))> {
    static auto get(const T &t) -> decltype(std::begin(t)) { return std::begin(t); }
};

// A plan handle (shared engine plan) plus device and pinned staging buffers.
class Engine {
  public:
    Engine() = default;
    Engine(const Engine &) = delete;
    Engine &operator=(const Engine &) = delete;
    ~Engine() { release(); }

    // Builds the new plan and buffers aside and takes them only once all
    // succeeded: on an error (FFTError) the engine keeps its previous state.
    void plan(size_t n, int kind, int prec, size_t in_bytes, size_t out_bytes) {
        Engine next;
        if (n) {
            check(sfft_plan_create(&next.plan_, n, kind, prec));
            check(sfft_malloc_device(&next.din_, in_bytes));
            check(sfft_malloc_device(&next.dout_, out_bytes));
            check(sfft_malloc_host(&next.hin_, in_bytes));
            check(sfft_malloc_host(&next.hout_, out_bytes));
            next.in_bytes_ = in_bytes;
            next.out_bytes_ = out_bytes;
        }
        swap(next);
    }
    bool ready() const { return plan_ != nullptr; }
    void *host_in() const { return hin_; }
    void *host_out() const { return hout_; }

    // host staging -> device -> transform -> device -> host staging
    void run_staged(int direction) {
        check(sfft_memcpy(din_, hin_, in_bytes_, SFFT_COPY_H2D, nullptr));
        check(sfft_execute(plan_, din_, dout_, 1, direction, nullptr));
        check(sfft_memcpy(hout_, dout_, out_bytes_, SFFT_COPY_D2H, nullptr));
        check(sfft_stream_synchronize(nullptr));
    }
    void run_device(const void *in, void *out, size_t batch, int direction, void *stream) {
        check(sfft_execute(plan_, in, out, batch, direction, stream));
    }

  private:
    void swap(Engine &o) {
        std::swap(plan_, o.plan_);
        std::swap(din_, o.din_);
        std::swap(dout_, o.dout_);
        std::swap(hin_, o.hin_);
        std::swap(hout_, o.hout_);
        std::swap(in_bytes_, o.in_bytes_);
        std::swap(out_bytes_, o.out_bytes_);
    }
    void release() {
        if (plan_) sfft_plan_destroy(plan_);
        if (din_) sfft_free_device(din_);
        if (dout_) sfft_free_device(dout_);
        if (hin_) sfft_free_host(hin_);
        if (hout_) sfft_free_host(hout_);
        plan_ = nullptr;
        din_ = dout_ = hin_ = hout_ = nullptr;
    }
    sfft_plan plan_ = nullptr;
    void *din_ = nullptr, *dout_ = nullptr, *hin_ = nullptr, *hout_ = nullptr;
    size_t in_bytes_ = 0, out_bytes_ = 0;
};

}  // namespace gpu_detail

template <typename V>
class FFT {
    static_assert(std::is_same<V, float>::value || std::is_same<V, double>::value,
                  "FFT<V> runs on the GPU for float and double");
    using complex = std::complex<V>;
    size_t _size = 0;
    gpu_detail::Engine engine;

    template <bool inverse, typename InputIterator, typename OutputIterator>
    void run(InputIterator input, OutputIterator output) {
        if (_size == 0) return;
        complex *hi = static_cast<complex *>(engine.host_in());
        for (size_t i = 0; i < _size; ++i) hi[i] = input[i];
        engine.run_staged(inverse ? SFFT_INVERSE : SFFT_FORWARD);
        const complex *ho = static_cast<const complex *>(engine.host_out());
        for (size_t i = 0; i < _size; ++i) output[i] = ho[i];
    }

  public:
    // signalsmith-fft.h:327-348 (2^a * {1, 3, 9}, never a multiple of 5)
    static size_t sizeMinimum(size_t size) { return sfft_size_minimum(size); }
    static size_t sizeMaximum(size_t size) { return sfft_size_maximum(size); }

    FFT(size_t size, int fastDirection = 0) {
        if (fastDirection > 0) size = sizeMinimum(size);
        if (fastDirection < 0) size = sizeMaximum(size);
        this->setSize(size);
    }

    // Replans only on a change; returns the size (signalsmith-fft.h:356-363).
    size_t setSize(size_t size) {
        if (size != _size) {
            engine.plan(size, SFFT_C2C, gpu_detail::Precision<V>::value, size * sizeof(complex),
                        size * sizeof(complex));
            _size = size;
        }
        return _size;
    }
    size_t setSizeMinimum(size_t size) { return setSize(sizeMinimum(size)); }
    size_t setSizeMaximum(size_t size) { return setSize(sizeMaximum(size)); }
    const size_t &size() const { return _size; }

    template <typename InputIterator, typename OutputIterator>
    void fft(InputIterator &&input, OutputIterator &&output) {
        run<false>(gpu_detail::GetIterator<InputIterator>::get(input),
                   gpu_detail::GetIterator<OutputIterator>::get(output));
    }
    template <typename InputIterator, typename OutputIterator>
    void ifft(InputIterator &&input, OutputIterator &&output) {
        run<true>(gpu_detail::GetIterator<InputIterator>::get(input),
                  gpu_detail::GetIterator<OutputIterator>::get(output));
    }

    // Device pointers (HBM), `batch` transforms back to back, asynchronous on `stream`.
    void fftDevice(const complex *input, complex *output, size_t batch = 1, void *stream = nullptr) {
        if (_size) engine.run_device(input, output, batch, SFFT_FORWARD, stream);
    }
    void ifftDevice(const complex *input, complex *output, size_t batch = 1,
                    void *stream = nullptr) {
        if (_size) engine.run_device(input, output, batch, SFFT_INVERSE, stream);
    }
};

struct FFTOptions {
    static constexpr int halfFreqShift = 1;
};

template <typename V, int optionFlags = 0>
class RealFFT {
    static_assert(std::is_same<V, float>::value || std::is_same<V, double>::value,
                  "RealFFT<V> runs on the GPU for float and double");
    static constexpr bool modified = (optionFlags & FFTOptions::halfFreqShift) != 0;
    using complex = std::complex<V>;
    size_t _half = 0;
    gpu_detail::Engine engine;

  public:
    // signalsmith-fft.h:403-408 (the maximum may exceed its argument)
    static size_t sizeMinimum(size_t size) { return sfft_real_size_minimum(size); }
    static size_t sizeMaximum(size_t size) { return sfft_real_size_maximum(size); }

    RealFFT(size_t size, int fastDirection = 0) {
        if (fastDirection > 0) size = sizeMinimum(size);
        if (fastDirection < 0) size = sizeMaximum(size);
        this->setSize(size);
    }

    // Rebuilds and returns the HALF size, as the reference does
    // (signalsmith-fft.h:416-435).
    size_t setSize(size_t size) {
        const size_t half = size / 2;
        engine.plan(2 * half, modified ? SFFT_R2C_HALFSHIFT : SFFT_R2C,
                    gpu_detail::Precision<V>::value, 2 * half * sizeof(V), half * sizeof(complex));
        _half = half;
        return _half;
    }
    size_t setSizeMinimum(size_t size) { return setSize(sizeMinimum(size)); }
    size_t setSizeMaximum(size_t size) { return setSize(sizeMaximum(size)); }
    size_t size() const { return 2 * _half; }

    // N reals -> N/2 bins; bin 0 = (DC, Nyquist) unless half-shifted.  Only
    // output[0 .. N/2) is written (signalsmith-fft.h:446-473).
    template <typename InputIterator, typename OutputIterator>
    void fft(InputIterator &&input, OutputIterator &&output) {
        if (_half == 0) return;
        auto in = gpu_detail::GetIterator<InputIterator>::get(input);
        auto out = gpu_detail::GetIterator<OutputIterator>::get(output);
        V *hi = static_cast<V *>(engine.host_in());
        for (size_t i = 0; i < 2 * _half; ++i) hi[i] = in[i];
        engine.run_staged(SFFT_FORWARD);
        const complex *ho = static_cast<const complex *>(engine.host_out());
        for (size_t i = 0; i < _half; ++i) out[i] = ho[i];
    }

    // N/2 bins -> N reals, unnormalised (signalsmith-fft.h:475-502).
    template <typename InputIterator, typename OutputIterator>
    void ifft(InputIterator &&input, OutputIterator &&output) {
        if (_half == 0) return;
        auto in = gpu_detail::GetIterator<InputIterator>::get(input);
        auto out = gpu_detail::GetIterator<OutputIterator>::get(output);
        // the inverse reads N/2 complex and writes N reals: swap staging roles
        complex *hi = static_cast<complex *>(engine.host_in());
        for (size_t i = 0; i < _half; ++i) hi[i] = in[i];
        engine.run_staged(SFFT_INVERSE);
        const V *ho = static_cast<const V *>(engine.host_out());
        for (size_t i = 0; i < 2 * _half; ++i) out[i] = ho[i];
    }

    void fftDevice(const V *input, complex *output, size_t batch = 1, void *stream = nullptr) {
        if (_half) engine.run_device(input, output, batch, SFFT_FORWARD, stream);
    }
    void ifftDevice(const complex *input, V *output, size_t batch = 1, void *stream = nullptr) {
        if (_half) engine.run_device(input, output, batch, SFFT_INVERSE, stream);
    }
};

template <typename V>
struct ModifiedRealFFT : public RealFFT<V, FFTOptions::halfFreqShift> {
    using RealFFT<V, FFTOptions::halfFreqShift>::RealFFT;
};

}  // namespace SIGNALSMITH_FFT_NAMESPACE

#undef SIGNALSMITH_FFT_NAMESPACE
#endif  // SIGNALSMITH_FFT_GPU_H
