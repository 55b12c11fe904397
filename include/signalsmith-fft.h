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
// V is float, double or long double.  float and double run in their own
// precision; long double, which the GPU has no arithmetic for, runs in FP64
// (values are converted when they are staged), so FFT<long double> has
// FFT<double>'s accuracy.  Other V do not compile.
//
// fft/ifft accept any random-access iterator or container of std::complex<V>
// (V for real input/output), as the reference does (signalsmith-fft.h:55-67).
// The data is staged through page-locked host memory: small one-pass lengths
// (up to 256 KiB per transform) are transformed by a kernel that reads and
// writes the staging buffers directly; larger ones are copied to and from
// HBM in chunks, each chunk's copy overlapping the host's staging of the
// next (input) or previous (output) chunk.
// Additions: fftDevice / ifftDevice run on device pointers (optionally a
// batch and a hipStream_t), and every method reports engine errors by
// throwing signalsmith::FFTError (the reference has no error path).
//
// Differences the reference leaves undefined, defined here: in-place use
// (input == output) works on host iterators (staging separates them) and on
// device pointers (two-pass lengths use a workspace); odd RealFFT sizes
// round down to even.  Link with -lsfft.
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

// The device type and engine precision a value type V runs in.
template <typename V>
struct Device {
    static constexpr bool supported = false;
};
template <>
struct Device<float> {
    static constexpr bool supported = true;
    using type = float;
    static constexpr int precision = SFFT_F32;
};
template <>
struct Device<double> {
    static constexpr bool supported = true;
    using type = double;
    static constexpr int precision = SFFT_F64;
};
template <>
struct Device<long double> {
    static constexpr bool supported = true;
    using type = double;
    static constexpr int precision = SFFT_F64;
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

// A plan handle (shared engine plan), page-locked staging buffers, and for
// lengths that are not transformed in place in the staging buffers, device
// buffers, a stream and per-chunk events.
class Engine {
  public:
    // Host-staged transforms up to this many bytes per buffer on one-pass
    // plans read and write the staging buffers from the kernel.
    static constexpr size_t kDirectBytes = 256 << 10;
    // Larger transforms copy in this many chunks (from 1 MiB up).
    static constexpr int kChunks = 8;
    static constexpr size_t kChunkFrom = 1 << 20;

    Engine() = default;
    Engine(const Engine &) = delete;
    Engine &operator=(const Engine &) = delete;
    ~Engine() { release(); }

    // Builds the new plan and buffers aside and takes them only once all
    // succeeded: on an error (FFTError) the engine keeps its previous state.
    // bytes: one transform's input, and also its output (equal for every kind).
    void plan(size_t n, int kind, int prec, size_t bytes) {
        Engine next;
        if (n) {
            check(sfft_plan_create(&next.plan_, n, kind, prec));
            sfft_plan_info info;
            check(sfft_plan_get_info(next.plan_, &info));
            check(sfft_malloc_host(&next.hin_, bytes));
            check(sfft_malloc_host(&next.hout_, bytes));
            next.bytes_ = bytes;
            next.direct_ = info.passes == 1 && info.chirp_length == 0 && bytes <= kDirectBytes;
            if (next.direct_) {
                check(sfft_host_device_pointer(&next.hin_dev_, next.hin_));
                check(sfft_host_device_pointer(&next.hout_dev_, next.hout_));
            } else {
                check(sfft_malloc_device(&next.din_, bytes));
                check(sfft_malloc_device(&next.dout_, bytes));
                next.chunks_ = bytes >= kChunkFrom ? kChunks : 1;
                for (int k = 0; k < next.chunks_; ++k) check(sfft_event_create(&next.ev_[k]));
            }
            // chirp-z lengths: the workspace of a host-staged call, allocated
            // once here so that fft/ifft allocate nothing
            check(sfft_plan_workspace_size(next.plan_, 1, 0, &next.ws_bytes_));
            if (next.ws_bytes_) check(sfft_malloc_device(&next.ws_, next.ws_bytes_));
            check(sfft_stream_create(&next.stream_));
        }
        swap(next);
    }
    void *host_in() const { return hin_; }
    void *host_out() const { return hout_; }

    // One transform of host data.  copy_in(i0, i1) writes input elements
    // [i0, i1) (of in_bytes each) into host_in(); copy_out(i0, i1) reads
    // output elements [i0, i1) (of out_bytes each) from host_out().
    template <class CopyIn, class CopyOut>
    void run_staged(int direction, size_t in_elems, size_t in_bytes, size_t out_elems,
                    size_t out_bytes, CopyIn &&copy_in, CopyOut &&copy_out) {
        if (direct_) {
            copy_in(0, in_elems);
            check(sfft_execute_workspace(plan_, hin_dev_, hout_dev_, 1, direction, ws_, ws_bytes_,
                                         stream_));
            check(sfft_stream_synchronize(stream_));
            copy_out(0, out_elems);
            return;
        }
        char *hi = static_cast<char *>(hin_), *ho = static_cast<char *>(hout_);
        char *di = static_cast<char *>(din_), *dout = static_cast<char *>(dout_);
        for (int k = 0; k < chunks_; ++k) {  // host copy of chunk k+1 overlaps the DMA of k
            const size_t e0 = in_elems * k / chunks_, e1 = in_elems * (k + 1) / chunks_;
            if (e1 == e0) continue;
            copy_in(e0, e1);
            check(sfft_memcpy(di + e0 * in_bytes, hi + e0 * in_bytes, (e1 - e0) * in_bytes,
                              SFFT_COPY_H2D, stream_));
        }
        check(sfft_execute_workspace(plan_, din_, dout_, 1, direction, ws_, ws_bytes_, stream_));
        for (int k = 0; k < chunks_; ++k) {
            const size_t e0 = out_elems * k / chunks_, e1 = out_elems * (k + 1) / chunks_;
            if (e1 > e0)
                check(sfft_memcpy(ho + e0 * out_bytes, dout + e0 * out_bytes,
                                  (e1 - e0) * out_bytes, SFFT_COPY_D2H, stream_));
            check(sfft_event_record(ev_[k], stream_));
        }
        for (int k = 0; k < chunks_; ++k) {  // host copy of chunk k overlaps the DMA of k+1
            const size_t e0 = out_elems * k / chunks_, e1 = out_elems * (k + 1) / chunks_;
            check(sfft_event_synchronize(ev_[k]));
            if (e1 > e0) copy_out(e0, e1);
        }
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
        std::swap(hin_dev_, o.hin_dev_);
        std::swap(hout_dev_, o.hout_dev_);
        std::swap(stream_, o.stream_);
        for (int k = 0; k < kChunks; ++k) std::swap(ev_[k], o.ev_[k]);
        std::swap(ws_, o.ws_);
        std::swap(ws_bytes_, o.ws_bytes_);
        std::swap(bytes_, o.bytes_);
        std::swap(direct_, o.direct_);
        std::swap(chunks_, o.chunks_);
    }
    void release() {
        if (stream_) sfft_stream_synchronize(stream_);
        if (plan_) sfft_plan_destroy(plan_);
        if (din_) sfft_free_device(din_);
        if (dout_) sfft_free_device(dout_);
        if (ws_) sfft_free_device(ws_);
        if (hin_) sfft_free_host(hin_);
        if (hout_) sfft_free_host(hout_);
        for (int k = 0; k < kChunks; ++k)
            if (ev_[k]) sfft_event_destroy(ev_[k]);
        if (stream_) sfft_stream_destroy(stream_);
        plan_ = nullptr;
        din_ = dout_ = hin_ = hout_ = hin_dev_ = hout_dev_ = stream_ = ws_ = nullptr;
        ws_bytes_ = 0;
        for (int k = 0; k < kChunks; ++k) ev_[k] = nullptr;
    }
    sfft_plan plan_ = nullptr;
    void *din_ = nullptr, *dout_ = nullptr, *hin_ = nullptr, *hout_ = nullptr;
    void *hin_dev_ = nullptr, *hout_dev_ = nullptr;
    void *stream_ = nullptr;
    void *ws_ = nullptr;  // workspace of a host-staged call (chirp-z lengths)
    size_t ws_bytes_ = 0;
    void *ev_[kChunks] = {};
    size_t bytes_ = 0;
    bool direct_ = false;
    int chunks_ = 1;
};

}  // namespace gpu_detail

template <typename V>
class FFT {
    static_assert(gpu_detail::Device<V>::supported,
                  "FFT<V> runs on the GPU for float, double and long double (as FP64)");
    using D = typename gpu_detail::Device<V>::type;
    using complex = std::complex<V>;
    using dcomplex = std::complex<D>;
    size_t _size = 0;
    gpu_detail::Engine engine;

    template <bool inverse, typename InputIterator, typename OutputIterator>
    void run(InputIterator input, OutputIterator output) {
        if (_size == 0) return;
        dcomplex *hi = static_cast<dcomplex *>(engine.host_in());
        const dcomplex *ho = static_cast<const dcomplex *>(engine.host_out());
        engine.run_staged(
            inverse ? SFFT_INVERSE : SFFT_FORWARD, _size, sizeof(dcomplex), _size,
            sizeof(dcomplex),
            [&](size_t i0, size_t i1) {
                for (size_t i = i0; i < i1; ++i) hi[i] = dcomplex(input[i]);
            },
            [&](size_t i0, size_t i1) {
                for (size_t i = i0; i < i1; ++i) output[i] = complex(ho[i]);
            });
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
            engine.plan(size, SFFT_C2C, gpu_detail::Device<V>::precision, size * sizeof(dcomplex));
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

    // Device pointers (HBM), `batch` transforms back to back, asynchronous on
    // `stream` (values of the device type: double for FFT<long double>).
    void fftDevice(const dcomplex *input, dcomplex *output, size_t batch = 1,
                   void *stream = nullptr) {
        if (_size) engine.run_device(input, output, batch, SFFT_FORWARD, stream);
    }
    void ifftDevice(const dcomplex *input, dcomplex *output, size_t batch = 1,
                    void *stream = nullptr) {
        if (_size) engine.run_device(input, output, batch, SFFT_INVERSE, stream);
    }
};

struct FFTOptions {
    static constexpr int halfFreqShift = 1;
};

template <typename V, int optionFlags = 0>
class RealFFT {
    static_assert(gpu_detail::Device<V>::supported,
                  "RealFFT<V> runs on the GPU for float, double and long double (as FP64)");
    static constexpr bool modified = (optionFlags & FFTOptions::halfFreqShift) != 0;
    using D = typename gpu_detail::Device<V>::type;
    using complex = std::complex<V>;
    using dcomplex = std::complex<D>;
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
    // (signalsmith-fft.h:416-435).  N reals and N/2 complex values take the
    // same bytes, so one staging size serves both directions.
    size_t setSize(size_t size) {
        const size_t half = size / 2;
        engine.plan(2 * half, modified ? SFFT_R2C_HALFSHIFT : SFFT_R2C,
                    gpu_detail::Device<V>::precision, 2 * half * sizeof(D));
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
        D *hi = static_cast<D *>(engine.host_in());
        const dcomplex *ho = static_cast<const dcomplex *>(engine.host_out());
        engine.run_staged(
            SFFT_FORWARD, 2 * _half, sizeof(D), _half, sizeof(dcomplex),
            [&](size_t i0, size_t i1) {
                for (size_t i = i0; i < i1; ++i) hi[i] = D(in[i]);
            },
            [&](size_t i0, size_t i1) {
                for (size_t i = i0; i < i1; ++i) out[i] = complex(ho[i]);
            });
    }

    // N/2 bins -> N reals, unnormalised (signalsmith-fft.h:475-502).
    template <typename InputIterator, typename OutputIterator>
    void ifft(InputIterator &&input, OutputIterator &&output) {
        if (_half == 0) return;
        auto in = gpu_detail::GetIterator<InputIterator>::get(input);
        auto out = gpu_detail::GetIterator<OutputIterator>::get(output);
        // the inverse reads N/2 complex and writes N reals: swap staging roles
        dcomplex *hi = static_cast<dcomplex *>(engine.host_in());
        const D *ho = static_cast<const D *>(engine.host_out());
        engine.run_staged(
            SFFT_INVERSE, _half, sizeof(dcomplex), 2 * _half, sizeof(D),
            [&](size_t i0, size_t i1) {
                for (size_t i = i0; i < i1; ++i) hi[i] = dcomplex(in[i]);
            },
            [&](size_t i0, size_t i1) {
                for (size_t i = i0; i < i1; ++i) out[i] = V(ho[i]);
            });
    }

    void fftDevice(const D *input, dcomplex *output, size_t batch = 1, void *stream = nullptr) {
        if (_half) engine.run_device(input, output, batch, SFFT_FORWARD, stream);
    }
    void ifftDevice(const dcomplex *input, D *output, size_t batch = 1, void *stream = nullptr) {
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
