// test_facade.cpp -- the reference's own tests (tests/00-fft.cpp,
// tests/01-real.cpp) re-expressed against the drop-in header, which runs
// every transform on the GPU through libsfft.so.  Exit status 0 = pass.
//
// Same checks and tolerances as the reference: closeEnough (rel. error below
// 1e-6 with the reference's energy term, tests/tests-common.h:17-33),
// single-bin closed forms, linearity, unnormalised inverse, size helpers,
// RealFFT against the complex FFT (bins, bin-0 packing, untouched second
// half, inverse) and ModifiedRealFFT against the half-bin rotation.
#include <signalsmith-fft.h>

#include <cmath>
#include <complex>
#include <cstdio>
#include <cstdlib>
#include <vector>

using cd = std::complex<double>;

static int failures = 0;
static const int FFTOptions_halfFreqShift = signalsmith::FFTOptions::halfFreqShift;
#define CHECK(cond, ...)                                   \
    do {                                                   \
        if (!(cond)) {                                     \
            std::printf("FAIL %s:%d: ", __FILE__, __LINE__); \
            std::printf(__VA_ARGS__);                      \
            std::printf("\n");                             \
            ++failures;                                    \
        }                                                  \
    } while (0)

static const double kPi = 3.14159265358979323846;

// deterministic inputs (no reliance on rand())
static unsigned long long g_state = 0x2545F4914F6CDD1Dull;
static double uniform_half() {  // U[-0.5, 0.5)
    g_state ^= g_state >> 12;
    g_state ^= g_state << 25;
    g_state ^= g_state >> 27;
    return (double)((g_state * 0x2545F4914F6CDD1Dull) >> 11) / 9007199254740992.0 - 0.5;
}

static bool close_enough(const std::vector<cd> &a, const std::vector<cd> &b) {
    if (a.size() != b.size()) return false;
    double err = 0, energy = 0;
    for (size_t i = 0; i < a.size(); ++i) {
        err += std::norm(a[i] - b[i]);
        energy += std::norm(a[i] * a[i]) + std::norm(b[i] * b[i]);
    }
    if (energy == 0) return true;
    return std::sqrt(err / energy) < 1e-6;
}

static std::vector<int> test_sizes() {  // tests/00-fft.cpp:8-16
    return {1, 2, 4, 8, 16, 32, 64, 128, 256, 3, 6, 9, 12, 18, 24, 5, 10, 15, 20, 25,
            7, 14, 21, 28, 49, 11, 13, 17, 19, 22, 23};
}

static void individual_bins() {
    for (int n : test_sizes()) {
        signalsmith::FFT<double> fft(n);
        std::vector<cd> in(n), copy(n), out(n), expect(n);
        for (int bin = 0; bin < n; ++bin) {
            for (int i = 0; i < n; ++i) {
                const double ph = 2 * kPi * i * bin / n;
                in[i] = copy[i] = cd(std::cos(ph), std::sin(ph));
                expect[i] = (i == bin) ? n : 0;
            }
            fft.fft(in, out);
            CHECK(close_enough(in, copy), "input changed, N=%d", n);
            CHECK(close_enough(out, expect), "bin %d of N=%d", bin, n);
        }
    }
}

static void linearity() {
    for (int n : test_sizes()) {
        signalsmith::FFT<double> fft(n);
        std::vector<cd> a(n), b(n), ab(n), fa(n), fb(n), fab(n), sum(n);
        for (int i = 0; i < n; ++i) {
            a[i] = cd(uniform_half(), uniform_half());
            b[i] = cd(uniform_half(), uniform_half());
            ab[i] = a[i] + b[i];
        }
        fft.fft(a, fa);
        fft.fft(b, fb);
        fft.fft(ab, fab);
        for (int i = 0; i < n; ++i) sum[i] = fa[i] + fb[i];
        CHECK(close_enough(fab, sum), "linearity N=%d", n);
    }
}

static void inverse(int harmonic) {
    for (int n : test_sizes()) {
        std::vector<cd> in(n), spec(n), out(n), expect(n);
        for (int i = 0; i < n; ++i) {
            if (harmonic >= 0) {
                const double ph = 2 * kPi * i * harmonic / n;
                in[i] = cd(std::cos(ph), std::sin(ph));
            } else {
                in[i] = cd(uniform_half(), uniform_half());
            }
            expect[i] = in[i] * (double)n;
        }
        signalsmith::FFT<double> fft(n);
        fft.fft(in, spec);
        fft.ifft(spec, out);
        CHECK(close_enough(out, expect), "inverse N=%d harmonic %d", n, harmonic);
    }
}

static void sizes() {  // tests/00-fft.cpp:168-187
    auto fast = [](size_t n, int &threes_fives) {
        threes_fives = 0;
        while (n % 2 == 0) n /= 2;
        while (n % 3 == 0) n /= 3, ++threes_fives;
        while (n % 5 == 0) n /= 5, ++threes_fives;
        return n == 1;
    };
    for (size_t i = 1; i < 1000; ++i) {
        const size_t above = signalsmith::FFT<double>::sizeMinimum(i);
        const size_t below = signalsmith::FFT<double>::sizeMaximum(i);
        int ca = 0, cb = 0;
        CHECK(above >= i && below <= i, "bounds at %zu", i);
        CHECK(fast(above, ca) && fast(below, cb), "fast sizes at %zu", i);
        CHECK(ca <= 2 && cb <= 2, "complexity at %zu", i);
    }
}

template <bool modified>
static void real_tests() {  // tests/01-real.cpp:17-100
    for (int n = 2; n < 100; n += 2) {
        if (!sfft_size_supported((size_t)n, SFFT_R2C)) continue;
        std::vector<cd> cin(n), cmid(n), cout(n), rmid(n);
        std::vector<double> rin(n), rout(n);
        signalsmith::FFT<double> fft(n);
        typename std::conditional<modified, signalsmith::ModifiedRealFFT<double>,
                                  signalsmith::RealFFT<double>>::type rfft(n);
        for (int i = 0; i < n; ++i) {
            const double v = uniform_half();
            cin[i] = v;
            rin[i] = v;
        }
        if (modified)
            for (int i = 0; i < n; ++i) cin[i] *= cd(std::cos(-kPi * i / n), std::sin(-kPi * i / n));
        for (int i = n / 2; i < n; ++i) rmid[i] = cd(52, 21);
        fft.fft(cin, cmid);
        rfft.fft(rin, rmid);
        if (!modified) {
            CHECK(cmid[0].imag() <= 1e-6, "DC imag N=%d", n);
            CHECK(std::abs(cmid[0].real() - rmid[0].real()) <= 1e-6, "DC N=%d", n);
            CHECK(std::abs(cmid[n / 2].real() - rmid[0].imag()) <= 1e-6, "Nyquist N=%d", n);
        }
        for (int i = modified ? 0 : 1; i < n / 2; ++i)
            CHECK(std::abs(cmid[i] - rmid[i]) <= n * 1e-6, "bin %d N=%d", i, n);
        for (int i = n / 2; i < n; ++i) CHECK(rmid[i] == cd(52, 21), "second half N=%d", n);
        fft.ifft(cmid, cout);
        rfft.ifft(rmid, rout);
        if (modified)
            for (int i = 0; i < n; ++i) cout[i] *= cd(std::cos(kPi * i / n), std::sin(kPi * i / n));
        for (int i = 0; i < n; ++i) {
            CHECK(cout[i].imag() <= n * 1e-6, "inverse imag N=%d", n);
            CHECK(std::abs(cout[i].real() - rout[i]) <= n * 1e-6, "inverse N=%d i=%d", n, i);
        }
    }
}

static void return_values() {  // setSize quirks (signalsmith-fft.h:356-372, :416-444)
    signalsmith::FFT<float> f(1);
    CHECK(f.setSize(1920) == 1920 && f.size() == 1920, "FFT setSize");
    CHECK(f.setSizeMinimum(1000) == 1024, "FFT setSizeMinimum");
    CHECK(f.setSizeMaximum(1000) == 768, "FFT setSizeMaximum");
    signalsmith::RealFFT<float> r(2);
    CHECK(r.setSize(2048) == 1024 && r.size() == 2048, "RealFFT setSize returns half");
    CHECK(signalsmith::RealFFT<float>::sizeMaximum(10) == 12, "RealFFT sizeMaximum quirk");
    CHECK(signalsmith::FFT<double>(100, 1).size() == 128, "fastDirection +1");
    CHECK(signalsmith::FFT<double>(100, -1).size() == 96, "fastDirection -1");
    CHECK(signalsmith::FFT<double>(17).size() == 17, "prime length plans (chirp-z)");
    bool threw = false;
    try {
        signalsmith::FFT<double> bad((size_t(1) << 28) + 1);
    } catch (const signalsmith::FFTError &e) {
        threw = e.code == SFFT_ERR_UNSUPPORTED_SIZE;
    }
    CHECK(threw, "oversize length reports SFFT_ERR_UNSUPPORTED_SIZE");
}

static void float_device_batch() {  // device pointers, a batch, FFT<float>
    const int n = 4096, batch = 8;
    signalsmith::FFT<float> fft(n);
    std::vector<std::complex<float>> host(n * batch), back(n * batch);
    for (auto &v : host) v = std::complex<float>((float)uniform_half(), (float)uniform_half());
    void *din = nullptr, *dout = nullptr, *dback = nullptr;
    const size_t bytes = host.size() * sizeof(host[0]);
    sfft_malloc_device(&din, bytes);
    sfft_malloc_device(&dout, bytes);
    sfft_malloc_device(&dback, bytes);
    sfft_memcpy(din, host.data(), bytes, SFFT_COPY_H2D, nullptr);
    fft.fftDevice((const std::complex<float> *)din, (std::complex<float> *)dout, batch);
    fft.ifftDevice((const std::complex<float> *)dout, (std::complex<float> *)dback, batch);
    sfft_memcpy(back.data(), dback, bytes, SFFT_COPY_D2H, nullptr);
    sfft_stream_synchronize(nullptr);
    double err = 0, ref = 0;
    for (size_t i = 0; i < host.size(); ++i) {
        err += std::norm(std::complex<double>(back[i]) / (double)n - std::complex<double>(host[i]));
        ref += std::norm(std::complex<double>(host[i]));
    }
    CHECK(std::sqrt(err / ref) < 1e-6, "float round trip rel L2 %g", std::sqrt(err / ref));
    sfft_free_device(din);
    sfft_free_device(dout);
    sfft_free_device(dback);
}

// Host staging, both routes (the kernel on the page-locked buffers for small
// one-pass lengths, chunked copies for large ones), against the device-pointer
// path on the same values: bitwise equal.
template <typename V>
static void staged_matches_device(size_t n) {
    using C = std::complex<V>;
    signalsmith::FFT<V> fft(n);
    std::vector<C> in(n), out(n), back(n), dev_out(n);
    for (auto &v : in) v = C((V)uniform_half(), (V)uniform_half());
    fft.fft(in, out);
    void *din = nullptr, *dout = nullptr;
    const size_t bytes = n * sizeof(C);
    sfft_malloc_device(&din, bytes);
    sfft_malloc_device(&dout, bytes);
    sfft_memcpy(din, in.data(), bytes, SFFT_COPY_H2D, nullptr);
    fft.fftDevice((const C *)din, (C *)dout);
    sfft_memcpy(dev_out.data(), dout, bytes, SFFT_COPY_D2H, nullptr);
    sfft_stream_synchronize(nullptr);
    bool same = true;
    for (size_t i = 0; i < n; ++i) same = same && out[i] == dev_out[i];
    CHECK(same, "host-staged and device results differ, N=%zu", n);
    // in place on the host vector, and the inverse back
    std::vector<C> inplace = in;
    fft.fft(inplace, inplace);
    same = true;
    for (size_t i = 0; i < n; ++i) same = same && inplace[i] == out[i];
    CHECK(same, "in-place host transform differs, N=%zu", n);
    fft.ifft(out, back);
    double err = 0, ref = 0;
    for (size_t i = 0; i < n; ++i) {
        err += std::norm(std::complex<double>(back[i]) / (double)n - std::complex<double>(in[i]));
        ref += std::norm(std::complex<double>(in[i]));
    }
    CHECK(std::sqrt(err / ref) < (sizeof(V) == 4 ? 1e-6 : 1e-13), "round trip N=%zu: %g", n,
          std::sqrt(err / ref));
    sfft_free_device(din);
    sfft_free_device(dout);
}

template <int flags>
static void real_staged_matches_device(size_t n) {
    signalsmith::RealFFT<float, flags> rfft(n);
    std::vector<float> in(n), back(n);
    std::vector<std::complex<float>> out(n / 2), dev_out(n / 2);
    for (auto &v : in) v = (float)uniform_half();
    rfft.fft(in, out);
    void *din = nullptr, *dout = nullptr;
    sfft_malloc_device(&din, n * sizeof(float));
    sfft_malloc_device(&dout, n * sizeof(float));
    sfft_memcpy(din, in.data(), n * sizeof(float), SFFT_COPY_H2D, nullptr);
    rfft.fftDevice((const float *)din, (std::complex<float> *)dout);
    sfft_memcpy(dev_out.data(), dout, n * sizeof(float), SFFT_COPY_D2H, nullptr);
    sfft_stream_synchronize(nullptr);
    bool same = true;
    for (size_t i = 0; i < n / 2; ++i) same = same && out[i] == dev_out[i];
    CHECK(same, "real host-staged and device results differ, N=%zu", n);
    rfft.ifft(out, back);
    double err = 0, ref = 0;
    for (size_t i = 0; i < n; ++i) {
        err += (back[i] / (double)n - in[i]) * (back[i] / (double)n - in[i]);
        ref += (double)in[i] * in[i];
    }
    CHECK(std::sqrt(err / ref) < 1e-6, "real round trip N=%zu: %g", n, std::sqrt(err / ref));
    sfft_free_device(din);
    sfft_free_device(dout);
}

// FFT<long double> runs in FP64: the same values as FFT<double> on the same
// (double-representable) input, and the single-bin closed form.
static void long_double() {
    const int n = 96;
    signalsmith::FFT<long double> fl(n);
    signalsmith::FFT<double> fd(n);
    std::vector<std::complex<long double>> in(n), out(n);
    std::vector<cd> ind(n), outd(n);
    for (int i = 0; i < n; ++i) {
        ind[i] = cd(uniform_half(), uniform_half());
        in[i] = std::complex<long double>(ind[i]);
    }
    fl.fft(in, out);
    fd.fft(ind, outd);
    bool same = true;
    for (int i = 0; i < n; ++i) same = same && cd(out[i]) == outd[i];
    CHECK(same, "FFT<long double> differs from FFT<double>");
    const int bin = 5;
    for (int i = 0; i < n; ++i) {
        const long double ph = 2 * 3.14159265358979323846264338327950288L * i * bin / n;
        in[i] = std::complex<long double>(std::cos(ph), std::sin(ph));
    }
    fl.fft(in, out);
    double err = 0;
    for (int i = 0; i < n; ++i) err += std::norm(cd(out[i]) - cd(i == bin ? n : 0, 0));
    CHECK(std::sqrt(err) / n < 1e-13, "FFT<long double> single bin: %g", std::sqrt(err) / n);
    signalsmith::RealFFT<long double> rl(64);
    std::vector<long double> rin(64);
    std::vector<std::complex<long double>> rout(32);
    for (auto &v : rin) v = uniform_half();
    rl.fft(rin, rout);
    CHECK(std::abs(cd(rout[0]).real() - [&] {
              double s = 0;
              for (auto v : rin) s += (double)v;
              return s;
          }()) < 1e-12,
          "RealFFT<long double> DC");
}

// A failed setSize leaves the object as it was (strong exception guarantee).
static void failed_set_size_keeps_state() {
    signalsmith::FFT<float> f(256);
    bool threw = false;
    try {
        f.setSize((size_t(1) << 28) + 1);
    } catch (const signalsmith::FFTError &) {
        threw = true;
    }
    CHECK(threw, "oversize setSize throws");
    CHECK(f.size() == 256, "size kept after a failed setSize");
    std::vector<std::complex<float>> in(256, std::complex<float>(1, 0)), out(256);
    f.fft(in, out);
    CHECK(std::abs(out[0] - std::complex<float>(256, 0)) < 1e-3, "transform after a failed setSize");
}

int main() {
    individual_bins();
    linearity();
    inverse(1);
    inverse(-1);
    sizes();
    real_tests<false>();
    real_tests<true>();
    return_values();
    float_device_batch();
    staged_matches_device<float>(4096);        // kernel on the staging buffers
    staged_matches_device<float>(1 << 17);     // 1 MiB: chunked copies, two passes
    staged_matches_device<double>(3 << 14);    // chunked copies, one pass
    staged_matches_device<float>(1009);        // chirp-z
    real_staged_matches_device<0>(2048);
    real_staged_matches_device<FFTOptions_halfFreqShift>(1 << 19);
    long_double();
    failed_set_size_keeps_state();
    if (failures) {
        std::printf("%d failures\n", failures);
        return 1;
    }
    std::printf("all facade tests passed\n");
    return 0;
}
