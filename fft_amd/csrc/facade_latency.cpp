// facade_latency.cpp -- host-staged latency of the drop-in header
// (include/signalsmith-fft.h): FFT<V>::fft of N complex values from one
// std::vector into another, then ifft back, timed on the wall clock; the
// median of repeated round trips is printed as one JSON object.
//
//     facade_latency [N] [repeats]
#include <signalsmith-fft.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <complex>
#include <cstdio>
#include <cstdlib>
#include <vector>

namespace {

unsigned long long g_state = 4096;
double uniform_pm1() {  // splitmix64 -> U[-1, 1)
    unsigned long long z = (g_state += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    return (double)(z >> 11) * (2.0 / 9007199254740992.0) - 1.0;
}

template <typename V>
void round_trip(const char *name, size_t n, int reps) {
    signalsmith::FFT<V> fft(n);
    std::vector<std::complex<V>> x(n), y(n), z(n);
    for (auto &v : x) v = std::complex<V>((V)uniform_pm1(), (V)uniform_pm1());
    for (int i = 0; i < 10; ++i) {
        fft.fft(x, y);
        fft.ifft(y, z);
    }
    std::vector<double> us;
    for (int i = 0; i < reps; ++i) {
        const auto t0 = std::chrono::steady_clock::now();
        fft.fft(x, y);
        fft.ifft(y, z);
        const auto t1 = std::chrono::steady_clock::now();
        us.push_back(std::chrono::duration<double, std::micro>(t1 - t0).count());
    }
    std::sort(us.begin(), us.end());
    double err = 0, ref = 0;
    for (size_t i = 0; i < n; ++i) {
        const std::complex<double> a(z[i]), b(x[i]);
        err += std::norm(a / (double)n - b);
        ref += std::norm(b);
    }
    std::printf("\"%s\": {\"host_staged_round_trip_us\": %.2f, \"round_trip_rel_l2\": %.3e}", name,
                us[us.size() / 2], std::sqrt(err / ref));
}

}  // namespace

int main(int argc, char **argv) {
    const size_t n = argc > 1 ? (size_t)std::strtoull(argv[1], nullptr, 10) : 4096;
    const int reps = argc > 2 ? std::atoi(argv[2]) : 200;
    try {
        std::printf("{\"n\": %zu, \"repeats\": %d, ", n, reps);
        round_trip<float>("f32", n, reps);
        std::printf(", ");
        round_trip<double>("f64", n, reps);
        std::printf("}\n");
    } catch (const signalsmith::FFTError &e) {
        std::printf("\n{\"error\": \"%s\"}\n", e.what());
        return 1;
    }
    return 0;
}
