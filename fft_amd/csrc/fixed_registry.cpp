// fixed_registry.cpp -- lookup table of the compiled fixed-plan kernels.
#include <mutex>
#include <vector>

#include "fixed.h"

namespace sfft {

static std::vector<FixedKernel> &registry() {
    static std::vector<FixedKernel> r;
    return r;
}
static std::mutex &registry_mutex() {
    static std::mutex m;
    return m;
}

void register_fixed(const FixedKernel &k) {
    std::lock_guard<std::mutex> lock(registry_mutex());
    registry().push_back(k);
}

const FixedKernel *find_fixed(int n, int prec, int real, int variant) {
    std::lock_guard<std::mutex> lock(registry_mutex());
    const int want = variant < 0 ? 0 : variant;
    for (const auto &k : registry())
        if (k.n == n && k.prec == prec && k.real == real && k.variant == want) return &k;
    return nullptr;
}

int fixed_variant_count(int n, int prec, int real) {
    std::lock_guard<std::mutex> lock(registry_mutex());
    int c = 0;
    for (const auto &k : registry())
        if (k.n == n && k.prec == prec && k.real == real) ++c;
    return c;
}

}  // namespace sfft
