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

const FixedKernel *find_fixed(int n, int prec, int real) {
    std::lock_guard<std::mutex> lock(registry_mutex());
    for (const auto &k : registry())
        if (k.n == n && k.prec == prec && k.real == real) return &k;
    return nullptr;
}

}  // namespace sfft
