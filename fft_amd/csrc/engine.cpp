// engine.cpp -- planner, device tables, plan cache and pass execution.
#include "engine.h"
#include "tune.h"

#include <sfft.h>

#include <algorithm>
#include <cmath>
#include <map>
#include <mutex>
#include <cstdlib>
#include <tuple>
#include <type_traits>

namespace sfft {

void check_hip(hipError_t e, const char *what) {
    if (e != hipSuccess)
        throw Error(SFFT_ERR_DEVICE, std::string(what) + ": " + hipGetErrorString(e));
}

const DeviceInfo &device_info() {
    static DeviceInfo info;
    static std::once_flag once;
    std::call_once(once, []() {
        int dev = 0;
        check_hip(hipGetDevice(&dev), "hipGetDevice");
        hipDeviceProp_t p;
        check_hip(hipGetDeviceProperties(&p, dev), "hipGetDeviceProperties");
        info.device = dev;
        info.cu_count = p.multiProcessorCount;
        info.lds_bytes = p.sharedMemPerBlock;
        info.arch = p.gcnArchName;
    });
    return info;
}

// ------------------------------------------------------------ size helpers --
// Same integer steps as the reference (signalsmith-fft.h:317-348): mantissas
// below 32 in {0,1,2,3,4,6,8,9,12,16,18,24} are accepted.

static bool accepted_mantissa(size_t m) {
    switch (m) {
    case 0: case 1: case 2: case 3: case 4: case 6: case 8: case 9: case 12: case 16: case 18:
    case 24:
        return true;
    default:
        return false;
    }
}

size_t size_minimum(size_t n) {
    size_t scale = 1;
    for (; n >= 32; scale *= 2) n = (n - 1) / 2 + 1;
    while (n < 32 && !accepted_mantissa(n)) ++n;
    return scale * n;
}

size_t size_maximum(size_t n) {
    size_t scale = 1;
    for (; n >= 32; scale *= 2) n /= 2;
    while (n > 1 && !accepted_mantissa(n)) --n;
    return scale * n;
}

// signalsmith-fft.h:403-408 (the maximum is built on sizeMinimum of the half).
size_t real_size_minimum(size_t n) { return (size_minimum((n - 1) / 2) + 1) * 2; }
size_t real_size_maximum(size_t n) { return size_minimum(n / 2) * 2; }

// --------------------------------------------------------------- planning --

static const int kRadixSet[] = {25, 16, 15, 13, 12, 11, 10, 9, 8, 7, 6, 5, 4, 3, 2};

static int butterflies_per_thread(int radix, int ept) { return std::max(1, ept / radix); }

// Threads per slot so every stage fits its per-thread butterfly budget.
static int threads_for(int n, const std::vector<int> &radices, int ept) {
    int tpt = 1;
    for (int r : radices) {
        const int nb = n / r;
        const int b = butterflies_per_thread(r, ept);
        tpt = std::max(tpt, (nb + b - 1) / b);
    }
    return tpt;
}

// Cost: sum over stages of thread-slot capacity / work (1.0 = every thread
// busy), i.e. the number of stages inflated by idle lanes.
static double radix_cost(int n, const std::vector<int> &radices, int ept, int tpt) {
    double c = 0;
    for (int r : radices) c += (double)tpt * butterflies_per_thread(r, ept) * r / n;
    return c;
}

static void search_radices(int rem, int max_r, std::vector<int> &cur, int n, int ept,
                           int max_threads, double &best_cost, std::vector<int> &best) {
    // 25 complex values exceed the register budget of the 16-element instance
    // in FP64 and of small per-thread budgets: only offered when ept >= 16 and
    // the caller allows it (max_r).
    if (rem == 1) {
        const int tpt = threads_for(n, cur, ept);
        if (tpt > max_threads) return;
        const double c = radix_cost(n, cur, ept, tpt);
        if (c < best_cost - 1e-9) {
            best_cost = c;
            best = cur;
        }
        return;
    }
    if ((int)cur.size() >= kMaxStages) return;
    for (int r : kRadixSet) {
        if (r > max_r || rem % r) continue;
        cur.push_back(r);
        search_radices(rem / r, r, cur, n, ept, max_threads, best_cost, best);
        cur.pop_back();
    }
}

bool choose_radices(int n, int ept, std::vector<int> &radices, int max_radix, int max_threads) {
    radices.clear();
    if (n < 1) return false;
    if (n == 1) {
        radices.push_back(1);
        return true;
    }
    std::vector<int> cur;
    double best_cost = 1e30;
    search_radices(n, max_radix, cur, n, ept, max_threads, best_cost, radices);
    return !radices.empty();
}

static bool only_small_primes(size_t n) {
    for (size_t p : {2, 3, 5, 7, 11, 13})
        while (n % p == 0) n /= p;
    return n == 1;
}

bool supported_length(size_t n) { return n >= 1 && n <= (1ull << 28) && only_small_primes(n); }

// Largest length a chirp-z plan takes (its convolution length stays <= 2^28).
constexpr size_t kMaxBluestein = 1ull << 27;

bool plannable_length(size_t n) { return supported_length(n) || (n >= 1 && n <= kMaxBluestein); }

// Smallest convolution length >= 2N - 1 made of the factors 2, 3, 5.
static size_t bluestein_length(size_t n) {
    const size_t lo = 2 * n - 1;
    size_t best = 0;
    for (size_t p2 = 1; p2 < 4 * lo; p2 *= 2)
        for (size_t p3 = p2; p3 < 4 * lo; p3 *= 3)
            for (size_t p5 = p3; p5 < 4 * lo; p5 *= 5)
                if (p5 >= lo && (!best || p5 < best)) best = p5;
    return best;
}

static size_t elem_bytes(int prec) { return arith_f64(prec) ? 16 : 8; }
// Two-pass plans up to this many bytes of elements keep a full table of
// their four-step factors (one entry per element, held in L2 / the
// last-level cache while a batch streams through).  Larger plans use the
// two-level fine x coarse table, whose gathers cost more per element but
// whose size does not grow with N (a full table at 2^24 F32 measured 1.43 ->
// 1.19 TB/s: the single transform reads its 128 MB table once).
constexpr size_t kFullPostTwBytes = 8u << 20;
// Complex FP32 length from which a runtime-planned single pass gives way to
// a two-pass split over compiled column kernels (Plan::Plan).
constexpr size_t kTwoPassFrom = 16000;

// Measured first-pass column lengths N1 of complex FP32 two-pass plans:
// every split of the fast lengths in [16000, 2^18] whose two column lengths
// have compiled kernels was timed (batched, 256 MB launches), and a length
// is listed where its fastest split beat the planner's scored choice by
// more than 5 % (81 of 185 lengths; up to 1.22x, median of all lengths 2.04
// -> 2.17 TB/s).  The first five lengths, below 16000, have no compiled
// single pass and measured faster split than in one runtime-planned pass
// (10935: 1.81 -> 2.14 TB/s, 12150: 2.25 -> 2.39, 13122: 1.61 -> 2.06,
// 14580: 2.08 -> 2.37, 15552: 1.79 -> 2.29); 16000 measured 2.37 -> 2.62
// TB/s with a first pass of 128 after the 125-length column kernel's change,
// and a re-sweep of the lengths divisible by 125 then set ten more entries
// (16875, 18000, 18750, 20000, 48000, 50000, 60000, 90000, 93750, 180000:
// 1.06-1.20x over the previous choice); a final re-sweep of all of them with
// the final kernels set ten more (31104, 32400, 36864, 46080, 48600, 54000,
// 55296, 147456, 184320, 196608: 1.07-1.17x).
// The last two, 2^21 and 2^22, measured 3-5 % faster with first-pass
// columns of 4096 in two runs each (2^21 x 8: 1.69-1.74 -> 1.79-1.81;
// 2^22 x 4: 1.64-1.70 -> 1.74).  Sorted by length.
struct MeasuredSplit {
    uint32_t n, n1;
};
static const MeasuredSplit kMeasuredSplits[] = {
    {10935, 135}, {12150, 81}, {13122, 81}, {14580, 81}, {15552, 192}, {16000, 128},
    {16875, 135}, {18000, 144}, {18225, 81}, {18750, 150}, {19683, 243}, {20000, 160},
    {24000, 160}, {31104, 128}, {32400, 144}, {34560, 192}, {34992, 144}, {36000, 160},
    {36864, 144}, {38880, 160}, {39366, 243}, {43200, 192}, {46080, 256}, {46656, 192},
    {46875, 375}, {48000, 160}, {48600, 243}, {50000, 250}, {51200, 256}, {54000, 240},
    {54675, 243}, {55296, 128}, {56250, 250}, {60000, 150}, {64000, 160}, {64800, 160},
    {67500, 300}, {69120, 160}, {72000, 144}, {72900, 150}, {73728, 512}, {77760, 160},
    {78125, 625}, {80000, 160}, {81000, 360}, {81920, 512}, {82944, 192}, {84375, 375},
    {86400, 360}, {87480, 180}, {90000, 144}, {91125, 405}, {92160, 360}, {93312, 192},
    {93750, 150}, {96000, 400}, {97200, 432}, {98304, 512}, {100000, 500}, {101250, 450},
    {102400, 256}, {103680, 432}, {104976, 432}, {108000, 432}, {109350, 729}, {110592, 256},
    {112500, 500}, {115200, 480}, {116640, 288}, {120000, 500}, {121500, 540}, {122880, 512},
    {124416, 288}, {125000, 625}, {129600, 432}, {131220, 729}, {135000, 625}, {139968, 288},
    {140625, 625}, {144000, 288}, {145800, 729}, {147456, 144}, {155520, 360}, {156250, 625},
    {157464, 729}, {163840, 160}, {164025, 729}, {174960, 432}, {180000, 288}, {182250, 375},
    {184320, 360}, {196608, 192}, {202500, 500}, {204800, 1024}, {216000, 432}, {218700, 729},
    {225000, 360}, {243000, 500}, {245760, 1024}, {248832, 512}, {253125, 625}, {2097152, 4096},
    {4194304, 4096},
};

// The same for real FP32 two-pass plans (RealFFT, not the half-bin shifted
// kind), keyed by the complex half length h: every split of the even fast
// lengths in [32000, 2^19] was timed over R2C forward and C2R inverse
// together, and 107 of 177 half lengths keep a split that beat the scored
// choice by more than 5 % (median of all 1.08x, up to 1.32x).  A re-sweep of
// the half lengths divisible by 125 after the 125-length column change
// added 30000 (1.07x), and a final re-sweep of all 177 set 23040 and 186624
// (1.07x, 1.09x).
static const MeasuredSplit kMeasuredRealSplits[] = {
    {19440, 240}, {19683, 243}, {20000, 160}, {23040, 128}, {24000, 160}, {25000, 200},
    {25920, 192}, {27648, 192}, {28125, 225}, {29160, 360}, {30000, 240}, {30720, 128},
    {31104, 128}, {32805, 243}, {34560, 192}, {34992, 144}, {36000, 240}, {36450, 243},
    {36864, 256}, {38880, 160}, {39366, 243}, {40000, 160}, {40960, 256}, {43200, 192},
    {45000, 225}, {46080, 256}, {46656, 192}, {46875, 375}, {48000, 240}, {48600, 243},
    {51200, 256}, {51840, 192}, {55296, 256}, {61440, 256}, {62208, 256}, {62500, 500},
    {64000, 256}, {67500, 450}, {69984, 144}, {72000, 480}, {72900, 180}, {75000, 500},
    {77760, 192}, {78125, 625}, {80000, 160}, {81000, 450}, {82944, 256}, {86400, 480},
    {87480, 180}, {90000, 450}, {91125, 225}, {92160, 512}, {93312, 192}, {93750, 625},
    {96000, 240}, {97200, 240}, {100000, 160}, {101250, 450}, {102400, 256}, {103680, 256},
    {104976, 144}, {108000, 480}, {109350, 225}, {110592, 256}, {112500, 500}, {115200, 512},
    {116640, 480}, {120000, 240}, {121500, 225}, {122880, 512}, {124416, 256}, {125000, 625},
    {128000, 256}, {129600, 240}, {131220, 180}, {135000, 450}, {138240, 256}, {139968, 192},
    {140625, 625}, {144000, 480}, {147456, 1024}, {150000, 240}, {151875, 625}, {153600, 1024},
    {155520, 480}, {156250, 625}, {157464, 729}, {160000, 256}, {163840, 1024}, {165888, 512},
    {168750, 625}, {172800, 480}, {174960, 240}, {180000, 450}, {182250, 450}, {184320, 1024},
    {186624, 256}, {187500, 625}, {192000, 480}, {194400, 480}, {196608, 1024}, {204800, 1024},
    {207360, 480}, {221184, 512}, {234375, 625}, {245760, 1024}, {248832, 512}, {253125, 625},
    {256000, 512}, {262144, 1024},
};

// The same for the power-of-two two-pass lengths of the other precisions
// (complex, about 128 MB per launch, every split timed twice; listed where
// both timings beat both of the scored split's by more than 5 %, and at 2^20
// for 16-bit storage and F32X, where 2^20 x 16 measured f16 110-113 -> 105-106
// us and F32X 222-225 -> 196 us).  16-bit storage (IEEE half and bfloat16:
// 2^18 1.16-1.22x, 2^19 1.11-1.18x, 2^21 1.07-1.08x):
static const MeasuredSplit kMeasuredSplits16[] = {
    {262144, 2048}, {524288, 2048}, {1048576, 4096}, {2097152, 2048},
};
// FP32 storage under FP64 arithmetic (F32X; 2^19 1.13x):
static const MeasuredSplit kMeasuredSplitsF32X[] = {
    {524288, 2048}, {1048576, 4096},
};
// FP64 (2^21 1.17x, 2^22 1.10x):
static const MeasuredSplit kMeasuredSplitsF64[] = {
    {2097152, 4096}, {4194304, 4096},
};

template <size_t K>
static size_t lookup_split(const MeasuredSplit (&t)[K], size_t n) {
    const auto *b = std::begin(t), *e = std::end(t);
    const auto *it = std::lower_bound(b, e, n, [](const MeasuredSplit &m, size_t v) { return m.n < v; });
    return it != e && it->n == n ? it->n1 : 0;
}
static int max_block(int prec) {
    return stored_16(prec) ? kMaxBlock16 : (arith_f64(prec) ? kMaxBlock<double>() : kMaxBlock<float>());
}

// Dynamic LDS a layout may use: the device limit less room for the kernels'
// static __shared__ variables.
static size_t lds_budget() { return device_info().lds_bytes - kLdsReserve; }

static int floor_pow2(int x) {
    int p = 1;
    while (p * 2 <= x) p *= 2;
    return p;
}

// Fills the stage part of a PassDesc for length n with the given radices.
static void fill_stages(PassDesc &pd, int n, const std::vector<int> &radices) {
    pd.n = n;
    pd.nstages = (int)radices.size();
    int ns = 1;
    for (size_t s = 0; s < radices.size(); ++s) {
        pd.st[s].radix = radices[s];
        pd.st[s].ns = ns;
        pd.st[s].tw_off = 0;
        pd.st[s].div = make_fastdiv((uint32_t)ns);
        ns *= radices[s];
    }
}

// Single-pass layout for sub-transform length n; false if it does not fit.
static bool layout_single(int n, int prec, bool uses_lds_always, const PlanHints *h,
                          PassPlan &pp) {
    const size_t eb = elem_bytes(prec);
    std::vector<int> radices;
    int ept = 0;
    const int epts[2] = {kEptSmall, kEptLarge};
    for (int e : epts) {
        if (arith_f64(prec) && e > kEptSmall) continue;
        if (h && h->ept && h->ept != e) continue;
        if (h && !h->radices.empty()) {
            int prod = 1;
            for (int r : h->radices) prod *= r;
            if (prod != n) return false;
            radices = h->radices;
            // radix-32 stages run in the 32-element instances only
            if (e < kEptLarge && *std::max_element(radices.begin(), radices.end()) > 25) continue;
            if (threads_for(n, radices, e) > max_block(prec)) continue;
        } else if (!choose_radices(n, e, radices, arith_f64(prec) ? 16 : 25, max_block(prec))) {
            continue;
        }
        ept = e;
        break;
    }
    if (!ept) return false;
    const int tpt = threads_for(n, radices, ept);
    const int stride = lds_padded(n);
    const bool uses_lds = uses_lds_always || radices.size() > 1;
    const size_t slot_lds = uses_lds ? (size_t)stride * eb : 0;
    if (slot_lds > lds_budget()) return false;
    int tpw = std::max(1, 256 / tpt);
    if (h && h->tpw) tpw = h->tpw;
    if (tpw * tpt > max_block(prec)) tpw = std::max(1, max_block(prec) / tpt);
    if (tpt > max_block(prec)) return false;
    if (slot_lds) tpw = std::min<int>(tpw, (int)(lds_budget() / slot_lds));
    if (tpw < 1) return false;
    pp = PassPlan{};
    fill_stages(pp.pd, n, radices);
    pp.pd.tpt = tpt;
    pp.pd.tpw = tpw;
    pp.pd.lds_stride = stride;
    pp.ept = ept;
    pp.block = (unsigned)(tpt * tpw);
    pp.lds_bytes = slot_lds * tpw;
    return true;
}

// Column-pass layout: `cols` adjacent slots per block along the contiguous
// dimension (coalesced column access); cols is capped by LDS and 1024 threads.
static bool layout_columns(int n, int prec, int max_cols, const PlanHints *h, PassPlan &pp) {
    const size_t eb = elem_bytes(prec);
    std::vector<int> radices;
    int ept = 0;
    for (int e : {kEptSmall, kEptLarge}) {
        if (arith_f64(prec) && e > kEptSmall) continue;
        if (h && h->ept && h->ept != e) continue;
        if (choose_radices(n, e, radices, arith_f64(prec) ? 16 : 25, max_block(prec))) {
            ept = e;
            break;
        }
    }
    if (!ept) return false;
    const int tpt = threads_for(n, radices, ept);
    const int stride = lds_padded(n);
    const size_t slot_lds = (size_t)stride * eb;
    int cols = std::min(max_cols, 16);
    cols = std::min<int>(cols, (int)(lds_budget() / slot_lds));
    cols = std::min(cols, max_block(prec) / tpt);
    if (cols < 1) return false;
    cols = floor_pow2(cols);
    pp = PassPlan{};
    fill_stages(pp.pd, n, radices);
    pp.pd.tpt = tpt;
    pp.pd.tpw = cols;
    pp.pd.lds_stride = stride;
    pp.ept = ept;
    pp.block = (unsigned)(tpt * cols);
    pp.lds_bytes = slot_lds * cols;
    return true;
}

// A compiled column kernel for a pass of column length n (first pass
// kModeColRow, second kModeColCol): the pass takes its radices, threads per
// column and columns per block, so the runtime-planned kernel (used by uneven
// batches and strided calls) computes the same bits.
static void use_fixed_columns(PassPlan &pp, size_t n, int prec, int mode) {
    const FixedKernel *fk = find_fixed((int)n, prec, 2, -1);
    if (!fk) return;
    // FP64 compiled columns measured faster than the runtime-planned pass
    // only with 512 or more columns per transform (2^18 = 512 x 512: 1.81
    // -> 2.26 TB/s; 2^16 = 128 x 512: 2.47 runtime-planned, 2.13 compiled)
    if (arith_f64(prec) && pp.slots_per_transform < 512) return;
    std::vector<int> rs(fk->radices, fk->radices + fk->nst);
    int ept = 0;
    for (int e : {kEptSmall, kEptLarge})
        if (!ept && threads_for((int)n, rs, e) <= fk->tpt) ept = e;
    if (!ept || fk->tpt * fk->tpw > max_block(prec)) return;
    const size_t slot_lds = pp.lds_bytes / (size_t)pp.pd.tpw;
    PassDesc pd = pp.pd;
    fill_stages(pd, (int)n, rs);
    pd.tpt = fk->tpt;
    pd.tpw = fk->tpw;
    pp.pd = pd;
    pp.ept = ept;
    pp.block = (unsigned)(fk->tpt * fk->tpw);
    pp.lds_bytes = slot_lds * (size_t)fk->tpw;
    pp.fixed = fk;
    pp.fixed_grid[mode] = (unsigned)(fk->blocks_per_cu(mode) * device_info().cu_count);
}

// 16-bit storage: run the pass' radix-16 stages on MFMA when every wave of
// the block takes whole groups of 32 butterflies (stockham.h stage16_mfma).
static void enable_mfma(PassPlan &pp) {
    const int n = pp.pd.n;
    bool has16 = false;
    for (int s = 0; s < pp.pd.nstages; ++s) has16 |= pp.pd.st[s].radix == 16;
    if (!has16 || n % 16 || pp.ept != kEptSmall) return;
    const int nf = pp.pd.tpw * (n / 16);
    const int waves = (int)(pp.block / 64);
    pp.pd.mfma = (pp.block % 64 == 0) && (nf % 32 == 0) && (nf / 32 <= waves * (pp.ept / 8));
}

Plan::Plan(size_t n, int kind, int prec, const PlanHints *h) : n_(n), kind_(kind), prec_(prec) {
    if (kind != kC2C && kind != kR2C && kind != kR2CHalf)
        throw Error(SFFT_ERR_INVALID_ARG, "unknown transform kind");
    if (prec != kF32 && prec != kF64 && prec != kF16 && prec != kBF16 && prec != kF32X)
        throw Error(SFFT_ERR_INVALID_ARG, "unknown precision");
    if (mixed_storage(prec) && kind != kC2C)
        throw Error(SFFT_ERR_UNSUPPORTED_SIZE,
                    "16-bit storage and F32X run complex transforms only (real kinds: F32 or "
                    "F64)");
    if (n == 0) throw Error(SFFT_ERR_INVALID_ARG, "size 0");
    if (kind != kC2C && (n % 2))
        throw Error(SFFT_ERR_UNSUPPORTED_SIZE, "real transforms need an even size");
    cn_ = kind == kC2C ? n : n / 2;
    if (!supported_length(cn_)) {
        if (mixed_storage(prec))
            throw Error(SFFT_ERR_UNSUPPORTED_SIZE,
                        "16-bit storage and F32X: only lengths with factors 2, 3, 5, 7, 11, 13");
        if (cn_ > kMaxBluestein)
            throw Error(SFFT_ERR_UNSUPPORTED_SIZE, "size " + std::to_string(n) + " too large");
        bs_m_ = bluestein_length(cn_);
        try {
            bs_inner_.reset(new Plan(bs_m_, kC2C, prec));
        } catch (const Error &e) {
            throw Error(e.code, "size " + std::to_string(n) + " needs a chirp-z length " +
                                    std::to_string(bs_m_) + ": " + e.what());
        }
        if (prec == kF64) {
            build_tables<double>();
            build_bluestein<double>();
        } else {
            build_tables<float>();
            build_bluestein<float>();
        }
        return;
    }

    PassPlan single;
    const bool force_two = h && h->split > 1;
    const bool force_one = h && h->split == 1;
    bool ok = false;
    // A compile-time planned kernel fixes the radices of this length.
    // (16-bit plans asking for MFMA stages keep the runtime-planned pass)
    const FixedKernel *fk = (h && (!h->radices.empty() || h->fixed_variant == -2 ||
                                   (stored_16(prec) && h->mfma == 1)))
                                ? nullptr
                                : find_fixed((int)cn_, prec, kind != kC2C ? 1 : 0,
                                             h ? h->fixed_variant : -1);
    if (h && h->fixed_variant >= 0 && !fk)
        throw Error(SFFT_ERR_INVALID_ARG, "no such compiled variant");
    if (fk && !force_two) {
        PlanHints fh;
        if (h) fh = *h;
        fh.radices.assign(fk->radices, fk->radices + fk->nst);
        ok = layout_single((int)cn_, prec, kind != kC2C, &fh, single);
        if (ok) {
            single.fixed = fk;
            const int cus = device_info().cu_count;
            if (kind == kC2C) {
                single.fixed_grid[kModeC2C] = (unsigned)(fk->blocks_per_cu(kModeC2C) * cus);
            } else {
                single.fixed_grid[kModeR2C] = (unsigned)(fk->blocks_per_cu(kModeR2C) * cus);
                single.fixed_grid[kModeC2R] = (unsigned)(fk->blocks_per_cu(kModeC2R) * cus);
            }
        }
    }
    if (!ok && !force_two && cn_ <= (1u << 20))
        ok = layout_single((int)cn_, prec, kind != kC2C, h, single);
    // A long runtime-planned single pass (no compiled kernel, one transform
    // per CU) runs slower than a two-pass split over compiled column kernels
    // (C2C F32, 256 MB batches: 16200 1.46 -> 2.23 TB/s, 16000 1.59 -> 2.10,
    // 18432 1.70 -> 2.54; below 16000 the two were mixed, 15625 0.65x).  From
    // kTwoPassFrom such lengths take a split whose two column lengths both
    // have compiled kernels, where one exists; hinted plans keep the rule
    // below (one pass where it fits).  Shorter lengths listed in the
    // measured split table (below) measured faster as two passes too.
    const bool compiled_split = ok && !single.fixed && !h && kind == kC2C && prec == kF32 &&
                                (cn_ >= kTwoPassFrom || lookup_split(kMeasuredSplits, cn_));
    size_t best_n1 = 0;
    if (!ok || compiled_split) {
        if (!ok && force_one) throw Error(SFFT_ERR_UNSUPPORTED_SIZE, "does not fit one pass");
        // Two passes: cn = N1 * N2.  Pass 1: N2 column FFTs of length N1
        // (stride N2) with W_N^{n2 k1}, written transposed as rows; pass 2:
        // N1 column FFTs of length N2 in place (stride N1).
        double best_score = 1e30;
        // column tile width: a hint's transforms_per_block caps it.  8 columns
        // (64 B rows in F32) measured faster than 16 on most lengths: smaller
        // tiles give more resident blocks per CU.
        const size_t max_cols = (h && h->tpw) ? (size_t)h->tpw : 8;
        // the measured split where the table has one (if its layout fails,
        // the scored search runs over every split)
        auto search = [&](size_t only_n1) {
            for (size_t n1 = 32; n1 * 16 <= cn_; ++n1) {
                if (cn_ % n1) continue;
                if (h && h->split > 1 && n1 != h->split) continue;
                if (only_n1 && n1 != only_n1) continue;
                const size_t n2 = cn_ / n1;
                if (compiled_split &&
                    (!find_fixed((int)n1, prec, 2, -1) || !find_fixed((int)n2, prec, 2, -1)))
                    continue;
                PassPlan p1, p2;
                if (!layout_columns((int)n1, prec, (int)std::min(n2, max_cols), h, p1)) continue;
                if (!layout_columns((int)n2, prec, (int)std::min(n1, max_cols), h, p2)) continue;
                if (p1.pd.nstages < 2) continue;
                // prefer wide column tiles (coalescing), then balance
                const double score = -std::log2((double)std::min(p1.pd.tpw, p2.pd.tpw)) * 4.0 +
                                     std::fabs(std::log2((double)n1) - std::log2((double)n2));
                if (score < best_score) {
                    best_score = score;
                    best_n1 = n1;
                }
            }
        };
        // (hints that fix neither the split nor the layout, such as a
        // variant choice, keep the table's split: a plan and its
        // runtime-planned variant compute the same factorisation)
        const bool free_split = !h || (h->split == 0 && h->radices.empty() && !h->tpw);
        size_t table_n1 = 0;
        if (free_split && prec == kF32)
            table_n1 = kind == kC2C   ? lookup_split(kMeasuredSplits, cn_)
                       : kind == kR2C ? lookup_split(kMeasuredRealSplits, cn_)
                                      : 0;
        else if (free_split && kind == kC2C)
            table_n1 = stored_16(prec)  ? lookup_split(kMeasuredSplits16, cn_)
                       : prec == kF32X ? lookup_split(kMeasuredSplitsF32X, cn_)
                       : prec == kF64  ? lookup_split(kMeasuredSplitsF64, cn_)
                                       : 0;
        search(table_n1);
        if (!best_n1 && table_n1) search(0);
        if (!best_n1 && !ok)
            throw Error(SFFT_ERR_UNSUPPORTED_SIZE,
                        "size " + std::to_string(n) + ": no two-pass split fits LDS (F64 plans "
                        "reach 2^26, F32 2^28)");
    }
    if (!best_n1) {
        passes_.push_back(single);
    } else {
        const size_t max_cols = (h && h->tpw) ? (size_t)h->tpw : 8;
        const size_t n1 = best_n1, n2 = cn_ / n1;
        PassPlan p1, p2;
        layout_columns((int)n1, prec, (int)std::min(n2, max_cols), h, p1);
        layout_columns((int)n2, prec, (int)std::min(n1, max_cols), h, p2);
        p1.slots_per_transform = (int)n2;
        p1.in_slot_stride = 1;
        p1.in_elem_stride = (int64_t)n2;
        p1.out_slot_stride = (int64_t)n1;
        p1.out_elem_stride = 1;
        p1.pd.in_col = 1;
        p1.pd.out_col = 0;
        p1.post_tw = 1;
        p2.slots_per_transform = (int)n1;
        p2.in_from_out = 1;
        p2.in_slot_stride = 1;
        p2.in_elem_stride = (int64_t)n1;
        p2.out_slot_stride = 1;
        p2.out_elem_stride = (int64_t)n1;
        p2.pd.in_col = 1;
        p2.pd.out_col = 1;
        // compiled column kernels (plans without layout hints): the pass takes
        // the kernel's radices and tile; it runs where the pass is a plain
        // column pass for the direction (see run)
        // (16-bit plans asking for MFMA stages keep the runtime-planned pass)
        if (!(h && (h->tpw || h->fixed_variant == -2 || (stored_16(prec) && h->mfma == 1)))) {
            use_fixed_columns(p1, n1, prec, kModeColRow);
            use_fixed_columns(p2, n2, prec, kModeColCol);
        }
        passes_.push_back(p1);
        passes_.push_back(p2);
    }
    // 16-bit storage: radix-16 stages on MFMA when asked for (measured slower
    // than float butterflies here: the operand splits and the larger register
    // footprint outweigh the matrix-core arithmetic in an HBM-bound pass)
    if (stored_16(prec) && h && h->mfma == 1) {
        for (auto &pp : passes_) enable_mfma(pp);
    }
    if (arith_f64(prec))
        build_tables<double>();
    else
        build_tables<float>();
}

int Plan::mfma_stages() const {
    int c = 0;
    for (const auto &pp : passes_)
        if (pp.pd.mfma)
            for (int s = 0; s < pp.pd.nstages; ++s) c += pp.pd.st[s].radix == 16;
    return c;
}

template <typename T>
void Plan::build_bluestein() {
    const size_t eb = sizeof(C2<T>);
    void *tmp = nullptr;
    auto alloc = [&](void **p, size_t bytes) {
        if (hipMalloc(p, bytes) != hipSuccess)
            throw Error(SFFT_ERR_OUT_OF_MEMORY, "device allocation of plan tables failed");
    };
    alloc(&d_chirp_, cn_ * eb);
    alloc(&d_bhat_, bs_m_ * eb);
    table_bytes_ += (cn_ + bs_m_) * eb;
    alloc(&tmp, bs_m_ * eb);
    struct Free {
        void *p;
        ~Free() { (void)hipFree(p); }
    } free_tmp{tmp};
    check_hip(bs_gen_chirp<T>((C2<T> *)d_chirp_, (C2<T> *)tmp, cn_, bs_m_, nullptr),
              "chirp generation");
    bs_inner_->execute(tmp, bs_m_, d_bhat_, bs_m_, 1, 0, nullptr);
    check_hip(bs_scale<T>((C2<T> *)d_bhat_, bs_m_, 1.0 / (double)bs_m_, nullptr), "chirp scale");
    check_hip(hipDeviceSynchronize(), "chirp tables");
}

template <typename T>
void Plan::run_bluestein(const void *in, int64_t din, void *out, int64_t dout, size_t batch,
                         int inverse, hipStream_t stream, void *ws) const {
    const size_t row = bs_m_ * sizeof(C2<T>);
    C2<T> *a = (C2<T> *)ws, *c = (C2<T> *)((char *)ws + row * batch);
    const bool real = kind_ != kC2C;
    const int mode = !real ? kModeC2C : (inverse ? kModeC2RLoad : kModeR2CLoad);
    const int hs = kind_ == kR2CHalf;
    check_hip(bs_pre<T>((const C2<T> *)in, din, a, bs_m_, cn_, batch, (const C2<T> *)d_chirp_,
                        mode, !real && inverse, hs, (const C2<T> *)d_real_tw_,
                        (const C2<T> *)d_rot_, stream),
              "chirp-z pre-multiply");
    bs_inner_->execute(a, bs_m_, c, bs_m_, batch, 0, stream);
    check_hip(bs_mul<T>(c, (const C2<T> *)d_bhat_, bs_m_, batch, stream), "chirp-z multiply");
    bs_inner_->execute(c, bs_m_, a, bs_m_, batch, 1, stream);
    check_hip(bs_post<T>(a, bs_m_, (C2<T> *)out, dout, cn_, batch, (const C2<T> *)d_chirp_,
                         inverse, (real && inverse && hs) ? (const C2<T> *)d_rot_ : nullptr,
                         stream),
              "chirp-z post-multiply");
    if (real && !inverse)
        check_hip(launch_real_post<T>((C2<T> *)out, dout, (int)cn_, batch,
                                      (const C2<T> *)d_real_tw_, hs, stream),
                  "real post-process launch");
}

Plan::~Plan() {
    if (d_chirp_) (void)hipFree(d_chirp_);
    if (d_bhat_) (void)hipFree(d_bhat_);
    if (d_tw_) (void)hipFree(d_tw_);
    if (d_real_tw_) (void)hipFree(d_real_tw_);
    if (d_rot_) (void)hipFree(d_rot_);
    if (d_ptw_fine_) (void)hipFree(d_ptw_fine_);
    if (d_ptw_coarse_) (void)hipFree(d_ptw_coarse_);
    if (d_ptw_full_) (void)hipFree(d_ptw_full_);
}

template <typename T>
void Plan::build_tables() {
    // stage twiddle offsets
    size_t total = 0;
    for (auto &p : passes_) {
        for (int s = 0; s < p.pd.nstages; ++s) {
            StageDesc &st = p.pd.st[s];
            if (st.ns > 1) {
                st.tw_off = (int32_t)total;
                total += (size_t)st.ns * tw_group(st.radix);
            }
        }
    }
    tw_count_ = total;
    hipStream_t s;
    check_hip(hipStreamCreateWithFlags(&s, hipStreamNonBlocking), "hipStreamCreate");
    struct StreamGuard {
        hipStream_t s;
        ~StreamGuard() { (void)hipStreamDestroy(s); }
    } guard{s};
    auto alloc = [&](void **p, size_t bytes) {
        if (hipMalloc(p, bytes) != hipSuccess)
            throw Error(SFFT_ERR_OUT_OF_MEMORY, "device allocation of plan tables failed");
        table_bytes_ += bytes;
    };
    if (total) {
        alloc(&d_tw_, total * sizeof(C2<T>));
        for (auto &p : passes_)
            for (int st = 0; st < p.pd.nstages; ++st)
                if (p.pd.st[st].ns > 1)
                    check_hip(gen_stage_twiddles<T>((C2<T> *)d_tw_, p.pd.st[st].tw_off,
                                                    p.pd.st[st].ns, p.pd.st[st].radix, s),
                              "twiddle generation");
    }
    if (kind_ != kC2C) {
        const size_t nreal = n_;
        alloc(&d_real_tw_, (nreal / 4 + 1) * sizeof(C2<T>));
        if (kind_ == kR2CHalf) alloc(&d_rot_, (nreal / 2) * sizeof(C2<T>));
        check_hip(gen_real_twiddles<T>((C2<T> *)d_real_tw_, (C2<T> *)d_rot_, nreal,
                                       kind_ == kR2CHalf, s),
                  "real twiddle generation");
    }
    if (passes_.size() > 1 && cn_ * sizeof(C2<T>) <= kFullPostTwBytes) {
        // one entry per element: the first pass reads its factors with the
        // same lane pattern as its stores, instead of two gathers each
        alloc(&d_ptw_full_, cn_ * sizeof(C2<T>));
        check_hip(gen_post_twiddles_full<T>((C2<T> *)d_ptw_full_, (size_t)passes_[0].pd.n, cn_, s),
                  "post twiddle generation");
    } else if (passes_.size() > 1) {
        int shift = 0;
        while ((1ull << (2 * shift)) < cn_) ++shift;
        ptw_shift_ = shift;
        const size_t nf = 1ull << shift, nc = (cn_ + nf - 1) >> shift;
        alloc(&d_ptw_fine_, nf * sizeof(C2<T>));
        alloc(&d_ptw_coarse_, nc * sizeof(C2<T>));
        check_hip(gen_post_twiddles<T>((C2<T> *)d_ptw_fine_, (C2<T> *)d_ptw_coarse_, shift, cn_, s),
                  "post twiddle generation");
    }
    check_hip(hipStreamSynchronize(s), "plan table sync");
}

size_t Plan::default_in_dist(int inverse) const {
    if (kind_ == kC2C) return n_;
    return inverse ? cn_ : n_;
}
size_t Plan::default_out_dist(int inverse) const {
    if (kind_ == kC2C) return n_;
    return inverse ? n_ : cn_;
}

size_t Plan::in_bytes(size_t batch, int inverse) const {
    const size_t rb = real_bytes(prec_);
    if (kind_ == kC2C || inverse) return batch * cn_ * 2 * rb;
    return batch * n_ * rb;
}
size_t Plan::out_bytes(size_t batch, int inverse) const {
    const size_t rb = real_bytes(prec_);
    if (kind_ == kC2C || !inverse) return batch * cn_ * 2 * rb;
    return batch * n_ * rb;
}

template <typename T>
PassArgs<T> Plan::pass_args(size_t p, int inverse) const {
    const PassPlan &pp = passes_[p];
    const size_t np = passes_.size();
    const int64_t din = (int64_t)default_in_dist(inverse) / ((kind_ != kC2C && !inverse) ? 2 : 1);
    const int64_t dout = (int64_t)default_out_dist(inverse) / ((kind_ != kC2C && inverse) ? 2 : 1);
    PassArgs<T> a{};
    a.pd = pp.pd;
    const FastDiv pb = make_fastdiv((uint32_t)pp.slots_per_transform);
    a.iin = IoDesc{pp.in_from_out ? dout : din, pp.in_slot_stride, pp.in_elem_stride, pb};
    a.iout = IoDesc{dout, pp.out_slot_stride, pp.out_elem_stride, pb};
    a.tw = (const C2<T> *)d_tw_;
    a.conj_in = inverse && p == 0;
    a.conj_out = inverse && p + 1 == np;
    if (kind_ == kC2C) {
        a.mode = kModeC2C;
    } else if (np == 1) {
        a.mode = inverse ? kModeC2R : kModeR2C;
    } else if (p == 0) {
        // two-pass real: pre-process / rotation fused into the first pass' loads
        a.mode = inverse ? kModeC2RLoad : kModeR2CLoad;
        if (inverse) a.conj_in = 0;  // the fused pre-process stores conj(.) itself
    } else {
        a.mode = inverse ? kModeC2RStore : kModeC2C;
    }
    a.real_h = (int32_t)cn_;
    a.half_shift = kind_ == kR2CHalf;
    a.post_tw = pp.post_tw;
    a.post_tw_shift = ptw_shift_;
    a.real_tw = (const C2<T> *)d_real_tw_;
    a.rot = (const C2<T> *)d_rot_;
    a.ptw_fine = (const C2<T> *)d_ptw_fine_;
    a.ptw_coarse = (const C2<T> *)d_ptw_coarse_;
    a.ptw_full = (const C2<T> *)d_ptw_full_;
    a.in_from_out = pp.in_from_out;
    return a;
}
template PassArgs<float> Plan::pass_args<float>(size_t, int) const;
template PassArgs<double> Plan::pass_args<double>(size_t, int) const;

size_t Plan::slot_lds_bytes(size_t p) const {
    const PassPlan &pp = passes_[p];
    return pp.lds_bytes / (size_t)pp.pd.tpw;
}

template <typename T>
void Plan::run(const void *in, size_t in_dist, void *out, size_t out_dist, size_t batch,
               int inverse, hipStream_t stream, void *ws) const {
    if (passes_.size() == 1 && passes_[0].fixed && in_dist == default_in_dist(inverse) && out_dist == default_out_dist(inverse)) {
        run_fixed<T>(in, out, batch, inverse, stream);
        return;
    }
    // distances in complex elements
    int64_t din = (int64_t)in_dist, dout = (int64_t)out_dist;
    if (kind_ != kC2C) {
        if (inverse) {
            if (out_dist % 2) throw Error(SFFT_ERR_INVALID_ARG, "real output distance must be even");
            dout = (int64_t)(out_dist / 2);
        } else {
            if (in_dist % 2) throw Error(SFFT_ERR_INVALID_ARG, "real input distance must be even");
            din = (int64_t)(in_dist / 2);
        }
    }
    if (bs_inner_) {
        run_bluestein<T>(in, din, out, dout, batch, inverse, stream, ws);
        return;
    }
    run_passes<T>(in, din, out, dout, batch, inverse, stream, ws);
}

// The passes of a direct plan over one batch (distances in complex values).
template <typename T>
void Plan::run_passes(const void *in, int64_t din, void *out, int64_t dout, size_t batch,
                      int inverse, hipStream_t stream, void *ws) const {
    const size_t np = passes_.size();
    // In place on a two-pass length the first pass writes its intermediate to
    // the workspace (cn values per transform) and the second pass reads it
    // there; out of place the output buffer holds the intermediate.
    const int64_t dws = (int64_t)cn_;
    for (size_t p = 0; p < np; ++p) {
        const PassPlan &pp = passes_[p];
        const bool last = p + 1 == np;
        const void *src = p == 0 ? in : (ws ? ws : out);
        const int64_t dsrc = p == 0 ? din : (ws ? dws : dout);
        void *dst = (!last && ws) ? ws : out;
        const int64_t ddst = (!last && ws) ? dws : dout;
        const size_t slots = batch * (size_t)pp.slots_per_transform;
        if (slots > 0x7fffffffull) throw Error(SFFT_ERR_INVALID_ARG, "batch too large");
        if (pp.fixed && pp.fixed->real == 2) {
            // compiled column kernel where the pass is a plain column pass:
            // complex passes, and real passes with no pre-process or rotation
            // fused into their loads / stores for this direction
            const PassArgs<T> m = pass_args<T>(p, inverse);
            const bool plain = m.mode == kModeC2C ||
                               ((m.mode == kModeR2CLoad || m.mode == kModeC2RStore) && !m.half_shift);
            if (plain) {
                run_fixed_columns<T>(p, src, dsrc, dst, ddst, batch, inverse, stream);
                continue;
            }
        }
        PassArgs<T> a = pass_args<T>(p, inverse);
        a.iin.batch_stride = dsrc;
        a.iout.batch_stride = ddst;
        a.in = (const C2<T> *)src;
        a.out = (C2<T> *)dst;
        a.total_slots = (int32_t)slots;
        const unsigned grid = (unsigned)((slots + pp.pd.tpw - 1) / pp.pd.tpw);
        if constexpr (std::is_same<T, float>::value) {
            if (stored_16(prec_)) {
                check_hip(launch_pass_16(a, prec_ == kF16 ? kStorageF16 : kStorageBF16, pp.ept,
                                         grid, pp.block, pp.lds_bytes, stream),
                          "pass launch");
                continue;
            }
        }
        if constexpr (std::is_same<T, double>::value) {
            if (prec_ == kF32X) {
                check_hip(launch_pass_f32x(a, grid, pp.block, pp.lds_bytes, stream), "pass launch");
                continue;
            }
        }
        check_hip(launch_pass<T>(a, pp.ept, grid, pp.block, pp.lds_bytes, stream), "pass launch");
    }
    // two-pass forward real: pairwise post-process in place on the output
    if (kind_ != kC2C && !inverse && np > 1)
        check_hip(launch_real_post<T>((C2<T> *)out, dout, (int)cn_, batch,
                                      (const C2<T> *)d_real_tw_, kind_ == kR2CHalf, stream),
                  "real post-process launch");
}

// Compile-time planned kernel over a contiguous batch.  It addresses HBM
// with 32-bit byte offsets from the launch base, so the batch is split into
// launches spanning less than 4 GiB.
template <typename T>
void Plan::run_fixed(const void *in, void *out, size_t batch, int inverse,
                     hipStream_t stream) const {
    const PassPlan &pp = passes_[0];
    const FixedKernel &fk = *pp.fixed;
    const int mode = kind_ == kC2C ? kModeC2C : (inverse ? kModeC2R : kModeR2C);
    const size_t eb = 2 * real_bytes(prec_);  // bytes per complex value in HBM
    const size_t per = cn_ * eb;  // bytes per transform (input and output)
    const size_t chunk = std::max<size_t>(1, ((1ull << 32) - 1) / per);
    const uint32_t flip = inverse ? 0x80000000u : 0u;
    for (size_t c0 = 0; c0 < batch; c0 += chunk) {
        const size_t cb = std::min(chunk, batch - c0);
        FixedArgs<T> a{};
        a.in = (const C2<T> *)((const char *)in + c0 * per);
        a.out = (C2<T> *)((char *)out + c0 * per);
        a.tw = (const C2<T> *)d_tw_;
        a.real_tw = (const C2<T> *)d_real_tw_;
        a.rot = (const C2<T> *)d_rot_;
        a.total_slots = (int32_t)cb;
        a.flip_in = mode == kModeC2C ? flip : 0u;
        a.flip_out = flip;
        a.half_shift = kind_ == kR2CHalf;
        const size_t groups = (cb + fk.tpw - 1) / fk.tpw;
        const unsigned grid = (unsigned)std::min<size_t>(groups, pp.fixed_grid[mode]);
        check_hip(fk.launch(&a, mode, grid, stream), "fixed-plan launch");
    }
}
template void Plan::run_fixed<float>(const void *, void *, size_t, int, hipStream_t) const;

// Pass p of a two-pass complex plan through its compiled column kernel.  The
// kernel addresses HBM with 32-bit byte offsets from the launch base, so the
// batch is split into launches spanning less than 4 GiB.
template <typename T>
void Plan::run_fixed_columns(size_t p, const void *src, int64_t dsrc, void *dst, int64_t ddst,
                             size_t batch, int inverse, hipStream_t stream) const {
    const PassPlan &pp = passes_[p];
    const FixedKernel &fk = *pp.fixed;
    const int mode = p == 0 ? kModeColRow : kModeColCol;
    const size_t eb = 2 * real_bytes(prec_);  // bytes per complex value in HBM
    const int64_t dmax = std::max(dsrc, ddst);
    const size_t chunk = std::max<size_t>(1, ((1ull << 32) - 1) / ((size_t)dmax * eb));
    if ((size_t)pp.slots_per_transform * std::min(chunk, batch) > 0x7fffffffull)
        throw Error(SFFT_ERR_INVALID_ARG, "batch too large");
    const uint32_t flip = 0x80000000u;
    const PassArgs<T> conj = pass_args<T>(p, inverse);  // the pass' conjugations
    // first stage with twiddles (stage 1): the pass' part of the dense table
    const int32_t tw0 = pp.pd.nstages > 1 ? pp.pd.st[1].tw_off : 0;
    for (size_t c0 = 0; c0 < batch; c0 += chunk) {
        const size_t cb = std::min(chunk, batch - c0);
        FixedArgs<T> a{};
        a.in = (const C2<T> *)((const char *)src + (int64_t)c0 * dsrc * (int64_t)eb);
        a.out = (C2<T> *)((char *)dst + (int64_t)c0 * ddst * (int64_t)eb);
        a.tw = (const C2<T> *)d_tw_ + tw0;
        a.total_slots = (int32_t)(cb * (size_t)pp.slots_per_transform);
        a.flip_in = conj.conj_in ? flip : 0u;
        a.flip_out = conj.conj_out ? flip : 0u;
        a.spt = make_fastdiv((uint32_t)pp.slots_per_transform);
        a.in_bs = (int32_t)dsrc;
        a.in_ss = (int32_t)pp.in_slot_stride;
        a.in_es = (int32_t)pp.in_elem_stride;
        a.out_bs = (int32_t)ddst;
        a.out_ss = (int32_t)pp.out_slot_stride;
        a.out_es = (int32_t)pp.out_elem_stride;
        a.post_tw = pp.post_tw;
        a.ptw_shift = ptw_shift_;
        a.ptw_fine = (const C2<T> *)d_ptw_fine_;
        a.ptw_coarse = (const C2<T> *)d_ptw_coarse_;
        a.ptw_full = (const C2<T> *)d_ptw_full_;
        const size_t groups = ((size_t)a.total_slots + fk.tpw - 1) / fk.tpw;
        const unsigned grid = (unsigned)std::min<size_t>(groups, pp.fixed_grid[mode]);
        check_hip(fk.launch(&a, mode, grid, stream), "column-pass launch");
    }
}
template void Plan::run_fixed_columns<float>(size_t, const void *, int64_t, void *, int64_t,
                                             size_t, int, hipStream_t) const;
template void Plan::run_fixed_columns<double>(size_t, const void *, int64_t, void *, int64_t,
                                              size_t, int, hipStream_t) const;
template void Plan::run_fixed<double>(const void *, void *, size_t, int, hipStream_t) const;

size_t Plan::workspace_bytes(size_t batch, bool in_place) const {
    const size_t cb = 2 * real_bytes(prec_);  // one complex value in HBM
    if (bs_inner_) return 2 * bs_m_ * cb * batch;
    if (in_place && passes_.size() > 1) return cn_ * cb * batch;
    return 0;
}

void Plan::execute(const void *in, size_t in_dist, void *out, size_t out_dist, size_t batch,
                   int inverse, hipStream_t stream, void *ws, size_t ws_bytes) const {
    if (batch == 0) return;
    if (!in || !out) throw Error(SFFT_ERR_INVALID_ARG, "null buffer");
    // overlap check on the byte ranges actually touched
    const size_t rb = real_bytes(prec_);
    const size_t in_unit = (kind_ == kC2C || inverse) ? 2 * rb : rb;
    const size_t out_unit = (kind_ == kC2C || !inverse) ? 2 * rb : rb;
    const size_t in_len = ((batch - 1) * in_dist + default_in_dist(inverse)) * in_unit;
    const size_t out_len = ((batch - 1) * out_dist + default_out_dist(inverse)) * out_unit;
    const uintptr_t i0 = (uintptr_t)in, o0 = (uintptr_t)out;
    // in place, same layout: one-pass plans read every transform whole before
    // writing it (one workgroup per slot, reads before the first barrier),
    // chirp-z plans consume the input into their workspace, and two-pass
    // plans keep their intermediate in a workspace
    const bool in_place = i0 == o0 && in_dist * in_unit == out_dist * out_unit;
    if (!in_place && i0 < o0 + out_len && o0 < i0 + in_len)
        throw Error(SFFT_ERR_ALIASING, "input and output overlap other than exactly in place");
    if (in_dist < default_in_dist(inverse) || out_dist < default_out_dist(inverse))
        throw Error(SFFT_ERR_INVALID_ARG, "batch distance shorter than one transform");
    if (bs_inner_ && batch > (0x7fffffffull / bs_m_))
        throw Error(SFFT_ERR_INVALID_ARG, "batch too large");
    const size_t need = workspace_bytes(batch, in_place);
    struct Owned {
        void *p = nullptr;
        hipStream_t s = nullptr;
        ~Owned() {
            if (p) (void)hipFreeAsync(p, s);
        }
    } owned;
    if (need && ws) {
        if (ws_bytes < need)
            throw Error(SFFT_ERR_INVALID_ARG, "workspace of " + std::to_string(ws_bytes) +
                                                  " bytes, the call needs " + std::to_string(need));
        const uintptr_t w0 = (uintptr_t)ws;
        if ((w0 < i0 + in_len && i0 < w0 + need) || (w0 < o0 + out_len && o0 < w0 + need))
            throw Error(SFFT_ERR_ALIASING, "workspace overlaps the input or the output");
    } else if (need) {
        if (hipMallocAsync(&owned.p, need, stream) != hipSuccess) {
            (void)hipGetLastError();
            owned.p = nullptr;
            throw Error(SFFT_ERR_OUT_OF_MEMORY, "workspace allocation failed");
        }
        owned.s = stream;
        ws = owned.p;
    } else {
        ws = nullptr;
    }
    if (arith_f64(prec_))
        run<double>(in, in_dist, out, out_dist, batch, inverse, stream, ws);
    else
        run<float>(in, in_dist, out, out_dist, batch, inverse, stream, ws);
}

static std::mutex g_cache_mu;
static std::map<std::tuple<size_t, int, int>, std::shared_ptr<Plan>> g_cache;

std::shared_ptr<Plan> get_plan(size_t n, int kind, int prec) {
    // tuned choices named by SFFT_TUNE_FILE, read once (before the cache lock:
    // loading drops cached plans)
    static std::once_flag tune_file_once;
    std::call_once(tune_file_once, [] {
        const char *f = std::getenv("SFFT_TUNE_FILE");
        if (!f || !*f) return;
        try {
            (void)tune_load(f, true);
        } catch (const Error &) {
            // an unreadable tuning file leaves the planner's defaults
        }
    });
    std::lock_guard<std::mutex> lock(g_cache_mu);
    auto key = std::make_tuple(n, kind, prec);
    auto it = g_cache.find(key);
    if (it != g_cache.end()) return it->second;
    PlanHints h;
    std::shared_ptr<Plan> p;
    if (tuned_hints(n, kind, prec, h)) {
        try {
            p = std::make_shared<Plan>(n, kind, prec, &h);
        } catch (const Error &) {
            p = nullptr;  // a stale tuned entry never blocks the default plan
        }
    }
    if (!p) p = std::make_shared<Plan>(n, kind, prec);
    g_cache[key] = p;
    return p;
}

void drop_cached_plan(size_t n, int kind, int prec) {
    std::lock_guard<std::mutex> lock(g_cache_mu);
    g_cache.erase(std::make_tuple(n, kind, prec));
}

// Every radix multiset (non-increasing order) for length n within the
// thread budget, with its cost.
static void collect_radices(int rem, int max_r, std::vector<int> &cur, int n, int ept,
                            int max_threads, std::vector<std::pair<double, std::vector<int>>> &out) {
    if (rem == 1) {
        const int tpt = threads_for(n, cur, ept);
        if (tpt <= max_threads) out.push_back({radix_cost(n, cur, ept, tpt), cur});
        return;
    }
    if ((int)cur.size() >= kMaxStages || out.size() > 4096) return;
    for (int r : kRadixSet) {
        if (r > max_r || r < 2 || rem % r) continue;
        cur.push_back(r);
        collect_radices(rem / r, r, cur, n, ept, max_threads, out);
        cur.pop_back();
    }
}

std::vector<PlanHints> plan_candidates(size_t n, int kind, int prec) {
    std::vector<PlanHints> out;
    out.push_back(PlanHints{});  // the planner's choice
    const size_t cn = kind == kC2C ? n : n / 2;
    if (!supported_length(cn) || cn < 2) return out;
    const int real = kind != kC2C;
    const int nv = fixed_variant_count((int)cn, prec, real);
    for (int v = 0; v < nv; ++v) {
        PlanHints h;
        h.fixed_variant = v;
        out.push_back(h);
    }
    if (stored_16(prec)) {
        PlanHints h;
        h.mfma = 1;
        out.push_back(h);
    }
    const int maxr = arith_f64(prec) ? 16 : 25;
    if (cn <= (1u << 20)) {
        for (int ept : {kEptSmall, kEptLarge}) {
            if (arith_f64(prec) && ept > kEptSmall) continue;
            std::vector<std::pair<double, std::vector<int>>> all;
            std::vector<int> cur;
            collect_radices((int)cn, maxr, cur, (int)cn, ept, max_block(prec), all);
            std::stable_sort(all.begin(), all.end(),
                             [](const auto &a, const auto &b) { return a.first < b.first; });
            const size_t keep = std::min<size_t>(all.size(), 3);
            for (size_t i = 0; i < keep; ++i) {
                std::vector<int> desc = all[i].second, asc(desc.rbegin(), desc.rend());
                for (const auto &order : {desc, asc}) {
                    PlanHints h;
                    h.radices = order;
                    h.ept = ept;
                    out.push_back(h);
                    if (order == asc && asc == desc) break;
                }
                if (i == 0) {  // tile widths around the default for the best multiset
                    const int tpt = threads_for((int)cn, desc, ept);
                    const int def = std::max(1, 256 / tpt);
                    for (int tpw : {def / 2, def * 2}) {
                        if (tpw < 1 || tpw * tpt > max_block(prec) || tpw == def) continue;
                        PlanHints h;
                        h.radices = desc;
                        h.ept = ept;
                        h.tpw = tpw;
                        out.push_back(h);
                    }
                }
            }
        }
    }
    // two-pass splits (and one pass where both exist)
    std::vector<std::pair<double, size_t>> splits;
    for (size_t n1 = 32; n1 * 16 <= cn; ++n1) {
        if (cn % n1) continue;
        const size_t n2 = cn / n1;
        PassPlan p1, p2;
        if (!layout_columns((int)n1, prec, (int)std::min<size_t>(n2, 16), nullptr, p1)) continue;
        if (!layout_columns((int)n2, prec, (int)std::min<size_t>(n1, 16), nullptr, p2)) continue;
        if (p1.pd.nstages < 2) continue;
        const double score = -std::log2((double)std::min(p1.pd.tpw, p2.pd.tpw)) * 4.0 +
                             std::fabs(std::log2((double)n1) - std::log2((double)n2));
        splits.push_back({score, n1});
    }
    std::stable_sort(splits.begin(), splits.end());
    for (size_t i = 0; i < std::min<size_t>(splits.size(), 3); ++i) {
        // column tile width (0: the planner's, 8); 32 elements per thread
        // measured slower on every column pass, so it is not offered
        for (int cols : {0, 16, 4}) {
            PlanHints h;
            h.split = splits[i].second;
            h.tpw = cols;
            out.push_back(h);
        }
    }
    return out;
}

}  // namespace sfft
