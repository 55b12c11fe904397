// engine.h -- host side of the transform engine: planner, device tables,
// plan cache and pass execution.  The C ABI (capi.cpp) is a thin layer on top.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "fixed.h"
#include "kernels.h"
#include "plan_desc.h"

namespace sfft {

enum Kind : int { kC2C = 0, kR2C = 1, kR2CHalf = 2 };
// kF16 / kBF16: float arithmetic, 16-bit HBM storage (complex-to-complex only).
// kF32X: FP32 HBM storage (the FFT<float> layout), FP64 twiddles, LDS and
// arithmetic (complex-to-complex only).
enum Prec : int { kF32 = 0, kF64 = 1, kF16 = 2, kBF16 = 3, kF32X = 4 };

inline bool stored_16(int prec) { return prec == kF16 || prec == kBF16; }
// The arithmetic (registers, LDS, twiddles) is FP64.
inline bool arith_f64(int prec) { return prec == kF64 || prec == kF32X; }
// HBM storage differs from the arithmetic type.
inline bool mixed_storage(int prec) { return stored_16(prec) || prec == kF32X; }

// Bytes of one real component in HBM.
inline size_t real_bytes(int prec) { return prec == kF64 ? 8 : (stored_16(prec) ? 2 : 4); }

// Error with an ABI status code.
struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

struct DeviceInfo {
    int device = 0;
    int cu_count = 256;
    size_t lds_bytes = 160 * 1024;
    std::string arch;
};
const DeviceInfo &device_info();

// Integer-exact size helpers (reference signalsmith-fft.h:317-348, :403-408).
size_t size_minimum(size_t n);
size_t size_maximum(size_t n);
size_t real_size_minimum(size_t n);
size_t real_size_maximum(size_t n);

// One launch over `slots` sub-transforms.
struct PassPlan {
    PassDesc pd{};
    int ept = kEptSmall;
    int slots_per_transform = 1;  // per_batch of the slot -> address map
    int in_from_out = 0;          // reads the output buffer (second pass)
    int post_tw = 0;
    // address map; batch strides are filled at execute time
    int64_t in_slot_stride = 0, in_elem_stride = 1;
    int64_t out_slot_stride = 0, out_elem_stride = 1;
    unsigned block = 0;
    size_t lds_bytes = 0;
    // compile-time planned kernel with the same radices (contiguous batches)
    const FixedKernel *fixed = nullptr;
    unsigned fixed_grid[8] = {};  // resident blocks per mode (kModeC2C ... kModeColCol)
};

// Overrides the planner's choice (autotuner candidates).
struct PlanHints {
    std::vector<int> radices;  // first-pass radix order (empty: planner's)
    int tpw = 0;               // slots per block (0: planner's)
    int ept = 0;               // elements-per-thread instance (0: planner's)
    size_t split = 0;          // N1 of a two-pass plan (0: planner's; 1: force single pass)
    int fixed_variant = -1;    // -1 default, -2 no fixed kernel, k >= 0 that variant
    int mfma = -1;             // 16-bit storage: 1 radix-16 stages on MFMA; -1 / 0 float butterflies
};

class Plan {
  public:
    Plan(size_t n, int kind, int prec, const PlanHints *hints = nullptr);
    ~Plan();
    Plan(const Plan &) = delete;
    Plan &operator=(const Plan &) = delete;

    size_t n() const { return n_; }
    int kind() const { return kind_; }
    int prec() const { return prec_; }
    size_t complex_len() const { return cn_; }
    const std::vector<PassPlan> &passes() const { return passes_; }
    size_t table_bytes() const { return table_bytes_; }
    int mfma_stages() const;
    // in == out (same layout) runs without a workspace: one-pass and chirp-z
    // plans read each transform whole before writing it; two-pass plans run
    // in place through a workspace for the intermediate.
    bool in_place_direct() const { return bs_inner_ != nullptr || passes_.size() == 1; }
    // Device workspace an execute call of `batch` transforms needs: chirp-z
    // rows (2 M values per transform) or, in place on a two-pass length, the
    // intermediate (cn values per transform); 0 if none.
    size_t workspace_bytes(size_t batch, bool in_place) const;
    // Chirp-z plan: the length has a prime factor the kernels do not plan;
    // the transform runs as a convolution through FFTs of length bluestein_m().
    bool bluestein() const { return bs_inner_ != nullptr; }
    size_t bluestein_m() const { return bs_m_; }
    const Plan *bluestein_inner() const { return bs_inner_.get(); }

    // in/out: device pointers; dists in elements of the in/out type
    // (complex for C2C, real/complex for the real kinds).  ws / ws_bytes: a
    // caller's device workspace of at least workspace_bytes(batch, in == out)
    // bytes; without one, a needed workspace is allocated stream-ordered
    // (hipMallocAsync) and freed after the call.
    void execute(const void *in, size_t in_dist, void *out, size_t out_dist, size_t batch,
                 int inverse, hipStream_t stream, void *ws = nullptr, size_t ws_bytes = 0) const;

    // Static launch arguments of pass p for a direction (in/out/total_slots
    // left empty; batch distances = contiguous defaults).
    template <typename T>
    PassArgs<T> pass_args(size_t p, int inverse) const;
    // LDS bytes one slot of pass p needs (0 if the pass never touches LDS).
    size_t slot_lds_bytes(size_t p) const;

    // Default distances for a contiguous batch.
    size_t default_in_dist(int inverse) const;
    size_t default_out_dist(int inverse) const;
    size_t in_bytes(size_t batch, int inverse) const;
    size_t out_bytes(size_t batch, int inverse) const;

  private:
    template <typename T>
    void build_tables();
    // ws: the workspace when one is needed (chirp-z rows, or the
    // intermediate of a two-pass plan run in place), else null
    template <typename T>
    void run(const void *in, size_t in_dist, void *out, size_t out_dist, size_t batch, int inverse,
             hipStream_t stream, void *ws) const;
    template <typename T>
    void run_passes(const void *in, int64_t din, void *out, int64_t dout, size_t batch,
                    int inverse, hipStream_t stream, void *ws) const;
    template <typename T>
    void run_fixed(const void *in, void *out, size_t batch, int inverse, hipStream_t stream) const;
    template <typename T>
    void run_fixed_columns(size_t p, const void *src, int64_t dsrc, void *dst, int64_t ddst,
                           size_t batch, int inverse, hipStream_t stream) const;
    template <typename T>
    void build_bluestein();
    template <typename T>
    void run_bluestein(const void *in, int64_t din, void *out, int64_t dout, size_t batch,
                       int inverse, hipStream_t stream, void *ws) const;

    size_t n_;
    int kind_, prec_;
    size_t cn_;  // complex FFT length (n, or n/2 for real kinds)
    std::vector<PassPlan> passes_;
    std::vector<std::vector<int>> stage_ns_;  // per pass, per stage span (for tables)
    void *d_tw_ = nullptr;
    size_t tw_count_ = 0;
    void *d_real_tw_ = nullptr;
    void *d_rot_ = nullptr;
    void *d_ptw_fine_ = nullptr;  // C2<T>, the plan's arithmetic type
    void *d_ptw_coarse_ = nullptr;
    void *d_ptw_full_ = nullptr;  // full four-step table, small plans only
    int ptw_shift_ = 0;
    size_t table_bytes_ = 0;
    std::unique_ptr<Plan> bs_inner_;
    size_t bs_m_ = 0;
    void *d_chirp_ = nullptr;
    void *d_bhat_ = nullptr;
};

// Candidate planner overrides for the autotuner: the planner's default
// first, then compiled variants, radix multisets and orders, tile widths,
// elements per thread, two-pass splits (and MFMA off for 16-bit storage).
std::vector<PlanHints> plan_candidates(size_t n, int kind, int prec);

// Tuned overrides (tune.cpp): get_plan builds with them when present.
bool tuned_hints(size_t n, int kind, int prec, PlanHints &out);
// Drops the cached plan of (n, kind, prec) so the next get_plan rebuilds it.
void drop_cached_plan(size_t n, int kind, int prec);

// Cached plans keyed by (n, kind, precision).
std::shared_ptr<Plan> get_plan(size_t n, int kind, int prec);

// Radix choice for a single pass of length n (exposed for tests/tuning).
bool choose_radices(int n, int ept, std::vector<int> &radices, int max_radix = 25,
                    int max_threads = 1024);
bool supported_length(size_t n);
// Any length a plan exists for in F32 / F64: direct, or chirp-z (up to 2^27).
bool plannable_length(size_t n);

void check_hip(hipError_t e, const char *what);

// Message of the last failed C ABI call on this thread.
std::string &last_error();

}  // namespace sfft

// The object behind the C ABI's sfft_plan handle.
struct sfft_plan_s {
    std::shared_ptr<sfft::Plan> plan;
};
