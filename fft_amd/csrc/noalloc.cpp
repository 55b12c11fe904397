// noalloc.cpp -- sfft_execute_noalloc: sfft_execute without any device
// allocation.  The workspace some calls need (chirp-z rows; the
// intermediate of a two-pass length run in place) comes from a file-scope
// device array of fixed size instead of hipMallocAsync, for callers that
// cannot allocate (code under HIP graph capture, tight serving loops).
//
// Each GPU holds its own copy of the array for the life of the process.
// Every call on a GPU shares that GPU's copy, so calls that overlap on one
// GPU (other streams, other threads) race; the array's contents are left
// from the previous call and nothing relies on them (every workspace byte a
// call reads it has written first in the same call).
#include <sfft.h>

#include <mutex>
#include <string>
#include <vector>

#include "engine.h"

namespace {

constexpr size_t kScratchBytes = SFFT_NOALLOC_SCRATCH_BYTES;
static_assert(kScratchBytes % 256 == 0, "scratch size");

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

std::mutex g_mu;
std::vector<void *> g_scratch;  // device address of `scratch`, per device

// The current device's address of the array, looked up once per device with
// that device current.  The array's name is only a host-side stand-in: host
// code passes this address, never the name.
void *scratch_on_current_device() {
    int dev = 0;
    sfft::check_hip(hipGetDevice(&dev), "hipGetDevice");
    std::lock_guard<std::mutex> lock(g_mu);
    if ((size_t)dev >= g_scratch.size()) g_scratch.resize((size_t)dev + 1, nullptr);
    if (!g_scratch[(size_t)dev]) {
        void *ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr) {
            (void)hipGetLastError();
            throw sfft::Error(SFFT_ERR_DEVICE,
                              std::string("scratch address lookup failed: ") + hipGetErrorString(e));
        }
        g_scratch[(size_t)dev] = ptr;
    }
    return g_scratch[(size_t)dev];
}

}  // namespace

extern "C" int sfft_execute_noalloc(sfft_plan plan, const void *in, void *out, size_t batch,
                                    int direction, void *stream) {
    try {
        sfft::last_error().clear();
        if (!plan) throw sfft::Error(SFFT_ERR_INVALID_ARG, "plan is NULL");
        if (direction != SFFT_FORWARD && direction != SFFT_INVERSE)
            throw sfft::Error(SFFT_ERR_INVALID_ARG,
                              "direction must be SFFT_FORWARD or SFFT_INVERSE");
        const sfft::Plan &p = *plan->plan;
        const bool in_place = in != nullptr && in == out;
        const size_t need = p.workspace_bytes(batch, in_place);
        if (need > kScratchBytes)
            throw sfft::Error(SFFT_ERR_BUDGET,
                              "sfft_execute_noalloc: this call needs " + std::to_string(need) +
                                  " bytes of workspace, the scratch array holds " +
                                  std::to_string(kScratchBytes) + " (split the batch)");
        void *ws = need ? scratch_on_current_device() : nullptr;
        p.execute(in, p.default_in_dist(direction), out, p.default_out_dist(direction), batch,
                  direction, (hipStream_t)stream, ws, need ? kScratchBytes : 0);
        return SFFT_OK;
    } catch (const sfft::Error &e) {
        sfft::last_error() = e.what();
        return e.code;
    } catch (const std::exception &e) {
        sfft::last_error() = e.what();
        return SFFT_ERR_DEVICE;
    }
}
