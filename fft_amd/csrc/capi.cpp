// capi.cpp -- the C ABI (include/sfft.h) over the engine.
#include <sfft.h>

#include <algorithm>
#include <cstring>
#include <memory>
#include <string>

#include "batch.h"
#include "engine.h"
#include "tune.h"

namespace sfft {
std::string &last_error() {
    thread_local std::string msg;
    return msg;
}
}  // namespace sfft

namespace {

int fail(int code, const std::string &msg) {
    sfft::last_error() = msg;
    return code;
}

template <typename F>
int guarded(F &&f) {
    try {
        sfft::last_error().clear();
        return f();
    } catch (const sfft::Error &e) {
        return fail(e.code, e.what());
    } catch (const std::bad_alloc &) {
        return fail(SFFT_ERR_OUT_OF_MEMORY, "host allocation failed");
    } catch (const std::exception &e) {
        return fail(SFFT_ERR_DEVICE, e.what());
    }
}

}  // namespace

struct sfft_batch_s {
    std::unique_ptr<sfft::UnevenBatch> impl;
};

extern "C" {

int sfft_init(void) {
    return guarded([] {
        (void)sfft::device_info();
        return SFFT_OK;
    });
}

const char *sfft_last_error(void) { return sfft::last_error().c_str(); }

int sfft_device_info(int *cu_count, size_t *lds_bytes, char *arch, size_t arch_len) {
    return guarded([&] {
        const sfft::DeviceInfo &d = sfft::device_info();
        if (cu_count) *cu_count = d.cu_count;
        if (lds_bytes) *lds_bytes = d.lds_bytes;
        if (arch && arch_len) {
            std::strncpy(arch, d.arch.c_str(), arch_len - 1);
            arch[arch_len - 1] = 0;
        }
        return SFFT_OK;
    });
}

size_t sfft_size_minimum(size_t n) { return sfft::size_minimum(n); }
size_t sfft_size_maximum(size_t n) { return sfft::size_maximum(n); }
size_t sfft_real_size_minimum(size_t n) { return sfft::real_size_minimum(n); }
size_t sfft_real_size_maximum(size_t n) { return sfft::real_size_maximum(n); }

int sfft_size_supported(size_t n, int kind) {
    if (kind == SFFT_C2C) return sfft::plannable_length(n) ? 1 : 0;
    if (kind == SFFT_R2C || kind == SFFT_R2C_HALFSHIFT)
        return (n >= 2 && n % 2 == 0 && sfft::plannable_length(n / 2)) ? 1 : 0;
    return 0;
}

int sfft_plan_create(sfft_plan *plan, size_t n, int kind, int precision) {
    return guarded([&] {
        if (!plan) return fail(SFFT_ERR_INVALID_ARG, "plan pointer is NULL");
        *plan = nullptr;
        auto p = sfft::get_plan(n, kind, precision);
        *plan = new sfft_plan_s{p};
        return SFFT_OK;
    });
}

static sfft::PlanHints from_abi(const sfft_plan_hints &a) {
    sfft::PlanHints h;
    h.fixed_variant = a.variant;
    for (int i = 0; i < a.nradices && i < 16; ++i) h.radices.push_back(a.radices[i]);
    h.tpw = a.transforms_per_block;
    h.split = a.split;
    h.ept = a.elements_per_thread;
    h.mfma = a.mfma;
    return h;
}

static sfft_plan_hints to_abi(const sfft::PlanHints &h) {
    sfft_plan_hints a;
    std::memset(&a, 0, sizeof(a));
    a.variant = h.fixed_variant;
    a.nradices = (int)std::min<size_t>(h.radices.size(), 16);
    for (int i = 0; i < a.nradices; ++i) a.radices[i] = h.radices[i];
    a.transforms_per_block = h.tpw;
    a.split = h.split;
    a.elements_per_thread = h.ept;
    a.mfma = h.mfma;
    return a;
}

int sfft_autotune(size_t n, int kind, int precision, size_t batch, sfft_tune_result *result) {
    return guarded([&] {
        const sfft::TuneResult t = sfft::autotune(n, kind, precision, batch, 5);
        if (result) {
            result->hints = to_abi(t.hints);
            result->ms = t.ms;
            result->gbs = t.gbs;
            result->batch = t.batch;
            result->candidates = t.candidates;
        }
        return SFFT_OK;
    });
}

int sfft_tune_save(const char *path) {
    return guarded([&] {
        if (!path) return fail(SFFT_ERR_INVALID_ARG, "NULL path");
        sfft::tune_save(path);
        return SFFT_OK;
    });
}

int sfft_tune_load(const char *path, int check_device, size_t *loaded) {
    return guarded([&] {
        if (!path) return fail(SFFT_ERR_INVALID_ARG, "NULL path");
        const size_t k = sfft::tune_load(path, check_device != 0);
        if (loaded) *loaded = k;
        return SFFT_OK;
    });
}

int sfft_tune_clear(void) {
    return guarded([&] {
        sfft::tune_clear();
        return SFFT_OK;
    });
}

int sfft_plan_create_hinted(sfft_plan *plan, size_t n, int kind, int precision,
                            const sfft_plan_hints *hints) {
    return guarded([&] {
        if (!plan) return fail(SFFT_ERR_INVALID_ARG, "plan pointer is NULL");
        *plan = nullptr;
        sfft::PlanHints h;
        if (hints) h = from_abi(*hints);
        auto p = std::make_shared<sfft::Plan>(n, kind, precision, &h);
        *plan = new sfft_plan_s{p};
        return SFFT_OK;
    });
}

int sfft_plan_variants(size_t n, int kind, int precision) {
    const size_t cn = kind == SFFT_C2C ? n : n / 2;
    return sfft::fixed_variant_count((int)cn, precision, kind == SFFT_C2C ? 0 : 1);
}

int sfft_plan_destroy(sfft_plan plan) {
    delete plan;
    return SFFT_OK;
}

int sfft_plan_get_info(sfft_plan plan, sfft_plan_info *info) {
    return guarded([&] {
        if (!plan || !info) return fail(SFFT_ERR_INVALID_ARG, "NULL argument");
        const sfft::Plan &p = *plan->plan;
        std::memset(info, 0, sizeof(*info));
        info->n = p.n();
        info->kind = p.kind();
        info->precision = p.prec();
        // a chirp-z plan reports the layout of its length-M inner FFT (run twice)
        const sfft::Plan &q = p.bluestein() ? *p.bluestein_inner() : p;
        info->passes = (int)q.passes().size() * (p.bluestein() ? 2 : 1);
        info->chirp_length = p.bluestein_m();
        const sfft::PassDesc &pd = q.passes()[0].pd;
        info->nradices = pd.nstages;
        for (int s = 0; s < pd.nstages && s < 16; ++s) info->radices[s] = pd.st[s].radix;
        info->threads_per_transform = pd.tpt;
        info->transforms_per_block = pd.tpw;
        info->workspace_bytes = p.workspace_bytes(1, true);
        info->table_bytes = p.table_bytes();
        info->mfma_stages = p.mfma_stages();
        info->in_place = 1;
        return SFFT_OK;
    });
}

int sfft_execute(sfft_plan plan, const void *in, void *out, size_t batch, int direction,
                 void *stream) {
    return guarded([&] {
        if (!plan) return fail(SFFT_ERR_INVALID_ARG, "plan is NULL");
        if (direction != SFFT_FORWARD && direction != SFFT_INVERSE)
            return fail(SFFT_ERR_INVALID_ARG, "direction must be SFFT_FORWARD or SFFT_INVERSE");
        const sfft::Plan &p = *plan->plan;
        p.execute(in, p.default_in_dist(direction), out, p.default_out_dist(direction), batch,
                  direction, (hipStream_t)stream);
        return SFFT_OK;
    });
}

int sfft_plan_workspace_size(sfft_plan plan, size_t batch, int in_place, size_t *bytes) {
    return guarded([&] {
        if (!plan || !bytes) return fail(SFFT_ERR_INVALID_ARG, "NULL argument");
        *bytes = plan->plan->workspace_bytes(batch, in_place != 0);
        return SFFT_OK;
    });
}

int sfft_execute_workspace(sfft_plan plan, const void *in, void *out, size_t batch,
                           int direction, void *workspace, size_t workspace_bytes,
                           void *stream) {
    return guarded([&] {
        if (!plan) return fail(SFFT_ERR_INVALID_ARG, "plan is NULL");
        if (direction != SFFT_FORWARD && direction != SFFT_INVERSE)
            return fail(SFFT_ERR_INVALID_ARG, "direction must be SFFT_FORWARD or SFFT_INVERSE");
        const sfft::Plan &p = *plan->plan;
        const bool in_place = in != nullptr && in == out;
        if (p.workspace_bytes(batch, in_place) && !workspace)
            return fail(SFFT_ERR_INVALID_ARG, "this call needs a workspace");
        p.execute(in, p.default_in_dist(direction), out, p.default_out_dist(direction), batch,
                  direction, (hipStream_t)stream, workspace, workspace_bytes);
        return SFFT_OK;
    });
}

int sfft_execute_strided(sfft_plan plan, const void *in, size_t in_dist, void *out,
                         size_t out_dist, size_t batch, int direction, void *stream) {
    return guarded([&] {
        if (!plan) return fail(SFFT_ERR_INVALID_ARG, "plan is NULL");
        if (direction != SFFT_FORWARD && direction != SFFT_INVERSE)
            return fail(SFFT_ERR_INVALID_ARG, "direction must be SFFT_FORWARD or SFFT_INVERSE");
        plan->plan->execute(in, in_dist, out, out_dist, batch, direction, (hipStream_t)stream);
        return SFFT_OK;
    });
}

int sfft_batch_create(sfft_batch *batch, const sfft_batch_item *items, size_t count,
                      int precision) {
    return guarded([&] {
        if (!batch || (!items && count)) return fail(SFFT_ERR_INVALID_ARG, "NULL argument");
        *batch = nullptr;
        auto impl = std::make_unique<sfft::UnevenBatch>(items, count, precision);
        *batch = new sfft_batch_s{std::move(impl)};
        return SFFT_OK;
    });
}

int sfft_batch_destroy(sfft_batch batch) {
    delete batch;
    return SFFT_OK;
}

int sfft_batch_execute(sfft_batch batch, const void *in_arena, void *out_arena, void *stream) {
    return guarded([&] {
        if (!batch) return fail(SFFT_ERR_INVALID_ARG, "batch is NULL");
        batch->impl->execute_persistent(in_arena, out_arena, (hipStream_t)stream);
        return SFFT_OK;
    });
}

int sfft_batch_execute_per_size(sfft_batch batch, const void *in_arena, void *out_arena,
                                void *stream) {
    return guarded([&] {
        if (!batch) return fail(SFFT_ERR_INVALID_ARG, "batch is NULL");
        batch->impl->execute_per_size(in_arena, out_arena, (hipStream_t)stream);
        return SFFT_OK;
    });
}

int sfft_malloc_device(void **ptr, size_t bytes) {
    return guarded([&] {
        if (!ptr) return fail(SFFT_ERR_INVALID_ARG, "NULL pointer");
        *ptr = nullptr;
        if (hipMalloc(ptr, bytes ? bytes : 1) != hipSuccess) {
            (void)hipGetLastError();
            return fail(SFFT_ERR_OUT_OF_MEMORY, "device allocation failed");
        }
        return SFFT_OK;
    });
}

int sfft_free_device(void *ptr) {
    return guarded([&] {
        if (ptr) sfft::check_hip(hipFree(ptr), "hipFree");
        return SFFT_OK;
    });
}

int sfft_malloc_host(void **ptr, size_t bytes) {
    return guarded([&] {
        if (!ptr) return fail(SFFT_ERR_INVALID_ARG, "NULL pointer");
        *ptr = nullptr;
        // mapped for device access, coherent: kernels that read and write it
        // directly see the host's latest values on every launch
        if (hipHostMalloc(ptr, bytes ? bytes : 1, hipHostMallocMapped | hipHostMallocCoherent) !=
            hipSuccess) {
            (void)hipGetLastError();
            return fail(SFFT_ERR_OUT_OF_MEMORY, "pinned host allocation failed");
        }
        return SFFT_OK;
    });
}

int sfft_free_host(void *ptr) {
    return guarded([&] {
        if (ptr) sfft::check_hip(hipHostFree(ptr), "hipHostFree");
        return SFFT_OK;
    });
}

int sfft_memcpy(void *dst, const void *src, size_t bytes, int kind, void *stream) {
    return guarded([&] {
        hipMemcpyKind k;
        switch (kind) {
        case SFFT_COPY_H2D: k = hipMemcpyHostToDevice; break;
        case SFFT_COPY_D2H: k = hipMemcpyDeviceToHost; break;
        case SFFT_COPY_D2D: k = hipMemcpyDeviceToDevice; break;
        default: return fail(SFFT_ERR_INVALID_ARG, "copy kind");
        }
        if (bytes) sfft::check_hip(hipMemcpyAsync(dst, src, bytes, k, (hipStream_t)stream), "copy");
        return SFFT_OK;
    });
}

int sfft_stream_synchronize(void *stream) {
    return guarded([&] {
        sfft::check_hip(hipStreamSynchronize((hipStream_t)stream), "stream synchronize");
        return SFFT_OK;
    });
}

int sfft_host_device_pointer(void **device_ptr, void *host_ptr) {
    return guarded([&] {
        if (!device_ptr || !host_ptr) return fail(SFFT_ERR_INVALID_ARG, "NULL pointer");
        *device_ptr = nullptr;
        sfft::check_hip(hipHostGetDevicePointer(device_ptr, host_ptr, 0),
                        "device address of pinned host memory");
        return SFFT_OK;
    });
}

int sfft_stream_create(void **stream) {
    return guarded([&] {
        if (!stream) return fail(SFFT_ERR_INVALID_ARG, "NULL pointer");
        hipStream_t s = nullptr;
        sfft::check_hip(hipStreamCreateWithFlags(&s, hipStreamNonBlocking), "stream create");
        *stream = s;
        return SFFT_OK;
    });
}

int sfft_stream_destroy(void *stream) {
    return guarded([&] {
        if (stream) sfft::check_hip(hipStreamDestroy((hipStream_t)stream), "stream destroy");
        return SFFT_OK;
    });
}

int sfft_event_create(void **event) {
    return guarded([&] {
        if (!event) return fail(SFFT_ERR_INVALID_ARG, "NULL pointer");
        hipEvent_t e = nullptr;
        sfft::check_hip(hipEventCreateWithFlags(&e, hipEventDisableTiming), "event create");
        *event = e;
        return SFFT_OK;
    });
}

int sfft_event_destroy(void *event) {
    return guarded([&] {
        if (event) sfft::check_hip(hipEventDestroy((hipEvent_t)event), "event destroy");
        return SFFT_OK;
    });
}

int sfft_event_record(void *event, void *stream) {
    return guarded([&] {
        if (!event) return fail(SFFT_ERR_INVALID_ARG, "NULL event");
        sfft::check_hip(hipEventRecord((hipEvent_t)event, (hipStream_t)stream), "event record");
        return SFFT_OK;
    });
}

int sfft_event_synchronize(void *event) {
    return guarded([&] {
        if (!event) return fail(SFFT_ERR_INVALID_ARG, "NULL event");
        sfft::check_hip(hipEventSynchronize((hipEvent_t)event), "event synchronize");
        return SFFT_OK;
    });
}

int sfft_fill(void *dst, size_t count, int precision, int dist, uint64_t seed, void *stream) {
    return guarded([&] {
        if (!dst && count) return fail(SFFT_ERR_INVALID_ARG, "NULL buffer");
        if (dist != 0 && dist != 1) return fail(SFFT_ERR_INVALID_ARG, "dist must be 0 or 1");
        sfft::check_hip(sfft::fill_synthetic(dst, count, precision == SFFT_F64, dist, seed,
                                             (hipStream_t)stream),
                        "fill");
        return SFFT_OK;
    });
}

}  // extern "C"
