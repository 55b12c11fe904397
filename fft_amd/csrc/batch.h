// batch.h -- uneven batches: many transform sizes in one persistent launch.
#pragma once

#include <hip/hip_runtime.h>
#include <sfft.h>

#include <memory>
#include <mutex>
#include <vector>

#include "engine.h"
#include "persistent.h"

namespace sfft {

class UnevenBatch {
  public:
    UnevenBatch(const sfft_batch_item *items, size_t count, int prec);
    ~UnevenBatch();
    UnevenBatch(const UnevenBatch &) = delete;
    UnevenBatch &operator=(const UnevenBatch &) = delete;

    // One persistent launch over every transform of every item, replayed
    // from a captured hipGraph (re-captured when the arenas change).  The
    // launch's work queue lives in the batch, so launches of one batch are
    // serialised on the device: each waits for the previous one (on any
    // stream, from any thread) before it starts.
    void execute_persistent(const void *in_arena, void *out_arena, hipStream_t stream);
    // The plain per-size loop: one batched launch sequence per item.
    void execute_per_size(const void *in_arena, void *out_arena, hipStream_t stream);

    size_t units() const { return nunits_; }
    unsigned grid() const { return grid_; }

  private:
    template <typename T>
    void build(size_t lds_cap);
    void launch(const void *in_arena, void *out_arena, hipStream_t stream);
    void release();

    int prec_;
    std::vector<sfft_batch_item> items_;
    std::vector<std::shared_ptr<Plan>> plans_;
    WorkUnit *d_units_ = nullptr;
    void *d_table_ = nullptr;
    unsigned *d_counters_ = nullptr;
    int64_t *d_gather_in_ = nullptr;
    int64_t *d_gather_out_ = nullptr;
    size_t nunits_ = 0;
    unsigned grid_ = 0;
    size_t lds_bytes_ = 0;
    // captured launch; mu_ guards the graph and the ordering event, which
    // marks the end of the last launch (launched_: recorded at least once)
    std::mutex mu_;
    hipEvent_t done_ = nullptr;
    bool launched_ = false;
    hipGraph_t graph_ = nullptr;
    hipGraphExec_t graph_exec_ = nullptr;
    const void *graph_in_ = nullptr;
    void *graph_out_ = nullptr;
};

}  // namespace sfft
