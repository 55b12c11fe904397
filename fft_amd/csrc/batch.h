// batch.h -- uneven batches: many transform sizes in one persistent launch.
#pragma once

#include <hip/hip_runtime.h>
#include <sfft.h>

#include <memory>
#include <vector>

#include "engine.h"

namespace sfft {

class UnevenBatch {
  public:
    UnevenBatch(const sfft_batch_item *items, size_t count, int prec);
    ~UnevenBatch();
    // One persistent launch over every transform of every item.
    void execute_persistent(const void *in_arena, void *out_arena, hipStream_t stream);
    // The plain per-size loop: one sfft_execute per item.
    void execute_per_size(const void *in_arena, void *out_arena, hipStream_t stream);

  private:
    int prec_;
    std::vector<sfft_batch_item> items_;
    std::vector<std::shared_ptr<Plan>> plans_;
};

}  // namespace sfft
