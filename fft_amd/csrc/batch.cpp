// batch.cpp -- uneven batches.
#include "batch.h"

namespace sfft {

UnevenBatch::UnevenBatch(const sfft_batch_item *items, size_t count, int prec)
    : prec_(prec), items_(items, items + count) {
    for (const auto &it : items_) {
        if (it.direction != SFFT_FORWARD && it.direction != SFFT_INVERSE)
            throw Error(SFFT_ERR_INVALID_ARG, "batch item direction");
        plans_.push_back(get_plan(it.n, it.kind, prec));
    }
}

UnevenBatch::~UnevenBatch() {}

void UnevenBatch::execute_per_size(const void *in_arena, void *out_arena, hipStream_t stream) {
    for (size_t i = 0; i < items_.size(); ++i) {
        const auto &it = items_[i];
        const Plan &p = *plans_[i];
        p.execute((const char *)in_arena + it.in_offset, p.default_in_dist(it.direction),
                  (char *)out_arena + it.out_offset, p.default_out_dist(it.direction), it.count,
                  it.direction, stream);
    }
}

void UnevenBatch::execute_persistent(const void *, void *, hipStream_t) {
    throw Error(SFFT_ERR_UNSUPPORTED_SIZE, "persistent uneven batch not built yet");
}

}  // namespace sfft
