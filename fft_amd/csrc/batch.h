// batch.h -- uneven batches: many transform sizes in one persistent launch.
#pragma once

#include <hip/hip_runtime.h>
#include <sfft.h>

#include <map>
#include <memory>
#include <mutex>
#include <utility>
#include <vector>

#include "engine.h"
#include "persistent.h"

namespace sfft {

// Work list of one persistent launch.  Transforms are added one item at a
// time; single-pass ones are gathered per (plan, direction) into units of
// one workgroup round, two-pass ones are a unit each; finish() orders the
// units longest first.  The pass descriptors come from a UnitTable, so a
// transform computes exactly what sfft_execute computes for it (same passes,
// radices and tables; only the slots per round differ).
template <typename T>
class UnitTable {
  public:
    // lds_bytes: LDS of one CU
    explicit UnitTable(size_t lds_bytes) : lds_bytes_(lds_bytes) {}
    // First descriptor index of (plan, direction), appending its passes.
    int entry(const Plan &p, int inverse);
    const std::vector<PassArgs<T>> &descs() const { return descs_; }
    // Persistent kernel class (persistent_class) of an entry's units, and the
    // LDS the units of each class need.
    int class_of(int e0) const { return cls_.at((size_t)e0); }
    size_t lds_need(int cls) const { return lds_need_[cls]; }

  private:
    size_t lds_bytes_;
    size_t lds_need_[kPersistentClasses] = {};
    std::vector<PassArgs<T>> descs_;
    std::vector<int> cls_;  // per descriptor: its entry's class
    std::map<std::pair<const Plan *, int>, int> first_;
};

template <typename T>
class UnitBuilder {
  public:
    // `count` transforms of plan p in direction `inverse`, the first at byte
    // offsets in_off / out_off of the arenas, the rest back to back.
    void add(UnitTable<T> &table, const Plan &p, int kind, int inverse, size_t count,
             size_t in_off, size_t out_off);
    // Units plus the gather lists (element offsets), grouped by kernel class
    // in class order (per_class[c] units of class c), longest first in each.
    void finish(std::vector<WorkUnit> &units, std::vector<int64_t> &gin,
                std::vector<int64_t> &gout, size_t (&per_class)[kPersistentClasses]);
    bool empty() const { return hu_.empty() && groups_.empty(); }

  private:
    struct HostUnit {
        WorkUnit u;
        double cost;
        int cls;
    };
    std::vector<HostUnit> hu_;
    std::map<int, std::vector<std::pair<int64_t, int64_t>>> groups_;
    std::map<int, double> group_cost_;
    std::map<int, int> group_tpw_;
    std::map<int, int> group_cls_;
};

class UnevenBatch {
  public:
    UnevenBatch(const sfft_batch_item *items, size_t count, int prec);
    ~UnevenBatch();
    UnevenBatch(const UnevenBatch &) = delete;
    UnevenBatch &operator=(const UnevenBatch &) = delete;

    // One persistent launch over every transform of every item (items of
    // chirp-z lengths follow it as their plan's own launches, in the same
    // graph), replayed from a captured hipGraph (re-captured when the arenas
    // change).  The launch's work queue lives in the batch, so launches of
    // one batch are serialised on the device: each waits for the previous one
    // (on any stream, from any thread) before it starts.
    void execute_persistent(const void *in_arena, void *out_arena, hipStream_t stream);
    // The plain per-size loop: one batched launch sequence per item.
    void execute_per_size(const void *in_arena, void *out_arena, hipStream_t stream);


  private:
    template <typename T>
    void build(size_t lds_bytes);
    void launch(const void *in_arena, void *out_arena, hipStream_t stream);
    void release();

    int prec_;
    std::vector<sfft_batch_item> items_;
    std::vector<std::shared_ptr<Plan>> plans_;
    WorkUnit *d_units_ = nullptr;
    void *d_table_ = nullptr;
    unsigned *d_counters_ = nullptr;  // one queue (2 counters) per kernel class
    int64_t *d_gather_in_ = nullptr;
    int64_t *d_gather_out_ = nullptr;
    // units grouped by kernel class (nunits_[c] of class c); one launch each
    size_t nunits_[kPersistentClasses] = {};
    unsigned grid_[kPersistentClasses] = {};
    size_t lds_bytes_[kPersistentClasses] = {};
    // workspace of the chirp-z items (run after the persistent launch)
    void *d_ws_ = nullptr;
    size_t ws_bytes_ = 0;
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

// Workgroups of a persistent launch of class cls whose descriptors need
// lds_bytes of LDS.
template <typename T>
unsigned persistent_grid(int cls, size_t lds_bytes, size_t nunits);

}  // namespace sfft
