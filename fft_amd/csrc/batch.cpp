// batch.cpp -- uneven batches: unit list, descriptor table, persistent launch.
#include "batch.h"

#include <algorithm>
#include <map>
#include <tuple>

namespace sfft {

UnevenBatch::UnevenBatch(const sfft_batch_item *items, size_t count, int prec)
    : prec_(prec), items_(items, items + count) {
    if (prec != kF32 && prec != kF64) throw Error(SFFT_ERR_INVALID_ARG, "precision");
    for (const auto &it : items_) {
        if (it.direction != SFFT_FORWARD && it.direction != SFFT_INVERSE)
            throw Error(SFFT_ERR_INVALID_ARG, "batch item direction");
        plans_.push_back(get_plan(it.n, it.kind, prec));
        // chirp-z lengths (prime factors above 13) run as their plan's own
        // launch sequence after the persistent launch, in the same graph,
        // through one workspace sized for the largest of them
        if (plans_.back()->bluestein())
            ws_bytes_ = std::max(ws_bytes_, plans_.back()->workspace_bytes(it.count, false));
    }

    const size_t lds_cap = device_info().lds_bytes;
    check_hip(hipEventCreateWithFlags(&done_, hipEventDisableTiming), "batch event");
    try {
        if (ws_bytes_ && hipMalloc(&d_ws_, ws_bytes_) != hipSuccess) {
            (void)hipGetLastError();
            d_ws_ = nullptr;
            throw Error(SFFT_ERR_OUT_OF_MEMORY, "chirp-z workspace allocation failed");
        }
        if (prec == kF64)
            build<double>(lds_cap);
        else
            build<float>(lds_cap);
    } catch (...) {
        release();
        throw;
    }
}

UnevenBatch::~UnevenBatch() { release(); }

void UnevenBatch::release() {
    // a launch still in flight reads the unit list and the counters
    if (done_ && launched_) (void)hipEventSynchronize(done_);
    if (done_) (void)hipEventDestroy(done_);
    done_ = nullptr;
    if (graph_exec_) (void)hipGraphExecDestroy(graph_exec_);
    if (graph_) (void)hipGraphDestroy(graph_);
    if (d_units_) (void)hipFree(d_units_);
    if (d_table_) (void)hipFree(d_table_);
    if (d_counters_) (void)hipFree(d_counters_);
    if (d_gather_in_) (void)hipFree(d_gather_in_);
    if (d_gather_out_) (void)hipFree(d_gather_out_);
    if (d_ws_) (void)hipFree(d_ws_);
    d_ws_ = nullptr;
    graph_exec_ = nullptr;
    graph_ = nullptr;
    d_units_ = nullptr;
    d_table_ = nullptr;
    d_counters_ = nullptr;
    d_gather_in_ = d_gather_out_ = nullptr;
}

template <typename T>
int UnitTable<T>::entry(const Plan &p, int inv) {
    const auto key = std::make_pair(&p, inv);
    auto it = first_.find(key);
    if (it != first_.end()) return it->second;
    // the first class (kernel instance) whose threads, elements per thread
    // and LDS share hold every pass of the plan
    for (int c = 0; c < persistent_classes<T>(); ++c) {
        const PersistentClass pc = persistent_class<T>(c);
        const size_t cap = lds_bytes_ / (size_t)pc.blocks_per_cu - kLdsReserve;
        std::vector<PassArgs<T>> ds;
        size_t need = 0;
        bool fits = true;
        for (size_t q = 0; q < p.passes().size() && fits; ++q) {
            PassArgs<T> a = p.pass_args<T>(q, inv);
            const size_t sl = p.slot_lds_bytes(q);
            if (p.passes()[q].ept > pc.ept || a.pd.tpt > pc.threads || sl > cap) {
                fits = false;
                break;
            }
            int tpw = pc.threads / a.pd.tpt;
            if (sl) tpw = std::min<int>(tpw, (int)(cap / sl));
            if (p.passes().size() == 1) {
                if (tpw < 2) {
                    // one slot per round (LDS-bound lengths): one butterfly per
                    // thread in every stage, up to the block (the bits do not
                    // depend on the thread layout)
                    int wide = 1;
                    for (int st = 0; st < a.pd.nstages; ++st)
                        wide = std::max(wide, a.pd.n / a.pd.st[st].radix);
                    if (wide <= pc.threads) a.pd.tpt = std::max(a.pd.tpt, wide);
                }
                a.pd.tpw = std::max(1, tpw);
            } else {
                // column passes: as many adjacent columns per round as the block
                // and LDS allow (at least the plan's own tile); the last round
                // of a transform masks the columns past its end
                tpw = std::min(tpw, p.passes()[q].slots_per_transform);
                int cols = 1;
                while (cols * 2 <= tpw) cols *= 2;
                a.pd.tpw = std::max(a.pd.tpw, cols);
                if (a.pd.tpw * a.pd.tpt > pc.threads || sl * (size_t)a.pd.tpw > cap) fits = false;
            }
            need = std::max(need, sl * (size_t)a.pd.tpw);
            ds.push_back(a);
        }
        if (!fits) continue;
        const int e0 = (int)descs_.size();
        descs_.insert(descs_.end(), ds.begin(), ds.end());
        cls_.resize(descs_.size(), c);
        lds_need_[c] = std::max(lds_need_[c], need);
        first_[key] = e0;
        return e0;
    }
    throw Error(SFFT_ERR_INVALID_ARG, "a transform of length " + std::to_string(p.n()) +
                                          " does not fit a persistent block");
}

template <typename T>
void UnitBuilder<T>::add(UnitTable<T> &table, const Plan &p, int kind, int inv, size_t count,
                         size_t in_off, size_t out_off) {
    const int e0 = table.entry(p, inv);
    const int npass = (int)p.passes().size();
    const size_t eb = sizeof(C2<T>);
    const size_t in_step = p.in_bytes(1, inv), out_step = p.out_bytes(1, inv);
    if ((in_off % eb) || (out_off % eb) || (in_step % eb) || (out_step % eb))
        throw Error(SFFT_ERR_INVALID_ARG, "batch item offsets must be element aligned");
    for (size_t c = 0; c < count; ++c) {
        const size_t io = in_off + c * in_step, oo = out_off + c * out_step;
        if (npass == 1) {
            groups_[e0].push_back({(int64_t)(io / eb), (int64_t)(oo / eb)});
            group_cost_[e0] = (double)(in_step + out_step);
            group_tpw_[e0] = table.descs()[(size_t)e0].pd.tpw;
            group_cls_[e0] = table.class_of(e0);
        } else {
            HostUnit h;
            const int post = (kind != SFFT_C2C && inv == 0) ? kUnitRealPost : 0;
            h.u = WorkUnit{e0, npass | post, 1, -1, (int64_t)io, (int64_t)oo};
            h.cost = (double)(in_step + out_step) * npass;
            h.cls = table.class_of(e0);
            hu_.push_back(h);
        }
    }
}

template <typename T>
void UnitBuilder<T>::finish(std::vector<WorkUnit> &units, std::vector<int64_t> &gin,
                            std::vector<int64_t> &gout,
                            size_t (&per_class)[kPersistentClasses]) {
    gin.clear();
    gout.clear();
    for (auto &g : groups_) {
        const int e0 = g.first;
        const size_t per_unit = (size_t)group_tpw_[e0];
        const auto &lst = g.second;
        for (size_t c = 0; c < lst.size(); c += per_unit) {
            const size_t cnt = std::min(per_unit, lst.size() - c);
            HostUnit h;
            h.u = WorkUnit{e0, 1, (int32_t)cnt, (int32_t)gin.size(), 0, 0};
            h.cost = (double)cnt * group_cost_[e0];
            h.cls = group_cls_[e0];
            for (size_t k = 0; k < cnt; ++k) {
                gin.push_back(lst[c + k].first);
                gout.push_back(lst[c + k].second);
            }
            hu_.push_back(h);
        }
    }
    // by class, longest first in each: big units start early, small ones
    // fill the tail
    std::stable_sort(hu_.begin(), hu_.end(), [](const HostUnit &a, const HostUnit &b) {
        if (a.cls != b.cls) return a.cls < b.cls;
        return a.cost > b.cost;
    });
    units.clear();
    units.reserve(hu_.size());
    for (int c = 0; c < kPersistentClasses; ++c) per_class[c] = 0;
    for (auto &h : hu_) {
        units.push_back(h.u);
        ++per_class[h.cls];
    }
    hu_.clear();
    groups_.clear();
    group_cost_.clear();
    group_tpw_.clear();
    group_cls_.clear();
}

template <typename T>
unsigned persistent_grid(int cls, size_t lds_bytes, size_t nunits) {
    const int per_cu = persistent_blocks_per_cu<T>(cls, lds_bytes);
    unsigned grid = (unsigned)(device_info().cu_count * per_cu);
    if (nunits && grid > nunits) grid = (unsigned)nunits;
    return grid ? grid : 1;
}

template class UnitTable<float>;
template class UnitTable<double>;
template class UnitBuilder<float>;
template class UnitBuilder<double>;
template unsigned persistent_grid<float>(int, size_t, size_t);
template unsigned persistent_grid<double>(int, size_t, size_t);

template <typename T>
void UnevenBatch::build(size_t lds_bytes) {
    // One descriptor set per (plan, direction); the passes and radices are
    // the plan's own, so every transform computes exactly what sfft_execute
    // computes for it.  Only the slots per round (tpw) and the threads per
    // slot change, to fill the block of the unit's kernel class.
    UnitTable<T> table(lds_bytes);
    UnitBuilder<T> builder;
    for (size_t i = 0; i < items_.size(); ++i) {
        const auto &it = items_[i];
        if (plans_[i]->bluestein()) continue;
        builder.add(table, *plans_[i], it.kind, it.direction, it.count, it.in_offset,
                    it.out_offset);
    }
    std::vector<WorkUnit> units;
    std::vector<int64_t> gin, gout;
    builder.finish(units, gin, gout, nunits_);
    for (int c = 0; c < kPersistentClasses; ++c) {
        lds_bytes_[c] = table.lds_need(c);
        grid_[c] = nunits_[c] ? persistent_grid<T>(c, lds_bytes_[c], nunits_[c]) : 0;
    }

    auto upload = [](void **dst, const void *src, size_t bytes, const char *what) {
        if (hipMalloc(dst, bytes ? bytes : 16) != hipSuccess)
            throw Error(SFFT_ERR_OUT_OF_MEMORY, std::string(what) + " allocation failed");
        if (bytes) check_hip(hipMemcpy(*dst, src, bytes, hipMemcpyHostToDevice), what);
    };
    const auto &table_h = table.descs();
    upload((void **)&d_units_, units.data(), units.size() * sizeof(WorkUnit), "units");
    upload(&d_table_, table_h.data(), table_h.size() * sizeof(PassArgs<T>), "descriptor table");
    upload((void **)&d_gather_in_, gin.data(), gin.size() * sizeof(int64_t), "gather list");
    upload((void **)&d_gather_out_, gout.data(), gout.size() * sizeof(int64_t), "gather list");
    const size_t cb = 2 * kPersistentClasses * sizeof(unsigned);
    if (hipMalloc((void **)&d_counters_, cb) != hipSuccess)
        throw Error(SFFT_ERR_OUT_OF_MEMORY, "counters allocation failed");
    check_hip(hipMemset(d_counters_, 0, cb), "counter init");
}

void UnevenBatch::launch(const void *in_arena, void *out_arena, hipStream_t stream) {
    // one persistent launch per kernel class, in class order
    size_t first = 0;
    for (int c = 0; c < kPersistentClasses; ++c) {
        if (!nunits_[c]) continue;
        const std::string what = "persistent launch (class " + std::to_string(c) + ", grid " +
                                 std::to_string(grid_[c]) + ", LDS " +
                                 std::to_string(lds_bytes_[c]) + " B)";
        const WorkUnit *u = d_units_ + first;
        first += nunits_[c];
        if (prec_ == kF64)
            check_hip(launch_persistent<double>(c, u, (int)nunits_[c],
                                                (const PassArgs<double> *)d_table_, d_gather_in_,
                                                d_gather_out_, d_counters_ + 2 * c, in_arena,
                                                out_arena, grid_[c], lds_bytes_[c], stream),
                      what.c_str());
        else
            check_hip(launch_persistent<float>(c, u, (int)nunits_[c],
                                               (const PassArgs<float> *)d_table_, d_gather_in_,
                                               d_gather_out_, d_counters_ + 2 * c, in_arena,
                                               out_arena, grid_[c], lds_bytes_[c], stream),
                      what.c_str());
    }
    for (size_t k = 0; k < items_.size(); ++k) {
        const auto &it = items_[k];
        const Plan &p = *plans_[k];
        if (!p.bluestein()) continue;
        p.execute((const char *)in_arena + it.in_offset, p.default_in_dist(it.direction),
                  (char *)out_arena + it.out_offset, p.default_out_dist(it.direction), it.count,
                  it.direction, stream, d_ws_, ws_bytes_);
    }
}

void UnevenBatch::execute_persistent(const void *in_arena, void *out_arena, hipStream_t stream) {
    if (!in_arena || !out_arena) throw Error(SFFT_ERR_INVALID_ARG, "NULL arena");
    if (in_arena == out_arena) throw Error(SFFT_ERR_ALIASING, "arenas must differ");
    std::lock_guard<std::mutex> lock(mu_);
    if (!graph_exec_ || graph_in_ != in_arena || graph_out_ != out_arena) {
        // the previous launch may still run the graph about to be replaced
        if (launched_) check_hip(hipEventSynchronize(done_), "previous batch launch");
        if (graph_exec_) (void)hipGraphExecDestroy(graph_exec_);
        if (graph_) (void)hipGraphDestroy(graph_);
        graph_exec_ = nullptr;
        graph_ = nullptr;
        hipStream_t cs;
        check_hip(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking), "capture stream");
        check_hip(hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal), "begin capture");
        launch(in_arena, out_arena, cs);
        hipError_t e = hipStreamEndCapture(cs, &graph_);
        (void)hipStreamDestroy(cs);
        check_hip(e, "end capture");
        check_hip(hipGraphInstantiate(&graph_exec_, graph_, nullptr, nullptr, 0),
                  "graph instantiate");
        graph_in_ = in_arena;
        graph_out_ = out_arena;
    }
    // the work queue is shared by every launch of this batch: start after the
    // previous launch has drained it (and reset its counters)
    if (launched_) check_hip(hipStreamWaitEvent(stream, done_, 0), "order batch launches");
    check_hip(hipGraphLaunch(graph_exec_, stream), "graph launch");
    check_hip(hipEventRecord(done_, stream), "batch launch event");
    launched_ = true;
}

void UnevenBatch::execute_per_size(const void *in_arena, void *out_arena, hipStream_t stream) {
    for (size_t i = 0; i < items_.size(); ++i) {
        const auto &it = items_[i];
        const Plan &p = *plans_[i];
        p.execute((const char *)in_arena + it.in_offset, p.default_in_dist(it.direction),
                  (char *)out_arena + it.out_offset, p.default_out_dist(it.direction), it.count,
                  it.direction, stream);
    }
}

}  // namespace sfft
