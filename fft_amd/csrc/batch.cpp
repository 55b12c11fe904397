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
        if (plans_.back()->bluestein())
            throw Error(SFFT_ERR_UNSUPPORTED_SIZE,
                        "size " + std::to_string(it.n) +
                            ": lengths with prime factors > 13 run through sfft_execute only");
    }
    const size_t lds_cap = device_info().lds_bytes - kLdsReserve;
    check_hip(hipEventCreateWithFlags(&done_, hipEventDisableTiming), "batch event");
    try {
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
    graph_exec_ = nullptr;
    graph_ = nullptr;
    d_units_ = nullptr;
    d_table_ = nullptr;
    d_counters_ = nullptr;
    d_gather_in_ = d_gather_out_ = nullptr;
}

template <typename T>
void UnevenBatch::build(size_t lds_cap) {
    // One descriptor set per (plan, direction); the passes and radices are
    // the plan's own, so every transform computes exactly what sfft_execute
    // computes for it.  Only the slots per round change (tpw): the block is
    // persistent_block<T>() threads wide.
    std::vector<PassArgs<T>> table;
    std::map<std::pair<const Plan *, int>, int> first_entry;
    size_t lds_need = 0;
    struct HostUnit {
        WorkUnit u;
        double cost;
    };
    std::vector<HostUnit> hu;
    // single-pass transforms grouped per (plan, direction): element offsets
    std::map<int, std::vector<std::pair<int64_t, int64_t>>> groups;
    std::map<int, double> group_cost;
    const size_t eb = sizeof(C2<T>);
    for (size_t i = 0; i < items_.size(); ++i) {
        const auto &it = items_[i];
        const Plan &p = *plans_[i];
        const int inv = it.direction;
        auto key = std::make_pair(&p, inv);
        if (!first_entry.count(key)) {
            first_entry[key] = (int)table.size();
            for (size_t q = 0; q < p.passes().size(); ++q) {
                PassArgs<T> a = p.pass_args<T>(q, inv);
                const size_t sl = p.slot_lds_bytes(q);
                if (p.passes().size() == 1) {
                    int tpw = persistent_block<T>() / a.pd.tpt;
                    if (sl) tpw = std::min<int>(tpw, (int)(lds_cap / sl));
                    a.pd.tpw = std::max(1, tpw);
                } else {
                    // column passes: as many adjacent columns per round as the
                    // block and LDS allow; the last round of a transform masks
                    // the columns past its end
                    const int spt = p.passes()[q].slots_per_transform;
                    int tpw = persistent_block<T>() / a.pd.tpt;
                    if (sl) tpw = std::min<int>(tpw, (int)(lds_cap / sl));
                    tpw = std::min(tpw, spt);
                    int c = 1;
                    while (c * 2 <= tpw) c *= 2;
                    a.pd.tpw = std::max(a.pd.tpw, c);
                }
                lds_need = std::max(lds_need, sl * (size_t)a.pd.tpw);
                table.push_back(a);
            }
        }
        const int e0 = first_entry[key];
        const int npass = (int)p.passes().size();
        const size_t in_step = p.in_bytes(1, inv), out_step = p.out_bytes(1, inv);
        if ((it.in_offset % eb) || (it.out_offset % eb) || (in_step % eb) || (out_step % eb))
            throw Error(SFFT_ERR_INVALID_ARG, "batch item offsets must be element aligned");
        for (size_t c = 0; c < it.count; ++c) {
            const size_t io = it.in_offset + c * in_step, oo = it.out_offset + c * out_step;
            if (npass == 1) {
                groups[e0].push_back({(int64_t)(io / eb), (int64_t)(oo / eb)});
                group_cost[e0] = (double)(in_step + out_step);
            } else {
                HostUnit h;
                const int post = (it.kind != SFFT_C2C && inv == 0) ? kUnitRealPost : 0;
                h.u = WorkUnit{e0, npass | post, 1, -1, (int64_t)io, (int64_t)oo};
                h.cost = (double)(in_step + out_step) * npass;
                hu.push_back(h);
            }
        }
    }
    std::vector<int64_t> gin, gout;
    for (auto &g : groups) {
        const int e0 = g.first;
        const size_t per_unit = (size_t)table[e0].pd.tpw;
        const auto &lst = g.second;
        for (size_t c = 0; c < lst.size(); c += per_unit) {
            const size_t cnt = std::min(per_unit, lst.size() - c);
            HostUnit h;
            h.u = WorkUnit{e0, 1, (int32_t)cnt, (int32_t)gin.size(), 0, 0};
            h.cost = (double)cnt * group_cost[e0];
            for (size_t k = 0; k < cnt; ++k) {
                gin.push_back(lst[c + k].first);
                gout.push_back(lst[c + k].second);
            }
            hu.push_back(h);
        }
    }
    // Longest first: big units start early, small ones fill the tail.
    std::stable_sort(hu.begin(), hu.end(),
                     [](const HostUnit &a, const HostUnit &b) { return a.cost > b.cost; });
    std::vector<WorkUnit> units;
    units.reserve(hu.size());
    for (auto &h : hu) units.push_back(h.u);
    nunits_ = units.size();
    lds_bytes_ = lds_need;
    const int per_cu = persistent_blocks_per_cu<T>(lds_bytes_);
    grid_ = (unsigned)(device_info().cu_count * per_cu);
    if (nunits_ && grid_ > nunits_) grid_ = (unsigned)nunits_;
    if (grid_ == 0) grid_ = 1;

    auto upload = [](void **dst, const void *src, size_t bytes, const char *what) {
        if (hipMalloc(dst, bytes ? bytes : 16) != hipSuccess)
            throw Error(SFFT_ERR_OUT_OF_MEMORY, std::string(what) + " allocation failed");
        if (bytes) check_hip(hipMemcpy(*dst, src, bytes, hipMemcpyHostToDevice), what);
    };
    upload((void **)&d_units_, units.data(), units.size() * sizeof(WorkUnit), "units");
    upload(&d_table_, table.data(), table.size() * sizeof(PassArgs<T>), "descriptor table");
    upload((void **)&d_gather_in_, gin.data(), gin.size() * sizeof(int64_t), "gather list");
    upload((void **)&d_gather_out_, gout.data(), gout.size() * sizeof(int64_t), "gather list");
    upload((void **)&d_counters_, nullptr, 0, "counters");
    check_hip(hipMemset(d_counters_, 0, 2 * sizeof(unsigned)), "counter init");
}

void UnevenBatch::launch(const void *in_arena, void *out_arena, hipStream_t stream) {
    const std::string what = "persistent launch (grid " + std::to_string(grid_) + ", LDS " +
                             std::to_string(lds_bytes_) + " B)";
    if (prec_ == kF64)
        check_hip(launch_persistent<double>(d_units_, (int)nunits_,
                                            (const PassArgs<double> *)d_table_, d_gather_in_,
                                            d_gather_out_, d_counters_,
                                            in_arena, out_arena, grid_, lds_bytes_, stream),
                  what.c_str());
    else
        check_hip(launch_persistent<float>(d_units_, (int)nunits_,
                                           (const PassArgs<float> *)d_table_, d_gather_in_,
                                           d_gather_out_, d_counters_,
                                           in_arena, out_arena, grid_, lds_bytes_, stream),
                  what.c_str());
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
