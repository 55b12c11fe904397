// serve.cpp -- serving loop: continuous batching of transform requests.
//
// A worker thread admits queued requests in arrival order while an in-flight
// slot is free and their buffers (input, output and any workspace) fit in the
// HBM pool; the head of the queue blocks the rest (no reordering, no
// starvation).  Everything admitted in one pass is launched together:
//   * small requests (each at most kMergeBytes of input + output) are merged
//     into ONE persistent launch (the uneven-batch kernel over a per-tick
//     work list, with one fill launch synthesising all their inputs), so many
//     requests of different sizes, kinds and directions cost two launches;
//   * larger requests run their plan's own kernels (compiled where one
//     exists), each with its own fill, round-robin over the device streams.
// Each stream has a waiter thread that blocks on the completion events of
// its requests in launch order and retires them (frees their buffers unless
// retained, records the completion time), which wakes the worker to admit
// more.  Plans, twiddles and descriptor tables stay device-resident.  The
// plain mode serves one request at a time synchronously (the comparison
// path).
#include <sfft.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstring>
#include <condition_variable>
#include <deque>
#include <map>
#include <mutex>
#include <set>
#include <thread>
#include <unordered_map>
#include <vector>

#include "batch.h"
#include "engine.h"
#include "kernels.h"

namespace sfft {

namespace {

// First-fit allocator over one device arena (offsets, 256-byte aligned).
class Pool {
  public:
    explicit Pool(size_t bytes) : size_(bytes) {
        if (hipMalloc(&base_, bytes) != hipSuccess) {
            (void)hipGetLastError();
            throw Error(SFFT_ERR_OUT_OF_MEMORY, "serving pool allocation failed");
        }
        free_[0] = bytes;
    }
    ~Pool() {
        if (base_) (void)hipFree(base_);
    }
    // returns offset or SIZE_MAX
    size_t alloc(size_t bytes) {
        bytes = (bytes + 255) & ~size_t(255);
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second >= bytes) {
                const size_t off = it->first, rest = it->second - bytes;
                free_.erase(it);
                if (rest) free_[off + bytes] = rest;
                used_ += bytes;
                return off;
            }
        }
        return SIZE_MAX;
    }
    void release(size_t off, size_t bytes) {
        bytes = (bytes + 255) & ~size_t(255);
        used_ -= bytes;
        auto it = free_.emplace(off, bytes).first;
        auto next = std::next(it);
        if (next != free_.end() && it->first + it->second == next->first) {
            it->second += next->second;
            free_.erase(next);
        }
        if (it != free_.begin()) {
            auto prev = std::prev(it);
            if (prev->first + prev->second == it->first) {
                prev->second += it->second;
                free_.erase(it);
            }
        }
    }
    char *base() const { return static_cast<char *>(base_); }
    size_t size() const { return size_; }
    size_t used() const { return used_; }

  private:
    void *base_ = nullptr;
    size_t size_ = 0, used_ = 0;
    std::map<size_t, size_t> free_;
};

struct Req {
    sfft_request r;
    double t_arrive = 0, t_admit = 0, t_done = 0;
    size_t in_bytes = 0, out_bytes = 0, ws_bytes = 0;
    size_t in_off = SIZE_MAX, out_off = SIZE_MAX, ws_off = SIZE_MAX;
    std::shared_ptr<Plan> plan;
    hipEvent_t ev = nullptr;
    int status = SFFT_OK;
};

size_t align256(size_t b) { return (b + 255) & ~size_t(255); }

}  // namespace

class Server {
  public:
    // A request of at most this many input + output bytes joins the tick's
    // merged persistent launch.
    static constexpr size_t kMergeBytes = 16ull << 20;
    // Capacity of one stream's merged work list.
    static constexpr size_t kMaxUnits = 1 << 16, kMaxGather = 1 << 17, kMaxFills = 1024;
    static constexpr size_t kMaxDescs = 8192;

    explicit Server(const sfft_server_config &c) : cfg_(c) {
        if (cfg_.hbm_budget_bytes == 0) cfg_.hbm_budget_bytes = 64ull << 30;
        if (cfg_.max_inflight <= 0) cfg_.max_inflight = 64;
        if (cfg_.streams <= 0) cfg_.streams = 4;
        if (cfg_.plain) {
            cfg_.max_inflight = 1;
            cfg_.streams = 1;
        }
        if (cfg_.precision != kF32 && cfg_.precision != kF64)
            throw Error(SFFT_ERR_INVALID_ARG, "precision");
        pool_ = std::make_unique<Pool>(cfg_.hbm_budget_bytes);
        try {
            setup_device();
        } catch (...) {
            teardown_device();
            throw;
        }
        t0_ = std::chrono::steady_clock::now();
        if (!cfg_.plain)
            for (auto &st : streams_) st->waiter = std::thread([this, p = st.get()] { wait_loop(p); });
        worker_ = std::thread([this] { loop(); });
    }

    ~Server() {
        {
            std::lock_guard<std::mutex> lock(mu_);
            stop_ = true;
        }
        cv_.notify_all();
        worker_.join();
        {
            std::lock_guard<std::mutex> lock(mu_);
            stop_waiters_ = true;
        }
        for (auto &st : streams_) {
            st->cv.notify_all();
            if (st->waiter.joinable()) st->waiter.join();
        }
        for (auto &kv : retained_) pool_->release(kv.second.first, kv.second.second);
        teardown_device();
    }

    double now() const {
        return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0_).count();
    }

    void submit(const sfft_request &rq, double t_arrive) {
        auto *r = new Req;
        r->r = rq;
        r->t_arrive = t_arrive;
        try {
            r->plan = get_plan(rq.n, rq.kind, cfg_.precision);
            r->in_bytes = r->plan->in_bytes(rq.batch, rq.direction);
            r->out_bytes = r->plan->out_bytes(rq.batch, rq.direction);
            r->ws_bytes = r->plan->workspace_bytes(rq.batch, false);
            const size_t need =
                align256(r->in_bytes) + align256(r->out_bytes) + align256(r->ws_bytes);
            if (need > pool_->size())
                throw Error(SFFT_ERR_BUDGET, "request larger than the HBM budget");
        } catch (const Error &e) {
            r->status = e.code;
        }
        {
            std::lock_guard<std::mutex> lock(mu_);
            ++outstanding_;
            if (r->status != SFFT_OK) {
                r->t_admit = r->t_done = now();
                complete(r);
            } else {
                pending_.push_back(r);
            }
        }
        cv_.notify_all();
    }

    size_t poll(sfft_completion *out, size_t max) {
        std::lock_guard<std::mutex> lock(mu_);
        const size_t n = std::min(max, done_.size());
        for (size_t i = 0; i < n; ++i) out[i] = done_[i];
        done_.erase(done_.begin(), done_.begin() + (long)n);
        return n;
    }

    void drain() {
        std::unique_lock<std::mutex> lock(mu_);
        idle_cv_.wait(lock, [this] { return outstanding_ == 0; });
    }

    void fetch(uint64_t id, void *dst, size_t bytes) {
        std::pair<size_t, size_t> where;
        {
            std::lock_guard<std::mutex> lock(mu_);
            auto it = retained_.find(id);
            if (it == retained_.end()) throw Error(SFFT_ERR_INVALID_ARG, "no retained output");
            where = it->second;
        }
        if (bytes > where.second) throw Error(SFFT_ERR_INVALID_ARG, "fetch larger than the output");
        check_hip(hipMemcpy(dst, pool_->base() + where.first, bytes, hipMemcpyDeviceToHost),
                  "fetch");
    }

    void release(uint64_t id) {
        std::lock_guard<std::mutex> lock(mu_);
        auto it = retained_.find(id);
        if (it == retained_.end()) throw Error(SFFT_ERR_INVALID_ARG, "no retained output");
        pool_->release(it->second.first, it->second.second);
        retained_.erase(it);
        cv_.notify_all();
    }

  private:
    // One device stream, its launched-but-not-retired requests (in launch
    // order) and the buffers of its merged launches.
    struct StreamCtx {
        hipStream_t s = nullptr;
        std::deque<Req *> fifo;  // guarded by the server's mu_
        std::condition_variable cv;
        std::thread waiter;
        unsigned *d_counters = nullptr;  // persistent-launch queue counters
        char *h_block = nullptr;         // pinned staging of the work list
        char *d_block = nullptr;         // its device copy
        hipEvent_t copy_done = nullptr;  // the last staging copy has been read
        bool copied = false;
    };
    static constexpr size_t kBlockBytes =
        kMaxUnits * sizeof(WorkUnit) + 2 * kMaxGather * sizeof(int64_t) +
        kMaxFills * sizeof(FillEntry);

    void setup_device() {
        for (int i = 0; i < cfg_.streams; ++i) {
            streams_.push_back(std::make_unique<StreamCtx>());
            StreamCtx &st = *streams_.back();
            check_hip(hipStreamCreateWithFlags(&st.s, hipStreamNonBlocking), "stream create");
            if (cfg_.plain) continue;
            if (hipMalloc((void **)&st.d_counters, 2 * kPersistentClasses * sizeof(unsigned)) !=
                    hipSuccess ||
                hipMalloc((void **)&st.d_block, kBlockBytes) != hipSuccess ||
                hipHostMalloc((void **)&st.h_block, kBlockBytes, hipHostMallocDefault) !=
                    hipSuccess) {
                (void)hipGetLastError();
                throw Error(SFFT_ERR_OUT_OF_MEMORY, "serving work-list buffers");
            }
            check_hip(hipMemset(st.d_counters, 0, 2 * kPersistentClasses * sizeof(unsigned)),
                      "counter init");
            check_hip(hipEventCreateWithFlags(&st.copy_done, hipEventDisableTiming), "event");
        }
        if (cfg_.plain) return;
        const size_t eb = cfg_.precision == kF64 ? sizeof(PassArgs<double>) : sizeof(PassArgs<float>);
        if (hipMalloc(&d_table_, kMaxDescs * eb) != hipSuccess ||
            hipHostMalloc(&h_table_, kMaxDescs * eb, hipHostMallocDefault) != hipSuccess) {
            (void)hipGetLastError();
            throw Error(SFFT_ERR_OUT_OF_MEMORY, "serving descriptor table");
        }
        const size_t lds_cap = device_info().lds_bytes;
        if (cfg_.precision == kF64)
            table64_ = std::make_unique<UnitTable<double>>(lds_cap);
        else
            table32_ = std::make_unique<UnitTable<float>>(lds_cap);
    }

    void teardown_device() {
        for (auto &st : streams_) {
            if (st->s) (void)hipStreamSynchronize(st->s);
            if (st->d_counters) (void)hipFree(st->d_counters);
            if (st->d_block) (void)hipFree(st->d_block);
            if (st->h_block) (void)hipHostFree(st->h_block);
            if (st->copy_done) (void)hipEventDestroy(st->copy_done);
            if (st->s) (void)hipStreamDestroy(st->s);
        }
        streams_.clear();
        if (d_table_) (void)hipFree(d_table_);
        if (h_table_) (void)hipHostFree(h_table_);
        d_table_ = h_table_ = nullptr;
        for (auto e : event_pool_) (void)hipEventDestroy(e);
        event_pool_.clear();
    }

    hipEvent_t get_event() {
        {
            std::lock_guard<std::mutex> lock(ev_mu_);
            if (!event_pool_.empty()) {
                hipEvent_t e = event_pool_.back();
                event_pool_.pop_back();
                return e;
            }
        }
        hipEvent_t e;
        check_hip(hipEventCreateWithFlags(&e, hipEventDisableTiming), "event create");
        return e;
    }
    void put_event(hipEvent_t e) {
        std::lock_guard<std::mutex> lock(ev_mu_);
        event_pool_.push_back(e);
    }

    void complete(Req *r) {  // mu_ held
        sfft_completion c;
        c.id = r->r.id;
        c.status = r->status;
        c.t_arrive = r->t_arrive;
        c.t_admit = r->t_admit;
        c.t_done = r->t_done;
        c.bytes = r->status == SFFT_OK ? r->in_bytes + r->out_bytes : 0;
        done_.push_back(c);
        if (r->ev) put_event(r->ev);
        delete r;
        if (--outstanding_ == 0) idle_cv_.notify_all();
    }

    // Frees a finished request's buffers (keeping a retained output) and
    // records its completion.  mu_ held.
    void retire(Req *r, hipError_t q) {
        r->t_done = now();
        if (q != hipSuccess) {
            (void)hipGetLastError();
            r->status = SFFT_ERR_DEVICE;
        }
        pool_->release(r->in_off, r->in_bytes);
        if (r->ws_off != SIZE_MAX) pool_->release(r->ws_off, r->ws_bytes);
        if (r->r.retain && r->status == SFFT_OK)
            retained_[r->r.id] = {r->out_off, r->out_bytes};
        else
            pool_->release(r->out_off, r->out_bytes);
        --inflight_;
        complete(r);
    }

    // Waiter of one stream: blocks on its requests' events in launch order.
    void wait_loop(StreamCtx *st) {
        std::unique_lock<std::mutex> lock(mu_);
        while (true) {
            st->cv.wait(lock, [&] { return !st->fifo.empty() || stop_waiters_; });
            if (st->fifo.empty()) break;
            Req *r = st->fifo.front();
            lock.unlock();
            const hipError_t q = hipEventSynchronize(r->ev);
            lock.lock();
            st->fifo.pop_front();
            retire(r, q);
            cv_.notify_all();  // memory and a slot are free: admit more
        }
    }

    // Device work of one request on its own: fill, then the plan's kernels.
    void launch_single(Req *r, hipStream_t s) {
        char *in = pool_->base() + r->in_off;
        char *out = pool_->base() + r->out_off;
        void *ws = r->ws_off != SIZE_MAX ? pool_->base() + r->ws_off : nullptr;
        const size_t reals = r->in_bytes / (cfg_.precision == kF64 ? 8 : 4);
        check_hip(fill_synthetic(in, reals, cfg_.precision == kF64, 0, r->r.seed, s), "fill");
        r->plan->execute(in, r->plan->default_in_dist(r->r.direction), out,
                         r->plan->default_out_dist(r->r.direction), r->r.batch, r->r.direction,
                         s, ws, r->ws_bytes);
        r->ev = get_event();
        check_hip(hipEventRecord(r->ev, s), "event record");
    }

    // Newly appended descriptors reach the device table before any launch
    // can use them (synchronously: this happens once per plan and direction).
    template <typename T>
    void sync_table(UnitTable<T> &table) {
        const auto &d = table.descs();
        if (d.size() > kMaxDescs) throw Error(SFFT_ERR_BUDGET, "serving descriptor table full");
        if (d.size() == uploaded_descs_) return;
        PassArgs<T> *h = static_cast<PassArgs<T> *>(h_table_);
        for (size_t i = uploaded_descs_; i < d.size(); ++i) h[i] = d[i];
        check_hip(hipMemcpy(static_cast<PassArgs<T> *>(d_table_) + uploaded_descs_,
                            h + uploaded_descs_, (d.size() - uploaded_descs_) * sizeof(PassArgs<T>),
                            hipMemcpyHostToDevice),
                  "descriptor upload");
        uploaded_descs_ = d.size();
    }

    // Small requests in one persistent launch (plus one fill launch) on st.
    // Returns false if the work list does not fit (the caller runs them one
    // by one instead).
    template <typename T>
    bool launch_merged(const std::vector<Req *> &rs, StreamCtx &st, UnitTable<T> &table) {
        UnitBuilder<T> b;
        for (Req *r : rs) {
            // the table's descriptors hold the plan's device tables and are
            // keyed by its address: the plan stays alive while the table does
            table_plans_.insert(r->plan);
            b.add(table, *r->plan, r->r.kind, r->r.direction, r->r.batch, r->in_off, r->out_off);
        }
        std::vector<WorkUnit> units;
        std::vector<int64_t> gin, gout;
        size_t per_class[kPersistentClasses];
        b.finish(units, gin, gout, per_class);
        if (units.size() > kMaxUnits || gin.size() > kMaxGather || rs.size() > kMaxFills)
            return false;
        sync_table(table);
        // the staging block is rewritten only after its last copy was read
        if (st.copied) check_hip(hipEventSynchronize(st.copy_done), "staging reuse");
        char *h = st.h_block;
        const size_t ub = units.size() * sizeof(WorkUnit), gb = gin.size() * sizeof(int64_t);
        std::memcpy(h, units.data(), ub);
        std::memcpy(h + ub, gin.data(), gb);
        std::memcpy(h + ub + gb, gout.data(), gb);
        FillEntry *fe = reinterpret_cast<FillEntry *>(h + ub + 2 * gb);
        const size_t fb = rs.size() * sizeof(FillEntry);
        const int f64 = cfg_.precision == kF64;
        uint64_t max_count = 0;
        for (size_t i = 0; i < rs.size(); ++i) {
            fe[i] = FillEntry{pool_->base() + rs[i]->in_off,
                              (uint64_t)(rs[i]->in_bytes / (f64 ? 8 : 4)), rs[i]->r.seed, f64, 0};
            max_count = std::max(max_count, fe[i].count);
        }
        check_hip(hipMemcpyAsync(st.d_block, h, ub + 2 * gb + fb, hipMemcpyHostToDevice, st.s),
                  "work-list upload");
        check_hip(hipEventRecord(st.copy_done, st.s), "event record");
        st.copied = true;
        const WorkUnit *du = reinterpret_cast<const WorkUnit *>(st.d_block);
        const int64_t *dgi = reinterpret_cast<const int64_t *>(st.d_block + ub);
        const int64_t *dgo = reinterpret_cast<const int64_t *>(st.d_block + ub + gb);
        check_hip(fill_synthetic_many(reinterpret_cast<const FillEntry *>(st.d_block + ub + 2 * gb),
                                      (int)rs.size(), max_count, f64, st.s),
                  "fill");
        // one launch per kernel class, in class order
        size_t first = 0;
        for (int c = 0; c < kPersistentClasses; ++c) {
            if (!per_class[c]) continue;
            const size_t lds = table.lds_need(c);
            const unsigned grid = persistent_grid<T>(c, lds, per_class[c]);
            check_hip(launch_persistent<T>(c, du + first, (int)per_class[c],
                                           static_cast<const PassArgs<T> *>(d_table_), dgi, dgo,
                                           st.d_counters + 2 * c, pool_->base(), pool_->base(),
                                           grid, lds, st.s),
                      "merged launch");
            first += per_class[c];
        }
        for (Req *r : rs) {
            r->ev = get_event();
            check_hip(hipEventRecord(r->ev, st.s), "event record");
        }
        return true;
    }

    // Enqueues the device work of the requests admitted in one pass (worker
    // thread, mu_ not held).  Requests whose launch failed are returned in
    // `failed` after their stream has drained.
    void launch_group(std::vector<Req *> &rs, std::vector<std::pair<Req *, int>> &failed,
                      std::vector<std::pair<Req *, StreamCtx *>> &launched) {
        std::vector<Req *> merge, single;
        for (Req *r : rs) {
            const bool small = r->in_bytes + r->out_bytes <= kMergeBytes && r->ws_bytes == 0 &&
                               !r->plan->bluestein();
            (small && !cfg_.plain ? merge : single).push_back(r);
        }
        if (merge.size() < 2) {
            single.insert(single.end(), merge.begin(), merge.end());
            merge.clear();
        }
        if (!merge.empty()) {
            StreamCtx &st = *streams_[next_stream_++ % streams_.size()];
            int status = SFFT_OK;
            bool ok = false;
            try {
                ok = cfg_.precision == kF64 ? launch_merged<double>(merge, st, *table64_)
                                            : launch_merged<float>(merge, st, *table32_);
            } catch (const Error &e) {
                status = e.code;
            } catch (const std::exception &) {
                status = SFFT_ERR_DEVICE;
            }
            if (status != SFFT_OK) {
                (void)hipStreamSynchronize(st.s);  // no queued work still uses the buffers
                (void)hipGetLastError();
                for (Req *r : merge) failed.push_back({r, status});
            } else if (ok) {
                for (Req *r : merge) launched.push_back({r, &st});
            } else {
                single.insert(single.end(), merge.begin(), merge.end());
            }
        }
        for (Req *r : single) {
            StreamCtx &st = *streams_[next_stream_++ % streams_.size()];
            try {
                launch_single(r, st.s);
                if (cfg_.plain) check_hip(hipStreamSynchronize(st.s), "plain sync");
                launched.push_back({r, &st});
            } catch (const std::exception &e) {
                // part of the request may already be queued: its pool ranges
                // are released only once that work has finished
                (void)hipStreamSynchronize(st.s);
                (void)hipGetLastError();
                const auto *err = dynamic_cast<const Error *>(&e);
                failed.push_back({r, err ? err->code : SFFT_ERR_DEVICE});
            }
        }
    }

    bool try_admit(Req *r) {  // mu_ held
        const size_t a = pool_->alloc(r->in_bytes);
        if (a == SIZE_MAX) return false;
        const size_t b = pool_->alloc(r->out_bytes);
        if (b == SIZE_MAX) {
            pool_->release(a, r->in_bytes);
            return false;
        }
        size_t w = SIZE_MAX;
        if (r->ws_bytes) {
            w = pool_->alloc(r->ws_bytes);
            if (w == SIZE_MAX) {
                pool_->release(a, r->in_bytes);
                pool_->release(b, r->out_bytes);
                return false;
            }
        }
        r->in_off = a;
        r->out_off = b;
        r->ws_off = w;
        return true;
    }

    void loop() {
        std::unique_lock<std::mutex> lock(mu_);
        std::vector<Req *> group;
        std::vector<std::pair<Req *, int>> failed;
        std::vector<std::pair<Req *, StreamCtx *>> launched;
        while (true) {
            // admit in arrival order while slots and memory allow
            group.clear();
            while (!pending_.empty() && inflight_ + (int)group.size() < cfg_.max_inflight) {
                Req *r = pending_.front();
                if (!try_admit(r)) break;
                pending_.pop_front();
                r->t_admit = now();
                group.push_back(r);
            }
            if (!group.empty()) {
                inflight_ += (int)group.size();
                // the device work is enqueued without the lock: submit / poll /
                // release and the waiters are not held up
                failed.clear();
                launched.clear();
                lock.unlock();
                launch_group(group, failed, launched);
                lock.lock();
                for (auto &f : failed) {
                    f.first->status = f.second;
                    retire(f.first, hipSuccess);
                }
                for (auto &l : launched) {
                    if (cfg_.plain) {
                        retire(l.first, hipEventQuery(l.first->ev));
                    } else {
                        l.second->fifo.push_back(l.first);
                        l.second->cv.notify_one();
                    }
                }
                continue;
            }
            if (stop_ && pending_.empty() && inflight_ == 0) break;
            cv_.wait(lock);
        }
    }

    sfft_server_config cfg_;
    std::unique_ptr<Pool> pool_;
    std::vector<std::unique_ptr<StreamCtx>> streams_;
    size_t next_stream_ = 0;
    std::mutex ev_mu_;
    std::vector<hipEvent_t> event_pool_;
    // merged launches: device descriptor table (append-only) and its mirror
    void *d_table_ = nullptr, *h_table_ = nullptr;
    size_t uploaded_descs_ = 0;
    std::unique_ptr<UnitTable<float>> table32_;
    std::unique_ptr<UnitTable<double>> table64_;
    std::set<std::shared_ptr<Plan>> table_plans_;
    std::chrono::steady_clock::time_point t0_;
    std::mutex mu_;
    std::condition_variable cv_, idle_cv_;
    std::deque<Req *> pending_;
    int inflight_ = 0;
    std::vector<sfft_completion> done_;
    std::unordered_map<uint64_t, std::pair<size_t, size_t>> retained_;
    size_t outstanding_ = 0;
    bool stop_ = false, stop_waiters_ = false;
    std::thread worker_;
};

// ------------------------------------------------------- load generator --

namespace {
uint64_t mix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
// uniform [0, 1) from stream `seed`, element e
double unit(uint64_t seed, uint64_t e) {
    return (double)(mix64(mix64(seed) + (e + 1) * 0x9E3779B97F4A7C15ull) >> 11) * 0x1.0p-53;
}
}  // namespace

void run_poisson(Server &srv, const sfft_poisson_config &pc, int precision,
                 std::vector<sfft_completion> &out) {
    (void)precision;
    std::vector<size_t> fast;
    for (size_t p2 = 1; p2 <= pc.n_max; p2 *= 2)
        for (size_t p3 = p2; p3 <= pc.n_max; p3 *= 3)
            for (size_t p5 = p3; p5 <= pc.n_max; p5 *= 5)
                if (p5 >= pc.n_min) fast.push_back(p5);
    std::sort(fast.begin(), fast.end());
    if (fast.empty() || fast.back() < 2) throw Error(SFFT_ERR_INVALID_ARG, "no fast size in range");
    std::vector<double> logs;
    for (size_t f : fast) logs.push_back(std::log((double)f));
    // requests
    struct Plan_ {
        double t;
        sfft_request r;
    };
    std::vector<Plan_> reqs;
    double t = 0;
    for (uint64_t i = 0;; ++i) {
        t += -std::log(1.0 - unit(pc.seed, 6 * i)) / pc.rate;
        if (t > pc.duration) break;
        sfft_request r{};
        r.id = i;
        const double k = unit(pc.seed, 6 * i + 1);
        const double ln = std::log((double)pc.n_min) +
                          unit(pc.seed, 6 * i + 2) * (std::log((double)pc.n_max) - std::log((double)pc.n_min));
        const bool real = k >= 2.0 / 3;
        size_t best = SIZE_MAX;
        for (size_t j = 0; j < fast.size(); ++j) {
            if (real && (fast[j] % 2)) continue;  // real transforms need even sizes
            if (best == SIZE_MAX || std::fabs(logs[j] - ln) < std::fabs(logs[best] - ln)) best = j;
        }
        r.n = fast[best];
        if (k < 1.0 / 3) {
            r.kind = SFFT_C2C;
            r.direction = SFFT_FORWARD;
        } else if (k < 2.0 / 3) {
            r.kind = SFFT_C2C;
            r.direction = SFFT_INVERSE;
        } else {
            r.kind = SFFT_R2C;
            r.direction = SFFT_FORWARD;
        }
        r.batch = 1 + (size_t)(unit(pc.seed, 6 * i + 3) * (double)pc.batch_max);
        if (r.batch > pc.batch_max) r.batch = pc.batch_max;
        r.seed = i;
        r.retain = 0;
        reqs.push_back({t, r});
    }
    // plans are built before the clock starts (resident, as in a warm server)
    for (auto &q : reqs) {
        try {
            (void)get_plan(q.r.n, q.r.kind, precision);
        } catch (const Error &) {
        }
    }
    const double start = srv.now();
    for (auto &q : reqs) {
        const double at = start + q.t;
        while (srv.now() < at) {
            const double gap = at - srv.now();
            if (gap > 200e-6) std::this_thread::sleep_for(std::chrono::duration<double>(gap - 100e-6));
        }
        srv.submit(q.r, at);  // arrival stamped at its scheduled time
    }
    srv.drain();
    std::vector<sfft_completion> buf(4096);
    out.clear();
    while (true) {
        const size_t n = srv.poll(buf.data(), buf.size());
        if (!n) break;
        out.insert(out.end(), buf.begin(), buf.begin() + (long)n);
    }
    for (auto &c : out) {
        c.t_arrive -= start;
        c.t_admit -= start;
        c.t_done -= start;
    }
}

}  // namespace sfft

// ------------------------------------------------------------------ C ABI --

#include <string>

namespace {
template <typename F>
int srv_guard(F &&f) {
    try {
        sfft::last_error().clear();
        return f();
    } catch (const sfft::Error &e) {
        sfft::last_error() = e.what();
        return e.code;
    } catch (const std::exception &e) {
        sfft::last_error() = e.what();
        return SFFT_ERR_DEVICE;
    }
}
}  // namespace

struct sfft_server_s {
    std::unique_ptr<sfft::Server> impl;
    int precision;
};

extern "C" {

int sfft_server_create(sfft_server *server, const sfft_server_config *config) {
    return srv_guard([&] {
        if (!server) return SFFT_ERR_INVALID_ARG;
        sfft_server_config c{};
        if (config) c = *config;
        auto *s = new sfft_server_s;
        s->precision = c.precision;
        try {
            s->impl = std::make_unique<sfft::Server>(c);
        } catch (...) {
            delete s;
            throw;
        }
        *server = s;
        return SFFT_OK;
    });
}

int sfft_server_destroy(sfft_server server) {
    delete server;
    return SFFT_OK;
}

int sfft_server_submit(sfft_server server, const sfft_request *request) {
    return srv_guard([&] {
        if (!server || !request) return SFFT_ERR_INVALID_ARG;
        server->impl->submit(*request, server->impl->now());
        return SFFT_OK;
    });
}

int sfft_server_poll(sfft_server server, sfft_completion *out, size_t max, size_t *count) {
    return srv_guard([&] {
        if (!server || !count || (!out && max)) return SFFT_ERR_INVALID_ARG;
        *count = server->impl->poll(out, max);
        return SFFT_OK;
    });
}

int sfft_server_drain(sfft_server server) {
    return srv_guard([&] {
        if (!server) return SFFT_ERR_INVALID_ARG;
        server->impl->drain();
        return SFFT_OK;
    });
}

int sfft_server_fetch(sfft_server server, uint64_t id, void *host_dst, size_t bytes) {
    return srv_guard([&] {
        if (!server || (!host_dst && bytes)) return SFFT_ERR_INVALID_ARG;
        server->impl->fetch(id, host_dst, bytes);
        return SFFT_OK;
    });
}

int sfft_server_release(sfft_server server, uint64_t id) {
    return srv_guard([&] {
        if (!server) return SFFT_ERR_INVALID_ARG;
        server->impl->release(id);
        return SFFT_OK;
    });
}

double sfft_server_now(sfft_server server) { return server ? server->impl->now() : 0.0; }

int sfft_server_run_poisson(sfft_server server, const sfft_poisson_config *config,
                            sfft_completion *records, size_t max_records, size_t *count) {
    return srv_guard([&] {
        if (!server || !config || !count || config->rate <= 0 || config->duration <= 0 ||
            config->n_min < 1 || config->n_max < config->n_min || config->batch_max < 1)
            return SFFT_ERR_INVALID_ARG;
        std::vector<sfft_completion> out;
        sfft::run_poisson(*server->impl, *config, server->precision, out);
        const size_t n = std::min(max_records, out.size());
        for (size_t i = 0; i < n; ++i) records[i] = out[i];
        *count = out.size();
        return SFFT_OK;
    });
}

}  // extern "C"
