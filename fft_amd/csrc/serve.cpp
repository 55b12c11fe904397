// serve.cpp -- serving loop: continuous batching of transform requests.
//
// A worker thread owns the device side.  Each tick it
//   1. retires requests whose completion event has fired (frees their
//      buffers unless retained, records the completion time),
//   2. admits queued requests in arrival order while an in-flight slot is
//      free and their input + output buffers fit in the HBM pool (the head
//      of the queue blocks the rest: no reordering, no starvation),
//   3. enqueues each admitted request on the next device stream: synthesise
//      the input from its seed, run the plan, record an event.
// Admitted requests therefore overlap on the device, new arrivals are
// admitted as soon as slots and memory free up, and plans / twiddles stay
// resident in the plan cache.  The plain mode serves one request at a time
// synchronously (the comparison path).
#include <sfft.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <deque>
#include <map>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

#include "engine.h"

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
    size_t in_bytes = 0, out_bytes = 0, in_off = SIZE_MAX, out_off = SIZE_MAX;
    std::shared_ptr<Plan> plan;
    hipEvent_t ev = nullptr;
    int status = SFFT_OK;
};

}  // namespace

class Server {
  public:
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
        for (int i = 0; i < cfg_.streams; ++i) {
            hipStream_t s;
            check_hip(hipStreamCreateWithFlags(&s, hipStreamNonBlocking), "stream create");
            streams_.push_back(s);
        }
        t0_ = std::chrono::steady_clock::now();
        worker_ = std::thread([this] { loop(); });
    }

    ~Server() {
        {
            std::lock_guard<std::mutex> lock(mu_);
            stop_ = true;
        }
        cv_.notify_all();
        worker_.join();
        for (auto &kv : retained_) pool_->release(kv.second.first, kv.second.second);
        for (Req *r : inflight_) finish_wait(r);
        for (auto s : streams_) (void)hipStreamDestroy(s);
        for (auto e : event_pool_) (void)hipEventDestroy(e);
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
            const size_t need = ((r->in_bytes + 255) & ~size_t(255)) + ((r->out_bytes + 255) & ~size_t(255));
            if (need > pool_->size()) throw Error(SFFT_ERR_BUDGET, "request larger than the HBM budget");
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
    hipEvent_t get_event() {
        if (!event_pool_.empty()) {
            hipEvent_t e = event_pool_.back();
            event_pool_.pop_back();
            return e;
        }
        hipEvent_t e;
        check_hip(hipEventCreateWithFlags(&e, hipEventDisableTiming), "event create");
        return e;
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
        if (r->ev) event_pool_.push_back(r->ev);
        delete r;
        if (--outstanding_ == 0) idle_cv_.notify_all();
    }

    void finish_wait(Req *r) {
        if (r->ev) (void)hipEventSynchronize(r->ev);
    }

    // enqueue the device work of an admitted request (worker thread, mu_ not
    // held: uses only the streams, the event pool and the request)
    void launch(Req *r) {
        hipStream_t s = streams_[next_stream_++ % streams_.size()];
        char *in = pool_->base() + r->in_off;
        char *out = pool_->base() + r->out_off;
        const size_t reals = r->in_bytes / (cfg_.precision == kF64 ? 8 : 4);
        try {
            check_hip(fill_synthetic(in, reals, cfg_.precision == kF64, 0, r->r.seed, s),
                      "fill");
            r->plan->execute(in, r->plan->default_in_dist(r->r.direction), out,
                             r->plan->default_out_dist(r->r.direction), r->r.batch,
                             r->r.direction, s);
            r->ev = get_event();
            check_hip(hipEventRecord(r->ev, s), "event record");
        } catch (...) {
            // part of the request may already be queued on s: its pool ranges
            // are released by the caller only once that work has finished
            (void)hipStreamSynchronize(s);
            (void)hipGetLastError();
            throw;
        }
        if (cfg_.plain) check_hip(hipStreamSynchronize(s), "plain sync");
    }

    void loop() {
        std::unique_lock<std::mutex> lock(mu_);
        while (true) {
            bool progress = false;
            // 1. retire
            for (size_t i = 0; i < inflight_.size();) {
                Req *r = inflight_[i];
                const hipError_t q = hipEventQuery(r->ev);
                if (q == hipErrorNotReady) {
                    ++i;
                    continue;
                }
                r->t_done = now();
                if (q != hipSuccess) {
                    (void)hipGetLastError();
                    r->status = SFFT_ERR_DEVICE;
                }
                pool_->release(r->in_off, r->in_bytes);
                if (r->r.retain && r->status == SFFT_OK)
                    retained_[r->r.id] = {r->out_off, r->out_bytes};
                else
                    pool_->release(r->out_off, r->out_bytes);
                inflight_[i] = inflight_.back();
                inflight_.pop_back();
                complete(r);
                progress = true;
            }
            // 2. admit in arrival order
            while (!pending_.empty() && (int)inflight_.size() < cfg_.max_inflight) {
                Req *r = pending_.front();
                const size_t a = pool_->alloc(r->in_bytes);
                if (a == SIZE_MAX) break;
                const size_t b = pool_->alloc(r->out_bytes);
                if (b == SIZE_MAX) {
                    pool_->release(a, r->in_bytes);
                    break;
                }
                pending_.pop_front();
                r->in_off = a;
                r->out_off = b;
                r->t_admit = now();
                // The device work is enqueued (and, in plain mode, waited
                // for) without the lock: launch() touches only this thread's
                // state, so submit / poll / release are not held up.
                int status = SFFT_OK;
                lock.unlock();
                try {
                    launch(r);
                } catch (const Error &e) {
                    status = e.code;
                } catch (const std::exception &) {
                    status = SFFT_ERR_DEVICE;
                }
                lock.lock();
                if (status == SFFT_OK) {
                    inflight_.push_back(r);
                } else {
                    pool_->release(a, r->in_bytes);
                    pool_->release(b, r->out_bytes);
                    r->status = status;
                    r->t_done = now();
                    complete(r);
                }
                progress = true;
            }
            if (stop_ && pending_.empty() && inflight_.empty()) break;
            if (!progress) {
                if (inflight_.empty() && pending_.empty())
                    cv_.wait(lock);
                else
                    cv_.wait_for(lock, std::chrono::microseconds(20));
            }
        }
    }

    sfft_server_config cfg_;
    std::unique_ptr<Pool> pool_;
    std::vector<hipStream_t> streams_;
    size_t next_stream_ = 0;
    std::vector<hipEvent_t> event_pool_;
    std::chrono::steady_clock::time_point t0_;
    std::mutex mu_;
    std::condition_variable cv_, idle_cv_;
    std::deque<Req *> pending_;
    std::vector<Req *> inflight_;
    std::vector<sfft_completion> done_;
    std::unordered_map<uint64_t, std::pair<size_t, size_t>> retained_;
    size_t outstanding_ = 0;
    bool stop_ = false;
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
