// tune.cpp -- the autotuner: times the planner's candidates for one
// (length, kind, precision) on synthetic data resident in HBM, keeps the
// fastest as the length's plan, and saves / loads the choices as JSON keyed
// by the GPU architecture and the HIP runtime version.
#include <sfft.h>

#include <algorithm>
#include <cctype>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <map>
#include <mutex>
#include <sstream>
#include <string>
#include <tuple>
#include <vector>

#include "engine.h"
#include "tune.h"

namespace sfft {

namespace {

std::mutex g_mu;
std::map<std::tuple<size_t, int, int>, TuneResult> g_tuned;

int hip_runtime_version() {
    int v = 0;
    if (hipRuntimeGetVersion(&v) != hipSuccess) v = 0;
    return v;
}

struct DevBuf {
    void *p = nullptr;
    explicit DevBuf(size_t bytes) {
        if (hipMalloc(&p, bytes ? bytes : 1) != hipSuccess) {
            (void)hipGetLastError();
            throw Error(SFFT_ERR_OUT_OF_MEMORY, "autotune buffer allocation failed");
        }
    }
    ~DevBuf() { (void)hipFree(p); }
};

struct Stream {
    hipStream_t s = nullptr;
    Stream() { check_hip(hipStreamCreateWithFlags(&s, hipStreamNonBlocking), "stream"); }
    ~Stream() { (void)hipStreamDestroy(s); }
};

struct Events {
    hipEvent_t a = nullptr, b = nullptr;
    Events() {
        check_hip(hipEventCreate(&a), "event");
        check_hip(hipEventCreate(&b), "event");
    }
    ~Events() {
        (void)hipEventDestroy(a);
        (void)hipEventDestroy(b);
    }
};

}  // namespace

bool tuned_hints(size_t n, int kind, int prec, PlanHints &out) {
    std::lock_guard<std::mutex> lock(g_mu);
    auto it = g_tuned.find(std::make_tuple(n, kind, prec));
    if (it == g_tuned.end()) return false;
    out = it->second.hints;
    return true;
}

TuneResult autotune(size_t n, int kind, int prec, size_t batch, int reps) {
    const std::vector<PlanHints> cands = plan_candidates(n, kind, prec);
    Plan probe(n, kind, prec);  // validates the arguments, sizes the buffers
    if (batch == 0) {
        // about 1 GB of input + output per launch: well beyond the 256 MB
        // memory-side cache, so the timing is HBM's
        const size_t per = probe.in_bytes(1, 0) + probe.out_bytes(1, 0);
        batch = std::max<size_t>(1, (1ull << 30) / per);
    }
    if (reps < 1) reps = 5;
    const size_t in_b = std::max(probe.in_bytes(batch, 0), probe.in_bytes(batch, 1));
    const size_t out_b = std::max(probe.out_bytes(batch, 0), probe.out_bytes(batch, 1));
    DevBuf x(in_b), y(out_b);
    Stream st;
    Events ev;
    const int f64 = prec == kF64;
    if (stored_16(prec))  // 16-bit values: zeros (the timing does not depend on them)
        check_hip(hipMemsetAsync(x.p, 0, in_b, st.s), "clear");
    else
        check_hip(fill_synthetic(x.p, in_b / (f64 ? 8 : 4), f64, 0, 0x5EEDull, st.s), "fill");
    TuneResult best;
    best.candidates = 0;
    best.ms = 1e30;
    for (const PlanHints &h : cands) {
        std::unique_ptr<Plan> p;
        try {
            p.reset(new Plan(n, kind, prec, &h));
        } catch (const Error &) {
            continue;  // a candidate that does not fit this length
        }
        double total = 0;
        for (int dir = 0; dir < 2; ++dir) {
            p->execute(x.p, p->default_in_dist(dir), y.p, p->default_out_dist(dir), batch, dir,
                       st.s);
            check_hip(hipEventRecord(ev.a, st.s), "event");
            for (int r = 0; r < reps; ++r)
                p->execute(x.p, p->default_in_dist(dir), y.p, p->default_out_dist(dir), batch,
                           dir, st.s);
            check_hip(hipEventRecord(ev.b, st.s), "event");
            check_hip(hipEventSynchronize(ev.b), "event sync");
            float ms = 0;
            check_hip(hipEventElapsedTime(&ms, ev.a, ev.b), "elapsed");
            total += ms / reps;
        }
        ++best.candidates;
        const double ms = total / 2;
        if (ms < best.ms) {
            best.ms = ms;
            best.hints = h;
        }
    }
    if (!best.candidates) throw Error(SFFT_ERR_UNSUPPORTED_SIZE, "no candidate plan");
    const double bytes = (double)(probe.in_bytes(batch, 0) + probe.out_bytes(batch, 0));
    best.gbs = bytes / (best.ms * 1e-3) / 1e9;
    best.batch = batch;
    {
        std::lock_guard<std::mutex> lock(g_mu);
        g_tuned[std::make_tuple(n, kind, prec)] = best;
    }
    drop_cached_plan(n, kind, prec);
    return best;
}

// Cached plans are dropped after g_mu is released: get_plan takes the plan
// cache lock first and then g_mu (through tuned_hints).
void tune_clear() {
    std::vector<std::tuple<size_t, int, int>> keys;
    {
        std::lock_guard<std::mutex> lock(g_mu);
        for (auto &e : g_tuned) keys.push_back(e.first);
        g_tuned.clear();
    }
    for (auto &k : keys) drop_cached_plan(std::get<0>(k), std::get<1>(k), std::get<2>(k));
}

size_t tune_count() {
    std::lock_guard<std::mutex> lock(g_mu);
    return g_tuned.size();
}

// ------------------------------------------------------------------ JSON --

void tune_save(const std::string &path) {
    std::ostringstream o;
    o << "{\n \"format\": \"sfft-tune-1\",\n \"arch\": \"" << device_info().arch
      << "\",\n \"hip_runtime\": " << hip_runtime_version() << ",\n \"entries\": [";
    std::lock_guard<std::mutex> lock(g_mu);
    bool first = true;
    for (const auto &e : g_tuned) {
        const TuneResult &t = e.second;
        o << (first ? "\n" : ",\n") << "  {\"n\": " << std::get<0>(e.first)
          << ", \"kind\": " << std::get<1>(e.first) << ", \"precision\": " << std::get<2>(e.first)
          << ", \"variant\": " << t.hints.fixed_variant << ", \"radices\": [";
        for (size_t i = 0; i < t.hints.radices.size(); ++i)
            o << (i ? ", " : "") << t.hints.radices[i];
        o << "], \"tpw\": " << t.hints.tpw << ", \"ept\": " << t.hints.ept
          << ", \"split\": " << t.hints.split << ", \"mfma\": " << t.hints.mfma
          << ", \"ms\": " << t.ms << ", \"gbs\": " << t.gbs << "}";
        first = false;
    }
    o << "\n ]\n}\n";
    std::ofstream f(path);
    if (!f) throw Error(SFFT_ERR_INVALID_ARG, "cannot write " + path);
    f << o.str();
    if (!f) throw Error(SFFT_ERR_INVALID_ARG, "write failed: " + path);
}

namespace {

// Minimal reader for the format written above: objects, arrays, numbers,
// strings.  Values are kept as text or as nested nodes.
struct Node {
    enum { kNum, kStr, kArr, kObj } type = kNum;
    double num = 0;
    std::string str;
    std::vector<Node> arr;
    std::vector<std::pair<std::string, Node>> obj;
    const Node *get(const std::string &k) const {
        for (const auto &kv : obj)
            if (kv.first == k) return &kv.second;
        return nullptr;
    }
};

struct Parser {
    const std::string &s;
    size_t i = 0;
    explicit Parser(const std::string &t) : s(t) {}
    [[noreturn]] void fail(const char *what) const {
        throw Error(SFFT_ERR_INVALID_ARG,
                    std::string("tuning file: ") + what + " at byte " + std::to_string(i));
    }
    void ws() {
        while (i < s.size() && std::isspace((unsigned char)s[i])) ++i;
    }
    bool eat(char c) {
        ws();
        if (i < s.size() && s[i] == c) {
            ++i;
            return true;
        }
        return false;
    }
    std::string string() {
        if (!eat('"')) fail("expected string");
        std::string out;
        while (i < s.size() && s[i] != '"') {
            if (s[i] == '\\' && i + 1 < s.size()) ++i;
            out += s[i++];
        }
        if (i >= s.size()) fail("unterminated string");
        ++i;
        return out;
    }
    Node value(int depth = 0) {
        if (depth > 8) fail("nesting too deep");
        ws();
        Node v;
        if (i >= s.size()) fail("unexpected end");
        if (s[i] == '{') {
            ++i;
            v.type = Node::kObj;
            if (eat('}')) return v;
            do {
                std::string k = string();
                if (!eat(':')) fail("expected ':'");
                v.obj.push_back({k, value(depth + 1)});
            } while (eat(','));
            if (!eat('}')) fail("expected '}'");
        } else if (s[i] == '[') {
            ++i;
            v.type = Node::kArr;
            if (eat(']')) return v;
            do v.arr.push_back(value(depth + 1));
            while (eat(','));
            if (!eat(']')) fail("expected ']'");
        } else if (s[i] == '"') {
            v.type = Node::kStr;
            v.str = string();
        } else {
            char *end = nullptr;
            v.num = std::strtod(s.c_str() + i, &end);
            if (end == s.c_str() + i) fail("expected a value");
            i = (size_t)(end - s.c_str());
        }
        return v;
    }
};

long num_field(const Node &e, const char *k, long dflt) {
    const Node *v = e.get(k);
    return (v && v->type == Node::kNum) ? (long)v->num : dflt;
}

}  // namespace

size_t tune_load(const std::string &path, bool check_device) {
    std::ifstream f(path);
    if (!f) throw Error(SFFT_ERR_INVALID_ARG, "cannot read " + path);
    std::stringstream buf;
    buf << f.rdbuf();
    const std::string text = buf.str();
    Parser ps(text);
    const Node root = ps.value();
    if (root.type != Node::kObj) throw Error(SFFT_ERR_INVALID_ARG, "tuning file: not an object");
    const Node *fmt = root.get("format");
    if (!fmt || fmt->type != Node::kStr || fmt->str != "sfft-tune-1")
        throw Error(SFFT_ERR_INVALID_ARG, "tuning file: unknown format");
    if (check_device) {
        const Node *arch = root.get("arch");
        if (!arch || arch->str != device_info().arch ||
            num_field(root, "hip_runtime", -1) != hip_runtime_version())
            return 0;  // tuned on another device or runtime: ignored
    }
    const Node *entries = root.get("entries");
    if (!entries || entries->type != Node::kArr)
        throw Error(SFFT_ERR_INVALID_ARG, "tuning file: no entries");
    std::vector<std::pair<std::tuple<size_t, int, int>, TuneResult>> loaded;
    for (const Node &e : entries->arr) {
        if (e.type != Node::kObj) throw Error(SFFT_ERR_INVALID_ARG, "tuning file: bad entry");
        const long n = num_field(e, "n", 0);
        const long kind = num_field(e, "kind", -1), prec = num_field(e, "precision", -1);
        if (n < 1 || kind < 0 || kind > 2 || prec < 0 || prec > kF32X)
            throw Error(SFFT_ERR_INVALID_ARG, "tuning file: bad key");
        TuneResult t;
        t.hints.fixed_variant = (int)num_field(e, "variant", -1);
        if (const Node *r = e.get("radices")) {
            if (r->type != Node::kArr || r->arr.size() > (size_t)kMaxStages)
                throw Error(SFFT_ERR_INVALID_ARG, "tuning file: bad radices");
            for (const Node &x : r->arr) t.hints.radices.push_back((int)x.num);
        }
        t.hints.tpw = (int)num_field(e, "tpw", 0);
        t.hints.ept = (int)num_field(e, "ept", 0);
        t.hints.split = (size_t)num_field(e, "split", 0);
        t.hints.mfma = (int)num_field(e, "mfma", -1);
        if (const Node *v = e.get("ms")) t.ms = v->num;
        if (const Node *v = e.get("gbs")) t.gbs = v->num;
        loaded.push_back({std::make_tuple((size_t)n, (int)kind, (int)prec), t});
    }
    {
        std::lock_guard<std::mutex> lock(g_mu);
        for (auto &kv : loaded) g_tuned[kv.first] = kv.second;
    }
    for (auto &kv : loaded)
        drop_cached_plan(std::get<0>(kv.first), std::get<1>(kv.first), std::get<2>(kv.first));
    return loaded.size();
}

}  // namespace sfft
