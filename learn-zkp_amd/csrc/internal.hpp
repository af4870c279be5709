// Shared internals of the host runtime: error handling, device buffers,
// kernel timing, the opaque handle types of the C ABI and the helpers the
// plain, batched and serving paths have in common.
#pragma once
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <array>
#include <atomic>
#include <condition_variable>
#include <exception>
#include <functional>
#include <chrono>
#include <thread>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/bfd.h"
#include "gf128.hpp"
#include "host.hpp"
#include "kernels.hpp"

using namespace bfd;

// ---------------------------------------------------------------------------
// errors
// ---------------------------------------------------------------------------
inline thread_local std::string g_err;

namespace bfd {
namespace rt {

struct Error {
    int code;
    std::string msg;
};

[[noreturn]] inline void fail(int code, const std::string &msg) { throw Error{code, msg}; }

inline void hip_check(hipError_t e, const char *what) {
    if (e != hipSuccess) fail(BFD_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

template <class F>
int guarded(F &&f) {
    try {
        f();
        return BFD_OK;
    } catch (const Error &e) {
        g_err = e.msg;
        return e.code;
    } catch (const std::exception &e) {
        g_err = e.what();
        return BFD_ERR_ARG;
    } catch (...) {
        g_err = "unknown error";
        return BFD_ERR_ARG;
    }
}

// ---------------------------------------------------------------------------
// device memory: stream-ordered allocations from the device's default pool
// ---------------------------------------------------------------------------
// Freed blocks of one stream kept for reuse: a block waits here for the next
// request of exactly its size on the same stream, so
// repeated commits, opens and sumchecks of one shape make no allocator calls
// (a stream-ordered free costs the calling thread 15-30 us, even for the few
// kilobytes of a transcript image).  Reuse is safe by stream order (the
// block's last use and its next use are on the same stream).  Holds at most
// `cap` bytes; a block that does not fit empties the cache first.
struct BlockCache {
    static constexpr size_t kMin = 1;
    std::mutex m;
    std::multimap<size_t, void *> blocks;
    size_t held = 0, cap = 0;
    hipStream_t st = nullptr;
    void *take(size_t n) {
        if (n < kMin) return nullptr;
        std::lock_guard<std::mutex> lk(m);
        auto it = blocks.find(n);
        if (it == blocks.end()) return nullptr;
        void *p = it->second;
        blocks.erase(it);
        held -= n;
        return p;
    }
    // false when the caller must free the block itself
    bool put(void *p, size_t n) {
        if (n < kMin || n > cap) return false;
        std::lock_guard<std::mutex> lk(m);
        if (held + n > cap) drop_locked();
        blocks.emplace(n, p);
        held += n;
        return true;
    }
    void flush() {
        std::lock_guard<std::mutex> lk(m);
        drop_locked();
    }
    // the owner flushes while the stream still exists; this is a fallback
    ~BlockCache() {
        for (auto &kv : blocks) (void)hipFree(kv.second);
    }

  private:
    void drop_locked() {
        for (auto &kv : blocks) (void)hipFreeAsync(kv.second, st);
        blocks.clear();
        held = 0;
    }
};

struct DevBuf {
    void *p = nullptr;
    size_t bytes = 0;
    hipStream_t st = nullptr;
    BlockCache *cache = nullptr;
    DevBuf() = default;
    DevBuf(size_t n, hipStream_t s, BlockCache *c = nullptr) : bytes(n), st(s), cache(c) {
        if (!n) return;
        if (c && c->st == s && (p = c->take(n))) return;
        hipError_t e = hipMallocAsync(&p, n, s);
        if (e != hipSuccess && c) {  // cached blocks may be what is missing
            (void)hipGetLastError();
            c->flush();
            e = hipMallocAsync(&p, n, s);
        }
        hip_check(e, "hipMallocAsync");
    }
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    DevBuf(DevBuf &&o) noexcept { *this = std::move(o); }
    DevBuf &operator=(DevBuf &&o) noexcept {
        release();
        p = o.p; bytes = o.bytes; st = o.st; cache = o.cache;
        o.p = nullptr; o.bytes = 0;
        return *this;
    }
    ~DevBuf() { release(); }
    void release() {
        if (p && !(cache && cache->st == st && cache->put(p, bytes))) (void)hipFreeAsync(p, st);
        p = nullptr;
        bytes = 0;
    }
    template <class T>
    T *as() const { return reinterpret_cast<T *>(p); }
};

struct PinnedBuf {
    void *p = nullptr;
    size_t bytes = 0;
    void ensure(size_t n) {
        if (n <= bytes) return;
        if (p) (void)hipHostFree(p);
        p = nullptr;
        hip_check(hipHostMalloc(&p, n, hipHostMallocDefault), "hipHostMalloc");
        bytes = n;
    }
    ~PinnedBuf() {
        if (p) (void)hipHostFree(p);
    }
};

inline double now_ms() {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// Optional per-kernel timing with HIP events on the context's stream.
struct KernelTimer {
    bool enabled = false;
    // when set, only launches of this name are timed, and a run of
    // consecutive launches of it is bracketed by one pair of events (closed
    // by the next timed launch of another name): every event recorded
    // between launches costs the stream about 5 us, so a timed run that
    // needs one kernel's average duration brackets runs of that kernel alone
    std::string only;
    struct Rec {
        std::string name;
        hipEvent_t a, b;
        uint32_t count;  // launches between a and b
    };
    std::vector<Rec> pending;
    std::vector<hipEvent_t> pool;
    std::map<std::string, std::pair<double, uint32_t>> totals;
    bool run_open = false;
    Rec cur{};
    hipStream_t run_st = nullptr;
    hipEvent_t get() {
        if (!pool.empty()) {
            hipEvent_t e = pool.back();
            pool.pop_back();
            return e;
        }
        hipEvent_t e;
        hip_check(hipEventCreate(&e), "hipEventCreate");
        return e;
    }
    void close_run() {
        if (!run_open) return;
        cur.b = get();
        hip_check(hipEventRecord(cur.b, run_st), "hipEventRecord");
        pending.push_back(cur);
        run_open = false;
    }
    template <class F>
    void run(hipStream_t st, const char *name, F &&launch) {
        if (!enabled) {
            hip_check(launch(), name);
            return;
        }
        if (!only.empty()) {
            if (only != name) {
                close_run();
                hip_check(launch(), name);
                return;
            }
            if (run_open && run_st != st) close_run();
            if (!run_open) {
                cur = Rec{name, get(), nullptr, 0};
                hip_check(hipEventRecord(cur.a, st), "hipEventRecord");
                run_open = true;
                run_st = st;
            }
            hip_check(launch(), name);
            cur.count++;
            return;
        }
        Rec r{name, get(), get(), 1};
        hip_check(hipEventRecord(r.a, st), "hipEventRecord");
        hip_check(launch(), name);
        hip_check(hipEventRecord(r.b, st), "hipEventRecord");
        pending.push_back(r);
    }
    // host-side phase time, reported next to the kernels as "host:<name>"
    void host(const char *name, double ms) {
        if (!enabled) return;
        auto &t = totals[std::string("host:") + name];
        t.first += ms;
        t.second += 1;
    }
    // adds the launches that have completed; launches still in flight (work
    // enqueued after the synchronisation the caller waited for) stay pending
    void collect() {
        // a run still open here ended right before the caller's
        // synchronisation, with no event behind it: it is not timed, and
        // counted under "untimed:<name>" so the caller sees the launches it
        // is missing
        if (run_open) {
            auto &t = totals[std::string("untimed:") + cur.name];
            t.second += cur.count;
            pool.push_back(cur.a);
            run_open = false;
        }
        std::vector<Rec> later;
        for (auto &r : pending) {
            if (hipEventQuery(r.b) != hipSuccess) {
                later.push_back(r);
                continue;
            }
            float ms = 0;
            hip_check(hipEventElapsedTime(&ms, r.a, r.b), "hipEventElapsedTime");
            auto &t = totals[r.name];
            t.first += ms;
            t.second += r.count;
            pool.push_back(r.a);
            pool.push_back(r.b);
        }
        pending.swap(later);
    }
    ~KernelTimer() {
        for (auto &r : pending) { (void)hipEventDestroy(r.a); (void)hipEventDestroy(r.b); }
        for (auto e : pool) (void)hipEventDestroy(e);
    }
};

// ---------------------------------------------------------------------------
// launch shapes (bfd_tuning; chosen per codeword size by tune.cpp)
// ---------------------------------------------------------------------------
constexpr uint32_t kTuneSizes = 48;  // log2_m = 0..47

inline bfd_tuning default_tuning() {
    bfd_tuning t;
    memset(&t, 0, sizeof t);
    t.ntt_order = 0;
    t.ntt_block = 256;
    t.ntt_waves_per_cu = 32;  // 8 blocks of 256 per compute unit
    t.ntt_occupancy = 2;
    t.hash_block = 256;
    t.hash_waves_per_cu = 64;  // 16 blocks of 256 per compute unit
    t.subtree_from_log2 = 17;
    t.subtree_levels = 9;
    t.seg_blocks_per_cu = 8;
    return t;
}

// empty string when t is a valid shape, else what is wrong with it
inline std::string tuning_problem(const bfd_tuning &t) {
    const auto blk = [](uint32_t b) { return b == 64 || b == 128 || b == 256; };
    if (t.ntt_order > 2) return "ntt_order must be 0, 1 or 2";
    if (!blk(t.ntt_block)) return "ntt_block must be 64, 128 or 256";
    if (t.ntt_waves_per_cu < 4 || t.ntt_waves_per_cu > 1024) return "ntt_waves_per_cu must be 4..1024";
    if (t.ntt_occupancy != 2 && t.ntt_occupancy != 3) return "ntt_occupancy must be 2 or 3";
    if (!blk(t.hash_block)) return "hash_block must be 64, 128 or 256";
    if (t.hash_waves_per_cu < 4 || t.hash_waves_per_cu > 1024) return "hash_waves_per_cu must be 4..1024";
    if (t.subtree_from_log2 < 1 || t.subtree_from_log2 > 40) return "subtree_from_log2 must be 1..40";
    if (t.subtree_levels < 1 || t.subtree_levels > (uint32_t)launch::subtree_max_levels())
        return "subtree_levels must be 1..9";
    if (t.seg_blocks_per_cu < 1 || t.seg_blocks_per_cu > 64) return "seg_blocks_per_cu must be 1..64";
    for (uint32_t r : t.reserved)
        if (r) return "reserved fields must be zero";
    return "";
}

inline bool tuning_equal(const bfd_tuning &a, const bfd_tuning &b) { return memcmp(&a, &b, sizeof a) == 0; }

// grid caps in blocks for a device of `cus` compute units
inline launch::Shape shape_of(const bfd_tuning &t, int cus) {
    launch::Shape sh;
    sh.ntt_block = (int)t.ntt_block;
    sh.ntt_cap = std::max<size_t>(1, (size_t)cus * t.ntt_waves_per_cu * 64 / t.ntt_block);
    sh.ntt_waves = (int)t.ntt_occupancy;
    sh.hash_block = (int)t.hash_block;
    sh.hash_cap = std::max<size_t>(1, (size_t)cus * t.hash_waves_per_cu * 64 / t.hash_block);
    return sh;
}

// The NTT passes over layers 1..n-1 as (first layer, layers in the pass)
// pairs: order 0 runs three layers per pass with the shorter remainder pass
// last, order 1 the remainder first, order 2 two layers per pass (and one
// single layer when n-1 is odd).  Any order applies every layer once, in
// increasing order, which is all the data-order NTT needs.
inline void ntt_passes(int n, uint32_t order, std::vector<std::pair<int, int>> &out) {
    out.clear();
    const int L = n > 1 ? n - 1 : 0;
    if (order == 2) {
        for (int s = 1; s <= L; s += 2) out.push_back({s, std::min(2, L - s + 1)});
        return;
    }
    const int rem = L % 3;
    int s = 1;
    if (order == 1 && rem) {
        out.push_back({s, rem});
        s += rem;
    }
    for (; s + 3 <= L + 1; s += 3) out.push_back({s, 3});
    if (order != 1 && rem) out.push_back({s, rem});
}

}  // namespace rt
}  // namespace bfd

// ---------------------------------------------------------------------------
// handles
// ---------------------------------------------------------------------------
struct bfd_code;
namespace bfd {
namespace rt {
class Lane;
struct LaneDeleter {
    void operator()(Lane *l) const;  // engine.cpp
};
}  // namespace rt
}  // namespace bfd

struct bfd_ctx {
    int device = 0;
    // device address of the allocation-free entry points' scratch array on
    // this device (noalloc.cpp), taken when the context is created
    gf *scratch = nullptr;
    std::string scratch_error;
    // codes made on demand by the batched and serving paths, keyed by
    // (lambda, log2_c, d); owned by the context
    std::map<uint64_t, bfd_code *> code_cache;
    hipStream_t stream = nullptr;
    uint64_t budget = 0;
    uint64_t resident = 0;
    std::recursive_mutex mu;
    bfd::rt::PinnedBuf pinned;      // per-round read-backs (roots, partial sums)
    bfd::rt::PinnedBuf gather_pin;  // opened values of a proof, device -> host
    bfd::rt::PinnedBuf seg_pin;     // segment tables of batched launches, host -> device
    bfd::rt::PinnedBuf eq_pin;      // eq-table halves, host -> device
    bfd::rt::PinnedBuf scalar_pin;  // random scalars of a width > 1 opening, host -> device
    bfd::rt::PinnedBuf verify_pin;  // proofs of a verification batch, host -> device
    bfd::rt::KernelTimer timer;
    // blocks of the plain path on `stream`, reused across calls
    bfd::rt::BlockCache cache;
    // engine lane on `stream` for batched calls (made on first use)
    std::unique_ptr<bfd::rt::Lane, bfd::rt::LaneDeleter> lane;
    // launch shapes: tuning[log2 of the codeword length] for the plain path,
    // batch_tuning for the engine (batched calls and servers made later)
    int cus = 256;
    std::array<bfd_tuning, bfd::rt::kTuneSizes> tuning;
    bfd_tuning batch_tuning;
    const bfd_tuning &tuning_for(size_t M) const {
        const int l = M ? 63 - __builtin_clzll((unsigned long long)M) : 0;
        return tuning[(size_t)l < tuning.size() ? (size_t)l : tuning.size() - 1];
    }
};

struct bfd_code {
    bfd_ctx *ctx = nullptr;
    uint32_t lambda = 0, log2_c = 0, d = 0;
    uint64_t q = 0;
    bfd::rt::DevBuf ft;                     // fold-order twiddles, level s at ft_off[s]
    std::vector<size_t> ft_off;
    // host copy of the XOR basis the table is built from: ft[s][j] = XOR over
    // the set bits p of j of bt[bt_off[s] + p]
    std::vector<gf> bt;
    std::vector<size_t> bt_off;
};

struct bfd_commitment {
    bfd_ctx *ctx = nullptr;
    const bfd_code *code = nullptr;
    uint32_t n = 0, b = 0;
    size_t N = 0, M = 0;
    size_t w = 1;                           // polynomials (columns of the row-major matrices)
    bfd::rt::DevBuf data;                   // N x w evaluations
    bfd::rt::DevBuf cw;                     // M x w codeword elements
    bfd::rt::DevBuf tree;                   // 2M-1 digests, leaf layer first
    uint8_t root[32];
    uint64_t bytes = 0;
};

struct bfd_challenger {
    bfd::Challenger c;
};

namespace bfd {
namespace rt {

// device address of the allocation-free scratch array on `device`, the
// current device; taken once per device (noalloc.cpp)
gf *noalloc_scratch(int device);

// BinaryReedSolomonCode::new with its device twiddle table (api.cpp)
bfd_code *create_code(bfd_ctx *ctx, uint32_t lambda, uint32_t log2_c, uint32_t d);
// cached code for the batched and serving paths
bfd_code *cached_code(bfd_ctx *ctx, uint32_t lambda, uint32_t log2_c, uint32_t d);

// widest commitment (columns); the per-column launches put columns on the
// grid's y dimension
constexpr size_t kMaxWidth = 65535;

inline void set_device(bfd_ctx *ctx) { hip_check(hipSetDevice(ctx->device), "hipSetDevice"); }

// a device buffer on the context stream, drawn from and returned to its cache
inline DevBuf ctx_buf(bfd_ctx *ctx, size_t bytes) { return DevBuf(bytes, ctx->stream, &ctx->cache); }

inline size_t num_queries(uint32_t lambda, uint32_t log2_c) {
    double c = (double)(1ull << log2_c);
    double delta = (1.0 + 1.0 / c) / 2.0;
    double q = ceil((double)lambda / -log2(delta));
    if (!(q > 0)) return 0;  // Rust's saturating float->usize cast
    return (size_t)q;
}

inline gf load_elem(const uint8_t *p) {
    gf v;
    memcpy(&v.lo, p, 8);
    memcpy(&v.hi, p + 8, 8);
    return v;
}
inline void store_elem(uint8_t *p, gf v) {
    memcpy(p, &v.lo, 8);
    memcpy(p + 8, &v.hi, 8);
}

// host-built eq table eq(r[0..m], .) (multilinear.rs:41-64)
inline std::vector<gf> host_eq(const gf *r, int m) {
    std::vector<gf> ev((size_t)1 << m);
    ev[0] = gf_one();
    for (int i = 0; i < m; i++) {
        size_t h = (size_t)1 << i;
        for (size_t k = 0; k < h; k++) {
            ev[h + k] = hmul(ev[k], r[i]);
            ev[k] = gf_add(ev[k], ev[h + k]);
        }
    }
    return ev;
}

// Builds eq(r[0..m], .) of 2^m entries on the device as the tensor product of
// two host-built halves.
// The halves go through ctx->eq_pin without a stream synchronisation: every
// caller synchronises the stream before it returns, so the staging buffer is
// idle again whenever device_eq is entered.
inline void device_eq(bfd_ctx *ctx, const gf *r, int m, gf *out) {
    hipStream_t st = ctx->stream;
    int lo = (m + 1) / 2, hi = m - lo;
    std::vector<gf> A = host_eq(r, lo), B = host_eq(r + lo, hi);
    const size_t na = A.size(), nb = B.size();
    ctx->eq_pin.ensure(16 * (na + nb));
    gf *h = (gf *)ctx->eq_pin.p;
    memcpy(h, A.data(), 16 * na);
    memcpy(h + na, B.data(), 16 * nb);
    DevBuf d(16 * (na + nb), st);
    hip_check(hipMemcpyAsync(d.p, h, 16 * (na + nb), hipMemcpyHostToDevice, st), "copy eq halves");
    ctx->timer.run(st, "k_outer", [&] { return launch::outer(st, d.as<gf>(), lo, d.as<gf>() + na, out, (size_t)1 << m); });
}

// The node levels of a Merkle tree over a power-of-two number of leaf digests,
// as launches.  A level whose input layer holds more than `from` digests gets
// a k_nodes launch of its own; from there up, each launch hashes up to
// subtree_max_levels() levels per block through LDS (k_subtree), so the narrow
// top of the tree costs two launches instead of one per level.  For one tree,
// from = kSubtreeFrom = 2^17 makes the first subtree launch 256 blocks, one
// per CU, so every narrow level runs as one wave per SIMD at one Keccak-f
// latency; with T trees hashed together (segmented launches) subtree_from()
// divides it by T so that the first subtree launch still has about 256 blocks.
struct TreeStep {
    size_t prev;  // offset in digests of the input layer from the leaf layer
    size_t h;     // digests in the input layer
    int lv;       // levels hashed; 0 = one k_nodes level
};
constexpr size_t kSubtreeFrom = (size_t)1 << 17;
inline size_t subtree_from(size_t trees, size_t base = kSubtreeFrom) {
    size_t from = base;
    while (trees > 1 && from > 2) {
        from >>= 1;
        trees = (trees + 1) >> 1;
    }
    return from;
}
// steps with lv == 0 come first, then the subtree steps of at most max_lv
// levels each
inline void tree_steps(size_t leaves, std::vector<TreeStep> &out, size_t from = kSubtreeFrom,
                       int max_lv = launch::subtree_max_levels()) {
    if (leaves & (leaves - 1)) fail(BFD_ERR_ARG, "Merkle leaf count must be a power of two");
    out.clear();
    size_t off = 0, h = leaves;
    while (h > 1) {
        if (h > from) {
            out.push_back({off, h, 0});
            off += h;
            h >>= 1;
            continue;
        }
        const int lv = std::min(std::min(max_lv, launch::subtree_max_levels()), __builtin_ctzll(h));
        out.push_back({off, h, lv});
        for (int i = 0; i < lv; i++) {
            off += h;
            h >>= 1;
        }
    }
}

// MLE value of every column of a rows x w row-major matrix at `point`
// (basefold.rs:126-131) in one pass over the data (k_mle_dot): enqueue()
// stages the eq factors through ctx->eq_pin and the partial sums back into
// ctx->pinned on the context stream; after the caller's synchronisation,
// finish() adds them up.  Shares device_eq's invariant: the staging buffers
// are idle whenever a call starts, because every call synchronises before it
// returns.
struct ColumnEval {
    size_t w = 1;
    unsigned nb = 0;
    int n = 0;
    DevBuf dev;
    const gf *single = nullptr;  // n == 0: the values are the data row itself

    void enqueue(bfd_ctx *ctx, const gf *data, size_t rows, size_t w_, const uint8_t *point, int n_) {
        hipStream_t st = ctx->stream;
        w = w_;
        n = n_;
        const unsigned max_blocks = std::max(1u, std::min(512u, (unsigned)(4096 / w)));
        ctx->pinned.ensure(16 * (size_t)max_blocks * w + 64);
        if (n == 0) {
            hip_check(hipMemcpyAsync(ctx->pinned.p, data, 16 * w, hipMemcpyDeviceToHost, st), "copy values");
            return;
        }
        std::vector<gf> x(n);
        for (int i = 0; i < n; i++) x[i] = load_elem(point + 16 * i);
        const int lo = (n + 1) / 2, hi = n - lo;
        std::vector<gf> A = host_eq(x.data(), lo), B = host_eq(x.data() + lo, hi);
        const size_t na = A.size(), nbq = B.size();
        ctx->eq_pin.ensure(16 * (na + nbq));
        gf *h = (gf *)ctx->eq_pin.p;
        memcpy(h, A.data(), 16 * na);
        memcpy(h + na, B.data(), 16 * nbq);
        dev = ctx_buf(ctx, 16 * (na + nbq) + 16 * (size_t)max_blocks * w);
        gf *dA = dev.as<gf>(), *dpart = dA + na + nbq;
        hip_check(hipMemcpyAsync(dA, h, 16 * (na + nbq), hipMemcpyHostToDevice, st), "copy eq factors");
        ctx->timer.run(st, "k_mle_dot", [&] {
            return launch::mle_dot(st, data, w, dA, lo, dA + na, rows, dpart, max_blocks, &nb);
        });
        hip_check(hipMemcpyAsync(ctx->pinned.p, dpart, 16 * (size_t)nb * w, hipMemcpyDeviceToHost, st), "copy sums");
    }
    void finish(bfd_ctx *ctx, uint8_t *out) const {
        const uint8_t *pin = (const uint8_t *)ctx->pinned.p;
        if (n == 0) {
            memcpy(out, pin, 16 * w);
            return;
        }
        for (size_t c = 0; c < w; c++) {
            gf s = gf_zero();
            for (unsigned u = 0; u < nb; u++) s = gf_add(s, load_elem(pin + 16 * (c * nb + u)));
            store_elem(out + 16 * c, s);
        }
    }
};

// the first `count` of a tree's steps (tree_steps) as launches
inline void tree_step_launches(bfd_ctx *ctx, uint64_t *layer0, const std::vector<TreeStep> &steps,
                               const launch::Shape &sh, size_t count) {
    hipStream_t st = ctx->stream;
    for (size_t q = 0; q < count && q < steps.size(); q++) {
        const TreeStep &s = steps[q];
        uint64_t *prev = layer0 + 4 * s.prev;
        if (s.lv == 0)
            ctx->timer.run(st, "k_nodes", [&] { return launch::nodes(st, prev, prev + 4 * s.h, s.h / 2, sh); });
        else
            ctx->timer.run(st, "k_subtree", [&] { return launch::subtree(st, prev, s.h, s.lv, ctx->cus); });
    }
}

// all Merkle levels above `leaves` digests at layer0 (layers contiguous),
// cut into launches by the shape tn
inline void tree_nodes(bfd_ctx *ctx, uint64_t *layer0, size_t leaves, const bfd_tuning &tn) {
    std::vector<TreeStep> steps;
    tree_steps(leaves, steps, (size_t)1 << tn.subtree_from_log2, (int)tn.subtree_levels);
    tree_step_launches(ctx, layer0, steps, shape_of(tn, ctx->cus), steps.size());
}

// RS encode of a width-w row-major matrix (binary_reed_solomon.rs:67-82):
// layer 0, then the remaining layers in the passes of the shape's plan
inline void encode_w(bfd_ctx *ctx, const bfd_code *code, const gf *data, gf *cw, int b, int n, size_t w,
                     const bfd_tuning &tn) {
    hipStream_t st = ctx->stream;
    const gf *ft = code->ft.as<gf>();
    if (n == 0) {  // c copies of the single row
        for (size_t u = 0; u < ((size_t)1 << b); u++)
            hip_check(hipMemcpyAsync(cw + u * w, data, 16 * w, hipMemcpyDeviceToDevice, st), "replicate");
        return;
    }
    const launch::Shape sh = shape_of(tn, ctx->cus);
    ctx->timer.run(st, "k_ntt_first_w",
                   [&] { return launch::ntt_first_w(st, data, cw, ft + code->ft_off[0], b, n, w, sh); });
    std::vector<std::pair<int, int>> passes;
    ntt_passes(n, tn.ntt_order, passes);
    for (const auto &p : passes) {
        const int s = p.first, R = p.second;
        const gf *t0 = ft + code->ft_off[s];
        const gf *t1 = R >= 2 ? ft + code->ft_off[s + 1] : nullptr, *t2 = R >= 3 ? ft + code->ft_off[s + 2] : nullptr;
        ctx->timer.run(st, "k_ntt_multi_w",
                       [&] { return launch::ntt_multi_w(st, cw, t0, t1, t2, R, s, b, n, w, sh); });
    }
}

inline void encode(bfd_ctx *ctx, const bfd_code *code, const gf *data, gf *cw, int b, int n, const bfd_tuning &tn) {
    hipStream_t st = ctx->stream;
    const gf *ft = code->ft.as<gf>();
    const launch::Shape sh = shape_of(tn, ctx->cus);
    ctx->timer.run(st, "k_ntt_first", [&] { return launch::ntt_first(st, data, cw, ft + code->ft_off[0], b, n, sh); });
    // the remaining layers in passes of up to three (one pass over the
    // codeword per pass)
    std::vector<std::pair<int, int>> passes;
    ntt_passes(n, tn.ntt_order, passes);
    for (const auto &p : passes) {
        const int s = p.first, R = p.second;
        const gf *t0 = ft + code->ft_off[s];
        const gf *t1 = R >= 2 ? ft + code->ft_off[s + 1] : nullptr, *t2 = R >= 3 ? ft + code->ft_off[s + 2] : nullptr;
        if (R == 1)
            ctx->timer.run(st, "k_ntt_layer", [&] { return launch::ntt_layer(st, cw, t0, s, b, n, sh); });
        else
            ctx->timer.run(st, "k_ntt_multi", [&] { return launch::ntt_multi(st, cw, t0, t1, t2, R, s, b, n, sh); });
    }
}

// Basefold::commit; with a point, also the claims of every column
// (Basefold::eval) from the same stream pass and synchronisation.
inline void do_commit(bfd_ctx *ctx, const bfd_code *code, const void *evals, bool on_device, size_t height,
               size_t width, bfd_commitment **out, uint8_t root[32], const uint8_t *point = nullptr,
               uint8_t *values = nullptr) {
    if (!ctx || !code || !out || !evals || !root) fail(BFD_ERR_ARG, "null argument");
    if (code->ctx->device != ctx->device) fail(BFD_ERR_ARG, "the code's twiddles live on another device");
    if (width < 1 || width > kMaxWidth) fail(BFD_ERR_ARG, "width must be 1..65535");
    if (height != ((size_t)1 << code->d)) fail(BFD_ERR_ARG, "height must equal 2^d of the code");
    std::lock_guard<std::recursive_mutex> lk(ctx->mu);
    set_device(ctx);
    hipStream_t st = ctx->stream;
    auto cm = std::make_unique<bfd_commitment>();
    cm->ctx = ctx;
    cm->code = code;
    cm->n = code->d;
    cm->b = code->log2_c;
    cm->N = height;
    cm->M = height << code->log2_c;
    cm->w = width;
    const size_t w = width;
    cm->bytes = 16 * cm->N * w + 16 * cm->M * w + 32 * (2 * cm->M - 1);
    if (ctx->budget && ctx->resident + cm->bytes > ctx->budget)
        fail(BFD_ERR_BUDGET, "commitment exceeds the context's HBM budget");
    const double t0 = now_ms();
    cm->data = ctx_buf(ctx, 16 * cm->N * w);
    cm->cw = ctx_buf(ctx, 16 * cm->M * w);
    cm->tree = ctx_buf(ctx, 32 * (2 * cm->M - 1));
    ctx->timer.host("commit_alloc", now_ms() - t0);
    hip_check(hipMemcpyAsync(cm->data.p, evals, 16 * cm->N * w,
                             on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, st),
              "copy evaluations");
    gf *data = cm->data.as<gf>(), *cw = cm->cw.as<gf>();
    uint64_t *tree = cm->tree.as<uint64_t>();
    const bfd_tuning tn = ctx->tuning_for(cm->M);
    const launch::Shape sh = shape_of(tn, ctx->cus);
    if (w == 1) {
        encode(ctx, code, data, cw, (int)cm->b, (int)cm->n, tn);
        ctx->timer.run(st, "k_leaves_elem", [&] { return launch::leaves_elem(st, cw, tree, cm->M, sh); });
    } else {
        encode_w(ctx, code, data, cw, (int)cm->b, (int)cm->n, w, tn);
        ctx->timer.run(st, "k_leaves_rows", [&] { return launch::leaves_rows(st, cw, w, tree, cm->M, sh); });
    }
    tree_nodes(ctx, tree, cm->M, tn);
    hip_check(hipMemcpyAsync(cm->root, tree + 4 * (2 * cm->M - 2), 32, hipMemcpyDeviceToHost, st), "copy root");
    ColumnEval ev;
    if (values) ev.enqueue(ctx, data, cm->N, w, point, (int)cm->n);
    hip_check(hipStreamSynchronize(st), "commit");
    ctx->timer.collect();
    if (values) ev.finish(ctx, values);
    memcpy(root, cm->root, 32);
    ctx->resident += cm->bytes;
    *out = cm.release();
}

// challenges the random scalars of a width-w opening take: log2 of w rounded
// up (sample_random_scalars_by_eq, basefold.rs:398-408)
inline size_t scalar_challenges(size_t w) {
    size_t lw = 0;
    while (((size_t)1 << lw) < w) lw++;
    return w > 1 ? lw : 0;
}

// Width > 1 (basefold.rs:166-183): the random scalars eq(rs, .) truncated to
// w combine the claims (into claim_b), the evaluation rows (N) and the
// codeword rows (M) into one polynomial and pi_0, held in comb; poly0 and
// pi0 point at them afterwards.  Enqueued on the context stream.
inline void combine_columns(bfd_ctx *ctx, const bfd_commitment *cm, const std::vector<gf> &rs,
                            const uint8_t *values_b, uint8_t claim_b[16], DevBuf &comb, const gf *&poly0,
                            const gf *&pi0) {
    hipStream_t st = ctx->stream;
    const size_t w = cm->w;
    std::vector<gf> sc = host_eq(rs.data(), (int)rs.size());  // MultiPoly::eq, truncated to w
    gf claim = gf_zero();
    for (size_t t = 0; t < w; t++) claim = gf_add(claim, hmul(sc[t], load_elem(values_b + 16 * t)));
    store_elem(claim_b, claim);
    comb = ctx_buf(ctx, 16 * (cm->N + cm->M + w));
    gf *p_poly = comb.as<gf>(), *p_pi = p_poly + cm->N, *p_sc = p_pi + cm->M;
    ctx->scalar_pin.ensure(16 * w);
    memcpy(ctx->scalar_pin.p, sc.data(), 16 * w);
    hip_check(hipMemcpyAsync(p_sc, ctx->scalar_pin.p, 16 * w, hipMemcpyHostToDevice, st), "copy scalars");
    ctx->timer.run(st, "k_rowdot", [&] { return launch::rowdot(st, cm->data.as<gf>(), w, p_sc, p_poly, cm->N); });
    ctx->timer.run(st, "k_rowdot", [&] { return launch::rowdot(st, cm->cw.as<gf>(), w, p_sc, p_pi, cm->M); });
    poly0 = p_poly;
    pi0 = p_pi;
}

inline void put_u64(std::vector<uint8_t> &v, uint64_t x) {
    uint8_t b[8];
    memcpy(b, &x, 8);
    v.insert(v.end(), b, b + 8);
}
inline void put_elem(std::vector<uint8_t> &v, gf x) {
    uint8_t b[16];
    store_elem(b, x);
    v.insert(v.end(), b, b + 16);
}
inline void put_bytes(std::vector<uint8_t> &v, const void *p, size_t n) {
    const uint8_t *b = (const uint8_t *)p;
    v.insert(v.end(), b, b + n);
}

}  // namespace rt
}  // namespace bfd

namespace bfd {
namespace rt {

// ---------------------------------------------------------------------------
// Host side of one width-1 Basefold opening: the Eval sumcheck
// (sumcheck/src/function/{eval,eval_impr}.rs) interleaved with the folding
// rounds (poly-commit/src/basefold.rs:335-369), the query sampling
// (:200-203) and the bincode proof (:49-66, :205-231).  The device side only
// supplies per-round pi-tree roots and inner products; plain and batched
// openings share this so their transcripts are the same code.
// ---------------------------------------------------------------------------
struct OpenRounds {
    int n = 0, b = 0, d = 0, L = 0;
    size_t M = 0, q = 0;
    size_t w = 1;  // elements per opened codeword row
    std::vector<gf> r, rinv, cf;
    gf claim{}, subclaim{}, inner_sub{}, s_delta{}, final_poly{};
    std::vector<std::array<gf, 2>> comp;
    std::vector<std::array<uint8_t, 32>> pi_roots;
    std::vector<uint64_t> idx;

    void init(int n_, int log2_c, size_t q_, const uint8_t *point_b, const uint8_t *claim_b) {
        n = d = n_;
        b = log2_c;
        L = b + d;
        M = (size_t)1 << L;
        q = q_;
        r.assign(n, gf_zero());
        for (int i = 0; i < n; i++) {
            r[i] = load_elem(point_b + 16 * i);
            if (gf_eq(r[i], gf_zero()) || gf_eq(r[i], gf_one()))
                fail(BFD_ERR_ARG, "point coordinates 0 and 1 are not invertible in the Eval prover");
        }
        claim = load_elem(claim_b);
        // r^-1 (EvalImpr::new, eval_impr.rs:20-27) and (1 - r_j)^-1 (HalfEq::new,
        // :69-79) with one batched inversion
        const int hn = n > 0 ? n - 1 : 0;
        std::vector<gf> inv(r);
        for (int j = 0; j < hn; j++) inv.push_back(gf_add(gf_one(), r[j]));
        batch_inverse(inv);
        rinv.assign(inv.begin(), inv.begin() + n);
        // correcting_factor(i) = prod_{j=i}^{n-2} (1 - r_j)^-1  (eval_impr.rs:81-88)
        cf.assign(n + 1, gf_one());
        for (int j = hn - 1; j >= 0; j--) cf[j] = hmul(cf[j + 1], inv[n + j]);
        subclaim = claim;
        inner_sub = gf_zero();
        s_delta = gf_one();
        comp.assign(d, {gf_zero(), gf_zero()});
        pi_roots.assign(d, {});
    }
    int half_eq_vars() const { return n > 0 ? n - 1 : 0; }

    // The Eval prover's round polynomial of round k (level i = d-1-k) from
    // the inner product S = sum half_eq[b] * poly_lo[b]: the inner linear
    // polynomial [coeff0, coeff1] and the outer [rp0, rp1, rp2].
    struct RoundPoly {
        gf coeff0, coeff1, rp0, rp1, rp2;
    };
    RoundPoly round_poly(int k, gf S) const {
        const int i = d - 1 - k;
        const gf sub = (i == n - 1) ? subclaim : inner_sub;             // eval.rs:90-94
        RoundPoly p;
        p.coeff0 = hmul(cf[i], S);
        const gf eval1 = hmul(gf_add(sub, hmul(gf_add(gf_one(), r[i]), p.coeff0)), rinv[i]);  // eval_impr.rs:41-43
        p.coeff1 = gf_add(eval1, p.coeff0);
        // delta (eval_impr.rs:45-58): scalar = eq(r[i+1..], x[i+1..]); in
        // characteristic 2 the two factors are s (1 + r_i) and s
        const gf d0 = hmul(s_delta, gf_add(gf_one(), r[i]));
        const gf d1 = s_delta;
        p.rp0 = hmul(p.coeff0, d0);
        p.rp1 = gf_add(hmul(p.coeff0, d1), hmul(p.coeff1, d0));
        p.rp2 = hmul(p.coeff1, d1);
        return p;
    }
    // After the challenge x of round k: both subclaims and the delta scalar
    // advance (fix_last_var and SumcheckSubclaim::reduce).
    void advance(int k, const RoundPoly &p, gf x) {
        const int i = d - 1 - k;
        inner_sub = gf_add(p.coeff0, hmul(p.coeff1, x));                // eval.rs:110
        subclaim = gf_add(p.rp0, hmul(x, gf_add(p.rp1, hmul(x, p.rp2))));  // lib.rs:100 (horner)
        s_delta = hmul(s_delta, gf_add(gf_one(), gf_add(r[i], x)));
    }

    // Round k with the transcript on the host: observe the pi-tree root,
    // observe the compressed round polynomial, sample the challenge and
    // advance both subclaims.
    gf round(int k, const uint8_t root[32], gf S, Challenger &ch) {
        memcpy(pi_roots[k].data(), root, 32);
        ch.observe(root, 32);                                           // basefold.rs:356
        const RoundPoly p = round_poly(k, S);
        comp[k] = {p.rp0, p.rp2};                                       // eval.rs:64-66
        ch.observe_elem(p.rp0);                                         // sumcheck/src/lib.rs:94
        ch.observe_elem(p.rp2);
        const gf x = ch.sample();                                       // lib.rs:96
        advance(k, p, x);
        return x;
    }

    // basefold.rs:194-203: observe the final polynomial, sample the queries
    void finish(gf fin, Challenger &ch) {
        final_poly = fin;
        ch.observe_elem(fin);
        idx.resize(q);
        for (size_t t = 0; t < q; t++) idx[t] = ch.sample_bits(L);
    }
};

// ---------------------------------------------------------------------------
// Device transcript of an opening (FsState region, kernels.hpp): the host
// writes the starting image (the challenger's bytes and the Eval prover's
// constants), the round kernel advances it on the device, and after the
// last round the host reads back the round data and the challenger state.
// ---------------------------------------------------------------------------
// input capacity: the starting input, or the 32-byte digest that replaces
// it at the first sample, plus the 64 bytes a round observes; a multiple of 8
inline uint32_t fs_in_cap(size_t in_len) { return (uint32_t)(((in_len > 32 ? in_len : 32) + 64 + 7) & ~(size_t)7); }

// Writes the starting image to dst; returns the bytes to upload.
inline size_t fs_write(const OpenRounds &o, const Challenger &ch, uint32_t in_cap, uint8_t *dst) {
    if (ch.input.size() + 64 > in_cap || ch.output.size() > 32) fail(BFD_ERR_ARG, "transcript state too large");
    FsState h;
    memset(&h, 0, sizeof h);
    h.in_len = (uint32_t)ch.input.size();
    h.out_len = (uint32_t)ch.output.size();
    h.in_cap = in_cap;
    h.n = (uint32_t)o.n;
    h.sub = o.subclaim;
    h.inner = o.inner_sub;
    h.s_delta = o.s_delta;
    if (!ch.output.empty()) memcpy(h.out, ch.output.data(), ch.output.size());
    memcpy(dst, &h, sizeof h);
    gf *c = (gf *)(dst + fs_consts_off());
    for (int i = 0; i < o.n; i++) {
        c[i] = o.r[i];
        c[o.n + i] = o.rinv[i];
    }
    for (int i = 0; i <= o.n; i++) c[2 * o.n + i] = o.cf[i];
    const size_t in_off = fs_input_off((uint32_t)o.n);
    memset(dst + fs_rounds_off((uint32_t)o.n), 0, in_off - fs_rounds_off((uint32_t)o.n));
    if (!ch.input.empty()) memcpy(dst + in_off, ch.input.data(), ch.input.size());
    return in_off + ch.input.size();
}

// bytes to read back after the last round (the input is then one digest)
inline size_t fs_readback_bytes(uint32_t n) { return fs_input_off(n) + 32; }

// Takes the round polynomials, the pi-tree roots and the challenger state
// from a read-back image into o and ch (the state OpenRounds::round would
// have left on the host).
inline void fs_read(const uint8_t *img, OpenRounds &o, Challenger &ch) {
    FsState h;
    memcpy(&h, img, sizeof h);
    if ((int)h.n != o.n || h.in_len > 32 || h.out_len > 32) fail(BFD_ERR_HIP, "device transcript state is corrupt");
    const FsRound *rd = (const FsRound *)(img + fs_rounds_off(h.n));
    for (int k = 0; k < o.d; k++) {
        o.comp[k] = {rd[k].rp0, rd[k].rp2};
        memcpy(o.pi_roots[k].data(), rd[k].root, 32);
    }
    const uint8_t *in = img + fs_input_off(h.n);
    ch.input.assign(in, in + h.in_len);
    ch.output.assign(h.out, h.out + h.out_len);
    o.subclaim = h.sub;
    o.inner_sub = h.inner;
    o.s_delta = h.s_delta;
}

// Segment of the round kernel for round k of one opening: its pi-tree root,
// its transcript region and its inner-product sums partial[blk0, blk0 + nb).
inline Seg fs_round_seg(const uint64_t *root, void *fs, int k, uint32_t blk0, uint32_t nb) {
    Seg g;
    memset(&g, 0, sizeof g);
    g.in = root;
    g.out = fs;
    g.count = (uint64_t)k;
    g.x = gf{blk0, nb};
    return g;
}

// Device address of the challenge slot of a transcript region.
inline const gf *fs_challenge(const void *fs) { return (const gf *)((const uint8_t *)fs + offsetof(FsState, x)); }

// Device addresses of one polynomial's opening data.
struct OpenLayout {
    const gf *cw = nullptr;          // pi of round 0: the codeword (width 1) or its scalar combination
    const gf *rows = nullptr;        // codeword matrix, w elements per row (nullptr: cw, w = 1)
    size_t w = 1;
    const uint64_t *tree = nullptr;  // main tree, 2M-1 digests, leaf layer first
    const gf *pis = nullptr;         // pi of rounds 1..d, at pi_off[k]
    const uint64_t *pitrees = nullptr;  // pi trees of rounds 0..d-1, at tree_off[k]
    std::vector<size_t> pi_off, tree_off;

    void plan(size_t M, int d) {
        pi_off.assign(d + 1, 0);
        tree_off.assign(d + 1, 0);
        for (int k = 2; k <= d; k++) pi_off[k] = pi_off[k - 1] + (M >> (k - 1));
        size_t t = 0;
        for (int k = 0; k < d; k++) {
            tree_off[k] = t;
            t += (M >> k) - 1;
        }
        tree_off[d] = t;
    }
    size_t pis_elems(size_t M, size_t c) const { return M - c + 1; }
    size_t pitree_digests() const { return tree_off.empty() ? 1 : tree_off.back() + 1; }
    const gf *pi(int k) const { return k == 0 ? cw : pis + pi_off[k]; }
    const uint64_t *pitree(int k) const { return pitrees + 4 * tree_off[k]; }
};

// Bytes of one query record of BasefoldProof (basefold.rs:49-66): the
// codeword row (w elements) and its L = b + d main-tree siblings, then per
// round a pi sibling value and its b + i pi-tree siblings, each vector with
// its u64 length.
inline size_t query_record_bytes(int b, int d, size_t w) {
    size_t r = 8 + 16 * w + 8 + 32 * (size_t)(b + d) + 8;
    for (int k = 0; k < d; k++) r += 16 + 8 + 32 * (size_t)(b + d - 1 - k);
    return r;
}

// Bytes of BasefoldProof before its query records: the d compressed round
// polynomials, the final polynomial, the d pi-tree roots and the query count.
inline size_t proof_header_bytes(int d) { return 8 + 40 * (size_t)d + 8 + 16 + 8 + 32 * (size_t)d + 8; }

// Exact bincode size of BasefoldProof for d rounds, L = b + d, q queries and
// w elements per opened codeword row.
inline size_t proof_size(int b, int d, size_t q, size_t w = 1) {
    return proof_header_bytes(d) + q * query_record_bytes(b, d, w);
}

// The header of BasefoldProof (basefold.rs:49-66) into out, which holds
// proof_header_bytes(o.d) bytes; the query records follow it.
inline void write_proof_header(const OpenRounds &o, uint8_t *out) {
    uint8_t *p = out;
    const auto u64 = [&](uint64_t v) { memcpy(p, &v, 8); p += 8; };
    const auto elem = [&](gf v) { store_elem(p, v); p += 16; };
    u64((uint64_t)o.d);
    for (int k = 0; k < o.d; k++) {
        u64(2);
        elem(o.comp[k][0]);
        elem(o.comp[k][1]);
    }
    u64(1);
    elem(o.final_poly);
    u64((uint64_t)o.d);
    for (int k = 0; k < o.d; k++) {
        memcpy(p, o.pi_roots[k].data(), 32);
        p += 32;
    }
    u64((uint64_t)o.q);
}

// k_proof_records' description of o's query records: indices at d_idx, the
// records (o.q of query_record_bytes each) written to d_out; rec0 = the
// job's first record in the launch.
inline ProofJob proof_job(const OpenRounds &o, const OpenLayout &lay, const uint64_t *d_idx, uint8_t *d_out,
                          uint32_t rec0) {
    ProofJob J;
    memset(&J, 0, sizeof J);
    J.idx = d_idx;
    J.cw = lay.cw;
    J.rows = lay.rows;
    J.tree = lay.tree;
    J.pis = lay.pis;
    J.pitrees = lay.pitrees;
    J.out = d_out;
    J.w = lay.rows ? lay.w : 1;
    J.rec_bytes = query_record_bytes(o.b, o.d, J.w);
    J.b = (uint32_t)o.b;
    J.d = (uint32_t)o.d;
    J.q = (uint32_t)o.q;
    J.rec0 = rec0;
    return J;
}

// Persistent host worker pool for per-polynomial host work (proof planning
// and assembly, transcript set-up).  run() hands out indices through an
// atomic counter; the calling thread works too.  One run at a time.  An
// exception thrown by f stops the hand-out; run() waits for every worker and
// then rethrows the first one in the calling thread.
class HostPool {
  public:
    static HostPool &get() {
        static HostPool pool;
        return pool;
    }
    template <class F>
    void run(size_t n, F &&f) {
        if (n == 0) return;
        if (n == 1 || workers_.empty()) {
            for (size_t i = 0; i < n; i++) f(i);
            return;
        }
        std::lock_guard<std::mutex> one(run_mu_);
        std::function<void(size_t)> fn = f;
        {
            std::lock_guard<std::mutex> lk(mu_);
            fn_ = &fn;
            n_ = n;
            next_.store(0);
            busy_ = workers_.size();
            gen_++;
        }
        cv_.notify_all();
        work(fn, n);
        std::unique_lock<std::mutex> lk(mu_);
        done_cv_.wait(lk, [&] { return busy_ == 0; });
        fn_ = nullptr;
        if (err_) {
            std::exception_ptr e = err_;
            err_ = nullptr;
            std::rethrow_exception(e);
        }
    }
    ~HostPool() {
        {
            std::lock_guard<std::mutex> lk(mu_);
            stop_ = true;
            gen_++;
        }
        cv_.notify_all();
        for (auto &t : workers_) t.join();
    }

  private:
    HostPool() {
        unsigned t = std::thread::hardware_concurrency();
        if (t > 16) t = 16;
        for (unsigned k = 1; k < t; k++) workers_.emplace_back([this] { loop(); });
    }
    void loop() {
        uint64_t seen = 0;
        for (;;) {
            std::function<void(size_t)> *fn;
            size_t n;
            {
                std::unique_lock<std::mutex> lk(mu_);
                cv_.wait(lk, [&] { return stop_ || gen_ != seen; });
                if (stop_) return;
                seen = gen_;
                fn = fn_;
                n = n_;
            }
            work(*fn, n);
            std::lock_guard<std::mutex> lk(mu_);
            if (--busy_ == 0) done_cv_.notify_all();
        }
    }
    void work(const std::function<void(size_t)> &fn, size_t n) {
        try {
            for (size_t i; (i = next_.fetch_add(1)) < n;) fn(i);
        } catch (...) {
            next_.store(n);  // the other threads take no further index
            std::lock_guard<std::mutex> lk(mu_);
            if (!err_) err_ = std::current_exception();
        }
    }
    std::vector<std::thread> workers_;
    std::mutex mu_, run_mu_;
    std::condition_variable cv_, done_cv_;
    std::function<void(size_t)> *fn_ = nullptr;
    size_t n_ = 0, busy_ = 0;
    uint64_t gen_ = 0;
    bool stop_ = false;
    std::exception_ptr err_;
    std::atomic<size_t> next_{0};
};

// Runs f(i) for i in [0, n) on the host pool.
template <class F>
void parallel_for(size_t n, F &&f) {
    HostPool::get().run(n, std::forward<F>(f));
}

// Host copies cut into chunks for the pool: add() records a copy, run()
// performs every recorded copy (one pool run, so it must not be called from
// inside another).
struct ChunkCopy {
    static constexpr size_t kChunk = (size_t)256 << 10;
    struct Piece {
        uint8_t *dst;
        const uint8_t *src;
        size_t n;
    };
    std::vector<Piece> pieces;
    void add(uint8_t *dst, const uint8_t *src, size_t n) {
        for (size_t o = 0; o < n; o += kChunk) pieces.push_back({dst + o, src + o, std::min(kChunk, n - o)});
    }
    void run() {
        parallel_for(pieces.size(), [&](size_t i) { memcpy(pieces[i].dst, pieces[i].src, pieces[i].n); });
        pieces.clear();
    }
    // zero-fills n bytes at dst on the pool (touches fresh pages)
    static void zero(uint8_t *dst, size_t n) {
        parallel_for((n + kChunk - 1) / kChunk, [&](size_t c) { memset(dst + c * kChunk, 0, std::min(kChunk, n - c * kChunk)); });
    }
};

// One opening's proof into dst (sz = proof_size bytes, pages already
// touched): the header from o on the host, then the query records at o.idx,
// written by the device in their bincode layout (k_proof_records) into
// ctx->gather_pin and copied in behind the header.  Stages the indices
// through ctx->pinned and synchronises the context stream.
inline void write_proof(bfd_ctx *ctx, const OpenRounds &o, const OpenLayout &lay, size_t sz, uint8_t *dst) {
    hipStream_t st = ctx->stream;
    write_proof_header(o, dst);
    const size_t hb = proof_header_bytes(o.d), body = sz - hb;
    if (!o.q) {
        hip_check(hipStreamSynchronize(st), "opening");
        ctx->timer.collect();
        return;
    }
    const size_t up = sizeof(ProofJob) + 8 * o.q;
    ctx->pinned.ensure(up);
    uint8_t *hu = (uint8_t *)ctx->pinned.p;
    // the records go from the kernel straight into pinned host memory: one
    // pass over PCIe, no device buffer and no copy-engine transfer behind it
    ctx->gather_pin.ensure(body);
    uint8_t *hout = nullptr;
    hip_check(hipHostGetDevicePointer((void **)&hout, ctx->gather_pin.p, 0), "hipHostGetDevicePointer");
    DevBuf dj = ctx_buf(ctx, up);
    const ProofJob J = proof_job(o, lay, reinterpret_cast<const uint64_t *>(dj.as<uint8_t>() + sizeof(ProofJob)), hout, 0);
    memcpy(hu, &J, sizeof J);
    memcpy(hu + sizeof J, o.idx.data(), 8 * o.q);
    hip_check(hipMemcpyAsync(dj.p, hu, up, hipMemcpyHostToDevice, st), "copy query indices");
    ctx->timer.run(st, "k_proof_records", [&] { return launch::proof_records(st, dj.as<ProofJob>(), 1, (uint32_t)o.q); });
    const double t_wait = now_ms();
    hip_check(hipStreamSynchronize(st), "query records");
    ctx->timer.host("proof_records_wait", now_ms() - t_wait);
    ctx->timer.collect();
    const double t_copy = now_ms();
    ChunkCopy cc;
    cc.add(dst + hb, (const uint8_t *)ctx->gather_pin.p, body);
    cc.run();
    ctx->timer.host("proof_records_copy", now_ms() - t_copy);
}

}  // namespace rt
}  // namespace bfd
