// Shared internals of the host runtime: error handling, device buffers,
// kernel timing, the opaque handle types of the C ABI and the helpers the
// plain, batched and serving paths have in common.
#pragma once
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <array>
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
// (kept cached across calls so repeated commits/opens do not hit the driver)
// ---------------------------------------------------------------------------
struct DevBuf {
    void *p = nullptr;
    size_t bytes = 0;
    hipStream_t st = nullptr;
    DevBuf() = default;
    DevBuf(size_t n, hipStream_t s) : bytes(n), st(s) {
        if (n) hip_check(hipMallocAsync(&p, n, s), "hipMallocAsync");
    }
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    DevBuf(DevBuf &&o) noexcept { *this = std::move(o); }
    DevBuf &operator=(DevBuf &&o) noexcept {
        release();
        p = o.p; bytes = o.bytes; st = o.st;
        o.p = nullptr; o.bytes = 0;
        return *this;
    }
    ~DevBuf() { release(); }
    void release() {
        if (p) (void)hipFreeAsync(p, st);
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

// Optional per-kernel timing with HIP events on the context's stream.
struct KernelTimer {
    bool enabled = false;
    struct Rec {
        std::string name;
        hipEvent_t a, b;
    };
    std::vector<Rec> pending;
    std::vector<hipEvent_t> pool;
    std::map<std::string, std::pair<double, uint32_t>> totals;
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
    template <class F>
    void run(hipStream_t st, const char *name, F &&launch) {
        if (!enabled) {
            hip_check(launch(), name);
            return;
        }
        Rec r{name, get(), get()};
        hip_check(hipEventRecord(r.a, st), "hipEventRecord");
        hip_check(launch(), name);
        hip_check(hipEventRecord(r.b, st), "hipEventRecord");
        pending.push_back(r);
    }
    // call after a stream synchronisation
    void collect() {
        for (auto &r : pending) {
            float ms = 0;
            hip_check(hipEventElapsedTime(&ms, r.a, r.b), "hipEventElapsedTime");
            auto &t = totals[r.name];
            t.first += ms;
            t.second += 1;
            pool.push_back(r.a);
            pool.push_back(r.b);
        }
        pending.clear();
    }
    ~KernelTimer() {
        for (auto &r : pending) { (void)hipEventDestroy(r.a); (void)hipEventDestroy(r.b); }
        for (auto e : pool) (void)hipEventDestroy(e);
    }
};

}  // namespace rt
}  // namespace bfd

// ---------------------------------------------------------------------------
// handles
// ---------------------------------------------------------------------------
struct bfd_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    uint64_t budget = 0;
    uint64_t resident = 0;
    std::recursive_mutex mu;
    bfd::rt::PinnedBuf pinned;
    bfd::rt::KernelTimer timer;
};

struct bfd_code {
    bfd_ctx *ctx = nullptr;
    uint32_t lambda = 0, log2_c = 0, d = 0;
    uint64_t q = 0;
    bfd::rt::DevBuf ft;                     // fold-order twiddles, level s at ft_off[s]
    std::vector<size_t> ft_off;
};

struct bfd_commitment {
    bfd_ctx *ctx = nullptr;
    const bfd_code *code = nullptr;
    uint32_t n = 0, b = 0;
    size_t N = 0, M = 0;
    bfd::rt::DevBuf data;                   // N evaluations
    bfd::rt::DevBuf cw;                     // M codeword elements
    bfd::rt::DevBuf tree;                   // 2M-1 digests, leaf layer first
    uint8_t root[32];
    uint64_t bytes = 0;
};

struct bfd_challenger {
    bfd::Challenger c;
};

namespace bfd {
namespace rt {

inline void set_device(bfd_ctx *ctx) { hip_check(hipSetDevice(ctx->device), "hipSetDevice"); }

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
inline void device_eq(bfd_ctx *ctx, const gf *r, int m, gf *out) {
    hipStream_t st = ctx->stream;
    int lo = m < 12 ? m : 12, hi = m - lo;
    std::vector<gf> A = host_eq(r, lo), B = host_eq(r + lo, hi);
    DevBuf dA(A.size() * 16, st), dB(B.size() * 16, st);
    hip_check(hipMemcpyAsync(dA.p, A.data(), A.size() * 16, hipMemcpyHostToDevice, st), "copy eq A");
    hip_check(hipMemcpyAsync(dB.p, B.data(), B.size() * 16, hipMemcpyHostToDevice, st), "copy eq B");
    ctx->timer.run(st, "k_outer", [&] { return launch::outer(st, dA.as<gf>(), lo, dB.as<gf>(), out, (size_t)1 << m); });
    // the pageable host halves must outlive the copies
    hip_check(hipStreamSynchronize(st), "eq table");
}

// all Merkle levels above `leaves` digests at layer0 (layers contiguous)
inline void tree_nodes(bfd_ctx *ctx, uint64_t *layer0, size_t leaves) {
    hipStream_t st = ctx->stream;
    uint64_t *prev = layer0;
    for (size_t h = leaves; h > 1; h >>= 1) {
        uint64_t *next = prev + 4 * h;
        ctx->timer.run(st, "k_nodes", [&] { return launch::nodes(st, prev, next, h / 2); });
        prev = next;
    }
}

inline void encode(bfd_ctx *ctx, const bfd_code *code, const gf *data, gf *cw, int b, int n) {
    hipStream_t st = ctx->stream;
    const gf *ft = code->ft.as<gf>();
    ctx->timer.run(st, "k_ntt_first", [&] { return launch::ntt_first(st, data, cw, ft + code->ft_off[0], b, n); });
    for (int s = 1; s < n; s++)
        ctx->timer.run(st, "k_ntt_layer", [&] { return launch::ntt_layer(st, cw, ft + code->ft_off[s], s, b, n); });
}

inline void do_commit(bfd_ctx *ctx, const bfd_code *code, const void *evals, bool on_device, size_t height,
               size_t width, bfd_commitment **out, uint8_t root[32]) {
    if (!ctx || !code || !out || !evals) fail(BFD_ERR_ARG, "null argument");
    if (width != 1) fail(BFD_ERR_ARG, "only width-1 commitments are supported");
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
    cm->bytes = 16 * cm->N + 16 * cm->M + 32 * (2 * cm->M - 1);
    if (ctx->budget && ctx->resident + cm->bytes > ctx->budget)
        fail(BFD_ERR_BUDGET, "commitment exceeds the context's HBM budget");
    cm->data = DevBuf(16 * cm->N, st);
    cm->cw = DevBuf(16 * cm->M, st);
    cm->tree = DevBuf(32 * (2 * cm->M - 1), st);
    hip_check(hipMemcpyAsync(cm->data.p, evals, 16 * cm->N,
                             on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, st),
              "copy evaluations");
    gf *data = cm->data.as<gf>(), *cw = cm->cw.as<gf>();
    encode(ctx, code, data, cw, (int)cm->b, (int)cm->n);
    uint64_t *tree = cm->tree.as<uint64_t>();
    ctx->timer.run(st, "k_leaves_elem", [&] { return launch::leaves_elem(st, cw, tree, cm->M); });
    tree_nodes(ctx, tree, cm->M);
    hip_check(hipMemcpyAsync(cm->root, tree + 4 * (2 * cm->M - 2), 32, hipMemcpyDeviceToHost, st), "copy root");
    hip_check(hipStreamSynchronize(st), "commit");
    ctx->timer.collect();
    memcpy(root, cm->root, 32);
    ctx->resident += cm->bytes;
    *out = cm.release();
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
