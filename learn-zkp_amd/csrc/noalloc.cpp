// bfd_eval_data_noalloc: bfd_eval_data on device-resident evaluations with
// no device allocation, for callers that cannot allocate (code under HIP
// graph capture, tight serving loops, harnesses that forbid hipMalloc).
//
// Its working memory (the two eq factors of the point, the per-block
// partial sums and the per-column values) lives in `scratch`, a file-scope
// device array of kScratchElems field elements (4 MiB) that every GPU the
// library runs on holds until the process ends.  Its device address is taken
// once per device with hipGetSymbolAddress when a context is created on that
// device (bfd_ctx_create); host code uses only that address, never the
// array's name.  The eq factors are built on the device from the point,
// passed by value, so no host staging buffer is involved either.  Every call
// writes each scratch word it reads before reading it, so nothing has to
// start at zero; two calls that overlap on one GPU share the array and race.
#include "internal.hpp"

using namespace bfd::rt;

constexpr int kMaxVars = 32;                     // 2^16 + 2^16 eq-factor entries
constexpr size_t kFactor = (size_t)1 << 16;      // entries per eq factor
constexpr size_t kPartials = 65536;              // max_blocks * width <= max(4096, 65535)
constexpr size_t kValues = 65536;                // one value per column, width <= 65535
constexpr size_t kScratchElems = 2 * kFactor + kPartials + kValues;

// 16 bytes per field element, as two 64-bit words
static __device__ __align__(256) uint64_t scratch[2 * kScratchElems];

namespace {

struct EqPoint {
    gf r[kMaxVars];
    int lo, hi;  // A = eq(r[0..lo], .), B = eq(r[lo..lo+hi], .)
};

// A[x] = prod over bits j < lo of x of (r_j if set, 1 + r_j if clear); B the
// same over r[lo..lo+hi] (multilinear.rs:41-64 entry by entry)
__global__ __launch_bounds__(128) void k_eq_factors(EqPoint p, gf *__restrict__ A, gf *__restrict__ B) {
    const size_t na = (size_t)1 << p.lo, nb = (size_t)1 << p.hi;
    for (size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x; t < na + nb; t += (size_t)gridDim.x * blockDim.x) {
        const bool first = t < na;
        const size_t x = first ? t : t - na;
        const int m = first ? p.lo : p.hi, base = first ? 0 : p.lo;
        gf acc = gf_one();
        for (int j = 0; j < m; j++) {
            const gf rj = p.r[base + j];
            acc = gf_mul_reg(acc, ((x >> j) & 1) ? rj : gf_add(rj, gf_one()));
        }
        gf_st(first ? A + x : B + x, acc);
    }
}

// out[c] = XOR of partial[c * nb + u], u < nb (the column sums of k_mle_dot)
__global__ __launch_bounds__(128) void k_col_xor(const gf *__restrict__ partial, unsigned nb, size_t w,
                                                 gf *__restrict__ out) {
    for (size_t c = blockIdx.x * (size_t)blockDim.x + threadIdx.x; c < w; c += (size_t)gridDim.x * blockDim.x) {
        gf s = gf_zero();
        for (unsigned u = 0; u < nb; u++) s = gf_add(s, gf_ld(partial + c * nb + u));
        gf_st(out + c, s);
    }
}

std::mutex g_scratch_mu;
std::map<int, gf *> g_scratch_ptr;

}  // namespace

namespace bfd {
namespace rt {

// The device address of `scratch` on `device` (the current device), taken
// once per device.
gf *noalloc_scratch(int device) {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    auto it = g_scratch_ptr.find(device);
    if (it != g_scratch_ptr.end()) return it->second;
    void *p = nullptr;
    hip_check(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
    if (!p) fail(BFD_ERR_HIP, "scratch has no device address");
    g_scratch_ptr[device] = (gf *)p;
    return (gf *)p;
}

}  // namespace rt
}  // namespace bfd

extern "C" {

int bfd_eval_data_noalloc(bfd_ctx *ctx, const void *d_evals, size_t height, size_t width, const uint8_t *point,
                          uint8_t *values) {
    return guarded([&] {
        if (!ctx || !d_evals || !values || (!point && height > 1)) fail(BFD_ERR_ARG, "null argument");
        if (width < 1 || width > 65535) fail(BFD_ERR_ARG, "width must be 1..65535");
        if (height == 0 || (height & (height - 1))) fail(BFD_ERR_ARG, "height must be a power of two");
        const int n = __builtin_ctzll((unsigned long long)height);
        if (n > kMaxVars) fail(BFD_ERR_ARG, "the allocation-free eval takes at most 2^32 rows");
        if (!ctx->scratch) fail(BFD_ERR_HIP, "no scratch address on this device: " + ctx->scratch_error);
        std::lock_guard<std::recursive_mutex> lk(ctx->mu);
        set_device(ctx);
        hipStream_t st = ctx->stream;
        const gf *data = (const gf *)d_evals;
        if (n == 0) {  // the values are the single row
            hip_check(hipMemcpyAsync(values, data, 16 * width, hipMemcpyDeviceToHost, st), "copy values");
            hip_check(hipStreamSynchronize(st), "eval");
            return;
        }
        gf *A = ctx->scratch, *B = A + kFactor, *partial = B + kFactor, *out = partial + kPartials;
        EqPoint p;
        memset(&p, 0, sizeof p);
        p.lo = (n + 1) / 2;
        p.hi = n - p.lo;
        for (int i = 0; i < n; i++) p.r[i] = load_elem(point + 16 * i);
        const size_t nab = ((size_t)1 << p.lo) + ((size_t)1 << p.hi);
        k_eq_factors<<<(unsigned)((nab + 127) / 128), 128, 0, st>>>(p, A, B);
        hip_check(hipGetLastError(), "k_eq_factors");
        // the same launch and block count as ColumnEval (bfd_eval_data)
        const unsigned max_blocks = std::max(1u, std::min(512u, (unsigned)(4096 / width)));
        if ((size_t)max_blocks * width > kPartials) fail(BFD_ERR_ARG, "partial sums exceed the scratch");
        unsigned nb = 0;
        ctx->timer.run(st, "k_mle_dot", [&] {
            return launch::mle_dot(st, data, width, A, p.lo, B, height, partial, max_blocks, &nb);
        });
        k_col_xor<<<(unsigned)((width + 127) / 128), 128, 0, st>>>(partial, nb, width, out);
        hip_check(hipGetLastError(), "k_col_xor");
        hip_check(hipMemcpyAsync(values, out, 16 * width, hipMemcpyDeviceToHost, st), "copy values");
        hip_check(hipStreamSynchronize(st), "eval");
        ctx->timer.collect();
    });
}

}  // extern "C"
