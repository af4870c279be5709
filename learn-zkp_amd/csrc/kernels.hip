// CDNA4 kernels of the Basefold commit/open path over POLYVAL GF(2^128).
//
// Data layout in HBM (all arrays in the reference's codeword order, 16 B per
// element, 32 B per digest):
//   codeword   cw[x*c + u]      u < c is the coset slot; slot u holds coset
//                               t = bitrev_b(u), which makes the final
//                               bit-reversal of encode_batch
//                               (binary_reed_solomon.rs:81) the identity.
//   fold table ft[s][j]         = twiddles[b+s][bitrev_{b+s}(j)], j < c*2^s:
//                               the fold twiddle of level s
//                               (binary_reed_solomon.rs:29-35, 88-96) and,
//                               at index (p << b) | u, the NTT twiddle of
//                               data-order layer s (binary.rs:98-137).
//   Merkle     layers stored leaf layer first, each layer contiguous.
//
// In data order (index x of the input, not of its bit reversal) the
// reference's coset NTT layer i = n-1-s pairs x and x + 2^s, so the n layers
// run with strides 1, 2, ..., N/2 over contiguous memory.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "gf128.hpp"
#include "keccak.hpp"
#include "kernels.hpp"

namespace bfd {
namespace {

constexpr int CB = 128;  // block size of kernels that multiply (LDS comb tables)
constexpr int KB = 256;  // block size of hashing kernels

__device__ __forceinline__ uint32_t *comb_slot(uint32_t *lds) { return lds + threadIdx.x; }

inline unsigned grid_for(size_t work, int block, unsigned cap) {
    size_t g = (work + block - 1) / block;
    if (g > cap) g = cap;
    if (g == 0) g = 1;
    return (unsigned)g;
}

// ---------------------------------------------------------------------------
// twiddle tables
// ---------------------------------------------------------------------------
// ft[j] = XOR over set bits p of j of bt[p], j < count
__global__ void k_xor_span(gf *ft, const gf *bt, int bits, size_t count) {
    for (size_t j = blockIdx.x * (size_t)blockDim.x + threadIdx.x; j < count;
         j += (size_t)gridDim.x * blockDim.x) {
        gf acc = gf_zero();
        for (int p = 0; p < bits; p++)
            if ((j >> p) & 1) acc = gf_add(acc, bt[p]);
        gf_st(ft + j, acc);
    }
}

// ---------------------------------------------------------------------------
// Reed-Solomon encode: NTT layers in data order
// ---------------------------------------------------------------------------
// One thread per (butterfly, slot): q = (hi * 2^s + p) * c + u, so lanes
// touch consecutive 16-byte elements and consecutive twiddles; the twiddle
// index (p << b) | u is simply q mod 2^(s+b).
// Layer s = 0 (stride 1) reads the input and replicates it into all slots.
__global__ __launch_bounds__(256) void k_ntt_first(const gf *__restrict__ data, gf *__restrict__ cw,
                                                   const gf *__restrict__ ft0, int b, size_t work) {
    const size_t cm = ((size_t)1 << b) - 1;
    for (size_t q = blockIdx.x * (size_t)blockDim.x + threadIdx.x; q < work;
         q += (size_t)gridDim.x * blockDim.x) {
        const size_t u = q & cm, pr = q >> b;
        gf lo = gf_ld(data + 2 * pr), hi = gf_ld(data + 2 * pr + 1);
        lo = gf_add(lo, gf_mul_reg(hi, gf_ld(ft0 + u)));
        hi = gf_add(hi, lo);
        gf_st(cw + ((2 * pr) << b) + u, lo);
        gf_st(cw + ((2 * pr + 1) << b) + u, hi);
    }
}

// general layer s >= 1, in place on the slot-interleaved codeword
__global__ __launch_bounds__(256) void k_ntt_layer(gf *__restrict__ cw, const gf *__restrict__ fts, int s, int b,
                                                   size_t work) {
    const size_t tmask = ((size_t)1 << (s + b)) - 1;
    for (size_t q = blockIdx.x * (size_t)blockDim.x + threadIdx.x; q < work;
         q += (size_t)gridDim.x * blockDim.x) {
        const size_t lo_i = ((q >> (s + b)) << (s + b + 1)) | (q & tmask);
        const size_t hi_i = lo_i + ((size_t)1 << (s + b));
        gf lo = gf_ld(cw + lo_i), hi = gf_ld(cw + hi_i);
        lo = gf_add(lo, gf_mul_reg(hi, gf_ld(fts + (q & tmask))));
        hi = gf_add(hi, lo);
        gf_st(cw + lo_i, lo);
        gf_st(cw + hi_i, hi);
    }
}

// R consecutive layers s .. s+R-1 in one pass (K = s + b): group g owns the
// 2^R elements base + i * 2^K, base = ((g >> K) << (K + R)) | (g mod 2^K).
// Layer s + r pairs i and i + 2^r with twiddle t_r[(g mod 2^K) | (i mod 2^r)
// << K] (the twiddle index is the low K + r bits of the element index), so
// layer s uses one twiddle for the whole group, layer s+1 two, layer s+2
// four, and each twiddle is prepared once for all its butterflies.  Two
// waves per SIMD (at most 256 VGPRs) hide one wave's loads behind the other's
// multiplies: 10.9 ms for the 23 layers of a 2^26-element codeword against
// 12.2 ms at the compiler's one-wave allocation.
// The group body, expanded in place in both kernels below: written out as a
// function and inlined, it makes the register allocator spill 650 bytes per
// lane at two waves per SIMD, against 20 bytes when it is in the kernel body.
#define NTT_GROUP(R, cw, t0, t1, t2, K, mK, stride, g, pitch, col)                        \
    do {                                                                                  \
        constexpr int E_ = 1 << (R);                                                      \
        const size_t low = (g) & (mK);                                                    \
        gf *p = (cw) + ((((g) >> (K)) << ((K) + (R))) | low) * (pitch) + (col);           \
        gf v[E_];                                                                         \
        _Pragma("unroll") for (int i = 0; i < E_; i++) v[i] = gf_ld(p + i * (stride) * (pitch)); \
        {                                                                                 \
            const gf_pre w = gf_prepare(gf_ld((t0) + low));                               \
            _Pragma("unroll") for (int i = 0; i < E_; i += 2) {                           \
                v[i] = gf_add(v[i], gf_mul_pre(v[i + 1], w));                             \
                v[i + 1] = gf_add(v[i + 1], v[i]);                                        \
            }                                                                             \
        }                                                                                 \
        if constexpr ((R) >= 2) {                                                         \
            _Pragma("unroll") for (int a = 0; a < 2; a++) {                               \
                const gf_pre w = gf_prepare(gf_ld((t1) + (low | ((size_t)a << (K)))));    \
                _Pragma("unroll") for (int i = a; i < E_; i += 4) {                       \
                    v[i] = gf_add(v[i], gf_mul_pre(v[i + 2], w));                         \
                    v[i + 2] = gf_add(v[i + 2], v[i]);                                    \
                }                                                                         \
            }                                                                             \
        }                                                                                 \
        if constexpr ((R) >= 3) {                                                         \
            _Pragma("unroll") for (int a = 0; a < 4; a++) {                               \
                v[a] = gf_add(v[a], gf_mul_reg(v[a + 4], gf_ld((t2) + (low | ((size_t)a << (K)))))); \
                v[a + 4] = gf_add(v[a + 4], v[a]);                                        \
            }                                                                             \
        }                                                                                 \
        _Pragma("unroll") for (int i = 0; i < E_; i++) gf_st(p + i * (stride) * (pitch), v[i]); \
    } while (0)

template <int R>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) void k_ntt_multi(
    gf *__restrict__ cw, const gf *__restrict__ t0, const gf *__restrict__ t1, const gf *__restrict__ t2, int K,
    size_t groups) {
    const size_t mK = ((size_t)1 << K) - 1, stride = (size_t)1 << K;
    for (size_t g = blockIdx.x * (size_t)blockDim.x + threadIdx.x; g < groups;
         g += (size_t)gridDim.x * blockDim.x)
        NTT_GROUP(R, cw, t0, t1, t2, K, mK, stride, g, 1, 0);
}

// The three-layer pass built for three waves per SIMD (at most 168 VGPRs):
// more loads in flight per SIMD, against whatever the tighter register
// budget costs; the autotuner picks between the two builds.
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3))) void k_ntt_multi3_w3(
    gf *__restrict__ cw, const gf *__restrict__ t0, const gf *__restrict__ t1, const gf *__restrict__ t2, int K,
    size_t groups) {
    const size_t mK = ((size_t)1 << K) - 1, stride = (size_t)1 << K;
    for (size_t g = blockIdx.x * (size_t)blockDim.x + threadIdx.x; g < groups;
         g += (size_t)gridDim.x * blockDim.x)
        NTT_GROUP(3, cw, t0, t1, t2, K, mK, stride, g, 1, 0);
}

// Width-w codeword (row-major, w elements per row): the same passes with one
// thread per (group, column); every column of a row uses the row's twiddles.
template <int R>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) void k_ntt_multi_w(
    gf *__restrict__ cw, const gf *__restrict__ t0, const gf *__restrict__ t1, const gf *__restrict__ t2, int K,
    size_t groups, size_t w) {
    const size_t mK = ((size_t)1 << K) - 1, stride = (size_t)1 << K;
    for (size_t q = blockIdx.x * (size_t)blockDim.x + threadIdx.x; q < groups * w;
         q += (size_t)gridDim.x * blockDim.x) {
        const size_t g = q / w, col = q - g * w;
        NTT_GROUP(R, cw, t0, t1, t2, K, mK, stride, g, w, col);
    }
}

// layer 0 of a width-w matrix: data rows 2pr, 2pr+1 into codeword rows
// (2pr << b) + u and ((2pr + 1) << b) + u, one thread per (pair, slot, column)
__global__ __launch_bounds__(256) void k_ntt_first_w(const gf *__restrict__ data, gf *__restrict__ cw,
                                                     const gf *__restrict__ ft0, int b, size_t work, size_t w) {
    const size_t cm = ((size_t)1 << b) - 1;
    for (size_t q = blockIdx.x * (size_t)blockDim.x + threadIdx.x; q < work * w;
         q += (size_t)gridDim.x * blockDim.x) {
        const size_t it = q / w, col = q - it * w;
        const size_t u = it & cm, pr = it >> b;
        gf lo = gf_ld(data + (2 * pr) * w + col), hi = gf_ld(data + (2 * pr + 1) * w + col);
        lo = gf_add(lo, gf_mul_reg(hi, gf_ld(ft0 + u)));
        hi = gf_add(hi, lo);
        gf_st(cw + (((2 * pr) << b) + u) * w + col, lo);
        gf_st(cw + (((2 * pr + 1) << b) + u) * w + col, hi);
    }
}

__global__ void k_replicate(const gf *__restrict__ data, gf *__restrict__ cw, int c, size_t n) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n * c;
         i += (size_t)gridDim.x * blockDim.x)
        gf_st(cw + i, gf_ld(data + i / c));
}

// ---------------------------------------------------------------------------
// Merkle hashing (Keccak-256)
// ---------------------------------------------------------------------------
__device__ __forceinline__ void st_digest(uint64_t *p, const uint64_t d[4]) {
    reinterpret_cast<ulonglong2 *>(p)[0] = ulonglong2{d[0], d[1]};
    reinterpret_cast<ulonglong2 *>(p)[1] = ulonglong2{d[2], d[3]};
}

// leaf i = Keccak256(elem[i]) for width-1 rows (16-byte leaves)
__global__ __launch_bounds__(KB) void k_leaves_elem(const gf *__restrict__ rows, uint64_t *__restrict__ out,
                                                    size_t count) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < count;
         i += (size_t)gridDim.x * blockDim.x) {
        gf v = gf_ld(rows + i);
        uint64_t m[2] = {v.lo, v.hi}, d[4];
        kk::keccak256_words<2>(m, d);
        st_digest(out + 4 * i, d);
    }
}

// leaf j = Keccak256(pi[j] || pi[j + half]) (pi-tree rows, basefold.rs:350-354)
__global__ __launch_bounds__(KB) void k_leaves_pair(const gf *__restrict__ pi, uint64_t *__restrict__ out,
                                                    size_t half) {
    for (size_t j = blockIdx.x * (size_t)blockDim.x + threadIdx.x; j < half;
         j += (size_t)gridDim.x * blockDim.x) {
        gf a = gf_ld(pi + j), b = gf_ld(pi + j + half);
        uint64_t m[4] = {a.lo, a.hi, b.lo, b.hi}, d[4];
        kk::keccak256_words<4>(m, d);
        st_digest(out + 4 * j, d);
    }
}

// leaf i = Keccak256 of row i of a width-w matrix (16w bytes, 2w 64-bit words)
__global__ __launch_bounds__(KB) void k_leaves_rows(const gf *__restrict__ rows, size_t w, uint64_t *__restrict__ out,
                                                    size_t count) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < count;
         i += (size_t)gridDim.x * blockDim.x) {
        uint64_t d[4];
        kk::keccak256_var(reinterpret_cast<const uint64_t *>(rows + i * w), 2 * w, d);
        st_digest(out + 4 * i, d);
    }
}

// out[j] = sum_t scalars[t] * m[j * w + t]  (row-wise dot_product,
// basefold.rs:168-183)
__global__ __launch_bounds__(CB) void k_rowdot(const gf *__restrict__ m, size_t w, const gf *__restrict__ scalars,
                                               gf *__restrict__ out, size_t rows) {
    for (size_t j = blockIdx.x * (size_t)blockDim.x + threadIdx.x; j < rows;
         j += (size_t)gridDim.x * blockDim.x) {
        const gf *row = m + j * w;
        gf acc = gf_zero();
        for (size_t t = 0; t < w; t++) acc = gf_add(acc, gf_mul_reg(gf_ld(row + t), gf_ld(scalars + t)));
        gf_st(out + j, acc);
    }
}

// node i = Keccak256(prev[2i] || prev[2i+1])
__global__ __launch_bounds__(KB) void k_nodes(const uint64_t *__restrict__ prev, uint64_t *__restrict__ out,
                                              size_t count) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < count;
         i += (size_t)gridDim.x * blockDim.x) {
        const ulonglong2 *p = reinterpret_cast<const ulonglong2 *>(prev + 8 * i);
        ulonglong2 a = p[0], b = p[1], c = p[2], e = p[3];
        uint64_t m[8] = {a.x, a.y, b.x, b.y, c.x, c.y, e.x, e.y}, d[4];
        kk::keccak256_words<8>(m, d);
        st_digest(out + 4 * i, d);
    }
}

// Upper Merkle levels, several per launch.  `prev` is a layer of h digests
// (h a power of two) and the next levels follow it contiguously.  Subtree
// `sub` covers the C = 2^lv children [sub * C, (sub + 1) * C), lv <=
// kSubMaxLevels: the block copies them into LDS, hashes level after level
// inside LDS (one call site per permutation form, so each straight-line
// permutation is fetched into the instruction cache once), and finally
// writes every level's digests out with coalesced 16-byte stores.  buf holds the children at
// [0, C) and level l at [2C - 2^(lv-l+1), 2C - 2^(lv-l)).  Node t of a level
// runs on thread t, so a narrow level occupies as few waves as possible: a
// block's waves share SIMDs, and a level spread over all of them runs slower
// than the same level packed into one wave.  A level of at most pair_w nodes
// per block runs node t on the lane pair 2t, 2t+1 instead (keccak_f_pair):
// its latency drops by about a third for 1.36 times the instructions, which
// pays while the level leaves SIMDs idle (the launcher chooses pair_w so that
// paired levels fill at most one wave per SIMD).
constexpr int kSubThreads = 256;
constexpr int kSubMaxLevels = 9;               // 2^(levels-1) level-1 nodes = kSubThreads
constexpr int kSubLds = 2 << kSubMaxLevels;    // digests of LDS per block (2C - 1 used)
__device__ __forceinline__ void subtree(uint64_t *__restrict__ prev, size_t h, int lv, size_t sub,
                                        ulonglong2 (*buf)[2], int pair_w) {
    const int tid = threadIdx.x;
    const int C = 1 << lv;
    const ulonglong2 *src = reinterpret_cast<const ulonglong2 *>(prev + 4 * (sub << lv));
    for (int k = tid; k < 2 * C; k += kSubThreads) buf[k >> 1][k & 1] = src[k];
    __syncthreads();
    int in = 0, out = C;
    for (int l = 1; l <= lv; l++) {
        const int w = C >> l;
        if (w <= pair_w) {  // block-uniform; both lanes of every pair are active
            if (tid < 2 * w) {
                const int t = tid >> 1, odd = tid & 1;
                const ulonglong2 a = buf[in + 2 * t][0], b = buf[in + 2 * t][1];
                const ulonglong2 c = buf[in + 2 * t + 1][0], e = buf[in + 2 * t + 1][1];
                uint64_t m[8] = {a.x, a.y, b.x, b.y, c.x, c.y, e.x, e.y}, d[4];
                kk::keccak256_words_pair<8>(m, d, odd);
                buf[out + t][odd] = odd ? ulonglong2{d[2], d[3]} : ulonglong2{d[0], d[1]};
            }
        } else if (tid < w) {
            // a latency-bound level: the two-step theta, whose single-wave
            // latency is lower (7.9 against 9.1 us)
            const ulonglong2 a = buf[in + 2 * tid][0], b = buf[in + 2 * tid][1];
            const ulonglong2 c = buf[in + 2 * tid + 1][0], e = buf[in + 2 * tid + 1][1];
            uint64_t m[8] = {a.x, a.y, b.x, b.y, c.x, c.y, e.x, e.y}, d[4];
            kk::keccak256_words<8, true>(m, d);
            buf[out + tid][0] = ulonglong2{d[0], d[1]};
            buf[out + tid][1] = ulonglong2{d[2], d[3]};
        }
        __syncthreads();
        in = out;
        out += w;
    }
    uint64_t *lvl = prev + 4 * h;
    size_t cnt = h >> 1;
    int off = C;
    for (int l = 1; l <= lv; l++) {
        const int w = C >> l;
        ulonglong2 *dst = reinterpret_cast<ulonglong2 *>(lvl + 4 * (sub * (size_t)w));
        for (int k = tid; k < 2 * w; k += kSubThreads) dst[k] = buf[off + (k >> 1)][k & 1];
        lvl += 4 * cnt;
        cnt >>= 1;
        off += w;
    }
    __syncthreads();  // buf is refilled by the block's next subtree
}

__global__ __launch_bounds__(kSubThreads) void k_subtree(uint64_t *__restrict__ prev, size_t h, int lv, int pair_w) {
    __shared__ ulonglong2 buf[kSubLds][2];
    const size_t subs = h >> lv;
    for (size_t sub = blockIdx.x; sub < subs; sub += gridDim.x) subtree(prev, h, lv, sub, buf, pair_w);
}

// ---------------------------------------------------------------------------
// open: sumcheck inner product, fold, fix_last_var, eq table
// ---------------------------------------------------------------------------
__device__ __forceinline__ gf wave_xor(gf v) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
        v.lo ^= __shfl_xor(v.lo, off);
        v.hi ^= __shfl_xor(v.hi, off);
    }
    return v;
}

// partial[block] = XOR_k a[k] * b[k]; with `root` set, block 0 also copies
// that 32-byte digest in front of the sums (partial[-2..-1]), so a round's
// pi-tree root and inner-product sums come back in one read-back
__global__ __launch_bounds__(CB) void k_dot(const gf *__restrict__ a, const gf *__restrict__ b, size_t n,
                                            gf *__restrict__ partial, const uint64_t *__restrict__ root) {
    if (root && blockIdx.x == 0 && threadIdx.x < 2) {
        const ulonglong2 v = reinterpret_cast<const ulonglong2 *>(root)[threadIdx.x];
        reinterpret_cast<ulonglong2 *>(partial - 2)[threadIdx.x] = v;
    }
    __shared__ gf red[CB / 64];
    gf acc = gf_zero();
    for (size_t k = blockIdx.x * (size_t)blockDim.x + threadIdx.x; k < n;
         k += (size_t)gridDim.x * blockDim.x)
        acc = gf_add(acc, gf_mul_reg(gf_ld(a + k), gf_ld(b + k)));
    acc = wave_xor(acc);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        gf s = gf_zero();
        for (int w = 0; w < CB / 64; w++) s = gf_add(s, red[w]);
        gf_st(partial + blockIdx.x, s);
    }
}

// out[j] = interpolate(ft[j], w[j], w[j+mid], r)  (binary_reed_solomon.rs:103-109)
__global__ __launch_bounds__(CB) void k_fold(const gf *__restrict__ w, gf *__restrict__ out,
                                             const gf *__restrict__ ft, size_t mid, gf r) {
    const gf_pre pr = gf_prepare(r);
    for (size_t j = blockIdx.x * (size_t)blockDim.x + threadIdx.x; j < mid;
         j += (size_t)gridDim.x * blockDim.x) {
        gf lo = gf_ld(w + j), hi = gf_ld(w + j + mid);
        hi = gf_add(hi, lo);
        lo = gf_add(lo, gf_mul_reg(hi, gf_ld(ft + j)));
        gf_st(out + j, gf_add(lo, gf_mul_pre(gf_add(hi, lo), pr)));
    }
}

// fold of pi and fix_last_var of the table in one launch (both take the
// round's challenge): blocks [0, nfold) fold, the others fix; the fix may run
// in place (out == in), each thread reading only the j and j + half it owns
// TAB: the products by x go through an LDS window table of x built by each
// block (worth it when every thread has a few outputs)
template <bool TAB>
__global__ __launch_bounds__(CB) void k_fold_fix(const gf *__restrict__ w, gf *__restrict__ wout,
                                                 const gf *__restrict__ ft, size_t mid, const gf *in, gf *out,
                                                 size_t half, const gf *__restrict__ x, unsigned nfold) {
    __shared__ GfTab T;
    __shared__ gf u[128];
    gf_pre px;
    if constexpr (TAB) gf_tab_build(T, u, gf_ld(x));
    else px = gf_prepare(gf_ld(x));
    const auto byx = [&](gf v) {
        if constexpr (TAB) return gf_tab_mul(v, T);
        else return gf_mul_pre(v, px);
    };
    if (blockIdx.x < nfold) {
        for (size_t j = blockIdx.x * (size_t)blockDim.x + threadIdx.x; j < mid; j += (size_t)nfold * blockDim.x) {
            gf lo = gf_ld(w + j), hi = gf_ld(w + j + mid);
            hi = gf_add(hi, lo);
            lo = gf_add(lo, gf_mul_reg(hi, gf_ld(ft + j)));
            gf_st(wout + j, gf_add(lo, byx(gf_add(hi, lo))));
        }
    } else {
        const size_t nb = gridDim.x - nfold;
        for (size_t j = (blockIdx.x - nfold) * (size_t)blockDim.x + threadIdx.x; j < half; j += nb * blockDim.x) {
            const gf lo = gf_ld(in + j), hi = gf_ld(in + j + half);
            gf_st(out + j, gf_add(lo, byx(gf_add(hi, lo))));
        }
    }
}

// out[j] = in[j] + x * (in[j+half] - in[j])  (multilinear.rs:183-198)
template <bool TAB>
__global__ __launch_bounds__(CB) void k_fix(const gf *__restrict__ in, gf *__restrict__ out, size_t half, gf x) {
    __shared__ GfTab T;
    __shared__ gf u[128];
    gf_pre px;
    if constexpr (TAB) gf_tab_build(T, u, x);
    else px = gf_prepare(x);
    for (size_t j = blockIdx.x * (size_t)blockDim.x + threadIdx.x; j < half;
         j += (size_t)gridDim.x * blockDim.x) {
        gf lo = gf_ld(in + j), hi = gf_ld(in + j + half);
        const gf d = gf_add(hi, lo);
        if constexpr (TAB) gf_st(out + j, gf_add(lo, gf_tab_mul(d, T)));
        else gf_st(out + j, gf_add(lo, gf_mul_pre(d, px)));
    }
}

// MLE evaluation of every column of a rows x w row-major matrix in one pass
// (basefold.rs:126-131 as the sum over x of eq(point, x) * data[x]):
// column blockIdx.y, eq(point, x) = A[x mod 2^m] * B[x >> m] rebuilt per
// element from the two host-built factors; partial[col * gridDim.x + block]
__global__ __launch_bounds__(CB) void k_mle_dot(const gf *__restrict__ data, size_t w, const gf *__restrict__ A,
                                                int m, const gf *__restrict__ B, size_t rows,
                                                gf *__restrict__ partial) {
    __shared__ gf red[CB / 64];
    const size_t col = blockIdx.y, mask = ((size_t)1 << m) - 1;
    gf acc = gf_zero();
    for (size_t x = blockIdx.x * (size_t)blockDim.x + threadIdx.x; x < rows; x += (size_t)gridDim.x * blockDim.x) {
        const gf e = gf_mul_reg(gf_ld(A + (x & mask)), gf_ld(B + (x >> m)));
        acc = gf_add(acc, gf_mul_reg(e, gf_ld(data + x * w + col)));
    }
    acc = wave_xor(acc);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        gf s = gf_zero();
        for (int v = 0; v < CB / 64; v++) s = gf_add(s, red[v]);
        gf_st(partial + col * gridDim.x + blockIdx.x, s);
    }
}

// eq table as the tensor product of two host-built factors:
// out[lo + (hi << m)] = A[lo] * B[hi]
__global__ __launch_bounds__(CB) void k_outer(const gf *__restrict__ A, int m, const gf *__restrict__ B,
                                              gf *__restrict__ out, size_t n) {
    const size_t mask = ((size_t)1 << m) - 1;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x)
        gf_st(out + i, gf_mul_reg(gf_ld(A + (i & mask)), gf_ld(B + (i >> m))));
}

// ---------------------------------------------------------------------------
// degree-3 product sumcheck  sum_b f(b) g(b) eq(r, b)
//
// With the top variable X split off, eq(r, (b, X)) = T[b] * cf * eq(r_i, X)
// (T the halved eq table, cf its correcting factor), so the round polynomial
// is s * ((1 + r_i) + X) * q(X) with q(X) = cf * sum_b T[b] f_b(X) g_b(X).
// Per pair the kernel accumulates A0 = sum T f0 g0, A1 = sum T f1 g1 and
// A2 = sum T (f0+f1)(g0+g1); the host turns them into q's coefficients.
// ---------------------------------------------------------------------------
__device__ __forceinline__ void block_xor3(gf a0, gf a1, gf a2, gf *partial) {
    __shared__ gf red[3][CB / 64];
    a0 = wave_xor(a0);
    a1 = wave_xor(a1);
    a2 = wave_xor(a2);
    if ((threadIdx.x & 63) == 0) {
        red[0][threadIdx.x >> 6] = a0;
        red[1][threadIdx.x >> 6] = a1;
        red[2][threadIdx.x >> 6] = a2;
    }
    __syncthreads();
    if (threadIdx.x < 3) {
        gf s = gf_zero();
        for (int w = 0; w < CB / 64; w++) s = gf_add(s, red[threadIdx.x][w]);
        gf_st(partial + 3 * blockIdx.x + threadIdx.x, s);
    }
}

// Counts the block once the values it published with device-scope atomics
// are performed (done[0..nd) set after those atomics returned, so no release
// fence, which would write back the whole L2, is needed); true in the block
// that counts last, which then reads the values with device-scope atomic
// loads (gf_ld_published).
__device__ __forceinline__ bool block_counted(unsigned *counter, unsigned total, const unsigned *done, int nd) {
    __shared__ unsigned last;
    if (threadIdx.x == 0) {
        unsigned inc = 1u;
        for (int t = 0; t < nd; t++) inc &= done[t];
        last = __hip_atomic_fetch_add(counter, inc, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == total - 1;
    }
    __syncthreads();
    return last;
}

// Counts the block in two levels (kCountGroups, kernels.hpp): block bid of
// nblocks counts in its group's counter, and the block that counts last in
// its group counts in the launch counter; true in the block that counts last
// there.  Called by one thread with inc = 1 made data-dependent on the
// atomics the count vouches for (so they are performed first); a group's
// last count returns after every member's, and the launch count after it.
__device__ __forceinline__ bool grid_last(unsigned *cnt, unsigned nblocks, unsigned bid, unsigned inc) {
    const unsigned G = nblocks < (unsigned)kCountGroups ? nblocks : (unsigned)kCountGroups;
    const unsigned g = bid % G, members = (nblocks - g + G - 1) / G;
    if (__hip_atomic_fetch_add(cnt + kCountLine * (1 + g), inc, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) !=
        members - 1)
        return false;
    return __hip_atomic_fetch_add(cnt, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == G - 1;
}

// a gf published with device-scope atomics, read in the last block
__device__ __forceinline__ gf gf_ld_published(const gf *p) {
    const unsigned long long *q = (const unsigned long long *)p;
    return gf{(uint64_t)__hip_atomic_load(q, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT),
              (uint64_t)__hip_atomic_load(q + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)};
}

// lane src's value of v, on every lane of the wave
__device__ __forceinline__ gf lane_gf(gf v, int src) { return gf{(uint64_t)__shfl((long long)v.lo, src), (uint64_t)__shfl((long long)v.hi, src)}; }
// v from lane (this lane ^ off)
__device__ __forceinline__ gf gf_shfl_xor(gf v, int off) {
    return gf{(uint64_t)__shfl_xor((long long)v.lo, off), (uint64_t)__shfl_xor((long long)v.hi, off)};
}
// c ? a : b word by word (a select of whole elements on a lane-dependent
// condition can otherwise be lowered through a stack copy)
__device__ __forceinline__ gf gf_sel(bool c, gf a, gf b) { return gf{c ? a.lo : b.lo, c ? a.hi : b.hi}; }

// word v of lane Q of this lane's quad (DPP quad_perm [Q, Q, Q, Q])
template <int Q>
__device__ __forceinline__ uint32_t quad_lane(uint32_t v) {
    return (uint32_t)__builtin_amdgcn_mov_dpp((int)v, Q * 0x55, 0xF, 0xF, true);
}

// a * b with four lanes per product, for the serial chains of the
// transcript steps: lanes base + l, l < 16, get a * b for the operands they
// hold (the other lanes' results are unspecified; every lane of the wave
// must take part).  Lane 4l + u forms one 64x64 Karatsuba product of lane
// base + l's operands (u = 0 the low halves, 1 the high halves, 2 and 3
// their sums), the quad exchanges them with DPP, and lane base + l takes the
// folded result back: about 170 instructions on the chain instead of about
// 380.
__device__ __forceinline__ gf gf_mul_wide(gf a, gf b, int base = 0) {
    const int lane = threadIdx.x & 63, u = lane & 3;
    const gf A = lane_gf(a, base + (lane >> 2)), B = lane_gf(b, base + (lane >> 2));
    const uint64_t x = u == 0 ? A.lo : (u == 1 ? A.hi : A.lo ^ A.hi);
    const uint64_t y = u == 0 ? B.lo : (u == 1 ? B.hi : B.lo ^ B.hi);
    const pieces x0 = split4((uint32_t)x), x1 = split4((uint32_t)(x >> 32));
    const pieces y0 = split4((uint32_t)y), y1 = split4((uint32_t)(y >> 32));
    uint32_t r[4], L[4], H[4], Mm[4];
    kara64(x0, x1, pxor(x0, x1), y0, y1, pxor(y0, y1), r);
#pragma unroll
    for (int q = 0; q < 4; q++) {
        L[q] = quad_lane<0>(r[q]);
        H[q] = quad_lane<1>(r[q]);
        Mm[q] = quad_lane<2>(r[q]);
    }
    return lane_gf(gf_combine(L, H, Mm), 4 * (lane - base));
}

// Pair sums of the product sumcheck with the pair's eq weight t multiplied
// into f first: tf0 = t f0, tf1 = t f1, then A0 += tf0 g0, A1 += tf1 g1 and
// A2 += (tf0 + tf1)(g0 + g1), five products per pair.
__device__ __forceinline__ void prod_terms(gf f0, gf f1, gf g0, gf g1, gf t, gf &a0, gf &a1, gf &a2) {
    const gf_pre pt = gf_prepare(t);
    const gf tf0 = gf_mul_pre(f0, pt), tf1 = gf_mul_pre(f1, pt);
    a0 = gf_add(a0, gf_mul_reg(tf0, g0));
    a1 = gf_add(a1, gf_mul_reg(tf1, g1));
    a2 = gf_add(a2, gf_mul_reg(gf_add(tf0, tf1), gf_add(g0, g1)));
}

// fold on the top variable: lo + x (hi - lo)  (multilinear.rs:183-198)
__device__ __forceinline__ gf fold_top(gf lo, gf hi, const gf_pre &px) { return gf_add(lo, gf_mul_pre(gf_add(hi, lo), px)); }

// The eq scalars s_k, K_k, L_k of round k (see ProdFs): round 0 takes them as
// uploaded; a later round moves s_{k-1}, P1, P2 on by t = e0[i + 1] +
// x_{k-1} with one product on lanes 0..2 of the calling wave.  Every lane
// gets all three.
__device__ __forceinline__ void prod_scalars(const ProdFs *fs, const gf *e0s, int k, gf &s, gf &K, gf &L) {
    if (k == 0) {
        s = fs->s;
        K = fs->K;
        L = fs->L;
        return;
    }
    const int lane = threadIdx.x & 63;
    const gf t = gf_add(gf_ld(e0s + (fs->n - k)), fs->x);
    const gf p = gf_mul_wide(gf_sel(lane == 0, fs->s, gf_sel(lane == 1, fs->P1, fs->P2)), t);
    s = lane_gf(p, 0);
    K = lane_gf(p, 1);
    L = lane_gf(p, 2);
}

// The round polynomial p(X) = s (e0 + X) cf (A0 + Q1 X + A2 X^2) from the
// pair sums A0, A1, A2 (Q1 = A0 + A1 + A2) and the eq scalars K = s e0 cf,
// L = s cf: c0 = K A0, c1 = L A0 + K Q1, c2 = L Q1 + K A2, c3 = L A2, six
// independent products on lanes 0..5 of the calling wave; every lane gets
// the four.
struct ProdPoly {
    gf c0, c1, c2, c3;
};
__device__ __forceinline__ ProdPoly prod_coeffs(gf A0, gf A1, gf A2, gf K, gf L) {
    const int lane = threadIdx.x & 63;
    const gf Q1 = gf_add(A0, gf_add(A1, A2));
    const gf u = gf_sel(lane & 1, L, K);
    const gf v = gf_sel(lane < 2, A0, gf_sel(lane < 4, Q1, A2));
    const gf p = gf_mul_wide(u, v);  // lane 0: K A0, 1: L A0, 2: K Q1, 3: L Q1, 4: K A2, 5: L A2
    return ProdPoly{lane_gf(p, 0), gf_add(lane_gf(p, 1), lane_gf(p, 2)), gf_add(lane_gf(p, 3), lane_gf(p, 4)),
                    lane_gf(p, 5)};
}

// One round's transcript step of the product sumcheck, run by one whole wave
// (prove_sumcheck_round, sumcheck/src/lib.rs:81-103), given the round
// polynomial: observe [claim,] c0, c2, c3 (the claim p(0) + p(1) =
// c1 + c2 + c3 in round 0, sumcheck/src/lib.rs:67) and sample x.  With
// have_prev the challenger's input is the previous round's 32-byte digest
// prev (every round after the first); otherwise the input bytes come from the
// region (the first round, after whatever the caller observed; lane 0
// appends to them there).  Every lane gets the results; the caller stores
// them.
struct ProdRound {
    gf c0, c2, c3, claim, x;
    uint64_t d[4];
};
struct Digest {
    uint64_t w[4];
};
__device__ __forceinline__ ProdRound prod_observe(ProdFs *st, int k, const ProdPoly &c, bool have_prev, Digest prev) {
    const int lane = threadIdx.x & 63;
    uint8_t *in = (uint8_t *)st + prod_input_off(st->n);
    ProdRound R;
    R.c0 = c.c0;
    R.c2 = c.c2;
    R.c3 = c.c3;
    R.claim = gf_add(c.c1, gf_add(c.c2, c.c3));
    if (have_prev) {
        const uint64_t m[10] = {prev.w[0], prev.w[1], prev.w[2], prev.w[3], R.c0.lo, R.c0.hi,
                                R.c2.lo,   R.c2.hi,   R.c3.lo,   R.c3.hi};
        kk::keccak256_words_pair<10>(m, R.d, lane & 1);
    } else {
        if (lane == 0) {
            uint32_t len = st->in_len;
            const auto put = [&](uint64_t w) {
#pragma unroll 1
                for (int t = 0; t < 8; t++) in[len + t] = (uint8_t)(w >> (8 * t));
                len += 8;
            };
            if (k == 0) {  // the claim is observed first (sumcheck/src/lib.rs:67)
                put(R.claim.lo);
                put(R.claim.hi);
            }
            put(R.c0.lo);
            put(R.c0.hi);
            put(R.c2.lo);
            put(R.c2.hi);
            put(R.c3.lo);
            put(R.c3.hi);
            kk::keccak256_bytes(in, len, R.d);
        }
#pragma unroll
        for (int q = 0; q < 4; q++) R.d[q] = (uint64_t)__shfl((long long)R.d[q], 0);
    }
    // the challenge is the last 16 digest bytes, last byte first
    R.x = gf{__builtin_bswap64(R.d[3]), __builtin_bswap64(R.d[2])};
    return R;
}
// the polynomial from the pair sums and eq scalars, then the step
__device__ __forceinline__ ProdRound prod_transcript(ProdFs *st, int k, gf A0, gf A1, gf A2, gf K, gf L, bool have_prev,
                                     Digest prev) {
    return prod_observe(st, k, prod_coeffs(A0, A1, A2, K, L), have_prev, prev);
}

// The challenger state after a round into the region (one lane): input =
// output = the digest, 16 output bytes left unread, and the challenge.
__device__ __forceinline__ void prod_store_round(ProdFs *st, const ProdRound &R) {
    uint64_t *in64 = reinterpret_cast<uint64_t *>((uint8_t *)st + prod_input_off(st->n));
    uint64_t *out64 = reinterpret_cast<uint64_t *>(st->out);
#pragma unroll
    for (int q = 0; q < 4; q++) {
        in64[q] = R.d[q];
        out64[q] = R.d[q];
    }
    st->in_len = 32;
    st->out_len = 16;
    st->x = R.x;
}

// Rounds after the first: fold the previous tables (4 q entries, q = half)
// on their top variable with x and form the pair sums of pairs [j0, j1).
// Entries j and j + q of the folded tables come from j, j + 2q and j + q,
// j + 3q: each thread reads its four before writing its two, so in place is
// safe.  TAB: the products by x through the window table T.
template <bool TAB>
__device__ __forceinline__ void prod_fold_sums(const gf *fsrc, const gf *gsrc, gf *fdst, gf *gdst,
                                               const EqFactors &eq, size_t q, size_t j0, size_t j1, size_t amask,
                                               const GfTab *T, const gf_pre &px, gf &a0, gf &a1, gf &a2) {
    const auto fold = [&](gf lo, gf hi) {
        if constexpr (TAB) return gf_add(lo, gf_tab_mul(gf_add(hi, lo), *T));
        else return fold_top(lo, hi, px);
    };
    for (size_t j = j0 + threadIdx.x; j < j1; j += CB) {
        const gf fa = gf_ld(fsrc + j), fb = gf_ld(fsrc + j + q), fc = gf_ld(fsrc + j + 2 * q),
                 fd = gf_ld(fsrc + j + 3 * q);
        const gf ga = gf_ld(gsrc + j), gb = gf_ld(gsrc + j + q), gc = gf_ld(gsrc + j + 2 * q),
                 gd = gf_ld(gsrc + j + 3 * q);
        const gf f0 = fold(fa, fc), f1 = fold(fb, fd);
        const gf g0 = fold(ga, gc), g1 = fold(gb, gd);
        gf_st(fdst + j, f0);
        gf_st(fdst + j + q, f1);
        gf_st(gdst + j, g0);
        gf_st(gdst + j + q, g1);
        prod_terms(f0, f1, g0, g1, gf_ld(eq.A + (j & amask)), a0, a1, a2);
    }
}

// Round k of the product sumcheck (see launch::prod_step).  MODE 0: round 0,
// sums over the caller's tables; 1: fold the caller's tables into the working
// ones, then the sums; 2: fold the working tables in place, then the sums.
// Block b takes pairs [b C, (b + 1) C), C = half / gridDim.x,
// which share one entry of the eq factor B: pair j is weighted by
// A[j mod 2^lo] alone and the block's sums by that B entry.  A block
// publishes its share of the round polynomial, (K B, L B) times its sums
// (prod_coeffs), so the block that finishes last only adds the shares
// before its transcript step; meanwhile its second wave forms the eq
// scalars for the next round (P1, P2, ProdFs).
template <int MODE>
__global__ __launch_bounds__(CB) void k_prod_round(const gf *fsrc, const gf *gsrc, gf *fdst, gf *gdst, EqFactors eq,
                                                   size_t half, int k, ProdFs *fs) {
    __shared__ gf sc[3];  // K B, L B of this block, s_k
    __shared__ gf red[3][CB / 64];
    __shared__ unsigned last;
    const uint32_t nv = fs->n;
    const int iv = (int)nv - 1 - k;
    const int tid = threadIdx.x, wave = __builtin_amdgcn_readfirstlane(tid >> 6), lane = tid & 63;
    const gf *cfs = (const gf *)((const uint8_t *)fs + prod_cf_off()), *e0s = cfs + nv, *ecfs = e0s + nv;
    uint8_t *base = (uint8_t *)fs;
    const uint32_t in_cap = fs->in_cap;
    unsigned long long *acc0 = (unsigned long long *)(base + prod_acc_off(nv, in_cap));
    // the next round's slots and counters start at zero (this launch ends
    // before it runs)
    if (k + 1 < (int)nv) {
        unsigned long long *nxt = acc0 + (size_t)(k + 1) * (kProdAccRound / 8);
        for (size_t w = blockIdx.x * (size_t)CB + tid; w < kProdAccRound / 8; w += (size_t)gridDim.x * CB)
            nxt[w] = 0;
    }
    const size_t C = half / gridDim.x, j0 = blockIdx.x * C, j1 = j0 + C;
    const size_t amask = ((size_t)1 << eq.lo) - 1;
    const gf Bv = gf_ld(eq.B + (j0 >> eq.lo));
    // the eq scalars on the second wave, before its pairs: s_k, K_k, L_k,
    // then K_k B and L_k B
    if (wave == 1) {
        gf s, K, L;
        prod_scalars(fs, e0s, k, s, K, L);
        const gf p = gf_mul_reg(gf_sel(lane == 0, K, L), Bv);
        const gf kb = lane_gf(p, 0), lb = lane_gf(p, 1);
        if (lane == 0) {
            sc[0] = kb;
            sc[1] = lb;
            sc[2] = s;
        }
    }
    gf a0 = gf_zero(), a1 = gf_zero(), a2 = gf_zero();
    if constexpr (MODE == 0) {
        for (size_t j = j0 + tid; j < j1; j += CB)
            prod_terms(gf_ld(fsrc + j), gf_ld(fsrc + j + half), gf_ld(gsrc + j), gf_ld(gsrc + j + half),
                       gf_ld(eq.A + (j & amask)), a0, a1, a2);
    } else if (C >= 2 * (size_t)CB) {
        // two pairs or more per thread: the folds multiply by x through a
        // window table
        __shared__ GfTab T;
        __shared__ gf u[128];
        gf_tab_build(T, u, fs->x);
        prod_fold_sums<true>(fsrc, gsrc, fdst, gdst, eq, half, j0, j1, amask, &T, gf_pre{}, a0, a1, a2);
    } else {
        prod_fold_sums<false>(fsrc, gsrc, fdst, gdst, eq, half, j0, j1, amask, nullptr, gf_prepare(fs->x), a0,
                              a1, a2);
    }
    a0 = wave_xor(a0);
    a1 = wave_xor(a1);
    a2 = wave_xor(a2);
    if (lane == 0) {
        red[0][wave] = a0;
        red[1][wave] = a1;
        red[2][wave] = a2;
    }
    __syncthreads();
    const uint32_t n = nv;
    unsigned long long *acc = acc0 + (size_t)k * (kProdAccRound / 8);
    unsigned *counter = (unsigned *)(acc + kProdSlotBytes / 8);
    // The block's share of the polynomial goes out as device-scope atomics
    // whose results it waits for, so they are performed before its count
    // below is issued; every block's shares are therefore in the slots before
    // the last count returns.  No block needs a release fence (which would
    // write back the whole L2): the folded tables are read only by the next
    // launch.
    if (wave == 0) {
        gf A[3];
#pragma unroll
        for (int c = 0; c < 3; c++) {
            A[c] = gf_zero();
            for (int w = 0; w < CB / 64; w++) A[c] = gf_add(A[c], red[c][w]);
        }
        const ProdPoly cs = prod_coeffs(A[0], A[1], A[2], sc[0], sc[1]);
        // c0..c3 on lanes 0..3
        const gf v = gf_sel(lane < 2, gf_sel(lane == 0, cs.c0, cs.c1), gf_sel(lane == 2, cs.c2, cs.c3));
        unsigned done = 1u;
        if (lane < 4) {
            unsigned long long *slot = acc + 8 * (blockIdx.x % kProdSlots) + 2 * lane;
            const unsigned long long o0 =
                __hip_atomic_fetch_xor(slot, (unsigned long long)v.lo, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            const unsigned long long o1 =
                __hip_atomic_fetch_xor(slot + 1, (unsigned long long)v.hi, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            done = (unsigned)(o0 ^ o1) | 1u;
        }
        // 1, after lanes 0..3's atomics returned
        const unsigned inc = done & (unsigned)__shfl((int)done, 1) & (unsigned)__shfl((int)done, 2) &
                             (unsigned)__shfl((int)done, 3) & 1u;
        if (lane == 0) last = grid_last(counter, gridDim.x, blockIdx.x, inc);
    }
    __syncthreads();
    if (!last) return;
    if (wave == 1) {
        // the eq scalars for the next round (independent of its challenge)
        if (iv > 0) {
            const gf sk = sc[2];
            const gf p = gf_mul_reg(sk, (lane & 1) ? gf_ld(cfs + iv - 1) : gf_ld(ecfs + iv - 1));
            if (lane == 0) {
                fs->s = sk;
                fs->P1 = p;
            } else if (lane == 1) {
                fs->P2 = p;
            }
        }
        return;
    }
    if (wave != 0) return;
    // lane l XORs slots l, l + 64, ... (the slots a smaller grid used), all
    // loads issued before the first use; then a wave reduction
    unsigned long long v[8 * (kProdSlots / 64)];
#pragma unroll
    for (int t = 0; t < kProdSlots / 64; t++)
#pragma unroll
        for (int c = 0; c < 8; c++)
            v[8 * t + c] = (size_t)(lane + 64 * t) < gridDim.x
                               ? __hip_atomic_load(acc + 8 * (lane + 64 * t) + c, __ATOMIC_RELAXED,
                                                   __HIP_MEMORY_SCOPE_AGENT)
                               : 0ull;
    gf cv[4];
#pragma unroll
    for (int c = 0; c < 4; c++) {
        cv[c] = gf_zero();
#pragma unroll
        for (int t = 0; t < kProdSlots / 64; t++) cv[c] = gf_add(cv[c], gf{v[8 * t + 2 * c], v[8 * t + 2 * c + 1]});
        cv[c] = wave_xor(cv[c]);
    }
    const ProdPoly cs{cv[0], cv[1], cv[2], cv[3]};
    // after the first round the challenger's input is the last digest
    Digest prev;
    const uint64_t *in64 = reinterpret_cast<const uint64_t *>(base + prod_input_off(n));
#pragma unroll
    for (int w = 0; w < 4; w++) prev.w[w] = k > 0 ? in64[w] : 0;
    const ProdRound R = prod_observe(fs, k, cs, k > 0, prev);
    if (lane == 0) {
        gf *polys = (gf *)(base + prod_polys_off(n)), *chal = (gf *)(base + prod_chal_off(n));
        polys[3 * k] = R.c0;
        polys[3 * k + 1] = R.c2;
        polys[3 * k + 2] = R.c3;
        chal[k] = R.x;
        if (k == 0) fs->claim = R.claim;
        prod_store_round(fs, R);
    }
}

// Rounds k >= 1 of at most kProdLookMax pairs with one round of lookahead
// (see launch::prod_look).  With q the pairs of round k, eight lanes take pair
// j < q/2 of round k+1: lane u folds entry j + (u & 3) q/2 of f (u < 4) or g
// into round k's tables, then the f lanes multiply by A[j] (the block's B
// entry enters its sums once), giving a0, a1, b0, b1 and c0, c1, d0, d1 as in
// k_prod_tail, and the nine Karatsuba products p0..p8 of round k+1's
// quadratics (p0 = a0 c0, p1 = b0 d0, p2 = (a0 + b0)(c0 + d0), p3..p5 the
// same with index 1, p6..p8 for the sums of both).  The same products are
// round k's pair sums: pair j has A0 = p0, A1 = p2, A2 = p1 (a0 + b0 = T V2
// and c0 + d0 = W2), and pair j + q/2, weighted by rho T[j] (rho the ratio of
// eq entries whose indices differ in the bit of q/2), A0 = p3, A1 = p5,
// A2 = p4.  Each worker block XORs B times its nine sums into slot (block mod
// kLookSlots) of the set for round k+1.  FIRST: every block works; the last
// (two-level count) reads round k+1's set, forms round k's sums with rho and
// runs round k's transcript step.  Otherwise block 0 runs round k's step at
// once, from the set the previous launch left and x_{k-1} (Horner, as in
// k_prod_tail), and clears that set, while blocks 1.. work.
constexpr size_t kProdLookMax = (size_t)1 << 17;
constexpr int kLookGroups = CB / 8;  // pairs of round k+1 per worker block

// NS lookahead sums (lane u of wave 0 holds sum u) into slot (wb mod
// kLookSlots) of a set; with wait, done is made nonzero from the atomics'
// results (so a count issued after it follows them)
template <int NS>
__device__ __forceinline__ void look_publish(gf *set, unsigned wb, gf v, bool wait, unsigned &done) {
    const int lane = threadIdx.x & 63;
    done = 1u;
    if (lane < NS) {
        unsigned long long *q = (unsigned long long *)(set + NS * (wb % kLookSlots) + lane);
        const unsigned long long o0 =
            __hip_atomic_fetch_xor(q, (unsigned long long)v.lo, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        const unsigned long long o1 =
            __hip_atomic_fetch_xor(q + 1, (unsigned long long)v.hi, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (wait) done = (unsigned)(o0 ^ o1) | 1u;
    }
}

// the NS sums of a set on lanes 0..NS-1 of the calling wave: lane l reads sum
// l % NS of slots l / NS, l / NS + 64 / NS, ... (all loads issued before the
// first use), then lanes 0..NS-1 add the values of their sum; with clear the
// slots are zeroed
template <int NS>
__device__ __forceinline__ gf look_gather(gf *set, bool clear) {
    constexpr int per = 64 / NS, passes = (kLookSlots + per - 1) / per;
    const int lane = threadIdx.x & 63, c = lane % NS;
    gf v[passes];
#pragma unroll
    for (int r = 0; r < passes; r++) {
        const int sl = per * r + lane / NS;
        v[r] = gf_zero();
        if (lane < per * NS && sl < kLookSlots) v[r] = gf_ld_published(set + NS * sl + c);
    }
    if (clear) {
#pragma unroll
        for (int r = 0; r < passes; r++) {
            const int sl = per * r + lane / NS;
            if (lane < per * NS && sl < kLookSlots) gf_st(set + NS * sl + c, gf_zero());
        }
    }
    gf P = gf_zero();
    const int src = lane < NS ? lane : 0;
#pragma unroll
    for (int r = 0; r < passes; r++)
#pragma unroll
        for (int g = 0; g < per; g++) P = gf_add(P, lane_gf(v[r], src + NS * g));
    return gf_sel(lane < NS, P, gf_zero());
}

template <bool FIRST>
__global__ __launch_bounds__(CB) void k_prod_look(const gf *fsrc, const gf *gsrc, gf *fdst, gf *gdst, EqFactors eq,
                                                  size_t half, int k, ProdFs *fs) {
    __shared__ gf red[9][CB / 64];
    __shared__ unsigned last;
    const uint32_t n = fs->n;
    const int iv = (int)n - 1 - k;
    const int tid = threadIdx.x, wave = __builtin_amdgcn_readfirstlane(tid >> 6), lane = tid & 63;
    uint8_t *base = (uint8_t *)fs;
    const uint32_t in_cap = fs->in_cap;
    const gf *cfs = (const gf *)(base + prod_cf_off()), *e0s = cfs + n, *ecfs = e0s + n, *rhos = ecfs + n;
    gf *chal = (gf *)(base + prod_chal_off(n)), *polys = (gf *)(base + prod_polys_off(n));
    gf *look = (gf *)(base + prod_look_off(n, in_cap));
    gf *set_in = look + (size_t)9 * kLookSlots * (k & 1), *set_out = look + (size_t)9 * kLookSlots * ((k + 1) & 1);
    const gf xp = gf_ld(chal + k - 1);  // x_{k-1}
    if (!FIRST && blockIdx.x == 0) {
        // round k's transcript step from the lookahead sums and x_{k-1}
        if (wave != 0) return;
        const gf P = look_gather<9>(set_in, true);
        const int c = lane < 3 ? lane : 0;
        const gf al = lane_gf(P, 3 * c), ga = lane_gf(P, 3 * c + 1), mid = lane_gf(P, 3 * c + 2);
        const gf be = gf_add(mid, gf_add(al, ga));
        const gf t = gf_add(gf_ld(e0s + iv + 1), xp);
        const gf s0 = fs->s, P1 = fs->P1, P2 = fs->P2;
        // lanes 0..2 x gamma_c, lanes 3..5 {s, P1, P2} t
        const gf r1 = gf_mul_wide(gf_sel(lane < 3, xp, gf_sel(lane == 3, s0, gf_sel(lane == 4, P1, P2))),
                                 gf_sel(lane < 3, ga, t));
        const gf tc = gf_add(be, r1);
        const gf s = lane_gf(r1, 3), K = lane_gf(r1, 4), L = lane_gf(r1, 5);
        // lanes 0..2 x (beta_c + x gamma_c), lanes 3, 4 s {ecf, cf}[iv - 1]
        const gf r2 = gf_mul_wide(gf_sel(lane < 3, xp, s),
                                 gf_sel(lane < 3, tc, iv > 0 ? gf_sel(lane == 3, gf_ld(ecfs + iv - 1), gf_ld(cfs + iv - 1))
                                                            : gf_zero()));
        const gf Ac = gf_add(al, r2);
        const gf P1n = lane_gf(r2, 3), P2n = lane_gf(r2, 4);
        Digest prev;
        const uint64_t *in64 = reinterpret_cast<const uint64_t *>(base + prod_input_off(n));
#pragma unroll
        for (int w = 0; w < 4; w++) prev.w[w] = in64[w];
        const ProdRound R =
            prod_observe(fs, k, prod_coeffs(lane_gf(Ac, 0), lane_gf(Ac, 1), lane_gf(Ac, 2), K, L), true, prev);
        if (lane == 0) {
            polys[3 * k] = R.c0;
            polys[3 * k + 1] = R.c2;
            polys[3 * k + 2] = R.c3;
            chal[k] = R.x;
            prod_store_round(fs, R);
            fs->s = s;
            if (iv > 0) {
                fs->P1 = P1n;
                fs->P2 = P2n;
            }
        }
        return;
    }
    // worker: pair j of round k+1
    const unsigned wb = FIRST ? blockIdx.x : blockIdx.x - 1;
    const size_t q = half, qh = q / 2;
    const size_t j0 = (size_t)wb * kLookGroups, j = j0 + (size_t)(tid >> 3);
    const int u = lane & 7, tb = u >> 2, i = u & 3, gb = lane & ~7;
    const size_t e = j + (size_t)i * qh;
    const gf *src = tb ? gsrc : fsrc;
    gf *dst = tb ? gdst : fdst;
    const gf lo = gf_ld(src + e);
    const gf V = gf_add(lo, gf_mul_reg(gf_add(gf_ld(src + e + 2 * q), lo), xp));
    gf_st(dst + e, V);
    const gf X1 = gf_sel(i >= 2, gf_add(V, lane_gf(V, lane ^ 2)), V);
    const gf M = gf_mul_reg(X1, gf_ld(eq.A + (j & (((size_t)1 << eq.lo) - 1))));
    const gf X = gf_sel(tb == 0, M, X1);
    const int s1 = (u == 0 || u == 2 || u == 6) ? 0 : ((u == 1 || u == 7) ? 2 : ((u == 3 || u == 5) ? 1 : 3));
    const int s2 = u == 2 ? 2 : (u == 5 ? 3 : (u == 6 ? 1 : (u == 7 ? 3 : -1)));
    const gf xa = lane_gf(X, gb + s1), xb = lane_gf(X, gb + (s2 < 0 ? s1 : s2));
    const gf ya = lane_gf(X, gb + 4 + s1), yb = lane_gf(X, gb + 4 + (s2 < 0 ? s1 : s2));
    gf a = gf_mul_reg(gf_sel(s2 < 0, xa, gf_add(xa, xb)), gf_sel(s2 < 0, ya, gf_add(ya, yb)));
    gf Y = gf_add(X, gf_shfl_xor(X, 1));
    Y = gf_add(Y, gf_shfl_xor(Y, 2));
    gf a8 = gf_mul_reg(lane_gf(Y, gb), lane_gf(Y, gb + 4));
    a8 = gf_sel(u == 0, a8, gf_zero());
#pragma unroll
    for (int off = 8; off < 64; off <<= 1) {
        a = gf_add(a, gf_shfl_xor(a, off));
        a8 = gf_add(a8, gf_shfl_xor(a8, off));
    }
    if (lane < 8) red[lane][wave] = a;
    if (lane == 0) red[8][wave] = a8;
    __syncthreads();
    if (wave == 0) {
        gf v = gf_zero();
        if (lane < 9)
            for (int w = 0; w < CB / 64; w++) v = gf_add(v, red[lane][w]);
        v = gf_mul_reg(v, gf_ld(eq.B + (j0 >> eq.lo)));
        unsigned done;
        look_publish<9>(set_out, wb, v, FIRST, done);
        if (FIRST) {
            // 1, after lanes 0..8's atomics returned
            unsigned inc = done;
#pragma unroll
            for (int l = 1; l < 9; l++) inc &= (unsigned)__shfl((int)done, l);
            unsigned long long *acc = (unsigned long long *)(base + prod_acc_off(n, in_cap)) + (size_t)k * (kProdAccRound / 8);
            if (lane == 0) last = grid_last((unsigned *)(acc + kProdSlotBytes / 8), gridDim.x, blockIdx.x, inc & 1u);
        }
    }
    if (!FIRST) return;
    __syncthreads();
    if (!last || wave != 0) return;
    // round k's transcript step: its pair sums from round k+1's set
    const gf P = look_gather<9>(set_out, false);
    // lanes 0..2: A0 = p0 + rho p3, A1 = p2 + rho p5, A2 = p1 + rho p4;
    // lanes 3..5: s, K, L
    const int hi_i = lane == 0 ? 3 : (lane == 1 ? 5 : 4), lo_i = lane == 0 ? 0 : (lane == 1 ? 2 : 1);
    const gf Ph = lane_gf(P, lane < 3 ? hi_i : 0), Pl = lane_gf(P, lane < 3 ? lo_i : 0);
    const gf t = gf_add(gf_ld(e0s + iv + 1), xp);
    const gf rho = gf_ld(rhos + (n - 2 - k));
    const gf r1 = gf_mul_wide(gf_sel(lane < 3, Ph, gf_sel(lane == 3, fs->s, gf_sel(lane == 4, fs->P1, fs->P2))),
                             gf_sel(lane < 3, rho, t));
    const gf Ac = gf_add(Pl, r1);
    const gf K = lane_gf(r1, 4), L = lane_gf(r1, 5);
    Digest prev;
    const uint64_t *in64 = reinterpret_cast<const uint64_t *>(base + prod_input_off(n));
#pragma unroll
    for (int w = 0; w < 4; w++) prev.w[w] = in64[w];
    const ProdRound R = prod_observe(fs, k, prod_coeffs(lane_gf(Ac, 0), lane_gf(Ac, 1), lane_gf(Ac, 2), K, L), true, prev);
    // the eq scalars for round k + 1 (s_k, then P1, P2)
    const gf s = lane_gf(r1, 3);
    const gf p = gf_mul_wide(s, iv > 0 ? gf_sel(lane & 1, gf_ld(cfs + iv - 1), gf_ld(ecfs + iv - 1)) : gf_zero());
    if (lane == 0) {
        polys[3 * k] = R.c0;
        polys[3 * k + 1] = R.c2;
        polys[3 * k + 2] = R.c3;
        chal[k] = R.x;
        prod_store_round(fs, R);
        fs->s = s;
        if (iv > 0) fs->P1 = p;
    } else if (lane == 1 && iv > 0) {
        fs->P2 = p;
    }
}

// The rounds from k0 on in one block of kTailThreads threads, tables in LDS,
// with one round of lookahead so that a round's table work runs while the
// round before it samples its challenge.  Round k+1's pair sums are
// quadratics in x_k: with q the pairs of round k, j < q/2 a pair of round
// k+1, V_i = F_k[j + i q/2] and W_i = G_k[j + i q/2] (i < 4), the folded
// entries are f0' = V0 + x (V2 + V0), f1' = V1 + x (V3 + V1) (the same for g
// with W), so with T = T[j] (the halved eq table)
//   A0' = sum (a0 + x b0)(c0 + x d0),   A1' = sum (a1 + x b1)(c1 + x d1),
//   A2' = sum ((a0 + a1) + x (b0 + b1))((c0 + c1) + x (d0 + d1)),
//   a0 = T V0, b0 = T (V2 + V0), a1 = T V1, b1 = T (V3 + V1),
//   c0 = W0, d0 = W2 + W0, c1 = W1, d1 = W3 + W1,
// and three Karatsuba products per sum give its coefficients of 1, x, x^2.
// Wave 0 runs the transcript: in round k it evaluates the three quadratics
// at x_{k-1} and moves the eq scalars (two product steps on six lanes),
// forms the round polynomial (one step) and samples x_k.  Meanwhile the
// worker waves fold F_{k-1} into F_k with x_{k-1} and form round k+1's
// coefficient sums from F_k, eight lanes per pair of round k+1: one fold
// each, one product by T on the f lanes, then the nine products.  Wave 4,
// which shares wave 0's SIMD, sits the rounds out.  Round k0's sums are
// formed directly before the first round, four lanes per pair.  (The
// transcript step needs about 160 VGPRs, so the block stays at 512 threads.)
constexpr int kTailThreads = 512;
constexpr int kTailWaves = kTailThreads / 64;
constexpr size_t kTailPairs = 256;  // pairs of the first tail round (4 * kTailPairs entries per table)


// prod_transcript behind a call: its Keccak-f keeps its own registers
// instead of sharing an allocation with the tail kernel's table work
__device__ __noinline__ ProdRound prod_transcript_call(ProdFs *st, int k, gf A0, gf A1, gf A2, gf K, gf L,
                                                       bool have_prev, Digest prev) {
    return prod_transcript(st, k, A0, A1, A2, K, L, have_prev, prev);
}

// The region's first `bytes` bytes (a multiple of 8) into pinned host memory,
// by the whole block, after the block's own writes to the region: every
// thread fences and meets the others, then reads through L2 (agent-scope
// loads, so no stale line of this CU's L1 is used) and stores to the host.
// Takes the place of a device-to-host copy after the launch.
__device__ __forceinline__ void region_to_host(const uint8_t *region, uint8_t *host, uint32_t bytes) {
    __threadfence();
    __syncthreads();
    const uint64_t *src = reinterpret_cast<const uint64_t *>(region);
    uint64_t *dst = reinterpret_cast<uint64_t *>(host);
    for (uint32_t w = threadIdx.x; w < bytes / 8; w += blockDim.x)
        dst[w] = __hip_atomic_load(src + w, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__global__ __launch_bounds__(kTailThreads) void k_prod_tail(const gf *fsrc, const gf *gsrc, gf *fout, gf *gout,
                                                            EqFactors eq, size_t half, int k0, ProdFs *fs, uint8_t *host, uint32_t host_bytes) {
    // F_k of round k lives in Fb when k - k0 is even, in Fa when odd; Fa
    // first holds the input tables when k0 > 0 (4 half entries)
    __shared__ gf Fa[4 * kTailPairs], Ga[4 * kTailPairs], Fb[2 * kTailPairs], Gb[2 * kTailPairs];
    __shared__ gf Ts[kTailPairs];
    __shared__ gf redA[3][kTailWaves];     // round k0's pair sums per wave
    __shared__ gf red[2][9][kTailWaves];   // the next round's coefficient sums per wave (by round parity)
    __shared__ gf sx[2];                   // challenge of the round just sampled (by round parity)
    __shared__ gf cst[3][32];              // cf, e0, ecf
    __shared__ gf outs[32][4];             // per round: c0, c2, c3, x (stored to the region at the end)
    const uint32_t n = fs->n;
    const int tid = threadIdx.x, wave = __builtin_amdgcn_readfirstlane(tid >> 6), lane = tid & 63;
    const size_t q0 = half;
    if (k0 == 0) {
        for (size_t e = tid; e < 2 * q0; e += kTailThreads) {
            Fb[e] = gf_ld(fsrc + e);
            Gb[e] = gf_ld(gsrc + e);
        }
    } else {
        for (size_t e = tid; e < 4 * q0; e += kTailThreads) {
            Fa[e] = gf_ld(fsrc + e);
            Ga[e] = gf_ld(gsrc + e);
        }
    }
    for (size_t e = tid; e < q0; e += kTailThreads)
        Ts[e] = gf_mul_reg(gf_ld(eq.A + (e & (((size_t)1 << eq.lo) - 1))), gf_ld(eq.B + (e >> eq.lo)));
    if (tid < 3 * (int)n) cst[tid / n][tid % n] = gf_ld((const gf *)((const uint8_t *)fs + prod_cf_off()) + tid);
    __syncthreads();
    // round k0's pair sums (after folding the input with x_{k0-1} into F_k0
    // when k0 > 0), four lanes per pair: f0, f1, g0, g1, then t f0, t f1,
    // then the three products
    {
        const gf_pre px = gf_prepare(fs->x);
        const int v = lane & 3, t = v >> 1, i = v & 1, base = lane & ~3;
        gf acc = gf_zero();
#pragma unroll 1
        for (size_t g0 = 0; g0 < q0; g0 += kTailThreads / 4) {
            const size_t j = g0 + (size_t)(tid >> 2);
            const bool valid = j < q0;
            const size_t e = j + (size_t)i * q0;
            gf V = gf_zero();
            if (valid) {
                if (k0 > 0) {
                    const gf *P = t ? Ga : Fa;
                    V = fold_top(P[e], P[e + 2 * q0], px);
                    (t ? Gb : Fb)[e] = V;
                } else {
                    V = (t ? Gb : Fb)[e];
                }
            }
            const gf M = gf_mul_reg(V, valid ? Ts[j] : gf_zero());
            const gf X = gf_sel(t == 0, M, V);
            const gf X0 = lane_gf(X, base), X1 = lane_gf(X, base + 1), X2 = lane_gf(X, base + 2),
                     X3 = lane_gf(X, base + 3);
            const gf o1 = gf_sel(v == 0, X0, gf_sel(v == 1, X1, gf_add(X0, X1)));
            const gf o2 = gf_sel(v == 0, X2, gf_sel(v == 1, X3, gf_add(X2, X3)));
            const gf p = gf_mul_reg(o1, o2);
            if (valid && v < 3) acc = gf_add(acc, p);
        }
#pragma unroll
        for (int off = 4; off < 64; off <<= 1) acc = gf_add(acc, gf_shfl_xor(acc, off));
        if (lane < 3) redA[lane][wave] = acc;
    }
    // wave 0's running state, kept in LDS between rounds: eq scalars s, K, L
    // of the round being run, P1 and P2 for moving them past it, the last
    // challenge and digest, the claim
    __shared__ gf ws[7];  // s, K, L, P1, P2, x, claim
    __shared__ uint64_t wdig[4];
    if (wave == 0) {
        gf s0, K0, L0;
        prod_scalars(fs, (const gf *)((const uint8_t *)fs + prod_cf_off()) + n, k0, s0, K0, L0);
        if (lane == 0) {
            ws[0] = s0;
            ws[1] = K0;
            ws[2] = L0;
            ws[3] = gf_zero();
            ws[4] = gf_zero();
            ws[5] = fs->x;
            ws[6] = gf_zero();
            const uint64_t *in64 = reinterpret_cast<const uint64_t *>((const uint8_t *)fs + prod_input_off(n));
            for (int w = 0; w < 4; w++) wdig[w] = in64[w];
        }
    }
    __syncthreads();
#pragma unroll 1
    for (int k = k0; k < (int)n; k++) {
        const size_t q = q0 >> (k - k0);
        const int iv = (int)n - 1 - k;
        const bool odd = (k - k0) & 1;
        gf *Fc = odd ? Fa : Fb, *Gc = odd ? Ga : Gb;
        const gf *Fp = odd ? Fb : Fa, *Gp = odd ? Gb : Ga;
        if (wave == 0) {
            gf s = ws[0], K = ws[1], L = ws[2], P1 = ws[3], P2 = ws[4];
            const gf xprev = ws[5];
            gf A0, A1, A2;
            if (k == k0) {
                gf a = gf_zero();
                if (lane < 3)
                    for (int w = 0; w < kTailWaves; w++) a = gf_add(a, redA[lane][w]);
                A0 = lane_gf(a, 0);
                A1 = lane_gf(a, 1);
                A2 = lane_gf(a, 2);
                if (iv > 0) {  // P1, P2 for leaving round k0 (lanes 0, 1)
                    const gf pp = gf_mul_wide(s, (lane & 1) ? cst[0][iv - 1] : cst[2][iv - 1]);
                    P1 = lane_gf(pp, 0);
                    P2 = lane_gf(pp, 1);
                }
            } else {
                // round k's coefficient sums (9 on lanes 0..8), then
                // alpha_c, gamma_c and the middle product on lane c < 3
                gf P = gf_zero();
                if (lane < 9)
                    for (int w = 1; w < kTailWaves; w++) P = gf_add(P, red[k & 1][lane][w]);
                const int c = lane < 3 ? lane : 0;
                const gf al = lane_gf(P, 3 * c), ga = lane_gf(P, 3 * c + 1), mid = lane_gf(P, 3 * c + 2);
                const gf be = gf_add(mid, gf_add(al, ga));
                const gf t = gf_add(cst[1][iv + 1], xprev);  // eq(r_i, x_{k-1}) of the variable just fixed
                // step 1: lanes 0..2 x gamma_c, lanes 3..5 {s, P1, P2} t
                const gf u1 = gf_sel(lane < 3, xprev, gf_sel(lane == 3, s, gf_sel(lane == 4, P1, P2)));
                const gf r1 = gf_mul_wide(u1, gf_sel(lane < 3, ga, t));
                const gf tc = gf_add(be, r1);
                s = lane_gf(r1, 3);
                K = lane_gf(r1, 4);
                L = lane_gf(r1, 5);
                // step 2: lanes 0..2 x (beta_c + x gamma_c), lanes 3, 4 s {ecf, cf}[iv - 1]
                const gf u2 = gf_sel(lane < 3, xprev, s);
                const gf v2 = gf_sel(lane < 3, tc, iv > 0 ? gf_sel(lane == 3, cst[2][iv - 1], cst[0][iv - 1]) : gf_zero());
                const gf r2 = gf_mul_wide(u2, v2);
                const gf Ac = gf_add(al, r2);
                A0 = lane_gf(Ac, 0);
                A1 = lane_gf(Ac, 1);
                A2 = lane_gf(Ac, 2);
                if (iv > 0) {
                    P1 = lane_gf(r2, 3);
                    P2 = lane_gf(r2, 4);
                }
            }
            Digest prev;
#pragma unroll
            for (int w = 0; w < 4; w++) prev.w[w] = wdig[w];
            const ProdRound R = prod_transcript_call(fs, k, A0, A1, A2, K, L, k > 0, prev);
            if (lane == 0) {
                ws[0] = s;
                ws[1] = K;
                ws[2] = L;
                ws[3] = P1;
                ws[4] = P2;
                ws[5] = R.x;
                if (k == 0) ws[6] = R.claim;
                for (int w = 0; w < 4; w++) wdig[w] = R.d[w];
                sx[k & 1] = R.x;
                outs[k][0] = R.c0;
                outs[k][1] = R.c2;
                outs[k][2] = R.c3;
                outs[k][3] = R.x;
            }
        } else if (wave & 3) {
            // worker lanes: wl over the waves off wave 0's SIMD
            const int wl = 64 * (wave - 1 - (wave >> 2)) + lane;
            constexpr int kWorkLanes = 64 * (3 * kTailWaves / 4);
            const gf xw = k > k0 ? sx[(k - 1) & 1] : gf_zero();
            if (q >= 2) {
                const size_t qh = q / 2;
                const int u = lane & 7, t = u >> 2, i = u & 3, base = lane & ~7;
                gf acc = gf_zero(), acc8 = gf_zero();
#pragma unroll 1
                for (size_t g0 = 0; g0 < qh; g0 += kWorkLanes / 8) {
                    const size_t j = g0 + (size_t)(wl >> 3);
                    const bool valid = j < qh;
                    const size_t e = j + (size_t)i * qh;
                    gf V = gf_zero();
                    if (valid) {
                        if (k > k0) {
                            const gf *P = t ? Gp : Fp;
                            V = gf_add(P[e], gf_mul_reg(gf_add(P[e + 2 * q], P[e]), xw));
                            (t ? Gc : Fc)[e] = V;
                        } else {
                            V = (t ? Gc : Fc)[e];
                        }
                    }
                    const gf Vx = lane_gf(V, lane ^ 2);
                    const gf X1 = gf_sel(i >= 2, gf_add(V, Vx), V);
                    const gf M = gf_mul_reg(X1, valid ? Ts[j] : gf_zero());
                    const gf X = gf_sel(t == 0, M, X1);
                    // lane u: product u of a0c0, b0d0, (a0+b0)(c0+d0), a1c1, b1d1,
                    // (a1+b1)(c1+d1), (a0+a1)(c0+c1), (b0+b1)(d0+d1) (X of group
                    // lanes 0..7 = a0, a1, b0, b1, c0, c1, d0, d1); every lane also
                    // forms the ninth, (a0+a1+b0+b1)(c0+c1+d0+d1)
                    const int s1 = (u == 0 || u == 2 || u == 6) ? 0 : ((u == 1 || u == 7) ? 2 : ((u == 3 || u == 5) ? 1 : 3));
                    const int s2 = u == 2 ? 2 : (u == 5 ? 3 : (u == 6 ? 1 : (u == 7 ? 3 : -1)));
                    const gf xa = lane_gf(X, base + s1), xb = lane_gf(X, base + (s2 < 0 ? s1 : s2));
                    const gf ya = lane_gf(X, base + 4 + s1), yb = lane_gf(X, base + 4 + (s2 < 0 ? s1 : s2));
                    const gf o1 = gf_sel(s2 < 0, xa, gf_add(xa, xb)), o2 = gf_sel(s2 < 0, ya, gf_add(ya, yb));
                    const gf p = gf_mul_reg(o1, o2);
                    __builtin_amdgcn_sched_barrier(0);  // one product's temporaries live at a time
                    // sums of group lanes 0..3 and 4..7
                    gf Y = gf_add(X, gf_shfl_xor(X, 1));
                    Y = gf_add(Y, gf_shfl_xor(Y, 2));
                    const gf p8 = gf_mul_reg(lane_gf(Y, base), lane_gf(Y, base + 4));
                    if (valid) {
                        acc = gf_add(acc, p);
                        if (u == 0) acc8 = gf_add(acc8, p8);
                    }
                }
#pragma unroll
                for (int off = 8; off < 64; off <<= 1) {
                    acc = gf_add(acc, gf_shfl_xor(acc, off));
                    acc8 = gf_add(acc8, gf_shfl_xor(acc8, off));
                }
                gf(*rd)[kTailWaves] = red[(k + 1) & 1];
                if (lane < 8) rd[lane][wave] = acc;
                if (lane == 0) rd[8][wave] = acc8;
            } else if (k > k0 && wl < 4) {
                // the last round: its two entries of each table
                const int t = wl >> 1, e = wl & 1;
                const gf *P = t ? Gp : Fp;
                (t ? Gc : Fc)[e] = gf_add(P[e], gf_mul_reg(gf_add(P[e + 2], P[e]), xw));
            }
        } else if (q >= 2) {
            // wave 4 contributes zero coefficient sums
            gf(*rd)[kTailWaves] = red[(k + 1) & 1];
            if (lane < 9) rd[lane][wave] = gf_zero();
        }
        __syncthreads();
    }
    // the last round's two entries of each table, for the host's final fold,
    // and the rounds' outputs and the challenger state into the region
    const bool odd = ((int)n - 1 - k0) & 1;
    if (tid < 2) {
        gf_st(fout + tid, (odd ? Fa : Fb)[tid]);
        gf_st(gout + tid, (odd ? Ga : Gb)[tid]);
    }
    uint8_t *base = (uint8_t *)fs;
    gf *polys = (gf *)(base + prod_polys_off(n)), *chal = (gf *)(base + prod_chal_off(n));
    for (int k = k0 + tid; k < (int)n; k += kTailThreads) {
        polys[3 * k] = outs[k][0];
        polys[3 * k + 1] = outs[k][1];
        polys[3 * k + 2] = outs[k][2];
        chal[k] = outs[k][3];
    }
    if (tid == 0) {
        if (k0 == 0) fs->claim = ws[6];
        ProdRound R;
        R.x = ws[5];
        for (int w = 0; w < 4; w++) R.d[w] = wdig[w];
        prod_store_round(fs, R);
    }
    if (host) region_to_host(base, host, host_bytes);
}

// ---------------------------------------------------------------------------
// Quadratic sumcheck  sum_b sum_k s_k P_{a_k}(b) P_{b_k}(b)
// (sumcheck/src/function/quadratic.rs:105-127): the m tables are stored
// table-major with a fixed stride and fixed in place; pair blockIdx.y sums
// A0 = sum f0 g0, A1 = sum f1 g1, A2 = sum (f0+f1)(g0+g1) over the current
// halves (A2 is the X^2 coefficient: (f1-f0)(g1-g0) in characteristic 2);
// the host applies the scalars.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(CB) void k_quad_sums(const gf *__restrict__ P, size_t stride,
                                                  const uint32_t *__restrict__ pairs, size_t h,
                                                  gf *__restrict__ partial) {
    const size_t a = pairs[2 * blockIdx.y], b = pairs[2 * blockIdx.y + 1];
    const gf *f = P + a * stride, *g = P + b * stride;
    gf a0 = gf_zero(), a1 = gf_zero(), a2 = gf_zero();
    for (size_t j = blockIdx.x * (size_t)blockDim.x + threadIdx.x; j < h; j += (size_t)gridDim.x * blockDim.x) {
        const gf f0 = gf_ld(f + j), f1 = gf_ld(f + j + h), g0 = gf_ld(g + j), g1 = gf_ld(g + j + h);
        a0 = gf_add(a0, gf_mul_reg(f0, g0));
        a1 = gf_add(a1, gf_mul_reg(f1, g1));
        a2 = gf_add(a2, gf_mul_reg(gf_add(f0, f1), gf_add(g0, g1)));
    }
    block_xor3(a0, a1, a2, partial + 3 * (size_t)gridDim.x * blockIdx.y);
}

// One round's transcript step of a Quadratic sumcheck, run by one whole wave
// (prove_sumcheck_round, sumcheck/src/lib.rs:81-103): observe [claim,] c0, c2
// (the compressed polynomial, quadratic.rs:62-64) and sample x.  In round 0
// the claim, the sum s (A0 + A1) over the pairs, is observed first
// (sumcheck/src/lib.rs:67).  With have_prev the challenger's input is the
// previous round's 32-byte digest prev; otherwise the input bytes come from
// the region (the call's first round, after whatever the caller observed;
// lane 0 appends to them there).  Every lane gets the results.  The
// subclaim (c1 = subclaim + c2 in characteristic 2, quadratic.rs:70-80) is
// carried on the host from the polynomials and challenges.
struct QuadRound {
    gf x;
    uint64_t d[4];
};
__device__ QuadRound quad_observe(QuadFs *st, int k, gf c0, gf c2, gf both, bool have_prev, Digest prev) {
    const int lane = threadIdx.x & 63;
    uint8_t *in = (uint8_t *)st + quad_input_off(st->npairs, st->n);
    QuadRound R;
    if (have_prev) {
        const uint64_t m[8] = {prev.w[0], prev.w[1], prev.w[2], prev.w[3], c0.lo, c0.hi, c2.lo, c2.hi};
        kk::keccak256_words_pair<8>(m, R.d, lane & 1);
    } else {
        if (lane == 0) {
            uint32_t len = st->in_len;
            const auto put = [&](uint64_t v) {
#pragma unroll 1
                for (int t = 0; t < 8; t++) in[len + t] = (uint8_t)(v >> (8 * t));
                len += 8;
            };
            if (k == 0) {
                put(both.lo);
                put(both.hi);
            }
            put(c0.lo);
            put(c0.hi);
            put(c2.lo);
            put(c2.hi);
            kk::keccak256_bytes(in, len, R.d);
        }
#pragma unroll
        for (int q = 0; q < 4; q++) R.d[q] = (uint64_t)__shfl((long long)R.d[q], 0);
    }
    // the challenge is the last 16 digest bytes, last byte first
    R.x = gf{__builtin_bswap64(R.d[3]), __builtin_bswap64(R.d[2])};
    return R;
}

// Round k's outputs into the region (one lane): polynomial, challenge, claim
// (round 0) and the challenger state after the sample (input = output = the
// digest, 16 output bytes left unread).
__device__ __forceinline__ void quad_store(QuadFs *st, int k, gf c0, gf c2, gf both, const QuadRound &R) {
    uint8_t *base = (uint8_t *)st;
    const uint32_t n = st->n, np = st->npairs;
    gf *polys = (gf *)(base + quad_polys_off(np)), *chal = (gf *)(base + quad_chal_off(np, n));
    uint64_t *in64 = reinterpret_cast<uint64_t *>(base + quad_input_off(np, n));
    uint64_t *out64 = reinterpret_cast<uint64_t *>(st->out);
#pragma unroll
    for (int q = 0; q < 4; q++) {
        in64[q] = R.d[q];
        out64[q] = R.d[q];
    }
    st->in_len = 32;
    st->out_len = 16;
    if (k == 0) st->claim = both;
    st->x = R.x;
    polys[2 * k] = c0;
    polys[2 * k + 1] = c2;
    chal[k] = R.x;
}

// Adds the pair scalars' products c0 += s A0, c2 += s A2, both += s (A0 + A1)
// (quadratic.rs:100-122) for the cnt <= 64 pairs base.. whose sums
// P0, P1, P2 sit on lanes 0..cnt-1, one product per lane: three lanes per
// pair, 21 pairs per pass.  Each lane adds to one of the three; the caller
// reduces them over the wave.
__device__ void quad_scale(const gf *sc, uint32_t base, uint32_t cnt, gf P0, gf P1, gf P2, gf &c0, gf &c2, gf &both) {
    const int lane = threadIdx.x & 63, comp = lane % 3;
    for (uint32_t i = 0; i < cnt; i += 21) {
        const uint32_t src = i + (uint32_t)lane / 3;
        const int sl = (int)(src & 63);
        const gf B0 = lane_gf(P0, sl), B1 = lane_gf(P1, sl), B2 = lane_gf(P2, sl);
        if (lane < 63 && src < cnt) {
            const gf v = gf_sel(comp == 0, B0, gf_sel(comp == 1, B2, gf_add(B0, B1)));
            const gf p = gf_mul_reg(v, gf_ld(sc + base + src));
            if (comp == 0) c0 = gf_add(c0, p);
            else if (comp == 1) c2 = gf_add(c2, p);
            else both = gf_add(both, p);
        }
    }
}

// Pair sums with the round's transcript step in the block that finishes
// last.  Pair blockIdx.y sums A0 = sum f0 g0, A1 = sum f1 g1 and
// A2 = sum (f0+f1)(g0+g1) over its share of the halves; the block publishes
// its share of the round's compressed polynomial, s A0, s A2 and s (A0 + A1)
// with the pair's scalar s (quadratic.rs:100-122), XORed with device-scope
// atomics into slot (block mod kQuadSlots); the last block (two-level count,
// grid_last) adds the slots, clears them and the counters for the next round
// and runs the transcript step.
__global__ __launch_bounds__(CB) void k_quad_step(const gf *__restrict__ P, size_t stride,
                                                  const uint32_t *__restrict__ pairs, size_t h, int k, QuadFs *fs,
                                                  gf *slots, unsigned *counter) {
    const size_t a = pairs[2 * blockIdx.y], b = pairs[2 * blockIdx.y + 1];
    const gf *f = P + a * stride, *g = P + b * stride;
    const int lane = threadIdx.x & 63, wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    gf a0 = gf_zero(), a1 = gf_zero(), a2 = gf_zero();
    for (size_t j = blockIdx.x * (size_t)blockDim.x + threadIdx.x; j < h; j += (size_t)gridDim.x * blockDim.x) {
        const gf f0 = gf_ld(f + j), f1 = gf_ld(f + j + h), g0 = gf_ld(g + j), g1 = gf_ld(g + j + h);
        a0 = gf_add(a0, gf_mul_reg(f0, g0));
        a1 = gf_add(a1, gf_mul_reg(f1, g1));
        a2 = gf_add(a2, gf_mul_reg(gf_add(f0, f1), gf_add(g0, g1)));
    }
    __shared__ gf red[3][CB / 64];
    __shared__ unsigned last;
    a0 = wave_xor(a0);
    a1 = wave_xor(a1);
    a2 = wave_xor(a2);
    if (lane == 0) {
        red[0][wave] = a0;
        red[1][wave] = a1;
        red[2][wave] = a2;
    }
    __syncthreads();
    const unsigned nblocks = gridDim.x * gridDim.y, bid = blockIdx.y * gridDim.x + blockIdx.x;
    const uint32_t np = fs->npairs;
    if (wave == 0) {
        gf A[3];
#pragma unroll
        for (int c = 0; c < 3; c++) {
            A[c] = gf_zero();
            for (int w = 0; w < CB / 64; w++) A[c] = gf_add(A[c], red[c][w]);
        }
        // lanes 0, 1, 2: s A0, s A2, s (A0 + A1)
        const gf sp = gf_ld((const gf *)((const uint8_t *)fs + quad_scalars_off()) + blockIdx.y);
        const gf v = gf_mul_wide(sp, gf_sel(lane == 0, A[0], gf_sel(lane == 1, A[2], gf_add(A[0], A[1]))));
        unsigned done = 1u;
        if (lane < 3) {
            unsigned long long *slot = (unsigned long long *)(slots + 3 * (bid % kQuadSlots) + lane);
            const unsigned long long o0 =
                __hip_atomic_fetch_xor(slot, (unsigned long long)v.lo, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            const unsigned long long o1 =
                __hip_atomic_fetch_xor(slot + 1, (unsigned long long)v.hi, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            done = (unsigned)(o0 ^ o1) | 1u;
        }
        // 1, after lanes 0..2's atomics returned
        const unsigned inc = done & (unsigned)__shfl((int)done, 1) & (unsigned)__shfl((int)done, 2) & 1u;
        if (lane == 0) last = grid_last(counter, nblocks, bid, inc);
    }
    __syncthreads();
    if (!last || wave != 0) return;
    // lane l adds slots l, l + 64, ... (those the grid used), all loads
    // issued before the first use, clears them, then a wave reduction
    const unsigned used = nblocks < (unsigned)kQuadSlots ? nblocks : (unsigned)kQuadSlots;
    gf c0 = gf_zero(), c2 = gf_zero(), both = gf_zero();
#pragma unroll
    for (int t = 0; t < kQuadSlots / 64; t++) {
        const unsigned sl = (unsigned)lane + 64 * t;
        if (sl < used) {
            gf *q = slots + 3 * sl;
            c0 = gf_add(c0, gf_ld_published(q));
            c2 = gf_add(c2, gf_ld_published(q + 1));
            both = gf_add(both, gf_ld_published(q + 2));
            gf_st(q, gf_zero());
            gf_st(q + 1, gf_zero());
            gf_st(q + 2, gf_zero());
        }
    }
    c0 = wave_xor(c0);
    c2 = wave_xor(c2);
    both = wave_xor(both);
    // every block has counted: the counters start the next round at zero
    for (int w = lane; w < (int)kCountWords; w += 64) counter[w] = 0u;
    Digest prev;
    const uint64_t *in64 = reinterpret_cast<const uint64_t *>((const uint8_t *)fs + quad_input_off(np, fs->n));
#pragma unroll
    for (int w = 0; w < 4; w++) prev.w[w] = k > 0 ? in64[w] : 0;
    const QuadRound R = quad_observe(fs, k, c0, c2, both, k > 0, prev);
    if (lane == 0) quad_store(fs, k, c0, c2, both, R);
}

// fix_last_var of table blockIdx.y with the challenge at *xp: 2 half entries
// at src (tables src_stride apart) to half entries at dst (dst_stride apart;
// in place when dst == src)
template <bool TAB>
__global__ __launch_bounds__(CB) void k_fix_multi_dev(const gf *src, size_t src_stride, gf *dst, size_t dst_stride,
                                                      size_t half, const gf *__restrict__ xp) {
    __shared__ GfTab T;
    __shared__ gf u[128];
    const gf *f = src + blockIdx.y * src_stride;
    gf *o = dst + blockIdx.y * dst_stride;
    gf_pre px;
    if constexpr (TAB) gf_tab_build(T, u, gf_ld(xp));
    else px = gf_prepare(gf_ld(xp));
    for (size_t j = blockIdx.x * (size_t)blockDim.x + threadIdx.x; j < half;
         j += (size_t)gridDim.x * blockDim.x) {
        const gf lo = gf_ld(f + j), hi = gf_ld(f + j + half);
        const gf d = gf_add(hi, lo);
        if constexpr (TAB) gf_st(o + j, gf_add(lo, gf_tab_mul(d, T)));
        else gf_st(o + j, gf_add(lo, gf_mul_pre(d, px)));
    }
}

// Quadratic sumcheck rounds k >= 1 with one round of lookahead (see
// launch::quad_look).  Rows 0..npairs-1 are the pairs.  Every thread first
// folds its share of the m tables' entries with x_{k-1} from src into round
// k's tables at dst.  Then eight lanes take pair p (blockIdx.y) at position
// j < h/2 of round k+1: lane u folds entry j + (u & 3) h/2 of the pair's
// first (u < 4) or second table again, in registers, giving a0, a1, b0, b1
// and c0, c1, d0, d1 (a_i = V_i, b_i = V_{i+2} + V_i, likewise c, d) as in
// k_prod_look without eq weights, and the products p0 = a0 c0, p1 = b0 d0,
// p2 = (a0 + b0)(c0 + d0) (round k+1's A0 as a quadratic in its challenge),
// p3 = a1 c1, p4 = b1 d1, and p5..p7, the same three for A2 with a0 + a1 and
// so on.  p0 + p3 and p1 + p4 are round k's A0 and A2 of the pair at
// positions j and j + h/2.  Each block XORs s_p times its eight sums into
// slot (block mod kLookSlots) of the set for round k+1.  FIRST: the last
// block (two-level count) reads that set and runs round k's transcript step
// with c0 = sum (p0 + p3), c2 = sum (p1 + p4); otherwise row npairs' first
// block runs it from the set the previous launch left (c0 and c2 as
// quadratics in x_{k-1}) and clears that set.
constexpr size_t kQuadLookMax = (size_t)1 << 12;
template <bool FIRST>
__global__ __launch_bounds__(CB) void k_quad_look(const gf *src, size_t sstride, gf *dst, size_t dstride,
                                                  const uint32_t *__restrict__ pairs, uint32_t m, size_t h, int k,
                                                  QuadFs *fs) {
    __shared__ gf red[8][CB / 64];
    __shared__ unsigned last;
    const int tid = threadIdx.x, wave = __builtin_amdgcn_readfirstlane(tid >> 6), lane = tid & 63;
    uint8_t *base = (uint8_t *)fs;
    const uint32_t n = fs->n, np = fs->npairs, in_cap = fs->in_cap;
    gf *chal = (gf *)(base + quad_chal_off(np, n));
    gf *look = (gf *)(base + quad_look_off(np, n, in_cap, m));
    gf *set_in = look + (size_t)8 * kLookSlots * (k & 1), *set_out = look + (size_t)8 * kLookSlots * ((k + 1) & 1);
    const gf xp = gf_ld(chal + k - 1);  // x_{k-1}
    Digest prev;
    const uint64_t *in64 = reinterpret_cast<const uint64_t *>(base + quad_input_off(np, n));
    if (!FIRST && blockIdx.y == np) {
        if (blockIdx.x != 0 || wave != 0) return;
        const gf P = look_gather<8>(set_in, true);
        // lane 0: c0 from sums 0, 1, 2; lane 1: c2 from sums 5, 6, 7
        const int b = lane == 1 ? 5 : 0;
        const gf al = lane_gf(P, b), ga = lane_gf(P, b + 1), mid = lane_gf(P, b + 2);
        const gf be = gf_add(mid, gf_add(al, ga));
        const gf r2 = gf_mul_wide(xp, gf_add(be, gf_mul_wide(xp, ga)));
        const gf cv = gf_add(al, r2);
        const gf c0 = lane_gf(cv, 0), c2 = lane_gf(cv, 1);
#pragma unroll
        for (int w = 0; w < 4; w++) prev.w[w] = in64[w];
        const QuadRound R = quad_observe(fs, k, c0, c2, gf_zero(), true, prev);
        if (lane == 0) quad_store(fs, k, c0, c2, gf_zero(), R);
        return;
    }
    // every entry of round k's tables, folded once
    const unsigned nb = gridDim.x * np, bid = blockIdx.y * gridDim.x + blockIdx.x;
    const size_t w2 = 2 * h, total = (size_t)nb * CB;
    for (size_t f = (size_t)bid * CB + tid; f < (size_t)m * w2; f += total) {
        const size_t t = f / w2, e = f % w2;
        const gf *s0 = src + t * sstride;
        const gf lo = gf_ld(s0 + e);
        gf_st(dst + t * dstride + e, gf_add(lo, gf_mul_reg(gf_add(gf_ld(s0 + e + w2), lo), xp)));
    }
    // pair p, position j of round k+1
    const uint32_t p = blockIdx.y;
    const size_t qh = h / 2, j = (size_t)blockIdx.x * kLookGroups + (size_t)(tid >> 3);
    const int u = lane & 7, tb = u >> 2, i = u & 3, gb = lane & ~7;
    const gf *s0 = src + (size_t)pairs[2 * p + tb] * sstride;
    const size_t e = j + (size_t)i * qh;
    const gf lo = gf_ld(s0 + e);
    const gf V = gf_add(lo, gf_mul_reg(gf_add(gf_ld(s0 + e + w2), lo), xp));
    const gf X = gf_sel(i >= 2, gf_add(V, lane_gf(V, lane ^ 2)), V);
    // lane u: p_u with operands X[s1] (+ X[s2]) and the same of the second
    // table; u = 7 takes the sums of all four
    const int s1 = (u == 0 || u == 2 || u == 5) ? 0 : ((u == 1 || u == 6) ? 2 : (u == 3 ? 1 : 3));
    const int s2 = u == 2 ? 2 : (u == 5 ? 1 : (u == 6 ? 3 : -1));
    const gf xa = lane_gf(X, gb + s1), xb = lane_gf(X, gb + (s2 < 0 ? s1 : s2));
    const gf ya = lane_gf(X, gb + 4 + s1), yb = lane_gf(X, gb + 4 + (s2 < 0 ? s1 : s2));
    gf Y = gf_add(X, gf_shfl_xor(X, 1));
    Y = gf_add(Y, gf_shfl_xor(Y, 2));
    const gf ys = lane_gf(Y, gb), zs = lane_gf(Y, gb + 4);
    const gf o1 = gf_sel(u == 7, ys, gf_sel(s2 < 0, xa, gf_add(xa, xb)));
    const gf o2 = gf_sel(u == 7, zs, gf_sel(s2 < 0, ya, gf_add(ya, yb)));
    gf a = gf_mul_reg(o1, o2);
#pragma unroll
    for (int off = 8; off < 64; off <<= 1) a = gf_add(a, gf_shfl_xor(a, off));
    if (lane < 8) red[lane][wave] = a;
    __syncthreads();
    if (wave == 0) {
        gf v = gf_zero();
        if (lane < 8)
            for (int w = 0; w < CB / 64; w++) v = gf_add(v, red[lane][w]);
        v = gf_mul_reg(v, gf_ld((const gf *)(base + quad_scalars_off()) + p));
        unsigned done;
        look_publish<8>(set_out, bid, v, FIRST, done);
        if (FIRST) {
            unsigned inc = done;
#pragma unroll
            for (int l = 1; l < 8; l++) inc &= (unsigned)__shfl((int)done, l);
            unsigned *counter = (unsigned *)(base + quad_counters_off(np, n, in_cap, m));
            if (lane == 0) last = grid_last(counter, nb, bid, inc & 1u);
        }
    }
    if (!FIRST) return;
    __syncthreads();
    if (!last || wave != 0) return;
    // round k's transcript step from its sums in round k+1's set
    const gf P = look_gather<8>(set_out, false);
    const gf c0 = gf_add(lane_gf(P, 0), lane_gf(P, 3)), c2 = gf_add(lane_gf(P, 1), lane_gf(P, 4));
    unsigned *counter = (unsigned *)(base + quad_counters_off(np, n, in_cap, m));
    for (int w = lane; w < (int)kCountWords; w += 64) counter[w] = 0u;
#pragma unroll
    for (int w = 0; w < 4; w++) prev.w[w] = in64[w];
    const QuadRound R = quad_observe(fs, k, c0, c2, gf_zero(), true, prev);
    if (lane == 0) quad_store(fs, k, c0, c2, gf_zero(), R);
}

// The Quadratic sumcheck's rounds from k0 on in one block, the m tables in
// LDS (2 h0 entries each, read from src once).  Per round: the pair sums,
// one wave per pair (lanes over the halves), into LDS; the transcript step
// on the first wave; the fix of every table in place in LDS.  The digest,
// challenge and subclaim stay in LDS between rounds; each table's final
// entry is written to fin (tables fin_stride apart) and the challenger state
// to the region at the end.
constexpr int kQuadTailThreads = 512;
constexpr size_t kQuadTailElems = 3584;  // table entries in LDS (56 KiB)
constexpr uint32_t kQuadTailPairs = 128;
__global__ __launch_bounds__(kQuadTailThreads) void k_quad_tail(const gf *src, size_t src_stride, gf *fin,
                                                                size_t fin_stride, const uint32_t *__restrict__ pairs,
                                                                uint32_t m, size_t h0, int k0, QuadFs *fs,
                                                                uint8_t *host, uint32_t host_bytes) {
    __shared__ gf F[kQuadTailElems];
    __shared__ gf S3[3][kQuadTailPairs];
    __shared__ gf xs;
    __shared__ uint64_t dig[4];
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const uint32_t n = fs->n, np = fs->npairs;
    const size_t w = 2 * h0;
    for (size_t e = tid; e < (size_t)m * w; e += kQuadTailThreads) F[e] = gf_ld(src + (e / w) * src_stride + e % w);
    if (tid == 0) {
        const uint64_t *in64 = reinterpret_cast<const uint64_t *>((const uint8_t *)fs + quad_input_off(np, n));
        for (int q = 0; q < 4; q++) dig[q] = in64[q];
    }
    __syncthreads();
    const gf *sc = (const gf *)((const uint8_t *)fs + quad_scalars_off());
    size_t h = h0;
    for (int k = k0; k < (int)n; k++, h >>= 1) {
        for (uint32_t t = wave; t < np; t += kQuadTailThreads / 64) {
            const size_t fa = pairs[2 * t] * w, gb = pairs[2 * t + 1] * w;
            gf a0 = gf_zero(), a1 = gf_zero(), a2 = gf_zero();
            for (size_t j = lane; j < h; j += 64) {
                const gf f0 = F[fa + j], f1 = F[fa + j + h], g0 = F[gb + j], g1 = F[gb + j + h];
                a0 = gf_add(a0, gf_mul_reg(f0, g0));
                a1 = gf_add(a1, gf_mul_reg(f1, g1));
                a2 = gf_add(a2, gf_mul_reg(gf_add(f0, f1), gf_add(g0, g1)));
            }
            a0 = wave_xor(a0);
            a1 = wave_xor(a1);
            a2 = wave_xor(a2);
            if (lane == 0) {
                S3[0][t] = a0;
                S3[1][t] = a1;
                S3[2][t] = a2;
            }
        }
        __syncthreads();
        if (wave == 0) {
            gf c0 = gf_zero(), c2 = gf_zero(), both = gf_zero();
            for (uint32_t c = 0; c < np; c += 64) {
                const uint32_t t = c + lane;
                const gf P0 = t < np ? S3[0][t] : gf_zero(), P1 = t < np ? S3[1][t] : gf_zero(),
                         P2 = t < np ? S3[2][t] : gf_zero();
                quad_scale(sc, c, np - c < 64 ? np - c : 64, P0, P1, P2, c0, c2, both);
            }
            c0 = wave_xor(c0);
            c2 = wave_xor(c2);
            both = wave_xor(both);
            const Digest prev{{dig[0], dig[1], dig[2], dig[3]}};
            const QuadRound R = quad_observe(fs, k, c0, c2, both, k > 0, prev);
            if (lane == 0) {
                uint8_t *base = (uint8_t *)fs;
                gf *polys = (gf *)(base + quad_polys_off(np)), *chal = (gf *)(base + quad_chal_off(np, n));
                polys[2 * k] = c0;
                polys[2 * k + 1] = c2;
                chal[k] = R.x;
                if (k == 0) fs->claim = both;
                dig[0] = R.d[0];
                dig[1] = R.d[1];
                dig[2] = R.d[2];
                dig[3] = R.d[3];
                xs = R.x;
            }
        }
        __syncthreads();
        const gf_pre px = gf_prepare(xs);
        for (size_t idx = tid; idx < (size_t)m * h; idx += kQuadTailThreads) {
            const size_t t = idx / h, j = idx % h;
            const gf lo = F[t * w + j], hi = F[t * w + j + h];
            F[t * w + j] = gf_add(lo, gf_mul_pre(gf_add(hi, lo), px));
        }
        __syncthreads();
    }
    for (size_t t = tid; t < m; t += kQuadTailThreads) gf_st(fin + t * fin_stride, F[t * w]);
    if (tid == 0) {
        uint64_t *in64 = reinterpret_cast<uint64_t *>((uint8_t *)fs + quad_input_off(np, n));
        uint64_t *out64 = reinterpret_cast<uint64_t *>(fs->out);
        for (int q = 0; q < 4; q++) {
            in64[q] = dig[q];
            out64[q] = dig[q];
        }
        fs->in_len = 32;
        fs->out_len = 16;
        fs->x = xs;
    }
    if (host) region_to_host((const uint8_t *)fs, host, host_bytes);
}

__global__ __launch_bounds__(CB) void k_fix_multi(gf *__restrict__ P, size_t stride, size_t half, gf x) {
    gf *f = P + blockIdx.y * stride;
    const gf_pre px = gf_prepare(x);
    for (size_t j = blockIdx.x * (size_t)blockDim.x + threadIdx.x; j < half;
         j += (size_t)gridDim.x * blockDim.x) {
        const gf lo = gf_ld(f + j), hi = gf_ld(f + j + half);
        gf_st(f + j, gf_add(lo, gf_mul_pre(gf_add(hi, lo), px)));
    }
}

// ---------------------------------------------------------------------------
// Verifier (basefold.rs:234-325) over many proofs of one code.  A proof's
// query records are read in place from its bincode bytes (8-byte aligned
// copy in device memory; every field is a multiple of 8 bytes).  Record
// layout: u64 w | w elements | u64 L | L digests | u64 d | per round k
// (level i = d-1-k): element | u64 (b+i) | (b+i) digests.
// ---------------------------------------------------------------------------
__device__ __forceinline__ gf ld8(const uint8_t *p) {
    const uint64_t *q = reinterpret_cast<const uint64_t *>(p);
    return gf{q[0], q[1]};
}
__device__ __forceinline__ gf interpolate(gf t, gf lo, gf hi, gf r) {  // binary_reed_solomon.rs:103-109
    hi = gf_add(hi, lo);
    lo = gf_add(lo, gf_mul_reg(hi, t));
    return gf_add(lo, gf_mul_reg(gf_add(hi, lo), r));
}

// thread per (proof, query): the fold chain of the opened row through every
// round; vals[(p*q + t)*d + k] = the two pi values of round k, fold_ok = the
// final value equals final_poly (pi_0 = encode0(final_poly) is constant)
__global__ __launch_bounds__(CB) void k_verify_fold(const VerifyJob *__restrict__ jobs, uint32_t count,
                                                    uint32_t q, int b, int d, size_t w, const gf *__restrict__ ft,
                                                    gf *__restrict__ vals, uint8_t *__restrict__ fold_ok) {
    const size_t L = (size_t)(b + d);
    for (size_t it = blockIdx.x * (size_t)blockDim.x + threadIdx.x; it < (size_t)count * q;
         it += (size_t)gridDim.x * blockDim.x) {
        const VerifyJob &J = jobs[it / q];
        const size_t t = it % q;
        const uint8_t *rec = J.proof + J.open_off + t * J.rec_bytes;
        const uint64_t idx = J.idx[t];
        gf folded = gf_zero();
        for (size_t c = 0; c < w; c++) folded = gf_add(folded, gf_mul_reg(gf_ld(J.scalars + c), ld8(rec + 8 + 16 * c)));
        size_t off = 24 + 16 * w + 32 * L;
        for (int k = 0; k < d; k++) {
            const int i = d - 1 - k;
            const uint64_t ni = (uint64_t)1 << (b + i);
            const uint64_t sib = ((idx >> (b + i)) & 1) ^ 1, row = idx & (ni - 1);
            const gf sv = ld8(rec + off);
            const gf v0 = sib == 0 ? sv : folded, v1 = sib == 0 ? folded : sv;
            gf_st(vals + 2 * (it * d + k), v0);
            gf_st(vals + 2 * (it * d + k) + 1, v1);
            const gf *lvl = ft + (((size_t)1 << b) * (((size_t)1 << i) - 1));  // fold table of level i
            folded = interpolate(gf_ld(lvl + row), v0, v1, gf_ld(J.chal + i));
            off += 24 + 32 * (size_t)(b + i);
        }
        fold_ok[it] = gf_eq(folded, J.final_poly) ? 1 : 0;
    }
}

// thread per (proof, query, path): path 0 climbs the main tree from the
// opened row, path k+1 the pi tree of round k from its two values
__global__ __launch_bounds__(KB) void k_verify_paths(const VerifyJob *__restrict__ jobs, uint32_t count, uint32_t q,
                                                     int b, int d, size_t w, const gf *__restrict__ vals,
                                                     uint8_t *__restrict__ path_ok) {
    const size_t L = (size_t)(b + d), P = (size_t)d + 1;
    for (size_t it = blockIdx.x * (size_t)blockDim.x + threadIdx.x; it < (size_t)count * q * P;
         it += (size_t)gridDim.x * blockDim.x) {
        const size_t pq = it / P, j = it % P;
        const VerifyJob &J = jobs[pq / q];
        const size_t t = pq % q;
        const uint8_t *rec = J.proof + J.open_off + t * J.rec_bytes;
        const uint64_t idx = J.idx[t];
        uint64_t cur[4];
        const uint8_t *sib, *root;
        uint64_t index;
        size_t len;
        if (j == 0) {
            kk::keccak256_var(reinterpret_cast<const uint64_t *>(rec + 8), 2 * w, cur);
            sib = rec + 16 + 16 * w;
            len = L;
            index = idx;
            root = J.root;
        } else {
            const int k = (int)j - 1, i = d - 1 - k;
            size_t off = 24 + 16 * w + 32 * L;
            for (int kk2 = 0; kk2 < k; kk2++) off += 24 + 32 * (size_t)(b + d - 1 - kk2);
            const gf v0 = gf_ld(vals + 2 * (pq * d + k)), v1 = gf_ld(vals + 2 * (pq * d + k) + 1);
            uint64_t m[4] = {v0.lo, v0.hi, v1.lo, v1.hi};
            kk::keccak256_words<4>(m, cur);
            sib = rec + off + 24;
            len = (size_t)(b + i);
            index = idx & ((((uint64_t)1) << (b + i)) - 1);
            root = J.proof + J.roots_off + 32 * (size_t)k;
        }
        for (size_t u = 0; u < len; u++, index >>= 1) {
            const uint64_t *s8 = reinterpret_cast<const uint64_t *>(sib + 32 * u);
            uint64_t m[8];
            if ((index & 1) == 0) {
                m[0] = cur[0]; m[1] = cur[1]; m[2] = cur[2]; m[3] = cur[3];
                m[4] = s8[0]; m[5] = s8[1]; m[6] = s8[2]; m[7] = s8[3];
            } else {
                m[0] = s8[0]; m[1] = s8[1]; m[2] = s8[2]; m[3] = s8[3];
                m[4] = cur[0]; m[5] = cur[1]; m[6] = cur[2]; m[7] = cur[3];
            }
            kk::keccak256_words<8>(m, cur);
        }
        const uint64_t *r8 = reinterpret_cast<const uint64_t *>(root);
        path_ok[it] = (cur[0] == r8[0] && cur[1] == r8[1] && cur[2] == r8[2] && cur[3] == r8[3]) ? 1 : 0;
    }
}

// ---------------------------------------------------------------------------
// Segmented forms: one launch covers the same stage of many polynomials.
// Segment g owns blocks [blk0, blk0 + nblk) and grid-strides over its own
// `count` items; a block finds its segment by binary search on blk0.
// ---------------------------------------------------------------------------
// The block's segment, by value: its fields stay in scalar registers instead
// of being reloaded after every store of the grid-stride loop.
__device__ __forceinline__ Seg my_seg(const Seg *segs, int nseg) {
    int lo = 0, hi = nseg - 1;
    const uint32_t bx = blockIdx.x;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (segs[mid].blk0 <= bx) lo = mid;
        else hi = mid - 1;
    }
    return segs[lo];
}
#define SEG_LOOP(g, i)                                                                                  \
    for (size_t i = (size_t)(blockIdx.x - g.blk0) * blockDim.x + threadIdx.x; i < g.count;               \
         i += (size_t)g.nblk * blockDim.x)

// in = evaluations, out = codeword, aux = level-0 twiddles, arg = b
__global__ __launch_bounds__(256) void ks_ntt_first(const Seg *segs, int nseg) {
    const Seg g = my_seg(segs, nseg);
    const int b = (int)g.arg;
    const size_t cm = ((size_t)1 << b) - 1;
    const gf *data = (const gf *)g.in, *ft0 = (const gf *)g.aux;
    gf *cw = (gf *)g.out;
    SEG_LOOP(g, q) {
        const size_t u = q & cm, pr = q >> b;
        gf lo = gf_ld(data + 2 * pr), hi = gf_ld(data + 2 * pr + 1);
        lo = gf_add(lo, gf_mul_reg(hi, gf_ld(ft0 + u)));
        hi = gf_add(hi, lo);
        gf_st(cw + ((2 * pr) << b) + u, lo);
        gf_st(cw + ((2 * pr + 1) << b) + u, hi);
    }
}

// out = codeword (in place), aux = level-s twiddles, arg = s + b
__global__ __launch_bounds__(256) void ks_ntt_layer(const Seg *segs, int nseg) {
    const Seg g = my_seg(segs, nseg);
    const int sb = (int)g.arg;
    const size_t tmask = ((size_t)1 << sb) - 1;
    gf *cw = (gf *)g.out;
    const gf *fts = (const gf *)g.aux;
    SEG_LOOP(g, q) {
        const size_t lo_i = ((q >> sb) << (sb + 1)) | (q & tmask);
        const size_t hi_i = lo_i + ((size_t)1 << sb);
        gf lo = gf_ld(cw + lo_i), hi = gf_ld(cw + hi_i);
        lo = gf_add(lo, gf_mul_reg(hi, gf_ld(fts + (q & tmask))));
        hi = gf_add(hi, lo);
        gf_st(cw + lo_i, lo);
        gf_st(cw + hi_i, hi);
    }
}

// out = codeword (in place), aux = table of layer s, arg = K = s + b,
// count = groups of 2^R elements; the tables of layers s+1 and s+2 follow
// that of layer s contiguously (2^K and 2^(K+1) entries after it)
template <int R>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) void ks_ntt_multi(const Seg *segs,
                                                                                            int nseg) {
    const Seg g = my_seg(segs, nseg);
    const int K = (int)g.arg;
    const gf *__restrict__ t0 = (const gf *)g.aux;
    const gf *__restrict__ t1 = t0 + ((size_t)1 << K);
    const gf *__restrict__ t2 = t1 + ((size_t)2 << K);
    gf *__restrict__ cw = (gf *)g.out;
    const size_t mK = ((size_t)1 << K) - 1, stride = (size_t)1 << K;
    SEG_LOOP(g, q) NTT_GROUP(R, cw, t0, t1, t2, K, mK, stride, q, 1, 0);
}

__global__ __launch_bounds__(KB) void ks_leaves_elem(const Seg *segs, int nseg) {
    const Seg g = my_seg(segs, nseg);
    const gf *rows = (const gf *)g.in;
    uint64_t *out = (uint64_t *)g.out;
    SEG_LOOP(g, i) {
        gf v = gf_ld(rows + i);
        uint64_t m[2] = {v.lo, v.hi}, d[4];
        kk::keccak256_words<2, true>(m, d);
        st_digest(out + 4 * i, d);
    }
}


__global__ __launch_bounds__(KB) void ks_nodes(const Seg *segs, int nseg) {
    const Seg g = my_seg(segs, nseg);
    const uint64_t *prev = (const uint64_t *)g.in;
    uint64_t *out = (uint64_t *)g.out;
    SEG_LOOP(g, i) {
        const ulonglong2 *p = reinterpret_cast<const ulonglong2 *>(prev + 8 * i);
        ulonglong2 a = p[0], b = p[1], c = p[2], e = p[3];
        uint64_t m[8] = {a.x, a.y, b.x, b.y, c.x, c.y, e.x, e.y}, d[4];
        kk::keccak256_words<8, true>(m, d);
        st_digest(out + 4 * i, d);
    }
}

// in = layer of arg = h digests, count = h >> lv subtrees (both powers of two,
// so lv = log2(h / count)), x.lo = pair_w; see subtree()
__global__ __launch_bounds__(kSubThreads) void ks_subtree(const Seg *segs, int nseg) {
    __shared__ ulonglong2 buf[kSubLds][2];
    const Seg g = my_seg(segs, nseg);
    uint64_t *prev = (uint64_t *)g.in;
    const int lv = __builtin_ctzll(g.arg) - __builtin_ctzll(g.count);
    const int pair_w = (int)g.x.lo;
    for (size_t sub = blockIdx.x - g.blk0; sub < g.count; sub += g.nblk) subtree(prev, g.arg, lv, sub, buf, pair_w);
}

// in = halved eq table, aux = poly; partial[blockIdx.x] = block sum
// A round's pi-tree leaves (segments with an output: leaf j = Keccak-256 of
// pi[j] || pi[j + arg]) and its inner products (segments without one: in =
// halved eq table, aux = polynomial, count elements; block b of the segment
// leaves its sum in partial[x.lo + b]) in one launch, so the small inner
// products run beside the leaf hashing instead of as a launch of their own.
__global__ __launch_bounds__(KB) void ks_leaves_pair_dot(const Seg *segs, int nseg, gf *__restrict__ partial) {
    const Seg g = my_seg(segs, nseg);
    if (g.out) {
        const gf *pi = (const gf *)g.in;
        uint64_t *out = (uint64_t *)g.out;
        const size_t half = g.arg;
        SEG_LOOP(g, j) {
            gf a = gf_ld(pi + j), b = gf_ld(pi + j + half);
            uint64_t m[4] = {a.lo, a.hi, b.lo, b.hi}, d[4];
            kk::keccak256_words<4>(m, d);
            st_digest(out + 4 * j, d);
        }
        return;
    }
    __shared__ gf red[KB / 64];
    const gf *a = (const gf *)g.in, *b = (const gf *)g.aux;
    gf acc = gf_zero();
    SEG_LOOP(g, k) acc = gf_add(acc, gf_mul_reg(gf_ld(a + k), gf_ld(b + k)));
    acc = wave_xor(acc);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        gf sum = gf_zero();
        for (int w = 0; w < KB / 64; w++) sum = gf_add(sum, red[w]);
        gf_st(partial + g.x.lo + (blockIdx.x - g.blk0), sum);
    }
}


// the segment's challenge: at device address arg when arg is set (written
// there by the round's transcript kernel), else the value x
__device__ __forceinline__ gf seg_challenge(const Seg &g) { return g.arg ? gf_ld((const gf *)g.arg) : g.x; }

// in = table (2*count), out = fixed table (count), challenge per
// seg_challenge
__global__ __launch_bounds__(CB) void ks_fix(const Seg *segs, int nseg) {
    const Seg g = my_seg(segs, nseg);
    const gf *in = (const gf *)g.in;
    gf *out = (gf *)g.out;
    const size_t half = g.count;
    const gf_pre px = gf_prepare(seg_challenge(g));
    SEG_LOOP(g, j) {
        gf lo = gf_ld(in + j), hi = gf_ld(in + j + half);
        gf_st(out + j, gf_add(lo, gf_mul_pre(gf_add(hi, lo), px)));
    }
}

// fold (binary_reed_solomon.rs:88-96: in = pi, out = next pi, aux = level
// twiddles) and fix segments (as ks_fix) in one launch: a segment with a
// twiddle table (aux) folds, one without fixes.  A block serves one segment,
// so the two kinds share the device instead of running one after the other.
template <bool TAB>
__device__ __forceinline__ void seg_fold_fix(const Seg &g, const GfTab &T, const gf_pre &px) {
    const auto byx = [&](gf v) {
        if constexpr (TAB) return gf_tab_mul(v, T);
        else return gf_mul_pre(v, px);
    };
    if (g.aux) {
        const gf *w = (const gf *)g.in, *ft = (const gf *)g.aux;
        gf *out = (gf *)g.out;
        const size_t mid = g.count;
        SEG_LOOP(g, j) {
            gf lo = gf_ld(w + j), hi = gf_ld(w + j + mid);
            hi = gf_add(hi, lo);
            lo = gf_add(lo, gf_mul_reg(hi, gf_ld(ft + j)));
            gf_st(out + j, gf_add(lo, byx(gf_add(hi, lo))));
        }
    } else {
        const gf *in = (const gf *)g.in;
        gf *out = (gf *)g.out;
        const size_t half = g.count;
        SEG_LOOP(g, j) {
            gf lo = gf_ld(in + j), hi = gf_ld(in + j + half);
            gf_st(out + j, gf_add(lo, byx(gf_add(hi, lo))));
        }
    }
}
// a segment whose threads have two outputs or more multiplies by its
// challenge through a window table
__global__ __launch_bounds__(CB) void ks_fold_fix(const Seg *segs, int nseg) {
    __shared__ GfTab T;
    __shared__ gf u[128];
    const Seg g = my_seg(segs, nseg);
    const gf x = seg_challenge(g);
    if (g.count >= 2 * (size_t)g.nblk * CB) {
        gf_tab_build(T, u, x);
        seg_fold_fix<true>(g, T, gf_pre{});
    } else {
        seg_fold_fix<false>(g, T, gf_prepare(x));
    }
}


// One round of the Eval sumcheck's transcript for one opening, run by one
// wave (OpenRounds::round on the device; basefold.rs:356, eval.rs:64-110,
// eval_impr.rs:41-58, sumcheck/src/lib.rs:94-100): turn the inner product S
// into the compressed round polynomial, observe the pi-tree root and
// [rp0, rp2], sample the challenge and advance the subclaims.  Independent
// products run on different lanes at once (lane t takes operand pair t), so
// the round costs six multiplication latencies instead of twelve; lane 0
// writes the results.
__device__ void fs_round(FsState *st, int k, gf S, const uint64_t *root) {
    const int lane = threadIdx.x & 63;
    const uint32_t n = st->n;
    const int i = (int)n - 1 - k;
    uint8_t *base = (uint8_t *)st;
    const gf *r = (const gf *)(base + fs_consts_off()), *rinv = r + n, *cf = rinv + n;
    FsRound *rounds = (FsRound *)(base + fs_rounds_off(n));
    uint8_t *in = base + fs_input_off(n);
    uint64_t *in64 = reinterpret_cast<uint64_t *>(in);
    const gf sub = (i == (int)n - 1) ? st->sub : st->inner;
    const gf one_r = gf_add(gf_one(), gf_ld(r + i)), sd = st->s_delta;
    // coeff0 = cf_i S (eval_impr.rs:187-197 scaled by the correcting factor)
    // and d0 = s_delta (1 + r_i) at once
    gf p = gf_mul_wide(lane == 1 ? sd : gf_ld(cf + i), lane == 1 ? one_r : S);
    const gf coeff0 = lane_gf(p, 0), d0 = lane_gf(p, 1);
    // eval_1 = (sub + (1 + r_i) coeff0) / r_i  (eval_impr.rs:41-43)
    const gf eval1 = gf_mul_wide(gf_add(sub, gf_mul_wide(one_r, coeff0)), gf_ld(rinv + i));
    const gf coeff1 = gf_add(eval1, coeff0);
    // the four products of (coeff0 + coeff1 X)(d0 + s_delta X)
    p = gf_mul_wide((lane & 2) ? coeff1 : coeff0, (lane & 1) ? sd : d0);
    const gf rp0 = lane_gf(p, 0), rp1 = gf_add(lane_gf(p, 1), lane_gf(p, 2)), rp2 = lane_gf(p, 3);
    // observe root || rp0 || rp2 (little-endian words; observing drops the
    // unread output) and hash the whole input for the sample
    uint64_t d[4];
    if (st->in_len == 32) {
        // every round after the first: the input is the previous digest
        const uint64_t m[12] = {in64[0], in64[1], in64[2], in64[3], root[0], root[1],
                                root[2],  root[3],  rp0.lo, rp0.hi, rp2.lo, rp2.hi};
        kk::keccak256_words_pair<12>(m, d, lane & 1);
    } else {
        if (lane == 0) {
            const uint64_t w[8] = {root[0], root[1], root[2], root[3], rp0.lo, rp0.hi, rp2.lo, rp2.hi};
            const uint32_t len = st->in_len;
#pragma unroll
            for (int q = 0; q < 8; q++) {
#pragma unroll 1
                for (int t = 0; t < 8; t++) in[len + 8 * q + t] = (uint8_t)(w[q] >> (8 * t));
            }
            kk::keccak256_bytes(in, len + 64, d);
        }
#pragma unroll
        for (int q = 0; q < 4; q++) d[q] = (uint64_t)__shfl((long long)d[q], 0);
    }
    // 16 bytes are taken from the end of the digest (byte j of the challenge
    // = digest byte 31 - j); the digest becomes the input and the output,
    // whose bytes 0..15 stay unread
    const gf x{__builtin_bswap64(d[3]), __builtin_bswap64(d[2])};
    // x rp2, coeff1 x and s_delta (1 + r_i + x) at once (eval.rs:110,
    // lib.rs:100, eval_impr.rs:45-58)
    p = gf_mul_wide(lane == 0 ? x : (lane == 1 ? coeff1 : sd), lane == 0 ? rp2 : (lane == 1 ? x : gf_add(one_r, x)));
    const gf xrp2 = lane_gf(p, 0), cx = lane_gf(p, 1), sd_next = lane_gf(p, 2);
    const gf sub_next = gf_add(rp0, gf_mul_wide(x, gf_add(rp1, xrp2)));
    if (lane != 0) return;
    uint64_t *out64 = reinterpret_cast<uint64_t *>(st->out);
#pragma unroll
    for (int q = 0; q < 4; q++) {
        in64[q] = d[q];
        out64[q] = d[q];
    }
    st->in_len = 32;
    st->out_len = 16;
    st->sub = sub_next;
    st->inner = gf_add(coeff0, cx);
    st->s_delta = sd_next;
    st->x = x;
    FsRound &rec = rounds[k];
    rec.rp0 = rp0;
    rec.rp2 = rp2;
    for (int t = 0; t < 4; t++) rec.root[t] = root[t];
}

// One block (one wave) per opening: in = the round's pi-tree root, out =
// FsState, count = round k, x = (first block, blocks) of the round's inner
// product sums in `partial`
__global__ __launch_bounds__(64) void ks_round(const Seg *segs, int nseg, const gf *__restrict__ partial) {
    const Seg g = my_seg(segs, nseg);
    const uint32_t b0 = (uint32_t)g.x.lo, nb = (uint32_t)g.x.hi;
    gf acc = gf_zero();
    for (uint32_t t = threadIdx.x; t < nb; t += 64) acc = gf_add(acc, gf_ld(partial + b0 + t));
    acc = wave_xor(acc);
    fs_round((FsState *)g.out, (int)g.count, acc, (const uint64_t *)g.in);
}

// The last Merkle launch of a round's pi tree, the round's inner product and
// its Fiat-Shamir step in one launch.  Blocks [0, nsub) hash the top levels
// exactly as k_subtree; blocks [nsub, nsub + ndot) each sum a slice of
// a[k] b[k] (the halved eq table against the low half of the polynomial,
// eval_impr.rs:187-197) into partial[block - nsub].  The top of a tree is a
// chain of Keccak-f latencies on a few CUs, so the inner product runs on the
// idle ones for nothing.  Every block then counts itself done; the block
// that counts last runs the transcript step (fs_round) with its first wave
// on the root and the reduced sums.  Each block publishes its writes with a
// fence before it counts, and the last one fences before it reads them; no
// block waits for another.  *counter is zero at launch.
// prev is not __restrict__: the last block reads the root written through it
// by the subtree block (round->in points into the same tree).
__global__ __launch_bounds__(kSubThreads) void k_top_round(uint64_t *prev, size_t h, int lv,
                                                            unsigned nsub, const gf *__restrict__ a,
                                                            const gf *__restrict__ b, size_t n, gf *partial,
                                                            unsigned ndot, const Seg *round, unsigned *counter,
                                                            int pair_w) {
    __shared__ ulonglong2 buf[kSubLds][2];
    __shared__ gf red[kSubThreads / 64];
    __shared__ unsigned done[1];
    const unsigned blk = blockIdx.x;
    if (blk < nsub) {
        // the tree levels (the root among them) are plain stores: this block
        // releases them with a fence before it is counted
        subtree(prev, h, lv, blk, buf, pair_w);
        __threadfence();
        if (threadIdx.x == 0) done[0] = 1u;
        __syncthreads();
    } else {
        // the dot blocks publish their sums with device-scope atomics they
        // wait for, so they need no release fence (a write-back of the L2)
        const unsigned db = blk - nsub;
        gf acc = gf_zero();
        for (size_t k = db * (size_t)kSubThreads + threadIdx.x; k < n; k += (size_t)ndot * kSubThreads)
            acc = gf_add(acc, gf_mul_reg(gf_ld(a + k), gf_ld(b + k)));
        acc = wave_xor(acc);
        if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
        __syncthreads();
        if (threadIdx.x == 0) {
            gf sum = gf_zero();
            for (int w = 0; w < kSubThreads / 64; w++) sum = gf_add(sum, red[w]);
            unsigned long long *p = (unsigned long long *)(partial + db);
            const unsigned long long o0 =
                __hip_atomic_exchange(p, (unsigned long long)sum.lo, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            const unsigned long long o1 =
                __hip_atomic_exchange(p + 1, (unsigned long long)sum.hi, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            done[0] = (unsigned)(o0 ^ o1) | 1u;
        }
        __syncthreads();
    }
    if (!block_counted(counter, nsub + ndot, done, 1) || threadIdx.x >= 64) return;
    __threadfence();  // acquire: the subtree block's levels, the root among them
    const Seg &g = *round;
    gf acc = gf_zero();
    for (unsigned t = threadIdx.x; t < ndot; t += 64) acc = gf_add(acc, gf_ld_published(partial + t));
    acc = wave_xor(acc);
    fs_round((FsState *)g.out, (int)g.count, acc, (const uint64_t *)g.in);
}

// in = A (2^m), aux = B, out = table, arg = m
// A segment without aux is a copy of count 16-byte units from in to out (the
// starting transcript images of the jobs entering their first round).
__global__ __launch_bounds__(CB) void ks_outer(const Seg *segs, int nseg) {
    const Seg g = my_seg(segs, nseg);
    const gf *A = (const gf *)g.in, *B = (const gf *)g.aux;
    gf *out = (gf *)g.out;
    if (!B) {
        SEG_LOOP(g, i) gf_st(out + i, gf_ld(A + i));
        return;
    }
    const int m = (int)g.arg;
    const size_t mask = ((size_t)1 << m) - 1;
    SEG_LOOP(g, i) gf_st(out + i, gf_mul_reg(gf_ld(A + (i & mask)), gf_ld(B + (i >> m))));
}

// out[i] = the 16- or 32-byte record at absolute device address src[i]
__global__ void k_gather_abs(const uint64_t *__restrict__ src, uint8_t *__restrict__ out, size_t n, int rb) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const ulonglong2 *p = reinterpret_cast<const ulonglong2 *>(src[i]);
        ulonglong2 *o = reinterpret_cast<ulonglong2 *>(out + (size_t)rb * i);
        o[0] = p[0];
        if (rb == 32) o[1] = p[1];
    }
}

// ---------------------------------------------------------------------------
// proof records
// ---------------------------------------------------------------------------
__device__ __forceinline__ void st_elem8(uint8_t *o, const gf *src) {
    const gf v = gf_ld(src);
    uint64_t *q = reinterpret_cast<uint64_t *>(o);
    q[0] = v.lo;
    q[1] = v.hi;
}
__device__ __forceinline__ void st_digest8(uint8_t *o, const uint64_t *src) {
    const ulonglong2 a = reinterpret_cast<const ulonglong2 *>(src)[0], c = reinterpret_cast<const ulonglong2 *>(src)[1];
    uint64_t *q = reinterpret_cast<uint64_t *>(o);
    q[0] = a.x;
    q[1] = a.y;
    q[2] = c.x;
    q[3] = c.y;
}
// digests before layer t of a Merkle tree over `leaves` leaves
__device__ __forceinline__ size_t tree_layer_off(size_t leaves, uint32_t t) { return 2 * leaves - 2 * (leaves >> t); }

// The query records of BasefoldProof in their bincode layout (ProofJob),
// read from where the commitment and the opening rounds left the data
// (basefold.rs:205-222: the codeword row and its main-tree path, then per
// round the pi sibling and its pi-tree path).  Blocks stride over the
// records of all jobs; a block's threads take the items of one record (row
// elements, main-tree siblings, then per round the sibling value and its
// tree siblings), each written with 8-byte stores (records are 8-aligned).
constexpr int kRecThreads = 256;
__global__ __launch_bounds__(kRecThreads) void k_proof_records(const ProofJob *__restrict__ jobs, uint32_t njobs,
                                                               uint32_t records) {
    for (uint32_t r = blockIdx.x; r < records; r += gridDim.x) {
        uint32_t lo = 0, hi = njobs - 1;
        while (lo < hi) {
            const uint32_t mid = (lo + hi + 1) >> 1;
            if (jobs[mid].rec0 <= r) lo = mid;
            else hi = mid - 1;
        }
        const ProofJob &J = jobs[lo];
        const uint32_t t = r - J.rec0;
        const uint64_t id = J.idx[t];
        const size_t w = J.w;
        const uint32_t b = J.b, d = J.d, L = b + d;
        const size_t M = (size_t)1 << L;
        uint8_t *o = J.out + (size_t)t * J.rec_bytes;
        const size_t main_off = 16 + 16 * w, rounds_off = 24 + 16 * w + 32 * (size_t)L;
        // items: w row elements, L main siblings, per round k one value and
        // b + d - 1 - k siblings
        const size_t items = w + L + d + (size_t)d * (b + d - 1) - (size_t)d * (d - 1) / 2;
        if (threadIdx.x == 0) {
            *reinterpret_cast<uint64_t *>(o) = w;
            *reinterpret_cast<uint64_t *>(o + 8 + 16 * w) = L;
            *reinterpret_cast<uint64_t *>(o + 16 + 16 * w + 32 * (size_t)L) = d;
        }
        for (size_t i = threadIdx.x; i < items; i += kRecThreads) {
            if (i < w) {
                st_elem8(o + 8 + 16 * i, J.rows ? J.rows + id * w + i : J.cw + id);
            } else if (i < w + L) {
                const uint32_t u = (uint32_t)(i - w);
                st_digest8(o + main_off + 32 * (size_t)u, J.tree + 4 * (tree_layer_off(M, u) + ((id >> u) ^ 1)));
            } else {
                size_t j = i - w - L;
                uint32_t k = 0;
                while (j >= (size_t)(b + d - k)) {  // round k holds 1 + (b + d - 1 - k) items
                    j -= b + d - k;
                    k++;
                }
                const uint32_t bi = b + d - 1 - k;  // b + i for level i = d - 1 - k
                const size_t ni = (size_t)1 << bi;
                const size_t sib = ((id >> bi) & 1) ^ 1, row = id & (ni - 1);
                uint8_t *ro = o + rounds_off + 24 * (size_t)k +
                              32 * ((size_t)k * (b + d - 1) - (size_t)k * (k - 1) / 2);
                if (j == 0) {
                    const gf *pk = k == 0 ? J.cw : J.pis + (M - (M >> (k - 1)));
                    st_elem8(ro, pk + row + sib * ni);
                    *reinterpret_cast<uint64_t *>(ro + 16) = bi;
                } else {
                    const uint32_t u = (uint32_t)(j - 1);
                    const uint64_t *tk = J.pitrees + (k == 0 ? 0 : 4 * (2 * M - (M >> (k - 1)) - k));
                    st_digest8(ro + 24 + 32 * (size_t)u, tk + 4 * (tree_layer_off(ni, u) + ((row >> u) ^ 1)));
                }
            }
        }
    }
}

// ---------------------------------------------------------------------------
// micro-benchmarks of the two multiply forms and of Keccak-f
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(CB) void k_bench_mul_comb(gf *io, size_t n, int iters) {
    __shared__ uint32_t lds[80 * CB];
    uint32_t *tab = comb_slot(lds);
    gf_comb_init(tab, CB);
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        gf a = gf_ld(io + 2 * i), b = gf_ld(io + 2 * i + 1);
        for (int k = 0; k < iters; k++) {
            gf c = gf_mul(a, b, tab, CB);
            b = a;
            a = c;
        }
        gf_st(io + 2 * i, a);
    }
}
__global__ __launch_bounds__(CB) void k_bench_mul_reg(gf *io, size_t n, int iters) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        gf a = gf_ld(io + 2 * i), b = gf_ld(io + 2 * i + 1);
        for (int k = 0; k < iters; k++) {
            gf c = gf_mul_reg(a, b);
            b = a;
            a = c;
        }
        gf_st(io + 2 * i, a);
    }
}
// `iters` chained products in the four-lane form (gf_mul_wide), whole waves:
// lanes 0..15 of each wave chain their own operands (the single-wave latency
// of a transcript step's product at n = 64)
__global__ __launch_bounds__(CB) void k_bench_mul_wide(gf *io, size_t n, int iters) {
    const int lane = threadIdx.x & 63;
    for (size_t i0 = blockIdx.x * (size_t)blockDim.x + threadIdx.x - lane; i0 < n;
         i0 += (size_t)gridDim.x * blockDim.x) {
        const size_t i = i0 + lane;
        gf a = i < n ? gf_ld(io + 2 * i) : gf_zero(), b = i < n ? gf_ld(io + 2 * i + 1) : gf_zero();
        for (int k = 0; k < iters; k++) {
            gf c = gf_mul_wide(a, b);
            b = a;
            a = c;
        }
        if (i < n && lane < 16) gf_st(io + 2 * i, a);
    }
}
// `iters` chained products by one factor c = io[1] per thread through the
// block's window table of c (the products by a round challenge)
__global__ __launch_bounds__(CB) void k_bench_mul_tab(gf *io, size_t n, int iters) {
    __shared__ GfTab T;
    __shared__ gf u[128];
    gf_tab_build(T, u, gf_ld(io + 1));
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        gf a = gf_ld(io + 2 * i);
        for (int k = 0; k < iters; k++) a = gf_tab_mul(a, T);
        gf_st(io + 2 * i, a);
    }
}
// `iters` chained Keccak-256 hashes of 64-byte messages per thread (the node
// hash of k_nodes: next message = digest || digest ^ previous digest), the
// messages stored word-major (io[k * n + i]) so loads and stores are
// coalesced: the memory-free ceiling of the Merkle hashing kernels
__global__ __launch_bounds__(KB) void k_bench_keccak(uint64_t *io, size_t n, int iters) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        uint64_t m[8], d[4];
#pragma unroll
        for (int k = 0; k < 8; k++) m[k] = io[k * n + i];
        for (int k = 0; k < iters; k++) {
            kk::keccak256_words<8>(m, d);
#pragma unroll
            for (int w = 0; w < 4; w++) {
                m[4 + w] = m[w] ^ d[w];
                m[w] = d[w];
            }
        }
#pragma unroll
        for (int k = 0; k < 8; k++) io[k * n + i] = m[k];
    }
}

// `iters` chained Keccak-256 hashes of 64-byte messages on one lane pair
// (keccak256_words_pair, the form of the sumchecks' and open rounds'
// transcript steps): one block of one wave, io[0..8) the message; the time
// per hash is the latency of one transcript step's permutation
__global__ __launch_bounds__(64) void k_bench_keccak_pair(uint64_t *io, int iters) {
    const bool odd = threadIdx.x & 1;
    uint64_t m[8], d[4];
#pragma unroll
    for (int k = 0; k < 8; k++) m[k] = io[k];
    for (int k = 0; k < iters; k++) {
        kk::keccak256_words_pair<8>(m, d, odd);
#pragma unroll
        for (int w = 0; w < 4; w++) {
            m[4 + w] = m[w] ^ d[w];
            m[w] = d[w];
        }
    }
    if (threadIdx.x == 0)
#pragma unroll
        for (int k = 0; k < 8; k++) io[k] = m[k];
}

// out[i] = a[i] * b[i] (gf_mul_reg), or with shared_b a[i] * b[0] through
// b's masked pieces prepared once (gf_prepare / gf_mul_pre), the form the
// fold, fix and NTT kernels use for a challenge or a twiddle
__global__ __launch_bounds__(CB) void k_mul(const gf *__restrict__ a, const gf *__restrict__ b, gf *__restrict__ out,
                                            size_t n, int form) {
    if (form == 2) {  // four lanes per product, as the transcript steps; whole waves take part
        const int lane = threadIdx.x & 63;
        for (size_t i0 = blockIdx.x * (size_t)blockDim.x + threadIdx.x - lane; i0 < n;
             i0 += (size_t)gridDim.x * blockDim.x) {
            const size_t i = i0 + lane;
            const gf x = i < n ? gf_ld(a + i) : gf_zero(), y = i < n ? gf_ld(b + i) : gf_zero();
            gf r = gf_zero();
#pragma unroll
            for (int q = 0; q < 4; q++) r = gf_sel((lane >> 4) == q, gf_mul_wide(x, y, 16 * q), r);
            if (i < n) gf_st(out + i, r);
        }
        return;
    }
    if (form == 1) {  // through the window table of b[0], as the kernels that multiply by a round challenge
        __shared__ GfTab T;
        __shared__ gf u[128];
        gf_tab_build(T, u, gf_ld(b));
        for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
            gf_st(out + i, gf_tab_mul(gf_ld(a + i), T));
        return;
    }
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        gf_st(out + i, gf_mul_reg(gf_ld(a + i), gf_ld(b + i)));
}

// plain 16-byte-per-thread copy: the reference point for HBM byte counters
// p[i] = splitmix64(seed + i): synthetic input with the bit statistics of
// random field elements (timing runs; the switching activity of the data
// moves the clock under a power limit)
__global__ void k_fill_mix(uint64_t *__restrict__ p, size_t n, uint64_t seed) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        uint64_t z = seed + 0x9e3779b97f4a7c15ull * (i + 1);
        z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
        z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
        p[i] = z ^ (z >> 31);
    }
}

__global__ void k_copy16(const gf *__restrict__ src, gf *__restrict__ dst, size_t n) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        gf_st(dst + i, gf_ld(src + i));
}

}  // namespace

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
namespace launch {

static const unsigned kCapMul = 256 * 8;
static const unsigned kCapHash = 256 * 16;

static bool shape_ok(const Shape &sh) {
    const auto blk = [](int b) { return b == 64 || b == 128 || b == 256; };
    return blk(sh.ntt_block) && blk(sh.hash_block) && sh.ntt_cap > 0 && sh.hash_cap > 0 &&
           (sh.ntt_waves == 2 || sh.ntt_waves == 3);
}

hipError_t xor_span(hipStream_t st, gf *ft, const gf *bt, int bits, size_t count) {
    k_xor_span<<<grid_for(count, 256, 65535), 256, 0, st>>>(ft, bt, bits, count);
    return hipGetLastError();
}

hipError_t ntt_first(hipStream_t st, const gf *data, gf *cw, const gf *ft0, int b, int n, const Shape &sh) {
    if (!shape_ok(sh)) return hipErrorInvalidValue;
    const int c = 1 << b;
    if (n == 0) {
        k_replicate<<<1, 64, 0, st>>>(data, cw, c, 1);
        return hipGetLastError();
    }
    const size_t work = ((size_t)1 << n) / 2 * c;
    k_ntt_first<<<grid_for(work, sh.ntt_block, sh.ntt_cap), sh.ntt_block, 0, st>>>(data, cw, ft0, b, work);
    return hipGetLastError();
}
hipError_t ntt_layer(hipStream_t st, gf *cw, const gf *fts, int s, int b, int n, const Shape &sh) {
    if (!shape_ok(sh)) return hipErrorInvalidValue;
    const size_t work = ((size_t)1 << n) / 2 << b;
    k_ntt_layer<<<grid_for(work, sh.ntt_block, sh.ntt_cap), sh.ntt_block, 0, st>>>(cw, fts, s, b, work);
    return hipGetLastError();
}
hipError_t ntt_multi(hipStream_t st, gf *cw, const gf *t0, const gf *t1, const gf *t2, int R, int s, int b, int n,
                     const Shape &sh) {
    if (R < 2 || R > 3 || s + R > n || !shape_ok(sh)) return hipErrorInvalidValue;
    const size_t groups = ((size_t)1 << (n + b)) >> R;
    const int bs = sh.ntt_block;
    const unsigned g = grid_for(groups, bs, sh.ntt_cap);
    if (R == 2) k_ntt_multi<2><<<g, bs, 0, st>>>(cw, t0, t1, t2, s + b, groups);
    else if (sh.ntt_waves == 3) k_ntt_multi3_w3<<<g, bs, 0, st>>>(cw, t0, t1, t2, s + b, groups);
    else k_ntt_multi<3><<<g, bs, 0, st>>>(cw, t0, t1, t2, s + b, groups);
    return hipGetLastError();
}

hipError_t ntt_first_w(hipStream_t st, const gf *data, gf *cw, const gf *ft0, int b, int n, size_t w,
                       const Shape &sh) {
    if (n < 1 || w < 1 || !shape_ok(sh)) return hipErrorInvalidValue;
    const size_t work = ((size_t)1 << n) / 2 * ((size_t)1 << b);
    k_ntt_first_w<<<grid_for(work * w, sh.ntt_block, sh.ntt_cap), sh.ntt_block, 0, st>>>(data, cw, ft0, b, work, w);
    return hipGetLastError();
}
hipError_t ntt_multi_w(hipStream_t st, gf *cw, const gf *t0, const gf *t1, const gf *t2, int R, int s, int b, int n,
                       size_t w, const Shape &sh) {
    if (R < 1 || R > 3 || s + R > n || w < 1 || !shape_ok(sh)) return hipErrorInvalidValue;
    const size_t groups = ((size_t)1 << (n + b)) >> R;
    const int bs = sh.ntt_block;
    const unsigned g = grid_for(groups * w, bs, sh.ntt_cap);
    if (R == 1) k_ntt_multi_w<1><<<g, bs, 0, st>>>(cw, t0, t1, t2, s + b, groups, w);
    else if (R == 2) k_ntt_multi_w<2><<<g, bs, 0, st>>>(cw, t0, t1, t2, s + b, groups, w);
    else k_ntt_multi_w<3><<<g, bs, 0, st>>>(cw, t0, t1, t2, s + b, groups, w);
    return hipGetLastError();
}
hipError_t leaves_rows(hipStream_t st, const gf *rows, size_t w, uint64_t *out, size_t count, const Shape &sh) {
    if (!shape_ok(sh)) return hipErrorInvalidValue;
    k_leaves_rows<<<grid_for(count, sh.hash_block, sh.hash_cap), sh.hash_block, 0, st>>>(rows, w, out, count);
    return hipGetLastError();
}
hipError_t rowdot(hipStream_t st, const gf *m, size_t w, const gf *scalars, gf *out, size_t rows) {
    k_rowdot<<<grid_for(rows, CB, kCapMul), CB, 0, st>>>(m, w, scalars, out, rows);
    return hipGetLastError();
}
hipError_t leaves_elem(hipStream_t st, const gf *rows, uint64_t *out, size_t count, const Shape &sh) {
    if (!shape_ok(sh)) return hipErrorInvalidValue;
    k_leaves_elem<<<grid_for(count, sh.hash_block, sh.hash_cap), sh.hash_block, 0, st>>>(rows, out, count);
    return hipGetLastError();
}
hipError_t leaves_pair(hipStream_t st, const gf *pi, uint64_t *out, size_t half, const Shape &sh) {
    if (!shape_ok(sh)) return hipErrorInvalidValue;
    k_leaves_pair<<<grid_for(half, sh.hash_block, sh.hash_cap), sh.hash_block, 0, st>>>(pi, out, half);
    return hipGetLastError();
}
hipError_t nodes(hipStream_t st, const uint64_t *prev, uint64_t *next, size_t count, const Shape &sh) {
    if (!shape_ok(sh)) return hipErrorInvalidValue;
    k_nodes<<<grid_for(count, sh.hash_block, sh.hash_cap), sh.hash_block, 0, st>>>(prev, next, count);
    return hipGetLastError();
}
// Widest level (nodes per block) that runs on lane pairs: paired levels of
// `blocks` blocks fill at most one wave per SIMD of `cus` compute units.
int pair_width(unsigned blocks, int cus) {
    if (cus <= 0 || blocks == 0) return 0;
    return (int)std::min<size_t>(kSubThreads / 2, (size_t)(kSubThreads / 2) * (size_t)cus / blocks);
}

hipError_t subtree(hipStream_t st, uint64_t *prev, size_t h, int lv, int cus) {
    if (lv < 1 || lv > kSubMaxLevels || (h & (h - 1)) || (h >> lv) == 0) return hipErrorInvalidValue;
    const unsigned g = grid_for(h >> lv, 1, kCapHash);
    k_subtree<<<g, kSubThreads, 0, st>>>(prev, h, lv, pair_width(g, cus));
    return hipGetLastError();
}

unsigned top_round_dot_blocks(size_t n) { return grid_for(n, kSubThreads, 512); }
// blocks of k_top_round resident per compute unit (occupancy query, once)
static unsigned top_round_blocks_per_cu() {
    static unsigned cached = 0;
    if (!cached) {
        int nb = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_top_round, kSubThreads, 0) != hipSuccess || nb < 1)
            nb = 1;
        cached = (unsigned)nb;
    }
    return cached;
}
hipError_t top_round(hipStream_t st, uint64_t *prev, size_t h, int lv, const gf *a, const gf *b, size_t n,
                     gf *partial, const Seg *round, unsigned *counter, int cus) {
    if (lv < 1 || lv > kSubMaxLevels || (h & (h - 1)) || (h >> lv) == 0 || (h >> lv) > 4096) return hipErrorInvalidValue;
    // every block resident at once: a dot block left over for a second
    // wave of blocks would start only when another finishes, doubling the
    // launch (the kernel's registers allow few blocks per compute unit)
    const unsigned nsub = (unsigned)(h >> lv);
    const unsigned resident = cus > 0 ? (unsigned)cus * top_round_blocks_per_cu() : ~0u;
    const unsigned ndot = std::max(1u, std::min(top_round_dot_blocks(n), resident > nsub ? resident - nsub : 1u));
    k_top_round<<<nsub + ndot, kSubThreads, 0, st>>>(prev, h, lv, nsub, a, b, n, partial, ndot, round, counter,
                                                     pair_width(nsub, cus));
    return hipGetLastError();
}

hipError_t dot(hipStream_t st, const gf *a, const gf *b, size_t n, gf *partial, unsigned *nblocks,
               const uint64_t *root) {
    const unsigned g = dot_blocks(n);
    k_dot<<<g, CB, 0, st>>>(a, b, n, partial, root);
    *nblocks = g;
    return hipGetLastError();
}
unsigned dot_blocks(size_t n) { return grid_for(n, CB, 512); }
hipError_t fold(hipStream_t st, const gf *w, gf *out, const gf *ft, size_t mid, gf r) {
    k_fold<<<grid_for(mid, CB, kCapMul), CB, 0, st>>>(w, out, ft, mid, r);
    return hipGetLastError();
}
hipError_t fold_fix(hipStream_t st, const gf *w, gf *wout, const gf *ft, size_t mid, const gf *in, gf *out,
                    size_t half, const gf *x) {
    const unsigned nfold = grid_for(mid, CB, kCapMul), nfix = grid_for(half, CB, kCapMul);
    // the window table pays once every fold thread has two outputs or more
    if (mid >= 2 * (size_t)nfold * CB)
        k_fold_fix<true><<<nfold + nfix, CB, 0, st>>>(w, wout, ft, mid, in, out, half, x, nfold);
    else
        k_fold_fix<false><<<nfold + nfix, CB, 0, st>>>(w, wout, ft, mid, in, out, half, x, nfold);
    return hipGetLastError();
}
hipError_t fix(hipStream_t st, const gf *in, gf *out, size_t half, gf x) {
    const unsigned g = grid_for(half, CB, kCapMul);
    if (half >= 2 * (size_t)g * CB) k_fix<true><<<g, CB, 0, st>>>(in, out, half, x);
    else k_fix<false><<<g, CB, 0, st>>>(in, out, half, x);
    return hipGetLastError();
}
hipError_t outer(hipStream_t st, const gf *A, int m, const gf *B, gf *out, size_t n) {
    k_outer<<<grid_for(n, CB, kCapMul), CB, 0, st>>>(A, m, B, out, n);
    return hipGetLastError();
}
// blocks of k_prod_round<M> resident per compute unit (occupancy query, once)
static unsigned prod_blocks_per_cu() {
    static unsigned cached = 0;
    if (!cached) {
        int nb = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_prod_round<2>, CB, 0) != hipSuccess || nb < 1) nb = 4;
        cached = (unsigned)nb;
    }
    return cached;
}
hipError_t prod_step(hipStream_t st, const gf *fsrc, const gf *gsrc, gf *fdst, gf *gdst, EqFactors eq, size_t half,
                     int k, ProdFs *fs, int cus) {
    if (half <= kTailPairs || (half & (half - 1)) || eq.lo < 7 || eq.lo > 62) return hipErrorInvalidValue;
    // pairs per block: a power of two from CB up to 2^lo (one B entry per
    // block), the smallest that keeps the grid within the resident blocks
    const size_t cap = (size_t)std::max(1, cus) * prod_blocks_per_cu();
    size_t C = CB;
    while (half / C > cap && C < ((size_t)1 << eq.lo)) C <<= 1;
    const size_t g = half / C;
    if (g > 0x7fffffffu) return hipErrorInvalidValue;
    if (k == 0) k_prod_round<0><<<(unsigned)g, CB, 0, st>>>(fsrc, gsrc, fdst, gdst, eq, half, k, fs);
    else if (fsrc != fdst) k_prod_round<1><<<(unsigned)g, CB, 0, st>>>(fsrc, gsrc, fdst, gdst, eq, half, k, fs);
    else k_prod_round<2><<<(unsigned)g, CB, 0, st>>>(fsrc, gsrc, fdst, gdst, eq, half, k, fs);
    return hipGetLastError();
}
size_t prod_tail_max_pairs() { return kTailPairs; }
size_t prod_look_max_pairs() { return kProdLookMax; }
hipError_t prod_look(hipStream_t st, const gf *fsrc, const gf *gsrc, gf *fdst, gf *gdst, EqFactors eq, size_t half,
                     int k, ProdFs *fs, bool first) {
    // q / 64 worker blocks of kLookGroups pairs of round k+1, within one
    // entry of B
    if (k < 1 || half <= kTailPairs || half > kProdLookMax || (half & (half - 1)) || eq.lo < 5 || eq.lo > 62)
        return hipErrorInvalidValue;
    const unsigned wblocks = (unsigned)(half / (2 * kLookGroups));
    if (first) k_prod_look<true><<<wblocks, CB, 0, st>>>(fsrc, gsrc, fdst, gdst, eq, half, k, fs);
    else k_prod_look<false><<<wblocks + 1, CB, 0, st>>>(fsrc, gsrc, fdst, gdst, eq, half, k, fs);
    return hipGetLastError();
}
hipError_t prod_tail(hipStream_t st, const gf *fsrc, const gf *gsrc, gf *fout, gf *gout, EqFactors eq, size_t half,
                     int k0, ProdFs *fs, uint8_t *host, uint32_t host_bytes) {
    if (!half || half > kTailPairs || (half & (half - 1)) || eq.lo < 0 || eq.lo > 62 || (host_bytes & 7))
        return hipErrorInvalidValue;
    k_prod_tail<<<1, kTailThreads, 0, st>>>(fsrc, gsrc, fout, gout, eq, half, k0, fs, host, host_bytes);
    return hipGetLastError();
}
hipError_t mle_dot(hipStream_t st, const gf *data, size_t w, const gf *A, int m, const gf *B, size_t rows, gf *partial,
                   unsigned max_blocks, unsigned *nblocks) {
    if (w == 0 || w > 65535 || max_blocks == 0) return hipErrorInvalidValue;
    const unsigned gx = grid_for(rows, CB, max_blocks);
    k_mle_dot<<<dim3(gx, (unsigned)w), CB, 0, st>>>(data, w, A, m, B, rows, partial);
    *nblocks = gx;
    return hipGetLastError();
}
hipError_t quad_sums(hipStream_t st, const gf *P, size_t stride, const uint32_t *pairs, uint32_t npairs, size_t h,
                     gf *partial, unsigned max_blocks, unsigned *nblocks) {
    if (npairs == 0 || npairs > 65535 || max_blocks == 0) return hipErrorInvalidValue;
    const unsigned gx = grid_for(h, CB, max_blocks);
    k_quad_sums<<<dim3(gx, npairs), CB, 0, st>>>(P, stride, pairs, h, partial);
    *nblocks = gx;
    return hipGetLastError();
}
hipError_t fix_multi(hipStream_t st, gf *P, size_t stride, uint32_t m, size_t half, gf x) {
    if (m == 0 || m > 65535) return hipErrorInvalidValue;
    k_fix_multi<<<dim3(grid_for(half, CB, kCapMul / 4), m), CB, 0, st>>>(P, stride, half, x);
    return hipGetLastError();
}
hipError_t fix_multi_dev(hipStream_t st, const gf *src, size_t src_stride, gf *dst, size_t dst_stride, uint32_t m,
                         size_t half, const gf *x) {
    if (m == 0 || m > 65535) return hipErrorInvalidValue;
    const unsigned g = grid_for(half, CB, kCapMul / 4);
    if (half >= 2 * (size_t)g * CB)
        k_fix_multi_dev<true><<<dim3(g, m), CB, 0, st>>>(src, src_stride, dst, dst_stride, half, x);
    else
        k_fix_multi_dev<false><<<dim3(g, m), CB, 0, st>>>(src, src_stride, dst, dst_stride, half, x);
    return hipGetLastError();
}
hipError_t quad_step(hipStream_t st, const gf *P, size_t stride, const uint32_t *pairs, uint32_t npairs, size_t h,
                     int k, QuadFs *fs, gf *slots, unsigned max_blocks, unsigned *counter) {
    if (npairs == 0 || npairs > 65535 || max_blocks == 0) return hipErrorInvalidValue;
    const unsigned gx = grid_for(h, CB, max_blocks);
    k_quad_step<<<dim3(gx, npairs), CB, 0, st>>>(P, stride, pairs, h, k, fs, slots, counter);
    return hipGetLastError();
}
size_t quad_look_max_half() { return kQuadLookMax; }
hipError_t quad_look(hipStream_t st, const gf *src, size_t src_stride, gf *dst, size_t dst_stride,
                     const uint32_t *pairs, uint32_t npairs, uint32_t m, size_t half, int k, QuadFs *fs, bool first) {
    // half / 64 blocks per pair of kLookGroups positions of round k+1
    if (k < 1 || npairs == 0 || npairs > 65534 || m == 0 || half < 2 * kLookGroups || half > kQuadLookMax ||
        (half & (half - 1)))
        return hipErrorInvalidValue;
    const unsigned gx = (unsigned)(half / (2 * kLookGroups));
    if (first)
        k_quad_look<true><<<dim3(gx, npairs), CB, 0, st>>>(src, src_stride, dst, dst_stride, pairs, m, half, k, fs);
    else
        k_quad_look<false><<<dim3(gx, npairs + 1), CB, 0, st>>>(src, src_stride, dst, dst_stride, pairs, m, half, k,
                                                                 fs);
    return hipGetLastError();
}
size_t quad_tail_max_half(uint32_t m, uint32_t npairs) {
    if (npairs == 0 || npairs > kQuadTailPairs || m == 0) return 0;
    size_t h = 256;
    while (h && (size_t)m * 2 * h > kQuadTailElems) h >>= 1;
    return h;
}
hipError_t quad_tail(hipStream_t st, const gf *src, size_t src_stride, gf *fin, size_t fin_stride,
                     const uint32_t *pairs, uint32_t npairs, uint32_t m, size_t h0, int k0, QuadFs *fs,
                     uint8_t *host, uint32_t host_bytes) {
    if (!h0 || (h0 & (h0 - 1)) || h0 > quad_tail_max_half(m, npairs) || (host_bytes & 7)) return hipErrorInvalidValue;
    k_quad_tail<<<1, kQuadTailThreads, 0, st>>>(src, src_stride, fin, fin_stride, pairs, m, h0, k0, fs, host,
                                                host_bytes);
    return hipGetLastError();
}
hipError_t verify_fold(hipStream_t st, const VerifyJob *jobs, uint32_t count, uint32_t q, int b, int d, size_t w,
                       const gf *ft, gf *vals, uint8_t *fold_ok) {
    const size_t n = (size_t)count * q;
    if (!n) return hipSuccess;
    k_verify_fold<<<grid_for(n, CB, kCapMul), CB, 0, st>>>(jobs, count, q, b, d, w, ft, vals, fold_ok);
    return hipGetLastError();
}
hipError_t verify_paths(hipStream_t st, const VerifyJob *jobs, uint32_t count, uint32_t q, int b, int d, size_t w,
                        const gf *vals, uint8_t *path_ok) {
    const size_t n = (size_t)count * q * (size_t)(d + 1);
    if (!n) return hipSuccess;
    k_verify_paths<<<grid_for(n, KB, kCapHash), KB, 0, st>>>(jobs, count, q, b, d, w, vals, path_ok);
    return hipGetLastError();
}

hipError_t seg_launch(hipStream_t st, SegKernel which, const Seg *d_segs, int nseg, uint32_t nblocks, gf *partial) {
    if (nseg <= 0 || nblocks == 0) return hipSuccess;
    switch (which) {
    case SegKernel::NttFirst: ks_ntt_first<<<nblocks, 256, 0, st>>>(d_segs, nseg); break;
    case SegKernel::NttLayer: ks_ntt_layer<<<nblocks, 256, 0, st>>>(d_segs, nseg); break;
    case SegKernel::NttMulti2: ks_ntt_multi<2><<<nblocks, 256, 0, st>>>(d_segs, nseg); break;
    case SegKernel::NttMulti3: ks_ntt_multi<3><<<nblocks, 256, 0, st>>>(d_segs, nseg); break;
    case SegKernel::LeavesElem: ks_leaves_elem<<<nblocks, KB, 0, st>>>(d_segs, nseg); break;
    case SegKernel::LeavesPairDot: ks_leaves_pair_dot<<<nblocks, KB, 0, st>>>(d_segs, nseg, partial); break;
    case SegKernel::Nodes: ks_nodes<<<nblocks, KB, 0, st>>>(d_segs, nseg); break;
    case SegKernel::Subtree: ks_subtree<<<nblocks, kSubThreads, 0, st>>>(d_segs, nseg); break;
    case SegKernel::Fix: ks_fix<<<nblocks, CB, 0, st>>>(d_segs, nseg); break;
    case SegKernel::FoldFix: ks_fold_fix<<<nblocks, CB, 0, st>>>(d_segs, nseg); break;
    case SegKernel::Outer: ks_outer<<<nblocks, CB, 0, st>>>(d_segs, nseg); break;
    case SegKernel::Round: ks_round<<<nblocks, 64, 0, st>>>(d_segs, nseg, partial); break;
    }
    return hipGetLastError();
}
int seg_block_size(SegKernel which) {
    switch (which) {
    case SegKernel::NttFirst:
    case SegKernel::NttLayer:
    case SegKernel::NttMulti2:
    case SegKernel::NttMulti3: return 256;
    case SegKernel::LeavesElem:
    case SegKernel::LeavesPairDot:
    case SegKernel::Nodes: return KB;
    case SegKernel::Subtree: return kSubThreads;
    case SegKernel::Round: return 64;
    default: return CB;
    }
}
int subtree_max_levels() { return kSubMaxLevels; }
hipError_t gather_abs(hipStream_t st, const uint64_t *src, void *out, size_t n, int rb) {
    if (!n) return hipSuccess;
    k_gather_abs<<<grid_for(n, 256, 4096), 256, 0, st>>>(src, (uint8_t *)out, n, rb);
    return hipGetLastError();
}
hipError_t proof_records(hipStream_t st, const ProofJob *jobs, uint32_t njobs, uint32_t records) {
    if (!records) return hipSuccess;
    if (!njobs) return hipErrorInvalidValue;
    k_proof_records<<<grid_for(records, 1, 4096), kRecThreads, 0, st>>>(jobs, njobs, records);
    return hipGetLastError();
}
hipError_t bench_mul(hipStream_t st, int variant, gf *io, size_t n, int iters) {
    if (variant == 0)
        k_bench_mul_comb<<<grid_for(n, CB, 1 << 20), CB, 0, st>>>(io, n, iters);
    else if (variant == 4)
        k_bench_mul_tab<<<grid_for(n, CB, 1 << 20), CB, 0, st>>>(io, n, iters);
    else if (variant == 6)
        k_bench_mul_wide<<<grid_for(n, CB, 1 << 20), CB, 0, st>>>(io, n, iters);
    else
        k_bench_mul_reg<<<grid_for(n, CB, 1 << 20), CB, 0, st>>>(io, n, iters);
    return hipGetLastError();
}
hipError_t mul(hipStream_t st, const gf *a, const gf *b, gf *out, size_t n, int form) {
    if (!n) return hipSuccess;
    if (form < 0 || form > 2) return hipErrorInvalidValue;
    k_mul<<<grid_for(n, CB, kCapMul), CB, 0, st>>>(a, b, out, n, form);
    return hipGetLastError();
}
hipError_t copy16(hipStream_t st, const gf *src, gf *dst, size_t n) {
    k_copy16<<<grid_for(n, 256, 1 << 16), 256, 0, st>>>(src, dst, n);
    return hipGetLastError();
}
hipError_t fill_mix(hipStream_t st, uint64_t *p, size_t n, uint64_t seed) {
    k_fill_mix<<<grid_for(n, 256, 1 << 14), 256, 0, st>>>(p, n, seed);
    return hipGetLastError();
}
hipError_t bench_keccak(hipStream_t st, uint64_t *io, size_t n, int iters) {
    k_bench_keccak<<<grid_for(n, KB, 1 << 20), KB, 0, st>>>(io, n, iters);
    return hipGetLastError();
}
hipError_t bench_keccak_pair(hipStream_t st, uint64_t *io, int iters) {
    k_bench_keccak_pair<<<1, 64, 0, st>>>(io, iters);
    return hipGetLastError();
}

}  // namespace launch
}  // namespace bfd
