// POLYVAL GF(2^128) arithmetic for CDNA4 (gfx950) and the host.
//
// Field: GF(2)[X] / (X^128 + X^127 + X^126 + X^121 + 1).  Elements are kept in
// the reference's Montgomery form (value a*X^128, p3/src/polyval.rs:30-34,
// p3/src/polyval/portable.rs:20-28): 16 bytes, little-endian, low 64 bits first.
// Addition is XOR; multiplication is the POLYVAL "dot" a*b*X^-128.
//
// CDNA4 has no carry-less multiply.  The device multiply used by every kernel
// is gf_mul_reg / gf_mul_pre (register Karatsuba over 32x32 carry-less
// products built from masked integer multiply-adds, below); about 466 VALU
// instructions per product.  The 256-bit product is reduced with the two-step
// Montgomery fold of the reference backends (x86_64_avx2.rs:41-68).
//
// gf_comb_* is a second device multiply kept for the micro-benchmark only: a
// 4-bit-window comb with a per-lane 16-entry table of v*b in LDS (entries of
// 131 bits = 5 dwords), indexed by the 32 nibbles of a, eight nibble
// positions per 32-bit word together with one 256-bit shift by 4 between
// positions.  It measures about 30% slower than the register form.
#pragma once
#include <stdint.h>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define BFD_HD __host__ __device__ __forceinline__
#define BFD_D __device__ __forceinline__
#else
#define BFD_HD inline
#define BFD_D inline
#endif

struct gf {
    uint64_t lo, hi;
};

BFD_HD gf gf_zero() { return gf{0, 0}; }
BFD_HD gf gf_add(gf a, gf b) { return gf{a.lo ^ b.lo, a.hi ^ b.hi}; }
BFD_HD bool gf_eq(gf a, gf b) { return a.lo == b.lo && a.hi == b.hi; }
// ONE = from_canonical(1) in Montgomery form
BFD_HD gf gf_one() { return gf{1ull, 0xc200000000000000ull}; }

// Montgomery reduction of the 256-bit product (v0 least significant 64-bit word):
// returns v * X^-128 mod P.
BFD_HD gf gf_mont_reduce(uint64_t v0, uint64_t v1, uint64_t v2, uint64_t v3) {
    v2 ^= v0 ^ (v0 >> 1) ^ (v0 >> 2) ^ (v0 >> 7);
    v1 ^= (v0 << 63) ^ (v0 << 62) ^ (v0 << 57);
    v3 ^= v1 ^ (v1 >> 1) ^ (v1 >> 2) ^ (v1 >> 7);
    v2 ^= (v1 << 63) ^ (v1 << 62) ^ (v1 << 57);
    return gf{v2, v3};
}

#if defined(__HIPCC__)

// ---------------------------------------------------------------------------
// Device comb multiply.
//
// `tab` points at this lane's table slot in LDS: entry e dword k is at
// tab[(e * 5 + k) * stride].  With stride = block size the 64 lanes of a wave
// hit 64 consecutive dwords for any common (e, k) and distinct e give distinct
// bank offsets, so table reads are conflict-free.  Entry 0 must be zero; the
// kernel zeroes it once (gf_comb_init).
// ---------------------------------------------------------------------------
BFD_D void gf_comb_init(uint32_t *tab, int stride) {
#pragma unroll
    for (int k = 0; k < 5; k++) tab[k * stride] = 0;
}

BFD_D void gf_comb_table(uint32_t *tab, int stride, gf b) {
    uint32_t b0 = (uint32_t)b.lo, b1 = (uint32_t)(b.lo >> 32), b2 = (uint32_t)b.hi,
             b3 = (uint32_t)(b.hi >> 32);
    // t1 = b, t2 = b<<1, t4 = b<<2, t8 = b<<3 as 5-dword values
    uint32_t t[16][5];
    t[1][0] = b0; t[1][1] = b1; t[1][2] = b2; t[1][3] = b3; t[1][4] = 0;
#pragma unroll
    for (int s = 1; s < 4; s++) {
        int e = 1 << s;
        t[e][0] = b0 << s;
        t[e][1] = __builtin_amdgcn_alignbit(b1, b0, 32 - s);
        t[e][2] = __builtin_amdgcn_alignbit(b2, b1, 32 - s);
        t[e][3] = __builtin_amdgcn_alignbit(b3, b2, 32 - s);
        t[e][4] = b3 >> (32 - s);
    }
#pragma unroll
    for (int e = 3; e < 16; e++) {
        if ((e & (e - 1)) == 0) continue;
        int hb = 1 << (31 - __builtin_clz(e));
#pragma unroll
        for (int k = 0; k < 5; k++) t[e][k] = t[hb][k] ^ t[e - hb][k];
    }
#pragma unroll
    for (int e = 1; e < 16; e++)
#pragma unroll
        for (int k = 0; k < 5; k++) tab[(e * 5 + k) * stride] = t[e][k];
}

// a * (table operand) with the table already built for this lane.
BFD_D gf gf_comb_mul(gf a, const uint32_t *tab, int stride) {
    uint32_t aw[4] = {(uint32_t)a.lo, (uint32_t)(a.lo >> 32), (uint32_t)a.hi,
                      (uint32_t)(a.hi >> 32)};
    uint32_t acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
#pragma unroll
    for (int j = 7; j >= 0; j--) {
#pragma unroll
        for (int w = 0; w < 4; w++) {
            uint32_t v = (aw[w] >> (4 * j)) & 15u;
            const uint32_t *e = tab + v * 5 * stride;
#pragma unroll
            for (int k = 0; k < 5; k++) acc[w + k] ^= e[k * stride];
        }
        if (j) {
#pragma unroll
            for (int k = 7; k >= 1; k--) acc[k] = __builtin_amdgcn_alignbit(acc[k], acc[k - 1], 28);
            acc[0] <<= 4;
        }
    }
    uint64_t v0 = (uint64_t)acc[0] | ((uint64_t)acc[1] << 32);
    uint64_t v1 = (uint64_t)acc[2] | ((uint64_t)acc[3] << 32);
    uint64_t v2 = (uint64_t)acc[4] | ((uint64_t)acc[5] << 32);
    uint64_t v3 = (uint64_t)acc[6] | ((uint64_t)acc[7] << 32);
    return gf_mont_reduce(v0, v1, v2, v3);
}

// General a*b through this lane's table slot.
BFD_D gf gf_mul(gf a, gf b, uint32_t *tab, int stride) {
    gf_comb_table(tab, stride, b);
    return gf_comb_mul(a, tab, stride);
}

// ---------------------------------------------------------------------------
// Register multiply (no LDS).  A 32x32 carry-less product is four masked
// 32x32->64 integer multiply-adds per class: operands are split by the masks
// 0x1111.. << k so every data bit is followed by a three-bit hole; an
// integer product of two such pieces sums at most 8 terms per position, so
// carries stay inside the hole and bit 4i+c of the sum over the class-c
// pieces is the GF(2) coefficient.  128x128 uses two Karatsuba levels
// (9 such products); sums of three terms use v_bitop3_b32.
// ---------------------------------------------------------------------------
BFD_D uint32_t x3(uint32_t a, uint32_t b, uint32_t c) { return __builtin_amdgcn_bitop3_b32(a, b, c, 0x96); }

// masked pieces of a 32-bit operand
struct pieces {
    uint32_t p[4];
};
BFD_D pieces split4(uint32_t x) {
    const uint32_t m0 = 0x11111111u;
    return pieces{{x & m0, x & (m0 << 1), x & (m0 << 2), x & (m0 << 3)}};
}
BFD_D pieces pxor(pieces a, pieces b) {
    return pieces{{a.p[0] ^ b.p[0], a.p[1] ^ b.p[1], a.p[2] ^ b.p[2], a.p[3] ^ b.p[3]}};
}

struct u64p {
    uint32_t lo, hi;
};

// m ? a : b per bit, and (a ^ b) & m, each one v_bitop3_b32 (truth-table
// index 4a + 2b + m)
BFD_D uint32_t bsel(uint32_t m, uint32_t a, uint32_t b) { return __builtin_amdgcn_bitop3_b32(a, b, m, 0xE4); }
BFD_D uint32_t xand(uint32_t a, uint32_t b, uint32_t m) { return __builtin_amdgcn_bitop3_b32(a, b, m, 0x28); }

// One 32-bit word of a 32x32 carry-less product from its class sums: class
// c's terms (c0: 3, c1: 2, c2: 3, c3: 4 products) are XORed and bit 4i + c
// taken from class c, in eight bitwise operations (the masks are disjoint
// and cover the word)
BFD_D uint32_t clmul_pick(uint32_t c0a, uint32_t c0b, uint32_t c0c, uint32_t c1a, uint32_t c1b, uint32_t c2a,
                          uint32_t c2b, uint32_t c2c, uint32_t c3a, uint32_t c3b, uint32_t c3c, uint32_t c3d) {
    const uint32_t m0 = 0x11111111u, m1 = m0 << 1, m2 = m0 << 2, m3 = m0 << 3;
    const uint32_t t1 = xand(c1a, c1b, m1);
    const uint32_t t0 = bsel(m0, x3(c0a, c0b, c0c), t1);
    const uint32_t t2 = bsel(m2, x3(c2a, c2b, c2c), t0);
    return bsel(m3, x3(c3a, c3b, c3c) ^ c3d, t2);
}

// Class c collects the products x.p[i] * y.p[j] with i + j = c (mod 4).  Two
// products of one class whose i + j differ by 4 may be ADDED (one multiply-
// add) instead of XORed: at position 4u + c the first contributes at most
// min(u + 1, 15 - u) terms and the second at most min(u, 16 - u), together at
// most 15, so every count still fits its four-bit field and no carry reaches
// the next position of the class.  That merges (0,0)+(1,3), (0,1)+(2,3),
// (1,0)+(3,2) and (0,2)+(3,3) and leaves 10 XORs per product instead of 16.
BFD_D u64p clmul32p(const pieces &x, const pieces &y) {
    const uint64_t a00 = (uint64_t)x.p[1] * y.p[3] + (uint64_t)x.p[0] * y.p[0];
    const uint64_t a01 = (uint64_t)x.p[2] * y.p[3] + (uint64_t)x.p[0] * y.p[1];
    const uint64_t a10 = (uint64_t)x.p[3] * y.p[2] + (uint64_t)x.p[1] * y.p[0];
    const uint64_t a02 = (uint64_t)x.p[3] * y.p[3] + (uint64_t)x.p[0] * y.p[2];
    const uint64_t p22 = (uint64_t)x.p[2] * y.p[2], p31 = (uint64_t)x.p[3] * y.p[1];
    const uint64_t p11 = (uint64_t)x.p[1] * y.p[1], p20 = (uint64_t)x.p[2] * y.p[0];
    const uint64_t p03 = (uint64_t)x.p[0] * y.p[3], p12 = (uint64_t)x.p[1] * y.p[2];
    const uint64_t p21 = (uint64_t)x.p[2] * y.p[1], p30 = (uint64_t)x.p[3] * y.p[0];
    return u64p{clmul_pick((uint32_t)a00, (uint32_t)p22, (uint32_t)p31, (uint32_t)a01, (uint32_t)a10, (uint32_t)a02,
                           (uint32_t)p11, (uint32_t)p20, (uint32_t)p03, (uint32_t)p12, (uint32_t)p21, (uint32_t)p30),
                clmul_pick((uint32_t)(a00 >> 32), (uint32_t)(p22 >> 32), (uint32_t)(p31 >> 32), (uint32_t)(a01 >> 32),
                           (uint32_t)(a10 >> 32), (uint32_t)(a02 >> 32), (uint32_t)(p11 >> 32), (uint32_t)(p20 >> 32),
                           (uint32_t)(p03 >> 32), (uint32_t)(p12 >> 32), (uint32_t)(p21 >> 32), (uint32_t)(p30 >> 32))};
}

// b's masked pieces for the nine Karatsuba products, reusable across
// multiplies by the same b
struct gf_pre {
    pieces q[9];
};
BFD_D gf_pre gf_prepare(gf b) {
    pieces b0 = split4((uint32_t)b.lo), b1 = split4((uint32_t)(b.lo >> 32)), b2 = split4((uint32_t)b.hi),
           b3 = split4((uint32_t)(b.hi >> 32));
    pieces b01 = pxor(b0, b1), b23 = pxor(b2, b3), b02 = pxor(b0, b2), b13 = pxor(b1, b3), bx = pxor(b02, b13);
    return gf_pre{{b0, b1, b01, b2, b3, b23, b02, b13, bx}};
}

// 64x64 Karatsuba from three 32x32 products: returns 4 dwords
BFD_D void kara64(const pieces &a0, const pieces &a1, const pieces &a01, const pieces &b0, const pieces &b1,
                  const pieces &b01, uint32_t r[4]) {
    u64p l = clmul32p(a0, b0), h = clmul32p(a1, b1), m = clmul32p(a01, b01);
    uint32_t ml = x3(m.lo, l.lo, h.lo), mh = x3(m.hi, l.hi, h.hi);
    r[0] = l.lo;
    r[1] = l.hi ^ ml;
    r[2] = h.lo ^ mh;
    r[3] = h.hi;
}

// The product from its three 64x64 Karatsuba products (L of the low halves,
// H of the high halves, Mm of their sums), then the Montgomery fold.
BFD_D gf gf_combine(const uint32_t L[4], const uint32_t H[4], const uint32_t Mm[4]) {
    // 256-bit product w0..w7
    uint32_t w[8];
    w[0] = L[0];
    w[1] = L[1];
    w[2] = x3(L[2], Mm[0], x3(L[0], H[0], 0u));
    w[3] = x3(L[3], Mm[1], x3(L[1], H[1], 0u));
    w[4] = x3(H[0], Mm[2], x3(L[2], H[2], 0u));
    w[5] = x3(H[1], Mm[3], x3(L[3], H[3], 0u));
    w[6] = H[2];
    w[7] = H[3];
    // Montgomery fold (x86_64_avx2.rs:41-68) on 64-bit words v_k = w[2k] | w[2k+1] << 32
    // v2 ^= v0 ^ v0>>1 ^ v0>>2 ^ v0>>7 ; v1 ^= v0<<63 ^ v0<<62 ^ v0<<57
    uint32_t s1l = __builtin_amdgcn_alignbit(w[1], w[0], 1), s2l = __builtin_amdgcn_alignbit(w[1], w[0], 2),
             s7l = __builtin_amdgcn_alignbit(w[1], w[0], 7);
    w[4] = x3(x3(w[4], w[0], s1l), s2l, s7l);
    w[5] = x3(x3(w[5], w[1], w[1] >> 1), w[1] >> 2, w[1] >> 7);
    w[3] = x3(x3(w[3], w[0] << 31, w[0] << 30), w[0] << 25, 0u);
    // v3 ^= v1 ^ v1>>1 ^ v1>>2 ^ v1>>7 ; v2 ^= v1<<63 ^ v1<<62 ^ v1<<57
    uint32_t t1l = __builtin_amdgcn_alignbit(w[3], w[2], 1), t2l = __builtin_amdgcn_alignbit(w[3], w[2], 2),
             t7l = __builtin_amdgcn_alignbit(w[3], w[2], 7);
    w[6] = x3(x3(w[6], w[2], t1l), t2l, t7l);
    w[7] = x3(x3(w[7], w[3], w[3] >> 1), w[3] >> 2, w[3] >> 7);
    w[5] = x3(x3(w[5], w[2] << 31, w[2] << 30), w[2] << 25, 0u);
    return gf{(uint64_t)w[4] | ((uint64_t)w[5] << 32), (uint64_t)w[6] | ((uint64_t)w[7] << 32)};
}

BFD_D gf gf_mul_pre(gf a, const gf_pre &pb) {
    pieces a0 = split4((uint32_t)a.lo), a1 = split4((uint32_t)(a.lo >> 32)), a2 = split4((uint32_t)a.hi),
           a3 = split4((uint32_t)(a.hi >> 32));
    pieces a01 = pxor(a0, a1), a23 = pxor(a2, a3), a02 = pxor(a0, a2), a13 = pxor(a1, a3), ax = pxor(a02, a13);
    uint32_t L[4], H[4], Mm[4];
    kara64(a0, a1, a01, pb.q[0], pb.q[1], pb.q[2], L);
    kara64(a2, a3, a23, pb.q[3], pb.q[4], pb.q[5], H);
    kara64(a02, a13, ax, pb.q[6], pb.q[7], pb.q[8], Mm);
    return gf_combine(L, H, Mm);
}

BFD_D gf gf_mul_reg(gf a, gf b) { return gf_mul_pre(a, gf_prepare(b)); }

// ---------------------------------------------------------------------------
// Multiplication by a factor c shared by a whole block, through an LDS window
// table.  dot(a, c) is F2-linear in a, and a = sum_k nibble_k(a) X^(4k), so
// dot(a, c) = XOR_k T[16k + nibble_k(a)] with T[16k + v] = dot(v X^(4k), c):
// 32 windows of 16 entries (8 KiB).  One read of window k by a wave touches
// only that window's 16 entries, which occupy the 64 LDS banks exactly once,
// so the 32 reads of a product are free of bank conflicts; a product then
// costs about 130 VALU operations instead of about 390 for gf_mul_pre.
// ---------------------------------------------------------------------------
struct alignas(16) GfTab {
    gf e[512];
};

// a * X in GF(2^128) = GF(2)[X]/(X^128 + X^127 + X^126 + X^121 + 1)
BFD_D gf gf_times_x(gf a) {
    const uint64_t top = a.hi >> 63;
    gf r{a.lo << 1, (a.hi << 1) | (a.lo >> 63)};
    r.lo ^= top;
    r.hi ^= (0 - top) & 0xC200000000000000ull;
    return r;
}

// T for the factor c, built by the whole block (at least 32 threads):
// dot(X^(4t), c) for t < 32 by one product each, times X for the three
// windows' other unit entries, then every entry as an XOR of unit entries.
// Ends with a barrier; u is LDS scratch of 128 elements.
BFD_D void gf_tab_build(GfTab &T, gf *u, gf c) {
    const int t = threadIdx.x;
    if (t < 32) {
        const int i = 4 * t;
        gf v = gf_mul_reg(i < 64 ? gf{1ull << i, 0} : gf{0, 1ull << (i - 64)}, c);
        u[i] = v;
        v = gf_times_x(v);
        u[i + 1] = v;
        v = gf_times_x(v);
        u[i + 2] = v;
        u[i + 3] = gf_times_x(v);
    }
    __syncthreads();
    for (int e = t; e < 512; e += blockDim.x) {
        const int k = e >> 4, v = e & 15;
        gf acc = gf_zero();
        if (v & 1) acc = gf_add(acc, u[4 * k]);
        if (v & 2) acc = gf_add(acc, u[4 * k + 1]);
        if (v & 4) acc = gf_add(acc, u[4 * k + 2]);
        if (v & 8) acc = gf_add(acc, u[4 * k + 3]);
        T.e[e] = acc;
    }
    __syncthreads();
}

// dot(a, c) through the table of c
BFD_D gf gf_tab_mul(gf a, const GfTab &T) {
    const uint32_t w[4] = {(uint32_t)a.lo, (uint32_t)(a.lo >> 32), (uint32_t)a.hi, (uint32_t)(a.hi >> 32)};
    uint32_t r0 = 0, r1 = 0, r2 = 0, r3 = 0;
#pragma unroll
    for (int k = 0; k < 32; k += 2) {
        const gf e0 = T.e[16 * k + ((w[k >> 3] >> (4 * (k & 7))) & 15)];
        const gf e1 = T.e[16 * (k + 1) + ((w[(k + 1) >> 3] >> (4 * ((k + 1) & 7))) & 15)];
        r0 = x3(r0, (uint32_t)e0.lo, (uint32_t)e1.lo);
        r1 = x3(r1, (uint32_t)(e0.lo >> 32), (uint32_t)(e1.lo >> 32));
        r2 = x3(r2, (uint32_t)e0.hi, (uint32_t)e1.hi);
        r3 = x3(r3, (uint32_t)(e0.hi >> 32), (uint32_t)(e1.hi >> 32));
    }
    return gf{(uint64_t)r0 | ((uint64_t)r1 << 32), (uint64_t)r2 | ((uint64_t)r3 << 32)};
}

BFD_D gf gf_ld(const gf *p) {
    const uint4 v = *reinterpret_cast<const uint4 *>(p);
    return gf{(uint64_t)v.x | ((uint64_t)v.y << 32), (uint64_t)v.z | ((uint64_t)v.w << 32)};
}
BFD_D void gf_st(gf *p, gf a) {
    *reinterpret_cast<uint4 *>(p) =
        uint4{(uint32_t)a.lo, (uint32_t)(a.lo >> 32), (uint32_t)a.hi, (uint32_t)(a.hi >> 32)};
}
#endif
