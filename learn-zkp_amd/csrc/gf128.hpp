// POLYVAL GF(2^128) arithmetic for CDNA4 (gfx950) and the host.
//
// Field: GF(2)[X] / (X^128 + X^127 + X^126 + X^121 + 1).  Elements are kept in
// the reference's Montgomery form (value a*X^128, p3/src/polyval.rs:30-34,
// p3/src/polyval/portable.rs:20-28): 16 bytes, little-endian, low 64 bits first.
// Addition is XOR; multiplication is the POLYVAL "dot" a*b*X^-128.
//
// CDNA4 has no carry-less multiply.  The device multiply is a 4-bit-window
// comb: per lane, a 16-entry table of v*b (v a 4-bit polynomial, entries of
// 131 bits = 5 dwords) lives in LDS; the 32 nibbles of a index it, eight
// nibble positions per 32-bit word are processed together with one 256-bit
// shift by 4 between positions.  The 256-bit product is then reduced with
// the two-step Montgomery fold of the reference backends (x86_64_avx2.rs:41-68).
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

// Register-only 32x32 carry-less product: integer products of operands
// masked to every fourth bit; carries stay inside the three-bit holes.
BFD_D uint64_t clmul32_holes(uint32_t x, uint32_t y) {
    const uint32_t m0 = 0x11111111u, m1 = m0 << 1, m2 = m0 << 2, m3 = m0 << 3;
    uint32_t x0 = x & m0, x1 = x & m1, x2 = x & m2, x3 = x & m3;
    uint32_t y0 = y & m0, y1 = y & m1, y2 = y & m2, y3 = y & m3;
    uint64_t z0 = ((uint64_t)x0 * y0) ^ ((uint64_t)x1 * y3) ^ ((uint64_t)x2 * y2) ^ ((uint64_t)x3 * y1);
    uint64_t z1 = ((uint64_t)x0 * y1) ^ ((uint64_t)x1 * y0) ^ ((uint64_t)x2 * y3) ^ ((uint64_t)x3 * y2);
    uint64_t z2 = ((uint64_t)x0 * y2) ^ ((uint64_t)x1 * y1) ^ ((uint64_t)x2 * y0) ^ ((uint64_t)x3 * y3);
    uint64_t z3 = ((uint64_t)x0 * y3) ^ ((uint64_t)x1 * y2) ^ ((uint64_t)x2 * y1) ^ ((uint64_t)x3 * y0);
    const uint64_t M0 = 0x1111111111111111ull;
    return (z0 & M0) | (z1 & (M0 << 1)) | (z2 & (M0 << 2)) | (z3 & (M0 << 3));
}

BFD_D void clmul64_holes(uint64_t a, uint64_t b, uint64_t &lo, uint64_t &hi) {
    uint32_t a0 = (uint32_t)a, a1 = (uint32_t)(a >> 32), b0 = (uint32_t)b, b1 = (uint32_t)(b >> 32);
    uint64_t l = clmul32_holes(a0, b0), h = clmul32_holes(a1, b1);
    uint64_t m = clmul32_holes(a0 ^ a1, b0 ^ b1) ^ l ^ h;
    lo = l ^ (m << 32);
    hi = h ^ (m >> 32);
}

// Register-only multiply (two-level Karatsuba over clmul32_holes).
BFD_D gf gf_mul_reg(gf a, gf b) {
    uint64_t l0, l1, h0, h1, m0, m1;
    clmul64_holes(a.lo, b.lo, l0, l1);
    clmul64_holes(a.hi, b.hi, h0, h1);
    clmul64_holes(a.lo ^ a.hi, b.lo ^ b.hi, m0, m1);
    m0 ^= l0 ^ h0;
    m1 ^= l1 ^ h1;
    return gf_mont_reduce(l0, l1 ^ m0, h0 ^ m1, h1);
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
