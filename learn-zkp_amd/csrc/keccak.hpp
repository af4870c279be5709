// Keccak-f[1600] and Keccak-256 (original padding 0x01, rate 136 bytes) for
// CDNA4.  This is the hash behind the reference's Merkle commitments and
// transcript: p3/src/merkle_tree.rs:8-21 (leaf = Keccak256(row bytes), node =
// Keccak256(left || right)) over tiny-keccak's Keccak::v256.
//
// gfx950 specifics: every 64-bit lane is a pair of 32-bit halves; rotations
// are two v_alignbit_b32; the five-way column parity of theta is two
// three-input XORs and chi (a ^ (~b & c)) is one three-input LUT op, both via
// v_bitop3_b32, and so is theta's A ^ C[x-1] ^ rotl(C[x+1], 1) (D is never
// formed).  A round is ~174 VALU instructions; the 24 rounds are fully
// unrolled so round constants become immediates and the zero padding lanes
// of one-block messages fold away in the first round.
#pragma once
#include <stdint.h>
#include <hip/hip_runtime.h>

namespace kk {

__device__ __constant__ static const uint32_t RC_LO[24] = {
    0x00000001u, 0x00008082u, 0x0000808Au, 0x80008000u, 0x0000808Bu, 0x80000001u,
    0x80008081u, 0x00008009u, 0x0000008Au, 0x00000088u, 0x80008009u, 0x8000000Au,
    0x8000808Bu, 0x0000008Bu, 0x00008089u, 0x00008003u, 0x00008002u, 0x00000080u,
    0x0000800Au, 0x8000000Au, 0x80008081u, 0x00008080u, 0x80000001u, 0x80008008u};
__device__ __constant__ static const uint32_t RC_HI[24] = {
    0x00000000u, 0x00000000u, 0x80000000u, 0x80000000u, 0x00000000u, 0x00000000u,
    0x80000000u, 0x80000000u, 0x00000000u, 0x00000000u, 0x00000000u, 0x00000000u,
    0x00000000u, 0x80000000u, 0x80000000u, 0x80000000u, 0x80000000u, 0x80000000u,
    0x00000000u, 0x80000000u, 0x80000000u, 0x80000000u, 0x00000000u, 0x80000000u};

// v_bitop3_b32 truth tables: bit (a<<2 | b<<1 | c) of the immediate is
// f(a, b, c) for source operands (a, b, c); 0xF0/0xCC/0xAA select a/b/c.
constexpr uint32_t LUT_XOR3 = (0xF0 ^ 0xCC ^ 0xAA) & 0xFF;          // a ^ b ^ c
constexpr uint32_t LUT_CHI = (0xF0 ^ (~0xCC & 0xAA)) & 0xFF;        // a ^ (~b & c)

__device__ __forceinline__ uint32_t xor3(uint32_t a, uint32_t b, uint32_t c) {
    return __builtin_amdgcn_bitop3_b32(a, b, c, LUT_XOR3);
}
__device__ __forceinline__ uint32_t chi(uint32_t a, uint32_t b, uint32_t c) {
    return __builtin_amdgcn_bitop3_b32(a, b, c, LUT_CHI);
}

struct lane {
    uint32_t lo, hi;
};

// rotate left by a compile-time constant
template <int R>
__device__ __forceinline__ lane rotl(lane v) {
    if constexpr (R == 0) {
        return v;
    } else if constexpr (R == 32) {
        return lane{v.hi, v.lo};
    } else if constexpr (R < 32) {
        return lane{__builtin_amdgcn_alignbit(v.lo, v.hi, 32 - R), __builtin_amdgcn_alignbit(v.hi, v.lo, 32 - R)};
    } else {
        return lane{__builtin_amdgcn_alignbit(v.hi, v.lo, 64 - R), __builtin_amdgcn_alignbit(v.lo, v.hi, 64 - R)};
    }
}

__device__ __forceinline__ lane lx(lane a, lane b) { return lane{a.lo ^ b.lo, a.hi ^ b.hi}; }

// one permutation on lanes s[x + 5y].  SPLIT_D forms D[x] first and XORs it
// into each lane with a second operation (ten more instructions per round);
// the segmented hashing kernels measure faster with it (their generated code
// issues more VALU per cycle with that form, about 1.07 against 0.94 VALU
// instructions per CU per cycle), the other kernels with the merged form.
template <bool SPLIT_D = false>
__device__ __forceinline__ void keccak_f(lane s[25]) {
#pragma unroll
    for (int round = 0; round < 24; round++) {
        lane c[5];
#pragma unroll
        for (int x = 0; x < 5; x++) {
            c[x].lo = xor3(xor3(s[x].lo, s[x + 5].lo, s[x + 10].lo), s[x + 15].lo, s[x + 20].lo);
            c[x].hi = xor3(xor3(s[x].hi, s[x + 5].hi, s[x + 10].hi), s[x + 15].hi, s[x + 20].hi);
        }
        // D[x] = C[x-1] ^ rotl(C[x+1], 1) is never formed: each state lane
        // takes both terms in one three-input XOR (one operation per half
        // instead of two)
        lane rc[5], d[5];
#pragma unroll
        for (int x = 0; x < 5; x++) rc[x] = rotl<1>(c[x]);
#pragma unroll
        for (int x = 0; x < 5; x++) d[x] = lx(c[(x + 4) % 5], rc[(x + 1) % 5]);
        const auto td = [&](int i) {
            const int x = i % 5;
            if constexpr (SPLIT_D) return lx(s[i], d[x]);
            return lane{xor3(s[i].lo, c[(x + 4) % 5].lo, rc[(x + 1) % 5].lo),
                        xor3(s[i].hi, c[(x + 4) % 5].hi, rc[(x + 1) % 5].hi)};
        };
        // theta + rho + pi: B[y + 5*((2x+3y)%5)] = rotl(A[x+5y] ^ D[x], r[x][y])
        lane b[25];
        b[0] = td(0);
        b[10] = rotl<1>(td(1));
        b[20] = rotl<62>(td(2));
        b[5] = rotl<28>(td(3));
        b[15] = rotl<27>(td(4));
        b[16] = rotl<36>(td(5));
        b[1] = rotl<44>(td(6));
        b[11] = rotl<6>(td(7));
        b[21] = rotl<55>(td(8));
        b[6] = rotl<20>(td(9));
        b[7] = rotl<3>(td(10));
        b[17] = rotl<10>(td(11));
        b[2] = rotl<43>(td(12));
        b[12] = rotl<25>(td(13));
        b[22] = rotl<39>(td(14));
        b[23] = rotl<41>(td(15));
        b[8] = rotl<45>(td(16));
        b[18] = rotl<15>(td(17));
        b[3] = rotl<21>(td(18));
        b[13] = rotl<8>(td(19));
        b[14] = rotl<18>(td(20));
        b[24] = rotl<2>(td(21));
        b[9] = rotl<61>(td(22));
        b[19] = rotl<56>(td(23));
        b[4] = rotl<14>(td(24));
        // chi: A[x+5y] = B[x+5y] ^ (~B[x+1+5y] & B[x+2+5y])
#pragma unroll
        for (int y = 0; y < 25; y += 5)
#pragma unroll
            for (int x = 0; x < 5; x++) {
                s[y + x].lo = chi(b[y + x].lo, b[y + (x + 1) % 5].lo, b[y + (x + 2) % 5].lo);
                s[y + x].hi = chi(b[y + x].hi, b[y + (x + 1) % 5].hi, b[y + (x + 2) % 5].hi);
            }
        // iota
        s[0].lo ^= RC_LO[round];
        s[0].hi ^= RC_HI[round];
    }
}

// ---------------------------------------------------------------------------
// One permutation on a lane pair, for latency-bound hashing (the narrow top
// levels of a Merkle tree, the transcript's one sample per round).  The even
// lane of the pair (lanes 2t, 2t+1 of a wave, both active) holds the low
// 32-bit halves of all 25 state lanes, the odd lane the high halves.  Theta's
// parities, the XOR with D and chi only combine words of the same half.  A
// 64-bit rotation by R needs both halves of its word: each lane takes its
// partner's word with one DPP move (quad_perm [1,0,3,2]) and forms its own
// half with one v_alignbit_b32, the same instruction on both lanes
// (R < 32: alignbit(own, partner, 32 - R), the low half of the even lane and
// the high half of the odd one; R > 32: alignbit(partner, own, 64 - R)).
// Per round and lane: 10 XOR3 (parities), 5 moves + 5 alignbits + 5 XORs (D),
// 25 XORs (A ^ D), 24 moves + 24 alignbits (rho) and 25 chi, about 123
// instructions against 174 for the one-lane permutation, so a pair finishes a
// permutation in about two thirds of the time at 1.4 times the instructions.
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint32_t pair_swap(uint32_t v) {
    // bound_ctrl set: no lane of the pair is ever out of range, and the
    // destination's old value need not be materialised
    return (uint32_t)__builtin_amdgcn_mov_dpp((int)v, 0xB1, 0xF, 0xF, true);
}

template <int R>
__device__ __forceinline__ uint32_t rotl_half(uint32_t w) {
    if constexpr (R == 0) {
        return w;
    } else if constexpr (R == 32) {
        return pair_swap(w);
    } else if constexpr (R < 32) {
        return __builtin_amdgcn_alignbit(w, pair_swap(w), 32 - R);
    } else {
        return __builtin_amdgcn_alignbit(pair_swap(w), w, 64 - R);
    }
}

// s[x + 5y]: this lane's half of state lane (x, y); hm = ~0 on the odd lane
// of the pair, 0 on the even lane.  D is formed first and XORed into each
// lane with a second operation: a single wave finishes the permutation in
// 5.2 us that way against 5.55 us with the one-lane permutation's merged
// three-input form, which has five instructions fewer per round.
__device__ __forceinline__ void keccak_f_pair(uint32_t s[25], uint32_t hm) {
#pragma unroll
    for (int round = 0; round < 24; round++) {
        uint32_t c[5];
#pragma unroll
        for (int x = 0; x < 5; x++) c[x] = xor3(xor3(s[x], s[x + 5], s[x + 10]), s[x + 15], s[x + 20]);
        uint32_t d[5];
#pragma unroll
        for (int x = 0; x < 5; x++) d[x] = c[(x + 4) % 5] ^ rotl_half<1>(c[(x + 1) % 5]);
        const auto td = [&](int i) { return s[i] ^ d[i % 5]; };
        // theta + rho + pi, the one-lane permutation's table
        uint32_t b[25];
        b[0] = td(0);
        b[10] = rotl_half<1>(td(1));
        b[20] = rotl_half<62>(td(2));
        b[5] = rotl_half<28>(td(3));
        b[15] = rotl_half<27>(td(4));
        b[16] = rotl_half<36>(td(5));
        b[1] = rotl_half<44>(td(6));
        b[11] = rotl_half<6>(td(7));
        b[21] = rotl_half<55>(td(8));
        b[6] = rotl_half<20>(td(9));
        b[7] = rotl_half<3>(td(10));
        b[17] = rotl_half<10>(td(11));
        b[2] = rotl_half<43>(td(12));
        b[12] = rotl_half<25>(td(13));
        b[22] = rotl_half<39>(td(14));
        b[23] = rotl_half<41>(td(15));
        b[8] = rotl_half<45>(td(16));
        b[18] = rotl_half<15>(td(17));
        b[3] = rotl_half<21>(td(18));
        b[13] = rotl_half<8>(td(19));
        b[14] = rotl_half<18>(td(20));
        b[24] = rotl_half<2>(td(21));
        b[9] = rotl_half<61>(td(22));
        b[19] = rotl_half<56>(td(23));
        b[4] = rotl_half<14>(td(24));
#pragma unroll
        for (int y = 0; y < 25; y += 5)
#pragma unroll
            for (int x = 0; x < 5; x++) s[y + x] = chi(b[y + x], b[y + (x + 1) % 5], b[y + (x + 2) % 5]);
        s[0] ^= (RC_LO[round] & ~hm) | (RC_HI[round] & hm);
    }
}

// keccak256_words (below) on a lane pair: both lanes pass the same message
// words and get the whole digest.  odd = lane index & 1.
template <int NW>
__device__ __forceinline__ void keccak256_words_pair(const uint64_t *m, uint64_t out[4], bool odd) {
    const uint32_t hm = odd ? ~0u : 0u;
    uint32_t s[25];
#pragma unroll
    for (int i = 0; i < 25; i++) {
        uint64_t v = i < NW ? m[i] : 0ull;
        if (i == NW) v ^= 0x01ull;
        if (i == 16) v ^= 0x8000000000000000ull;
        s[i] = ((uint32_t)v & ~hm) | ((uint32_t)(v >> 32) & hm);
    }
    keccak_f_pair(s, hm);
#pragma unroll
    for (int i = 0; i < 4; i++) {
        const uint32_t p = pair_swap(s[i]);
        const uint32_t lo = (s[i] & ~hm) | (p & hm), hi = (p & ~hm) | (s[i] & hm);
        out[i] = (uint64_t)lo | ((uint64_t)hi << 32);
    }
}

// Keccak-256 of a message of NW 64-bit words (NW*8 < 136): one block.
// Writes the 4-word digest.
template <int NW, bool SPLIT_D = false>
__device__ __forceinline__ void keccak256_words(const uint64_t *m, uint64_t out[4]) {
    lane s[25];
#pragma unroll
    for (int i = 0; i < 25; i++) {
        uint64_t v = i < NW ? m[i] : 0ull;
        if (i == NW) v ^= 0x01ull;                    // Keccak padding start
        if (i == 16) v ^= 0x8000000000000000ull;      // last bit of the 136-byte rate
        s[i] = lane{(uint32_t)v, (uint32_t)(v >> 32)};
    }
    keccak_f<SPLIT_D>(s);
#pragma unroll
    for (int i = 0; i < 4; i++) out[i] = (uint64_t)s[i].lo | ((uint64_t)s[i].hi << 32);
}

// Keccak-256 of nw 64-bit words (8-byte aligned, runtime length): full
// 17-word blocks, then the last block with the remaining words and the
// padding (0x01 after the message, 0x80 in the last rate byte)
__device__ __forceinline__ void keccak256_var(const uint64_t *m, size_t nw, uint64_t out[4]) {
    const size_t full = nw / 17, rem = nw % 17;
    lane s[25];
#pragma unroll
    for (int k = 0; k < 25; k++) s[k] = lane{0u, 0u};
    for (size_t blk = 0; blk < full; blk++, m += 17) {
#pragma unroll
        for (int k = 0; k < 17; k++) {
            const uint64_t v = m[k];
            s[k].lo ^= (uint32_t)v;
            s[k].hi ^= (uint32_t)(v >> 32);
        }
        keccak_f(s);
    }
#pragma unroll
    for (int k = 0; k < 17; k++) {
        uint64_t v = (size_t)k < rem ? m[k] : 0ull;
        if ((size_t)k == rem) v ^= 0x01ull;
        if (k == 16) v ^= 0x8000000000000000ull;
        s[k].lo ^= (uint32_t)v;
        s[k].hi ^= (uint32_t)(v >> 32);
    }
    keccak_f(s);
#pragma unroll
    for (int k = 0; k < 4; k++) out[k] = (uint64_t)s[k].lo | ((uint64_t)s[k].hi << 32);
}

// Keccak-256 of len bytes at m, for the one-lane transcript sponge.  m is
// 8-byte aligned and readable up to len rounded up to a multiple of 8 (the
// bytes past len are masked off); the padding (0x01 after the message, 0x80
// in the last rate byte) goes into the block that holds the end of the
// message, a block of its own when len is a multiple of 136.
__device__ inline void keccak256_bytes(const uint8_t *m, uint32_t len, uint64_t out[4]) {
    const uint64_t *m64 = reinterpret_cast<const uint64_t *>(m);
    lane s[25];
#pragma unroll
    for (int k = 0; k < 25; k++) s[k] = lane{0u, 0u};
    for (uint32_t off = 0;; off += 136) {
        const uint32_t rest = len - off;
        const bool last = rest < 136;
#pragma unroll
        for (int k = 0; k < 17; k++) {
            const uint32_t wo = 8 * k;
            uint64_t v = 0;
            if (wo < rest) {
                v = m64[(off + wo) >> 3];
                if (rest - wo < 8) v &= (1ull << (8 * (rest - wo))) - 1;
            }
            if (last && (rest >> 3) == (uint32_t)k) v ^= 1ull << (8 * (rest & 7));
            if (last && k == 16) v ^= 0x8000000000000000ull;
            s[k].lo ^= (uint32_t)v;
            s[k].hi ^= (uint32_t)(v >> 32);
        }
        keccak_f(s);
        if (last) break;
    }
#pragma unroll
    for (int k = 0; k < 4; k++) out[k] = (uint64_t)s[k].lo | ((uint64_t)s[k].hi << 32);
}

// permutation on 64-bit words (used by the micro-benchmark)
__device__ __forceinline__ void keccak_f(uint64_t w[25]) {
    lane s[25];
#pragma unroll
    for (int i = 0; i < 25; i++) s[i] = lane{(uint32_t)w[i], (uint32_t)(w[i] >> 32)};
    keccak_f(s);
#pragma unroll
    for (int i = 0; i < 25; i++) w[i] = (uint64_t)s[i].lo | ((uint64_t)s[i].hi << 32);
}

}  // namespace kk
