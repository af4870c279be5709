// Host-side scalar pieces of the Basefold path: POLYVAL multiply (PCLMUL),
// inversion, Keccak-256, the Fiat-Shamir challenger, the ChaCha12 input
// generator and the additive-NTT twiddle basis.  These run once per round or
// once per setup; the bulk work is on the GPU.
#pragma once
#include <stdint.h>
#include <string.h>

#include <stdexcept>
#include <string>
#include <vector>
#include <wmmintrin.h>
#include <emmintrin.h>

#include "gf128.hpp"

namespace bfd {

// ---------------------------------------------------------------------------
// Field (p3/src/polyval/x86_64_avx2.rs:17-73 semantics: Montgomery dot)
// ---------------------------------------------------------------------------
__attribute__((target("pclmul,sse4.1"))) inline gf hmul(gf a, gf b) {
    __m128i x = _mm_set_epi64x((long long)a.hi, (long long)a.lo);
    __m128i y = _mm_set_epi64x((long long)b.hi, (long long)b.lo);
    __m128i lo = _mm_clmulepi64_si128(x, y, 0x00);
    __m128i hi = _mm_clmulepi64_si128(x, y, 0x11);
    __m128i mid = _mm_clmulepi64_si128(_mm_xor_si128(x, _mm_srli_si128(x, 8)),
                                       _mm_xor_si128(y, _mm_srli_si128(y, 8)), 0x00);
    mid = _mm_xor_si128(mid, _mm_xor_si128(lo, hi));
    uint64_t l[2], h[2], m[2];
    _mm_storeu_si128((__m128i *)l, lo);
    _mm_storeu_si128((__m128i *)h, hi);
    _mm_storeu_si128((__m128i *)m, mid);
    return gf_mont_reduce(l[0], l[1] ^ m[0], h[0] ^ m[1], h[1]);
}

inline gf hsq(gf a) { return hmul(a, a); }

// a^(2^128 - 2); zero maps to zero (Field::try_inverse semantics aside).
inline gf hinv(gf a) {
    gf r = gf_one();
    for (int bit = 127; bit >= 0; bit--) {
        r = hsq(r);
        if (bit >= 1) r = hmul(r, a);
    }
    return r;
}

// In-place inversion of every element with one field inversion
// (prefix products; all inputs must be nonzero).
inline void batch_inverse(std::vector<gf> &v) {
    if (v.empty()) return;
    std::vector<gf> pre(v.size());
    gf acc = gf_one();
    for (size_t i = 0; i < v.size(); i++) {
        pre[i] = acc;
        acc = hmul(acc, v[i]);
    }
    gf inv = hinv(acc);
    for (size_t i = v.size(); i-- > 0;) {
        gf vi = v[i];
        v[i] = hmul(inv, pre[i]);
        inv = hmul(inv, vi);
    }
}

inline gf hfrom_canonical(uint64_t lo, uint64_t hi) {
    // R^2 = X^256 mod P in canonical form (p3/src/polyval/portable.rs:10-13)
    return hmul(gf{lo, hi}, gf{0x4563df92ea7081b5ull, 0x1e563df92ea7081bull});
}

// ---------------------------------------------------------------------------
// Keccak-256 (tiny-keccak Keccak::v256: rate 136, padding 0x01 ... 0x80)
// ---------------------------------------------------------------------------
namespace hk {
static const uint64_t RC[24] = {
    0x0000000000000001ull, 0x0000000000008082ull, 0x800000000000808Aull, 0x8000000080008000ull,
    0x000000000000808Bull, 0x0000000080000001ull, 0x8000000080008081ull, 0x8000000000008009ull,
    0x000000000000008Aull, 0x0000000000000088ull, 0x0000000080008009ull, 0x000000008000000Aull,
    0x000000008000808Bull, 0x800000000000008Bull, 0x8000000000008089ull, 0x8000000000008003ull,
    0x8000000000008002ull, 0x8000000000000080ull, 0x000000000000800Aull, 0x800000008000000Aull,
    0x8000000080008081ull, 0x8000000000008080ull, 0x0000000080000001ull, 0x8000000080008008ull};
// rho offsets indexed by lane x + 5y
static const int ROT[25] = {0, 1, 62, 28, 27, 36, 44, 6, 55, 20, 3, 10, 43,
                            25, 39, 41, 45, 15, 21, 8, 18, 2, 61, 56, 14};
inline uint64_t rotl(uint64_t v, int r) { return r ? (v << r) | (v >> (64 - r)) : v; }
inline void permute(uint64_t a[25]) {
    for (int round = 0; round < 24; round++) {
        uint64_t c[5], d[5], b[25];
        for (int x = 0; x < 5; x++) c[x] = a[x] ^ a[x + 5] ^ a[x + 10] ^ a[x + 15] ^ a[x + 20];
        for (int x = 0; x < 5; x++) d[x] = c[(x + 4) % 5] ^ rotl(c[(x + 1) % 5], 1);
        for (int x = 0; x < 5; x++)
            for (int y = 0; y < 5; y++)
                b[y + 5 * ((2 * x + 3 * y) % 5)] = rotl(a[x + 5 * y] ^ d[x], ROT[x + 5 * y]);
        for (int y = 0; y < 5; y++)
            for (int x = 0; x < 5; x++)
                a[x + 5 * y] = b[x + 5 * y] ^ (~b[(x + 1) % 5 + 5 * y] & b[(x + 2) % 5 + 5 * y]);
        a[0] ^= RC[round];
    }
}
}  // namespace hk

inline void keccak256(const uint8_t *msg, size_t len, uint8_t out[32]) {
    uint64_t a[25] = {0};
    uint8_t blk[136];
    for (;;) {
        size_t take = len < 136 ? len : 136;
        memset(blk, 0, sizeof blk);
        memcpy(blk, msg, take);
        bool last = len < 136;
        if (last) {
            blk[take] ^= 0x01;
            blk[135] ^= 0x80;
        }
        for (int i = 0; i < 17; i++) {
            uint64_t w;
            memcpy(&w, blk + 8 * i, 8);
            a[i] ^= w;
        }
        hk::permute(a);
        if (last) break;
        msg += 136;
        len -= 136;
    }
    memcpy(out, a, 32);
}

// ---------------------------------------------------------------------------
// GenericChallenger<Polyval, Keccak256Hash> (p3/src/challenger.rs:24-90) over
// plonky3's HashChallenger<u8, _, 32>: observing appends bytes and drops any
// buffered output; sampling hashes the pending input when the output buffer
// is empty, keeps the digest as the next input (chaining) and hands out
// digest bytes from the last one backwards.
// ---------------------------------------------------------------------------
class Challenger {
  public:
    std::vector<uint8_t> input;
    std::vector<uint8_t> output;

    void observe(const uint8_t *p, size_t n) {
        output.clear();
        input.insert(input.end(), p, p + n);
    }
    void observe_elem(gf v) {
        uint8_t b[16];
        memcpy(b, &v.lo, 8);
        memcpy(b + 8, &v.hi, 8);
        observe(b, 16);
    }
    uint8_t sample_byte() {
        if (output.empty()) {
            uint8_t d[32];
            keccak256(input.data(), input.size(), d);
            input.assign(d, d + 32);
            output.assign(d, d + 32);
        }
        uint8_t v = output.back();
        output.pop_back();
        return v;
    }
    // Polyval::from_uniform_bytes: 16 sampled bytes, little-endian, used as
    // the raw Montgomery value (p3/src/polyval.rs:71-77).
    gf sample() {
        uint8_t b[16];
        for (int i = 0; i < 16; i++) b[i] = sample_byte();
        gf v;
        memcpy(&v.lo, b, 8);
        memcpy(&v.hi, b + 8, 8);
        return v;
    }
    uint64_t sample_bits(int bits) {
        uint8_t b[8];
        for (int i = 0; i < 8; i++) b[i] = sample_byte();
        uint64_t v;
        memcpy(&v, b, 8);
        return bits >= 64 ? v : (v & ((1ull << bits) - 1));
    }
};

// ---------------------------------------------------------------------------
// Polyval::random_vec(n, StdRng::seed_from_u64(seed)): PCG32 expansion of
// the seed into a ChaCha12 key, 64-bit block counter, zero stream; each
// element takes the next 16 bytes of keystream (from_uniform_bytes.rs:5-27).
// ---------------------------------------------------------------------------
inline void chacha_block(int rounds, const uint32_t key[8], uint64_t counter, uint32_t out[16]) {
    uint32_t s[16] = {0x61707865u, 0x3320646eu, 0x79622d32u, 0x6b206574u};
    memcpy(s + 4, key, 32);
    s[12] = (uint32_t)counter;
    s[13] = (uint32_t)(counter >> 32);
    s[14] = 0;
    s[15] = 0;
    uint32_t x[16];
    memcpy(x, s, sizeof x);
    auto rl = [](uint32_t v, int r) { return (v << r) | (v >> (32 - r)); };
    auto qr = [&](int a, int b, int c, int d) {
        x[a] += x[b]; x[d] ^= x[a]; x[d] = rl(x[d], 16);
        x[c] += x[d]; x[b] ^= x[c]; x[b] = rl(x[b], 12);
        x[a] += x[b]; x[d] ^= x[a]; x[d] = rl(x[d], 8);
        x[c] += x[d]; x[b] ^= x[c]; x[b] = rl(x[b], 7);
    };
    for (int i = 0; i < rounds; i += 2) {
        qr(0, 4, 8, 12); qr(1, 5, 9, 13); qr(2, 6, 10, 14); qr(3, 7, 11, 15);
        qr(0, 5, 10, 15); qr(1, 6, 11, 12); qr(2, 7, 8, 13); qr(3, 4, 9, 14);
    }
    for (int i = 0; i < 16; i++) out[i] = x[i] + s[i];
}

inline void seed_key(uint64_t seed, uint32_t key[8]) {
    uint64_t st = seed;
    for (int i = 0; i < 8; i++) {
        st = st * 6364136223846793005ull + 11634580456548032391ull;
        uint32_t xs = (uint32_t)(((st >> 18) ^ st) >> 27);
        uint32_t rot = (uint32_t)(st >> 59);
        key[i] = (xs >> rot) | (xs << ((32 - rot) & 31));
    }
}

// ---------------------------------------------------------------------------
// Additive-NTT twiddle basis (p3/src/field/binary.rs:23-53).
//
// The reference's layer K (2^K entries) of AddtiveNtt::new(L) is
//     Tab[K][k] = XOR over set bits q of k of  W_i(beta_{i+1+q}) / W_i(beta_i),
// with i = L-1-K, beta_j = X^j and W_{i+1}(x) = W_i(x) (W_i(x) + W_i(beta_i)).
// basis[K][q] = Tab[K][1 << q]; every twiddle is an XOR of basis entries.
// ---------------------------------------------------------------------------
inline std::vector<std::vector<gf>> ntt_basis(int L) {
    std::vector<std::vector<gf>> basis(L > 0 ? L : 0);
    if (L <= 0) return basis;
    std::vector<gf> ws;
    gf wbi = hfrom_canonical(1, 0);
    for (int j = 1; j < L; j++)
        ws.push_back(j < 64 ? hfrom_canonical(1ull << j, 0) : hfrom_canonical(0, 1ull << (j - 64)));
    for (int i = 0; i < L; i++) {
        int K = L - 1 - i;
        gf inv = hinv(wbi);
        basis[K].resize(ws.size());
        for (size_t q = 0; q < ws.size(); q++) basis[K][q] = hmul(ws[q], inv);
        if (ws.empty()) break;
        gf nb = gf_add(hsq(ws[0]), hmul(ws[0], wbi));
        std::vector<gf> nws;
        for (size_t q = 1; q < ws.size(); q++) nws.push_back(gf_add(hsq(ws[q]), hmul(ws[q], wbi)));
        ws.swap(nws);
        wbi = nb;
    }
    return basis;
}

inline size_t bitrev(size_t v, int bits) {
    size_t r = 0;
    for (int i = 0; i < bits; i++) r |= ((v >> i) & 1) << (bits - 1 - i);
    return r;
}

}  // namespace bfd
