/*
 * bfo.c — CPU restatement of the reference's Basefold prover path over the
 * POLYVAL field (han0110/learn-zkp).  TEST INFRASTRUCTURE: the checker for
 * the HIP library's parity tests and the CPU-baseline leg of bench.py.  The
 * product library never links or calls anything here.
 *
 * Citations are reference paths relative to the reference repository root.
 * How each part is pinned (independent of this file's own formulas):
 *   field mul / from_canonical : the 3,072-entry repeated-squaring table held
 *                                in p3/src/polyval/portable.rs:169-376, plus
 *                                three independent multiply algorithms that
 *                                must agree (portable, PCLMUL, bit-serial)
 *                                and the canonical-product identity.
 *   inverse                    : a * inv(a) == ONE on random inputs.
 *   Keccak-f[1600]             : SHA3-256 vs Python hashlib; Keccak-256("").
 *   ChaCha core                : RFC 8439 section 2.3.2 block vector.
 *   additive NTT               : brute-force definition (binary.rs:156-191),
 *                                forward/backward round trip, closed form of
 *                                the last twiddle layer.
 *   encode                     : closed form codeword[v] = P(point bitrev(v)).
 *   fold + sumcheck            : folding invariant (basefold.rs:198), per-round
 *                                p(0)+p(1) = claim, final-claim identity.
 *   Merkle / challenger /      : PARITY UNPINNED — these follow plonky3 at
 *   bincode / seed expansion     rev 784b7dd, which is not vendored; the proofs
 *                                are accepted by this file's verifier, which
 *                                is a restatement of basefold.rs:234-325.
 */
#include "bfo.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <wmmintrin.h>
#include <emmintrin.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef unsigned __int128 u128;
typedef uint64_t u64;

static int g_threads = 1;

void bfo_set_threads(int n) {
    g_threads = n < 1 ? 1 : n;
#ifdef _OPENMP
    omp_set_num_threads(g_threads);
#endif
}
int bfo_get_threads(void) { return g_threads; }

static inline u128 ld(const uint8_t *p) { u128 v; memcpy(&v, p, 16); return v; }
static inline void st(uint8_t *p, u128 v) { memcpy(p, &v, 16); }

/* ===================================================================== */
/* Field: POLYVAL, GF(2)[X]/(X^128+X^127+X^126+X^121+1), Montgomery R=X^128 */
/* ===================================================================== */

/* portable.rs:97-104 — 64-bit bit reversal */
static inline u64 rev64(u64 x) {
    x = ((x & 0x5555555555555555ull) << 1) | ((x >> 1) & 0x5555555555555555ull);
    x = ((x & 0x3333333333333333ull) << 2) | ((x >> 2) & 0x3333333333333333ull);
    x = ((x & 0x0f0f0f0f0f0f0f0full) << 4) | ((x >> 4) & 0x0f0f0f0f0f0f0f0full);
    x = ((x & 0x00ff00ff00ff00ffull) << 8) | ((x >> 8) & 0x00ff00ff00ff00ffull);
    x = ((x & 0x0000ffff0000ffffull) << 16) | ((x >> 16) & 0x0000ffff0000ffffull);
    return (x >> 32) | (x << 32);
}

/* portable.rs:69-95 — low 64 bits of a carry-less product, integer multiplies
 * on operands with three-bit holes so carries never reach a data bit. */
static inline u64 bmul64(u64 x, u64 y) {
    const u64 m0 = 0x1111111111111111ull, m1 = m0 << 1, m2 = m0 << 2, m3 = m0 << 3;
    u64 x0 = x & m0, x1 = x & m1, x2 = x & m2, x3 = x & m3;
    u64 y0 = y & m0, y1 = y & m1, y2 = y & m2, y3 = y & m3;
    u64 z0 = (x0 * y0) ^ (x1 * y3) ^ (x2 * y2) ^ (x3 * y1);
    u64 z1 = (x0 * y1) ^ (x1 * y0) ^ (x2 * y3) ^ (x3 * y2);
    u64 z2 = (x0 * y2) ^ (x1 * y1) ^ (x2 * y0) ^ (x3 * y3);
    u64 z3 = (x0 * y3) ^ (x1 * y2) ^ (x2 * y1) ^ (x3 * y0);
    return (z0 & m0) | (z1 & m1) | (z2 & m2) | (z3 & m3);
}

/* Two-step Montgomery reduction of the 256-bit product v0..v3 (64-bit words),
 * portable.rs:59-64 / x86_64_avx2.rs:41-68. */
static inline u128 mont_reduce(u64 v0, u64 v1, u64 v2, u64 v3) {
    v2 ^= v0 ^ (v0 >> 1) ^ (v0 >> 2) ^ (v0 >> 7);
    v1 ^= (v0 << 63) ^ (v0 << 62) ^ (v0 << 57);
    v3 ^= v1 ^ (v1 >> 1) ^ (v1 >> 2) ^ (v1 >> 7);
    v2 ^= (v1 << 63) ^ (v1 << 62) ^ (v1 << 57);
    return (u128)v2 | ((u128)v3 << 64);
}

/* portable.rs:27-67 — Karatsuba over bmul64, high halves via bit reversal */
static u128 mul_portable(u128 a, u128 b) {
    u64 h0 = (u64)a, h1 = (u64)(a >> 64), y0 = (u64)b, y1 = (u64)(b >> 64);
    u64 h0r = rev64(h0), h1r = rev64(h1), y0r = rev64(y0), y1r = rev64(y1);
    u64 h2 = h0 ^ h1, h2r = h0r ^ h1r, y2 = y0 ^ y1, y2r = y0r ^ y1r;
    u64 z0 = bmul64(y0, h0), z1 = bmul64(y1, h1), z2 = bmul64(y2, h2);
    u64 z0h = bmul64(y0r, h0r), z1h = bmul64(y1r, h1r), z2h = bmul64(y2r, h2r);
    z2 ^= z0 ^ z1;
    z2h ^= z0h ^ z1h;
    z0h = rev64(z0h) >> 1;
    z1h = rev64(z1h) >> 1;
    z2h = rev64(z2h) >> 1;
    return mont_reduce(z0, z0h ^ z2, z1 ^ z2h, z1h);
}

/* x86_64_avx2.rs:17-73 — PCLMULQDQ Karatsuba, same reduction */
__attribute__((target("pclmul,sse4.1")))
static u128 mul_pclmul(u128 a, u128 b) {
    __m128i h = _mm_set_epi64x((long long)(u64)(a >> 64), (long long)(u64)a);
    __m128i y = _mm_set_epi64x((long long)(u64)(b >> 64), (long long)(u64)b);
    __m128i t0 = _mm_clmulepi64_si128(y, h, 0x00);
    __m128i t1 = _mm_clmulepi64_si128(y, h, 0x11);
    __m128i t2 = _mm_clmulepi64_si128(_mm_xor_si128(y, _mm_shuffle_epi32(y, 0x0E)),
                                      _mm_xor_si128(h, _mm_shuffle_epi32(h, 0x0E)), 0x00);
    t2 = _mm_xor_si128(t2, _mm_xor_si128(t0, t1));
    u64 w0[2], w1[2], w2[2];
    _mm_storeu_si128((__m128i *)w0, t0);
    _mm_storeu_si128((__m128i *)w1, t1);
    _mm_storeu_si128((__m128i *)w2, t2);
    return mont_reduce(w0[0], w0[1] ^ w2[0], w1[0] ^ w2[1], w1[1]);
}

/* Textbook definition of the POLYVAL dot: bit-serial 256-bit product, then
 * 128 halvings with conditional addition of the modulus (a*b*X^-128 mod P). */
static const u128 P_LOW = ((u128)1 << 127) | ((u128)1 << 126) | ((u128)1 << 121) | 1;
static u128 mul_textbook(u128 a, u128 b) {
    u128 lo = 0, hi = 0;
    for (int i = 0; i < 128; i++)
        if ((b >> i) & 1) {
            lo ^= a << i;
            if (i) hi ^= a >> (128 - i);
        }
    for (int i = 0; i < 128; i++) {
        if (lo & 1) { lo ^= P_LOW; hi ^= 1; }
        lo = (lo >> 1) | (hi << 127);
        hi >>= 1;
    }
    return lo;
}

/* plain canonical product mod P (no Montgomery factor) */
static u128 polymul_mod(u128 a, u128 b) {
    u128 lo = 0, hi = 0;
    for (int i = 0; i < 128; i++)
        if ((b >> i) & 1) {
            lo ^= a << i;
            if (i) hi ^= a >> (128 - i);
        }
    for (int bit = 127; bit >= 0; bit--)
        if ((hi >> bit) & 1) {
            /* X^(128+bit) = X^bit * (X^127 + X^126 + X^121 + 1) */
            hi ^= (u128)1 << bit;
            u128 t = P_LOW;  /* multiply by X^bit */
            lo ^= t << bit;
            if (bit) hi ^= t >> (128 - bit);
        }
    return lo;
}

static int g_have_pclmul = -1;
static inline u128 fmul(u128 a, u128 b) {
    if (g_have_pclmul < 0) g_have_pclmul = __builtin_cpu_supports("pclmul") ? 1 : 0;
    return g_have_pclmul ? mul_pclmul(a, b) : mul_portable(a, b);
}

static const u128 R2 = ((u128)0x1e563df92ea7081bull << 64) | 0x4563df92ea7081b5ull;
static inline u128 from_canonical(u128 a) { return fmul(a, R2); }   /* portable.rs:10-13 */
static inline u128 to_canonical(u128 a) { return fmul(a, 1); }      /* portable.rs:15-18 */
#define ONE (((u128)0xc200000000000000ull << 64) | 1)               /* = from_canonical(1) */

/* a^(2^128 - 2) by square-and-multiply; the value equals invert_or_zero
 * (portable.rs:113-145), which is a different addition chain for the same power. */
static u128 finv(u128 a) {
    u128 r = ONE;
    for (int bit = 127; bit >= 0; bit--) {
        r = fmul(r, r);
        if (bit >= 1) r = fmul(r, a);
    }
    return r;
}

void bfo_mul_batch(const uint8_t *a, const uint8_t *b, uint8_t *out, size_t n, int impl) {
    for (size_t i = 0; i < n; i++) {
        u128 x = ld(a + 16 * i), y = ld(b + 16 * i), z;
        if (impl == 0) z = mul_portable(x, y);
        else if (impl == 1) z = mul_pclmul(x, y);
        else z = mul_textbook(x, y);
        st(out + 16 * i, z);
    }
}
void bfo_from_canonical_batch(const uint8_t *a, uint8_t *out, size_t n) {
    for (size_t i = 0; i < n; i++) st(out + 16 * i, from_canonical(ld(a + 16 * i)));
}
void bfo_to_canonical_batch(const uint8_t *a, uint8_t *out, size_t n) {
    for (size_t i = 0; i < n; i++) st(out + 16 * i, to_canonical(ld(a + 16 * i)));
}
void bfo_inv_batch(const uint8_t *a, uint8_t *out, size_t n) {
    for (size_t i = 0; i < n; i++) st(out + 16 * i, finv(ld(a + 16 * i)));
}
void bfo_polymul_mod_batch(const uint8_t *a, const uint8_t *b, uint8_t *out, size_t n) {
    for (size_t i = 0; i < n; i++) st(out + 16 * i, polymul_mod(ld(a + 16 * i), ld(b + 16 * i)));
}

/* ===================================================================== */
/* Keccak-f[1600] and the sponge (FIPS 202); constants derived from the spec */
/* ===================================================================== */
static u64 KRC[24];
static int KROT[25];
static int g_keccak_ready;

static void keccak_init(void) {
    if (g_keccak_ready) return;
    uint8_t lfsr = 1;
    for (int r = 0; r < 24; r++) {
        u64 rc = 0;
        for (int j = 0; j < 7; j++) {
            int out = lfsr & 1;
            lfsr = (lfsr & 0x80) ? (uint8_t)((lfsr << 1) ^ 0x71) : (uint8_t)(lfsr << 1);
            if (out) rc |= 1ull << ((1 << j) - 1);
        }
        KRC[r] = rc;
    }
    KROT[0] = 0;
    int x = 1, y = 0;
    for (int t = 0; t < 24; t++) {
        KROT[x + 5 * y] = ((t + 1) * (t + 2) / 2) % 64;
        int nx = y, ny = (2 * x + 3 * y) % 5;
        x = nx; y = ny;
    }
    g_keccak_ready = 1;
}

static inline u64 rotl64(u64 v, int r) { return r ? (v << r) | (v >> (64 - r)) : v; }

static void keccak_f(u64 A[25]) {
    for (int round = 0; round < 24; round++) {
        u64 C[5], D[5], B[25];
        for (int x = 0; x < 5; x++) C[x] = A[x] ^ A[x + 5] ^ A[x + 10] ^ A[x + 15] ^ A[x + 20];
        for (int x = 0; x < 5; x++) D[x] = C[(x + 4) % 5] ^ rotl64(C[(x + 1) % 5], 1);
        for (int i = 0; i < 25; i++) A[i] ^= D[i % 5];
        for (int x = 0; x < 5; x++)
            for (int y = 0; y < 5; y++)
                B[y + 5 * ((2 * x + 3 * y) % 5)] = rotl64(A[x + 5 * y], KROT[x + 5 * y]);
        for (int x = 0; x < 5; x++)
            for (int y = 0; y < 5; y++)
                A[x + 5 * y] = B[x + 5 * y] ^ (~B[(x + 1) % 5 + 5 * y] & B[(x + 2) % 5 + 5 * y]);
        A[0] ^= KRC[round];
    }
}

static void sponge256(const uint8_t *msg, size_t len, uint8_t domain, uint8_t out[32]) {
    keccak_init();
    const size_t rate = 136;
    u64 A[25];
    memset(A, 0, sizeof A);
    uint8_t block[136];
    for (;;) {
        size_t take = len < rate ? len : rate;
        memset(block, 0, rate);
        memcpy(block, msg, take);
        int last = len < rate;
        if (last) {
            block[take] ^= domain;
            block[rate - 1] ^= 0x80;
        }
        for (size_t i = 0; i < rate / 8; i++) {
            u64 w;
            memcpy(&w, block + 8 * i, 8);
            A[i] ^= w;
        }
        keccak_f(A);
        if (last) break;
        msg += rate;
        len -= rate;
    }
    memcpy(out, A, 32);
}

/* tiny-keccak Keccak::v256 (pad 0x01), the hash behind Keccak256Hash */
void bfo_keccak256(const uint8_t *msg, size_t len, uint8_t out[32]) { sponge256(msg, len, 0x01, out); }
void bfo_sha3_256(const uint8_t *msg, size_t len, uint8_t out[32]) { sponge256(msg, len, 0x06, out); }

/* ===================================================================== */
/* ChaCha (RFC 8439 core) and StdRng::seed_from_u64 (rand_core 0.6.4,      */
/* rand_chacha 0.3.1: ChaCha12, 64-bit block counter in words 12-13).      */
/* ===================================================================== */
static inline uint32_t rotl32(uint32_t v, int r) { return (v << r) | (v >> (32 - r)); }
#define QR(a, b, c, d)                      \
    a += b; d ^= a; d = rotl32(d, 16);      \
    c += d; b ^= c; b = rotl32(b, 12);      \
    a += b; d ^= a; d = rotl32(d, 8);       \
    c += d; b ^= c; b = rotl32(b, 7);

void bfo_chacha_block(int rounds, const uint32_t key[8], const uint32_t ctr_nonce[4], uint32_t out[16]) {
    uint32_t s[16] = {0x61707865, 0x3320646e, 0x79622d32, 0x6b206574};
    memcpy(s + 4, key, 32);
    memcpy(s + 12, ctr_nonce, 16);
    uint32_t x[16];
    memcpy(x, s, sizeof x);
    for (int i = 0; i < rounds; i += 2) {
        QR(x[0], x[4], x[8], x[12]) QR(x[1], x[5], x[9], x[13])
        QR(x[2], x[6], x[10], x[14]) QR(x[3], x[7], x[11], x[15])
        QR(x[0], x[5], x[10], x[15]) QR(x[1], x[6], x[11], x[12])
        QR(x[2], x[7], x[8], x[13]) QR(x[3], x[4], x[9], x[14])
    }
    for (int i = 0; i < 16; i++) out[i] = x[i] + s[i];
}

/* Polyval::random_vec(n, StdRng::seed_from_u64(seed)): each element is 16
 * bytes of the ChaCha12 stream (from_uniform_bytes.rs:5-27, polyval.rs:71-77). */
void bfo_random_vec(uint64_t seed, size_t n, uint8_t *out) {
    uint32_t key[8];
    u64 state = seed;
    for (int i = 0; i < 8; i++) { /* PCG32 expansion of the u64 seed */
        state = state * 6364136223846793005ull + 11634580456548032391ull;
        uint32_t xs = (uint32_t)(((state >> 18) ^ state) >> 27);
        uint32_t rot = (uint32_t)(state >> 59);
        key[i] = (xs >> rot) | (xs << ((32 - rot) & 31));
    }
    size_t words = 4 * n;
    uint32_t blk[16];
    for (size_t b = 0; b * 16 < words; b++) {
        uint32_t cn[4] = {(uint32_t)b, (uint32_t)((u64)b >> 32), 0, 0};
        bfo_chacha_block(12, key, cn, blk);
        size_t take = words - b * 16 < 16 ? words - b * 16 : 16;
        memcpy(out + 64 * b, blk, 4 * take);
    }
}

/* ===================================================================== */
/* Additive NTT (p3/src/field/binary.rs)                                   */
/* ===================================================================== */
static inline size_t bitrev(size_t v, int bits) {
    size_t r = 0;
    for (int i = 0; i < bits; i++) r |= ((v >> i) & 1) << (bits - 1 - i);
    return r;
}

/* binary.rs:23-53.  Layer K (2^K entries) lands at offset 2^K - 1.  Chain
 * element i carries W_i(beta_i) and W_i(beta_{i+1..}); twiddles are the
 * subset XORs of W_i(beta_{i+1+t}) / W_i(beta_i) and go to layer log_n-1-i. */
static void ntt_twiddles(int log_n, u128 *out) {
    if (log_n <= 0) return;
    u128 *ws = malloc(sizeof(u128) * (size_t)log_n);
    u128 wbi = from_canonical(1);
    int len = log_n - 1;
    for (int j = 0; j < len; j++) ws[j] = from_canonical((u128)1 << (j + 1));
    for (int i = 0; i < log_n; i++) {
        int K = log_n - 1 - i;
        u128 *tw = out + (((size_t)1 << K) - 1);
        u128 inv = finv(wbi);
        tw[0] = 0;
        for (int j = 0; j < len; j++) {
            u128 w = fmul(ws[j], inv);
            for (size_t k = 0; k < ((size_t)1 << j); k++) tw[((size_t)1 << j) + k] = tw[k] ^ w;
        }
        if (len == 0) break;
        u128 nb = fmul(ws[0], ws[0]) ^ fmul(ws[0], wbi);
        for (int j = 1; j < len; j++) ws[j - 1] = fmul(ws[j], ws[j]) ^ fmul(ws[j], wbi);
        len--;
        wbi = nb;
    }
    free(ws);
}

void bfo_ntt_twiddles(int log_n, uint8_t *out) {
    size_t total = ((size_t)1 << log_n) - 1;
    u128 *t = malloc(sizeof(u128) * (total ? total : 1));
    ntt_twiddles(log_n, t);
    memcpy(out, t, 16 * total);
    free(t);
}

/* binary.rs:98-137.  layers[K] is the K-th twiddle slice (2^K entries) of a
 * transform of log size L.  Forward pairs layer i = L-1..0 with layers[L-1-i];
 * chunk k of 2^(i+1) rows uses layers[L-1-i][k]: lo += t*hi; hi += lo.
 * Backward runs i = 0..L-1 with hi += lo; lo += t*hi. */
static void ntt_apply(const u128 *const *layers, int L, u128 *data, size_t width, int inverse) {
    for (int step = 0; step < L; step++) {
        int i = inverse ? step : L - 1 - step;
        const u128 *tw = layers[L - 1 - i];
        size_t half = (size_t)1 << i;
        size_t nbf = ((size_t)1 << (L - 1));
#pragma omp parallel for schedule(static) if (nbf >= 4096)
        for (size_t q = 0; q < nbf; q++) {
            size_t k = q >> i, o = q & (half - 1);
            size_t lo = (k << (i + 1)) + o, hi = lo + half;
            u128 t = tw[k];
            for (size_t c = 0; c < width; c++) {
                u128 *a = &data[lo * width + c], *b = &data[hi * width + c];
                if (!inverse) { *a ^= fmul(t, *b); *b ^= *a; }
                else { *b ^= *a; *a ^= fmul(t, *b); }
            }
        }
    }
}

static void ntt_whole(int log_n, uint8_t *data, size_t width, int inverse) {
    size_t total = ((size_t)1 << log_n) - 1;
    u128 *t = malloc(sizeof(u128) * (total ? total : 1));
    ntt_twiddles(log_n, t);
    const u128 **layers = malloc(sizeof(u128 *) * (size_t)(log_n > 0 ? log_n : 1));
    for (int K = 0; K < log_n; K++) layers[K] = t + (((size_t)1 << K) - 1);
    size_t n = ((size_t)1 << log_n) * width;
    u128 *d = malloc(sizeof(u128) * n);
    memcpy(d, data, 16 * n);
    ntt_apply(layers, log_n, d, width, inverse);
    memcpy(data, d, 16 * n);
    free(d); free(layers); free(t);
}
void bfo_ntt_forward(int log_n, uint8_t *data, size_t width) { ntt_whole(log_n, data, width, 0); }
void bfo_ntt_backward(int log_n, uint8_t *data, size_t width) { ntt_whole(log_n, data, width, 1); }

/* binary.rs:156-191 — the evaluation matrix built from the subspace
 * polynomials: row for point x is recurse(0, U_log_n, x), where each step maps
 * the point set through q_i and interleaves (v, v*x). */
static u128 def_w_beta(int i) {  /* prod_{u in U_i} (basis(i) - u) */
    u128 b = from_canonical((u128)1 << i), p = ONE;
    for (size_t u = 0; u < ((size_t)1 << i); u++) p = fmul(p, b ^ from_canonical(u));
    return p;
}
static u128 def_q(int i, u128 x, const u128 *coef) { return fmul(fmul(coef[i], x), x ^ ONE); }
static void def_recurse(int i, u128 *s, size_t slen, u128 x, const u128 *coef, u128 *out, size_t *olen) {
    if (slen == 1) { out[0] = ONE; *olen = 1; return; }
    u128 *ns = malloc(sizeof(u128) * slen);
    size_t nl = 0;
    for (size_t k = 0; k < slen; k++) {  /* map then drop consecutive duplicates */
        u128 v = def_q(i, s[k], coef);
        if (nl == 0 || ns[nl - 1] != v) ns[nl++] = v;
    }
    u128 *sub = malloc(sizeof(u128) * slen);
    size_t sl;
    def_recurse(i + 1, ns, nl, def_q(i, x, coef), coef, sub, &sl);
    for (size_t k = 0; k < sl; k++) { out[2 * k] = sub[k]; out[2 * k + 1] = fmul(sub[k], x); }
    *olen = 2 * sl;
    free(ns); free(sub);
}
void bfo_ntt_forward_definition(int log_n, const uint8_t *data, uint8_t *out) {
    size_t n = (size_t)1 << log_n;
    u128 coef[64];
    for (int i = 0; i < log_n; i++) {
        u128 a = def_w_beta(i), b = def_w_beta(i + 1);
        coef[i] = fmul(fmul(a, a), finv(b));
    }
    u128 *s = malloc(sizeof(u128) * n), *row = malloc(sizeof(u128) * n);
    for (size_t p = 0; p < n; p++) {
        for (size_t u = 0; u < n; u++) s[u] = from_canonical(u);
        size_t rl;
        def_recurse(0, s, n, from_canonical(p), coef, row, &rl);
        u128 acc = 0;
        for (size_t j = 0; j < n; j++) acc ^= fmul(row[j], ld(data + 16 * j));
        st(out + 16 * p, acc);
    }
    free(s); free(row);
}

/* ===================================================================== */
/* Binary Reed-Solomon code (binary_reed_solomon.rs)                       */
/* ===================================================================== */
/* :55-59 — ceil(lambda / -log2((1 + 1/c) / 2)), cast with Rust's saturation */
size_t bfo_num_queries(int lambda, int log2_c) {
    double c = (double)((u64)1 << log2_c);
    double delta = (1.0 + 1.0 / c) / 2.0;
    double q = ceil((double)lambda / -log2(delta));
    if (!(q > 0)) return 0;  /* NaN, -inf and negatives saturate to 0 */
    return (size_t)q;
}

typedef struct {
    int log2_c, d, L;
    u128 *tab;          /* AddtiveNtt::new(log2_c + d) layers, flattened */
    const u128 **layers;
} rs_code;

static void rs_new(rs_code *rs, int log2_c, int d) {
    rs->log2_c = log2_c; rs->d = d; rs->L = log2_c + d;
    size_t total = ((size_t)1 << rs->L) - 1;
    rs->tab = malloc(sizeof(u128) * (total ? total : 1));
    ntt_twiddles(rs->L, rs->tab);
    rs->layers = malloc(sizeof(u128 *) * (size_t)(rs->L > 0 ? rs->L : 1));
    for (int K = 0; K < rs->L; K++) rs->layers[K] = rs->tab + (((size_t)1 << K) - 1);
}
static void rs_free(rs_code *rs) { free(rs->tab); free(rs->layers); }

/* :67-82 — bitrev rows, zero-pad, replicate into the c cosets, coset NTTs
 * (coset t uses layers[b..] sliced to chunk t, binary.rs:55-68), bitrev rows. */
static void rs_encode(const rs_code *rs, const u128 *data, size_t width, u128 *out) {
    int b = rs->log2_c, d = rs->d;
    size_t N = (size_t)1 << d, c = (size_t)1 << b, M = N * c;
    u128 *m = malloc(sizeof(u128) * M * width);
#pragma omp parallel for schedule(static) if (N >= 4096)
    for (size_t j = 0; j < N; j++)
        memcpy(&m[j * width], &data[bitrev(j, d) * width], 16 * width);
    for (size_t t = 1; t < c; t++) memcpy(&m[t * N * width], m, 16 * N * width);
    const u128 **cl = malloc(sizeof(u128 *) * (size_t)(d ? d : 1));
    for (size_t t = 0; t < c; t++) {
        for (int K = 0; K < d; K++) cl[K] = rs->layers[K + b] + (t << K);
        ntt_apply(cl, d, m + t * N * width, width, 0);
    }
    free(cl);
    int L = b + d;
#pragma omp parallel for schedule(static) if (M >= 4096)
    for (size_t v = 0; v < M; v++)
        memcpy(&out[v * width], &m[bitrev(v, L) * width], 16 * width);
    free(m);
}

void bfo_rs_encode(int log2_c, int d, const uint8_t *data, size_t width, uint8_t *out) {
    rs_code rs;
    rs_new(&rs, log2_c, d);
    size_t N = (size_t)1 << d, M = N << log2_c;
    u128 *in = malloc(sizeof(u128) * N * width), *o = malloc(sizeof(u128) * M * width);
    memcpy(in, data, 16 * N * width);
    rs_encode(&rs, in, width, o);
    memcpy(out, o, 16 * M * width);
    free(in); free(o);
    rs_free(&rs);
}

/* :103-109 — inverse additive butterfly, then linear interpolation at r */
static inline u128 interpolate(u128 t, u128 lo, u128 hi, u128 r) {
    hi ^= lo;
    lo ^= fmul(hi, t);
    return lo ^ fmul(hi ^ lo, r);
}

/* :88-96 — w has 2^(b+i+1) entries; twiddle j is layers[b+i][bitrev_{b+i}(j)] */
static void rs_fold(const rs_code *rs, int i, u128 *w, u128 r) {
    int K = rs->log2_c + i;
    size_t mid = (size_t)1 << K;
    const u128 *tw = rs->layers[K];
#pragma omp parallel for schedule(static) if (mid >= 4096)
    for (size_t j = 0; j < mid; j++) w[j] = interpolate(tw[bitrev(j, K)], w[j], w[j + mid], r);
}

void bfo_rs_fold(int log2_c, int d, int i, const uint8_t *w, const uint8_t *r, uint8_t *out) {
    rs_code rs;
    rs_new(&rs, log2_c, d);
    size_t len = (size_t)1 << (log2_c + i + 1);
    u128 *buf = malloc(sizeof(u128) * len);
    memcpy(buf, w, 16 * len);
    rs_fold(&rs, i, buf, ld(r));
    memcpy(out, buf, 16 * (len / 2));
    free(buf);
    rs_free(&rs);
}

/* ===================================================================== */
/* Multilinear helpers (p3/src/poly/multilinear.rs)                        */
/* ===================================================================== */
/* :200-214 — recursive halving, last coordinate fixes the top index bit */
static u128 mle_eval(const u128 *ev, int n, const u128 *x) {
    if (n == 0) return ev[0];
    size_t h = (size_t)1 << (n - 1);
    u128 lo = mle_eval(ev, n - 1, x), hi = mle_eval(ev + h, n - 1, x);
    return fmul(x[n - 1], hi ^ lo) ^ lo;
}
void bfo_mle_evaluate(const uint8_t *evals, int n, const uint8_t *point, uint8_t out[16]) {
    size_t N = (size_t)1 << n;
    u128 *e = malloc(sizeof(u128) * N), x[64];
    memcpy(e, evals, 16 * N);
    for (int i = 0; i < n; i++) x[i] = ld(point + 16 * i);
    st(out, mle_eval(e, n, x));
    free(e);
}

/* :41-64 with :216-222 — eq table, index bit i <-> variable i */
static void eq_table(const u128 *r, int n, u128 scalar, u128 *ev) {
    ev[0] = scalar;
    for (int i = 0; i < n; i++) {
        size_t h = (size_t)1 << i;
        for (size_t k = 0; k < h; k++) {
            ev[h + k] = fmul(ev[k], r[i]);
            ev[k] ^= ev[h + k];
        }
    }
}
void bfo_eq_table(const uint8_t *r, int n, uint8_t *out) {
    u128 x[64];
    for (int i = 0; i < n; i++) x[i] = ld(r + 16 * i);
    size_t N = (size_t)1 << n;
    u128 *ev = malloc(sizeof(u128) * N);
    eq_table(x, n, ONE, ev);
    memcpy(out, ev, 16 * N);
    free(ev);
}

/* :224-229 — prod(2xy + 1 - x - y); in characteristic 2 this is prod(1 + x + y) */
static u128 eq_eval(const u128 *x, const u128 *y, size_t n) {
    u128 p = ONE;
    for (size_t i = 0; i < n; i++) p = fmul(p, ONE ^ x[i] ^ y[i]);
    return p;
}

/* ===================================================================== */
/* Merkle tree: leaf = Keccak256(row bytes), node = Keccak256(l || r)       */
/* (merkle_tree.rs:8-21, symmetric.rs:11-26; plonky3 MerkleTreeMmcs)       */
/* ===================================================================== */
typedef struct {
    size_t height;      /* power of two */
    int depth;          /* log2(height) */
    uint8_t **layer;    /* layer[k] has height >> k digests */
} merkle;

static void merkle_build(merkle *t, const uint8_t *rows, size_t height, size_t row_bytes) {
    int depth = 0;
    while (((size_t)1 << depth) < height) depth++;
    t->height = height;
    t->depth = depth;
    t->layer = malloc(sizeof(uint8_t *) * (size_t)(depth + 1));
    t->layer[0] = malloc(32 * height);
#pragma omp parallel for schedule(static) if (height >= 1024)
    for (size_t i = 0; i < height; i++) bfo_keccak256(rows + i * row_bytes, row_bytes, t->layer[0] + 32 * i);
    for (int k = 1; k <= depth; k++) {
        size_t h = height >> k;
        t->layer[k] = malloc(32 * h);
        const uint8_t *prev = t->layer[k - 1];
#pragma omp parallel for schedule(static) if (h >= 1024)
        for (size_t i = 0; i < h; i++) bfo_keccak256(prev + 64 * i, 64, t->layer[k] + 32 * i);
    }
}
static const uint8_t *merkle_root_ptr(const merkle *t) { return t->layer[t->depth]; }
static void merkle_free(merkle *t) {
    for (int k = 0; k <= t->depth; k++) free(t->layer[k]);
    free(t->layer);
}
void bfo_merkle_root(const uint8_t *rows, size_t height, size_t row_bytes, uint8_t root[32]) {
    merkle t;
    merkle_build(&t, rows, height, row_bytes);
    memcpy(root, merkle_root_ptr(&t), 32);
    merkle_free(&t);
}

/* verify_batch for one matrix: hash the row, then compress with siblings
 * (left/right chosen by the index bit) up to the root. */
static int merkle_verify(const uint8_t root[32], size_t height, size_t idx, const uint8_t *row,
                         size_t row_bytes, const uint8_t *siblings, size_t nsib) {
    int depth = 0;
    while (((size_t)1 << depth) < height) depth++;
    if (nsib != (size_t)depth) return 0;
    uint8_t h[32], buf[64];
    bfo_keccak256(row, row_bytes, h);
    for (size_t k = 0; k < nsib; k++) {
        if ((idx & 1) == 0) { memcpy(buf, h, 32); memcpy(buf + 32, siblings + 32 * k, 32); }
        else { memcpy(buf, siblings + 32 * k, 32); memcpy(buf + 32, h, 32); }
        bfo_keccak256(buf, 64, h);
        idx >>= 1;
    }
    return memcmp(h, root, 32) == 0;
}

/* ===================================================================== */
/* Challenger: GenericChallenger<_, Keccak256Hash> over HashChallenger      */
/* (challenger.rs:24-90): observe appends bytes and invalidates output;     */
/* sampling hashes the input when output is empty, chains the digest as the */
/* new input, and pops bytes from the end of the digest.                    */
/* ===================================================================== */
struct bfo_challenger {
    uint8_t *in;
    size_t in_len, in_cap;
    uint8_t out[32];
    int out_len;
};
bfo_challenger *bfo_challenger_new(void) { return calloc(1, sizeof(bfo_challenger)); }
void bfo_challenger_free(bfo_challenger *c) { if (c) { free(c->in); free(c); } }
void bfo_challenger_observe(bfo_challenger *c, const uint8_t *bytes, size_t len) {
    c->out_len = 0;
    if (c->in_len + len > c->in_cap) {
        c->in_cap = 2 * (c->in_len + len) + 64;
        c->in = realloc(c->in, c->in_cap);
    }
    memcpy(c->in + c->in_len, bytes, len);
    c->in_len += len;
}
static uint8_t ch_byte(bfo_challenger *c) {
    if (c->out_len == 0) {
        bfo_keccak256(c->in, c->in_len, c->out);
        if (c->in_cap < 32) { c->in_cap = 64; c->in = realloc(c->in, c->in_cap); }
        memcpy(c->in, c->out, 32);
        c->in_len = 32;
        c->out_len = 32;
    }
    return c->out[--c->out_len];
}
static u128 ch_sample(bfo_challenger *c) {
    uint8_t b[16];
    for (int i = 0; i < 16; i++) b[i] = ch_byte(c);
    return ld(b);
}
void bfo_challenger_sample(bfo_challenger *c, uint8_t out[16]) { st(out, ch_sample(c)); }
uint64_t bfo_challenger_sample_bits(bfo_challenger *c, int bits) {
    uint8_t b[8];
    for (int i = 0; i < 8; i++) b[i] = ch_byte(c);
    u64 v;
    memcpy(&v, b, 8);
    return bits >= 64 ? v : (v & (((u64)1 << bits) - 1));
}
static void ch_observe_elem(bfo_challenger *c, u128 v) { uint8_t b[16]; st(b, v); bfo_challenger_observe(c, b, 16); }

/* ===================================================================== */
/* bincode writer / reader (bincode 1.3 defaults: u64 LE lengths)           */
/* ===================================================================== */
typedef struct { uint8_t *p; size_t len, cap; } bbuf;
static void bb_put(bbuf *b, const void *src, size_t n) {
    if (b->len + n > b->cap) { b->cap = 2 * (b->len + n) + 256; b->p = realloc(b->p, b->cap); }
    memcpy(b->p + b->len, src, n);
    b->len += n;
}
static void bb_u64(bbuf *b, u64 v) { bb_put(b, &v, 8); }
static void bb_elem(bbuf *b, u128 v) { bb_put(b, &v, 16); }

typedef struct { const uint8_t *p; size_t len, pos; int err; } brd;
static u64 br_u64(brd *r) {
    u64 v = 0;
    if (r->pos + 8 > r->len) { r->err = 1; return 0; }
    memcpy(&v, r->p + r->pos, 8); r->pos += 8; return v;
}
static u128 br_elem(brd *r) {
    if (r->pos + 16 > r->len) { r->err = 1; return 0; }
    u128 v = ld(r->p + r->pos); r->pos += 16; return v;
}
static const uint8_t *br_bytes(brd *r, size_t n) {
    if (r->pos + n > r->len) { r->err = 1; return NULL; }
    const uint8_t *q = r->p + r->pos; r->pos += n; return q;
}

/* ===================================================================== */
/* Basefold prover (basefold.rs:133-232, 335-369) with the Eval sumcheck    */
/* (sumcheck/src/function/{eval,eval_impr}.rs, sumcheck/src/lib.rs:81-103) */
/* ===================================================================== */
static u128 horner(const u128 *c, int n, u128 x) {  /* univariate.rs:5-12 */
    u128 acc = 0;
    for (int k = n - 1; k >= 0; k--) acc = fmul(acc, x) ^ c[k];
    return acc;
}

/* inner product of the halved eq table with the low half of the table */
static u128 inner_product(const u128 *a, const u128 *b, size_t n) {
    u128 acc = 0;
#pragma omp parallel if (n >= 8192)
    {
        u128 part = 0;
#pragma omp for schedule(static)
        for (size_t k = 0; k < n; k++) part ^= fmul(a[k], b[k]);
#pragma omp critical
        acc ^= part;
    }
    return acc;
}


/* EvalProver::compute_round_poly (eval.rs:89-106) over EvalImprProver
 * (eval_impr.rs:187-197) for round variable i of n: sub is the subclaim the
 * round uses (the outer one in the first round, the inner one after,
 * eval.rs:90-94), chal[j] the challenge of variable j > i.  Writes the
 * degree-2 round polynomial rp and the inner linear one [coeff0, coeff1]. */
static void eval_round_poly(int n, int i, const u128 *r, const u128 *rinv, const u128 *omr_inv,
                            const u128 *half_eq, const u128 *poly, const u128 *chal, u128 sub, u128 rp[3],
                            u128 *coeff0_out, u128 *coeff1_out) {
    const int hn = n > 0 ? n - 1 : 0;
    u128 cf = ONE;                                /* correcting_factor(i), eval_impr.rs:81-88 */
    for (int j = i; j < hn; j++) cf = fmul(cf, omr_inv[j]);
    u128 coeff0 = fmul(cf, inner_product(half_eq, poly, (size_t)1 << i));
    u128 eval1 = fmul(sub ^ fmul(ONE ^ r[i], coeff0), rinv[i]);  /* eval_impr.rs:41-43 */
    u128 coeff1 = eval1 ^ coeff0;
    u128 s = ONE;                                 /* delta_scalar, eval_impr.rs:45-51 */
    if (i != n - 1) s = eq_eval(&r[i + 1], &chal[i + 1], (size_t)(n - 1 - i));
    u128 d0 = fmul(s, ONE ^ r[i]);                /* delta, eval_impr.rs:53-58 */
    u128 d1 = fmul(s, r[i] ^ (ONE ^ r[i]));
    rp[0] = fmul(coeff0, d0);
    rp[1] = fmul(coeff0, d1) ^ fmul(coeff1, d0);
    rp[2] = fmul(coeff1, d1);
    *coeff0_out = coeff0;
    *coeff1_out = coeff1;
}

typedef struct {
    int n, b, d, L;
    size_t N, M, q;
    rs_code rs;
    u128 *data, *cw;
    merkle main;
    u128 claim, final_poly;
    u128 comp[64][2];
    uint8_t (*pi_roots)[32];
    merkle *pi_trees;
    u128 **pi_rows;      /* rows (pi[j], pi[j+half]) of round k, interleaved */
} prover;

static void prover_free(prover *p) {
    for (int k = 0; k < p->d; k++) { merkle_free(&p->pi_trees[k]); free(p->pi_rows[k]); }
    free(p->pi_trees); free(p->pi_rows); free(p->pi_roots);
    merkle_free(&p->main);
    free(p->data); free(p->cw);
    rs_free(&p->rs);
}

/* sample_random_scalars_by_eq (basefold.rs:398-408): log2(next_pow2(w))
 * challenges, their eq table truncated to w entries */
static void scalars_by_eq(bfo_challenger *ch, size_t w, u128 *scalars) {
    int lw = 0;
    while (((size_t)1 << lw) < w) lw++;
    u128 rr[64];
    for (int i = 0; i < lw; i++) rr[i] = ch_sample(ch);     /* sample_ext_vec, challenger.rs:17-19 */
    u128 *e = malloc(sizeof(u128) * ((size_t)1 << lw));
    eq_table(rr, lw, ONE, e);
    memcpy(scalars, e, sizeof(u128) * w);
    free(e);
}

/* row-wise dot_product(&scalars, row) over a rows x w row-major matrix (basefold.rs:168-183) */
static void row_dots(const u128 *m, size_t rows, size_t w, const u128 *scalars, u128 *out) {
#pragma omp parallel for schedule(static) if (rows >= 1024)
    for (size_t j = 0; j < rows; j++) {
        u128 acc = 0;
        for (size_t t = 0; t < w; t++) acc ^= fmul(scalars[t], m[j * w + t]);
        out[j] = acc;
    }
}

int bfo_basefold_prove(int lambda, int log2_c, int n, const uint8_t *evals_b,
                       const uint8_t *point_b, uint8_t root_out[32], uint8_t claim_out[16],
                       uint8_t **proof_out, size_t *proof_len) {
    return bfo_basefold_prove_w(lambda, log2_c, n, 1, evals_b, point_b, root_out, claim_out, proof_out, proof_len);
}

int bfo_basefold_prove_w(int lambda, int log2_c, int n, size_t width, const uint8_t *evals_b,
                         const uint8_t *point_b, uint8_t root_out[32], uint8_t *claims_out,
                         uint8_t **proof_out, size_t *proof_len) {
    return bfo_basefold_prove_px(lambda, log2_c, n, width, evals_b, point_b, NULL, 0, root_out, claims_out,
                                 proof_out, proof_len);
}

int bfo_basefold_prove_px(int lambda, int log2_c, int n, size_t width, const uint8_t *evals_b,
                          const uint8_t *point_b, const uint8_t *pre, size_t pre_len, uint8_t root_out[32],
                          uint8_t *claims_out, uint8_t **proof_out, size_t *proof_len) {
    if (n < 0 || n > 40 || log2_c < 0 || log2_c > 8 || width < 1) return -1;
    const size_t w = width;
    prover P;
    memset(&P, 0, sizeof P);
    P.n = n; P.b = log2_c; P.d = n; P.L = log2_c + n;
    P.N = (size_t)1 << n; P.M = P.N << log2_c;
    P.q = bfo_num_queries(lambda, log2_c);
    rs_new(&P.rs, log2_c, n);
    int b = P.b, d = P.d;
    size_t N = P.N, M = P.M;
    P.data = malloc(sizeof(u128) * N * w);
    P.cw = malloc(sizeof(u128) * M * w);
    memcpy(P.data, evals_b, 16 * N * w);
    u128 r[64], rinv[64];
    for (int i = 0; i < n; i++) r[i] = ld(point_b + 16 * i);

    /* commit: encode_batch(clone(data)) then Merkle over the w-element rows (basefold.rs:112-120) */
    rs_encode(&P.rs, P.data, w, P.cw);
    merkle_build(&P.main, (const uint8_t *)P.cw, M, 16 * w);
    memcpy(root_out, merkle_root_ptr(&P.main), 32);

    /* harness transcript prefix: observe(root), observe_ext_slice(values) (poly-commit/src/lib.rs:82-92);
     * values = one MLE evaluation per column (basefold.rs:126-131) */
    bfo_challenger *ch = bfo_challenger_new();
    if (pre_len) bfo_challenger_observe(ch, pre, pre_len); /* whatever the caller observed first */
    bfo_challenger_observe(ch, merkle_root_ptr(&P.main), 32);
    u128 *claims = malloc(sizeof(u128) * w), *col = malloc(sizeof(u128) * N);
    for (size_t t = 0; t < w; t++) {
        for (size_t j = 0; j < N; j++) col[j] = P.data[j * w + t];
        claims[t] = mle_eval(col, n, r);
        st(claims_out + 16 * t, claims[t]);
        ch_observe_elem(ch, claims[t]);
    }
    free(col);

    /* open (basefold.rs:157-185): width 1 takes the data and the codeword as they are; otherwise
     * random scalars by eq combine the columns of data, claims and codeword rows */
    u128 *pi = malloc(sizeof(u128) * M);
    u128 *poly = malloc(sizeof(u128) * N);
    if (w == 1) {
        memcpy(pi, P.cw, 16 * M);
        memcpy(poly, P.data, 16 * N);
        P.claim = claims[0];
    } else {
        u128 *sc = malloc(sizeof(u128) * w);
        scalars_by_eq(ch, w, sc);
        row_dots(P.data, N, w, sc, poly);
        P.claim = 0;
        for (size_t t = 0; t < w; t++) P.claim ^= fmul(sc[t], claims[t]);
        row_dots(P.cw, M, w, sc, pi);
        free(sc);
    }
    free(claims);
    for (int i = 0; i < n; i++) {
        if (r[i] == 0 || r[i] == ONE) { fprintf(stderr, "bfo: degenerate point coordinate\n"); abort(); }
        rinv[i] = finv(r[i]);                         /* EvalImpr::new, eval_impr.rs:20-27 */
    }
    int hn = n > 0 ? n - 1 : 0;                       /* HalfEq::new, eval_impr.rs:69-79 */
    u128 omr_inv[64];
    for (int j = 0; j < hn; j++) omr_inv[j] = finv(ONE ^ r[j]);
    u128 *half_eq = malloc(sizeof(u128) * ((size_t)1 << hn));
    eq_table(r, hn, ONE, half_eq);

    u128 subclaim = P.claim, inner_sub = 0;
    u128 chal[64];                                    /* challenge of round i stored at chal[i] */
    P.pi_roots = malloc(32 * (size_t)(d ? d : 1));
    P.pi_trees = malloc(sizeof(merkle) * (size_t)(d ? d : 1));
    P.pi_rows = malloc(sizeof(u128 *) * (size_t)(d ? d : 1));

    /* prove_interleaved_folding_sumcheck (basefold.rs:335-369) */
    for (int i = d - 1; i >= 0; i--) {
        int k = d - 1 - i;
        size_t len = (size_t)1 << (b + i + 1), half = len / 2;
        u128 *rows = malloc(sizeof(u128) * len);
        for (size_t j = 0; j < half; j++) { rows[2 * j] = pi[j]; rows[2 * j + 1] = pi[j + half]; }
        P.pi_rows[k] = rows;
        merkle_build(&P.pi_trees[k], (const uint8_t *)rows, half, 32);
        memcpy(P.pi_roots[k], merkle_root_ptr(&P.pi_trees[k]), 32);
        bfo_challenger_observe(ch, P.pi_roots[k], 32);

        u128 rp[3], coeff0, coeff1;
        eval_round_poly(n, i, r, rinv, omr_inv, half_eq, poly, chal, (i == n - 1) ? subclaim : inner_sub, rp,
                        &coeff0, &coeff1);
        P.comp[k][0] = rp[0];                         /* compress: eval.rs:64-66 */
        P.comp[k][1] = rp[2];
        ch_observe_elem(ch, rp[0]);                   /* lib.rs:94 */
        ch_observe_elem(ch, rp[2]);
        u128 x = ch_sample(ch);                       /* lib.rs:96 */
        chal[i] = x;
        size_t ph = (size_t)1 << i;                   /* fix_last_var, multilinear.rs:183-198 */
        for (size_t j = 0; j < ph; j++) poly[j] = fmul(x, poly[j + ph] ^ poly[j]) ^ poly[j];
        inner_sub = coeff0 ^ fmul(coeff1, x);         /* eval.rs:108-111 */
        subclaim = horner(rp, 3, x);                  /* lib.rs:100 */
        rs_fold(&P.rs, i, pi, x);                     /* basefold.rs:365 */
    }
    P.final_poly = poly[0];
    ch_observe_elem(ch, P.final_poly);                /* basefold.rs:196 */
    for (size_t t = 0; t < ((size_t)1 << b); t++)     /* basefold.rs:198 invariant */
        if (pi[t] != P.final_poly) { fprintf(stderr, "bfo: fold invariant violated\n"); abort(); }

    /* proof (basefold.rs:49-66, 200-231) in bincode */
    bbuf pb = {0};
    bb_u64(&pb, (u64)d);
    for (int k = 0; k < d; k++) { bb_u64(&pb, 2); bb_elem(&pb, P.comp[k][0]); bb_elem(&pb, P.comp[k][1]); }
    bb_u64(&pb, 1);
    bb_elem(&pb, P.final_poly);
    bb_u64(&pb, (u64)d);
    for (int k = 0; k < d; k++) bb_put(&pb, P.pi_roots[k], 32);
    bb_u64(&pb, (u64)P.q);
    for (size_t qi = 0; qi < P.q; qi++) {
        size_t idx = (size_t)bfo_challenger_sample_bits(ch, P.L);
        bb_u64(&pb, (u64)w);                          /* leaves: Vec<F>, the codeword row */
        for (size_t t = 0; t < w; t++) bb_elem(&pb, P.cw[idx * w + t]);
        bb_u64(&pb, (u64)P.main.depth);
        for (int t = 0; t < P.main.depth; t++) bb_put(&pb, P.main.layer[t] + 32 * ((idx >> t) ^ 1), 32);
        bb_u64(&pb, (u64)d);
        for (int k = 0; k < d; k++) {
            int i = d - 1 - k;
            size_t ni = (size_t)1 << (b + i);
            size_t sib = ((idx / ni) & 1) ^ 1, row = idx % ni;
            bb_elem(&pb, P.pi_rows[k][2 * row + sib]);
            const merkle *t = &P.pi_trees[k];
            bb_u64(&pb, (u64)t->depth);
            for (int u = 0; u < t->depth; u++) bb_put(&pb, t->layer[u] + 32 * ((row >> u) ^ 1), 32);
        }
    }
    *proof_out = pb.p;
    *proof_len = pb.len;
    free(half_eq); free(poly); free(pi);
    bfo_challenger_free(ch);
    prover_free(&P);
    return 0;
}

void bfo_free(void *p) { free(p); }

/* prove_sumcheck (sumcheck/src/lib.rs:56-103) over EvalProver::new(Eval::new(
 * point), MultiPoly::base(poly)) (eval.rs:27-116) with a fresh keccak
 * challenger: observe the claim, then per round observe the compressed
 * [rp0, rp2], sample, fix_last_var and reduce both subclaims.  Outputs the
 * full round polynomials [rp0, rp1, rp2] (n x 3), the challenges in round
 * order, the final table value and the final subclaim. */
int bfo_sumcheck_eval(int n, const uint8_t *poly_b, const uint8_t *point_b, const uint8_t claim_b[16],
                      uint8_t *round_polys, uint8_t *challenges, uint8_t final_eval[16], uint8_t subclaim_out[16]) {
    if (n < 0 || n > 40) return -1;
    const size_t N = (size_t)1 << n;
    u128 r[64], rinv[64], omr_inv[64], chal[64];
    for (int i = 0; i < n; i++) {
        r[i] = ld(point_b + 16 * i);
        if (r[i] == 0 || r[i] == ONE) return -1;
        rinv[i] = finv(r[i]);                         /* EvalImpr::new, eval_impr.rs:20-27 */
    }
    const int hn = n > 0 ? n - 1 : 0;                 /* HalfEq::new, eval_impr.rs:69-79 */
    for (int j = 0; j < hn; j++) omr_inv[j] = finv(ONE ^ r[j]);
    u128 *half_eq = malloc(sizeof(u128) * ((size_t)1 << hn));
    u128 *poly = malloc(sizeof(u128) * N);
    eq_table(r, hn, ONE, half_eq);
    memcpy(poly, poly_b, 16 * N);
    u128 subclaim = ld(claim_b), inner_sub = 0;
    bfo_challenger *ch = bfo_challenger_new();
    if (n > 0) ch_observe_elem(ch, subclaim);         /* lib.rs:63-67 */
    for (int i = n - 1; i >= 0; i--) {
        const int k = n - 1 - i;
        u128 rp[3], coeff0, coeff1;
        eval_round_poly(n, i, r, rinv, omr_inv, half_eq, poly, chal, (i == n - 1) ? subclaim : inner_sub, rp,
                        &coeff0, &coeff1);
        for (int t = 0; t < 3; t++) st(round_polys + 48 * k + 16 * t, rp[t]);
        ch_observe_elem(ch, rp[0]);                   /* compress (eval.rs:64-66), lib.rs:94 */
        ch_observe_elem(ch, rp[2]);
        u128 x = ch_sample(ch);                       /* lib.rs:96 */
        chal[i] = x;
        st(challenges + 16 * k, x);
        size_t ph = (size_t)1 << i;                   /* fix_last_var, multilinear.rs:183-198 */
        for (size_t j = 0; j < ph; j++) poly[j] = fmul(x, poly[j + ph] ^ poly[j]) ^ poly[j];
        inner_sub = coeff0 ^ fmul(coeff1, x);         /* eval.rs:108-111 */
        subclaim = horner(rp, 3, x);                  /* lib.rs:100 */
    }
    st(final_eval, poly[0]);
    st(subclaim_out, subclaim);
    free(half_eq);
    free(poly);
    bfo_challenger_free(ch);
    return 0;
}

/* basefold.rs:234-325 with verify_interleaved_folding_sumcheck (:371-396) and
 * Eval::decompress_round_poly (eval.rs:68-79). */
int bfo_basefold_verify(int lambda, int log2_c, int n, const uint8_t root[32],
                        const uint8_t *point_b, const uint8_t claim_b[16],
                        const uint8_t *proof, size_t proof_len) {
    return bfo_basefold_verify_w(lambda, log2_c, n, 1, root, point_b, claim_b, proof, proof_len);
}

int bfo_basefold_verify_w(int lambda, int log2_c, int n, size_t width, const uint8_t root[32],
                          const uint8_t *point_b, const uint8_t *claims_b,
                          const uint8_t *proof, size_t proof_len) {
    return bfo_basefold_verify_px(lambda, log2_c, n, width, root, point_b, claims_b, NULL, 0, proof, proof_len);
}

int bfo_basefold_verify_px(int lambda, int log2_c, int n, size_t width, const uint8_t root[32],
                           const uint8_t *point_b, const uint8_t *claims_b, const uint8_t *pre, size_t pre_len,
                           const uint8_t *proof, size_t proof_len) {
    if (width < 1 || width > ((size_t)1 << 24)) return 4;
    const size_t w = width;
    int d = n, b = log2_c, L = b + d;
    size_t q = bfo_num_queries(lambda, log2_c);
    brd rd = {proof, proof_len, 0, 0};
    u128 r[64];
    for (int i = 0; i < n; i++) r[i] = ld(point_b + 16 * i);
    if (br_u64(&rd) != (u64)d) return 4;
    u128 comp[64][2];
    for (int k = 0; k < d; k++) {
        if (br_u64(&rd) != 2) return 4;
        comp[k][0] = br_elem(&rd); comp[k][1] = br_elem(&rd);
    }
    if (br_u64(&rd) != 1) return 4;
    u128 final_poly = br_elem(&rd);
    if (br_u64(&rd) != (u64)d) return 4;
    const uint8_t *pi_roots = br_bytes(&rd, 32 * (size_t)d);
    if (rd.err) return 4;

    bfo_challenger *ch = bfo_challenger_new();
    if (pre_len) bfo_challenger_observe(ch, pre, pre_len);
    bfo_challenger_observe(ch, root, 32);
    for (size_t t = 0; t < w; t++) ch_observe_elem(ch, ld(claims_b + 16 * t));
    /* scalars and claim (basefold.rs:252-258) */
    u128 *sc = malloc(sizeof(u128) * w), claim;
    if (w == 1) {
        sc[0] = ONE;
        claim = ld(claims_b);
    } else {
        scalars_by_eq(ch, w, sc);
        claim = 0;
        for (size_t t = 0; t < w; t++) claim ^= fmul(sc[t], ld(claims_b + 16 * t));
    }
    u128 subclaim = claim, chal[64];
    for (int k = 0; k < d; k++) {
        int i = d - 1 - k;
        bfo_challenger_observe(ch, pi_roots + 32 * k, 32);
        ch_observe_elem(ch, comp[k][0]);
        ch_observe_elem(ch, comp[k][1]);
        u128 c0 = comp[k][0], c2 = comp[k][1];
        u128 c1 = subclaim ^ c2;                      /* claim - 2*c0 - c2, char 2 */
        u128 x = ch_sample(ch);
        u128 rp[3] = {c0, c1, c2};
        subclaim = horner(rp, 3, x);
        chal[i] = x;
    }
    ch_observe_elem(ch, final_poly);
    /* final check (:273-276): evaluate(final, point_lo=[]) * eq_eval(point, r) */
    if (fmul(final_poly, eq_eval(r, chal, (size_t)n)) != subclaim) { bfo_challenger_free(ch); free(sc); return 1; }

    rs_code rs;
    rs_new(&rs, log2_c, n);
    int rc = 0;
    if (br_u64(&rd) != (u64)q || rd.err) { rc = 4; goto out; }
    for (size_t qi = 0; qi < q && !rc; qi++) {
        size_t idx = (size_t)bfo_challenger_sample_bits(ch, L);
        if (br_u64(&rd) != (u64)w) { rc = 4; break; }
        const uint8_t *lb = br_bytes(&rd, 16 * w);
        u64 np = br_u64(&rd);
        if (rd.err || np > 64) { rc = 4; break; }
        const uint8_t *sib = br_bytes(&rd, 32 * np);
        if (rd.err) { rc = 4; break; }
        if (!merkle_verify(root, (size_t)1 << L, idx, lb, 16 * w, sib, np)) { rc = 3; break; }
        u128 leaf = 0;                                /* dot_product(scalars, opening row) (:297) */
        for (size_t t = 0; t < w; t++) leaf ^= fmul(sc[t], ld(lb + 16 * t));
        if (br_u64(&rd) != (u64)d) { rc = 4; break; }
        u128 folded = leaf;
        for (int k = 0; k < d && !rc; k++) {
            int i = d - 1 - k;
            size_t ni = (size_t)1 << (b + i);
            size_t s = ((idx / ni) & 1) ^ 1, row = idx % ni;
            u128 sv = br_elem(&rd);
            u64 pl = br_u64(&rd);
            if (rd.err || pl > 64) { rc = 4; break; }
            const uint8_t *ps = br_bytes(&rd, 32 * pl);
            if (rd.err) { rc = 4; break; }
            u128 vals[2] = {folded, folded};
            vals[s] = sv;
            uint8_t rb[32];
            st(rb, vals[0]); st(rb + 16, vals[1]);
            if (!merkle_verify(pi_roots + 32 * k, ni, row, rb, 32, ps, pl)) { rc = 3; break; }
            folded = interpolate(rs.layers[b + i][bitrev(row, b + i)], vals[0], vals[1], chal[i]);
        }
        if (!rc && folded != final_poly) rc = 2;     /* pi_0[idx % n_0] == folded (:319-321) */
    }
out:
    rs_free(&rs);
    bfo_challenger_free(ch);
    free(sc);
    if (!rc && (rd.err || rd.pos != rd.len)) rc = 4;  /* bincode: truncated or trailing bytes */
    return rc;
}

/* ===================================================================== */
/* Degree-3 product sumcheck sum_b f(b) g(b) eq(r, b) — no reference        */
/* counterpart.  Driver follows sumcheck/src/lib.rs:56-103: observe claim,  */
/* per round observe the compressed poly [c0, c2, c3], sample x, fix the    */
/* top variable of every table (multilinear.rs:183-188), subclaim = p(x).   */
/* Straight coefficient expansion with an explicit eq table.               */
/* ===================================================================== */
int bfo_product_sumcheck(int n, const uint8_t *f_b, const uint8_t *g_b, const uint8_t *r_b,
                         uint8_t claim_out[16], uint8_t *round_polys, uint8_t *challenges,
                         uint8_t final_fg[32], uint8_t subclaim_out[16]) {
    size_t N = (size_t)1 << n;
    u128 *f = malloc(sizeof(u128) * N), *g = malloc(sizeof(u128) * N), *e = malloc(sizeof(u128) * N);
    memcpy(f, f_b, 16 * N);
    memcpy(g, g_b, 16 * N);
    u128 r[64];
    for (int i = 0; i < n; i++) r[i] = ld(r_b + 16 * i);
    eq_table(r, n, ONE, e);
    u128 claim = 0;
    for (size_t k = 0; k < N; k++) claim ^= fmul(fmul(f[k], g[k]), e[k]);
    st(claim_out, claim);
    bfo_challenger *ch = bfo_challenger_new();
    if (n > 0) ch_observe_elem(ch, claim);
    u128 sub = claim;
    for (int i = n - 1; i >= 0; i--) {
        int k = n - 1 - i;
        size_t h = (size_t)1 << i;
        u128 c[4] = {0, 0, 0, 0};
        for (size_t j = 0; j < h; j++) {
            u128 f0 = f[j], fd = f[j] ^ f[j + h], g0 = g[j], gd = g[j] ^ g[j + h], e0 = e[j], ed = e[j] ^ e[j + h];
            c[0] ^= fmul(fmul(f0, g0), e0);
            c[1] ^= fmul(fmul(f0, g0), ed) ^ fmul(fmul(f0, gd), e0) ^ fmul(fmul(fd, g0), e0);
            c[2] ^= fmul(fmul(f0, gd), ed) ^ fmul(fmul(fd, g0), ed) ^ fmul(fmul(fd, gd), e0);
            c[3] ^= fmul(fmul(fd, gd), ed);
        }
        if ((c[1] ^ c[2] ^ c[3]) != sub) { fprintf(stderr, "bfo: product sumcheck p(0)+p(1) != claim\n"); abort(); }
        st(round_polys + 48 * k, c[0]);
        st(round_polys + 48 * k + 16, c[2]);
        st(round_polys + 48 * k + 32, c[3]);
        ch_observe_elem(ch, c[0]);
        ch_observe_elem(ch, c[2]);
        ch_observe_elem(ch, c[3]);
        u128 x = ch_sample(ch);
        st(challenges + 16 * k, x);
        for (size_t j = 0; j < h; j++) {
            f[j] ^= fmul(x, f[j] ^ f[j + h]);
            g[j] ^= fmul(x, g[j] ^ g[j + h]);
            e[j] ^= fmul(x, e[j] ^ e[j + h]);
        }
        sub = horner(c, 4, x);
    }
    st(final_fg, f[0]);
    st(final_fg + 16, g[0]);
    st(subclaim_out, sub);
    bfo_challenger_free(ch);
    free(f); free(g); free(e);
    return 0;
}

/* ===================================================================== */
/* Quadratic sumcheck  sum_b sum_k s_k P_{a_k}(b) P_{b_k}(b)                */
/* (sumcheck/src/function/quadratic.rs:13-134) driven by prove_sumcheck     */
/* (sumcheck/src/lib.rs:56-103).  Per round (top variable first) the round  */
/* poly is c0 = sum_k s_k sum_b f0 g0, c2 = sum_k s_k sum_b (f1-f0)(g1-g0), */
/* c1 = subclaim - 2 c0 - c2 (quadratic.rs:70-80; 2 c0 = 0 here); the      */
/* compressed poly [c0, c2] is observed, x sampled, every table fixed.      */
/* ===================================================================== */
int bfo_sumcheck_quadratic(int n, int m, const uint8_t *polys_b, int npairs, const uint8_t *scalars_b,
                           const uint32_t *pair_idx, uint8_t claim_out[16], uint8_t *round_polys,
                           uint8_t *challenges, uint8_t *final_evals, uint8_t subclaim_out[16]) {
    if (n < 0 || n > 40 || m < 1 || npairs < 1) return -1;
    for (int k = 0; k < 2 * npairs; k++)
        if (pair_idx[k] >= (uint32_t)m) return -1;
    size_t N = (size_t)1 << n;
    u128 *P = malloc(sizeof(u128) * N * (size_t)m);
    memcpy(P, polys_b, 16 * N * (size_t)m);
    u128 *s = malloc(sizeof(u128) * (size_t)npairs);
    for (int k = 0; k < npairs; k++) s[k] = ld(scalars_b + 16 * k);
    /* claim = compute_sum (quadratic.rs:86-99) */
    u128 claim = 0;
    for (int k = 0; k < npairs; k++) {
        const u128 *f = P + N * pair_idx[2 * k], *g = P + N * pair_idx[2 * k + 1];
        u128 acc = 0;
        for (size_t j = 0; j < N; j++) acc ^= fmul(f[j], g[j]);
        claim ^= fmul(s[k], acc);
    }
    st(claim_out, claim);
    u128 sub = claim;
    bfo_challenger *ch = bfo_challenger_new();
    if (n > 0) ch_observe_elem(ch, claim);               /* lib.rs:67 */
    for (int i = n - 1; i >= 0; i--) {
        int k = n - 1 - i;
        size_t h = (size_t)1 << i;
        u128 c0 = 0, c2 = 0;
        for (int t = 0; t < npairs; t++) {
            const u128 *f = P + N * pair_idx[2 * t], *g = P + N * pair_idx[2 * t + 1];
            u128 a0 = 0, a2 = 0;
            for (size_t j = 0; j < h; j++) {
                a0 ^= fmul(f[j], g[j]);
                a2 ^= fmul(f[j + h] ^ f[j], g[j + h] ^ g[j]);
            }
            c0 ^= fmul(s[t], a0);
            c2 ^= fmul(s[t], a2);
        }
        u128 c1 = sub ^ c2;                              /* decompress_round_poly, char 2 */
        st(round_polys + 32 * k, c0);
        st(round_polys + 32 * k + 16, c2);
        ch_observe_elem(ch, c0);
        ch_observe_elem(ch, c2);
        u128 x = ch_sample(ch);
        st(challenges + 16 * k, x);
        for (int p = 0; p < m; p++) {                    /* fix_last_var (multilinear.rs:183-188) */
            u128 *f = P + N * (size_t)p;
            for (size_t j = 0; j < h; j++) f[j] ^= fmul(x, f[j] ^ f[j + h]);
        }
        u128 rp[3] = {c0, c1, c2};
        sub = horner(rp, 3, x);
    }
    for (int p = 0; p < m; p++) st(final_evals + 16 * p, P[N * (size_t)p]);
    st(subclaim_out, sub);
    bfo_challenger_free(ch);
    free(P);
    free(s);
    return 0;
}
