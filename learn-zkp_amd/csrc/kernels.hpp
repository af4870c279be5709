// Launchers of the CDNA4 kernels (kernels.hip).  All take the stream first
// and return the launch status.
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include "gf128.hpp"

namespace bfd {

// One polynomial's share of a segmented launch (meaning of the pointers and
// `arg` per kernel is documented at each segmented kernel).
struct Seg {
    const void *in;
    void *out;
    const void *aux;
    uint64_t count;   // work items of this segment
    uint64_t arg;
    gf x;             // per-segment challenge (fold / fix)
    uint32_t blk0;    // first block of the segment in the launch
    uint32_t nblk;    // blocks of the segment
};

// One proof of a verification batch (device copy; pointers into device memory).
struct VerifyJob {
    const uint8_t *proof;   // bincode bytes, 8-byte aligned
    const uint8_t *root;    // 32-byte main-tree root
    const uint64_t *idx;    // q query indices
    const gf *chal;         // chal[i] = challenge of level i
    const gf *scalars;      // w column scalars (ONE for width 1)
    gf final_poly;
    uint64_t open_off;      // byte offset of the first query record
    uint64_t rec_bytes;     // bytes per query record
    uint64_t roots_off;     // byte offset of the pi-tree roots
};

// One opening's query records for k_proof_records (device copy; pointers
// into device memory).  The records are BasefoldProof's query openings
// (basefold.rs:205-222) in bincode layout, record t at out + t * rec_bytes:
// u64 w, the w elements of codeword row idx[t], u64 L, the L main-tree
// siblings leaf to root, u64 d, then per round k (level i = d-1-k, ni =
// 2^(b+i)): the pi sibling value, u64 b+i and the b+i pi-tree siblings.
// Round k's pi sits at cw (k = 0) or pis + M - (M >> (k-1)); its tree at
// pitrees + 4 (2M - (M >> (k-1)) - k) digests (k >= 1; round 0's at
// pitrees).  rec0 is the job's first record in the launch (jobs sorted by it).
struct ProofJob {
    const uint64_t *idx;       // q query indices
    const gf *cw;              // round-0 pi (width 1: the codeword; else the scalar combination)
    const gf *rows;            // width > 1: the codeword matrix, w elements per row; else nullptr
    const uint64_t *tree;      // main tree, 2M-1 digests, leaf layer first
    const gf *pis;             // pi of rounds 1..d
    const uint64_t *pitrees;   // pi trees of rounds 0..d-1
    uint8_t *out;              // q records
    uint64_t w;
    uint64_t rec_bytes;
    uint32_t b, d, q, rec0;
};

// Device-resident Fiat-Shamir state of one opening: the byte buffers of
// GenericChallenger<_, Keccak256> (p3/src/challenger.rs:24-90 over plonky3's
// HashChallenger<u8, _, 32>) and the Eval-sumcheck scalars that change from
// round to round (sumcheck/src/function/eval.rs:90-110, eval_impr.rs:41-58).
// One region per opening, laid out as
//   FsState | gf consts[3n + 1] (r, r^-1, correcting factors) |
//   FsRound rounds[n] | uint8_t input[in_cap]
struct FsState {
    uint32_t in_len, out_len, in_cap, n;
    gf sub;      // subclaim of the Eval sumcheck (used in the first round)
    gf inner;    // inner subclaim (eval.rs:110)
    gf s_delta;  // eq(r[i+1..], x[i+1..]) of the delta polynomial
    gf x;        // last sampled challenge (read by the fold and fix kernels)
    uint8_t out[32];  // sampled digest; bytes out[0, out_len) are still unread
};
// What one round leaves for the proof: the compressed round polynomial
// (eval.rs:64-66) and the pi-tree root observed before it.
struct FsRound {
    gf rp0, rp2;
    uint64_t root[4];
};
__host__ __device__ inline size_t fs_consts_off() { return sizeof(FsState); }
__host__ __device__ inline size_t fs_rounds_off(uint32_t n) { return sizeof(FsState) + 16 * (3 * (size_t)n + 1); }
__host__ __device__ inline size_t fs_input_off(uint32_t n) { return fs_rounds_off(n) + sizeof(FsRound) * (size_t)n; }
__host__ __device__ inline size_t fs_bytes(uint32_t n, uint32_t in_cap) { return fs_input_off(n) + in_cap; }

// Device-resident transcript of a degree-3 product sumcheck (the
// prove_sumcheck driver, sumcheck/src/lib.rs:56-103, over f g eq(r, .)): the
// challenger's byte buffers and the running eq product.  One region laid out
// as
//   ProdFs | gf cf[n] | gf e0[n] | gf ecf[n] | gf rho[n] | gf polys[3n] |
//   gf chal[n] | uint8_t input[in_cap] | gf finals[4] |
//   gf look[2][kLookSlots][9] | per round k: uint64_t acc[8 kProdSlots],
//   then its block counters (kCountWords words)
// cf[i] is the correcting factor of the halved eq table at round variable i,
// e0[i] = 1 + r_i, ecf[i] = e0[i] * cf[i] and rho[i] = r_i / (1 + r_i), the
// ratio of halved eq table entries whose indices differ in bit i; look holds
// the coefficient sums of the lookahead rounds (kernels.hip k_prod_look),
// one set per round parity; polys and chal receive each
// round's compressed polynomial [c0, c2, c3] and challenge, finals the last
// round's two entries of f and of g.  Round k's acc holds its polynomial in
// kProdSlots slots of [c0, c1, c2, c3] that the blocks XOR their shares
// into; its counters count the finished blocks in two levels (kernels.hip
// grid_last).  The upload before the rounds carries round 0's slots and
// counters as zeros; round k clears round k + 1's.
struct ProdFs {
    uint32_t in_len, out_len, in_cap, n;
    // eq scalars.  Round k (variable i = n-1-k) needs s_k, the product of
    // eq(r_j, x_j) over the rounds done, K_k = s_k e0[i] cf[i] and
    // L_k = s_k cf[i].  Round 0 takes s = s_0, K = K_0, L = L_0 as uploaded;
    // before round k >= 1 the region holds s = s_{k-1} and P1 = s_{k-1}
    // ecf[i], P2 = s_{k-1} cf[i], and with t = e0[i + 1] + x_{k-1},
    // s_k = s_{k-1} t, K_k = P1 t, L_k = P2 t.
    gf s, K, L, P1, P2;
    gf x;      // last sampled challenge (read by the next round's fold)
    gf claim;  // sum over the cube, observed before the first round
    uint8_t out[32];
};
constexpr int kProdSlots = 256;  // accumulation slots per round (blocks spread their atomics over them)
constexpr int kLookSlots = 64;   // slots of one set of lookahead coefficient sums
__host__ __device__ inline size_t prod_cf_off() { return sizeof(ProdFs); }
__host__ __device__ inline size_t prod_polys_off(uint32_t n) { return sizeof(ProdFs) + 64 * (size_t)n; }
__host__ __device__ inline size_t prod_chal_off(uint32_t n) { return prod_polys_off(n) + 48 * (size_t)n; }
__host__ __device__ inline size_t prod_input_off(uint32_t n) { return prod_chal_off(n) + 16 * (size_t)n; }
__host__ __device__ inline size_t prod_finals_off(uint32_t n, uint32_t in_cap) {
    return (prod_input_off(n) + in_cap + 15) & ~(size_t)15;
}
__host__ __device__ inline size_t prod_look_off(uint32_t n, uint32_t in_cap) {
    return (prod_finals_off(n, in_cap) + 64 + 255) & ~(size_t)255;
}
__host__ __device__ inline size_t prod_acc_off(uint32_t n, uint32_t in_cap) {
    return prod_look_off(n, in_cap) + 2 * 9 * 16 * (size_t)kLookSlots;
}
// Two-level block counters of one launch: word 0 counts the groups done,
// word kCountLine (1 + g) the blocks of group g done; each on its own
// 128-byte line, so the atomics of different groups do not queue behind
// one another.
constexpr int kCountGroups = 16;
constexpr int kCountLine = 32;
constexpr size_t kCountWords = (size_t)kCountLine * (1 + kCountGroups);
constexpr size_t kProdSlotBytes = 8 * 8 * (size_t)kProdSlots;
constexpr size_t kProdAccRound = kProdSlotBytes + 4 * kCountWords;  // bytes of one round's slots and counters
// bytes uploaded before the rounds: everything up to and including acc[0]
__host__ __device__ inline size_t prod_upload_bytes(uint32_t n, uint32_t in_cap) {
    return prod_acc_off(n, in_cap) + kProdAccRound;
}
__host__ __device__ inline size_t prod_bytes(uint32_t n, uint32_t in_cap) {
    return prod_acc_off(n, in_cap) + kProdAccRound * (size_t)n;
}
// The halved eq table of a product sumcheck as two factors:
// T[j] = A[j mod 2^lo] B[j >> lo].  Round kernels multiply f by A per pair
// and each block's sums by the one B entry of its range, so T is never
// stored.
struct EqFactors {
    const gf *A;
    const gf *B;
    int lo;
};

// Device-resident transcript of a Quadratic sumcheck driven by
// prove_sumcheck (quadratic.rs:84-136, sumcheck/src/lib.rs:56-103): the
// challenger's byte buffers.  One region laid out as
//   QuadFs | gf scalars[npairs] | gf polys[2n] | gf chal[n] |
//   uint8_t input[in_cap] | gf finals[m] | uint32_t pairs[2 npairs] |
//   gf slots[kQuadSlots][3] | uint32_t counters[kCountWords] |
//   gf look[2][kLookSlots][8]
// polys and chal receive each round's compressed polynomial [c0, c2] and
// challenge, finals every table's last entry.  The round kernels XOR their
// shares of [c0, c2, claim] into the slots and count their blocks in the
// counters (both zero at upload; the last block of a round clears them);
// look holds the coefficient sums of the lookahead rounds (kernels.hip
// k_quad_look), one set per round parity.
struct QuadFs {
    uint32_t in_len, out_len, in_cap, n;
    uint32_t npairs, pad_[3];
    gf x;      // last sampled challenge (read by the round's fix)
    gf claim;  // sum over the cube, observed before the first round
    uint8_t out[32];
};
constexpr int kQuadSlots = 256;
__host__ __device__ inline size_t quad_scalars_off() { return sizeof(QuadFs); }
__host__ __device__ inline size_t quad_polys_off(uint32_t npairs) { return sizeof(QuadFs) + 16 * (size_t)npairs; }
__host__ __device__ inline size_t quad_chal_off(uint32_t npairs, uint32_t n) {
    return quad_polys_off(npairs) + 32 * (size_t)n;
}
__host__ __device__ inline size_t quad_input_off(uint32_t npairs, uint32_t n) {
    return quad_chal_off(npairs, n) + 16 * (size_t)n;
}
__host__ __device__ inline size_t quad_finals_off(uint32_t npairs, uint32_t n, uint32_t in_cap) {
    return (quad_input_off(npairs, n) + in_cap + 15) & ~(size_t)15;
}
__host__ __device__ inline size_t quad_pairs_off(uint32_t npairs, uint32_t n, uint32_t in_cap, uint32_t m) {
    return quad_finals_off(npairs, n, in_cap) + 16 * (size_t)m;
}
__host__ __device__ inline size_t quad_slots_off(uint32_t npairs, uint32_t n, uint32_t in_cap, uint32_t m) {
    return (quad_pairs_off(npairs, n, in_cap, m) + 8 * (size_t)npairs + 255) & ~(size_t)255;
}
__host__ __device__ inline size_t quad_counters_off(uint32_t npairs, uint32_t n, uint32_t in_cap, uint32_t m) {
    return quad_slots_off(npairs, n, in_cap, m) + 48 * (size_t)kQuadSlots;
}
__host__ __device__ inline size_t quad_look_off(uint32_t npairs, uint32_t n, uint32_t in_cap, uint32_t m) {
    return (quad_counters_off(npairs, n, in_cap, m) + 4 * kCountWords + 255) & ~(size_t)255;
}
__host__ __device__ inline size_t quad_bytes(uint32_t npairs, uint32_t n, uint32_t in_cap, uint32_t m) {
    return quad_look_off(npairs, n, in_cap, m) + 2 * 8 * 16 * (size_t)kLookSlots;
}

enum class SegKernel { NttFirst, NttLayer, NttMulti2, NttMulti3, LeavesElem, LeavesPairDot, Nodes, Subtree, Fix, FoldFix,
                       Outer, Round };

namespace launch {

// Launch shapes of the plain path's NTT and hashing kernels, chosen per
// codeword size by the autotuner (tune.cpp).  Grid caps are in blocks; the
// kernels loop over their work with a grid stride, so every shape computes
// the same result.  ntt_waves selects the build of the three-layer pass: 2
// (at most 256 VGPRs) or 3 (at most 168) waves per SIMD.
struct Shape {
    int ntt_block = 256;
    unsigned ntt_cap = 256 * 8;
    int ntt_waves = 2;
    int hash_block = 256;
    unsigned hash_cap = 256 * 16;
};

// segmented launch over nseg segments occupying nblocks blocks in total;
// `partial` receives the inner-product sums of SegKernel::LeavesPairDot and holds the
// sums SegKernel::Round reduces (one block per segment)
hipError_t seg_launch(hipStream_t st, SegKernel which, const Seg *d_segs, int nseg, uint32_t nblocks, gf *partial);
int seg_block_size(SegKernel which);
// out[i] = rb (16 or 32) bytes from absolute device address src[i]
hipError_t gather_abs(hipStream_t st, const uint64_t *src, void *out, size_t n, int rb);
// every query record of njobs openings (records rec0 .. rec0 + q - 1 of job
// a; records in total); jobs sorted by rec0
hipError_t proof_records(hipStream_t st, const ProofJob *jobs, uint32_t njobs, uint32_t records);


// ft[j] = XOR over set bits p of j of bt[p]
hipError_t xor_span(hipStream_t st, gf *ft, const gf *bt, int bits, size_t count);
// RS encode: layer 0 (reads the 2^n evaluations, replicates into the c slots)
// then layers s = 1..n-1 in place on the slot-interleaved codeword
hipError_t ntt_first(hipStream_t st, const gf *data, gf *cw, const gf *ft0, int b, int n, const Shape &sh = Shape{});
hipError_t ntt_layer(hipStream_t st, gf *cw, const gf *fts, int s, int b, int n, const Shape &sh = Shape{});
// R = 2 or 3 layers s .. s+R-1 in one pass; t_r is the table of layer s + r
hipError_t ntt_multi(hipStream_t st, gf *cw, const gf *t0, const gf *t1, const gf *t2, int R, int s, int b, int n,
                     const Shape &sh = Shape{});
// width-w forms (row-major matrix, w elements per row): layer 0, and R = 1..3
// layers per pass on the codeword matrix
hipError_t ntt_first_w(hipStream_t st, const gf *data, gf *cw, const gf *ft0, int b, int n, size_t w,
                       const Shape &sh = Shape{});
hipError_t ntt_multi_w(hipStream_t st, gf *cw, const gf *t0, const gf *t1, const gf *t2, int R, int s, int b, int n,
                       size_t w, const Shape &sh = Shape{});
// leaf i = Keccak256 of the 16w bytes of row i
hipError_t leaves_rows(hipStream_t st, const gf *rows, size_t w, uint64_t *out, size_t count, const Shape &sh = Shape{});
// out[j] = sum_t scalars[t] * m[j*w + t]
hipError_t rowdot(hipStream_t st, const gf *m, size_t w, const gf *scalars, gf *out, size_t rows);
hipError_t leaves_elem(hipStream_t st, const gf *rows, uint64_t *out, size_t count, const Shape &sh = Shape{});
hipError_t leaves_pair(hipStream_t st, const gf *pi, uint64_t *out, size_t half, const Shape &sh = Shape{});
// one Merkle level: next[i] = H(prev[2i] || prev[2i+1])
hipError_t nodes(hipStream_t st, const uint64_t *prev, uint64_t *next, size_t count, const Shape &sh = Shape{});
// lv Merkle levels above the layer `prev` of h digests (h a power of two,
// 1 <= lv <= subtree_max_levels(), lv <= log2 h), written to the layers that
// follow it.  With cus (the device's compute units) > 0, levels narrow enough
// to leave SIMDs idle hash each node on a lane pair (shorter latency); 0
// keeps one lane per node.
hipError_t subtree(hipStream_t st, uint64_t *prev, size_t h, int lv, int cus = 0);
// largest lv subtree() and SegKernel::Subtree accept
int subtree_max_levels();
// Widest subtree level (nodes per block) hashed on lane pairs when a launch
// of `blocks` subtree blocks runs on `cus` compute units (the x.lo of a
// SegKernel::Subtree segment); 0 = none
int pair_width(unsigned blocks, int cus);
// partial[block] = sum of a[k] b[k]; with root set, the 32 bytes at root are
// also copied to partial - 2 (two elements in front of the sums)
hipError_t dot(hipStream_t st, const gf *a, const gf *b, size_t n, gf *partial, unsigned *nblocks,
               const uint64_t *root = nullptr);
// the number of blocks (partial sums) dot() uses for n elements
unsigned dot_blocks(size_t n);
// The last subtree step of a round's pi tree (as subtree()) with the round's
// inner product sum a[k] b[k], k < n (partial sums in partial[0, ndot) with
// ndot at most top_round_dot_blocks(n), and with cus given at most the
// blocks that stay resident beside the subtree blocks), and the round's
// transcript step (the SegKernel::Round work of *round, on the reduced sum)
// by whichever block finishes last.  *counter must be zero.  cus as for
// subtree().
hipError_t top_round(hipStream_t st, uint64_t *prev, size_t h, int lv, const gf *a, const gf *b, size_t n,
                     gf *partial, const Seg *round, unsigned *counter, int cus = 0);
unsigned top_round_dot_blocks(size_t n);
// fold (binary_reed_solomon.rs:88-96) and fix_last_var with the same
// challenge in one launch; the fix may be in place.  The challenge is read
// from device memory (*x), where the round's transcript kernel wrote it.
hipError_t fold_fix(hipStream_t st, const gf *w, gf *wout, const gf *ft, size_t mid, const gf *in, gf *out,
                    size_t half, const gf *x);
hipError_t fold(hipStream_t st, const gf *w, gf *out, const gf *ft, size_t mid, gf r);
hipError_t fix(hipStream_t st, const gf *in, gf *out, size_t half, gf x);
hipError_t outer(hipStream_t st, const gf *A, int m, const gf *B, gf *out, size_t n);
// Round k of a product sumcheck with its transcript step on the device.
// Round 0 reads f, g (2 * half entries each) from fsrc, gsrc.  Round k >= 1
// first folds the previous tables (4 * half entries at fsrc, gsrc) on their
// top variable with the previous round's challenge, read from the ProdFs
// region, into 2 * half entries at fdst, gdst (which may equal the sources).
// Pair j is weighted by the halved eq table entry of eq (half > 256 and
// 2^eq.lo >= 128, so a block's range of pairs lies in one B entry).  Every
// block XORs its pair sums into the round's accumulation slots; the block
// that finishes last (counted in the region) reduces them and runs the
// round's transcript step, which writes the round's polynomial and
// challenge into the region.  cus sizes the grid to fill the device.
hipError_t prod_step(hipStream_t st, const gf *fsrc, const gf *gsrc, gf *fdst, gf *gdst, EqFactors eq, size_t half,
                     int k, ProdFs *fs, int cus);
// The product sumcheck's rounds from k0 on (round k0 has half pairs, at most
// prod_tail_max_pairs(), a power of two), all in one single-block launch:
// the tables (4 * half entries at fsrc, gsrc; 2 * half in round 0) are
// loaded once into the block's LDS and folded there.  The last round's two
// entries of each table are written to fout, gout.  With host set, the
// kernel ends by copying the region's first host_bytes bytes (a multiple of
// 8) into that pinned host memory, in place of a copy after the launch.
hipError_t prod_tail(hipStream_t st, const gf *fsrc, const gf *gsrc, gf *fout, gf *gout, EqFactors eq, size_t half,
                     int k0, ProdFs *fs, uint8_t *host = nullptr, uint32_t host_bytes = 0);
size_t prod_tail_max_pairs();
// Rounds of at most prod_look_max_pairs() pairs (and more than
// prod_tail_max_pairs()) with one round of lookahead: every block folds its
// share of the previous tables into round k's and forms round k+1's
// coefficient sums; with first set (the first such round) round k's own
// sums come from the same products and the block that finishes last runs
// its transcript step, otherwise one more block runs round k's transcript
// step from the coefficient sums the previous launch left, while the others
// work.
hipError_t prod_look(hipStream_t st, const gf *fsrc, const gf *gsrc, gf *fdst, gf *gdst, EqFactors eq, size_t half,
                     int k, ProdFs *fs, bool first);
size_t prod_look_max_pairs();
// MLE value of every column of a rows x w matrix at a point whose eq table is
// A (2^m entries) x B: partial[col * nblocks + block], one pass over the data
hipError_t mle_dot(hipStream_t st, const gf *data, size_t w, const gf *A, int m, const gf *B, size_t rows, gf *partial,
                   unsigned max_blocks, unsigned *nblocks);
// Quadratic sumcheck: per pair (blockIdx.y) and block, the three sums
// A0, A1, A2 over halves of h entries at partial[3 * (pair * nblocks + block)]
hipError_t quad_sums(hipStream_t st, const gf *P, size_t stride, const uint32_t *pairs, uint32_t npairs, size_t h,
                     gf *partial, unsigned max_blocks, unsigned *nblocks);
// fix the top variable of m tables (stride apart) from 2*half to half entries
hipError_t fix_multi(hipStream_t st, gf *P, size_t stride, uint32_t m, size_t half, gf x);
// the same with the challenge read from device memory (*x), from 2 * half
// entries at src (tables src_stride apart) to half entries at dst (dst_stride
// apart; in place when dst == src)
hipError_t fix_multi_dev(hipStream_t st, const gf *src, size_t src_stride, gf *dst, size_t dst_stride, uint32_t m,
                         size_t half, const gf *x);
// Round k of a Quadratic sumcheck with its transcript step on the device:
// the three sums of every pair over halves of h entries (as quad_sums), each
// block's share of the round polynomial XORed into the region's slots, and
// in the block that finishes last (two-level count in the region's counters)
// the round's Fiat-Shamir step into the QuadFs region.  max_blocks caps the
// blocks per pair.
hipError_t quad_step(hipStream_t st, const gf *P, size_t stride, const uint32_t *pairs, uint32_t npairs, size_t h,
                     int k, QuadFs *fs, gf *slots, unsigned max_blocks, unsigned *counter);
// The Quadratic sumcheck's rounds from k0 on (round k0 has h0 pairs of
// entries per table, a power of two) in one single-block launch with the m
// tables (2 * h0 entries each at src, tables src_stride apart) in LDS; each
// table's final entry goes to fin (tables fin_stride apart).  h0 is at most
// quad_tail_max_half(m, npairs), which is 0 when the tables or the pairs do
// not fit.
size_t quad_tail_max_half(uint32_t m, uint32_t npairs);
// Rounds k >= 1 of at most quad_look_max_half() pairs of entries (and more
// than the single-block rounds take) with one round of lookahead: the m
// tables (4 half entries each at src, src_stride apart) are folded with
// x_{k-1} into round k's (2 half entries at dst, dst_stride apart, a
// different buffer), and every pair's coefficient sums of round k+1 are
// formed from the same entries; with first set round k's own sums come from
// the same products and the block that finishes last runs its transcript
// step, otherwise one more block runs it from the sums the previous launch
// left.  max_blocks caps the blocks per pair.
hipError_t quad_look(hipStream_t st, const gf *src, size_t src_stride, gf *dst, size_t dst_stride,
                     const uint32_t *pairs, uint32_t npairs, uint32_t m, size_t half, int k, QuadFs *fs, bool first);
size_t quad_look_max_half();
// host, host_bytes: as for prod_tail
hipError_t quad_tail(hipStream_t st, const gf *src, size_t src_stride, gf *fin, size_t fin_stride,
                     const uint32_t *pairs, uint32_t npairs, uint32_t m, size_t h0, int k0, QuadFs *fs,
                     uint8_t *host = nullptr, uint32_t host_bytes = 0);
// verifier kernels: fold chains (thread per proof x query) and Merkle paths
// (thread per proof x query x path, d + 1 paths)
hipError_t verify_fold(hipStream_t st, const VerifyJob *jobs, uint32_t count, uint32_t q, int b, int d, size_t w,
                       const gf *ft, gf *vals, uint8_t *fold_ok);
hipError_t verify_paths(hipStream_t st, const VerifyJob *jobs, uint32_t count, uint32_t q, int b, int d, size_t w,
                        const gf *vals, uint8_t *path_ok);
// out[i] = a[i] * b[i] (form 0), a[i] * b[0] through b[0]'s window table
// (form 1), or a[i] * b[i] with four lanes per product (form 2, gf_mul_wide)
hipError_t mul(hipStream_t st, const gf *a, const gf *b, gf *out, size_t n, int form);
hipError_t bench_mul(hipStream_t st, int variant, gf *io, size_t n, int iters);
hipError_t copy16(hipStream_t st, const gf *src, gf *dst, size_t n);
hipError_t bench_keccak(hipStream_t st, uint64_t *io, size_t n, int iters);
hipError_t bench_keccak_pair(hipStream_t st, uint64_t *io, int iters);
// p[i] = splitmix64 of (seed, i) for n 64-bit words
hipError_t fill_mix(hipStream_t st, uint64_t *p, size_t n, uint64_t seed);

}  // namespace launch
}  // namespace bfd
