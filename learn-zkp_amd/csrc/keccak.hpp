// Keccak-f[1600] and Keccak-256 (original padding 0x01, rate 136 bytes) for
// CDNA4.  This is the hash behind the reference's Merkle commitments and
// transcript: p3/src/merkle_tree.rs:8-21 (leaf = Keccak256(row bytes), node =
// Keccak256(left || right)) over tiny-keccak's Keccak::v256.
//
// Every 64-bit lane lives in two VGPRs; rotations become v_alignbit_b32
// pairs, chi becomes v_bfi_b32 (a ^ (~b & c) == bfi(b, a, a ^ c)).  The 25
// lanes are plain scalars so the compiler keeps the whole state in registers
// and renames away the pi permutation.
#pragma once
#include <stdint.h>
#include <hip/hip_runtime.h>

namespace kk {

__device__ __constant__ static const uint64_t RC[24] = {
    0x0000000000000001ull, 0x0000000000008082ull, 0x800000000000808Aull, 0x8000000080008000ull,
    0x000000000000808Bull, 0x0000000080000001ull, 0x8000000080008081ull, 0x8000000000008009ull,
    0x000000000000008Aull, 0x0000000000000088ull, 0x0000000080008009ull, 0x000000008000000Aull,
    0x000000008000808Bull, 0x800000000000008Bull, 0x8000000000008089ull, 0x8000000000008003ull,
    0x8000000000008002ull, 0x8000000000000080ull, 0x000000000000800Aull, 0x800000008000000Aull,
    0x8000000080008081ull, 0x8000000000008080ull, 0x0000000080000001ull, 0x8000000080008008ull};

__device__ __forceinline__ uint64_t rotl(uint64_t v, int r) {
    return (v << r) | (v >> (64 - r));
}
// State lanes s[x + 5y].  One full permutation, 24 rounds.
__device__ __forceinline__ void keccak_f(uint64_t s[25]) {
#pragma unroll 1
    for (int round = 0; round < 24; round++) {
        uint64_t c0 = s[0] ^ s[5] ^ s[10] ^ s[15] ^ s[20];
        uint64_t c1 = s[1] ^ s[6] ^ s[11] ^ s[16] ^ s[21];
        uint64_t c2 = s[2] ^ s[7] ^ s[12] ^ s[17] ^ s[22];
        uint64_t c3 = s[3] ^ s[8] ^ s[13] ^ s[18] ^ s[23];
        uint64_t c4 = s[4] ^ s[9] ^ s[14] ^ s[19] ^ s[24];
        uint64_t d0 = c4 ^ rotl(c1, 1);
        uint64_t d1 = c0 ^ rotl(c2, 1);
        uint64_t d2 = c1 ^ rotl(c3, 1);
        uint64_t d3 = c2 ^ rotl(c4, 1);
        uint64_t d4 = c3 ^ rotl(c0, 1);
        // theta + rho + pi: B[y + 5*((2x+3y)%5)] = rotl(A[x+5y] ^ D[x], r[x][y])
        uint64_t b0 = s[0] ^ d0;
        uint64_t b10 = rotl(s[1] ^ d1, 1);
        uint64_t b20 = rotl(s[2] ^ d2, 62);
        uint64_t b5 = rotl(s[3] ^ d3, 28);
        uint64_t b15 = rotl(s[4] ^ d4, 27);
        uint64_t b16 = rotl(s[5] ^ d0, 36);
        uint64_t b1 = rotl(s[6] ^ d1, 44);
        uint64_t b11 = rotl(s[7] ^ d2, 6);
        uint64_t b21 = rotl(s[8] ^ d3, 55);
        uint64_t b6 = rotl(s[9] ^ d4, 20);
        uint64_t b7 = rotl(s[10] ^ d0, 3);
        uint64_t b17 = rotl(s[11] ^ d1, 10);
        uint64_t b2 = rotl(s[12] ^ d2, 43);
        uint64_t b12 = rotl(s[13] ^ d3, 25);
        uint64_t b22 = rotl(s[14] ^ d4, 39);
        uint64_t b23 = rotl(s[15] ^ d0, 41);
        uint64_t b8 = rotl(s[16] ^ d1, 45);
        uint64_t b18 = rotl(s[17] ^ d2, 15);
        uint64_t b3 = rotl(s[18] ^ d3, 21);
        uint64_t b13 = rotl(s[19] ^ d4, 8);
        uint64_t b14 = rotl(s[20] ^ d0, 18);
        uint64_t b24 = rotl(s[21] ^ d1, 2);
        uint64_t b9 = rotl(s[22] ^ d2, 61);
        uint64_t b19 = rotl(s[23] ^ d3, 56);
        uint64_t b4 = rotl(s[24] ^ d4, 14);
        // chi
        s[0] = b0 ^ (~b1 & b2);
        s[1] = b1 ^ (~b2 & b3);
        s[2] = b2 ^ (~b3 & b4);
        s[3] = b3 ^ (~b4 & b0);
        s[4] = b4 ^ (~b0 & b1);
        s[5] = b5 ^ (~b6 & b7);
        s[6] = b6 ^ (~b7 & b8);
        s[7] = b7 ^ (~b8 & b9);
        s[8] = b8 ^ (~b9 & b5);
        s[9] = b9 ^ (~b5 & b6);
        s[10] = b10 ^ (~b11 & b12);
        s[11] = b11 ^ (~b12 & b13);
        s[12] = b12 ^ (~b13 & b14);
        s[13] = b13 ^ (~b14 & b10);
        s[14] = b14 ^ (~b10 & b11);
        s[15] = b15 ^ (~b16 & b17);
        s[16] = b16 ^ (~b17 & b18);
        s[17] = b17 ^ (~b18 & b19);
        s[18] = b18 ^ (~b19 & b15);
        s[19] = b19 ^ (~b15 & b16);
        s[20] = b20 ^ (~b21 & b22);
        s[21] = b21 ^ (~b22 & b23);
        s[22] = b22 ^ (~b23 & b24);
        s[23] = b23 ^ (~b24 & b20);
        s[24] = b24 ^ (~b20 & b21);
        s[0] ^= RC[round];
    }
}

// Keccak-256 of a message of NW 64-bit words (NW*8 < 136): one block.
// Writes the 4-word digest.
template <int NW>
__device__ __forceinline__ void keccak256_words(const uint64_t *m, uint64_t out[4]) {
    uint64_t s[25];
#pragma unroll
    for (int i = 0; i < 25; i++) s[i] = i < NW ? m[i] : 0ull;
    s[NW] ^= 0x01ull;                  // Keccak padding start (domain bit)
    s[16] ^= 0x8000000000000000ull;    // final bit of the 136-byte rate
    keccak_f(s);
    out[0] = s[0]; out[1] = s[1]; out[2] = s[2]; out[3] = s[3];
}

}  // namespace kk
