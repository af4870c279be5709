/*
 * bfo — CPU restatement of the reference's Basefold prover path over POLYVAL
 * GF(2^128) (han0110/learn-zkp).  TEST INFRASTRUCTURE ONLY: the parity tests,
 * the smoke check and the CPU-baseline leg of the benchmark call it as the
 * checker.  Nothing in the product library links or calls it.
 *
 * Every element is a 16-byte little-endian u128 holding the Montgomery value
 * (p3/src/polyval.rs:30-34, bincode layout).  All functions take raw byte
 * pointers so they bind from ctypes without u128 support.
 *
 * Pinning (which check fixes which part) is listed per function in bfo.c.
 */
#ifndef BFO_H
#define BFO_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

void bfo_set_threads(int n);
int bfo_get_threads(void);

/* ---- field (p3/src/polyval.rs, p3/src/polyval/{portable,x86_64_avx2}.rs) ---- */
/* impl: 0 = portable bmul64 restatement, 1 = PCLMUL restatement, 2 = textbook bit-serial */
void bfo_mul_batch(const uint8_t *a, const uint8_t *b, uint8_t *out, size_t n, int impl);
void bfo_from_canonical_batch(const uint8_t *a, uint8_t *out, size_t n);
void bfo_to_canonical_batch(const uint8_t *a, uint8_t *out, size_t n);
void bfo_inv_batch(const uint8_t *a, uint8_t *out, size_t n);
/* canonical (plain polynomial) product mod X^128+X^127+X^126+X^121+1, for pinning */
void bfo_polymul_mod_batch(const uint8_t *a, const uint8_t *b, uint8_t *out, size_t n);

/* ---- hashing ---- */
void bfo_keccak256(const uint8_t *msg, size_t len, uint8_t out[32]);
void bfo_sha3_256(const uint8_t *msg, size_t len, uint8_t out[32]);

/* ---- RNG: StdRng = ChaCha12, seed_from_u64 (rand_core 0.6 / rand_chacha 0.3) ---- */
void bfo_chacha_block(int rounds, const uint32_t key[8], const uint32_t ctr_nonce[4], uint32_t out[16]);
void bfo_random_vec(uint64_t seed, size_t n, uint8_t *out);

/* ---- additive NTT (p3/src/field/binary.rs) ---- */
/* twiddles layer K (2^K entries) at element offset 2^K - 1; total 2^log_n - 1 */
void bfo_ntt_twiddles(int log_n, uint8_t *out);
void bfo_ntt_forward(int log_n, uint8_t *data, size_t width);
void bfo_ntt_backward(int log_n, uint8_t *data, size_t width);
/* brute-force definition (binary.rs:156-191), data length 2^log_n, width 1 */
void bfo_ntt_forward_definition(int log_n, const uint8_t *data, uint8_t *out);

/* ---- Reed-Solomon code (poly-commit/src/basefold/code/binary_reed_solomon.rs) ---- */
size_t bfo_num_queries(int lambda, int log2_c);
void bfo_rs_encode(int log2_c, int d, const uint8_t *data, size_t width, uint8_t *out);
void bfo_rs_fold(int log2_c, int d, int i, const uint8_t *w, const uint8_t *r, uint8_t *out);

/* ---- multilinear (p3/src/poly/multilinear.rs) ---- */
void bfo_mle_evaluate(const uint8_t *evals, int n, const uint8_t *point, uint8_t out[16]);
void bfo_eq_table(const uint8_t *r, int n, uint8_t *out);

/* ---- Merkle (p3/src/merkle_tree.rs over plonky3 MerkleTreeMmcs) ---- */
void bfo_merkle_root(const uint8_t *rows, size_t height, size_t row_bytes, uint8_t root[32]);

/* ---- challenger (p3/src/challenger.rs over plonky3 HashChallenger) ---- */
typedef struct bfo_challenger bfo_challenger;
bfo_challenger *bfo_challenger_new(void);
void bfo_challenger_free(bfo_challenger *c);
void bfo_challenger_observe(bfo_challenger *c, const uint8_t *bytes, size_t len);
void bfo_challenger_sample(bfo_challenger *c, uint8_t out[16]);
uint64_t bfo_challenger_sample_bits(bfo_challenger *c, int bits);

/* ---- Basefold (poly-commit/src/basefold.rs) ---- */
/*
 * Full harness flow of poly-commit/src/lib.rs:64-105 for one polynomial
 * (width 1): commit, observe(root), claim = eval, observe(claim), open.
 * Writes root and claim; *proof receives a malloc'd bincode BasefoldProof.
 * If chal_state is non-NULL the challenger used for open is also returned
 * through it (caller frees).
 */
int bfo_basefold_prove(int lambda, int log2_c, int n, const uint8_t *evals,
                       const uint8_t *point, uint8_t root[32], uint8_t claim[16],
                       uint8_t **proof, size_t *proof_len);
/* returns 0 ok, 1 InvalidFinalPoly, 2 InvalidQuery, 3 Mmcs, 4 malformed proof */
int bfo_basefold_verify(int lambda, int log2_c, int n, const uint8_t root[32],
                        const uint8_t *point, const uint8_t claim[16],
                        const uint8_t *proof, size_t proof_len);
/*
 * The same for a row-major 2^n x width matrix of evaluations (one polynomial
 * per column): claims = width values (one per column), observed in column
 * order by the harness; width > 1 takes the random-scalars-by-eq branch of
 * open/verify (basefold.rs:166-183, 252-258, 398-408).
 */
int bfo_basefold_prove_w(int lambda, int log2_c, int n, size_t width, const uint8_t *evals,
                         const uint8_t *point, uint8_t root[32], uint8_t *claims,
                         uint8_t **proof, size_t *proof_len);
int bfo_basefold_verify_w(int lambda, int log2_c, int n, size_t width, const uint8_t root[32],
                          const uint8_t *point, const uint8_t *claims,
                          const uint8_t *proof, size_t proof_len);
/*
 * The same with a transcript that has already observed pre_len bytes `pre`
 * before the harness prefix (a caller's own earlier observations).
 */
int bfo_basefold_prove_px(int lambda, int log2_c, int n, size_t width, const uint8_t *evals,
                          const uint8_t *point, const uint8_t *pre, size_t pre_len, uint8_t root[32],
                          uint8_t *claims, uint8_t **proof, size_t *proof_len);
int bfo_basefold_verify_px(int lambda, int log2_c, int n, size_t width, const uint8_t root[32],
                           const uint8_t *point, const uint8_t *claims, const uint8_t *pre, size_t pre_len,
                           const uint8_t *proof, size_t proof_len);
void bfo_free(void *p);

/*
 * prove_sumcheck (sumcheck/src/lib.rs:56-103) over the Basefold sumcheck
 * function EvalProver (eq(point, .) * poly, sumcheck/src/function/eval.rs
 * over eval_impr.rs) with a fresh keccak challenger; claim is observed first.
 * Outputs the full round polynomials [c0, c1, c2] (n x 3), the challenges in
 * round order, the final table value and the final subclaim.  Returns -1 for
 * a point coordinate 0 or 1.
 */
int bfo_sumcheck_eval(int n, const uint8_t *poly, const uint8_t *point, const uint8_t claim[16],
                      uint8_t *round_polys, uint8_t *challenges, uint8_t final_eval[16], uint8_t subclaim[16]);

/*
 * Degree-3 product sumcheck  sum_b f(b) g(b) eq(r, b)  (no reference
 * counterpart; follows the prove_sumcheck driver of sumcheck/src/lib.rs:56-103).
 * Fresh keccak challenger; observes claim first.  Outputs n x 3 compressed
 * round polys [c0, c2, c3], the n challenges (in round order), the final
 * f and g evaluations and the final subclaim.
 */
int bfo_product_sumcheck(int n, const uint8_t *f, const uint8_t *g, const uint8_t *r,
                         uint8_t claim[16], uint8_t *round_polys, uint8_t *challenges,
                         uint8_t final_fg[32], uint8_t subclaim[16]);

/*
 * Quadratic sumcheck  sum_b sum_k s_k P_{a_k}(b) P_{b_k}(b)  over m tables of
 * 2^n evaluations (polys: m x 2^n, table-major), pair_idx = [a_0, b_0, a_1,
 * b_1, ...]; prove_sumcheck driver (sumcheck/src/lib.rs:56-103) with a fresh
 * keccak challenger.  Outputs the claim, n x 2 compressed round polys
 * [c0, c2], the n challenges, the m final evaluations and the subclaim.
 */
int bfo_sumcheck_quadratic(int n, int m, const uint8_t *polys, int npairs, const uint8_t *scalars,
                           const uint32_t *pair_idx, uint8_t claim[16], uint8_t *round_polys,
                           uint8_t *challenges, uint8_t *final_evals, uint8_t subclaim[16]);

#ifdef __cplusplus
}
#endif
#endif
