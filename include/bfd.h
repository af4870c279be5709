/*
 * bfd.h — C ABI of the MI355X Basefold prover (POLYVAL GF(2^128), binary
 * Reed-Solomon code, Keccak-256 Merkle trees).
 *
 * Drop-in boundary for the reference's Rust traits (han0110/learn-zkp):
 *   PolyCommitScheme<F, E>            poly-commit/src/lib.rs:20-51
 *   Basefold::{commit, eval, open}    poly-commit/src/basefold.rs:112-232
 *   BinaryReedSolomonCode::new        poly-commit/src/basefold/code/binary_reed_solomon.rs:19-27
 *   GenericChallenger<_, Keccak256>   p3/src/challenger.rs:24-90
 *   prove_sumcheck                    sumcheck/src/lib.rs:56-79
 *
 * Conventions
 *   - A field element is 16 bytes: the little-endian u128 Montgomery value of
 *     the reference's `Polyval` (p3/src/polyval.rs:30-34), i.e. its bincode
 *     bytes.  Digests are 32 bytes.
 *   - Every function returns 0 on success or a BFD_ERR_* code; the message of
 *     the last failure on the calling thread is bfd_last_error().  Nothing
 *     unwinds across the ABI.  (The reference panics where these return
 *     errors; commit/open never fail on valid input.)
 *   - Host buffers belong to the caller.  Handles own their device memory.
 *     bfd_open does not consume the commitment (the reference borrows
 *     CommitmentData immutably), so a commitment can be opened repeatedly.
 *   - A context is internally synchronised; calls on one context serialise.
 *   - Proofs are returned as the bincode bytes of BasefoldProof
 *     (basefold.rs:49-66) and freed with bfd_proof_free.
 */
#ifndef BFD_H
#define BFD_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define BFD_OK 0
#define BFD_ERR_ARG 1        /* invalid argument or unsupported shape */
#define BFD_ERR_HIP 2        /* HIP runtime / kernel failure, or no GPU */
#define BFD_ERR_BUDGET 3     /* admission refused: HBM budget exceeded */
#define BFD_ERR_STATE 4      /* handle used in the wrong state */

typedef struct bfd_ctx bfd_ctx;
typedef struct bfd_code bfd_code;
typedef struct bfd_commitment bfd_commitment;
typedef struct bfd_challenger bfd_challenger;

const char *bfd_last_error(void);

/* ---- context: one GPU, one stream, an HBM budget for resident commitments ---- */
int bfd_ctx_create(int device, uint64_t hbm_budget_bytes, bfd_ctx **out);
int bfd_ctx_destroy(bfd_ctx *ctx);
int bfd_ctx_synchronize(bfd_ctx *ctx);
/* bytes currently held by commitments of this context */
uint64_t bfd_ctx_resident_bytes(const bfd_ctx *ctx);

/* ---- code: BinaryReedSolomonCode::new(lambda, log2_c, d)  (binary_reed_solomon.rs:19-27)
 * Builds the fold/NTT twiddle table on the device (setup, like AddtiveNtt::new). */
int bfd_code_create(bfd_ctx *ctx, uint32_t lambda, uint32_t log2_c, uint32_t d, bfd_code **out);
int bfd_code_destroy(bfd_code *code);
/* RandomFoldableCode::num_queries (binary_reed_solomon.rs:55-59) */
uint64_t bfd_code_num_queries(const bfd_code *code);

/* ---- challenger: GenericChallenger<Polyval, Keccak256Hash>  (p3/src/challenger.rs) ---- */
int bfd_challenger_create(bfd_challenger **out);            /* ::keccak256(), :45-49 */
int bfd_challenger_clone(const bfd_challenger *c, bfd_challenger **out);
int bfd_challenger_destroy(bfd_challenger *c);
int bfd_challenger_observe(bfd_challenger *c, const uint8_t *bytes, size_t len); /* CanObserve, :51-60 */
int bfd_challenger_sample(bfd_challenger *c, uint8_t out[16]);                     /* CanSample, :62-74 */
int bfd_challenger_sample_bits(bfd_challenger *c, uint32_t bits, uint64_t *out);   /* :76-83 */
/* The challenger's state as the two byte buffers of plonky3's
 * HashChallenger<u8, _, 32> (input_buffer, output_buffer; the bytes still to
 * be sampled are output[0, output_len), taken from the end), so a transcript
 * can cross the boundary either way.  export writes *input_len bytes when
 * input_cap suffices (else BFD_ERR_ARG with *input_len set); import makes a
 * new challenger in that state. */
int bfd_challenger_export(const bfd_challenger *c, uint8_t *input, size_t input_cap, size_t *input_len,
                          uint8_t output[32], size_t *output_len);
int bfd_challenger_import(const uint8_t *input, size_t input_len, const uint8_t *output, size_t output_len,
                          bfd_challenger **out);

/* ---- PolyCommitScheme ---- */
/* Basefold::commit (basefold.rs:112-120): evals is the height x width
 * row-major matrix of evaluations, one polynomial per column (width 1 ..
 * 65535), height = 2^d; each Merkle leaf hashes one 16*width-byte row. */
int bfd_commit(bfd_ctx *ctx, const bfd_code *code, const uint8_t *evals, size_t height, size_t width,
               bfd_commitment **out, uint8_t root[32]);
/* Same, with the evaluations already resident in device memory (not copied
 * out of; the commitment keeps its own copy of the data). */
int bfd_commit_device(bfd_ctx *ctx, const bfd_code *code, const void *d_evals, size_t height, size_t width,
                      bfd_commitment **out, uint8_t root[32]);
/* Basefold::commit followed by Basefold::eval at `point` (d elements) in one
 * call: values (width x 16 bytes) = the MLE of every column, computed in one
 * pass over the evaluations (sum of eq(point, x) * data[x]) in the same
 * stream pass as the commitment, one synchronisation for both.  No reference
 * counterpart as one call (basefold.rs:112-131 are two). */
int bfd_commit_eval(bfd_ctx *ctx, const bfd_code *code, const void *evals, int evals_on_device, size_t height,
                    size_t width, const uint8_t *point, bfd_commitment **out, uint8_t root[32], uint8_t *values);
int bfd_commitment_root(const bfd_commitment *comm, uint8_t root[32]);
int bfd_commitment_destroy(bfd_commitment *comm);

/* Basefold::eval (basefold.rs:126-131): values[col] = MLE(col)(point) for
 * every column (width x 16 bytes), point has d elements */
int bfd_eval(bfd_ctx *ctx, const bfd_commitment *comm, const uint8_t *point, uint8_t *values);

/* Basefold::eval (basefold.rs:126-131) of data that has no commitment: the
 * trait's eval takes &Self::Data, a height x width row-major matrix (host,
 * or device memory with evals_on_device; height = 2^n, n = the point's
 * length).  values[col] = MLE(col)(point), width x 16 bytes. */
int bfd_eval_data(bfd_ctx *ctx, const void *evals, int evals_on_device, size_t height, size_t width,
                  const uint8_t *point, uint8_t *values);

/* bfd_eval_data on device-resident evaluations that allocates nothing.  No
 * counterpart in the reference: it is for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  Its working memory is a fixed 4 MiB
 * device array per GPU, held until the process ends, so height is at most
 * 2^32 rows (width 1..65535 as for bfd_eval_data; larger is BFD_ERR_ARG).
 * Calls that overlap on one GPU share that array and race; calls on one
 * context serialise.  values: width x 16 bytes in host memory. */
int bfd_eval_data_noalloc(bfd_ctx *ctx, const void *d_evals, size_t height, size_t width, const uint8_t *point,
                          uint8_t *values);

/* Basefold::open (basefold.rs:133-232) for one PolyEvals {point, values}
 * (values: width x 16 bytes).  Width > 1 first samples the random scalars by
 * eq from the challenger (basefold.rs:166-183).  The challenger is advanced
 * in place (the reference takes it by value: pass a clone to keep the
 * original).  *proof is allocated by the library. */
int bfd_open(bfd_ctx *ctx, const bfd_code *code, const bfd_commitment *comm, const uint8_t *point,
             const uint8_t *values, bfd_challenger *challenger, uint8_t **proof, size_t *proof_len);
/* Same, writing into a caller buffer of `capacity` bytes (at least
 * bfd_proof_size_width); *proof_len receives the size written.  A smaller
 * buffer is refused (BFD_ERR_ARG) before the challenger is touched. */
int bfd_open_into(bfd_ctx *ctx, const bfd_code *code, const bfd_commitment *comm, const uint8_t *point,
                  const uint8_t *values, bfd_challenger *challenger, uint8_t *proof, size_t capacity,
                  size_t *proof_len);
void bfd_proof_free(uint8_t *proof);

/* ---- Basefold::open driven by the caller's own transcript ----
 * The trait's open (poly-commit/src/lib.rs:37-42) takes any
 * `impl FieldChallenger<F> + CanObserve<Commitment>`, whose state need not
 * be exportable (plonky3's HashChallenger keeps its buffers private).  These
 * calls run the same opening with every transcript step on the caller's
 * side, in the reference's order (basefold.rs:157-203, 335-369):
 *   begin      width > 1: the caller first samples ceil(log2 width) scalar
 *              challenges (sample_ext_vec, basefold.rs:398-408) and passes
 *              them; width 1 passes none
 *   d times:   bfd_open_round -> pi-tree root, compressed round poly
 *              [c0, c2]; the caller observes the root, then c0 and c2, and
 *              samples r; bfd_open_challenge(r)
 *   bfd_open_final_poly -> final_poly; the caller observes it and samples
 *              num_queries indices with sample_bits(log2_c + d)
 *   bfd_open_finish(indices) -> the bincode proof (free with bfd_proof_free)
 * The proof equals bfd_open's when the caller's transcript is a
 * GenericChallenger<_, Keccak256>.  The opening borrows the commitment and
 * the code until bfd_opening_destroy.  A call out of order is BFD_ERR_STATE.
 * Each round synchronises once; bfd_open, with the transcript on the device,
 * is the faster form for a challenger the library can take over. */
typedef struct bfd_opening bfd_opening;
int bfd_open_begin(bfd_ctx *ctx, const bfd_code *code, const bfd_commitment *comm, const uint8_t *point,
                   const uint8_t *values, const uint8_t *scalar_challenges, size_t num_scalar_challenges,
                   bfd_opening **out);
int bfd_open_rounds_left(const bfd_opening *op, uint32_t *rounds);
int bfd_open_round(bfd_opening *op, uint8_t root[32], uint8_t round_poly[32]);
int bfd_open_challenge(bfd_opening *op, const uint8_t x[16]);
int bfd_open_final_poly(bfd_opening *op, uint8_t out[16]);
int bfd_open_finish(bfd_opening *op, const uint64_t *indices, size_t count, uint8_t **proof, size_t *proof_len);
int bfd_opening_destroy(bfd_opening *op);
/* Exact bincode size of the proof for (lambda, log2_c, d), width 1, and for
 * a commitment of `width` columns (each opening carries a whole row). */
uint64_t bfd_proof_size(uint32_t lambda, uint32_t log2_c, uint32_t d);
uint64_t bfd_proof_size_width(uint32_t lambda, uint32_t log2_c, uint32_t d, uint64_t width);

/* ---- batched commit+open of independent polynomials (uneven sizes) ----
 * The reference commits one matrix height per Basefold instance
 * (basefold.rs:94 TODO), so a batch is `count` independent Basefold
 * commit+open runs, each with its own challenger; every proof is byte-equal
 * to the plain bfd_commit + bfd_open proof of the same polynomial.
 * evals[p]: 2^log_sizes[p] evaluations (device pointers if evals_on_device);
 * points[p]: log_sizes[p] coordinates; claims[p]: the evaluation at the point.
 * With observe_prefix set, each challenger first observes the root and the
 * claim (the harness order of poly-commit/src/lib.rs:77-89).  Every
 * polynomial needs its own challenger (BFD_ERR_ARG if two entries alias).
 * Outputs count x 32 roots and count proofs: a non-NULL proofs[p] on input is
 * a caller buffer of proof_lens[p] bytes (at least bfd_proof_size), otherwise
 * the library allocates it (free with bfd_proof_free).  A call with a smaller
 * buffer is refused (BFD_ERR_ARG) before any work, every challenger as it was. */
int bfd_commit_open_batch(bfd_ctx *ctx, uint32_t lambda, uint32_t log2_c, size_t count,
                          const uint32_t *log_sizes, const void *const *evals, int evals_on_device,
                          const uint8_t *const *points, const uint8_t *const *claims,
                          bfd_challenger *const *challengers, int observe_prefix, uint8_t *roots,
                          uint8_t **proofs, size_t *proof_lens);

/* ---- serving: continuous batching of commit-then-open requests ----
 * A worker thread runs commit ticks and open-round ticks on two streams; a
 * request is admitted when its resident footprint fits hbm_budget_bytes
 * (0 = unlimited) and at most max_active (0 = unlimited) are in flight.
 * Each request: evaluations already in device memory (kept valid until its
 * proof is collected), n point coordinates; the claim is evaluated on the
 * device and a fresh challenger observes root and claim before opening, so
 * the proof equals a plain commit + eval + open of the same polynomial. */
typedef struct bfd_server bfd_server;
int bfd_server_create(bfd_ctx *ctx, uint64_t hbm_budget_bytes, uint32_t max_active, bfd_server **out);
/* BFD_ERR_BUDGET if the request alone exceeds the budget */
int bfd_server_submit(bfd_server *srv, uint32_t lambda, uint32_t log2_c, uint32_t n, const void *d_evals,
                      const uint8_t *point, uint64_t *ticket);
/* Waits up to timeout_ms (< 0: forever; BFD_ERR_STATE on timeout) and hands
 * over the result: root, claim, proof (free with bfd_proof_free) and the
 * submit-to-done latency. */
int bfd_server_wait(bfd_server *srv, uint64_t ticket, int64_t timeout_ms, uint8_t root[32], uint8_t claim[16],
                    uint8_t **proof, size_t *proof_len, double *latency_ms);
int bfd_server_stats(bfd_server *srv, uint64_t *pending, uint64_t *active, uint64_t *resident_bytes,
                     uint64_t *completed);
/* out: commit ticks, open ticks, commit-lane busy ms, open-lane busy ms,
 * host ms in tick finishes, host ms enqueueing ticks, job-rounds in open ticks */
int bfd_server_timing(bfd_server *srv, double out[8]);
/* finishes every admitted and pending request, then stops */
int bfd_server_destroy(bfd_server *srv);

/* ---- degree-3 product sumcheck  sum_b f(b) g(b) eq(r, b)  (no reference
 * counterpart; driver semantics of prove_sumcheck, sumcheck/src/lib.rs:56-103):
 * observes the claim, then per round observes [c0, c2, c3] and samples the
 * challenge; tables are folded on the top variable first.
 * Outputs: claim, n x 3 compressed round polys, n challenges (round order),
 * final [f, g] evaluations and the final subclaim. */
int bfd_sumcheck_product(bfd_ctx *ctx, uint32_t n, const uint8_t *f, const uint8_t *g, const uint8_t *r,
                         bfd_challenger *challenger, uint8_t claim[16], uint8_t *round_polys,
                         uint8_t *challenges, uint8_t final_fg[32], uint8_t subclaim[16]);
/* Same, with f and g (2^n elements each) already in device memory; they are
 * copied into the working tables, not modified. */
int bfd_sumcheck_product_device(bfd_ctx *ctx, uint32_t n, const void *d_f, const void *d_g, const uint8_t *r,
                                bfd_challenger *challenger, uint8_t claim[16], uint8_t *round_polys,
                                uint8_t *challenges, uint8_t final_fg[32], uint8_t subclaim[16]);

/* ---- Quadratic sumcheck  sum_b sum_k s_k P_{a_k}(b) P_{b_k}(b)
 * (sumcheck/src/function/quadratic.rs:13-134) driven by prove_sumcheck
 * (sumcheck/src/lib.rs:56-103) over m tables of 2^n evaluations (polys:
 * m x 2^n x 16 bytes, table-major), npairs pairs (pair_idx = a_0, b_0, a_1,
 * b_1, ...) with one scalar each.  The claim is the function's sum
 * (compute_sum); it is observed first, then [c0, c2] per round.  Outputs
 * claim, n x 2 compressed round polys, n challenges (round order), the m
 * final evaluations and the final subclaim.  A Batch (sumcheck/src/function/
 * batch.rs) of Quadratics with equal num_vars is this call on the union of
 * their tables with the pairs' scalars multiplied by alpha^k. ---- */
int bfd_sumcheck_quadratic(bfd_ctx *ctx, uint32_t n, uint32_t m, const uint8_t *polys, uint32_t npairs,
                           const uint8_t *scalars, const uint32_t *pair_idx, bfd_challenger *challenger,
                           uint8_t claim[16], uint8_t *round_polys, uint8_t *challenges, uint8_t *final_evals,
                           uint8_t subclaim[16]);
/* Same, with the m tables already in device memory (table-major, m x 2^n
 * elements); they are copied into the working tables, not modified. */
int bfd_sumcheck_quadratic_device(bfd_ctx *ctx, uint32_t n, uint32_t m, const void *d_polys, uint32_t npairs,
                                  const uint8_t *scalars, const uint32_t *pair_idx, bfd_challenger *challenger,
                                  uint8_t claim[16], uint8_t *round_polys, uint8_t *challenges,
                                  uint8_t *final_evals, uint8_t subclaim[16]);

/* ---- RandomFoldableCode<Polyval> for the binary RS code (basefold/code.rs:
 * 16-84, code/binary_reed_solomon.rs:38-111) on a code's device tables.
 * bfd_code_encode = encode_batch: the 2^d x width row-major message (k_d rows)
 * to the 2^(log2_c + d) x width codeword in the reference's row order
 * (bit-reversed coset NTTs, :67-82); host or device buffers, which may not
 * overlap.  bfd_code_fold = fold(i, w, r): w of 2^(log2_c + i + 1) elements
 * to its first 2^(log2_c + i) (on the device, out may equal w; on the host,
 * out receives the folded half).  bfd_code_interpolate(i, j, lo, hi, r) is
 * interpolate (:103-109) at position j of level i, on the host.  encode0 is
 * c copies of the message (:61-65) and needs no call. ---- */
int bfd_code_encode(bfd_ctx *ctx, const bfd_code *code, const void *m, int m_on_device, size_t height, size_t width,
                    void *out, int out_on_device);
int bfd_code_fold(bfd_ctx *ctx, const bfd_code *code, uint32_t i, const void *w, int on_device, const uint8_t r[16],
                  void *out);
int bfd_code_interpolate(const bfd_code *code, uint32_t i, uint64_t j, const uint8_t lo[16], const uint8_t hi[16],
                         const uint8_t r[16], uint8_t out[16]);

/* ---- QuadraticProver, step by step (sumcheck/src/function/quadratic.rs:
 * 31-136): the SumcheckFunctionProver a caller's own prove_sumcheck driver
 * (sumcheck/src/lib.rs:56-103) calls once per round, with the tables
 * resident in device memory.  Same tables, pairs and scalars as
 * bfd_sumcheck_quadratic; polys is on the host or, with polys_on_device, in
 * device memory (copied either way).
 *   compute_sum        -> bfd_sc_compute_sum        (quadratic.rs:86-99)
 *   compute_round_poly -> bfd_sc_round_poly: [c0, c1, c2] with c1 from the
 *                         subclaim as decompress_round_poly does (:101-124)
 *   fix_last_var       -> bfd_sc_fix_last_var       (:126-130)
 *   evaluations        -> bfd_sc_evaluations: the m values once every
 *                         variable is fixed, else BFD_ERR_STATE (None) (:132-135)
 * compress_round_poly ([c0, c2]) and decompress_round_poly stay with the
 * caller; they touch no table. ---- */
typedef struct bfd_sc bfd_sc;
int bfd_sc_quadratic_create(bfd_ctx *ctx, uint32_t n, uint32_t m, const void *polys, int polys_on_device,
                            uint32_t npairs, const uint8_t *scalars, const uint32_t *pair_idx, bfd_sc **out);
/* EvalProver::new(Eval::new(point), MultiPoly::base(poly))
 * (sumcheck/src/function/eval.rs:27-47 over eval_impr.rs:13-207): the
 * function eq(point, .) * poly over n variables, the Basefold sumcheck
 * function.  poly: 2^n evaluations (host or device, copied); point: n
 * coordinates, none 0 or 1 (r and 1 - r are inverted, eval_impr.rs:20-27,
 * 69-79).  On this handle bfd_sc_round_poly returns the degree-2 round
 * polynomial [c0, c1, c2] of eval.rs:90-104 (the subclaim argument is used
 * in the first round; later rounds use the prover's inner subclaim, as the
 * reference does), fix_last_var also reduces the inner subclaim, and
 * evaluations gives the one table value.  compute_sum is eval.rs:86-88;
 * once every variable is fixed (and n > 0) it is BFD_ERR_STATE. */
int bfd_sc_eval_create(bfd_ctx *ctx, uint32_t n, const void *poly, int poly_on_device, const uint8_t *point,
                       bfd_sc **out);
int bfd_sc_num_vars(const bfd_sc *h, uint32_t *num_vars);  /* variables left */
int bfd_sc_compute_sum(bfd_sc *h, uint8_t out[16]);
int bfd_sc_round_poly(bfd_sc *h, const uint8_t subclaim[16], uint8_t coeffs[48]);
int bfd_sc_fix_last_var(bfd_sc *h, const uint8_t x[16]);
int bfd_sc_evaluations(bfd_sc *h, uint8_t *out);
int bfd_sc_destroy(bfd_sc *h);

/* ---- Basefold::verify (basefold.rs:234-325) of `count` proofs made with
 * `code`, each over a commitment of `width` columns: roots (count x 32
 * bytes), points[p] (d elements), values[p] (width elements), proofs[p] of
 * proof_lens[p] bytes, challengers[p] in the state open started from (the
 * harness prefix observed; advanced in place; BFD_ERR_ARG if two entries
 * alias).  The proof bytes are decoded
 * and the transcripts replayed on the host; every query's fold chain and
 * Merkle paths of all proofs are checked on the device in two launches.
 * results[p]: 0 accepted, 1 InvalidFinalPoly, 2 InvalidQuery, 3 Mmcs (a
 * Merkle path), 4 malformed proof bytes (the reference's first error in
 * query order). ---- */
int bfd_verify_batch(bfd_ctx *ctx, const bfd_code *code, size_t width, size_t count, const uint8_t *roots,
                     const uint8_t *const *points, const uint8_t *const *values, const uint8_t *const *proofs,
                     const size_t *proof_lens, bfd_challenger *const *challengers, int *results);
int bfd_verify(bfd_ctx *ctx, const bfd_code *code, size_t width, const uint8_t root[32], const uint8_t *point,
               const uint8_t *values, const uint8_t *proof, size_t proof_len, bfd_challenger *challenger,
               int *result);

/* ---- Polyval arithmetic on the host (p3/src/polyval.rs:30-252):
 * Montgomery-form product a*b*X^-128 (Mul) and inverse (invert_or_zero,
 * 0 for 0), little-endian 16-byte elements ---- */
int bfd_field_mul(const uint8_t a[16], const uint8_t b[16], uint8_t out[16]);
/* The same product on the device, in each of the forms the kernels use:
 * form 0, out[i] = a[i] * b[i] in registers; form 1, out[i] = a[i] * b[0]
 * through b[0]'s window table (the products by a round challenge); form 2,
 * out[i] = a[i] * b[i] with four lanes per product (the transcript steps'
 * serial products).  a, out: n elements; b: n (1 with form 1); host
 * buffers, or device memory with on_device.  Other forms: BFD_ERR_ARG. */
int bfd_field_mul_batch(bfd_ctx *ctx, const void *a, const void *b, int form, void *out, size_t n,
                        int on_device);
int bfd_field_inv(const uint8_t a[16], uint8_t out[16]);

/* ---- input generation: Polyval::random_vec(n, StdRng::seed_from_u64(seed))
 * (p3/src/field/from_uniform_bytes.rs:5-27) ---- */
int bfd_random_vec(uint64_t seed, size_t n, uint8_t *out);

/* ---- launch shapes and their autotuner.  No reference counterpart: the
 * reference runs on the CPU, where there is nothing to shape.  A shape
 * changes only how the work is cut into launches, never a result: every
 * shape gives the same commitments and proofs.  The context keeps one shape
 * per codeword size 2^log2_m (log2_m = d + log2_c, 0..47) for the plain path
 * and one (BFD_TUNING_BATCH) for the batched and serving engine. ---- */
#define BFD_TUNING_BATCH 0xfffffffeu /* log2_m selector of the engine's shape */
#define BFD_TUNING_ALL 0xffffffffu   /* bfd_tuning_set only: every plain-path size */
typedef struct bfd_tuning {
    uint32_t ntt_order;         /* NTT pass plan over layers 1..n-1: 0 three layers per pass, the shorter pass
                                   last; 1 the shorter pass first; 2 two layers per pass */
    uint32_t ntt_block;         /* threads per block of the NTT passes: 64, 128 or 256 */
    uint32_t ntt_waves_per_cu;  /* grid cap of the NTT passes, waves per compute unit: 4..1024 */
    uint32_t ntt_occupancy;     /* build of the three-layer pass: 2 or 3 waves per SIMD */
    uint32_t hash_block;        /* threads per block of the leaf and node kernels: 64, 128 or 256 */
    uint32_t hash_waves_per_cu; /* grid cap of the leaf and node kernels, waves per compute unit: 4..1024 */
    uint32_t subtree_from_log2; /* Merkle levels whose input holds more than 2^this digests get a launch each
                                   (1..40); the levels above run subtree_levels per launch through LDS.  The
                                   engine divides the threshold by the number of trees hashed together */
    uint32_t subtree_levels;    /* 1..9 */
    uint32_t seg_blocks_per_cu; /* engine only: most blocks one polynomial takes in a segmented launch, per
                                   compute unit: 1..64 */
    uint32_t reserved[7];       /* zero */
} bfd_tuning;
/* the built-in shape */
int bfd_tuning_default(bfd_tuning *out);
int bfd_tuning_get(const bfd_ctx *ctx, uint32_t log2_m, bfd_tuning *out);
int bfd_tuning_set(bfd_ctx *ctx, uint32_t log2_m, const bfd_tuning *t);
/* Tunes the plain path for `code`'s shape: times commit + open of a
 * synthetic polynomial of 2^d evaluations under candidate shapes, one field
 * at a time from the current shape (median of `reps` runs each; a candidate
 * must win by more than 0.5%), installs the fastest for log2_m = d + log2_c
 * and returns it with the medians of the built-in shape and of the chosen
 * one, re-measured alternately at the end. */
int bfd_tune(bfd_ctx *ctx, const bfd_code *code, uint32_t reps, bfd_tuning *chosen, double *ms_default,
             double *ms_tuned);
/* The same for the engine's shape, timing bfd_commit_open_batch on
 * synthetic polynomials of the given log2 sizes at rate 1/2^log2_c. */
int bfd_tune_batch(bfd_ctx *ctx, uint32_t lambda, uint32_t log2_c, size_t count, const uint32_t *log_sizes,
                   uint32_t reps, bfd_tuning *chosen, double *ms_default, double *ms_tuned);
/* Text file of the shapes that differ from the built-in one, one line per
 * size ("log2_m=26 ntt_order=0 ..."; "batch" for the engine's); load
 * applies every line of a file that save wrote. */
int bfd_tuning_save(const bfd_ctx *ctx, const char *path);
int bfd_tuning_load(bfd_ctx *ctx, const char *path);

/* ---- measurement helpers ---- */
/* Times `iters` dependent multiplies (variant 0: LDS comb, 1: register
 * Karatsuba, 4: by one shared factor through its LDS window table, 6: four
 * lanes per product, lanes 0..15 of each wave chaining), `iters` chained
 * Keccak-256 hashes of 64-byte messages (2, the Merkle node hash) on n
 * lanes, one copy of n 16-byte elements (3), or `iters` chained hashes on
 * one lane pair of one wave (5, the transcript step's form; n ignored);
 * returns milliseconds.  With n = 64 variants 1, 2 and 6 run on one wave,
 * so the time per iteration is the latency of one product or hash.  No
 * reference counterpart (measurement only). */
int bfd_microbench(bfd_ctx *ctx, int what, size_t n, int iters, double *ms);
/* Per-kernel timing of the last commit/open on this context: fills up to
 * `cap` (name, total ms, launches) records; returns the count in *count. */
int bfd_ctx_kernel_times(bfd_ctx *ctx, int enable, char *names, size_t name_len, double *ms,
                         uint32_t *launches, size_t cap, size_t *count);
/* Restricts that timing to launches of one kernel name (NULL or "": all).
 * An event costs the stream about 5 microseconds, so with a name set one
 * pair of events brackets each run of consecutive launches of that kernel
 * (closed by the next launch of another kernel); a run still open at a
 * synchronisation is not timed and is reported as "untimed:<name>". */
int bfd_ctx_kernel_filter(bfd_ctx *ctx, const char *only);

#ifdef __cplusplus
}
#endif
#endif
