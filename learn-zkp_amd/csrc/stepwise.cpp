// Basefold::open driven round by round by the caller's own transcript, and
// Basefold::eval on data that has no commitment.
//
// The reference's open (poly-commit/src/basefold.rs:133-232) takes any
// `impl FieldChallenger<F> + CanObserve<Commitment>`.  bfd_open runs the
// transcript of GenericChallenger<_, Keccak256> on the device, which needs
// the challenger's byte buffers; a challenger whose state the caller cannot
// export (plonky3's HashChallenger keeps its buffers private) drives the
// opening through these calls instead.  Per round the library returns the
// pi-tree root and the compressed round polynomial and takes the sampled
// challenge back; after the last round it returns the final polynomial and
// takes the sampled query indices.  The table work stays on the device; each
// round costs one synchronisation.
#include "internal.hpp"

using namespace bfd::rt;

struct bfd_opening {
    bfd_ctx *ctx = nullptr;
    const bfd_code *code = nullptr;
    const bfd_commitment *cm = nullptr;
    OpenRounds o;
    OpenLayout lay;
    DevBuf comb, half_eq, poly, pis, trees, partials, xdev;
    PinnedBuf pin;  // round read-backs (root, inner-product sums) and the challenge upload
    const gf *poly_cur = nullptr;
    OpenRounds::RoundPoly pending{};
    int k = 0;                // rounds completed
    bool awaiting_challenge = false;
    bool final_read = false;
};

namespace {

constexpr size_t kPinX = 0;                  // 16-byte challenge staged for upload
constexpr size_t kPinRound = 64;             // root (2 elements) + inner-product block sums
constexpr size_t kPinBytes = kPinRound + 16 * (2 + 512);

void begin(bfd_ctx *ctx, const bfd_code *code, const bfd_commitment *cm, const uint8_t *point_b,
           const uint8_t *values_b, const uint8_t *scal_b, size_t nscal, bfd_opening **out) {
    if (cm->code != code) fail(BFD_ERR_ARG, "commitment was made with another code");
    if (cm->ctx->device != ctx->device) fail(BFD_ERR_ARG, "the commitment lives on another device");
    const size_t w = cm->w;
    if (nscal != scalar_challenges(w))
        fail(BFD_ERR_ARG, "a width-w opening takes ceil(log2 w) scalar challenges (none for width 1)");
    std::lock_guard<std::recursive_mutex> lk(ctx->mu);
    set_device(ctx);
    hipStream_t st = ctx->stream;
    auto op = std::make_unique<bfd_opening>();
    op->ctx = ctx;
    op->code = code;
    op->cm = cm;
    uint8_t claim_b[16];
    const uint8_t *claim_p = values_b;
    const gf *poly0 = cm->data.as<gf>(), *pi0 = cm->cw.as<gf>();
    if (w > 1) {
        std::vector<gf> rs(nscal);
        for (size_t i = 0; i < nscal; i++) rs[i] = load_elem(scal_b + 16 * i);
        combine_columns(ctx, cm, rs, values_b, claim_b, op->comb, poly0, pi0);
        claim_p = claim_b;
    }
    OpenRounds &o = op->o;
    o.init((int)cm->n, (int)cm->b, code->q, point_b, claim_p);
    o.w = w;
    const size_t M = cm->M, c = (size_t)1 << o.b;
    const int hn = o.half_eq_vars();
    op->half_eq = ctx_buf(ctx, 16 * ((size_t)1 << hn));
    if (o.n > 0) device_eq(ctx, o.r.data(), hn, op->half_eq.as<gf>());
    op->poly = ctx_buf(ctx, 16 * (cm->N > 1 ? cm->N / 2 : 1));
    op->lay.plan(M, o.d);
    op->pis = ctx_buf(ctx, 16 * op->lay.pis_elems(M, c));
    op->trees = ctx_buf(ctx, 32 * op->lay.pitree_digests());
    op->lay.cw = pi0;
    if (w > 1) {
        op->lay.rows = cm->cw.as<gf>();
        op->lay.w = w;
    }
    op->lay.tree = cm->tree.as<uint64_t>();
    op->lay.pis = op->pis.as<gf>();
    op->lay.pitrees = op->trees.as<uint64_t>();
    op->partials = ctx_buf(ctx, 16 * (2 + 512));  // root in front of the sums (k_dot)
    op->xdev = ctx_buf(ctx, 16);
    op->pin.ensure(kPinBytes);
    op->poly_cur = poly0;
    // the staging buffers of device_eq and combine_columns are idle again
    // when the call returns
    hip_check(hipStreamSynchronize(st), "opening set-up");
    ctx->timer.collect();
    *out = op.release();
}

// Round k: the pi tree over rows (pi[j], pi[j + half]) (basefold.rs:350-355)
// and the inner product of the halved eq table with the low half of the
// polynomial (eval_impr.rs:187-197), read back together; the round
// polynomial follows on the host (eval.rs:84-104).
void round(bfd_opening *op, uint8_t root_out[32], uint8_t poly_out[32]) {
    bfd_ctx *ctx = op->ctx;
    OpenRounds &o = op->o;
    if (op->awaiting_challenge) fail(BFD_ERR_STATE, "the previous round is waiting for its challenge");
    if (op->k >= o.d) fail(BFD_ERR_STATE, "every round is done");
    std::lock_guard<std::recursive_mutex> lk(ctx->mu);
    set_device(ctx);
    hipStream_t st = ctx->stream;
    const int k = op->k, i = o.d - 1 - k;
    const size_t half = (o.M >> k) / 2, ph = (size_t)1 << i;
    const bfd_tuning tn = ctx->tuning_for(o.M);
    const launch::Shape sh = shape_of(tn, ctx->cus);
    const gf *pik = op->lay.pi(k);
    uint64_t *tk = op->trees.as<uint64_t>() + 4 * op->lay.tree_off[k];
    ctx->timer.run(st, "k_leaves_pair", [&] { return launch::leaves_pair(st, pik, tk, half, sh); });
    std::vector<TreeStep> steps;
    tree_steps(half, steps, (size_t)1 << tn.subtree_from_log2, (int)tn.subtree_levels);
    tree_step_launches(ctx, tk, steps, sh, steps.size());
    gf *sums = op->partials.as<gf>() + 2;
    unsigned nb = 0;
    const uint64_t *root = tk + 4 * (2 * half - 2);
    ctx->timer.run(st, "k_dot", [&] { return launch::dot(st, op->half_eq.as<gf>(), op->poly_cur, ph, sums, &nb, root); });
    uint8_t *rb = (uint8_t *)op->pin.p + kPinRound;
    hip_check(hipMemcpyAsync(rb, op->partials.p, 16 * (2 + (size_t)nb), hipMemcpyDeviceToHost, st), "copy round");
    hip_check(hipStreamSynchronize(st), "round");
    ctx->timer.collect();
    gf S = gf_zero();
    for (unsigned u = 0; u < nb; u++) S = gf_add(S, load_elem(rb + 32 + 16 * (size_t)u));
    memcpy(o.pi_roots[k].data(), rb, 32);
    op->pending = o.round_poly(k, S);
    o.comp[k] = {op->pending.rp0, op->pending.rp2};  // compress_round_poly (eval.rs:64-66)
    memcpy(root_out, rb, 32);
    store_elem(poly_out, op->pending.rp0);
    store_elem(poly_out + 16, op->pending.rp2);
    op->awaiting_challenge = true;
}

// The challenge of round k: subclaims advance (sumcheck/src/lib.rs:98-100,
// eval.rs:106-111), pi folds (binary_reed_solomon.rs:88-96) and the
// polynomial's top variable is fixed, in one launch.
void challenge(bfd_opening *op, const uint8_t x_b[16]) {
    bfd_ctx *ctx = op->ctx;
    OpenRounds &o = op->o;
    if (!op->awaiting_challenge) fail(BFD_ERR_STATE, "no round is waiting for a challenge");
    std::lock_guard<std::recursive_mutex> lk(ctx->mu);
    set_device(ctx);
    hipStream_t st = ctx->stream;
    const int k = op->k, i = o.d - 1 - k;
    const size_t half = (o.M >> k) / 2, ph = (size_t)1 << i;
    const gf x = load_elem(x_b);
    o.advance(k, op->pending, x);
    // the previous upload from this slot finished before the round's
    // synchronisation, so the slot is free
    memcpy((uint8_t *)op->pin.p + kPinX, x_b, 16);
    hip_check(hipMemcpyAsync(op->xdev.p, (uint8_t *)op->pin.p + kPinX, 16, hipMemcpyHostToDevice, st), "copy challenge");
    gf *pnext = op->pis.as<gf>() + op->lay.pi_off[k + 1];
    const gf *ftl = op->code->ft.as<gf>() + op->code->ft_off[i];
    gf *pout = op->poly.as<gf>();
    const gf *pik = op->lay.pi(k), *pin = op->poly_cur;
    ctx->timer.run(st, "k_fold_fix",
                   [&] { return launch::fold_fix(st, pik, pnext, ftl, half, pin, pout, ph, op->xdev.as<gf>()); });
    op->poly_cur = pout;
    op->awaiting_challenge = false;
    op->k++;
}

// The final polynomial (basefold.rs:194): the single entry left of the
// polynomial once every variable is fixed.
void final_poly(bfd_opening *op) {
    if (op->final_read) return;
    bfd_ctx *ctx = op->ctx;
    if (op->awaiting_challenge || op->k < op->o.d) fail(BFD_ERR_STATE, "rounds are left");
    std::lock_guard<std::recursive_mutex> lk(ctx->mu);
    set_device(ctx);
    hipStream_t st = ctx->stream;
    uint8_t *rb = (uint8_t *)op->pin.p + kPinRound;
    hip_check(hipMemcpyAsync(rb, op->poly_cur, 16, hipMemcpyDeviceToHost, st), "copy final");
    hip_check(hipStreamSynchronize(st), "final");
    ctx->timer.collect();
    op->o.final_poly = load_elem(rb);
    op->final_read = true;
}

}  // namespace

extern "C" {

int bfd_open_begin(bfd_ctx *ctx, const bfd_code *code, const bfd_commitment *comm, const uint8_t *point,
                   const uint8_t *values, const uint8_t *scalar_challenges, size_t num_scalar_challenges,
                   bfd_opening **out) {
    return guarded([&] {
        if (!ctx || !code || !comm || !values || !out || (!point && comm->n) ||
            (num_scalar_challenges && !scalar_challenges))
            fail(BFD_ERR_ARG, "null argument");
        begin(ctx, code, comm, point, values, scalar_challenges, num_scalar_challenges, out);
    });
}

int bfd_open_rounds_left(const bfd_opening *op, uint32_t *rounds) {
    return guarded([&] {
        if (!op || !rounds) fail(BFD_ERR_ARG, "null argument");
        *rounds = (uint32_t)(op->o.d - op->k);
    });
}

int bfd_open_round(bfd_opening *op, uint8_t root[32], uint8_t round_poly[32]) {
    return guarded([&] {
        if (!op || !root || !round_poly) fail(BFD_ERR_ARG, "null argument");
        round(op, root, round_poly);
    });
}

int bfd_open_challenge(bfd_opening *op, const uint8_t x[16]) {
    return guarded([&] {
        if (!op || !x) fail(BFD_ERR_ARG, "null argument");
        challenge(op, x);
    });
}

int bfd_open_final_poly(bfd_opening *op, uint8_t out[16]) {
    return guarded([&] {
        if (!op || !out) fail(BFD_ERR_ARG, "null argument");
        final_poly(op);
        store_elem(out, op->o.final_poly);
    });
}

int bfd_open_finish(bfd_opening *op, const uint64_t *indices, size_t count, uint8_t **proof, size_t *proof_len) {
    return guarded([&] {
        if (!op || !proof || !proof_len || (count && !indices)) fail(BFD_ERR_ARG, "null argument");
        OpenRounds &o = op->o;
        if (count != o.q) fail(BFD_ERR_ARG, "the proof takes num_queries query indices");
        for (size_t t = 0; t < count; t++)
            if (indices[t] >= o.M) fail(BFD_ERR_ARG, "query index beyond the codeword");
        final_poly(op);
        bfd_ctx *ctx = op->ctx;
        std::lock_guard<std::recursive_mutex> lk(ctx->mu);
        set_device(ctx);
        o.idx.assign(indices, indices + count);
        const size_t sz = proof_size(o.b, o.d, o.q, op->cm->w);
        std::unique_ptr<uint8_t, void (*)(void *)> owned((uint8_t *)malloc(sz ? sz : 1), free);
        if (!owned) fail(BFD_ERR_ARG, "host allocation failed");
        write_proof(ctx, o, op->lay, sz, owned.get());
        *proof = owned.release();
        *proof_len = sz;
    });
}

int bfd_opening_destroy(bfd_opening *op) {
    return guarded([&] {
        if (!op) return;
        {
            std::lock_guard<std::recursive_mutex> lk(op->ctx->mu);
            set_device(op->ctx);
            (void)hipStreamSynchronize(op->ctx->stream);
        }
        delete op;
    });
}

int bfd_eval_data(bfd_ctx *ctx, const void *evals, int evals_on_device, size_t height, size_t width,
                  const uint8_t *point, uint8_t *values) {
    return guarded([&] {
        if (!ctx || !evals || !values || (!point && height > 1)) fail(BFD_ERR_ARG, "null argument");
        if (width < 1 || width > kMaxWidth) fail(BFD_ERR_ARG, "width must be 1..65535");
        if (height == 0 || (height & (height - 1)) || height > ((size_t)1 << 40))
            fail(BFD_ERR_ARG, "height must be a power of two");
        const int n = __builtin_ctzll((unsigned long long)height);
        std::lock_guard<std::recursive_mutex> lk(ctx->mu);
        set_device(ctx);
        hipStream_t st = ctx->stream;
        DevBuf up;
        const gf *data = (const gf *)evals;
        if (!evals_on_device) {
            up = ctx_buf(ctx, 16 * height * width);
            hip_check(hipMemcpyAsync(up.p, evals, 16 * height * width, hipMemcpyHostToDevice, st), "copy evaluations");
            data = up.as<gf>();
        }
        ColumnEval ev;
        ev.enqueue(ctx, data, height, width, point, n);
        hip_check(hipStreamSynchronize(st), "eval");
        ctx->timer.collect();
        ev.finish(ctx, values);
    });
}

}  // extern "C"
