// Host runtime and C ABI of the MI355X Basefold prover.
//
// The host drives one HIP stream per context.  Bulk work (encode, Merkle
// hashing, folds, sumcheck reductions, the proof's query records) runs as
// kernels; the rounds' Fiat-Shamir steps run on the device too (in whichever
// block of a round's last tree launch finishes last), so an opening reads
// back once after its last round, samples the queries on the host and has
// the query records written on the device.
#include "internal.hpp"

using namespace bfd::rt;

namespace {

// Basefold::open for width 1 (basefold.rs:133-232) with the interleaved
// folding sumcheck (:335-369) and the Eval prover (eval.rs, eval_impr.rs).
void do_open(bfd_ctx *ctx, const bfd_code *code, const bfd_commitment *cm, const uint8_t *point_b,
             const uint8_t *values_b, Challenger &ch, uint8_t **proof, size_t *proof_len) {
    std::lock_guard<std::recursive_mutex> lk(ctx->mu);
    set_device(ctx);
    hipStream_t st = ctx->stream;
    double t_setup = now_ms();
    const size_t w = cm->w;
    // a caller buffer is checked before the challenger is touched
    const size_t sz = proof_size((int)cm->b, (int)cm->n, code->q, w);
    if (*proof && *proof_len < sz) fail(BFD_ERR_ARG, "proof buffer too small");
    // width > 1 (basefold.rs:166-183): random scalars by eq combine the
    // columns of the claims, the evaluations and the codeword rows; the rest
    // of the opening runs on the combination
    uint8_t claim_b[16];
    const uint8_t *claim_p = values_b;
    DevBuf comb;
    const gf *poly0 = cm->data.as<gf>(), *pi0 = cm->cw.as<gf>();
    if (w > 1) {
        std::vector<gf> rs(scalar_challenges(w));
        for (gf &v : rs) v = ch.sample();  // sample_ext_vec (challenger.rs:17-19)
        combine_columns(ctx, cm, rs, values_b, claim_b, comb, poly0, pi0);
        claim_p = claim_b;
    }
    OpenRounds o;
    o.init((int)cm->n, (int)cm->b, code->q, point_b, claim_p);
    o.w = w;
    const int d = o.d;
    const size_t M = cm->M, c = (size_t)1 << o.b;

    // device scratch for the whole opening
    OpenLayout lay;
    lay.plan(M, d);
    DevBuf pis = ctx_buf(ctx, 16 * lay.pis_elems(M, c)), trees = ctx_buf(ctx, 32 * lay.pitree_digests());
    lay.cw = pi0;
    if (w > 1) {
        lay.rows = cm->cw.as<gf>();
        lay.w = w;
    }
    lay.tree = cm->tree.as<uint64_t>();
    lay.pis = pis.as<gf>();
    lay.pitrees = trees.as<uint64_t>();
    std::vector<TreeStep> steps;
    const bfd_tuning tn = ctx->tuning_for(M);
    const launch::Shape sh = shape_of(tn, ctx->cus);
    // round k's pi tree over rows (pi[j], pi[j+half]) (basefold.rs:350-355),
    // up to its last launch, which also takes the round's inner product and
    // transcript step when it is a subtree launch (returns whether it is)
    const auto tree_base = [&](int k) {
        const size_t half = (M >> k) / 2;
        const gf *pik = lay.pi(k);
        uint64_t *tk = trees.as<uint64_t>() + 4 * lay.tree_off[k];
        ctx->timer.run(st, "k_leaves_pair", [&] { return launch::leaves_pair(st, pik, tk, half, sh); });
        tree_steps(half, steps, (size_t)1 << tn.subtree_from_log2, (int)tn.subtree_levels);
        const bool fused = !steps.empty() && steps.back().lv > 0;
        tree_step_launches(ctx, tk, steps, sh, fused ? steps.size() - 1 : steps.size());
        return fused;
    };
    // round 0's tree hashes pi_0, which the commitment holds: it goes first,
    // so the device hashes while the host builds the eq table and the
    // transcript image
    const bool fused0 = d > 0 && tree_base(0);
    const int hn = o.half_eq_vars();
    DevBuf half_eq = ctx_buf(ctx, 16 * ((size_t)1 << hn));
    if (o.n > 0) device_eq(ctx, o.r.data(), hn, half_eq.as<gf>());
    DevBuf poly = ctx_buf(ctx, 16 * (cm->N > 1 ? cm->N / 2 : 1));
    // inner-product block sums of the current round (dot_blocks() <= 512)
    DevBuf partials = ctx_buf(ctx, 16 * 512);
    gf *sums = partials.as<gf>();
    // the transcript runs on the device: the starting image and the round
    // kernel's segment table go up in one copy, the d rounds are enqueued
    // without a read-back, and one read-back after the last round returns
    // the round polynomials, the pi-tree roots and the challenger state
    // region: transcript | d round segments | d zeroed completion counters
    // of the rounds' fused top-of-tree launches (k_top_round)
    const uint32_t in_cap = fs_in_cap(ch.input.size());
    const size_t fsb = (fs_bytes((uint32_t)d, in_cap) + 255) & ~(size_t)255;
    const size_t nseg = (size_t)(d ? d : 1), img = fsb + sizeof(Seg) * nseg + 4 * nseg;
    DevBuf fs = ctx_buf(ctx, img);
    Seg *d_segs = (Seg *)(fs.as<uint8_t>() + fsb);
    unsigned *d_counters = (unsigned *)(fs.as<uint8_t>() + fsb + sizeof(Seg) * nseg);
    ctx->pinned.ensure(img + 64);
    uint8_t *pin = (uint8_t *)ctx->pinned.p;
    if (d > 0) {
        fs_write(o, ch, in_cap, pin);
        Seg *hs = (Seg *)(pin + fsb);
        for (int k = 0; k < d; k++) {
            const size_t half = (M >> k) / 2;
            const uint64_t *root_k = trees.as<uint64_t>() + 4 * lay.tree_off[k] + 4 * (2 * half - 2);
            hs[k] = fs_round_seg(root_k, fs.p, k, 0, launch::dot_blocks((size_t)1 << (d - 1 - k)));
        }
        memset(pin + fsb + sizeof(Seg) * nseg, 0, 4 * nseg);
        hip_check(hipMemcpyAsync(fs.p, pin, img, hipMemcpyHostToDevice, st), "copy transcript");
    }
    const gf *x_dev = fs_challenge(fs.p);

    const gf *poly_cur = poly0;
    ctx->timer.host("open_setup", now_ms() - t_setup);
    for (int k = 0; k < d; k++) {
        const int i = d - 1 - k;
        const size_t half = (M >> k) / 2, ph = (size_t)1 << i;
        const gf *pik = lay.pi(k);
        uint64_t *tk = trees.as<uint64_t>() + 4 * lay.tree_off[k];
        // the inner product of the halved eq table with the low half of poly
        // (eval_impr.rs:187-197) and the round's transcript step (observe
        // root and round poly, sample x) ride on the tree's last launch
        const bool fused = k == 0 ? fused0 : tree_base(k);
        if (fused) {
            const TreeStep &ts = steps.back();
            ctx->timer.run(st, "k_top_round", [&] {
                return launch::top_round(st, tk + 4 * ts.prev, ts.h, ts.lv, half_eq.as<gf>(), poly_cur, ph, sums,
                                         d_segs + k, d_counters + k, ctx->cus);
            });
        } else {
            unsigned nb = 0;
            ctx->timer.run(st, "k_dot", [&] { return launch::dot(st, half_eq.as<gf>(), poly_cur, ph, sums, &nb); });
            ctx->timer.run(st, "ks_round",
                           [&] { return launch::seg_launch(st, SegKernel::Round, d_segs + k, 1, 1, sums); });
        }
        // fold pi (binary_reed_solomon.rs:88-96) and fix the top variable of poly
        gf *pnext = pis.as<gf>() + lay.pi_off[k + 1];
        const gf *ftl = code->ft.as<gf>() + code->ft_off[i];
        gf *pout = poly.as<gf>();
        ctx->timer.run(st, "k_fold_fix",
                       [&] { return launch::fold_fix(st, pik, pnext, ftl, half, poly_cur, pout, ph, x_dev); });
        poly_cur = pout;
    }
    const size_t rbb = d > 0 ? fs_readback_bytes((uint32_t)d) : 0;
    if (d > 0) hip_check(hipMemcpyAsync(pin, fs.p, rbb, hipMemcpyDeviceToHost, st), "copy transcript back");
    hip_check(hipMemcpyAsync(pin + ((rbb + 15) & ~(size_t)15), poly_cur, 16, hipMemcpyDeviceToHost, st),
              "copy final");
    // while the rounds run: the proof buffer, with its pages touched so that
    // filling it later takes no page faults
    std::unique_ptr<uint8_t, void (*)(void *)> owned(nullptr, free);
    uint8_t *dst = *proof;
    if (!dst) {
        owned.reset((uint8_t *)malloc(sz ? sz : 1));
        if (!owned) fail(BFD_ERR_ARG, "host allocation failed");
        dst = owned.get();
        ChunkCopy::zero(dst, sz);
    }
    hip_check(hipStreamSynchronize(st), "final sync");
    const double th = now_ms();
    if (d > 0) fs_read(pin, o, ch);
    o.finish(load_elem(pin + ((rbb + 15) & ~(size_t)15)), ch);
    ctx->timer.host("round_transcript", now_ms() - th);

    double t0 = now_ms();
    write_proof(ctx, o, lay, sz, dst);
    owned.release();
    *proof = dst;
    *proof_len = sz;
    ctx->timer.host("proof", now_ms() - t0);
}

}  // namespace

namespace bfd {
namespace rt {

bfd_code *create_code(bfd_ctx *ctx, uint32_t lambda, uint32_t log2_c, uint32_t d) {
    if (log2_c > 8 || d > 32 || log2_c + d > 34) fail(BFD_ERR_ARG, "code too large");
    std::lock_guard<std::recursive_mutex> lk(ctx->mu);
    set_device(ctx);
    hipStream_t st = ctx->stream;
    auto code = std::make_unique<bfd_code>();
    code->ctx = ctx;
    code->lambda = lambda;
    code->log2_c = log2_c;
    code->d = d;
    code->q = num_queries(lambda, log2_c);
    const int L = (int)(log2_c + d);
    // level s table (fold order) = twiddles[b+s][bitrev_{b+s}(j)], j < 2^(b+s)
    // = XOR over bits p of j of basis[b+s][b+s-1-p]
    auto basis = ntt_basis(L);
    code->ft_off.assign(d + 1, 0);
    size_t total = 0;
    for (uint32_t s = 0; s < d; s++) {
        code->ft_off[s] = total;
        total += (size_t)1 << (log2_c + s);
    }
    code->ft_off[d] = total;
    code->ft = DevBuf(16 * (total ? total : 1), st);
    std::vector<gf> bt;
    std::vector<size_t> bt_off(d + 1, 0);
    for (uint32_t s = 0; s < d; s++) {
        const int K = (int)(log2_c + s);
        bt_off[s] = bt.size();
        for (int p = 0; p < K; p++) bt.push_back(basis[K][K - 1 - p]);
    }
    DevBuf dbt(16 * (bt.size() ? bt.size() : 1), st);
    if (!bt.empty())
        hip_check(hipMemcpyAsync(dbt.p, bt.data(), 16 * bt.size(), hipMemcpyHostToDevice, st), "copy basis");
    for (uint32_t s = 0; s < d; s++) {
        const int K = (int)(log2_c + s);
        hip_check(launch::xor_span(st, code->ft.as<gf>() + code->ft_off[s], dbt.as<gf>() + bt_off[s], K,
                                   (size_t)1 << K),
                  "xor_span");
    }
    hip_check(hipStreamSynchronize(st), "code setup");
    code->bt = std::move(bt);
    code->bt_off = std::move(bt_off);
    return code.release();
}

bfd_code *cached_code(bfd_ctx *ctx, uint32_t lambda, uint32_t log2_c, uint32_t d) {
    std::lock_guard<std::recursive_mutex> lk(ctx->mu);
    const uint64_t key = ((uint64_t)lambda << 32) | ((uint64_t)log2_c << 16) | d;
    auto it = ctx->code_cache.find(key);
    if (it != ctx->code_cache.end()) return it->second;
    bfd_code *c = create_code(ctx, lambda, log2_c, d);
    ctx->code_cache[key] = c;
    return c;
}

}  // namespace rt
}  // namespace bfd

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
extern "C" {

const char *bfd_last_error(void) { return g_err.c_str(); }

int bfd_ctx_create(int device, uint64_t hbm_budget_bytes, bfd_ctx **out) {
    return guarded([&] {
        if (!out) fail(BFD_ERR_ARG, "null out");
        int count = 0;
        hip_check(hipGetDeviceCount(&count), "hipGetDeviceCount");
        if (device < 0 || device >= count) fail(BFD_ERR_HIP, "no such HIP device");
        auto ctx = std::make_unique<bfd_ctx>();
        ctx->device = device;
        ctx->budget = hbm_budget_bytes;
        hip_check(hipSetDevice(device), "hipSetDevice");
        hip_check(hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking), "hipStreamCreate");
        // the stream is destroyed again if set-up fails below
        struct StreamGuard {
            hipStream_t s;
            ~StreamGuard() {
                if (s) (void)hipStreamDestroy(s);
            }
        } guard{ctx->stream};
        size_t free_b = 0, total_b = 0;
        hip_check(hipMemGetInfo(&free_b, &total_b), "hipMemGetInfo");
        ctx->cache.st = ctx->stream;
        ctx->tuning.fill(default_tuning());
        ctx->batch_tuning = default_tuning();
        hip_check(hipDeviceGetAttribute(&ctx->cus, hipDeviceAttributeMultiprocessorCount, device), "CU count");
        // the allocation-free entry points' scratch address on this device; a
        // failure here is reported by those entry points, not by the context
        try {
            ctx->scratch = noalloc_scratch(device);
        } catch (const Error &e) {
            (void)hipGetLastError();
            ctx->scratch_error = e.msg;
        }
        if (ctx->cus < 1) ctx->cus = 1;
        ctx->cache.cap = total_b / 4;
        if (hbm_budget_bytes && hbm_budget_bytes < ctx->cache.cap) ctx->cache.cap = hbm_budget_bytes;
        hipMemPool_t pool;
        if (hipDeviceGetDefaultMemPool(&pool, device) == hipSuccess) {
            uint64_t thr = UINT64_MAX;
            (void)hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &thr);
        }
        guard.s = nullptr;
        *out = ctx.release();
    });
}

int bfd_ctx_destroy(bfd_ctx *ctx) {
    return guarded([&] {
        if (!ctx) return;
        (void)hipSetDevice(ctx->device);
        ctx->lane.reset();
        for (auto &kv : ctx->code_cache) delete kv.second;
        ctx->code_cache.clear();
        ctx->cache.flush();
        (void)hipStreamSynchronize(ctx->stream);
        (void)hipStreamDestroy(ctx->stream);
        delete ctx;
    });
}

int bfd_ctx_synchronize(bfd_ctx *ctx) {
    return guarded([&] {
        if (!ctx) fail(BFD_ERR_ARG, "null ctx");
        set_device(ctx);
        hip_check(hipStreamSynchronize(ctx->stream), "synchronize");
    });
}

uint64_t bfd_ctx_resident_bytes(const bfd_ctx *ctx) { return ctx ? ctx->resident : 0; }

int bfd_code_create(bfd_ctx *ctx, uint32_t lambda, uint32_t log2_c, uint32_t d, bfd_code **out) {
    return guarded([&] {
        if (!ctx || !out) fail(BFD_ERR_ARG, "null argument");
        *out = create_code(ctx, lambda, log2_c, d);
    });
}

int bfd_code_destroy(bfd_code *code) {
    return guarded([&] {
        if (!code) return;
        std::lock_guard<std::recursive_mutex> lk(code->ctx->mu);
        set_device(code->ctx);
        code->ft.release();
        (void)hipStreamSynchronize(code->ctx->stream);
        delete code;
    });
}

uint64_t bfd_code_num_queries(const bfd_code *code) { return code ? code->q : 0; }

int bfd_challenger_create(bfd_challenger **out) {
    return guarded([&] {
        if (!out) fail(BFD_ERR_ARG, "null out");
        *out = new bfd_challenger();
    });
}
int bfd_challenger_clone(const bfd_challenger *c, bfd_challenger **out) {
    return guarded([&] {
        if (!c || !out) fail(BFD_ERR_ARG, "null argument");
        *out = new bfd_challenger(*c);
    });
}
int bfd_challenger_destroy(bfd_challenger *c) {
    delete c;
    return BFD_OK;
}
int bfd_challenger_observe(bfd_challenger *c, const uint8_t *bytes, size_t len) {
    return guarded([&] {
        if (!c || (!bytes && len)) fail(BFD_ERR_ARG, "null argument");
        c->c.observe(bytes, len);
    });
}
int bfd_challenger_sample(bfd_challenger *c, uint8_t out[16]) {
    return guarded([&] {
        if (!c || !out) fail(BFD_ERR_ARG, "null argument");
        store_elem(out, c->c.sample());
    });
}
int bfd_challenger_sample_bits(bfd_challenger *c, uint32_t bits, uint64_t *out) {
    return guarded([&] {
        if (!c || !out || bits > 64) fail(BFD_ERR_ARG, "bad argument");
        *out = c->c.sample_bits((int)bits);
    });
}

int bfd_challenger_export(const bfd_challenger *c, uint8_t *input, size_t input_cap, size_t *input_len,
                          uint8_t output[32], size_t *output_len) {
    return guarded([&] {
        if (!c || !input_len || !output_len || (input_cap && !input)) fail(BFD_ERR_ARG, "null argument");
        *input_len = c->c.input.size();
        *output_len = c->c.output.size();
        if (input_cap < c->c.input.size()) fail(BFD_ERR_ARG, "input buffer too small");
        if (c->c.output.size() > 32) fail(BFD_ERR_STATE, "output buffer holds more than one digest");
        if (!c->c.input.empty()) memcpy(input, c->c.input.data(), c->c.input.size());
        if (!c->c.output.empty()) {
            if (!output) fail(BFD_ERR_ARG, "null argument");
            memcpy(output, c->c.output.data(), c->c.output.size());
        }
    });
}
int bfd_challenger_import(const uint8_t *input, size_t input_len, const uint8_t *output, size_t output_len,
                          bfd_challenger **out) {
    return guarded([&] {
        if (!out || (input_len && !input) || (output_len && !output)) fail(BFD_ERR_ARG, "null argument");
        if (output_len > 32) fail(BFD_ERR_ARG, "the output buffer holds at most one 32-byte digest");
        auto c = std::make_unique<bfd_challenger>();
        c->c.input.assign(input, input + input_len);
        c->c.output.assign(output, output + output_len);
        *out = c.release();
    });
}

int bfd_commit(bfd_ctx *ctx, const bfd_code *code, const uint8_t *evals, size_t height, size_t width,
               bfd_commitment **out, uint8_t root[32]) {
    return guarded([&] { do_commit(ctx, code, evals, false, height, width, out, root); });
}
int bfd_commit_device(bfd_ctx *ctx, const bfd_code *code, const void *d_evals, size_t height, size_t width,
                      bfd_commitment **out, uint8_t root[32]) {
    return guarded([&] { do_commit(ctx, code, d_evals, true, height, width, out, root); });
}
int bfd_commit_eval(bfd_ctx *ctx, const bfd_code *code, const void *evals, int evals_on_device, size_t height,
                    size_t width, const uint8_t *point, bfd_commitment **out, uint8_t root[32], uint8_t *values) {
    return guarded([&] {
        if (!values || (!point && code && code->d)) fail(BFD_ERR_ARG, "null argument");
        do_commit(ctx, code, evals, evals_on_device != 0, height, width, out, root, point, values);
    });
}
int bfd_commitment_root(const bfd_commitment *comm, uint8_t root[32]) {
    return guarded([&] {
        if (!comm || !root) fail(BFD_ERR_ARG, "null argument");
        memcpy(root, comm->root, 32);
    });
}
int bfd_commitment_destroy(bfd_commitment *comm) {
    return guarded([&] {
        if (!comm) return;
        bfd_ctx *ctx = comm->ctx;
        std::lock_guard<std::recursive_mutex> lk(ctx->mu);
        set_device(ctx);
        ctx->resident -= comm->bytes;
        delete comm;
    });
}

int bfd_eval(bfd_ctx *ctx, const bfd_commitment *comm, const uint8_t *point, uint8_t *values) {
    return guarded([&] {
        if (!ctx || !comm || !point || !values) fail(BFD_ERR_ARG, "null argument");
        if (comm->ctx->device != ctx->device) fail(BFD_ERR_ARG, "the commitment lives on another device");
        std::lock_guard<std::recursive_mutex> lk(ctx->mu);
        set_device(ctx);
        hipStream_t st = ctx->stream;
        // one pass over the data (k_mle_dot) instead of n fix passes
        ColumnEval ev;
        ev.enqueue(ctx, comm->data.as<gf>(), comm->N, comm->w, point, (int)comm->n);
        hip_check(hipStreamSynchronize(st), "eval");
        ctx->timer.collect();
        ev.finish(ctx, values);
    });
}

int bfd_open(bfd_ctx *ctx, const bfd_code *code, const bfd_commitment *comm, const uint8_t *point,
             const uint8_t *values, bfd_challenger *challenger, uint8_t **proof, size_t *proof_len) {
    return guarded([&] {
        if (!ctx || !code || !comm || !values || !challenger || !proof || !proof_len || (!point && comm->n))
            fail(BFD_ERR_ARG, "null argument");
        if (comm->code != code) fail(BFD_ERR_ARG, "commitment was made with another code");
        if (comm->ctx->device != ctx->device) fail(BFD_ERR_ARG, "the commitment lives on another device");
        uint8_t *p = nullptr;
        size_t n = 0;
        do_open(ctx, code, comm, point, values, challenger->c, &p, &n);
        *proof = p;
        *proof_len = n;
    });
}

int bfd_open_into(bfd_ctx *ctx, const bfd_code *code, const bfd_commitment *comm, const uint8_t *point,
                  const uint8_t *values, bfd_challenger *challenger, uint8_t *proof, size_t capacity,
                  size_t *proof_len) {
    return guarded([&] {
        if (!ctx || !code || !comm || !values || !challenger || !proof || !proof_len || (!point && comm->n))
            fail(BFD_ERR_ARG, "null argument");
        if (comm->code != code) fail(BFD_ERR_ARG, "commitment was made with another code");
        if (comm->ctx->device != ctx->device) fail(BFD_ERR_ARG, "the commitment lives on another device");
        uint8_t *p = proof;
        size_t n = capacity;
        do_open(ctx, code, comm, point, values, challenger->c, &p, &n);
        *proof_len = n;
    });
}

uint64_t bfd_proof_size(uint32_t lambda, uint32_t log2_c, uint32_t d) {
    return proof_size((int)log2_c, (int)d, num_queries(lambda, log2_c));
}

uint64_t bfd_proof_size_width(uint32_t lambda, uint32_t log2_c, uint32_t d, uint64_t width) {
    return proof_size((int)log2_c, (int)d, num_queries(lambda, log2_c), width ? width : 1);
}

void bfd_proof_free(uint8_t *proof) { free(proof); }

}  // extern "C"

namespace {

// The degree-3 product sumcheck with its transcript on the device: one
// launch per round while a round has more than prod_tail_max_pairs() pairs
// (fold, pair sums and, in the block that finishes last, the round's
// Fiat-Shamir step), then one single-block launch for the remaining rounds
// with the tables in LDS; one upload of the eq factors and the transcript
// image before the rounds and one read-back after them.  Round 0 reads the
// caller's tables (device memory, or one upload of host tables) and round 1
// folds them into working tables of half their size, so the caller's tables
// are never copied or modified.  The halved eq table stays in its two
// factors (kernels.hpp EqFactors).
void product_sumcheck(bfd_ctx *ctx, uint32_t n, const void *f_src, const void *g_src, bool on_device,
                      const uint8_t *r_b, Challenger &ch, uint8_t claim_out[16], uint8_t *round_polys,
                      uint8_t *challenges, uint8_t final_fg[32], uint8_t subclaim_out[16]) {
    std::lock_guard<std::recursive_mutex> lk(ctx->mu);
    set_device(ctx);
    hipStream_t st = ctx->stream;
    const size_t N = (size_t)1 << n;
    if (n == 0) {  // prove_sumcheck returns before observing (sumcheck/src/lib.rs:63-65)
        uint8_t fg[32];
        if (on_device) {
            hip_check(hipMemcpyAsync(fg, f_src, 16, hipMemcpyDeviceToHost, st), "copy f");
            hip_check(hipMemcpyAsync(fg + 16, g_src, 16, hipMemcpyDeviceToHost, st), "copy g");
            hip_check(hipStreamSynchronize(st), "sync");
        } else {
            memcpy(fg, f_src, 16);
            memcpy(fg + 16, g_src, 16);
        }
        const gf f0 = load_elem(fg), g0 = load_elem(fg + 16), cl = hmul(f0, g0);
        store_elem(claim_out, cl);
        memcpy(final_fg, fg, 32);
        store_elem(subclaim_out, cl);
        return;
    }
    std::vector<gf> r(n);
    for (uint32_t i = 0; i < n; i++) r[i] = load_elem(r_b + 16 * i);
    const int hn = (int)n - 1;
    // correcting factors of the halved eq table (eval_impr.rs:81-111), with
    // one field inversion for all (1 + r_j)
    std::vector<gf> inv(hn);
    for (int j = 0; j < hn; j++) {
        inv[j] = gf_add(gf_one(), r[j]);
        if (gf_eq(inv[j], gf_zero())) fail(BFD_ERR_ARG, "r_j = 1 makes the halved eq table singular");
    }
    batch_inverse(inv);
    std::vector<gf> cf(n, gf_one());
    for (int j = hn - 1; j >= 0; j--) cf[j] = hmul(cf[j + 1], inv[j]);
    if (ch.output.size() > 32) fail(BFD_ERR_ARG, "transcript state too large");
    DevBuf up, W = ctx_buf(ctx, 2 * 16 * std::max<size_t>(2, N / 2));
    const gf *fs0 = (const gf *)f_src, *gs0 = (const gf *)g_src;
    if (!on_device) {
        up = ctx_buf(ctx, 2 * 16 * N);
        hip_check(hipMemcpyAsync(up.p, f_src, 16 * N, hipMemcpyHostToDevice, st), "copy f");
        hip_check(hipMemcpyAsync(up.as<gf>() + N, g_src, 16 * N, hipMemcpyHostToDevice, st), "copy g");
        fs0 = up.as<gf>();
        gs0 = fs0 + N;
    }
    gf *wf = W.as<gf>(), *wg = wf + std::max<size_t>(2, N / 2);
    // one upload: the two factors of the halved eq table, then the transcript
    // image (kernels.hpp ProdFs) with its counters and round 0's
    // accumulation slots zeroed.  The factors split the hn variables evenly,
    // with at least 7 in A so that a block of the round kernels (128 pairs
    // or more) always lies in one B entry
    const int lo = std::max((hn + 1) / 2, std::min(hn, 7)), hi = hn - lo;
    const std::vector<gf> A = host_eq(r.data(), lo), B = host_eq(r.data() + lo, hi);
    const size_t fb = (16 * (A.size() + B.size()) + 255) & ~(size_t)255;
    const uint32_t in_cap = (uint32_t)(((ch.input.size() > 32 ? ch.input.size() : 32) + 64 + 7) & ~(size_t)7);
    const size_t in_off = prod_input_off(n), img = prod_bytes(n, in_cap), upl = prod_upload_bytes(n, in_cap);
    const size_t fin_off = prod_finals_off(n, in_cap);
    DevBuf buf = ctx_buf(ctx, fb + img + 64);
    ctx->pinned.ensure(fb + upl + 64);
    uint8_t *pin = (uint8_t *)ctx->pinned.p;
    memcpy(pin, A.data(), 16 * A.size());
    memcpy(pin + 16 * A.size(), B.data(), 16 * B.size());
    uint8_t *hreg = pin + fb;
    memset(hreg, 0, upl);
    ProdFs h;
    memset(&h, 0, sizeof h);
    h.in_len = (uint32_t)ch.input.size();
    h.out_len = (uint32_t)ch.output.size();
    h.in_cap = in_cap;
    h.n = n;
    // round 0's eq scalars: s = 1, K = e0 cf, L = cf at variable n-1
    h.s = gf_one();
    h.K = hmul(gf_add(gf_one(), r[n - 1]), cf[n - 1]);
    h.L = cf[n - 1];
    if (!ch.output.empty()) memcpy(h.out, ch.output.data(), ch.output.size());
    memcpy(hreg, &h, sizeof h);
    gf *c = (gf *)(hreg + prod_cf_off());
    for (uint32_t i = 0; i < n; i++) {
        const gf e0 = gf_add(gf_one(), r[i]);  // eq(r_i, X) = (1 + r_i) + X
        c[i] = cf[i];
        c[n + i] = e0;
        c[2 * n + i] = hmul(e0, cf[i]);
        // halved eq table entries whose indices differ in bit i: ratio r_i / (1 + r_i)
        c[3 * n + i] = (int)i < hn ? hmul(r[i], inv[i]) : gf_zero();
    }
    if (!ch.input.empty()) memcpy(hreg + in_off, ch.input.data(), ch.input.size());
    hip_check(hipMemcpyAsync(buf.p, pin, fb + upl, hipMemcpyHostToDevice, st), "copy transcript");
    const EqFactors eq{buf.as<gf>(), buf.as<gf>() + A.size(), lo};
    ProdFs *dfs = (ProdFs *)(buf.as<uint8_t>() + fb);
    uint32_t k = 0;
    bool look = false;  // a lookahead round has run
    for (; k < n; k++) {
        const size_t half = (size_t)1 << (n - 1 - k);
        if (half <= launch::prod_tail_max_pairs()) break;
        const gf *fsrc = k <= 1 ? fs0 : wf, *gsrc = k <= 1 ? gs0 : wg;
        gf *fdst = k == 0 ? (gf *)fs0 : wf, *gdst = k == 0 ? (gf *)gs0 : wg;  // round 0 writes nothing
        if (k >= 1 && half <= launch::prod_look_max_pairs()) {
            ctx->timer.run(st, "k_prod_look", [&] {
                return launch::prod_look(st, fsrc, gsrc, fdst, gdst, eq, half, (int)k, dfs, !look);
            });
            look = true;
            continue;
        }
        ctx->timer.run(st, "k_prod_round", [&] {
            return launch::prod_step(st, fsrc, gsrc, fdst, gdst, eq, half, (int)k, dfs, ctx->cus);
        });
    }
    gf *fins = (gf *)((uint8_t *)dfs + fin_off);
    {
        // the single-block launch ends by writing the region up to the last
        // round's table entries into hreg (stream order puts it after the
        // upload that read hreg)
        const size_t half = (size_t)1 << (n - 1 - k);
        const gf *fsrc = k <= 1 ? fs0 : wf, *gsrc = k <= 1 ? gs0 : wg;
        uint8_t *dpin = nullptr;
        hip_check(hipHostGetDevicePointer((void **)&dpin, ctx->pinned.p, 0), "hipHostGetDevicePointer");
        ctx->timer.run(st, "k_prod_tail", [&] {
            return launch::prod_tail(st, fsrc, gsrc, fins, fins + 2, eq, half, (int)k, dfs, dpin + fb,
                                     (uint32_t)(fin_off + 64));
        });
    }
    hip_check(hipStreamSynchronize(st), "sumcheck");
    ctx->timer.collect();
    memcpy(&h, hreg, sizeof h);
    if (h.n != n || h.in_len != 32 || h.out_len != 16) fail(BFD_ERR_HIP, "device transcript state is corrupt");
    store_elem(claim_out, h.claim);
    memcpy(round_polys, hreg + prod_polys_off(n), 48 * (size_t)n);
    memcpy(challenges, hreg + prod_chal_off(n), 16 * (size_t)n);
    ch.input.assign(hreg + in_off, hreg + in_off + 32);
    ch.output.assign(h.out, h.out + h.out_len);
    pin = hreg;
    const size_t fo = fin_off;
    // the subclaim through the rounds (SumcheckSubclaim::reduce, lib.rs:98-100):
    // c1 = subclaim + c2 + c3 in characteristic 2, then Horner at x
    gf sub = h.claim;
    for (uint32_t t = 0; t < n; t++) {
        const uint8_t *p = pin + prod_polys_off(n) + 48 * (size_t)t;
        const gf c0 = load_elem(p), c2 = load_elem(p + 16), c3 = load_elem(p + 32);
        const gf c1 = gf_add(sub, gf_add(c2, c3));
        const gf x = load_elem(pin + prod_chal_off(n) + 16 * (size_t)t);
        sub = gf_add(c0, hmul(x, gf_add(c1, hmul(x, gf_add(c2, hmul(x, c3))))));
    }
    // the last fold of the two remaining entries (fix_last_var with x)
    const gf x = h.x;
    const gf f0 = load_elem(pin + fo), f1 = load_elem(pin + fo + 16);
    const gf g0 = load_elem(pin + fo + 32), g1 = load_elem(pin + fo + 48);
    store_elem(final_fg, gf_add(f0, hmul(x, gf_add(f1, f0))));
    store_elem(final_fg + 16, gf_add(g0, hmul(x, gf_add(g1, g0))));
    store_elem(subclaim_out, sub);
}

}  // namespace

extern "C" {

int bfd_sumcheck_product(bfd_ctx *ctx, uint32_t n, const uint8_t *f_b, const uint8_t *g_b, const uint8_t *r_b,
                         bfd_challenger *challenger, uint8_t claim_out[16], uint8_t *round_polys,
                         uint8_t *challenges, uint8_t final_fg[32], uint8_t subclaim_out[16]) {
    return guarded([&] {
        if (!ctx || !f_b || !g_b || !challenger || !claim_out || !final_fg || !subclaim_out ||
            (n && (!r_b || !round_polys || !challenges)))
            fail(BFD_ERR_ARG, "null argument");
        if (n > 32) fail(BFD_ERR_ARG, "too many variables");
        product_sumcheck(ctx, n, f_b, g_b, false, r_b, challenger->c, claim_out, round_polys, challenges, final_fg,
                         subclaim_out);
    });
}

int bfd_sumcheck_product_device(bfd_ctx *ctx, uint32_t n, const void *d_f, const void *d_g, const uint8_t *r_b,
                                bfd_challenger *challenger, uint8_t claim_out[16], uint8_t *round_polys,
                                uint8_t *challenges, uint8_t final_fg[32], uint8_t subclaim_out[16]) {
    return guarded([&] {
        if (!ctx || !d_f || !d_g || !challenger || !claim_out || !final_fg || !subclaim_out ||
            (n && (!r_b || !round_polys || !challenges)))
            fail(BFD_ERR_ARG, "null argument");
        if (n > 32) fail(BFD_ERR_ARG, "too many variables");
        product_sumcheck(ctx, n, d_f, d_g, true, r_b, challenger->c, claim_out, round_polys, challenges, final_fg,
                         subclaim_out);
    });
}

}  // extern "C"

namespace {

// prove_sumcheck for a Quadratic (quadratic.rs:84-136) with its transcript on
// the device: per round one launch for the pair sums (whose last block runs
// the round's Fiat-Shamir step) and one for the fix of every table with the
// challenge that step left in device memory, until the tables fit one
// block's LDS; then one single-block launch for the remaining rounds.  One
// upload of the transcript image before the rounds and one read-back after
// them.  Round 0 reads the caller's tables and its fix writes working tables
// of half their size, so the caller's tables are never copied or modified.
void quadratic_sumcheck(bfd_ctx *ctx, uint32_t n, uint32_t m, const void *polys, bool on_device, uint32_t npairs,
                        const uint8_t *scalars, const uint32_t *pair_idx, Challenger &ch, uint8_t claim_out[16],
                        uint8_t *round_polys, uint8_t *challenges, uint8_t *final_evals, uint8_t subclaim_out[16]) {
    std::lock_guard<std::recursive_mutex> lk(ctx->mu);
    set_device(ctx);
    hipStream_t st = ctx->stream;
    const size_t N = (size_t)1 << n;
    std::vector<gf> sc(npairs);
    for (uint32_t k = 0; k < npairs; k++) sc[k] = load_elem(scalars + 16 * k);
    if (n == 0) {  // prove_sumcheck returns before observing (sumcheck/src/lib.rs:63-65)
        std::vector<uint8_t> host(16 * (size_t)m);
        if (on_device) {
            hip_check(hipMemcpyAsync(host.data(), polys, 16 * (size_t)m, hipMemcpyDeviceToHost, st), "copy tables");
            hip_check(hipStreamSynchronize(st), "sync");
        } else {
            memcpy(host.data(), polys, 16 * (size_t)m);
        }
        gf cl = gf_zero();
        for (uint32_t k = 0; k < npairs; k++)
            cl = gf_add(cl, hmul(sc[k], hmul(load_elem(host.data() + 16 * pair_idx[2 * k]),
                                             load_elem(host.data() + 16 * pair_idx[2 * k + 1]))));
        store_elem(claim_out, cl);
        memcpy(final_evals, host.data(), 16 * (size_t)m);
        store_elem(subclaim_out, cl);
        return;
    }
    if (ch.output.size() > 32) fail(BFD_ERR_ARG, "transcript state too large");
    DevBuf up;
    const gf *src0 = (const gf *)polys;
    if (!on_device) {
        up = ctx_buf(ctx, 16 * N * m);
        hip_check(hipMemcpyAsync(up.p, polys, 16 * N * m, hipMemcpyHostToDevice, st), "copy tables");
        src0 = up.as<gf>();
    }
    const size_t NW = std::max<size_t>(1, N / 2);  // working table stride
    DevBuf W = ctx_buf(ctx, 16 * NW * m);
    // blocks per pair: enough for the device at the first round
    const unsigned max_blocks = std::max(1u, std::min(512u, 4096u / npairs));
    // one upload: the transcript image (kernels.hpp QuadFs) with the scalars,
    // the pair indices and zeroed slots and counters
    const uint32_t in_cap = (uint32_t)(((ch.input.size() > 32 ? ch.input.size() : 32) + 64 + 7) & ~(size_t)7);
    const size_t in_off = quad_input_off(npairs, n), fin_off = quad_finals_off(npairs, n, in_cap);
    const size_t img = quad_bytes(npairs, n, in_cap, m);
    DevBuf fs = ctx_buf(ctx, img);
    ctx->pinned.ensure(img + 64);
    uint8_t *pin = (uint8_t *)ctx->pinned.p;
    memset(pin, 0, img);
    QuadFs h;
    memset(&h, 0, sizeof h);
    h.in_len = (uint32_t)ch.input.size();
    h.out_len = (uint32_t)ch.output.size();
    h.in_cap = in_cap;
    h.n = n;
    h.npairs = npairs;
    if (!ch.output.empty()) memcpy(h.out, ch.output.data(), ch.output.size());
    memcpy(pin, &h, sizeof h);
    memcpy(pin + quad_scalars_off(), scalars, 16 * (size_t)npairs);
    if (!ch.input.empty()) memcpy(pin + in_off, ch.input.data(), ch.input.size());
    memcpy(pin + quad_pairs_off(npairs, n, in_cap, m), pair_idx, 8 * (size_t)npairs);
    hip_check(hipMemcpyAsync(fs.p, pin, img, hipMemcpyHostToDevice, st), "copy transcript");
    QuadFs *dfs = (QuadFs *)fs.p;
    uint8_t *dpin = nullptr;  // pin as the device addresses it
    hip_check(hipHostGetDevicePointer((void **)&dpin, pin, 0), "hipHostGetDevicePointer");
    const uint32_t *dpairs = (const uint32_t *)(fs.as<uint8_t>() + quad_pairs_off(npairs, n, in_cap, m));
    gf *slots = (gf *)(fs.as<uint8_t>() + quad_slots_off(npairs, n, in_cap, m));
    unsigned *counters = (unsigned *)(fs.as<uint8_t>() + quad_counters_off(npairs, n, in_cap, m));
    const gf *x_dev = (const gf *)(fs.as<uint8_t>() + offsetof(QuadFs, x));
    const size_t tail = launch::quad_tail_max_half(m, npairs), HL = launch::quad_look_max_half();
    // Rounds of at most HL pairs of entries (above the single-block rounds)
    // run with one round of lookahead; they fold the previous round's tables
    // themselves, into the other of two working buffers.  Every other round
    // needs its own tables folded first: `pending` says cur holds the
    // previous round's, still to be folded with that round's challenge.
    const auto lookahead = [&](uint32_t k, size_t half) { return k >= 1 && half <= HL && half >= 64 && half > tail; };
    DevBuf W2;
    const size_t NW2 = 2 * std::min(HL, std::max<size_t>(1, N / 4));
    const gf *cur = src0;
    size_t cstride = N;
    bool pending = false, look = false, tail_ran = false;
    const auto other = [&](gf *&o, size_t &ostride) {
        if (cur == W2.as<gf>()) {
            o = W.as<gf>();
            ostride = NW;
        } else {
            if (!W2.p) W2 = ctx_buf(ctx, 16 * NW2 * m);
            o = W2.as<gf>();
            ostride = NW2;
        }
    };
    for (uint32_t k = 0; k < n; k++) {
        const size_t half = (size_t)1 << (n - 1 - k);
        if (lookahead(k, half)) {
            gf *o;
            size_t os;
            other(o, os);
            ctx->timer.run(st, "k_quad_look", [&] {
                return launch::quad_look(st, cur, cstride, o, os, dpairs, npairs, m, half, (int)k, dfs, !look);
            });
            cur = o;
            cstride = os;
            pending = true;
            look = true;
            continue;
        }
        if (pending) {  // this round's tables from the previous round's
            gf *o;
            size_t os;
            other(o, os);
            ctx->timer.run(st, "k_fix_multi",
                           [&] { return launch::fix_multi_dev(st, cur, cstride, o, os, m, 2 * half, x_dev); });
            cur = o;
            cstride = os;
            pending = false;
        }
        if (half <= tail) {
            ctx->timer.run(st, "k_quad_tail", [&] {
                return launch::quad_tail(st, cur, cstride, (gf *)(fs.as<uint8_t>() + fin_off), 1, dpairs, npairs, m,
                                         half, (int)k, dfs, dpin, (uint32_t)(fin_off + 16 * (size_t)m));
            });
            tail_ran = true;
            break;
        }
        ctx->timer.run(st, "k_quad_step", [&] {
            return launch::quad_step(st, cur, cstride, dpairs, npairs, half, (int)k, dfs, slots, max_blocks, counters);
        });
        if (lookahead(k + 1, half / 2)) {
            pending = true;  // the lookahead round folds these tables itself
        } else {
            ctx->timer.run(st, "k_fix_multi",
                           [&] { return launch::fix_multi_dev(st, cur, cstride, W.as<gf>(), NW, m, half, x_dev); });
            cur = W.as<gf>();
            cstride = NW;
        }
    }
    // the single-block rounds end by writing the region up to the finals into
    // pin; otherwise one read-back of the region and one of the finals
    if (!tail_ran) {
        hip_check(hipMemcpyAsync(pin, fs.p, in_off + 32, hipMemcpyDeviceToHost, st), "copy transcript back");
        // the m final evaluations P_p[0] of the working tables
        hip_check(hipMemcpy2DAsync(final_evals, 16, W.p, 16 * NW, 16, m, hipMemcpyDeviceToHost, st), "copy finals");
    }
    hip_check(hipStreamSynchronize(st), "sumcheck");
    ctx->timer.collect();
    memcpy(&h, pin, sizeof h);
    if (h.n != n || h.in_len != 32 || h.out_len != 16) fail(BFD_ERR_HIP, "device transcript state is corrupt");
    if (tail_ran) memcpy(final_evals, pin + fin_off, 16 * (size_t)m);
    store_elem(claim_out, h.claim);
    memcpy(round_polys, pin + quad_polys_off(npairs), 32 * (size_t)n);
    memcpy(challenges, pin + quad_chal_off(npairs, n), 16 * (size_t)n);
    ch.input.assign(pin + in_off, pin + in_off + 32);
    ch.output.assign(h.out, h.out + h.out_len);
    // the subclaim through the rounds (decompress_round_poly, quadratic.rs:70-80,
    // then SumcheckSubclaim::reduce): c1 = subclaim + c2 in characteristic 2,
    // then Horner at x
    gf sub = h.claim;
    for (uint32_t t = 0; t < n; t++) {
        const uint8_t *p = pin + quad_polys_off(npairs) + 32 * (size_t)t;
        const gf c0 = load_elem(p), c2 = load_elem(p + 16), c1 = gf_add(sub, c2);
        const gf x = load_elem(pin + quad_chal_off(npairs, n) + 16 * (size_t)t);
        sub = gf_add(c0, hmul(x, gf_add(c1, hmul(x, c2))));
    }
    store_elem(subclaim_out, sub);
}

void quadratic_args(bfd_ctx *ctx, uint32_t n, uint32_t m, const void *polys, uint32_t npairs, const uint8_t *scalars,
                    const uint32_t *pair_idx, bfd_challenger *challenger, uint8_t *claim_out, uint8_t *round_polys,
                    uint8_t *challenges, uint8_t *final_evals, uint8_t *subclaim_out) {
    if (!ctx || !polys || !scalars || !pair_idx || !challenger || !claim_out || !final_evals || !subclaim_out ||
        (n && (!round_polys || !challenges)))
        fail(BFD_ERR_ARG, "null argument");
    if (n > 32) fail(BFD_ERR_ARG, "too many variables");
    if (m == 0 || m > 65535 || npairs == 0 || npairs > 65535) fail(BFD_ERR_ARG, "need 1..65535 tables and pairs");
    for (uint32_t k = 0; k < 2 * npairs; k++)
        if (pair_idx[k] >= m) fail(BFD_ERR_ARG, "pair index out of range");
}

}  // namespace

extern "C" {

int bfd_sumcheck_quadratic(bfd_ctx *ctx, uint32_t n, uint32_t m, const uint8_t *polys, uint32_t npairs,
                           const uint8_t *scalars, const uint32_t *pair_idx, bfd_challenger *challenger,
                           uint8_t claim_out[16], uint8_t *round_polys, uint8_t *challenges, uint8_t *final_evals,
                           uint8_t subclaim_out[16]) {
    return guarded([&] {
        quadratic_args(ctx, n, m, polys, npairs, scalars, pair_idx, challenger, claim_out, round_polys, challenges,
                       final_evals, subclaim_out);
        quadratic_sumcheck(ctx, n, m, polys, false, npairs, scalars, pair_idx, challenger->c, claim_out, round_polys,
                           challenges, final_evals, subclaim_out);
    });
}

int bfd_sumcheck_quadratic_device(bfd_ctx *ctx, uint32_t n, uint32_t m, const void *d_polys, uint32_t npairs,
                                  const uint8_t *scalars, const uint32_t *pair_idx, bfd_challenger *challenger,
                                  uint8_t claim_out[16], uint8_t *round_polys, uint8_t *challenges,
                                  uint8_t *final_evals, uint8_t subclaim_out[16]) {
    return guarded([&] {
        quadratic_args(ctx, n, m, d_polys, npairs, scalars, pair_idx, challenger, claim_out, round_polys, challenges,
                       final_evals, subclaim_out);
        quadratic_sumcheck(ctx, n, m, d_polys, true, npairs, scalars, pair_idx, challenger->c, claim_out, round_polys,
                           challenges, final_evals, subclaim_out);
    });
}

int bfd_code_encode(bfd_ctx *ctx, const bfd_code *code, const void *m, int m_on_device, size_t height, size_t width,
                    void *out, int out_on_device) {
    return guarded([&] {
        if (!ctx || !code || !m || !out) fail(BFD_ERR_ARG, "null argument");
        if (width < 1 || width > kMaxWidth) fail(BFD_ERR_ARG, "width must be 1..65535");
        if (height != ((size_t)1 << code->d)) fail(BFD_ERR_ARG, "height must equal k_d = 2^d of the code");
        const int b = (int)code->log2_c, n = (int)code->d;
        const size_t N = height, M = height << b, w = width;
        if (m_on_device && out_on_device) {
            const uint8_t *a0 = (const uint8_t *)m, *o0 = (const uint8_t *)out;
            if (a0 < o0 + 16 * M * w && o0 < a0 + 16 * N * w) fail(BFD_ERR_ARG, "input and codeword overlap");
        }
        std::lock_guard<std::recursive_mutex> lk(ctx->mu);
        set_device(ctx);
        hipStream_t st = ctx->stream;
        DevBuf din, dout;
        const gf *data = (const gf *)m;
        gf *cw = (gf *)out;
        if (!m_on_device) {
            din = ctx_buf(ctx, 16 * N * w);
            hip_check(hipMemcpyAsync(din.p, m, 16 * N * w, hipMemcpyHostToDevice, st), "copy message");
            data = din.as<gf>();
        }
        if (!out_on_device) {
            dout = ctx_buf(ctx, 16 * M * w);
            cw = dout.as<gf>();
        }
        const bfd_tuning tn = ctx->tuning_for(M);
        if (w == 1) encode(ctx, code, data, cw, b, n, tn);
        else encode_w(ctx, code, data, cw, b, n, w, tn);
        if (!out_on_device) hip_check(hipMemcpyAsync(out, cw, 16 * M * w, hipMemcpyDeviceToHost, st), "copy codeword");
        hip_check(hipStreamSynchronize(st), "encode");
        ctx->timer.collect();
    });
}

int bfd_code_fold(bfd_ctx *ctx, const bfd_code *code, uint32_t i, const void *w, int on_device, const uint8_t r[16],
                  void *out) {
    return guarded([&] {
        if (!ctx || !code || !w || !r || !out) fail(BFD_ERR_ARG, "null argument");
        if (i >= code->d) fail(BFD_ERR_ARG, "fold level must be below d");
        const size_t mid = (size_t)1 << (code->log2_c + i);
        std::lock_guard<std::recursive_mutex> lk(ctx->mu);
        set_device(ctx);
        hipStream_t st = ctx->stream;
        DevBuf dw;
        const gf *src = (const gf *)w;
        gf *dst = (gf *)out;
        if (!on_device) {
            dw = ctx_buf(ctx, 16 * 2 * mid);
            hip_check(hipMemcpyAsync(dw.p, w, 32 * mid, hipMemcpyHostToDevice, st), "copy word");
            src = dw.as<gf>();
            dst = dw.as<gf>();  // in place, as the reference truncates w
        }
        const gf *ft = code->ft.as<gf>() + code->ft_off[i];
        const gf rv = load_elem(r);
        ctx->timer.run(st, "k_fold", [&] { return launch::fold(st, src, dst, ft, mid, rv); });
        if (!on_device) hip_check(hipMemcpyAsync(out, dst, 16 * mid, hipMemcpyDeviceToHost, st), "copy fold");
        hip_check(hipStreamSynchronize(st), "fold");
        ctx->timer.collect();
    });
}

int bfd_code_interpolate(const bfd_code *code, uint32_t i, uint64_t j, const uint8_t lo[16], const uint8_t hi[16],
                         const uint8_t r[16], uint8_t out[16]) {
    return guarded([&] {
        if (!code || !lo || !hi || !r || !out) fail(BFD_ERR_ARG, "null argument");
        if (i >= code->d) fail(BFD_ERR_ARG, "fold level must be below d");
        const int K = (int)(code->log2_c + i);
        if (j >> K) fail(BFD_ERR_ARG, "position beyond the level");
        gf t = gf_zero();
        for (int p = 0; p < K; p++)
            if ((j >> p) & 1) t = gf_add(t, code->bt[code->bt_off[i] + p]);
        // binary_reed_solomon.rs:103-109: hi += lo; lo += hi t; lo + (hi - lo) r
        gf l = load_elem(lo), h = load_elem(hi);
        h = gf_add(h, l);
        l = gf_add(l, hmul(h, t));
        store_elem(out, gf_add(l, hmul(gf_add(h, l), load_elem(r))));
    });
}

int bfd_field_mul(const uint8_t a[16], const uint8_t b[16], uint8_t out[16]) {
    return guarded([&] {
        if (!a || !b || !out) fail(BFD_ERR_ARG, "null argument");
        store_elem(out, hmul(load_elem(a), load_elem(b)));
    });
}

int bfd_field_mul_batch(bfd_ctx *ctx, const void *a, const void *b, int form, void *out, size_t n,
                        int on_device) {
    return guarded([&] {
        if (!ctx || ((!a || !b || !out) && n)) fail(BFD_ERR_ARG, "null argument");
        if (form < 0 || form > 2) fail(BFD_ERR_ARG, "unknown product form");
        if (!n) return;
        std::lock_guard<std::recursive_mutex> lk(ctx->mu);
        set_device(ctx);
        hipStream_t st = ctx->stream;
        const bool shared_b = form == 1;
        const size_t nb = shared_b ? 1 : n;
        DevBuf buf;
        const gf *da = (const gf *)a, *db = (const gf *)b;
        gf *dout = (gf *)out;
        if (!on_device) {
            buf = ctx_buf(ctx, 16 * (2 * n + nb));
            gf *p = buf.as<gf>();
            hip_check(hipMemcpyAsync(p, a, 16 * n, hipMemcpyHostToDevice, st), "copy a");
            hip_check(hipMemcpyAsync(p + n, b, 16 * nb, hipMemcpyHostToDevice, st), "copy b");
            da = p;
            db = p + n;
            dout = p + n + nb;
        }
        ctx->timer.run(st, "k_mul", [&] { return launch::mul(st, da, db, dout, n, form); });
        if (!on_device) hip_check(hipMemcpyAsync(out, dout, 16 * n, hipMemcpyDeviceToHost, st), "copy product");
        hip_check(hipStreamSynchronize(st), "mul");
        ctx->timer.collect();
    });
}

int bfd_field_inv(const uint8_t a[16], uint8_t out[16]) {
    return guarded([&] {
        if (!a || !out) fail(BFD_ERR_ARG, "null argument");
        store_elem(out, hinv(load_elem(a)));
    });
}

int bfd_random_vec(uint64_t seed, size_t n, uint8_t *out) {
    return guarded([&] {
        if (!out && n) fail(BFD_ERR_ARG, "null out");
        uint32_t key[8];
        seed_key(seed, key);
        const size_t words = 4 * n;
        uint32_t blk[16];
        for (size_t bidx = 0; bidx * 16 < words; bidx++) {
            chacha_block(12, key, bidx, blk);
            const size_t take = words - bidx * 16 < 16 ? words - bidx * 16 : 16;
            memcpy(out + 64 * bidx, blk, 4 * take);
        }
    });
}

int bfd_microbench(bfd_ctx *ctx, int what, size_t n, int iters, double *ms) {
    return guarded([&] {
        if (!ctx || !ms) fail(BFD_ERR_ARG, "null argument");
        if (what < 0 || what > 6 || iters < 1) fail(BFD_ERR_ARG, "unknown micro-benchmark");
        std::lock_guard<std::recursive_mutex> lk(ctx->mu);
        set_device(ctx);
        hipStream_t st = ctx->stream;
        // Keccak: 8 message words per chain; the rest: two elements per item
        const size_t bytes = what == 2 ? 64 * n : (what == 5 ? 64 : 32 * n);
        DevBuf io(bytes, st);
        std::vector<uint64_t> h(bytes / 8);
        for (size_t i = 0; i < h.size(); i++) h[i] = (i + 1) * 0x9E3779B97F4A7C15ull ^ (i >> 5);
        hip_check(hipMemcpyAsync(io.p, h.data(), bytes, hipMemcpyHostToDevice, st), "copy");
        hip_check(hipStreamSynchronize(st), "copy");  // h goes out of scope before the launches end
        hipEvent_t a, e;
        hip_check(hipEventCreate(&a), "event");
        hip_check(hipEventCreate(&e), "event");
        auto go = [&](int it) {
            if (what == 2) return launch::bench_keccak(st, io.as<uint64_t>(), n, it);
            if (what == 5) return launch::bench_keccak_pair(st, io.as<uint64_t>(), it);
            if (what == 3) return launch::copy16(st, io.as<gf>(), io.as<gf>() + n, n);
            return launch::bench_mul(st, what, io.as<gf>(), n, it);
        };
        hip_check(go(1), "bench warm-up");
        hip_check(hipEventRecord(a, st), "record");
        hip_check(go(iters), "bench");
        hip_check(hipEventRecord(e, st), "record");
        hip_check(hipEventSynchronize(e), "sync");
        float t = 0;
        hip_check(hipEventElapsedTime(&t, a, e), "elapsed");
        (void)hipEventDestroy(a);
        (void)hipEventDestroy(e);
        *ms = t;
    });
}

int bfd_ctx_kernel_times(bfd_ctx *ctx, int enable, char *names, size_t name_len, double *ms,
                         uint32_t *launches, size_t cap, size_t *count) {
    return guarded([&] {
        if (!ctx) fail(BFD_ERR_ARG, "null ctx");
        std::lock_guard<std::recursive_mutex> lk(ctx->mu);
        size_t i = 0;
        for (auto &kv : ctx->timer.totals) {
            if (i >= cap) break;
            if (names && name_len) {
                strncpy(names + i * name_len, kv.first.c_str(), name_len - 1);
                names[i * name_len + name_len - 1] = 0;
            }
            if (ms) ms[i] = kv.second.first;
            if (launches) launches[i] = kv.second.second;
            i++;
        }
        if (count) *count = i;
        ctx->timer.totals.clear();
        ctx->timer.enabled = enable != 0;
    });
}

int bfd_ctx_kernel_filter(bfd_ctx *ctx, const char *only) {
    return guarded([&] {
        if (!ctx) fail(BFD_ERR_ARG, "null ctx");
        std::lock_guard<std::recursive_mutex> lk(ctx->mu);
        ctx->timer.close_run();
        ctx->timer.only = only ? only : "";
    });
}

}  // extern "C"
