// Batched commit+open of many independent polynomials of uneven sizes.
//
// The reference commits one matrix height per Basefold instance
// (poly-commit/src/basefold.rs:94), so a batch is a set of independent
// Basefold proofs, each with its own transcript; every proof equals the one
// the plain path produces for the same polynomial.  All jobs go through the
// continuous-batching engine together: one commit tick, then one open tick
// per round for the whole batch (segmented launches, one host round trip per
// tick), instead of one launch sequence and round trip per polynomial.
#include <algorithm>

#include "engine.hpp"

using namespace bfd::rt;

namespace bfd {
namespace rt {

Lane &ctx_lane(bfd_ctx *ctx) {
    if (!ctx->lane) ctx->lane.reset(new Lane(ctx->stream, &ctx->timer));
    return *ctx->lane;
}

void batch_commit_open(bfd_ctx *ctx, uint32_t lambda, uint32_t log2_c, size_t count, const uint32_t *log_sizes,
                       const void *const *evals, bool on_device, const uint8_t *const *points,
                       const uint8_t *const *claims, Challenger *const *chs, bool observe_prefix, uint8_t *roots,
                       uint8_t **proofs, size_t *proof_lens) {
    std::lock_guard<std::recursive_mutex> lk(ctx->mu);
    set_device(ctx);
    hipStream_t st = ctx->stream;
    std::vector<std::unique_ptr<Job>> jobs(count);
    std::vector<Job *> all(count);
    size_t bytes = 0, staged = 0;
    for (size_t p = 0; p < count; p++) {
        jobs[p].reset(new Job());
        all[p] = jobs[p].get();
        if (!on_device) staged += 16 * ((size_t)1 << log_sizes[p]);
    }
    DevBuf stage = ctx_buf(ctx, staged ? staged : 1);
    size_t off = 0;
    for (size_t p = 0; p < count; p++) {
        Job &j = *jobs[p];
        const void *d = evals[p];
        if (!on_device) {
            const size_t nb = 16 * ((size_t)1 << log_sizes[p]);
            d = stage.as<uint8_t>() + off;
            hip_check(hipMemcpyAsync((void *)d, evals[p], nb, hipMemcpyHostToDevice, st), "copy evaluations");
            off += nb;
        }
        job_prepare(ctx, j, lambda, log2_c, (int)log_sizes[p], d, points[p], claims ? claims[p] : nullptr);
        j.ch = chs[p];
        j.observe_prefix = observe_prefix;
        j.proof = proofs[p];
        j.proof_cap = proofs[p] ? proof_lens[p] : 0;
        bytes += j.bytes;
    }
    if (ctx->budget && ctx->resident + bytes > ctx->budget)
        fail(BFD_ERR_BUDGET, "batch exceeds the context's HBM budget");
    // one slab for every job, from the context's block cache: a repeated
    // batch of the same shape gets the same block back without touching the
    // allocator (taking it from the pool after plain calls have split the
    // pool's free memory would map new memory instead)
    DevBuf slab = ctx_buf(ctx, bytes);
    {
        size_t o = 0;
        for (Job *j : all) {
            j->base = slab.as<uint8_t>() + o;
            o += j->bytes;
        }
    }
    Lane &lane = ctx_lane(ctx);
    lane.commit_tick(all);
    hip_check(hipStreamSynchronize(st), "commit tick");
    ctx->timer.collect();
    lane.commit_finish(all);
    std::vector<Job *> active = all;
    while (!active.empty()) {
        lane.open_tick(active);
        hip_check(hipStreamSynchronize(st), "open tick");
        ctx->timer.collect();
        lane.open_finish(active);
        std::vector<Job *> next;
        for (Job *j : active)
            if (j->stage != Job::Done) next.push_back(j);
        active.swap(next);
    }
    for (size_t p = 0; p < count; p++) {
        Job &j = *jobs[p];
        memcpy(roots + 32 * p, j.root, 32);
        proofs[p] = j.proof;
        proof_lens[p] = j.proof_len;
        j.proof = nullptr;  // ownership passes to the caller
    }
}

}  // namespace rt
}  // namespace bfd

extern "C" int bfd_commit_open_batch(bfd_ctx *ctx, uint32_t lambda, uint32_t log2_c, size_t count,
                                     const uint32_t *log_sizes, const void *const *evals, int evals_on_device,
                                     const uint8_t *const *points, const uint8_t *const *claims,
                                     bfd_challenger *const *challengers, int observe_prefix, uint8_t *roots,
                                     uint8_t **proofs, size_t *proof_lens) {
    return guarded([&] {
        if (!ctx || !log_sizes || !evals || !points || !challengers || !roots || !proofs || !proof_lens)
            fail(BFD_ERR_ARG, "null argument");
        if (count == 0) return;
        std::vector<Challenger *> chs(count);
        for (size_t p = 0; p < count; p++) {
            if (!challengers[p] || !evals[p] || !points[p] || (claims && !claims[p])) fail(BFD_ERR_ARG, "null argument");
            chs[p] = &challengers[p]->c;
        }
        // each polynomial's transcript advances on its own (concurrently)
        std::vector<Challenger *> sorted(chs);
        std::sort(sorted.begin(), sorted.end());
        if (std::adjacent_find(sorted.begin(), sorted.end()) != sorted.end())
            fail(BFD_ERR_ARG, "every polynomial needs its own challenger");
        batch_commit_open(ctx, lambda, log2_c, count, log_sizes, evals, evals_on_device != 0, points, claims,
                          chs.data(), observe_prefix != 0, roots, proofs, proof_lens);
    });
}
