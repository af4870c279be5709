// Batched commit+open of many independent polynomials of uneven sizes.
//
// The reference commits one matrix height per Basefold instance
// (poly-commit/src/basefold.rs:94), so a batch is a set of independent
// Basefold proofs, each with its own transcript; every proof equals the one
// the plain path produces for the same polynomial.  All jobs go through the
// continuous-batching engine together: one commit tick, then one open tick
// per round for the whole batch (segmented launches, transcripts on the
// device), with the next tick enqueued while the previous one runs, instead
// of one launch sequence per polynomial.
#include <algorithm>
#include <deque>

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
    double t = now_ms();
    // host phases, reported as host:batch_* when the context's timer is on
    auto phase = [&](const char *name) {
        const double t1 = now_ms();
        ctx->timer.host(name, t1 - t);
        t = t1;
    };
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
        job_prepare(ctx, j, lambda, log2_c, (int)log_sizes[p], d, points[p], claims ? claims[p] : nullptr, chs[p],
                    observe_prefix);
        j.proof = proofs[p];
        j.proof_cap = proofs[p] ? proof_lens[p] : 0;
        bytes += j.bytes;
    }
    if (ctx->budget && ctx->resident + bytes > ctx->budget)
        fail(BFD_ERR_BUDGET, "batch exceeds the context's HBM budget");
    // caller buffers are checked before anything is enqueued, so a refused
    // call leaves every challenger as it was
    for (Job *j : all)
        if (j->proof && j->proof_cap < proof_size(j->b, j->n, j->code->q)) fail(BFD_ERR_ARG, "proof buffer too small");
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
    lane.set_tuning(ctx->batch_tuning, ctx->cus);
    phase("batch_prepare");
    lane.commit_tick(all);
    phase("batch_enqueue");
    // The proofs are written in the last ticks, when the device has little
    // left to do; touching their pages now, while the device runs the commit
    // tick, keeps the page faults of freshly allocated output buffers out of
    // those ticks.
    {
        std::vector<std::pair<uint8_t *, size_t>> chunks;
        for (Job *j : all) {
            const size_t sz = proof_size(j->b, j->n, j->code->q);
            if (!j->proof) {
                j->proof = (uint8_t *)malloc(sz ? sz : 1);
                if (!j->proof) fail(BFD_ERR_ARG, "host allocation failed");
                j->proof_owned = true;
                j->proof_cap = sz;
            }
            for (size_t o = 0; o < sz; o += (size_t)1 << 20)
                chunks.push_back({j->proof + o, std::min(sz - o, (size_t)1 << 20)});
        }
        parallel_for(chunks.size(), [&](size_t c) { memset(chunks[c].first, 0, chunks[c].second); });
        phase("batch_prefault");
    }
    hip_check(hipStreamSynchronize(st), "commit tick");
    phase("batch_wait");
    ctx->timer.collect();
    lane.commit_finish(all);
    phase("batch_finish");
    // open ticks, up to Lane::kSets in flight: while the oldest runs, the
    // next ones are enqueued, so the host's share of a tick (sampling
    // queries, assembling proofs) overlaps the device's
    std::deque<OpenTick> inflight;
    std::vector<Job *> work;
    // enqueues the next tick when a buffer set is free (the oldest tick's
    // set stays taken until its proofs are written)
    const auto enqueue = [&] {
        if (inflight.size() >= (size_t)Lane::kSets) return false;
        work.clear();
        for (Job *j : all)
            if (j->stage == Job::Open || j->stage == Job::Gather) work.push_back(j);
        if (work.empty()) return false;
        inflight.push_back(lane.open_tick(work));
        phase("batch_enqueue");
        return true;
    };
    for (;;) {
        if (enqueue()) continue;
        if (inflight.empty()) break;
        hip_check(hipEventSynchronize(inflight.front().done), "open tick");
        phase("batch_wait");
        ctx->timer.collect();
        // sample the finished jobs' queries first, so that the tick that
        // gathers their openings runs while the host writes proofs
        lane.open_finals(inflight.front());
        phase("batch_finish");
        enqueue();
        lane.open_proofs(inflight.front());
        inflight.pop_front();
        phase("batch_finish");
    }
    for (Job *j : all)
        if (j->stage != Job::Done) fail(BFD_ERR_HIP, "batched opening did not complete");
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
