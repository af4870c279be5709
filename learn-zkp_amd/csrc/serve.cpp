// Serving loop: continuous batching of commit-then-open requests.
//
// Requests arrive at any time (bfd_server_submit) with their evaluations in
// device memory and an opening point.  A worker thread drives two HIP
// streams (default priority): the commit lane runs one commit tick over the
// requests admitted since its last tick, the open lane runs one open tick
// (one round, or the final gather, of every committed request).  Both ticks
// are in flight together, so newly arriving work overlaps the rounds of
// earlier requests on the device; and since the rounds' transcripts run on
// the device, the open lane keeps a second tick enqueued behind the running
// one.  A request is admitted when its resident
// footprint (codeword, all Merkle layers, pi vectors and trees) fits the
// server's HBM budget; its memory is released when its proof is done.  Every
// request uses a fresh transcript with the harness prefix (observe root,
// observe claim), and its claim is evaluated on the device, so each proof is
// byte-equal to a plain commit + eval + open of the same polynomial.
#include <condition_variable>
#include <deque>

#include "engine.hpp"

using namespace bfd::rt;

struct bfd_server {
    bfd_ctx *ctx = nullptr;
    uint32_t lambda_default = 128;
    uint64_t budget = 0;
    uint32_t max_active = 0;
    hipStream_t s_commit = nullptr, s_open = nullptr;
    std::unique_ptr<Lane> commit_lane, open_lane;
    std::unique_ptr<Arena> arena;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<Job *> pending;                       // submitted, not admitted
    std::map<uint64_t, std::unique_ptr<Job>> jobs;   // until collected by wait()
    std::map<Job *, uint64_t> ticket_of;
    uint64_t next_ticket = 1;
    uint64_t resident = 0, completed = 0, active = 0;
    bool stopping = false;
    std::string fatal;
    // timing counters (ms): commit ticks, open ticks, lane busy time (launch to
    // completion observed), host time in finishes, host time enqueueing ticks
    double t_commit_busy = 0, t_open_busy = 0, t_finish = 0, t_enqueue = 0;
    uint64_t n_commit_ticks = 0, n_open_ticks = 0, n_open_jobs = 0;
    std::thread worker;
};

namespace {

void worker_loop(bfd_server *S) {
    try {
        hip_check(hipSetDevice(S->ctx->device), "hipSetDevice");
        hipEvent_t ev_commit;
        hip_check(hipEventCreateWithFlags(&ev_commit, hipEventDisableTiming), "event");
        std::vector<Job *> committing, opening, work;
        // open ticks in flight (at most Lane::kSets, one per buffer set: the
        // rounds need nothing from the host, so the next ticks are enqueued
        // while one runs) and the time each was enqueued
        std::deque<std::pair<OpenTick, double>> open_q;
        bool commit_busy = false;
        double commit_t0 = 0;
        for (;;) {
            // completions
            if (commit_busy && hipEventQuery(ev_commit) == hipSuccess) {
                const double t1 = now_ms();
                S->t_commit_busy += t1 - commit_t0;
                S->commit_lane->commit_finish(committing);
                S->t_finish += now_ms() - t1;
                opening.insert(opening.end(), committing.begin(), committing.end());
                committing.clear();
                commit_busy = false;
            }
            while (!open_q.empty() && hipEventQuery(open_q.front().first.done) == hipSuccess) {
                OpenTick &ot = open_q.front().first;
                const double t1 = now_ms();
                S->t_open_busy += t1 - open_q.front().second;
                S->open_lane->open_finish(ot);
                S->t_finish += now_ms() - t1;
                std::vector<Job *> done = ot.gathered;  // a tick's gathers finish their jobs
                if (!done.empty()) {
                    std::vector<Job *> keep;
                    for (Job *j : opening)
                        if (j->stage != Job::Done) keep.push_back(j);
                    opening.swap(keep);
                    std::lock_guard<std::mutex> lk(S->mu);
                    for (Job *j : done) {
                        S->arena->free(j->arena_off, j->bytes);  // the device is done with it
                        S->resident -= j->bytes;
                        S->active--;
                        S->completed++;
                        j->finished = true;
                    }
                    S->cv_done.notify_all();
                    S->cv_work.notify_all();
                }
                open_q.pop_front();
            }
            // admission and launches
            if (!commit_busy) {
                {
                    std::lock_guard<std::mutex> lk(S->mu);
                    while (!S->pending.empty()) {
                        Job *j = S->pending.front();
                        if (S->budget && S->resident + j->bytes > S->budget) break;  // FIFO: wait for memory
                        if (S->max_active && S->active >= S->max_active) break;
                        size_t off = 0;
                        if (!S->arena->alloc(j->bytes, off)) break;              // fragmented: wait
                        j->base = S->arena->base() + off;
                        j->arena_off = off;
                        S->pending.pop_front();
                        S->resident += j->bytes;
                        S->active++;
                        j->t_admit = now_ms();
                        committing.push_back(j);
                    }
                }
                if (!committing.empty()) {
                    commit_t0 = now_ms();
                    S->commit_lane->commit_tick(committing);
                    hip_check(hipEventRecord(ev_commit, S->s_commit), "record");
                    S->t_enqueue += now_ms() - commit_t0;
                    S->n_commit_ticks++;
                    commit_busy = true;
                }
            }
            if (open_q.size() < (size_t)Lane::kSets) {
                work.clear();
                for (Job *j : opening)
                    if (j->stage == Job::Open || j->stage == Job::Gather) work.push_back(j);
                if (!work.empty()) {
                    const double t0 = now_ms();
                    open_q.push_back({S->open_lane->open_tick(work), t0});
                    S->t_enqueue += now_ms() - t0;
                    S->n_open_ticks++;
                    S->n_open_jobs += work.size();
                }
            }
            if (!commit_busy && open_q.empty()) {
                std::unique_lock<std::mutex> lk(S->mu);
                if (S->stopping && S->pending.empty()) break;
                // nothing in flight: everything admitted has finished, so the
                // arena is empty and the head of the queue fits
                S->cv_work.wait(lk, [&] { return S->stopping || !S->pending.empty(); });
                if (S->stopping && S->pending.empty()) break;
            } else {
                std::this_thread::sleep_for(std::chrono::microseconds(20));
            }
        }
        (void)hipEventDestroy(ev_commit);
    } catch (const Error &e) {
        std::lock_guard<std::mutex> lk(S->mu);
        S->fatal = e.msg;
        S->cv_done.notify_all();
    } catch (const std::exception &e) {
        std::lock_guard<std::mutex> lk(S->mu);
        S->fatal = e.what();
        S->cv_done.notify_all();
    }
}

}  // namespace

extern "C" {

int bfd_server_create(bfd_ctx *ctx, uint64_t hbm_budget_bytes, uint32_t max_active, bfd_server **out) {
    return guarded([&] {
        if (!ctx || !out) fail(BFD_ERR_ARG, "null argument");
        set_device(ctx);
        auto S = std::make_unique<bfd_server>();
        S->ctx = ctx;
        S->budget = hbm_budget_bytes;
        S->max_active = max_active;
        // streams made here are destroyed again if set-up fails below
        struct StreamGuard {
            bfd_server *s;
            ~StreamGuard() {
                if (!s) return;
                if (s->s_commit) (void)hipStreamDestroy(s->s_commit);
                if (s->s_open) (void)hipStreamDestroy(s->s_open);
            }
        } guard{S.get()};
        hip_check(hipStreamCreateWithFlags(&S->s_commit, hipStreamNonBlocking), "hipStreamCreate");
        hip_check(hipStreamCreateWithFlags(&S->s_open, hipStreamNonBlocking), "hipStreamCreate");
        S->commit_lane.reset(new Lane(S->s_commit, nullptr));
        S->open_lane.reset(new Lane(S->s_open, nullptr));
        {
            std::lock_guard<std::recursive_mutex> lk(ctx->mu);
            S->commit_lane->set_tuning(ctx->batch_tuning, ctx->cus);
            S->open_lane->set_tuning(ctx->batch_tuning, ctx->cus);
        }
        // hand memory held for the plain path back to the device first
        {
            std::lock_guard<std::recursive_mutex> lk(ctx->mu);
            ctx->cache.flush();
            hip_check(hipStreamSynchronize(ctx->stream), "synchronize");
            hipMemPool_t pool;
            if (hipDeviceGetDefaultMemPool(&pool, ctx->device) == hipSuccess) (void)hipMemPoolTrimTo(pool, 0);
        }
        // reserve the budget once (unlimited budget: 3/4 of the free memory)
        size_t free_b = 0, total_b = 0;
        hip_check(hipMemGetInfo(&free_b, &total_b), "hipMemGetInfo");
        size_t arena = hbm_budget_bytes ? (size_t)hbm_budget_bytes : free_b / 4 * 3;
        if (arena > free_b / 20 * 19) fail(BFD_ERR_BUDGET, "HBM budget exceeds the free device memory");
        S->arena.reset(new Arena(arena));
        if (!hbm_budget_bytes) S->budget = arena;
        bfd_server *raw = S.get();
        S->worker = std::thread(worker_loop, raw);
        guard.s = nullptr;
        *out = S.release();
    });
}

int bfd_server_submit(bfd_server *S, uint32_t lambda, uint32_t log2_c, uint32_t n, const void *d_evals,
                      const uint8_t *point, uint64_t *ticket) {
    return guarded([&] {
        if (!S || !d_evals || !point || !ticket) fail(BFD_ERR_ARG, "null argument");
        auto j = std::make_unique<Job>();
        job_prepare(S->ctx, *j, lambda, log2_c, (int)n, d_evals, point, nullptr, &j->own_ch, true);
        if (S->budget && j->bytes > S->budget) fail(BFD_ERR_BUDGET, "request larger than the server's HBM budget");
        j->t_submit = now_ms();
        std::lock_guard<std::mutex> lk(S->mu);
        if (!S->fatal.empty()) fail(BFD_ERR_HIP, "server stopped: " + S->fatal);
        if (S->stopping) fail(BFD_ERR_STATE, "server is shutting down");
        const uint64_t t = S->next_ticket++;
        S->pending.push_back(j.get());
        S->ticket_of[j.get()] = t;
        S->jobs[t] = std::move(j);
        *ticket = t;
        S->cv_work.notify_all();
    });
}

int bfd_server_wait(bfd_server *S, uint64_t ticket, int64_t timeout_ms, uint8_t root[32], uint8_t claim[16],
                    uint8_t **proof, size_t *proof_len, double *latency_ms) {
    return guarded([&] {
        if (!S || !proof || !proof_len) fail(BFD_ERR_ARG, "null argument");
        std::unique_lock<std::mutex> lk(S->mu);
        auto it = S->jobs.find(ticket);
        if (it == S->jobs.end()) fail(BFD_ERR_ARG, "unknown ticket");
        Job *j = it->second.get();
        auto ready = [&] { return j->finished || !S->fatal.empty(); };
        if (timeout_ms < 0) S->cv_done.wait(lk, ready);
        else if (!S->cv_done.wait_for(lk, std::chrono::milliseconds(timeout_ms), ready))
            fail(BFD_ERR_STATE, "request not finished yet");
        if (!j->finished) fail(BFD_ERR_HIP, "server stopped: " + S->fatal);
        if (root) memcpy(root, j->root, 32);
        if (claim) store_elem(claim, j->claim);
        *proof = j->proof;
        *proof_len = j->proof_len;
        if (latency_ms) *latency_ms = j->t_done - j->t_submit;
        j->proof = nullptr;
        S->ticket_of.erase(j);
        S->jobs.erase(it);
    });
}

int bfd_server_stats(bfd_server *S, uint64_t *pending, uint64_t *active, uint64_t *resident_bytes,
                     uint64_t *completed) {
    return guarded([&] {
        if (!S) fail(BFD_ERR_ARG, "null server");
        std::lock_guard<std::mutex> lk(S->mu);
        if (pending) *pending = S->pending.size();
        if (active) *active = S->active;
        if (resident_bytes) *resident_bytes = S->resident;
        if (completed) *completed = S->completed;
    });
}

int bfd_server_timing(bfd_server *S, double out[8]) {
    return guarded([&] {
        if (!S || !out) fail(BFD_ERR_ARG, "null argument");
        std::lock_guard<std::mutex> lk(S->mu);
        out[0] = (double)S->n_commit_ticks;
        out[1] = (double)S->n_open_ticks;
        out[2] = S->t_commit_busy;
        out[3] = S->t_open_busy;
        out[4] = S->t_finish;
        out[5] = S->t_enqueue;
        out[6] = (double)S->n_open_jobs;
        out[7] = 0;
    });
}

int bfd_server_destroy(bfd_server *S) {
    return guarded([&] {
        if (!S) return;
        {
            std::lock_guard<std::mutex> lk(S->mu);
            S->stopping = true;
            S->cv_work.notify_all();
        }
        if (S->worker.joinable()) S->worker.join();
        (void)hipSetDevice(S->ctx->device);
        (void)hipStreamSynchronize(S->s_open);
        (void)hipStreamSynchronize(S->s_commit);
        S->jobs.clear();  // frees proofs never collected
        S->commit_lane.reset();
        S->open_lane.reset();
        (void)hipStreamSynchronize(S->s_open);
        (void)hipStreamSynchronize(S->s_commit);
        S->arena.reset();
        (void)hipStreamSynchronize(S->s_commit);
        (void)hipStreamDestroy(S->s_commit);
        (void)hipStreamDestroy(S->s_open);
        delete S;
    });
}

}  // extern "C"
