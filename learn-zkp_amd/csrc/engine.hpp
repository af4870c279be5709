// Continuous-batching engine behind the batched and serving paths.
//
// A Job is one independent Basefold commit+open (the reference commits one
// matrix height per instance, poly-commit/src/basefold.rs:94).  Jobs advance
// in ticks: one tick launches, for every job it holds, the next step of that
// job (the whole commitment; or one open round: fold/fix with the previous
// challenge, the next pi tree and inner product; or the final gather), each
// step as one segmented launch per kernel shared by all jobs, then a single
// stream synchronisation hands roots and sums to the per-job transcripts.
#pragma once
#include "internal.hpp"

namespace bfd {
namespace rt {

struct Job {
    // request
    int n = 0, b = 0, L = 0;
    size_t N = 0, M = 0;
    const bfd_code *code = nullptr;
    const gf *data = nullptr;          // evaluations in device memory
    std::vector<uint8_t> point;        // n x 16 bytes
    bool have_claim = false;           // claim given (else evaluated on device)
    gf claim{};
    bool observe_prefix = true;        // observe(root), observe(claim) first
    Challenger *ch = nullptr;          // transcript (owned by the caller or the job)
    Challenger own_ch;
    // device buffers: carved from `base` (an arena or a batch slab) when set,
    // else from an allocation of its own
    DevBuf slab;
    uint8_t *base = nullptr;
    size_t arena_off = 0;
    size_t bytes = 0;
    gf *cw = nullptr, *half_eq = nullptr, *poly = nullptr, *pis = nullptr;
    uint64_t *tree = nullptr, *pitrees = nullptr;
    OpenLayout lay;
    OpenRounds o;
    // progress
    enum Stage { Commit, Open, Gather, Done } stage = Commit;
    int k = 0;                         // next open round
    gf x_prev{};
    const gf *poly_cur = nullptr;
    std::vector<gf> eqA, eqB;
    int eq_lo = 0;
    // per-tick bookkeeping
    uint32_t dot_blk0 = 0, dot_nblk = 0;
    size_t slot = 0;
    size_t ne = 0, nd = 0;
    // results
    uint8_t root[32] = {0};
    uint8_t *proof = nullptr;          // caller buffer or library allocation
    size_t proof_cap = 0, proof_len = 0;
    bool proof_owned = false;
    bool finished = false;             // set under the server mutex once Done
    // timestamps (ms, steady clock)
    double t_submit = 0, t_admit = 0, t_done = 0;
};

// Device bytes a job holds while it is resident.
size_t job_bytes(const Job &j);
// Fills the request-derived fields of a job.
void job_prepare(bfd_ctx *ctx, Job &j, uint32_t lambda, uint32_t log2_c, int n, const void *d_evals,
                 const uint8_t *point, const uint8_t *claim);

// First-fit sub-allocator over one device allocation (serving: the HBM
// budget is reserved once instead of mapping memory per request).
class Arena {
  public:
    explicit Arena(size_t bytes);
    ~Arena();
    uint8_t *base() const { return base_; }
    size_t size() const { return size_; }
    // returns false when no free range fits (fragmentation or full)
    bool alloc(size_t bytes, size_t &off);
    void free(size_t off, size_t bytes);

  private:
    uint8_t *base_ = nullptr;
    size_t size_ = 0;
    std::map<size_t, size_t> free_;  // offset -> length, coalesced
};

// One execution lane: a stream with its own staging buffers.
class Lane {
  public:
    Lane(hipStream_t st, KernelTimer *timer) : st_(st), timer_(timer) {}
    hipStream_t stream() const { return st_; }
    // enqueue a commit tick / an open tick for `jobs`; the caller then waits
    // for the stream and calls the matching finish
    void commit_tick(std::vector<Job *> &jobs);
    void commit_finish(std::vector<Job *> &jobs);
    void open_tick(std::vector<Job *> &jobs);
    void open_finish(std::vector<Job *> &jobs);

  private:
    uint32_t seg(SegKernel k, std::vector<Seg> &segs, const char *name, gf *partial = nullptr,
                 uint32_t max_blocks_per_seg = 2048);
    void reset_ring();
    void flush();
    void tree_levels();
    hipStream_t st_;
    KernelTimer *timer_;
    // pinned staging; reused from tick to tick (a tick starts only after the
    // previous one has completed on the stream)
    PinnedBuf seg_pin_, rb_pin_, gather_pin_, stage_pin_, plan_pin_;
    DevBuf seg_dev_, rb_dev_, addr_dev_, partials_, gather_dev_, plan_dev_, eq_dev_;
    size_t seg_cap_ = 0, seg_used_ = 0, flushed_ = 0;
    struct Pending {  // a recorded segmented launch (ring entries at off)
        SegKernel k;
        size_t off;
        int nseg;
        uint32_t blocks;
        gf *partial;
        const char *name;
    };
    std::vector<Pending> pending_;
    std::vector<Seg> segs_;
    std::vector<std::pair<uint64_t *, size_t>> trees_;  // leaf layer, leaf count
    std::vector<std::vector<TreeStep>> plans_;
    std::vector<size_t> singles_;  // leading k_nodes-form steps of each plan
    std::vector<std::pair<size_t, size_t>> gather_at_;  // per job: first element / digest record
    size_t gather_NE_ = 0;
};

}  // namespace rt
}  // namespace bfd
