// Continuous-batching engine behind the batched and serving paths.
//
// A Job is one independent Basefold commit+open (the reference commits one
// matrix height per instance, poly-commit/src/basefold.rs:94).  Jobs advance
// in ticks: one tick launches, for every job it holds, the next step of that
// job (the whole commitment; or one open round: fold/fix with the previous
// challenge, the next pi tree, inner product and transcript step; or the
// final gather), each step as one segmented launch per kernel shared by all
// jobs.  The rounds' transcripts live on the device, so an open tick needs
// nothing from the host and the next one can be enqueued before it ends;
// the host steps in once per job, to sample the queries after its last
// round and to copy its query records behind the proof header after the
// gather.
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
    uint32_t fs_cap = 0;               // input capacity of the device transcript
    // device buffers: carved from `base` (an arena or a batch slab) when set,
    // else from an allocation of its own
    DevBuf slab;
    uint8_t *base = nullptr;
    size_t arena_off = 0;
    size_t bytes = 0;
    gf *cw = nullptr, *half_eq = nullptr, *poly = nullptr, *pis = nullptr;
    uint64_t *tree = nullptr, *pitrees = nullptr;
    uint8_t *fs = nullptr;             // device transcript region (FsState)
    OpenLayout lay;
    OpenRounds o;
    // progress
    // Open: rounds on the device; Final: its last round's read-back is in
    // flight; Gather: queries sampled, openings to gather; Gathering: the
    // gather is in flight
    enum Stage { Commit, Open, Final, Gather, Gathering, Done } stage = Commit;
    int k = 0;                         // next open round to enqueue (n: the final fold)
    const gf *poly_cur = nullptr;
    std::vector<gf> eqA, eqB;
    int eq_lo = 0;
    // per-tick bookkeeping
    size_t slot = 0;
    // results
    uint8_t root[32] = {0};
    uint8_t *proof = nullptr;          // caller buffer or library allocation
    size_t proof_cap = 0, proof_len = 0;
    bool proof_owned = false;
    bool finished = false;             // set under the server mutex once Done
    // timestamps (ms, steady clock)
    double t_submit = 0, t_admit = 0, t_done = 0;
    Job() = default;
    Job(const Job &) = delete;
    Job &operator=(const Job &) = delete;
    // a library-allocated proof not handed to the caller dies with the job
    ~Job() {
        if (proof_owned && proof) free(proof);
    }
};

// Device bytes a job holds while it is resident.
size_t job_bytes(const Job &j);
// Fills the request-derived fields of a job; ch is its transcript (its
// current input sizes the device transcript region).
void job_prepare(bfd_ctx *ctx, Job &j, uint32_t lambda, uint32_t log2_c, int n, const void *d_evals,
                 const uint8_t *point, const uint8_t *claim, Challenger *ch, bool observe_prefix);

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

// What the host does after an open tick has completed on the stream.
struct OpenTick {
    int par = 0;                                       // staging half the tick used
    hipEvent_t done = nullptr;                         // recorded after its last operation
    std::vector<std::pair<Job *, size_t>> finals;      // job, offset of its read-back in the staging
    std::vector<Job *> gathered;                       // jobs whose openings it gathered
    std::vector<size_t> gather_at;                     // per gathered job: offset of its query records
};

// One execution lane: a stream with its own staging buffers.
class Lane {
  public:
    // open ticks that may be in flight at once (each stages through its own
    // set of buffers): the host's share of a tick (sampling the queries of
    // jobs past their last round, copying finished proofs) overlaps the
    // device's next ticks, which at the end of a batch are short
    static constexpr int kSets = 3;
    Lane(hipStream_t st, KernelTimer *timer);
    ~Lane();
    hipStream_t stream() const { return st_; }
    // enqueue a commit tick for `jobs`; the caller then waits for the stream
    // and calls commit_finish
    void commit_tick(std::vector<Job *> &jobs);
    void commit_finish(std::vector<Job *> &jobs);
    // enqueue an open tick for the jobs with device work (stage Open or
    // Gather); at most kSets ticks may be in flight, finished in order with
    // open_finish once t.done has completed
    OpenTick open_tick(std::vector<Job *> &jobs);
    void open_finish(OpenTick &t);
    // open_finish in two parts: open_finals samples the queries of the jobs
    // past their last round (after it, the next tick can gather them);
    // open_proofs writes the proofs of the jobs this tick gathered, and
    // until it has run the tick's buffer set stays taken
    void open_finals(OpenTick &t);
    void open_proofs(OpenTick &t);
    // launch shape of the ticks enqueued from now on (its ntt_order,
    // subtree_from_log2, subtree_levels and seg_blocks_per_cu apply)
    void set_tuning(const bfd_tuning &t, int cus) {
        tune_ = t;
        cus_ = cus;
    }

  private:
    // max_blocks_per_seg = 0: the shape's seg_blocks_per_cu per compute unit
    uint32_t seg(SegKernel k, std::vector<Seg> &segs, const char *name, gf *partial = nullptr,
                 uint32_t max_blocks_per_seg = 0);
    void reset_ring();
    void flush();
    void tree_levels();
    hipStream_t st_;
    KernelTimer *timer_;
    bfd_tuning tune_ = default_tuning();
    int cus_ = 256;
    // staging in kSets sets, one per tick in turn: up to kSets open ticks
    // may be in flight, so a tick must not overwrite what the ones before it
    // have yet to read
    int cur_ = 0;
    hipEvent_t ev_[kSets] = {};
    PinnedBuf seg_pin_[kSets], rb_pin_[kSets], gather_pin_[kSets], stage_pin_[kSets], plan_pin_[kSets];
    DevBuf seg_dev_[kSets], rb_dev_[kSets], addr_dev_[kSets], plan_dev_[kSets], eq_dev_[kSets];
    DevBuf partials_;  // inner-product sums (stream-ordered: one tick at a time uses them)
    size_t seg_cap_[kSets] = {}, seg_used_[kSets] = {}, flushed_[kSets] = {};
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
};

}  // namespace rt
}  // namespace bfd
