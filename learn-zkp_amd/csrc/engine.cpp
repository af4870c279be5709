// Continuous-batching engine (see engine.hpp).
#include "engine.hpp"

namespace bfd {
namespace rt {

namespace {

size_t align_up(size_t v) { return (v + 255) & ~(size_t)255; }

Seg make_seg(const void *in, void *out, const void *aux, uint64_t count, uint64_t arg, gf x = gf_zero()) {
    Seg g;
    g.in = in;
    g.out = out;
    g.aux = aux;
    g.count = count;
    g.arg = arg;
    g.x = x;
    g.blk0 = g.nblk = 0;
    return g;
}

// grows a stream-ordered device buffer (contents are not preserved)
void ensure(DevBuf &b, size_t bytes, hipStream_t st) {
    if (b.bytes >= bytes) return;
    b = DevBuf(bytes + bytes / 4 + 4096, st);
}

}  // namespace

size_t job_bytes(const Job &j) {
    const size_t c = (size_t)1 << j.b;
    return align_up(16 * j.M) + align_up(32 * (2 * j.M - 1)) + align_up(16 * ((size_t)1 << (j.n - 1))) +
           align_up(16 * (j.N / 2)) + align_up(16 * j.lay.pis_elems(j.M, c)) + align_up(32 * j.lay.pitree_digests()) +
           align_up(fs_bytes((uint32_t)j.n, j.fs_cap));
}

void job_prepare(bfd_ctx *ctx, Job &j, uint32_t lambda, uint32_t log2_c, int n, const void *d_evals,
                 const uint8_t *point, const uint8_t *claim, Challenger *ch, bool observe_prefix) {
    if (n < 1 || n > 30) fail(BFD_ERR_ARG, "batched and served polynomials need 1..30 variables");
    if (log2_c > 8) fail(BFD_ERR_ARG, "log2_c too large");
    j.n = n;
    j.b = (int)log2_c;
    j.L = n + j.b;
    j.N = (size_t)1 << n;
    j.M = j.N << j.b;
    j.code = cached_code(ctx, lambda, log2_c, (uint32_t)n);
    j.data = (const gf *)d_evals;
    j.point.assign(point, point + 16 * (size_t)n);
    for (int i = 0; i < n; i++) {
        gf r = load_elem(point + 16 * i);
        if (gf_eq(r, gf_zero()) || gf_eq(r, gf_one()))
            fail(BFD_ERR_ARG, "point coordinates 0 and 1 are not invertible in the Eval prover");
    }
    j.have_claim = claim != nullptr;
    if (claim) j.claim = load_elem(claim);
    j.ch = ch;
    j.observe_prefix = observe_prefix;
    // the open starts from the transcript's current input plus the prefix
    // (32-byte root, 16-byte claim)
    j.fs_cap = fs_in_cap(ch->input.size() + (observe_prefix ? 48 : 0));
    j.lay.plan(j.M, n);
    j.bytes = job_bytes(j);
}

Arena::Arena(size_t bytes) : size_(bytes) {
    hip_check(hipMalloc((void **)&base_, bytes), "hipMalloc arena");
    free_[0] = bytes;
}
Arena::~Arena() {
    if (base_) (void)hipFree(base_);
}
bool Arena::alloc(size_t bytes, size_t &off) {
    bytes = align_up(bytes);
    for (auto it = free_.begin(); it != free_.end(); ++it) {
        if (it->second < bytes) continue;
        off = it->first;
        const size_t rest = it->second - bytes;
        free_.erase(it);
        if (rest) free_[off + bytes] = rest;
        return true;
    }
    return false;
}
void Arena::free(size_t off, size_t bytes) {
    bytes = align_up(bytes);
    auto it = free_.emplace(off, bytes).first;
    auto nx = std::next(it);
    if (nx != free_.end() && it->first + it->second == nx->first) {
        it->second += nx->second;
        free_.erase(nx);
    }
    if (it != free_.begin()) {
        auto pv = std::prev(it);
        if (pv->first + pv->second == it->first) {
            pv->second += it->second;
            free_.erase(it);
        }
    }
}

// Carves a job's buffers out of `j.base` when set, else out of one
// allocation of its own on `st`.
static void job_alloc(Job &j, hipStream_t st) {
    const size_t c = (size_t)1 << j.b;
    uint8_t *cur = j.base;
    if (!cur) {
        j.slab = DevBuf(j.bytes, st);
        cur = j.slab.as<uint8_t>();
    }
    auto take = [&](size_t n) {
        uint8_t *r = cur;
        cur += align_up(n);
        return r;
    };
    j.cw = (gf *)take(16 * j.M);
    j.tree = (uint64_t *)take(32 * (2 * j.M - 1));
    j.half_eq = (gf *)take(16 * ((size_t)1 << (j.n - 1)));
    j.poly = (gf *)take(16 * (j.N / 2));
    j.pis = (gf *)take(16 * j.lay.pis_elems(j.M, c));
    j.pitrees = (uint64_t *)take(32 * j.lay.pitree_digests());
    j.fs = take(fs_bytes((uint32_t)j.n, j.fs_cap));
    j.lay.cw = j.cw;
    j.lay.tree = j.tree;
    j.lay.pis = j.pis;
    j.lay.pitrees = j.pitrees;
}

void LaneDeleter::operator()(Lane *l) const { delete l; }

Lane::Lane(hipStream_t st, KernelTimer *timer) : st_(st), timer_(timer) {
    for (auto &e : ev_) hip_check(hipEventCreateWithFlags(&e, hipEventDisableTiming), "hipEventCreate");
}
Lane::~Lane() {
    for (auto &e : ev_)
        if (e) (void)hipEventDestroy(e);
}

void Lane::reset_ring() {
    flush();
    seg_used_[cur_] = 0;
    flushed_[cur_] = 0;
}

// Records a segmented launch: its table goes into the pinned ring now, the
// upload and the launch itself wait for flush(), so a tick's tables travel in
// one copy instead of one per launch.
uint32_t Lane::seg(SegKernel k, std::vector<Seg> &segs, const char *name, gf *partial, uint32_t max_blocks) {
    const uint64_t max_blocks_per_seg = max_blocks ? max_blocks : (uint64_t)tune_.seg_blocks_per_cu * (uint64_t)cus_;
    if (segs.empty()) return 0;
    const int c = cur_;
    if (seg_used_[c] + segs.size() > seg_cap_[c]) {
        // the tick outgrew its staging ring: run what it holds, drain the
        // stream, then start over
        flush();
        hip_check(hipStreamSynchronize(st_), "segment ring sync");
        seg_used_[c] = 0;
        flushed_[c] = 0;
        if (segs.size() > seg_cap_[c]) {
            seg_cap_[c] = 2 * segs.size() + 256;
            seg_pin_[c].ensure(sizeof(Seg) * seg_cap_[c]);
            seg_dev_[c] = DevBuf(sizeof(Seg) * seg_cap_[c], st_);
        }
    }
    // items per block: one subtree per block for Subtree, one opening per
    // block for Round, else one per thread
    const int bs = k == SegKernel::Subtree ? 1 : launch::seg_block_size(k);
    uint32_t blocks = 0;
    for (auto &g : segs) {
        uint64_t want = (g.count + bs - 1) / bs;
        if (want < 1 || k == SegKernel::Round) want = 1;
        if (want > max_blocks_per_seg) want = max_blocks_per_seg;
        if (g.nblk && want > g.nblk) want = g.nblk;  // the segment's own cap
        g.blk0 = blocks;
        g.nblk = (uint32_t)want;
        blocks += g.nblk;
    }
    // subtree levels narrow enough for the whole launch to leave SIMDs idle
    // hash on lane pairs
    if (k == SegKernel::Subtree) {
        const int pw = launch::pair_width(blocks, cus_);
        for (auto &g : segs) g.x = gf{(uint64_t)pw, 0};
    }
    memcpy((Seg *)seg_pin_[c].p + seg_used_[c], segs.data(), sizeof(Seg) * segs.size());
    pending_.push_back(Pending{k, seg_used_[c], (int)segs.size(), blocks, partial, name});
    seg_used_[c] += segs.size();
    return blocks;
}

// Uploads the ring entries recorded since the last flush in one copy and
// enqueues their launches in order.  Every stream operation that depends on
// a recorded launch (read-backs, gathers) comes after a flush.
void Lane::flush() {
    if (pending_.empty()) return;
    const int c = cur_;
    hip_check(hipMemcpyAsync(seg_dev_[c].as<Seg>() + flushed_[c], (Seg *)seg_pin_[c].p + flushed_[c],
                             sizeof(Seg) * (seg_used_[c] - flushed_[c]), hipMemcpyHostToDevice, st_),
              "copy segs");
    flushed_[c] = seg_used_[c];
    for (const Pending &p : pending_) {
        const Seg *d = seg_dev_[c].as<Seg>() + p.off;
        auto go = [&] { return launch::seg_launch(st_, p.k, d, p.nseg, p.blocks, p.partial); };
        if (timer_) timer_->run(st_, p.name, go);
        else hip_check(go(), p.name);
    }
    pending_.clear();
}

// Node levels of every tree in trees_ (leaf layer, leaf count).  The single
// levels (k_nodes form) run bottom-aligned, one segmented launch per level
// index shared by all trees; then the subtree steps of all trees run
// together, so the latency-bound tops of every tree share two or three
// launches instead of being spread over the tick.
void Lane::tree_levels() {
    const size_t from = subtree_from(trees_.size(), (size_t)1 << tune_.subtree_from_log2);
    plans_.resize(trees_.size());
    singles_.assign(trees_.size(), 0);
    size_t max_single = 0, max_sub = 0;
    for (size_t a = 0; a < trees_.size(); a++) {
        tree_steps(trees_[a].second, plans_[a], from, (int)tune_.subtree_levels);
        size_t k = 0;
        while (k < plans_[a].size() && plans_[a][k].lv == 0) k++;
        singles_[a] = k;
        max_single = std::max(max_single, k);
        max_sub = std::max(max_sub, plans_[a].size() - k);
    }
    std::vector<Seg> &segs = segs_;
    for (size_t t = 0; t < max_single; t++) {
        segs.clear();
        for (size_t a = 0; a < trees_.size(); a++) {
            if (t >= singles_[a]) continue;
            const TreeStep &s = plans_[a][t];
            uint64_t *prev = trees_[a].first + 4 * s.prev;
            segs.push_back(make_seg(prev, prev + 4 * s.h, nullptr, s.h / 2, 0));
        }
        seg(SegKernel::Nodes, segs, "ks_nodes");
    }
    for (size_t v = 0; v < max_sub; v++) {
        segs.clear();
        for (size_t a = 0; a < trees_.size(); a++) {
            if (singles_[a] + v >= plans_[a].size()) continue;
            const TreeStep &s = plans_[a][singles_[a] + v];
            segs.push_back(make_seg(trees_[a].first + 4 * s.prev, nullptr, nullptr, s.h >> s.lv, s.h));
        }
        seg(SegKernel::Subtree, segs, "ks_subtree");
    }
}

// ---------------------------------------------------------------------------
// commit tick: encode, Merkle tree and (when no claim was given) the claim
// ---------------------------------------------------------------------------
void Lane::commit_tick(std::vector<Job *> &jobs) {
    if (jobs.empty()) return;
    reset_ring();
    const int cr = cur_;
    const size_t need = 160 * jobs.size() + 256;
    if (seg_cap_[cr] < need) {
        seg_cap_[cr] = need;
        seg_pin_[cr].ensure(sizeof(Seg) * seg_cap_[cr]);
        seg_dev_[cr] = DevBuf(sizeof(Seg) * seg_cap_[cr], st_);
    }
    int max_n = 0;
    for (size_t a = 0; a < jobs.size(); a++) {
        Job &j = *jobs[a];
        if (!j.cw) job_alloc(j, st_);
        j.slot = a;
        max_n = std::max(max_n, j.n);
    }
    auto &segs = segs_;
    segs.clear();
    for (Job *jp : jobs) {
        Job &j = *jp;
        segs.push_back(make_seg(j.data, j.cw, j.code->ft.as<gf>() + j.code->ft_off[0], (j.N / 2) << j.b, (uint64_t)j.b));
    }
    seg(SegKernel::NttFirst, segs, "ks_ntt_first");
    // layers 1.. in the passes of the shape's plan (up to three layers per
    // pass, k_ntt_multi's grouping): the t-th pass of every job goes into one
    // launch per pass width, whatever layer it starts at
    std::vector<Seg> by_r[4];
    std::vector<std::vector<std::pair<int, int>>> plan(jobs.size());
    size_t max_passes = 0;
    for (size_t a = 0; a < jobs.size(); a++) {
        ntt_passes(jobs[a]->n, tune_.ntt_order, plan[a]);
        max_passes = std::max(max_passes, plan[a].size());
    }
    for (size_t t = 0; t < max_passes; t++) {
        for (auto &v : by_r) v.clear();
        for (size_t a = 0; a < jobs.size(); a++) {
            if (t >= plan[a].size()) continue;
            Job &j = *jobs[a];
            const int s = plan[a][t].first, R = plan[a][t].second;
            const gf *ft = j.code->ft.as<gf>() + j.code->ft_off[s];
            by_r[R].push_back(make_seg(nullptr, j.cw, ft, j.M >> (R == 1 ? 1 : R), (uint64_t)(s + j.b)));
        }
        seg(SegKernel::NttMulti3, by_r[3], "ks_ntt_multi");
        seg(SegKernel::NttMulti2, by_r[2], "ks_ntt_multi");
        seg(SegKernel::NttLayer, by_r[1], "ks_ntt_layer");
    }
    segs.clear();
    for (Job *jp : jobs) segs.push_back(make_seg(jp->cw, jp->tree, nullptr, jp->M, 0));
    seg(SegKernel::LeavesElem, segs, "ks_leaves_elem");
    trees_.clear();
    for (Job *jp : jobs) trees_.push_back({jp->tree, jp->M});
    tree_levels();
    // claim = MLE(evals)(point) by fixing the top variable repeatedly
    // (multilinear.rs:200-214), in the job's poly buffer
    for (int lvl = 0; lvl < max_n; lvl++) {
        segs.clear();
        for (Job *jp : jobs) {
            Job &j = *jp;
            if (j.have_claim || lvl >= j.n) continue;
            const gf x = load_elem(j.point.data() + 16 * (j.n - 1 - lvl));
            segs.push_back(make_seg(lvl == 0 ? j.data : j.poly, j.poly, nullptr, j.N >> (lvl + 1), 0, x));
        }
        seg(SegKernel::Fix, segs, "ks_fix");
    }
    flush();
    // read back roots (and claims)
    const size_t m = jobs.size();
    std::vector<uint64_t> addr(2 * m);
    for (size_t a = 0; a < m; a++) {
        const Job &j = *jobs[a];
        addr[a] = (uint64_t)(j.tree + 4 * (2 * j.M - 2));
        addr[m + a] = (uint64_t)(j.have_claim ? j.data : j.poly);
    }
    ensure(addr_dev_[cr], 16 * m, st_);
    ensure(rb_dev_[cr], 64 * m, st_);
    rb_pin_[cr].ensure(64 * m + 64);
    // the address list lives in pinned memory until the copy has run
    uint64_t *pa = (uint64_t *)((uint8_t *)rb_pin_[cr].p + 48 * m);
    memcpy(pa, addr.data(), 16 * m);
    hip_check(hipMemcpyAsync(addr_dev_[cr].p, pa, 16 * m, hipMemcpyHostToDevice, st_), "copy addr");
    hip_check(launch::gather_abs(st_, addr_dev_[cr].as<uint64_t>(), rb_dev_[cr].p, m, 32), "gather roots");
    hip_check(launch::gather_abs(st_, addr_dev_[cr].as<uint64_t>() + m, rb_dev_[cr].as<uint8_t>() + 32 * m, m, 16),
              "gather claims");
    hip_check(hipMemcpyAsync(rb_pin_[cr].p, rb_dev_[cr].p, 48 * m, hipMemcpyDeviceToHost, st_), "copy roots");
}

void Lane::commit_finish(std::vector<Job *> &jobs) {
    const size_t m = jobs.size();
    const uint8_t *rb = (const uint8_t *)rb_pin_[cur_].p;
    parallel_for(m, [&](size_t a) {
        Job &j = *jobs[a];
        memcpy(j.root, rb + 32 * a, 32);
        if (!j.have_claim) j.claim = load_elem(rb + 32 * m + 16 * a);
        uint8_t cb[16];
        store_elem(cb, j.claim);
        j.o.init(j.n, j.b, j.code->q, j.point.data(), cb);
        const int hn = j.o.half_eq_vars();
        j.eq_lo = (hn + 1) / 2;
        j.eqA = host_eq(j.o.r.data(), j.eq_lo);
        j.eqB = host_eq(j.o.r.data() + j.eq_lo, hn - j.eq_lo);
        if (j.observe_prefix) {                // poly-commit/src/lib.rs:77, 87-89
            j.ch->observe(j.root, 32);
            j.ch->observe_elem(j.claim);
        }
        j.stage = Job::Open;
        j.k = 0;
        j.poly_cur = j.data;
    });
}

// ---------------------------------------------------------------------------
// open tick: every job in the Open stage advances one round (or, after its
// last round, does the final fold and reads its transcript back); jobs in
// the Gather stage gather their openings
// ---------------------------------------------------------------------------
OpenTick Lane::open_tick(std::vector<Job *> &jobs) {
    OpenTick t;
    cur_ = (cur_ + 1) % kSets;
    const int c = cur_;
    t.par = c;
    t.done = ev_[c];
    reset_ring();
    auto &segs = segs_;
    const size_t m = jobs.size();
    const size_t need = 64 * m + 256;
    if (seg_cap_[c] < need) {
        seg_cap_[c] = need;
        seg_pin_[c].ensure(sizeof(Seg) * seg_cap_[c]);
        seg_dev_[c] = DevBuf(sizeof(Seg) * seg_cap_[c], st_);
    }
    for (size_t a = 0; a < m; a++) jobs[a]->slot = a;

    // (A) jobs entering their first round: halved eq tables and the starting
    // image of their device transcript
    {
        size_t elems = 0, img = 0;
        for (Job *jp : jobs)
            if (jp->stage == Job::Open && jp->k == 0) {
                elems += jp->eqA.size() + jp->eqB.size();
                img += align_up(fs_bytes((uint32_t)jp->n, jp->fs_cap));
            }
        if (elems + img) {
            // one upload for all of them: the eq halves, then each job's image
            // (256-byte aligned), which the same ks_outer launch copies into
            // the job's region as 16-byte units (one copy call per job cost
            // the host 10-40 us each)
            const size_t im0 = 16 * elems;
            ensure(eq_dev_[c], im0 + img + 16, st_);
            stage_pin_[c].ensure(im0 + img + 64);
            gf *h = (gf *)stage_pin_[c].p;
            uint8_t *im = (uint8_t *)stage_pin_[c].p + im0;
            size_t e = 0, io = 0;
            segs.clear();
            std::vector<Seg> copies;
            for (Job *jp : jobs) {
                Job &j = *jp;
                if (j.stage != Job::Open || j.k != 0) continue;
                const gf *A = eq_dev_[c].as<gf>() + e;
                memcpy(h + e, j.eqA.data(), 16 * j.eqA.size());
                e += j.eqA.size();
                const gf *B = eq_dev_[c].as<gf>() + e;
                memcpy(h + e, j.eqB.data(), 16 * j.eqB.size());
                e += j.eqB.size();
                segs.push_back(make_seg(A, j.half_eq, B, (size_t)1 << j.o.half_eq_vars(), (uint64_t)j.eq_lo));
                // the region holds align_up(fs_bytes) >= the image rounded up to 16
                const size_t nb = fs_write(j.o, *j.ch, j.fs_cap, im + io), n16 = (nb + 15) / 16;
                memset(im + io + nb, 0, 16 * n16 - nb);
                copies.push_back(make_seg(eq_dev_[c].as<uint8_t>() + im0 + io, j.fs, nullptr, n16, 0));
                io += align_up(fs_bytes((uint32_t)j.n, j.fs_cap));
            }
            segs.insert(segs.end(), copies.begin(), copies.end());
            hip_check(hipMemcpyAsync(eq_dev_[c].p, h, im0 + io, hipMemcpyHostToDevice, st_),
                      "copy eq halves and transcripts");
            seg(SegKernel::Outer, segs, "ks_outer");
        }
    }
    // (B) fold pi and fix poly with the previous round's challenge, which
    // the previous tick's transcript step left in the job's device region
    segs.clear();
    for (Job *jp : jobs) {
        Job &j = *jp;
        if (j.stage != Job::Open || j.k == 0) continue;
        const int kp = j.k - 1, i = j.n - 1 - kp;
        const size_t half = (j.M >> kp) / 2;
        segs.push_back(make_seg(j.lay.pi(kp), (void *)j.lay.pi(kp + 1), j.code->ft.as<gf>() + j.code->ft_off[i], half,
                                (uint64_t)fs_challenge(j.fs)));
    }
    // ... and in the same launch (segments without a twiddle table fix)
    for (Job *jp : jobs) {
        Job &j = *jp;
        if (j.stage != Job::Open || j.k == 0) continue;
        const int i = j.n - j.k;
        segs.push_back(make_seg(j.poly_cur, j.poly, nullptr, (size_t)1 << i, (uint64_t)fs_challenge(j.fs)));
        j.poly_cur = j.poly;
    }
    seg(SegKernel::FoldFix, segs, "ks_fold_fix");
    // (C) this round's pi tree, inner product and transcript step
    segs.clear();
    for (Job *jp : jobs) {
        Job &j = *jp;
        if (j.stage != Job::Open || j.k >= j.n) continue;
        const size_t half = (j.M >> j.k) / 2;
        segs.push_back(make_seg(j.lay.pi(j.k), (void *)j.lay.pitree(j.k), nullptr, half, half));
    }
    // the inner products ride in the same launch, at most 64 blocks (partial
    // sums) per job, job a's sums at partials_[64 a, 64 a + 64)
    const size_t nleaf = segs.size();
    std::vector<Job *> dotted;
    for (Job *jp : jobs) {
        Job &j = *jp;
        if (j.stage != Job::Open || j.k >= j.n) continue;
        const int i = j.n - 1 - j.k;
        Seg g = make_seg(j.half_eq, nullptr, j.poly_cur, (size_t)1 << i, 0, gf{64 * (uint64_t)dotted.size(), 0});
        g.nblk = 64;  // cap; seg() sets the count
        segs.push_back(g);
        dotted.push_back(jp);
    }
    ensure(partials_, 16 * 64 * m + 16, st_);
    seg(SegKernel::LeavesPairDot, segs, "ks_leaves_pair_dot", partials_.as<gf>());
    std::vector<Seg> dsegs(segs.begin() + nleaf, segs.end());
    trees_.clear();
    for (Job *jp : jobs) {
        Job &j = *jp;
        if (j.stage != Job::Open || j.k >= j.n) continue;
        trees_.push_back({(uint64_t *)j.lay.pitree(j.k), (j.M >> j.k) / 2});
    }
    tree_levels();
    std::vector<Seg> rsegs;
    for (size_t a = 0; a < dotted.size(); a++) {
        const Job &j = *dotted[a];
        const size_t half = (j.M >> j.k) / 2;
        const uint64_t *root = j.lay.pitree(j.k) + 4 * (2 * half - 2);
        rsegs.push_back(fs_round_seg(root, j.fs, j.k, (uint32_t)(64 * a), dsegs[a].nblk));
    }
    seg(SegKernel::Round, rsegs, "ks_round", partials_.as<gf>());
    flush();
    // (D) jobs past their last round: read back the transcript region and
    // the final polynomial, as 16-byte records gathered into one buffer
    size_t nrec = 0;
    for (Job *jp : jobs)
        if (jp->stage == Job::Open && jp->k == jp->n) nrec += fs_readback_bytes((uint32_t)jp->n) / 16 + 1;
    if (nrec) {
        rb_pin_[c].ensure(16 * nrec + 8 * nrec + 64);
        uint64_t *addr = (uint64_t *)((uint8_t *)rb_pin_[c].p + 16 * nrec);
        size_t r = 0;
        for (Job *jp : jobs) {
            Job &j = *jp;
            if (j.stage != Job::Open || j.k != j.n) continue;
            t.finals.push_back({jp, 16 * r});
            const size_t nb = fs_readback_bytes((uint32_t)j.n) / 16;
            for (size_t q = 0; q < nb; q++) addr[r++] = (uint64_t)(j.fs + 16 * q);
            addr[r++] = (uint64_t)j.poly_cur;
        }
        ensure(addr_dev_[c], 8 * nrec, st_);
        ensure(rb_dev_[c], 16 * nrec, st_);
        hip_check(hipMemcpyAsync(addr_dev_[c].p, addr, 8 * nrec, hipMemcpyHostToDevice, st_), "copy addr");
        hip_check(launch::gather_abs(st_, addr_dev_[c].as<uint64_t>(), rb_dev_[c].p, nrec, 16), "gather finals");
        hip_check(hipMemcpyAsync(rb_pin_[c].p, rb_dev_[c].p, 16 * nrec, hipMemcpyDeviceToHost, st_), "copy finals");
    }
    for (Job *jp : jobs) {
        Job &j = *jp;
        if (j.stage != Job::Open) continue;
        if (j.k == j.n) j.stage = Job::Final;
        j.k++;
    }
    // (E) openings of jobs whose queries are known: their query records,
    // written by the device in bincode layout (k_proof_records) into pinned
    // host memory, for open_proofs to copy behind each proof's header
    size_t body = 0, nidx = 0;
    for (Job *jp : jobs) {
        Job &j = *jp;
        if (j.stage != Job::Gather) continue;
        t.gathered.push_back(jp);
        t.gather_at.push_back(body);
        body += j.o.q * query_record_bytes(j.o.b, j.o.d, 1);
        nidx += j.o.q;
    }
    if (body) {
        const size_t G = t.gathered.size(), up = sizeof(ProofJob) * G + 8 * nidx;
        plan_pin_[c].ensure(up);
        ensure(plan_dev_[c], up, st_);
        // the records go from the kernel straight into pinned host memory
        // (one pass over PCIe, no device buffer and no copy behind it)
        gather_pin_[c].ensure(body);
        uint8_t *hout = nullptr;
        hip_check(hipHostGetDevicePointer((void **)&hout, gather_pin_[c].p, 0), "hipHostGetDevicePointer");
        ProofJob *hj = (ProofJob *)plan_pin_[c].p;
        uint64_t *hidx = reinterpret_cast<uint64_t *>(hj + G);
        const uint64_t *didx = reinterpret_cast<const uint64_t *>(plan_dev_[c].as<uint8_t>() + sizeof(ProofJob) * G);
        size_t io = 0;
        uint32_t rec = 0;
        for (size_t g = 0; g < G; g++) {
            const Job &j = *t.gathered[g];
            memcpy(hidx + io, j.o.idx.data(), 8 * j.o.q);
            hj[g] = proof_job(j.o, j.lay, didx + io, hout + t.gather_at[g], rec);
            io += j.o.q;
            rec += (uint32_t)j.o.q;
        }
        hip_check(hipMemcpyAsync(plan_dev_[c].p, hj, up, hipMemcpyHostToDevice, st_), "copy query indices");
        auto go = [&] { return launch::proof_records(st_, plan_dev_[c].as<ProofJob>(), (uint32_t)G, rec); };
        if (timer_) timer_->run(st_, "k_proof_records", go);
        else hip_check(go(), "proof records");
    }
    for (Job *jp : t.gathered) jp->stage = Job::Gathering;
    hip_check(hipEventRecord(t.done, st_), "hipEventRecord");
    return t;
}

void Lane::open_finish(OpenTick &t) {
    open_finals(t);
    open_proofs(t);
}

void Lane::open_proofs(OpenTick &t) {
    const int c = t.par;
    // proofs of the jobs this tick gathered: header from the host's round
    // data, query records from the kernel's writes to pinned memory
    const uint8_t *recs = (const uint8_t *)gather_pin_[c].p;
    ChunkCopy cc;
    for (size_t g = 0; g < t.gathered.size(); g++) {
        Job &j = *t.gathered[g];
        const size_t sz = proof_size(j.o.b, j.o.d, j.o.q);
        if (j.proof && j.proof_cap < sz) fail(BFD_ERR_ARG, "proof buffer too small");
        if (!j.proof) {
            j.proof = (uint8_t *)malloc(sz ? sz : 1);
            if (!j.proof) fail(BFD_ERR_ARG, "host allocation failed");
            j.proof_owned = true;
        }
        j.proof_len = sz;
        write_proof_header(j.o, j.proof);
        const size_t hb = proof_header_bytes(j.o.d);
        cc.add(j.proof + hb, recs + t.gather_at[g], sz - hb);
    }
    cc.run();
    const double now = now_ms();
    for (Job *jp : t.gathered) {
        jp->stage = Job::Done;
        jp->t_done = now;
    }
}

void Lane::open_finals(OpenTick &t) {
    const int c = t.par;
    // jobs past their last round: the round data and the challenger state
    // come from the read-back; the queries are sampled here
    // (basefold.rs:194-203); every job owns a distinct challenger
    const uint8_t *rb = (const uint8_t *)rb_pin_[c].p;
    parallel_for(t.finals.size(), [&](size_t a) {
        Job &j = *t.finals[a].first;
        const uint8_t *img = rb + t.finals[a].second;
        fs_read(img, j.o, *j.ch);
        j.o.finish(load_elem(img + fs_readback_bytes((uint32_t)j.n)), *j.ch);
        j.stage = Job::Gather;
    });
}

}  // namespace rt
}  // namespace bfd
