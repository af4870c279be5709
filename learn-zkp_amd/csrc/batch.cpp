// Batched commit+open of many independent polynomials of uneven sizes.
//
// The reference commits one matrix height per Basefold instance
// (poly-commit/src/basefold.rs:94), so a batch is a set of independent
// Basefold proofs, each with its own transcript; every proof equals the one
// the plain path produces for the same polynomial.
//
// Every stage runs as one segmented launch covering all polynomials that
// still have work at that stage (NTT layer s, Merkle level t, open round k),
// so a batch of P polynomials costs about max(n) + max(L) launches for the
// commitments and one host round trip per open round for the whole batch,
// instead of P times that.  The per-polynomial transcripts advance in
// lockstep on the host between rounds.
#include "internal.hpp"

using namespace bfd::rt;

namespace {

// Stages segment tables in pinned memory and uploads them on the stream.
// The pinned ring is reused after every stream synchronisation.
class SegLauncher {
  public:
    SegLauncher(bfd_ctx *ctx, size_t cap_segs) : ctx_(ctx), cap_(cap_segs) {
        ctx->seg_pin.ensure(sizeof(Seg) * cap_);
        host_ = (Seg *)ctx->seg_pin.p;
        dev_ = DevBuf(sizeof(Seg) * cap_, ctx->stream);
    }
    // After a stream synchronisation the staging ring can be reused.
    void synced() { used_ = 0; }

    // Assigns blocks to segments and launches; returns the total block count.
    uint32_t launch(SegKernel k, std::vector<Seg> &segs, const char *name, gf *partial = nullptr,
                    uint32_t max_blocks_per_seg = 2048) {
        if (segs.empty()) return 0;
        if (used_ + segs.size() > cap_) {
            hip_check(hipStreamSynchronize(ctx_->stream), "segment ring sync");
            used_ = 0;
        }
        const int bs = launch::seg_block_size(k);
        uint32_t blocks = 0;
        for (auto &g : segs) {
            uint64_t want = (g.count + bs - 1) / bs;
            if (want < 1) want = 1;
            if (want > max_blocks_per_seg) want = max_blocks_per_seg;
            g.blk0 = blocks;
            g.nblk = (uint32_t)want;
            blocks += g.nblk;
        }
        Seg *h = host_ + used_;
        memcpy(h, segs.data(), sizeof(Seg) * segs.size());
        Seg *d = dev_.as<Seg>() + used_;
        hip_check(hipMemcpyAsync(d, h, sizeof(Seg) * segs.size(), hipMemcpyHostToDevice, ctx_->stream), "copy segs");
        used_ += segs.size();
        hipStream_t st = ctx_->stream;
        ctx_->timer.run(st, name, [&] { return launch::seg_launch(st, k, d, (int)segs.size(), blocks, partial); });
        return blocks;
    }

  private:
    bfd_ctx *ctx_;
    size_t cap_;
    Seg *host_ = nullptr;
    DevBuf dev_;
    size_t used_ = 0;
};

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

struct Poly {
    int n = 0, L = 0;
    size_t N = 0, M = 0;
    const bfd_code *code = nullptr;
    const gf *data = nullptr;
    gf *cw = nullptr, *half_eq = nullptr, *poly = nullptr, *pis = nullptr;
    uint64_t *tree = nullptr, *pitrees = nullptr;
    OpenLayout lay;
    OpenRounds o;
    const gf *poly_cur = nullptr;
    Challenger *ch = nullptr;
    uint32_t dot_blk0 = 0, dot_nblk = 0;
    std::vector<gf> eqA, eqB;
    int eq_lo = 0;
};

size_t align_up(size_t v) { return (v + 255) & ~(size_t)255; }

}  // namespace

namespace bfd {
namespace rt {

void batch_commit_open(bfd_ctx *ctx, uint32_t lambda, uint32_t log2_c, size_t count, const uint32_t *log_sizes,
                       const void *const *evals, bool on_device, const uint8_t *const *points,
                       const uint8_t *const *claims, Challenger *const *chs, bool observe_prefix, uint8_t *roots,
                       uint8_t **proofs, size_t *proof_lens) {
    std::lock_guard<std::recursive_mutex> lk(ctx->mu);
    set_device(ctx);
    hipStream_t st = ctx->stream;
    const int b = (int)log2_c;
    const size_t c = (size_t)1 << b;
    std::vector<Poly> P(count);
    int max_n = 0, max_L = 0;
    size_t bytes = 0;
    for (size_t p = 0; p < count; p++) {
        Poly &q = P[p];
        q.n = (int)log_sizes[p];
        if (q.n < 1 || q.n > 30) fail(BFD_ERR_ARG, "batched polynomials need 1..30 variables");
        q.L = q.n + b;
        q.N = (size_t)1 << q.n;
        q.M = q.N << b;
        q.code = cached_code(ctx, lambda, log2_c, (uint32_t)q.n);
        q.ch = chs[p];
        q.lay.plan(q.M, q.n);
        max_n = std::max(max_n, q.n);
        max_L = std::max(max_L, q.L);
        bytes += align_up(on_device ? 0 : 16 * q.N) + align_up(16 * q.M) + align_up(32 * (2 * q.M - 1)) +
                 align_up(16 * ((size_t)1 << (q.n - 1))) + align_up(16 * (q.N / 2)) +
                 align_up(16 * q.lay.pis_elems(q.M, c)) + align_up(32 * q.lay.pitree_digests());
    }
    if (ctx->budget && ctx->resident + bytes > ctx->budget)
        fail(BFD_ERR_BUDGET, "batch exceeds the context's HBM budget");
    DevBuf slab(bytes, st);
    {
        uint8_t *cur = slab.as<uint8_t>();
        auto take = [&](size_t n) {
            uint8_t *r = cur;
            cur += align_up(n);
            return r;
        };
        for (size_t p = 0; p < count; p++) {
            Poly &q = P[p];
            if (on_device) {
                q.data = (const gf *)evals[p];
            } else {
                gf *d = (gf *)take(16 * q.N);
                hip_check(hipMemcpyAsync(d, evals[p], 16 * q.N, hipMemcpyHostToDevice, st), "copy evaluations");
                q.data = d;
            }
            q.cw = (gf *)take(16 * q.M);
            q.tree = (uint64_t *)take(32 * (2 * q.M - 1));
            q.half_eq = (gf *)take(16 * ((size_t)1 << (q.n - 1)));
            q.poly = (gf *)take(16 * (q.N / 2));
            q.pis = (gf *)take(16 * q.lay.pis_elems(q.M, c));
            q.pitrees = (uint64_t *)take(32 * q.lay.pitree_digests());
            q.lay.cw = q.cw;
            q.lay.tree = q.tree;
            q.lay.pis = q.pis;
            q.lay.pitrees = q.pitrees;
        }
    }
    SegLauncher segl(ctx, 64 * count + 64);
    std::vector<Seg> segs;

    // ---- commit: encode (NTT layers) and Merkle trees --------------------
    for (int s = 0; s < max_n; s++) {
        segs.clear();
        for (auto &q : P) {
            if (q.n <= s) continue;
            const gf *ft = q.code->ft.as<gf>() + q.code->ft_off[s];
            const size_t work = (q.N / 2) << b;
            if (s == 0) segs.push_back(make_seg(q.data, q.cw, ft, work, (uint64_t)b));
            else segs.push_back(make_seg(nullptr, q.cw, ft, work, (uint64_t)(s + b)));
        }
        segl.launch(s == 0 ? SegKernel::NttFirst : SegKernel::NttLayer, segs, s == 0 ? "ks_ntt_first" : "ks_ntt_layer");
    }
    segs.clear();
    for (auto &q : P) segs.push_back(make_seg(q.cw, q.tree, nullptr, q.M, 0));
    segl.launch(SegKernel::LeavesElem, segs, "ks_leaves_elem");
    for (int t = 0; t < max_L; t++) {
        segs.clear();
        for (auto &q : P)
            if (t < q.L)
                segs.push_back(make_seg(q.tree + 4 * layer_off(q.M, t), q.tree + 4 * layer_off(q.M, t + 1), nullptr,
                                        q.M >> (t + 1), 0));
        segl.launch(SegKernel::Nodes, segs, "ks_nodes");
    }
    // roots, gathered by absolute address
    const size_t pin_need = 48 * count + 16 * 64 * count + 64;
    ctx->pinned.ensure(pin_need);
    uint8_t *pin = (uint8_t *)ctx->pinned.p;
    std::vector<uint64_t> addr(count);
    DevBuf daddr(8 * 2 * count + 8, st), dgat(48 * count + 16, st);
    for (size_t p = 0; p < count; p++) addr[p] = (uint64_t)(P[p].tree + 4 * (2 * P[p].M - 2));
    hip_check(hipMemcpyAsync(daddr.p, addr.data(), 8 * count, hipMemcpyHostToDevice, st), "copy addr");
    hip_check(launch::gather_abs(st, daddr.as<uint64_t>(), dgat.p, count, 32), "gather roots");
    hip_check(hipMemcpyAsync(pin, dgat.p, 32 * count, hipMemcpyDeviceToHost, st), "copy roots");
    // host work that does not depend on the roots overlaps the commit kernels
    for (size_t p = 0; p < count; p++) {
        Poly &q = P[p];
        q.o.init(q.n, b, q.code->q, points[p], claims[p]);
        const int hn = q.o.half_eq_vars();
        q.eq_lo = (hn + 1) / 2;
        q.eqA = host_eq(q.o.r.data(), q.eq_lo);
        q.eqB = host_eq(q.o.r.data() + q.eq_lo, hn - q.eq_lo);
    }
    hip_check(hipStreamSynchronize(st), "commit sync");
    segl.synced();
    ctx->timer.collect();
    for (size_t p = 0; p < count; p++) {
        memcpy(roots + 32 * p, pin + 32 * p, 32);
        if (observe_prefix) {                        // poly-commit/src/lib.rs:77, 87-89
            P[p].ch->observe(pin + 32 * p, 32);
            P[p].ch->observe_elem(P[p].o.claim);
        }
    }

    // ---- halved eq tables (tensor product of host-built halves) ----------
    size_t eq_elems = 0;
    for (auto &q : P) eq_elems += q.eqA.size() + q.eqB.size();
    std::vector<gf> eqh;
    eqh.reserve(eq_elems);
    for (auto &q : P) {
        eqh.insert(eqh.end(), q.eqA.begin(), q.eqA.end());
        eqh.insert(eqh.end(), q.eqB.begin(), q.eqB.end());
    }
    DevBuf deq(16 * eq_elems, st);
    hip_check(hipMemcpyAsync(deq.p, eqh.data(), 16 * eq_elems, hipMemcpyHostToDevice, st), "copy eq halves");
    segs.clear();
    {
        const gf *cur = deq.as<gf>();
        for (auto &q : P) {
            const gf *A = cur, *B = cur + q.eqA.size();
            cur += q.eqA.size() + q.eqB.size();
            segs.push_back(make_seg(A, q.half_eq, B, (size_t)1 << q.o.half_eq_vars(), (uint64_t)q.eq_lo));
        }
    }
    segl.launch(SegKernel::Outer, segs, "ks_outer");
    for (auto &q : P) q.poly_cur = q.data;

    // ---- open rounds, all polynomials in lockstep -------------------------
    DevBuf partials(16 * 64 * count + 16, st);
    for (int k = 0; k < max_n; k++) {
        // pi trees
        segs.clear();
        for (auto &q : P)
            if (k < q.n) {
                const size_t half = (q.M >> k) / 2;
                segs.push_back(make_seg(q.lay.pi(k), (void *)q.lay.pitree(k), nullptr, half, half));
            }
        segl.launch(SegKernel::LeavesPair, segs, "ks_leaves_pair");
        for (int t = 0;; t++) {
            segs.clear();
            for (auto &q : P)
                if (k < q.n) {
                    const size_t half = (q.M >> k) / 2;
                    if ((half >> t) > 1) {
                        uint64_t *tk = (uint64_t *)q.lay.pitree(k);
                        segs.push_back(make_seg(tk + 4 * layer_off(half, t), tk + 4 * layer_off(half, t + 1), nullptr,
                                                half >> (t + 1), 0));
                    }
                }
            if (segs.empty()) break;
            segl.launch(SegKernel::Nodes, segs, "ks_nodes");
        }
        // inner products
        segs.clear();
        std::vector<size_t> active;
        for (size_t p = 0; p < count; p++) {
            Poly &q = P[p];
            if (k < q.n) {
                const int i = q.n - 1 - k;
                segs.push_back(make_seg(q.half_eq, nullptr, q.poly_cur, (size_t)1 << i, 0));
                active.push_back(p);
            }
        }
        const uint32_t nblk = segl.launch(SegKernel::Dot, segs, "ks_dot", partials.as<gf>(), 64);
        for (size_t a = 0; a < active.size(); a++) {
            P[active[a]].dot_blk0 = segs[a].blk0;
            P[active[a]].dot_nblk = segs[a].nblk;
        }
        // roots of this round's pi trees
        for (size_t a = 0; a < active.size(); a++) {
            Poly &q = P[active[a]];
            const size_t half = (q.M >> k) / 2;
            addr[a] = (uint64_t)(q.lay.pitree(k) + 4 * (2 * half - 2));
        }
        hip_check(hipMemcpyAsync(daddr.p, addr.data(), 8 * active.size(), hipMemcpyHostToDevice, st), "copy addr");
        hip_check(launch::gather_abs(st, daddr.as<uint64_t>(), dgat.p, active.size(), 32), "gather roots");
        hip_check(hipMemcpyAsync(pin, dgat.p, 32 * active.size(), hipMemcpyDeviceToHost, st), "copy roots");
        hip_check(hipMemcpyAsync(pin + 32 * count, partials.p, 16 * (size_t)nblk, hipMemcpyDeviceToHost, st),
                  "copy partials");
        hip_check(hipStreamSynchronize(st), "round sync");
        segl.synced();
        std::vector<gf> xs(active.size());
        for (size_t a = 0; a < active.size(); a++) {
            Poly &q = P[active[a]];
            gf S = gf_zero();
            for (uint32_t t = 0; t < q.dot_nblk; t++) S = gf_add(S, load_elem(pin + 32 * count + 16 * (q.dot_blk0 + t)));
            xs[a] = q.o.round(k, pin + 32 * a, S, *q.ch);
        }
        // folds and fixes
        segs.clear();
        for (size_t a = 0; a < active.size(); a++) {
            Poly &q = P[active[a]];
            const int i = q.n - 1 - k;
            const size_t half = (q.M >> k) / 2;
            segs.push_back(make_seg(q.lay.pi(k), (void *)q.lay.pi(k + 1), q.code->ft.as<gf>() + q.code->ft_off[i],
                                    half, 0, xs[a]));
        }
        segl.launch(SegKernel::Fold, segs, "ks_fold");
        segs.clear();
        for (size_t a = 0; a < active.size(); a++) {
            Poly &q = P[active[a]];
            const int i = q.n - 1 - k;
            segs.push_back(make_seg(q.poly_cur, q.poly, nullptr, (size_t)1 << i, 0, xs[a]));
            q.poly_cur = q.poly;
        }
        segl.launch(SegKernel::Fix, segs, "ks_fix");
    }

    // ---- final polynomials, queries and openings ---------------------------
    for (size_t p = 0; p < count; p++) addr[p] = (uint64_t)P[p].poly_cur;
    hip_check(hipMemcpyAsync(daddr.p, addr.data(), 8 * count, hipMemcpyHostToDevice, st), "copy addr");
    hip_check(launch::gather_abs(st, daddr.as<uint64_t>(), dgat.p, count, 16), "gather finals");
    hip_check(hipMemcpyAsync(pin, dgat.p, 16 * count, hipMemcpyDeviceToHost, st), "copy finals");
    hip_check(hipStreamSynchronize(st), "final sync");
    segl.synced();
    double t0 = now_ms();
    std::vector<size_t> e0(count + 1, 0), d0(count + 1, 0);
    // query sampling is independent per transcript
    parallel_for(count, [&](size_t p) { P[p].o.finish(load_elem(pin + 16 * p), *P[p].ch); });
    for (size_t p = 0; p < count; p++) {
        size_t ne = 0, nd = 0;
        gather_counts(P[p].o, ne, nd);
        e0[p + 1] = e0[p] + ne;
        d0[p + 1] = d0[p] + nd;
    }
    const size_t NE = e0[count], ND = d0[count];
    std::vector<uint64_t> plan(NE + ND);
    parallel_for(count, [&](size_t p) { plan_gathers(P[p].o, P[p].lay, plan.data() + e0[p], plan.data() + NE + d0[p]); });
    ctx->timer.host("plan", now_ms() - t0);
    PinnedBuf &hout = ctx->gather_pin;
    hout.ensure(16 * NE + 32 * ND + 64);
    DevBuf dix(8 * (NE + ND) + 8, st), dout(16 * NE + 32 * ND + 64, st);
    hip_check(hipMemcpyAsync(dix.p, plan.data(), 8 * (NE + ND), hipMemcpyHostToDevice, st), "copy gather plan");
    uint8_t *oe = dout.as<uint8_t>(), *od = oe + 16 * NE;
    ctx->timer.run(st, "k_gather_abs", [&] { return launch::gather_abs(st, dix.as<uint64_t>(), oe, NE, 16); });
    ctx->timer.run(st, "k_gather_abs", [&] { return launch::gather_abs(st, dix.as<uint64_t>() + NE, od, ND, 32); });
    hip_check(hipMemcpyAsync(hout.p, oe, 16 * NE + 32 * ND, hipMemcpyDeviceToHost, st), "copy gather");
    hip_check(hipStreamSynchronize(st), "gather sync");
    ctx->timer.collect();
    t0 = now_ms();
    const uint8_t *he = (const uint8_t *)hout.p, *hd = he + 16 * NE;
    for (size_t p = 0; p < count; p++) {
        const size_t sz = proof_size(P[p].o.b, P[p].o.d, P[p].o.q);
        if (proofs[p] && proof_lens[p] < sz) fail(BFD_ERR_ARG, "proof buffer too small");
        if (!proofs[p]) {
            proofs[p] = (uint8_t *)malloc(sz ? sz : 1);
            if (!proofs[p]) fail(BFD_ERR_ARG, "host allocation failed");
        }
        proof_lens[p] = sz;
    }
    parallel_for(count, [&](size_t p) { assemble_proof(P[p].o, he + 16 * e0[p], hd + 32 * d0[p], proofs[p]); });
    ctx->timer.host("assemble", now_ms() - t0);
}

}  // namespace rt
}  // namespace bfd

extern "C" int bfd_commit_open_batch(bfd_ctx *ctx, uint32_t lambda, uint32_t log2_c, size_t count,
                                     const uint32_t *log_sizes, const void *const *evals, int evals_on_device,
                                     const uint8_t *const *points, const uint8_t *const *claims,
                                     bfd_challenger *const *challengers, int observe_prefix, uint8_t *roots,
                                     uint8_t **proofs, size_t *proof_lens) {
    return guarded([&] {
        if (!ctx || !log_sizes || !evals || !points || !claims || !challengers || !roots || !proofs || !proof_lens)
            fail(BFD_ERR_ARG, "null argument");
        if (count == 0) return;
        std::vector<Challenger *> chs(count);
        for (size_t p = 0; p < count; p++) {
            if (!challengers[p] || !evals[p] || !points[p] || !claims[p]) fail(BFD_ERR_ARG, "null argument");
            chs[p] = &challengers[p]->c;
        }
        batch_commit_open(ctx, lambda, log2_c, count, log_sizes, evals, evals_on_device != 0, points, claims,
                          chs.data(), observe_prefix != 0, roots, proofs, proof_lens);
    });
}
