// Basefold::verify (poly-commit/src/basefold.rs:234-325) over many proofs of
// one code, with the bincode proof decoder.
//
// The host decodes every proof (layout and length prefixes of
// BasefoldProof, basefold.rs:49-66) and replays its transcript: the random
// scalars of a width > 1 commitment, the sumcheck rounds with the pi-tree
// roots (verify_interleaved_folding_sumcheck, :371-396), the final-poly
// identity (:271-276) and the query indices (:280-283).  The device then
// checks every query of every proof in two launches: the fold chain of the
// opened row through all rounds (thread per proof x query), and every
// Merkle path (thread per proof x query x path: the main tree and one pi
// tree per round), so a proof's latency is one path, not its 375 hashes in
// sequence.  The results map back to the reference's first error in query
// order: a main-tree or pi-tree path (Mmcs), then the folded value
// (InvalidQuery).
#include <algorithm>

#include "internal.hpp"

namespace {
using namespace bfd::rt;

struct Parsed {
    std::vector<std::array<gf, 2>> comp;
    gf final_poly{};
    size_t roots_off = 0, open_off = 0, rec_bytes = 0;
};

uint64_t rd64(const uint8_t *p) {
    uint64_t v;
    memcpy(&v, p, 8);
    return v;
}

// Decodes the bincode BasefoldProof for (b, d, q, w); false if its length
// or any length prefix differs from the layout of such a proof.
bool decode(const uint8_t *p, size_t len, int b, int d, size_t q, size_t w, Parsed &o) {
    if (len != proof_size(b, d, q, w)) return false;
    size_t pos = 0;
    if (rd64(p) != (uint64_t)d) return false;
    pos += 8;
    o.comp.assign(d, {});
    for (int k = 0; k < d; k++) {
        if (rd64(p + pos) != 2) return false;
        o.comp[k][0] = load_elem(p + pos + 8);
        o.comp[k][1] = load_elem(p + pos + 24);
        pos += 40;
    }
    if (rd64(p + pos) != 1) return false;
    o.final_poly = load_elem(p + pos + 8);
    pos += 24;
    if (rd64(p + pos) != (uint64_t)d) return false;
    o.roots_off = pos + 8;
    pos = o.roots_off + 32 * (size_t)d;
    if (rd64(p + pos) != q) return false;
    o.open_off = pos + 8;
    o.rec_bytes = query_record_bytes(b, d, w);
    const size_t L = (size_t)(b + d);
    for (size_t t = 0; t < q; t++) {
        const uint8_t *r = p + o.open_off + t * o.rec_bytes;
        if (rd64(r) != w || rd64(r + 8 + 16 * w) != L || rd64(r + 16 + 16 * w + 32 * L) != (uint64_t)d) return false;
        size_t off = 24 + 16 * w + 32 * L;
        for (int k = 0; k < d; k++) {
            if (rd64(r + off + 16) != (uint64_t)(b + d - 1 - k)) return false;
            off += 24 + 32 * (size_t)(b + d - 1 - k);
        }
    }
    return true;
}

}  // namespace

extern "C" {

int bfd_verify_batch(bfd_ctx *ctx, const bfd_code *code, size_t width, size_t count, const uint8_t *roots,
                     const uint8_t *const *points, const uint8_t *const *values, const uint8_t *const *proofs,
                     const size_t *proof_lens, bfd_challenger *const *challengers, int *results) {
    return guarded([&] {
        if (!ctx || !code || !results || (count && (!roots || !values || !proofs || !proof_lens || !challengers)))
            fail(BFD_ERR_ARG, "null argument");
        if (width < 1 || width > kMaxWidth) fail(BFD_ERR_ARG, "width must be 1..65535");
        if (code->d && count && !points) fail(BFD_ERR_ARG, "null points");
        std::lock_guard<std::recursive_mutex> lk(ctx->mu);
        set_device(ctx);
        hipStream_t st = ctx->stream;
        const int b = (int)code->log2_c, d = (int)code->d, L = b + d;
        const size_t q = code->q, w = width;
        std::vector<Parsed> parsed(count);
        if (code->ctx->device != ctx->device) fail(BFD_ERR_ARG, "the code's twiddles live on another device");
        for (size_t p = 0; p < count; p++)
            if (!proofs[p] || !challengers[p] || !values[p] || (d && !points[p])) fail(BFD_ERR_ARG, "null argument");
        // the transcripts are replayed in parallel, one challenger each
        {
            std::vector<const bfd_challenger *> seen(challengers, challengers + count);
            std::sort(seen.begin(), seen.end());
            if (std::adjacent_find(seen.begin(), seen.end()) != seen.end())
                fail(BFD_ERR_ARG, "every proof needs its own challenger");
        }
        // decode and replay every transcript, one proof per pool task
        std::vector<std::vector<uint64_t>> idx_p(count);
        std::vector<std::vector<gf>> chal_p(count), scal_p(count);
        parallel_for(count, [&](size_t p) {
            results[p] = 0;
            Parsed &P = parsed[p];
            if (!decode(proofs[p], proof_lens[p], b, d, q, w, P)) {
                results[p] = 4;
                return;
            }
            Challenger &ch = challengers[p]->c;
            // scalars and claim (basefold.rs:252-258)
            std::vector<gf> &sc = scal_p[p];
            sc.assign(w, gf_zero());
            gf claim = gf_zero();
            if (w == 1) {
                sc[0] = gf_one();
                claim = load_elem(values[p]);
            } else {
                int lw = 0;
                while (((size_t)1 << lw) < w) lw++;
                std::vector<gf> rs(lw);
                for (int i = 0; i < lw; i++) rs[i] = ch.sample();
                std::vector<gf> e = host_eq(rs.data(), lw);
                for (size_t c = 0; c < w; c++) {
                    sc[c] = e[c];
                    claim = gf_add(claim, hmul(e[c], load_elem(values[p] + 16 * c)));
                }
            }
            // sumcheck rounds interleaved with the pi-tree roots (:371-396),
            // Eval::decompress_round_poly: c1 = subclaim - 2 c0 - c2
            std::vector<gf> &ch_i = chal_p[p];
            ch_i.assign(d, gf_zero());
            gf sub = claim;
            for (int k = 0; k < d; k++) {
                ch.observe(proofs[p] + P.roots_off + 32 * (size_t)k, 32);
                ch.observe_elem(P.comp[k][0]);
                ch.observe_elem(P.comp[k][1]);
                const gf c0 = P.comp[k][0], c2 = P.comp[k][1], c1 = gf_add(sub, c2);
                const gf x = ch.sample();
                sub = gf_add(c0, hmul(x, gf_add(c1, hmul(x, c2))));
                ch_i[d - 1 - k] = x;
            }
            ch.observe_elem(P.final_poly);
            // evaluate(final_poly, point_lo = []) * eq_eval(point, r) (:273-276)
            gf eq = gf_one();
            for (int i = 0; i < d; i++)
                eq = hmul(eq, gf_add(gf_one(), gf_add(load_elem(points[p] + 16 * i), ch_i[i])));
            if (!gf_eq(hmul(P.final_poly, eq), sub)) {
                results[p] = 1;
                return;
            }
            idx_p[p].resize(q);
            for (size_t t = 0; t < q; t++) idx_p[p][t] = ch.sample_bits(L);
        });
        std::vector<size_t> live;  // proofs left for the device
        std::vector<uint64_t> idx;
        std::vector<gf> chal, scal;
        for (size_t p = 0; p < count; p++) {
            if (results[p]) continue;
            live.push_back(p);
            idx.insert(idx.end(), idx_p[p].begin(), idx_p[p].end());
            chal.insert(chal.end(), chal_p[p].begin(), chal_p[p].end());
            scal.insert(scal.end(), scal_p[p].begin(), scal_p[p].end());
        }
        const size_t nl = live.size();
        if (!nl || !q) return;
        // device copies of proofs, roots, indices, challenges and scalars; every
        // section starts 16-byte aligned (element loads are 16 bytes wide)
        auto up16 = [](size_t v) { return (v + 15) & ~(size_t)15; };
        std::vector<size_t> poff(nl);
        size_t pbytes = 0;
        for (size_t a = 0; a < nl; a++) {
            poff[a] = pbytes;
            pbytes += up16(proof_lens[live[a]]);
        }
        const size_t sz_idx = up16(8 * idx.size()), sz_ch = 16 * chal.size(), sz_sc = 16 * scal.size();
        const size_t sz_vals = 32 * nl * q * (size_t)(d > 0 ? d : 1);
        DevBuf dbuf = ctx_buf(ctx, pbytes + 32 * nl + sz_idx + sz_ch + sz_sc + sizeof(VerifyJob) * nl + 64);
        DevBuf dvals = ctx_buf(ctx, sz_vals), dok = ctx_buf(ctx, nl * q * (size_t)(d + 2));
        uint8_t *base = dbuf.as<uint8_t>();
        uint8_t *d_proofs = base, *d_roots = base + pbytes;
        uint64_t *d_idx = reinterpret_cast<uint64_t *>(d_roots + 32 * nl);
        gf *d_ch = reinterpret_cast<gf *>(reinterpret_cast<uint8_t *>(d_idx) + sz_idx);
        gf *d_sc = reinterpret_cast<gf *>(reinterpret_cast<uint8_t *>(d_ch) + sz_ch);
        VerifyJob *d_jobs = reinterpret_cast<VerifyJob *>(reinterpret_cast<uint8_t *>(d_sc) + sz_sc);
        // the proofs are staged in pinned memory by the pool and go up in one
        // copy; the small arrays through another
        ctx->verify_pin.ensure(pbytes);
        uint8_t *stage = (uint8_t *)ctx->verify_pin.p;
        parallel_for(nl, [&](size_t a) { memcpy(stage + poff[a], proofs[live[a]], proof_lens[live[a]]); });
        hip_check(hipMemcpyAsync(d_proofs, stage, pbytes, hipMemcpyHostToDevice, st), "copy proofs");
        std::vector<uint8_t> host(32 * nl + sz_idx + sz_ch + sz_sc, 0);
        std::vector<VerifyJob> jobs(nl);
        for (size_t a = 0; a < nl; a++) {
            const size_t p = live[a];
            memcpy(host.data() + 32 * a, roots + 32 * p, 32);
            VerifyJob &J = jobs[a];
            J.proof = d_proofs + poff[a];
            J.root = d_roots + 32 * a;
            J.idx = d_idx + q * a;
            J.chal = d_ch + (size_t)d * a;
            J.scalars = d_sc + w * a;
            J.final_poly = parsed[p].final_poly;
            J.open_off = parsed[p].open_off;
            J.rec_bytes = parsed[p].rec_bytes;
            J.roots_off = parsed[p].roots_off;
        }
        memcpy(host.data() + 32 * nl, idx.data(), 8 * idx.size());
        memcpy(host.data() + 32 * nl + sz_idx, chal.data(), sz_ch);
        memcpy(host.data() + 32 * nl + sz_idx + sz_ch, scal.data(), sz_sc);
        hip_check(hipMemcpyAsync(d_roots, host.data(), host.size(), hipMemcpyHostToDevice, st), "copy inputs");
        hip_check(hipMemcpyAsync(d_jobs, jobs.data(), sizeof(VerifyJob) * nl, hipMemcpyHostToDevice, st), "copy jobs");
        uint8_t *d_fold = dok.as<uint8_t>(), *d_path = d_fold + nl * q;
        ctx->timer.run(st, "k_verify_fold", [&] {
            return launch::verify_fold(st, d_jobs, (uint32_t)nl, (uint32_t)q, b, d, w, code->ft.as<gf>(),
                                       dvals.as<gf>(), d_fold);
        });
        ctx->timer.run(st, "k_verify_paths", [&] {
            return launch::verify_paths(st, d_jobs, (uint32_t)nl, (uint32_t)q, b, d, w, dvals.as<gf>(), d_path);
        });
        std::vector<uint8_t> ok(nl * q * (size_t)(d + 2));
        hip_check(hipMemcpyAsync(ok.data(), d_fold, ok.size(), hipMemcpyDeviceToHost, st), "copy results");
        hip_check(hipStreamSynchronize(st), "verify");
        ctx->timer.collect();
        const uint8_t *fold_ok = ok.data(), *path_ok = ok.data() + nl * q;
        const size_t P = (size_t)d + 1;
        for (size_t a = 0; a < nl; a++) {
            int r = 0;
            for (size_t t = 0; t < q && !r; t++) {
                const size_t pq = a * q + t;
                for (size_t j = 0; j < P && !r; j++)
                    if (!path_ok[pq * P + j]) r = 3;
                if (!r && !fold_ok[pq]) r = 2;
            }
            results[live[a]] = r;
        }
    });
}

int bfd_verify(bfd_ctx *ctx, const bfd_code *code, size_t width, const uint8_t root[32], const uint8_t *point,
               const uint8_t *values, const uint8_t *proof, size_t proof_len, bfd_challenger *challenger,
               int *result) {
    const uint8_t *pts[1] = {point}, *vals[1] = {values}, *prf[1] = {proof};
    bfd_challenger *chs[1] = {challenger};
    size_t lens[1] = {proof_len};
    return bfd_verify_batch(ctx, code, width, 1, root, pts, vals, prf, lens, chs, result);
}

}  // extern "C"
