// Step-wise SumcheckFunctionProvers on the device: QuadraticProver
// (sumcheck/src/function/quadratic.rs) and EvalProver
// (sumcheck/src/function/eval.rs over eval_impr.rs).
//
// The reference's sumcheck driver (prove_sumcheck, sumcheck/src/lib.rs:56-103)
// calls a SumcheckFunctionProver once per round: compute_round_poly, then
// fix_last_var with the sampled challenge, and evaluations at the end.  A
// bfd_sc handle is that prover with its tables resident in device memory, so
// a caller's own driver and challenger (a Rust impl of the trait, say) run the
// protocol while the table work stays on the GPU.  bfd_sumcheck_quadratic is
// the Quadratic function with the driver and transcript inside the library,
// and bfd_open runs the Eval prover inside the Basefold rounds.
#include "internal.hpp"

using namespace bfd::rt;

enum class ScKind { Quadratic, Eval };

struct bfd_sc {
    ScKind kind = ScKind::Quadratic;
    bfd_ctx *ctx = nullptr;
    uint32_t n = 0;      // num_vars at creation (table stride 2^n)
    uint32_t cur = 0;    // variables left (tables hold 2^cur live entries)
    uint32_t m = 0, npairs = 0;
    std::vector<gf> scalars;
    std::vector<uint32_t> pair_idx;  // a_0, b_0, a_1, b_1, ...
    DevBuf P, pairs, partials;
    unsigned max_blocks = 1;
    // Eval: the point's scalars and the inner and outer subclaims (the same
    // host state an opening keeps, round k = n - cur), the pending inner
    // round polynomial, the halved eq table and a second table buffer the
    // fixes alternate with
    OpenRounds ev;
    OpenRounds::RoundPoly pending{};
    DevBuf Q, half_eq;
    gf *live = nullptr;
};

namespace {

// A0 = sum f0 g0, A1 = sum f1 g1, A2 = sum (f0 + f1)(g0 + g1) of every pair
// over the halves of the live tables (2^cur entries, cur >= 1), each scaled
// by its pair's scalar and summed over the pairs.
void pair_sums(bfd_sc *h, gf &s0, gf &s1, gf &s2) {
    bfd_ctx *ctx = h->ctx;
    hipStream_t st = ctx->stream;
    const size_t half = (size_t)1 << (h->cur - 1);
    unsigned nb = 0;
    ctx->timer.run(st, "k_quad_sums", [&] {
        return launch::quad_sums(st, h->P.as<gf>(), (size_t)1 << h->n, h->pairs.as<uint32_t>(), h->npairs, half,
                                 h->partials.as<gf>(), h->max_blocks, &nb);
    });
    const size_t bytes = 48 * (size_t)nb * h->npairs;
    ctx->pinned.ensure(bytes + 64);
    uint8_t *pin = (uint8_t *)ctx->pinned.p;
    hip_check(hipMemcpyAsync(pin, h->partials.p, bytes, hipMemcpyDeviceToHost, st), "copy partials");
    hip_check(hipStreamSynchronize(st), "sumcheck sync");
    ctx->timer.collect();
    s0 = s1 = s2 = gf_zero();
    for (uint32_t t = 0; t < h->npairs; t++) {
        gf A0 = gf_zero(), A1 = gf_zero(), A2 = gf_zero();
        const uint8_t *q = pin + 48 * (size_t)nb * t;
        for (unsigned u = 0; u < nb; u++) {
            A0 = gf_add(A0, load_elem(q + 48 * u));
            A1 = gf_add(A1, load_elem(q + 48 * u + 16));
            A2 = gf_add(A2, load_elem(q + 48 * u + 32));
        }
        s0 = gf_add(s0, hmul(h->scalars[t], A0));
        s1 = gf_add(s1, hmul(h->scalars[t], A1));
        s2 = gf_add(s2, hmul(h->scalars[t], A2));
    }
}

// The m live values P_p[0] once every variable is fixed.
std::vector<gf> finals(bfd_sc *h) {
    hipStream_t st = h->ctx->stream;
    std::vector<gf> v(h->m);
    hip_check(hipMemcpy2DAsync(v.data(), 16, h->P.p, 16 * ((size_t)1 << h->n), 16, h->m, hipMemcpyDeviceToHost, st),
              "copy evaluations");
    hip_check(hipStreamSynchronize(st), "sumcheck sync");
    return v;
}

// sum over b < h of half_eq[b] * t[b] for the Eval prover (k_dot, block sums
// added on the host)
gf eval_dot(bfd_sc *h, const gf *t, size_t len) {
    bfd_ctx *ctx = h->ctx;
    hipStream_t st = ctx->stream;
    unsigned nb = 0;
    ctx->timer.run(st, "k_dot", [&] { return launch::dot(st, h->half_eq.as<gf>(), t, len, h->partials.as<gf>(), &nb); });
    ctx->pinned.ensure(16 * (size_t)nb + 64);
    uint8_t *pin = (uint8_t *)ctx->pinned.p;
    hip_check(hipMemcpyAsync(pin, h->partials.p, 16 * (size_t)nb, hipMemcpyDeviceToHost, st), "copy sums");
    hip_check(hipStreamSynchronize(st), "sumcheck sync");
    ctx->timer.collect();
    gf s = gf_zero();
    for (unsigned u = 0; u < nb; u++) s = gf_add(s, load_elem(pin + 16 * (size_t)u));
    return s;
}

// the single live entry of an Eval prover's table
gf eval_final(bfd_sc *h) {
    hipStream_t st = h->ctx->stream;
    uint8_t b[16];
    hip_check(hipMemcpyAsync(b, h->live, 16, hipMemcpyDeviceToHost, st), "copy evaluation");
    hip_check(hipStreamSynchronize(st), "sumcheck sync");
    return load_elem(b);
}

}  // namespace

extern "C" {

int bfd_sc_quadratic_create(bfd_ctx *ctx, uint32_t n, uint32_t m, const void *polys, int polys_on_device,
                            uint32_t npairs, const uint8_t *scalars, const uint32_t *pair_idx, bfd_sc **out) {
    return guarded([&] {
        if (!ctx || !polys || !scalars || !pair_idx || !out) fail(BFD_ERR_ARG, "null argument");
        if (n > 32) fail(BFD_ERR_ARG, "too many variables");
        if (m == 0 || m > 65535 || npairs == 0 || npairs > 65535) fail(BFD_ERR_ARG, "need 1..65535 tables and pairs");
        for (uint32_t k = 0; k < 2 * npairs; k++)
            if (pair_idx[k] >= m) fail(BFD_ERR_ARG, "pair index out of range");
        std::lock_guard<std::recursive_mutex> lk(ctx->mu);
        set_device(ctx);
        hipStream_t st = ctx->stream;
        auto h = std::make_unique<bfd_sc>();
        h->ctx = ctx;
        h->n = h->cur = n;
        h->m = m;
        h->npairs = npairs;
        h->scalars.resize(npairs);
        for (uint32_t k = 0; k < npairs; k++) h->scalars[k] = load_elem(scalars + 16 * k);
        h->pair_idx.assign(pair_idx, pair_idx + 2 * (size_t)npairs);
        const size_t N = (size_t)1 << n;
        h->P = DevBuf(16 * N * m, st);
        h->pairs = DevBuf(8 * (size_t)npairs, st);
        hip_check(hipMemcpyAsync(h->P.p, polys, 16 * N * m,
                                 polys_on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, st),
                  "copy tables");
        hip_check(hipMemcpyAsync(h->pairs.p, pair_idx, 8 * (size_t)npairs, hipMemcpyHostToDevice, st), "copy pairs");
        h->max_blocks = std::max(1u, std::min(512u, 4096u / npairs));
        h->partials = DevBuf(48 * (size_t)h->max_blocks * npairs, st);
        // the host arrays are read by the copies above before the call returns
        hip_check(hipStreamSynchronize(st), "sumcheck setup");
        *out = h.release();
    });
}

int bfd_sc_eval_create(bfd_ctx *ctx, uint32_t n, const void *poly, int poly_on_device, const uint8_t *point,
                       bfd_sc **out) {
    return guarded([&] {
        if (!ctx || !poly || !out || (n && !point)) fail(BFD_ERR_ARG, "null argument");
        if (n > 32) fail(BFD_ERR_ARG, "too many variables");
        std::lock_guard<std::recursive_mutex> lk(ctx->mu);
        set_device(ctx);
        hipStream_t st = ctx->stream;
        auto h = std::make_unique<bfd_sc>();
        h->kind = ScKind::Eval;
        h->ctx = ctx;
        h->n = h->cur = n;
        h->m = 1;
        // EvalImpr::new (r^-1) and HalfEq::new ((1 - r_j)^-1, correcting
        // factors); the claim is the caller's subclaim of the first round
        const uint8_t zero[16] = {0};
        h->ev.init((int)n, 0, 0, point, zero);
        const size_t N = (size_t)1 << n;
        h->P = DevBuf(16 * N, st);
        h->Q = DevBuf(16 * (N > 1 ? N / 2 : 1), st);
        hip_check(hipMemcpyAsync(h->P.p, poly, 16 * N, poly_on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice,
                                 st),
                  "copy table");
        const int hn = h->ev.half_eq_vars();
        h->half_eq = DevBuf(16 * ((size_t)1 << hn), st);
        if (n > 0) device_eq(ctx, h->ev.r.data(), hn, h->half_eq.as<gf>());
        h->partials = DevBuf(16 * 512, st);
        h->live = h->P.as<gf>();
        // the host table and the eq staging are read before the call returns
        hip_check(hipStreamSynchronize(st), "sumcheck setup");
        *out = h.release();
    });
}

int bfd_sc_num_vars(const bfd_sc *h, uint32_t *num_vars) {
    return guarded([&] {
        if (!h || !num_vars) fail(BFD_ERR_ARG, "null argument");
        *num_vars = h->cur;
    });
}

int bfd_sc_compute_sum(bfd_sc *h, uint8_t out[16]) {
    return guarded([&] {
        if (!h || !out) fail(BFD_ERR_ARG, "null argument");
        std::lock_guard<std::recursive_mutex> lk(h->ctx->mu);
        set_device(h->ctx);
        gf sum = gf_zero();
        if (h->kind == ScKind::Eval) {
            // EvalProver::compute_sum (eval.rs:86-88): the inner sum
            // cf(round) * evaluate([sum T lo, sum T hi], [r_round])
            // (eval_impr.rs:166-185) times the delta scalar
            if (h->n == 0) {
                sum = eval_final(h);  // eval_impr.rs:167-169
            } else {
                if (h->cur == 0) fail(BFD_ERR_STATE, "no variable left");
                const int round = (int)h->cur - 1;
                const size_t half = (size_t)1 << round;
                const gf lo = eval_dot(h, h->live, half), hi = eval_dot(h, h->live + half, half);
                const OpenRounds &o = h->ev;
                const gf inner = hmul(o.cf[round], gf_add(lo, hmul(o.r[round], gf_add(hi, lo))));
                sum = hmul(inner, o.s_delta);
            }
        } else if (h->cur == 0) {
            const std::vector<gf> v = finals(h);
            for (uint32_t t = 0; t < h->npairs; t++)
                sum = gf_add(sum, hmul(h->scalars[t], hmul(v[h->pair_idx[2 * t]], v[h->pair_idx[2 * t + 1]])));
        } else {
            gf s0, s1, s2;
            pair_sums(h, s0, s1, s2);
            sum = gf_add(s0, s1);  // both halves of sum_b f g
        }
        store_elem(out, sum);
    });
}

int bfd_sc_round_poly(bfd_sc *h, const uint8_t subclaim[16], uint8_t coeffs[48]) {
    return guarded([&] {
        if (!h || !subclaim || !coeffs) fail(BFD_ERR_ARG, "null argument");
        if (h->cur == 0) fail(BFD_ERR_STATE, "no variable left");
        std::lock_guard<std::recursive_mutex> lk(h->ctx->mu);
        set_device(h->ctx);
        if (h->kind == ScKind::Eval) {
            // EvalProver::compute_round_poly (eval.rs:90-104): the first round
            // takes the caller's subclaim, later rounds the inner one
            const int k = (int)(h->n - h->cur);
            if (k == 0) h->ev.subclaim = load_elem(subclaim);
            const gf S = eval_dot(h, h->live, (size_t)1 << (h->cur - 1));
            h->pending = h->ev.round_poly(k, S);
            store_elem(coeffs, h->pending.rp0);
            store_elem(coeffs + 16, h->pending.rp1);
            store_elem(coeffs + 32, h->pending.rp2);
            return;
        }
        gf c0, s1, c2;
        pair_sums(h, c0, s1, c2);
        // decompress_round_poly (quadratic.rs:70-80): c1 = subclaim - 2 c0 - c2,
        // and 2 c0 = 0 in characteristic 2
        const gf c1 = gf_add(load_elem(subclaim), c2);
        store_elem(coeffs, c0);
        store_elem(coeffs + 16, c1);
        store_elem(coeffs + 32, c2);
    });
}

int bfd_sc_fix_last_var(bfd_sc *h, const uint8_t x[16]) {
    return guarded([&] {
        if (!h || !x) fail(BFD_ERR_ARG, "null argument");
        if (h->cur == 0) fail(BFD_ERR_STATE, "no variable left");
        std::lock_guard<std::recursive_mutex> lk(h->ctx->mu);
        set_device(h->ctx);
        hipStream_t st = h->ctx->stream;
        const size_t half = (size_t)1 << (h->cur - 1);
        const gf xv = load_elem(x);
        if (h->kind == ScKind::Eval) {
            // poly.fix_last_var, half_eq.halve (a truncation: the table is
            // read over fewer entries), inner subclaim reduce (eval.rs:106-111)
            h->ev.advance((int)(h->n - h->cur), h->pending, xv);
            h->pending = OpenRounds::RoundPoly{};
            gf *dst = h->live == h->P.as<gf>() ? h->Q.as<gf>() : h->P.as<gf>();
            const gf *src = h->live;
            h->ctx->timer.run(st, "k_fix", [&] { return launch::fix(st, src, dst, half, xv); });
            h->live = dst;
            h->cur--;
            return;
        }
        h->ctx->timer.run(st, "k_fix_multi",
                          [&] { return launch::fix_multi(st, h->P.as<gf>(), (size_t)1 << h->n, h->m, half, xv); });
        h->cur--;
    });
}

int bfd_sc_evaluations(bfd_sc *h, uint8_t *out) {
    return guarded([&] {
        if (!h || !out) fail(BFD_ERR_ARG, "null argument");
        if (h->cur != 0) fail(BFD_ERR_STATE, "variables left: evaluations exist once all are fixed");
        std::lock_guard<std::recursive_mutex> lk(h->ctx->mu);
        set_device(h->ctx);
        if (h->kind == ScKind::Eval) {
            store_elem(out, eval_final(h));
            return;
        }
        const std::vector<gf> v = finals(h);
        for (uint32_t p = 0; p < h->m; p++) store_elem(out + 16 * (size_t)p, v[p]);
    });
}

int bfd_sc_destroy(bfd_sc *h) {
    return guarded([&] {
        if (!h) return;
        {
            std::lock_guard<std::recursive_mutex> lk(h->ctx->mu);
            set_device(h->ctx);
            (void)hipStreamSynchronize(h->ctx->stream);
        }
        delete h;
    });
}

}  // extern "C"
