// Launch shapes and their autotuner.
//
// A shape (bfd_tuning) decides how the plain path cuts its NTT and Merkle
// work into launches: the NTT pass plan (which layers share a pass over the
// codeword), block sizes, grid caps, the build of the three-layer pass, and
// where a Merkle tree switches from one launch per level to several levels
// per launch through LDS.  The engine's shape decides the same for batched
// and served work, plus how many blocks one polynomial may take in a
// segmented launch.  Every shape computes the same codewords, trees and
// proofs; only time changes.
//
// The tuner measures the real thing: commit + open (or a batched call) of
// synthetic polynomials of the caller's shape, timed on the host clock
// around calls that each end in a synchronisation.  Coordinate descent: from
// the current shape, each field in turn takes each of its candidate values
// with the others fixed; the median of `reps` runs decides, and a candidate
// replaces the incumbent only when it is more than 0.5% faster (below that,
// run-to-run noise decides).  The built-in and the chosen shapes are then
// re-measured alternately, so the reported gain carries no selection bias,
// and the chosen shape is installed only if it still wins there.
#include <algorithm>
#include <fstream>
#include <sstream>

#include "internal.hpp"

using namespace bfd::rt;

namespace {

struct Field {
    const char *name;
    uint32_t bfd_tuning::*f;
};
const Field kFields[] = {
    {"ntt_order", &bfd_tuning::ntt_order},
    {"ntt_block", &bfd_tuning::ntt_block},
    {"ntt_waves_per_cu", &bfd_tuning::ntt_waves_per_cu},
    {"ntt_occupancy", &bfd_tuning::ntt_occupancy},
    {"hash_block", &bfd_tuning::hash_block},
    {"hash_waves_per_cu", &bfd_tuning::hash_waves_per_cu},
    {"subtree_from_log2", &bfd_tuning::subtree_from_log2},
    {"subtree_levels", &bfd_tuning::subtree_levels},
    {"seg_blocks_per_cu", &bfd_tuning::seg_blocks_per_cu},
};

struct Knob {
    uint32_t bfd_tuning::*f;
    std::vector<uint32_t> values;
};

void check(int rc) {
    if (rc != BFD_OK) fail(rc, bfd_last_error());
}

void validate(const bfd_tuning &t) {
    const std::string why = tuning_problem(t);
    if (!why.empty()) fail(BFD_ERR_ARG, "invalid tuning: " + why);
}

double median(std::vector<double> v) {
    std::sort(v.begin(), v.end());
    const size_t m = v.size() / 2;
    return v.size() % 2 ? v[m] : 0.5 * (v[m - 1] + v[m]);
}

// Coordinate descent over `knobs` from `start`; `measure(t)` installs t and
// returns the median milliseconds of the workload under it.
template <class Measure>
bfd_tuning descend(const bfd_tuning &start, const std::vector<Knob> &knobs, Measure &&measure) {
    bfd_tuning best = start;
    double best_ms = measure(best);
    for (const Knob &k : knobs) {
        for (uint32_t v : k.values) {
            if (best.*(k.f) == v) continue;
            bfd_tuning t = best;
            t.*(k.f) = v;
            if (!tuning_problem(t).empty()) continue;
            const double ms = measure(t);
            if (ms < best_ms * 0.995) {
                best = t;
                best_ms = ms;
            }
        }
    }
    return best;
}

// Alternating re-measurement of the built-in and the chosen shape; the
// chosen one is kept only when it is still more than 0.5% faster (a descent
// step can win on noise), else the built-in shape is returned with its time.
template <class Measure>
bfd_tuning confirm(const bfd_tuning &chosen, Measure &&measure, double *ms_default, double *ms_tuned) {
    const bfd_tuning def = default_tuning();
    std::vector<double> a, b;
    for (int r = 0; r < 3; r++) {
        a.push_back(measure(def));
        if (!tuning_equal(chosen, def)) b.push_back(measure(chosen));
    }
    const double ta = median(a), tb = b.empty() ? ta : median(b);
    const bool keep = tb < ta * 0.995;
    if (ms_default) *ms_default = ta;
    if (ms_tuned) *ms_tuned = keep ? tb : ta;
    return keep ? chosen : def;
}

// synthetic evaluations: every field operation and hash runs the same
// instructions for any input, but the switching activity of the data moves
// the clock under the power limit, so the filler has the bit statistics of
// random elements (a constant fill runs about 2% faster than random data)
void fill_pattern(bfd_ctx *ctx, void *p, size_t bytes) {
    hip_check(launch::fill_mix(ctx->stream, (uint64_t *)p, bytes / 8, 0x5eed), "fill evaluations");
}

}  // namespace

extern "C" {

int bfd_tuning_default(bfd_tuning *out) {
    return guarded([&] {
        if (!out) fail(BFD_ERR_ARG, "null argument");
        *out = default_tuning();
    });
}

int bfd_tuning_get(const bfd_ctx *ctx, uint32_t log2_m, bfd_tuning *out) {
    return guarded([&] {
        if (!ctx || !out) fail(BFD_ERR_ARG, "null argument");
        if (log2_m == BFD_TUNING_BATCH) *out = ctx->batch_tuning;
        else if (log2_m < kTuneSizes) *out = ctx->tuning[log2_m];
        else fail(BFD_ERR_ARG, "log2_m must be 0..47 or BFD_TUNING_BATCH");
    });
}

int bfd_tuning_set(bfd_ctx *ctx, uint32_t log2_m, const bfd_tuning *t) {
    return guarded([&] {
        if (!ctx || !t) fail(BFD_ERR_ARG, "null argument");
        validate(*t);
        std::lock_guard<std::recursive_mutex> lk(ctx->mu);
        if (log2_m == BFD_TUNING_BATCH) ctx->batch_tuning = *t;
        else if (log2_m == BFD_TUNING_ALL) ctx->tuning.fill(*t);
        else if (log2_m < kTuneSizes) ctx->tuning[log2_m] = *t;
        else fail(BFD_ERR_ARG, "log2_m must be 0..47, BFD_TUNING_BATCH or BFD_TUNING_ALL");
    });
}

int bfd_tune(bfd_ctx *ctx, const bfd_code *code, uint32_t reps, bfd_tuning *chosen, double *ms_default,
             double *ms_tuned) {
    return guarded([&] {
        if (!ctx || !code) fail(BFD_ERR_ARG, "null argument");
        if (code->ctx != ctx) fail(BFD_ERR_ARG, "code belongs to another context");
        if (reps < 1 || reps > 99) fail(BFD_ERR_ARG, "reps must be 1..99");
        std::lock_guard<std::recursive_mutex> lk(ctx->mu);
        set_device(ctx);
        const uint32_t n = code->d, l = code->d + code->log2_c;
        if (l >= kTuneSizes) fail(BFD_ERR_ARG, "codeword too large to tune");
        const size_t N = (size_t)1 << n;
        DevBuf evals = ctx_buf(ctx, 16 * N);
        fill_pattern(ctx, evals.p, 16 * N);
        std::vector<uint8_t> point(16 * (size_t)(n ? n : 1));
        check(bfd_random_vec(7, n, point.data()));
        uint8_t claim[16];
        // the claim once (its evaluation is not part of what is tuned)
        {
            bfd_commitment *cm = nullptr;
            uint8_t root[32];
            check(bfd_commit_eval(ctx, code, evals.p, 1, N, 1, point.data(), &cm, root, claim));
            check(bfd_commitment_destroy(cm));
        }
        const bfd_tuning saved = ctx->tuning[l];
        auto run = [&]() {
            bfd_challenger *ch = nullptr;
            check(bfd_challenger_create(&ch));
            const double t0 = now_ms();
            bfd_commitment *cm = nullptr;
            uint8_t root[32];
            int rc = bfd_commit_device(ctx, code, evals.p, N, 1, &cm, root);
            uint8_t *proof = nullptr;
            size_t len = 0;
            if (rc == BFD_OK) rc = bfd_challenger_observe(ch, root, 32);
            if (rc == BFD_OK) rc = bfd_challenger_observe(ch, claim, 16);
            if (rc == BFD_OK) rc = bfd_open(ctx, code, cm, point.data(), claim, ch, &proof, &len);
            const double ms = now_ms() - t0;
            bfd_proof_free(proof);
            bfd_challenger_destroy(ch);
            if (cm) bfd_commitment_destroy(cm);
            check(rc);
            return ms;
        };
        auto measure = [&](const bfd_tuning &t) {
            ctx->tuning[l] = t;
            std::vector<double> v;
            for (uint32_t r = 0; r < reps; r++) v.push_back(run());
            return median(v);
        };
        try {
            measure(saved);  // warm-up: block cache, instruction cache
            std::vector<Knob> knobs = {
                {&bfd_tuning::ntt_order, {0, 1, 2}},
                {&bfd_tuning::ntt_occupancy, {2, 3}},
                {&bfd_tuning::ntt_block, {64, 128, 256}},
                {&bfd_tuning::ntt_waves_per_cu, {8, 16, 32, 64}},
                {&bfd_tuning::hash_block, {64, 128, 256}},
                {&bfd_tuning::hash_waves_per_cu, {16, 32, 64, 128}},
                {&bfd_tuning::subtree_from_log2, {}},
                {&bfd_tuning::subtree_levels, {6, 7, 8, 9}},
            };
            // thresholds below the largest tree (2^l leaves) change the plan
            for (uint32_t v = 14; v <= 18; v++)
                if (v < l) knobs[6].values.push_back(v);
            const bfd_tuning best = confirm(descend(saved, knobs, measure), measure, ms_default, ms_tuned);
            ctx->tuning[l] = best;
            if (chosen) *chosen = best;
        } catch (...) {
            ctx->tuning[l] = saved;
            throw;
        }
    });
}

int bfd_tune_batch(bfd_ctx *ctx, uint32_t lambda, uint32_t log2_c, size_t count, const uint32_t *log_sizes,
                   uint32_t reps, bfd_tuning *chosen, double *ms_default, double *ms_tuned) {
    return guarded([&] {
        if (!ctx || !log_sizes || count == 0) fail(BFD_ERR_ARG, "null argument or empty batch");
        if (reps < 1 || reps > 99) fail(BFD_ERR_ARG, "reps must be 1..99");
        std::lock_guard<std::recursive_mutex> lk(ctx->mu);
        set_device(ctx);
        size_t total = 0;
        for (size_t p = 0; p < count; p++) {
            if (log_sizes[p] > 30) fail(BFD_ERR_ARG, "polynomial too large to tune");
            total += (size_t)16 << log_sizes[p];
        }
        DevBuf slab = ctx_buf(ctx, total);
        fill_pattern(ctx, slab.p, total);
        std::vector<const void *> evals(count);
        std::vector<std::vector<uint8_t>> pts(count);
        std::vector<const uint8_t *> points(count);
        for (size_t p = 0, off = 0; p < count; p++) {
            evals[p] = slab.as<uint8_t>() + off;
            off += (size_t)16 << log_sizes[p];
            pts[p].resize(16 * (size_t)(log_sizes[p] ? log_sizes[p] : 1));
            check(bfd_random_vec(11 + p, log_sizes[p], pts[p].data()));
            points[p] = pts[p].data();
        }
        std::vector<uint8_t> roots(32 * count);
        std::vector<uint8_t *> proofs(count);
        std::vector<size_t> lens(count);
        std::vector<bfd_challenger *> chs(count, nullptr);
        auto run = [&]() {
            for (auto &c : chs) check(bfd_challenger_create(&c));
            std::fill(proofs.begin(), proofs.end(), nullptr);
            const double t0 = now_ms();
            const int rc = bfd_commit_open_batch(ctx, lambda, log2_c, count, log_sizes, evals.data(), 1,
                                                 points.data(), nullptr, chs.data(), 1, roots.data(),
                                                 proofs.data(), lens.data());
            const double ms = now_ms() - t0;
            for (auto &c : chs) {
                bfd_challenger_destroy(c);
                c = nullptr;
            }
            if (rc == BFD_OK)
                for (uint8_t *pr : proofs) bfd_proof_free(pr);
            check(rc);
            return ms;
        };
        const bfd_tuning saved = ctx->batch_tuning;
        auto measure = [&](const bfd_tuning &t) {
            ctx->batch_tuning = t;
            std::vector<double> v;
            for (uint32_t r = 0; r < reps; r++) v.push_back(run());
            return median(v);
        };
        try {
            measure(saved);
            const std::vector<Knob> knobs = {
                {&bfd_tuning::ntt_order, {0, 1, 2}},
                {&bfd_tuning::subtree_from_log2, {15, 16, 17, 18, 19}},
                {&bfd_tuning::subtree_levels, {7, 8, 9}},
                {&bfd_tuning::seg_blocks_per_cu, {2, 4, 8, 16, 32}},
            };
            const bfd_tuning best = confirm(descend(saved, knobs, measure), measure, ms_default, ms_tuned);
            ctx->batch_tuning = best;
            if (chosen) *chosen = best;
        } catch (...) {
            ctx->batch_tuning = saved;
            throw;
        }
    });
}

int bfd_tuning_save(const bfd_ctx *ctx, const char *path) {
    return guarded([&] {
        if (!ctx || !path) fail(BFD_ERR_ARG, "null argument");
        std::ostringstream os;
        auto line = [&](const std::string &key, const bfd_tuning &t) {
            os << key;
            for (const Field &f : kFields) os << ' ' << f.name << '=' << t.*(f.f);
            os << '\n';
        };
        const bfd_tuning def = default_tuning();
        for (uint32_t l = 0; l < kTuneSizes; l++)
            if (!tuning_equal(ctx->tuning[l], def)) line("log2_m=" + std::to_string(l), ctx->tuning[l]);
        if (!tuning_equal(ctx->batch_tuning, def)) line("batch", ctx->batch_tuning);
        std::ofstream f(path, std::ios::trunc);
        if (!f) fail(BFD_ERR_ARG, std::string("cannot write ") + path);
        f << os.str();
        if (!f.flush()) fail(BFD_ERR_ARG, std::string("cannot write ") + path);
    });
}

int bfd_tuning_load(bfd_ctx *ctx, const char *path) {
    return guarded([&] {
        if (!ctx || !path) fail(BFD_ERR_ARG, "null argument");
        std::ifstream f(path);
        if (!f) fail(BFD_ERR_ARG, std::string("cannot read ") + path);
        // parse everything first: a bad line changes nothing
        std::vector<std::pair<uint32_t, bfd_tuning>> entries;
        std::string ln;
        int lineno = 0;
        while (std::getline(f, ln)) {
            lineno++;
            std::istringstream is(ln);
            std::string tok;
            if (!(is >> tok)) continue;
            const std::string where = std::string(path) + ":" + std::to_string(lineno) + ": ";
            uint32_t key;
            if (tok == "batch") {
                key = BFD_TUNING_BATCH;
            } else if (tok.rfind("log2_m=", 0) == 0) {
                char *end = nullptr;
                const unsigned long v = strtoul(tok.c_str() + 7, &end, 10);
                if (!*(tok.c_str() + 7) || *end || v >= kTuneSizes) fail(BFD_ERR_ARG, where + "bad size " + tok);
                key = (uint32_t)v;
            } else {
                fail(BFD_ERR_ARG, where + "expected log2_m=<n> or batch");
            }
            bfd_tuning t = default_tuning();
            while (is >> tok) {
                const size_t eq = tok.find('=');
                const Field *fd = nullptr;
                for (const Field &x : kFields)
                    if (eq != std::string::npos && tok.compare(0, eq, x.name) == 0 && strlen(x.name) == eq) fd = &x;
                if (!fd) fail(BFD_ERR_ARG, where + "unknown field " + tok);
                char *end = nullptr;
                const unsigned long v = strtoul(tok.c_str() + eq + 1, &end, 10);
                if (!*(tok.c_str() + eq + 1) || *end || v > 0xffffffffu) fail(BFD_ERR_ARG, where + "bad value " + tok);
                t.*(fd->f) = (uint32_t)v;
            }
            const std::string why = tuning_problem(t);
            if (!why.empty()) fail(BFD_ERR_ARG, where + why);
            entries.push_back({key, t});
        }
        std::lock_guard<std::recursive_mutex> lk(ctx->mu);
        for (const auto &e : entries) {
            if (e.first == BFD_TUNING_BATCH) ctx->batch_tuning = e.second;
            else ctx->tuning[e.first] = e.second;
        }
    });
}

}  // extern "C"
