"""Benchmark of the Basefold commit+open hot path on one MI355X.

Workload: one multilinear polynomial with 2^24 evaluations over POLYVAL
GF(2^128), rate 1/4 (log_blowup 2), uniform random evaluations from
StdRng seed 0, random opening point from seed 1.  One step = commit (RS
encode via the additive NTT + Keccak Merkle tree) + the harness transcript
prefix + open (interleaved fold/sumcheck rounds, pi-tree commitments, 189
queries) with the evaluations already resident in HBM.  The claim is
evaluated once before timing, as the reference bench excludes `eval`
(poly-commit/benches/poly-commit.rs:147-158).

Prints one JSON line.  `value` = evaluations committed+opened per second.
"""
from __future__ import annotations

import argparse
import gc
import json
import os
import statistics
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

LOG_N = 24
LOG2_C = 2
LAMBDA = 128
HBM_PEAK_GBS = 8000.0


def algorithmic_bytes(n: int, log2_c: int) -> dict:
    """Compulsory HBM bytes per commit+open and per-kernel-family units.

    Each persistent or inter-round array is read once and written once; the
    leaf hash of a 16-byte element reads 16 B and writes a 32-byte digest, a
    node reads two digests and writes one."""
    N, c = 1 << n, 1 << log2_c
    M = N * c
    S, H = 16, 32
    commit = S * N + S * M + H * (M - 1)
    opn = 1.5 * S * (2 * M - 2 * c) + H * (M - c - n) + 2 * S * (2 * N - 2) + S * N / 2
    return {"commit": commit, "open": opn, "total": commit + opn}


def valu_work(n: int, log2_c: int) -> tuple:
    """Keccak-f permutations K and GF(2^128) multiplications of one
    commit+open: K = (2M-1) + (2M-2c-n) (leaves and nodes of the main tree and
    of every pi tree); G = n*M/2 (NTT) + (2M-2c) (two per fold output) +
    (2N-2) (fix_last_var and the inner products) + N/2 (eq table), returned
    as (K, G_reg, G_tab): G_tab are the products by a round challenge that go
    through its window table (the fold's second product and the fix, in the
    rounds whose fold has 2^19 outputs or more, where every thread of the
    capped grid has two or more), G_reg the rest."""
    N, c = 1 << n, 1 << log2_c
    M = N * c
    G = n * M // 2 + (2 * M - 2 * c) + (2 * N - 2) + N // 2
    G_tab = sum((M >> (k + 1)) + (N >> (k + 1)) for k in range(n) if (M >> (k + 1)) >= 1 << 19)
    return (2 * M - 1) + (2 * M - 2 * c - n), G - G_tab, G_tab


def valu_floor_s(K: int, G_reg: int, G_tab: int, rates) -> float:
    """Seconds the work would take at the isolated kernels' rates (Keccak-f,
    register multiply, window-table multiply)."""
    return K / rates[0] + G_reg / rates[1] + G_tab / rates[2]


def round_floor_s(rounds, rates) -> float:
    """Lower bound for a sumcheck on this device whose transcript steps run on
    it: a round's polynomial needs the previous round's challenge, and its
    challenge one Keccak-f permutation of the transcript, so every round takes
    at least the longer of its multiplications at the isolated rates and one
    permutation at the single-wave latency of the lane-pair form the
    transcript uses (rates[3]).  rounds = [(G_reg, G_tab)] per round."""
    return sum(max(gr / rates[1] + gt / rates[2], rates[3]) for gr, gt in rounds)


SUBTREE_FROM = 1 << 17


def single_level_nodes(leaves: int, subtree_from: int = SUBTREE_FROM) -> int:
    """Nodes hashed by k_nodes launches in a tree over `leaves` digests: every
    level whose input layer holds more than `subtree_from` digests (the
    shape's threshold)."""
    nodes, h = 0, leaves
    while h > subtree_from:
        nodes += h // 2
        h //= 2
    return nodes


def traffic_ratio(kernel: str):
    path = os.path.join(ROOT, "profiles", "basefold_2e24_rate4_kernels.md")
    try:
        for line in open(path):
            parts = line.split()
            if len(parts) == 3 and parts[0] == "traffic_ratio" and parts[1] == kernel:
                return float(parts[2])
    except OSError:
        pass
    return None


def valu_per_node(kernel: str):
    """VALU instructions per node hash per lane (SQ_INSTS_VALU x 64 / nodes)
    from the committed rocprof PMC summary of this kernel, when one exists."""
    path = os.path.join(ROOT, "profiles", "basefold_2e24_rate4_kernels.md")
    try:
        for line in open(path):
            parts = line.split()
            if len(parts) == 5 and parts[0] == "valu_issue" and parts[1] == kernel:
                return float(parts[4])
    except OSError:
        pass
    return None


# one wave64 VALU instruction per CU per cycle (4 SIMDs x 16 lanes, 4 cycles
# per instruction) on 256 CUs at 2.4 GHz
VALU_ISSUE_PEAK = 256 * 2.4e9


def cpu_baseline(args, gpu_proof=None) -> dict:
    """The CPU restatement (oracle/) with OpenMP on this host's cores, timed at
    the benchmark's own size (2^LOG_N evaluations at rate 1/4, same seeded
    evaluations and point): the median of `cpu_full_samples` runs after the
    smaller runs below have warmed the thread pool.  Its proof must equal the
    GPU proof of the timed steps.  Also reported: commit+open at 2^cpu_log_n
    (median of `cpu_samples` after a warm-up) and the restatement's verifier
    on one proof of the batched-verification workload (one core)."""
    import oracle as O
    cores = min(16, os.cpu_count() or 1)
    O.set_threads(cores)
    n = args.cpu_log_n
    ev, pt = O.random_vec(0, 1 << n), O.random_vec(1, n)
    O.basefold_prove(LAMBDA, LOG2_C, ev, pt)  # warm-up
    times = []
    for _ in range(args.cpu_samples):
        t0 = time.perf_counter()
        O.basefold_prove(LAMBDA, LOG2_C, ev, pt)
        times.append(time.perf_counter() - t0)
    t_small = statistics.median(times)
    del ev, pt
    full = None
    if args.cpu_full_samples > 0:
        nf = args.log_n
        evf, ptf = O.random_vec(0, 1 << nf), O.random_vec(1, nf)
        tf = []
        proof = None
        for _ in range(args.cpu_full_samples):
            t0 = time.perf_counter()
            _, _, proof = O.basefold_prove(LAMBDA, LOG2_C, evf, ptf)
            tf.append(time.perf_counter() - t0)
            print(f"cpu baseline sample 2^{nf}: {tf[-1]:.1f} s", file=sys.stderr, flush=True)
        full = {"log_n": nf, "seconds": tf, "median_s": statistics.median(tf),
                "proof_equals_gpu": (proof == gpu_proof) if gpu_proof is not None else None}
        if gpu_proof is not None and proof != gpu_proof:
            raise SystemExit("the CPU restatement's proof differs from the GPU proof at the benchmark size")
        del evf, ptf
    # the restatement's verifier on one proof of the batched-verification workload (2^16, rate 1/2)
    O.set_threads(1)
    ev1, pt1 = O.random_vec(3000, 1 << 16), O.random_vec(4000, 16)
    root1, claim1, proof1 = O.basefold_prove(LAMBDA, 1, ev1, pt1)
    vt = []
    for _ in range(3):
        t0 = time.perf_counter()
        assert O.basefold_verify(LAMBDA, 1, root1, pt1, claim1, proof1) == 0
        vt.append(time.perf_counter() - t0)
    O.set_threads(cores)
    host = f"C restatement of the reference CPU path, OpenMP {cores} threads, PCLMUL; host {cpu_model()}, nproc {os.cpu_count()}"
    if full:
        value = (1 << full["log_n"]) / full["median_s"]
        sample = (f"commit+open of 2^{full['log_n']} evaluations at rate 1/4, the benchmark's own inputs "
                  f"(median of {len(full['seconds'])} after warm-up runs at 2^{n}; proof byte-equal to the GPU "
                  f"proof), {host}")
    else:
        value = (1 << n) / t_small
        sample = f"commit+open of 2^{n} evaluations at rate 1/4 (median of {len(times)} after a warm-up), {host}"
    return {"value": value, "unit": "elems/s", "cores": cores, "kind": "port", "sample": sample,
            "full_size": full,
            "small": {"log_n": n, "median_s": t_small, "samples": len(times), "elems_per_s": (1 << n) / t_small},
            "verify_ms_per_proof": 1e3 * statistics.median(vt), "verify_cores": 1}


def stage(msg: str) -> None:
    """Progress on stderr (the JSON line is the only stdout output)."""
    print(f"bench: {msg}", file=sys.stderr, flush=True)


def cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown CPU"


def stepwise_open(bfd, pcs, cd, claim, point, reps: int) -> dict:
    """Basefold::open with the transcript on a host challenger the caller
    drives (bfd_open_begin .. bfd_open_finish: one device round trip per
    round), against the same opening with the transcript on the device, on
    one commitment; the two proofs must be byte-equal."""
    def fresh():
        ch = bfd.GenericChallenger.keccak256()
        ch.observe(cd.comm)
        ch.observe_ext_slice(claim)
        return ch

    ev = [bfd.PolyEvals(claim, point)]
    dev_ms, step_ms = [], []
    for _ in range(reps):
        t0 = time.perf_counter()
        p_dev = pcs.open(cd, ev, fresh())
        t1 = time.perf_counter()
        p_step = pcs.open_stepwise(cd, ev, fresh())
        t2 = time.perf_counter()
        if p_dev != p_step:
            raise SystemExit("step-wise open gave a proof different from the device-transcript open")
        dev_ms.append(1e3 * (t1 - t0))
        step_ms.append(1e3 * (t2 - t1))
    rounds = len(point)
    d, s = float(np.median(dev_ms)), float(np.median(step_ms))
    return {"workload": f"open of one 2^{rounds} commitment, host GenericChallenger driven from Python",
            "rounds": rounds, "device_transcript_ms": d, "stepwise_ms": s,
            "extra_ms_per_round": (s - d) / max(1, rounds), "reps": reps, "proofs_equal": True}


def product_sumcheck(bfd, ctx, torch, reps: int, rates=None) -> dict:
    """Config 3: degree-3 product sumcheck of f g eq(r, .) over 22 variables
    (f, g uniform from seeds 0 and 1, r from seed 2), tables resident in HBM;
    one call = claim + 22 rounds with the transcript (fresh challenger)."""
    n = 22
    f, g, r = bfd.random_vec(0, 1 << n), bfd.random_vec(1, 1 << n), bfd.random_vec(2, n)
    df = torch.from_numpy(f.view(np.int64).reshape(-1)).to("cuda:0")
    dg = torch.from_numpy(g.view(np.int64).reshape(-1)).to("cuda:0")
    torch.cuda.synchronize()
    ptrs = (df.data_ptr(), dg.data_ptr())

    def run():
        return bfd.prove_sumcheck_product(None, None, r, bfd.GenericChallenger.keccak256(), ctx, device_ptrs=ptrs)

    ref = run()  # warm-up
    ts = []
    for _ in range(reps):
        ctx.synchronize()
        t0 = time.perf_counter()
        out = run()
        ctx.synchronize()
        ts.append(time.perf_counter() - t0)
    assert all((a == b).all() for a, b in zip(out, ref)), "product sumcheck not deterministic"
    t = statistics.median(ts)
    # multiplications: five per pair in round 0 (A f0, A f1 with the eq
    # factor A, and three products) and nine after (four folds and five);
    # the eq factor B enters once per block, a few thousand products.  The
    # folds of rounds with more than 2^18 pairs (two or more per thread) go
    # through the challenge's window table.
    per_round = [(5 * (1 << (n - 1)), 0)]
    for k in range(1, n):
        h = 1 << (n - 1 - k)
        per_round.append((5 * h, 4 * h) if h > 1 << 18 else (9 * h, 0))
    G = sum(a + b for a, b in per_round)
    Gt = sum(b for _, b in per_round)
    pos = {"gf_muls": G, "gf_muls_by_table": Gt, "keccak_perms": n}
    if rates:
        pos["valu_floor_ms"] = 1e3 * valu_floor_s(n, G - Gt, Gt, rates)
        pos["valu_frac"] = pos["valu_floor_ms"] / (1e3 * t)
        pos["transcript_perm_latency_us"] = 1e6 * rates[3]
        pos["round_floor_ms"] = 1e3 * round_floor_s(per_round, rates)
        pos["round_floor_frac"] = pos["round_floor_ms"] / (1e3 * t)
        pos["note"] = ("valu_floor counts the multiplications at the isolated rates only; round_floor also gives "
                       "every round at least one transcript permutation at its single-wave latency, since the "
                       "rounds are sequential")
    return {"workload": "f g eq(r, .), 22 variables, tables resident, fresh challenger", "ms": 1e3 * t,
            "evals_per_s": (1 << n) / t, "position": pos}


def mix_position(jobs, seconds: float, rates) -> dict:
    """Roofline position of a set of commit+open jobs [(log2 n, log2 c)]
    completed in `seconds`: algorithmic HBM GB/s (and its fraction of 8 TB/s)
    and the VALU fraction (the time of its Keccak-f permutations and
    multiplications at the isolated kernels' rates over t; the engine's
    window-table products counted as on the plain path, an approximation)."""
    B = sum(algorithmic_bytes(n, c)["total"] for n, c in jobs)
    K = Gr = Gt = 0
    for n, c in jobs:
        k, gr, gt = valu_work(n, c)
        K += k
        Gr += gr
        Gt += gt
    out = {"hbm_gbs_algorithmic": B / seconds / 1e9, "hbm_frac": B / seconds / 1e9 / HBM_PEAK_GBS,
           "keccak_perms": K, "gf_muls": Gr + Gt, "gf_muls_by_table": Gt}
    if rates:
        f = valu_floor_s(K, Gr, Gt, rates)
        out["valu_floor_ms"] = 1e3 * f
        out["valu_frac"] = f / seconds
    return out


def quadratic_sumcheck(bfd, ctx, torch, reps: int, rates=None) -> dict:
    """prove_sumcheck of a Quadratic (the reference's general sumcheck
    function, quadratic.rs) over 8 tables of 2^20 evaluations with 8 random
    pairs and scalars (numpy seed 5), tables resident, fresh challenger."""
    n, m, npairs = 20, 8, 8
    rng = np.random.default_rng(5)
    polys = bfd.random_vec(50, m << n).reshape(m, 1 << n, 2)
    scalars = bfd.random_vec(51, npairs)
    pairs = [(scalars[k], int(rng.integers(0, m)), int(rng.integers(0, m))) for k in range(npairs)]
    f = bfd.Quadratic(polys, pairs)
    dP = torch.from_numpy(np.ascontiguousarray(polys).view(np.int64).reshape(-1)).to("cuda:0")
    torch.cuda.synchronize()

    def run():
        return bfd.prove_sumcheck_quadratic(f, bfd.GenericChallenger.keccak256(), ctx, device_ptr=dP.data_ptr(),
                                            shape=(m, 1 << n))

    ref = run()
    ts = []
    for _ in range(reps):
        ctx.synchronize()
        t0 = time.perf_counter()
        out = run()
        ctx.synchronize()
        ts.append(time.perf_counter() - t0)
    assert all((np.asarray(a) == np.asarray(b)).all() for a, b in zip(out, ref)), "quadratic sumcheck not deterministic"
    t = statistics.median(ts)
    # three products per pair and pair position, one per table entry for the
    # fix; the fixes of rounds of 2^17 pairs or more (two or more per thread)
    # go through the challenge's window table
    per_round = []
    for k in range(n):
        h = 1 << (n - 1 - k)
        per_round.append((3 * npairs * h, m * h) if h >= 1 << 17 else (3 * npairs * h + m * h, 0))
    G = sum(a + b for a, b in per_round)
    Gt = sum(b for _, b in per_round)
    pos = {"gf_muls": G, "gf_muls_by_table": Gt, "keccak_perms": n}
    if rates:
        pos["valu_floor_ms"] = 1e3 * valu_floor_s(n, G - Gt, Gt, rates)
        pos["valu_frac"] = pos["valu_floor_ms"] / (1e3 * t)
        pos["transcript_perm_latency_us"] = 1e6 * rates[3]
        pos["round_floor_ms"] = 1e3 * round_floor_s(per_round, rates)
        pos["round_floor_frac"] = pos["round_floor_ms"] / (1e3 * t)
    return {"workload": f"{m} tables of 2^{n}, {npairs} pairs, tables resident, fresh challenger", "ms": 1e3 * t,
            "table_evals_per_s": m * (1 << n) / t, "position": pos}


def batched_vs_plain(bfd, ctx, torch, reps: int, tune_reps: int = 0, rates=None) -> dict:
    """Config 4: 64 polynomials, log2 sizes uniform in {12..22} (numpy seed 0),
    rate 1/2, evaluations and points seeded per polynomial, resident in HBM.
    Times one batched call against 64 plain commit+open calls; both produce
    the same 64 proofs.  With tune_reps > 0 the plain path is first tuned for
    every size of the mix and the engine for the mix itself."""
    rng = np.random.default_rng(0)
    sizes = [int(v) for v in rng.integers(12, 23, size=64)]
    log2_c = 1
    evals = [bfd.random_vec(10 + i, 1 << n) for i, n in enumerate(sizes)]
    points = [bfd.random_vec(1000 + i, n) for i, n in enumerate(sizes)]
    dev = [torch.from_numpy(e.view(np.int64).reshape(-1)).to("cuda:0") for e in evals]
    torch.cuda.synchronize()
    ptrs = [d.data_ptr() for d in dev]
    pcs = {n: bfd.Basefold.setup(bfd.BasefoldConfig(bfd.BinaryReedSolomonCode(LAMBDA, log2_c, n)), ctx)
           for n in sorted(set(sizes))}
    claims = []
    for n, ptr, pt in zip(sizes, ptrs, points):
        cd = pcs[n].commit_device(ptr, 1 << n)
        claims.append(pcs[n].eval(cd, pt))
        del cd
    tuning = None
    if tune_reps > 0:
        t_tune = time.perf_counter()
        for n in sorted(set(sizes)):
            ctx.tune(pcs[n], reps=tune_reps)
        chosen, ms_def, ms_tuned = ctx.tune_batch(sizes, log2_c=log2_c, lambda_=LAMBDA, reps=tune_reps)
        tuning = {"engine_shape": chosen.as_dict(), "batch_ms_default": ms_def, "batch_ms_tuned": ms_tuned,
                  "tune_s": time.perf_counter() - t_tune}

    def plain():
        out = []
        for n, ptr, pt, cl in zip(sizes, ptrs, points, claims):
            cd = pcs[n].commit_device(ptr, 1 << n)
            ch = bfd.GenericChallenger.keccak256()
            ch.observe(cd.comm)
            ch.observe_ext_slice(cl)
            out.append(pcs[n].open(cd, [bfd.PolyEvals(cl, pt)], ch))
            del cd
        return out

    def batched():
        chs = [bfd.GenericChallenger.keccak256() for _ in sizes]
        return bfd.commit_open_batch(LAMBDA, log2_c, sizes, points, claims, chs, True, device_ptrs=ptrs, ctx=ctx)[1]

    p_ref, b_ref = plain(), batched()  # warm-up (also builds every code)
    assert p_ref == [b.tobytes() for b in b_ref], "batched proofs differ from the plain proofs"
    # each leg's proofs are dropped after its clock stops: freeing 64 fresh
    # output buffers (~100 MB, unmapped page by page) costs the caller
    # ~2.5 ms and is not the prover's work
    tp, tb = [], []
    for _ in range(reps):
        ctx.synchronize()
        t0 = time.perf_counter()
        out = plain()
        ctx.synchronize()
        tp.append(time.perf_counter() - t0)
        del out
        t0 = time.perf_counter()
        out = batched()
        ctx.synchronize()
        tb.append(time.perf_counter() - t0)
        del out
    total = sum(1 << n for n in sizes)
    t_p, t_b = statistics.median(tp), statistics.median(tb)
    return {"workload": "64 polys, log2 sizes U{12..22} seed 0, rate 1/2", "total_evals": total,
            "plain_ms": 1e3 * t_p, "batched_ms": 1e3 * t_b, "plain_elems_per_s": total / t_p,
            "batched_elems_per_s": total / t_b, "speedup": t_p / t_b, "proofs_identical": True,
            "batched_position": mix_position([(n, log2_c) for n in sizes], t_b, rates),
            "plain_position": mix_position([(n, log2_c) for n in sizes], t_p, rates),
            "tuning": tuning}


def reference_shapes(bfd, ctx, torch, reps: int) -> list:
    """The reference's own criterion benchmarks (poly-commit/benches/poly-commit.rs:95-96,
    147-158): commit, then open with a fresh GenericChallenger::keccak256(), of a
    2^log_n x 256 row-major matrix (one polynomial per column) at rate 1/2, log_n 16..19.
    Random device-resident data (the benchmark draws DenseMatrix::rand); medians of `reps`."""
    out = []
    w, log2_c = 256, 1
    gen = torch.Generator(device="cuda:0")
    gen.manual_seed(0)
    for log_n in range(16, 20):
        N = 1 << log_n
        data = torch.randint(-(2 ** 63), 2 ** 63 - 1, (N * w * 2,), dtype=torch.int64, device="cuda:0",
                             generator=gen)
        point = bfd.random_vec(log_n, log_n)
        pcs = bfd.Basefold.setup(bfd.BasefoldConfig(bfd.BinaryReedSolomonCode(LAMBDA, log2_c, log_n)), ctx)
        torch.cuda.synchronize()
        tc, to = [], []
        for _ in range(reps + 1):
            ctx.synchronize()
            t0 = time.perf_counter()
            cd = pcs.commit_device(data.data_ptr(), N, w)
            tc.append(time.perf_counter() - t0)
            values = pcs.eval(cd, point)
            ctx.synchronize()
            t0 = time.perf_counter()
            pcs.open(cd, [bfd.PolyEvals(values, point)], bfd.GenericChallenger.keccak256())
            to.append(time.perf_counter() - t0)
            del cd
        c, o = statistics.median(tc[1:]), statistics.median(to[1:])
        out.append({"log_n": log_n, "log_polys": 8, "rate": "1/2", "commit_ms": 1e3 * c, "open_ms": 1e3 * o,
                    "elems_per_s": N * w / (c + o)})
        del data
    return out


def verification(bfd, ctx, torch, count: int) -> dict:
    """Basefold::verify of `count` proofs of 2^16 evaluations at rate 1/2 (the config-1 shape,
    309 queries each) in one device call; cpu_baseline times the restatement's verifier on one."""
    n, log2_c = 16, 1
    evals = [bfd.random_vec(3000 + i, 1 << n) for i in range(count)]
    points = [bfd.random_vec(4000 + i, n) for i in range(count)]
    pcs = bfd.Basefold.setup(bfd.BasefoldConfig(bfd.BinaryReedSolomonCode(LAMBDA, log2_c, n)), ctx)
    roots, claims, proofs = [], [], []
    for e, pt in zip(evals, points):
        cd = pcs.commit(e)
        cl = pcs.eval(cd, pt)
        ch = bfd.GenericChallenger.keccak256()
        ch.observe(cd.comm)
        ch.observe_ext_slice(cl)
        proofs.append(pcs.open(cd, [bfd.PolyEvals(cl, pt)], ch))
        roots.append(cd.comm)
        claims.append(cl)
        del cd

    def challengers():
        out = []
        for r, cl in zip(roots, claims):
            ch = bfd.GenericChallenger.keccak256()
            ch.observe(r)
            ch.observe_ext_slice(cl)
            out.append(ch)
        return out

    evs = [bfd.PolyEvals(cl, pt) for cl, pt in zip(claims, points)]
    assert pcs.verify_batch(roots, evs, proofs, challengers()) == [0] * count
    ts = []
    for _ in range(3):
        chs = challengers()
        ctx.synchronize()
        t0 = time.perf_counter()
        res = pcs.verify_batch(roots, evs, proofs, chs)
        ts.append(time.perf_counter() - t0)
    assert res == [0] * count
    t = statistics.median(ts)
    return {"workload": f"{count} proofs of 2^16 evaluations, rate 1/2, 309 queries each", "batch_ms": 1e3 * t,
            "proofs_per_s": count / t}


def serving_trace(bfd, ctx, torch, n_req: int, rate: float, budget: int, rates_k=None) -> dict:
    """Config 5: Poisson arrivals (mean `rate` req/s, numpy seed 0) of
    commit-then-open requests, log2 size uniform over {14..24}, rate 1/2 or 1/4
    chosen 50/50, evaluations from a seeded pool resident in HBM (2 per size),
    random points; the server admits by HBM budget and batches continuously.
    Latency = submit to proof ready."""
    import threading
    rng = np.random.default_rng(0)
    gaps = rng.exponential(1.0 / rate, size=n_req)
    sizes = rng.integers(14, 25, size=n_req)
    rates = rng.choice([1, 2], size=n_req)
    pool = {}
    for n in range(14, 25):
        for j in range(2):
            e = bfd.random_vec(100 * n + j, 1 << n)
            pool[(n, j)] = torch.from_numpy(e.view(np.int64).reshape(-1)).to("cuda:0")
    points = [bfd.random_vec(1_000_000 + i, int(n)) for i, n in enumerate(sizes)]
    torch.cuda.synchronize()
    srv = bfd.Server(hbm_budget=budget, ctx=ctx)
    # warm-up: one request per size builds every code table
    warm = [srv.submit(pool[(n, 0)].data_ptr(), n, bfd.random_vec(7, n), c) for n in range(14, 25) for c in (1, 2)]
    for t in warm:
        srv.wait(t)
    tickets = [None] * n_req
    t0 = time.perf_counter()

    def client():
        at = 0.0
        for i in range(n_req):
            at += gaps[i]
            d = at - (time.perf_counter() - t0)
            if d > 0:
                time.sleep(d)
            n = int(sizes[i])
            tickets[i] = srv.submit(pool[(n, i % 2)].data_ptr(), n, points[i], int(rates[i]))

    # served proofs kept for the check after the clock: every 16th request and
    # the first eight of the largest size
    largest = [i for i in range(n_req) if int(sizes[i]) == int(sizes.max())][:8]
    check = sorted(set(range(0, n_req, 16)) | set(largest))
    kept = {}
    th = threading.Thread(target=client)
    th.start()
    lat = []
    peak_resident = 0
    for i in range(n_req):
        while tickets[i] is None:
            time.sleep(0.0005)
        root, claim, proof, ms = srv.wait(tickets[i])
        lat.append(ms)
        if i in check:
            kept[i] = (root, claim, proof)
        del proof
        peak_resident = max(peak_resident, srv.stats()["resident_bytes"])
    th.join()
    makespan = time.perf_counter() - t0
    timing = srv.timing()
    srv.close()
    total = int(sum(1 << int(n) for n in sizes))
    lat = np.array(lat)
    out = {"workload": f"{n_req} requests, Poisson {rate:g}/s, log2 n U{{14..24}}, rate 1/2|1/4 50/50",
           "budget_bytes": budget, "latency_samples": int(len(lat)),
           "p50_ms": float(np.percentile(lat, 50)), "p99_ms": float(np.percentile(lat, 99)),
           "p99_is_worst_of": int(round(0.01 * len(lat))), "mean_ms": float(lat.mean()),
           "elems_per_s": total / makespan, "makespan_s": makespan,
           "arrival_span_s": float(gaps.sum()), "requests_per_s_served": n_req / makespan,
           "position": mix_position([(int(n), int(c)) for n, c in zip(sizes, rates)], makespan, rates_k),
           "peak_resident_sampled_bytes": peak_resident, "server_timing": timing}
    out["served_proofs_checked"] = check_served(bfd, ctx, kept, sizes, rates, points, pool)
    return out


def check_served(bfd, ctx, kept, sizes, rates, points, pool) -> dict:
    """After the clock: the kept served proofs are verified on the device
    (Basefold::verify with the harness transcript, grouped by code), and the
    first four of at most 2^22 evaluations are compared byte for byte with the
    plain path's proof of the same polynomial."""
    by_code = {}
    for i in kept:
        by_code.setdefault((int(sizes[i]), int(rates[i])), []).append(i)
    verified, rejected = 0, []
    for (n, c), idx in sorted(by_code.items()):
        pcs = bfd.Basefold.setup(bfd.BasefoldConfig(bfd.BinaryReedSolomonCode(LAMBDA, c, n)), ctx)
        chs = []
        for i in idx:
            ch = bfd.GenericChallenger.keccak256()
            ch.observe(kept[i][0])
            ch.observe_ext_slice(kept[i][1])
            chs.append(ch)
        res = pcs.verify_batch([kept[i][0] for i in idx], [bfd.PolyEvals(kept[i][1], points[i]) for i in idx],
                               [kept[i][2] for i in idx], chs)
        verified += sum(1 for r in res if r == 0)
        rejected += [(i, r) for i, r in zip(idx, res) if r != 0]
        del pcs
    plain = 0
    for i in sorted(kept):
        n, c = int(sizes[i]), int(rates[i])
        if n > 22 or plain >= 4:
            continue
        pcs = bfd.Basefold.setup(bfd.BasefoldConfig(bfd.BinaryReedSolomonCode(LAMBDA, c, n)), ctx)
        cd = pcs.commit_device(pool[(n, i % 2)].data_ptr(), 1 << n)
        claim = pcs.eval(cd, points[i])
        ch = bfd.GenericChallenger.keccak256()
        ch.observe(cd.comm)
        ch.observe_ext_slice(claim)
        if pcs.open(cd, [bfd.PolyEvals(claim, points[i])], ch) != kept[i][2] or cd.comm != kept[i][0]:
            raise SystemExit(f"served proof {i} differs from the plain path's")
        plain += 1
        del cd, pcs
    if rejected:
        raise SystemExit(f"served proofs rejected by the verifier: {rejected[:8]}")
    return {"verified": verified, "byte_equal_to_plain": plain, "indices": "every 16th request and the first 8 "
            "of the largest size"}


def main() -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--log-n", type=int, default=LOG_N)
    ap.add_argument("--cpu-log-n", type=int, default=20)
    ap.add_argument("--cpu-samples", type=int, default=10, help="runs at 2^cpu-log-n (median)")
    ap.add_argument("--cpu-full-samples", type=int, default=3,
                    help="CPU runs at the benchmark's own size (median is cpu_baseline.value; 0: skip)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-batch", action="store_true", help="skip the batched-vs-plain figure")
    ap.add_argument("--no-ref-shapes", action="store_true",
                    help="skip the reference benchmark's own shapes (2^16..2^19 rows x 256 columns)")
    ap.add_argument("--serve-requests", type=int, default=2000, help="serving trace length (0: skip)")
    ap.add_argument("--serve-rate", type=float, default=200.0)
    ap.add_argument("--serve-sustain-rate", type=float, default=120.0,
                    help="second serving trace below the saturation rate (0: skip)")
    ap.add_argument("--serve-budget-gb", type=float, default=200.0)
    ap.add_argument("--batch-reps", type=int, default=2)
    ap.add_argument("--verify-proofs", type=int, default=64,
                    help="proofs in the batched-verification figure (0: skip)")
    ap.add_argument("--copy-ref", action="store_true",
                    help="run one k_copy16 over 2^24 elements first (HBM byte-counter calibration)")
    ap.add_argument("--no-tune", action="store_true",
                    help="keep the built-in launch shapes (default: autotune before timing)")
    ap.add_argument("--tune-reps", type=int, default=3)
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only: the library has no multi-GPU path "
              "(proofs are independent, so batches and requests shard by polynomial, one context per GPU)",
              file=sys.stderr)
        sys.exit(2)

    import torch
    from __graft_entry__ import load_package
    bfd = load_package()
    # Python's cyclic collector: with torch loaded a full collection walks
    # ~170k objects and pauses the calling thread ~50 ms, which lands inside a
    # timed step now and then.  The library itself runs no Python; objects
    # made so far are frozen out of the collector and collection stays off
    # for the measurements (reference counting still frees everything here).
    gc.collect()
    gc.freeze()
    gc.disable()

    n = args.log_n
    N = 1 << n
    ctx = bfd.Context(0)
    if args.copy_ref:
        ctx.microbench(3, 1 << 24, 1)
    pcs = bfd.Basefold.setup(bfd.BasefoldConfig(bfd.BinaryReedSolomonCode(LAMBDA, LOG2_C, n)), ctx)
    evals = bfd.random_vec(0, N)
    point = bfd.random_vec(1, n)
    dev = torch.from_numpy(evals.view(np.int64).reshape(-1)).to("cuda:0")
    torch.cuda.synchronize()
    ptr = dev.data_ptr()
    cd0 = pcs.commit_device(ptr, N)
    claim = pcs.eval(cd0, point)
    root0 = cd0.comm
    del cd0
    # launch shapes: autotuned for this size before anything is timed (the
    # tuner's own runs are not part of the timed steps)
    tuning = None
    if not args.no_tune:
        t_tune = time.perf_counter()
        chosen, ms_def, ms_tuned = ctx.tune(pcs, reps=args.tune_reps)
        tuning = {"shape": chosen.as_dict(), "default_shape": bfd.Tuning.default().as_dict(),
                  "commit_open_ms_default": ms_def, "commit_open_ms_tuned": ms_tuned,
                  "tune_s": time.perf_counter() - t_tune}

    phases = []  # (commit, transcript, open) wall times of the timed steps

    def step():
        ta = time.perf_counter()
        cd = pcs.commit_device(ptr, N)
        tb = time.perf_counter()
        ch = bfd.GenericChallenger.keccak256()
        ch.observe(cd.comm)
        ch.observe_ext_slice(claim)
        tc = time.perf_counter()
        proof = pcs.open(cd, [bfd.PolyEvals(claim, point)], ch)
        phases.append((tb - ta, tc - tb, time.perf_counter() - tc))
        return cd, proof

    for _ in range(args.warmup):
        cd, proof = step()
        assert cd.comm == root0
        del cd
    # one untimed step with every launch bracketed by events: the per-kernel
    # breakdown, and which kernel dominates
    ctx.synchronize()
    ctx.kernel_times(enable=True)
    cd, proof = step()
    del cd
    ctx.synchronize()
    kall = ctx.kernel_times(enable=False)
    name = max((kv for kv in kall.items() if not kv[0].startswith("host:")), key=lambda kv: kv[1][0])[0]
    # the timed steps time only the dominant kernel: one pair of HIP events
    # on the library's stream brackets each run of its consecutive launches
    # (a tree's single-level launches), as every event costs the stream about
    # 5 us; the average launch is the bracketed time over the launches in it
    ctx.kernel_filter(name)
    ctx.synchronize()
    torch.cuda.synchronize()
    ctx.kernel_times(enable=True)
    phases.clear()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        cd, proof = step()
        del cd
    ctx.synchronize()
    torch.cuda.synchronize()
    t = time.perf_counter() - t0
    ktimes = ctx.kernel_times(enable=False)
    ctx.kernel_filter(None)
    ms_per_step = 1e3 * t / args.steps

    # dominant kernel family and its roofline position
    if f"untimed:{name}" in ktimes:
        raise SystemExit(f"{ktimes[f'untimed:{name}'][1]} launches of {name} ended a run with no event behind them")
    kms, launches = ktimes[name]
    M = N << LOG2_C
    c = 1 << LOG2_C
    # Merkle levels whose input layer holds more than SUBTREE_FROM digests run
    # as one k_nodes launch each; the levels above them run in k_subtree (the
    # library's launch plan for one tree).  Trees per step: the main tree over
    # M leaves and the pi tree of every round k over (M >> k) / 2 leaves.
    trees = [M] + [(M >> k) // 2 for k in range(n)]
    sub_from = 1 << ctx.tuning(n + LOG2_C).subtree_from_log2
    node_launches = sum(max(0, t.bit_length() - 1 - (sub_from.bit_length() - 1)) for t in trees)
    per_step_units = {
        "k_nodes": sum(single_level_nodes(t, sub_from) for t in trees),
        "k_leaves_elem": M,
        "k_leaves_pair": sum((M >> k) // 2 for k in range(n)),
    }
    if "k_nodes" in ktimes and ktimes["k_nodes"][1] != node_launches * args.steps:
        raise SystemExit(f"k_nodes launches {ktimes['k_nodes'][1]} differ from the plan's {node_launches * args.steps}")
    bytes_per_unit = {"k_nodes": 96, "k_leaves_elem": 48, "k_leaves_pair": 64}
    if name in per_step_units:
        alg_bytes = per_step_units[name] * bytes_per_unit[name] * args.steps
    elif name == "k_ntt_layer":
        alg_bytes = (n - 1) * 2 * 16 * M * args.steps  # each layer reads and writes the codeword
    else:
        alg_bytes = None
    achieved = alg_bytes / (kms * 1e-3) / 1e9 if alg_bytes else None
    # HBM traffic per launch: algorithmic bytes per launch times the measured
    # (FETCH_SIZE + WRITE_SIZE) / algorithmic ratio of the committed rocprof
    # summary for this kernel, when one exists
    traffic = None
    ratio = traffic_ratio(name)
    if alg_bytes and ratio:
        traffic = ratio * alg_bytes / launches
    ab = algorithmic_bytes(n, LOG2_C)
    # integer-VALU position of the dominant Keccak kernel: permutations per
    # second inside it against an isolated, memory-free Keccak-f kernel
    perms_per_step = {"k_nodes": per_step_units["k_nodes"], "k_leaves_elem": M,
                      "k_leaves_pair": per_step_units["k_leaves_pair"]}.get(name)
    # isolated rates at launch sizes past the point where they stop rising
    # (a smaller launch pays its first and last waves' ramp: 9.1 G perm/s at
    # 2^20 chains, 10.1 at 2^22, 10.6 at 2^23; 90, 92, 94 G mul/s at 2^22,
    # 2^23, 2^24 products), the best of three back-to-back runs (the first
    # of a series runs ~4% slower than the ones after it)
    def best_rate(what, n):
        ctx.microbench(what, n, 2)
        return max(n * 64 / (ctx.microbench(what, n, 64) * 1e-3) for _ in range(3))

    r_k = best_rate(2, 1 << 23)
    r_g = best_rate(1, 1 << 24)
    r_t = best_rate(4, 1 << 22)
    # single-wave latency of one lane-pair Keccak-f (the transcript step's
    # form): the difference of two chains on one wave
    ctx.microbench(5, 64, 4)
    tau_k = (ctx.microbench(5, 64, 1100) - ctx.microbench(5, 64, 100)) * 1e-3 / 1000
    rates = (r_k, r_g, r_t, tau_k)
    valu = None
    if perms_per_step:
        rate = perms_per_step * args.steps / (kms * 1e-3)
        valu = {"keccak_perm_per_s": rate, "isolated_keccak_perm_per_s": r_k, "frac": rate / r_k}
        ipn = valu_per_node(name) if name == "k_nodes" else None
        if ipn:
            # wave instructions per second against the nominal issue rate
            valu["instr_per_node"] = ipn
            valu["issue_frac_nominal"] = rate * ipn / 64 / VALU_ISSUE_PEAK
    # whole-step VALU position: time the step's Keccak-f permutations and
    # GF(2^128) multiplications would take at the isolated kernels' rates,
    # over the measured step time
    K, Gr, Gt = valu_work(n, LOG2_C)
    t_step = t / args.steps
    valu_step = {"keccak_perms": K, "gf_muls": Gr + Gt, "gf_muls_by_table": Gt, "isolated_keccak_perm_per_s": r_k,
                 "isolated_gf_mul_per_s": r_g, "isolated_gf_mul_by_table_per_s": r_t,
                 "frac": valu_floor_s(K, Gr, Gt, rates) / t_step}

    result = {
        "metric": "Basefold commit+open field-elems/s (2^24 evals, rate 1/4)",
        "value": N / (t / args.steps),
        "unit": "elems/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "gf2_128",
        "data": "synthetic (ChaCha12 StdRng seed 0 evaluations, seed 1 point)",
        "config": {"workload": "basefold_commit_open_plain", "log_n": n, "log2_c": LOG2_C,
                   "lambda": LAMBDA, "num_queries": pcs.num_queries, "width": 1},
        "hbm_gbs_algorithmic": ab["total"] / (t / args.steps) / 1e9,
        "roofline": {"bound": "hbm", "kernel": name, "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": (achieved / HBM_PEAK_GBS) if achieved else None,
                     "traffic": traffic, "avg_launch_ms": kms / launches, "launches_per_step": launches / args.steps},
        "valu": valu,
        "valu_step": valu_step,
        "kernel_ms_per_step": {k: v[0] for k, v in sorted(kall.items())},
        "step_phases_ms": {k: [round(1e3 * p[i], 3) for p in phases]
                           for i, k in enumerate(("commit", "transcript", "open"))},
        "proof_bytes": len(proof),
        "tuning": tuning,
    }
    stage(f"main workload done: {ms_per_step:.2f} ms per step")
    cd = pcs.commit_device(ptr, N)
    result["stepwise_open"] = stepwise_open(bfd, pcs, cd, claim, point, 3)
    del cd
    if not args.no_ref_shapes:
        stage("reference shapes")
        result["reference_shapes"] = reference_shapes(bfd, ctx, torch, 3)
    stage("sumchecks")
    result["product_sumcheck"] = product_sumcheck(bfd, ctx, torch, 20, rates)
    result["quadratic_sumcheck"] = quadratic_sumcheck(bfd, ctx, torch, 20, rates)
    if not args.no_batch:
        stage("batched vs plain")
        result["batched"] = batched_vs_plain(bfd, ctx, torch, args.batch_reps,
                                             0 if args.no_tune else args.tune_reps, rates)
    if args.verify_proofs > 0:
        stage("verification")
        result["verify"] = verification(bfd, ctx, torch, args.verify_proofs)
    if args.serve_requests > 0:
        stage("serving")
        result["serving"] = serving_trace(bfd, ctx, torch, args.serve_requests, args.serve_rate,
                                          int(args.serve_budget_gb * 1e9), rates)
        # at 200 req/s the offered work exceeds what one GPU serves, so
        # latency grows with the trace; `requests_per_s_served` there is the
        # saturation throughput, and this trace shows latency below it
        if args.serve_sustain_rate > 0:
            stage("serving below saturation")
            result["serving_sustainable"] = serving_trace(bfd, ctx, torch, args.serve_requests,
                                                          args.serve_sustain_rate, int(args.serve_budget_gb * 1e9),
                                                          rates)
    if not args.no_cpu:
        stage("cpu baseline")
        result["cpu_baseline"] = cpu_baseline(args, proof if args.log_n == LOG_N else None)
    print(json.dumps(result))


if __name__ == "__main__":
    main()
