"""GPU parity: the HIP library against the CPU restatement (oracle/).

Bit-exact on every byte of the commitment root, claim and bincode proof at
sizes the restatement finishes in seconds; at the full 2^24 size the proof
is checked through verifier acceptance and claim agreement.  Mirrors
poly-commit/src/basefold.rs:480-490 (log2_c 0..2, num_vars 0..9) and the
harness transcript of poly-commit/src/lib.rs:64-105.
"""
import os
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def ctx(bfd):
    return bfd.default_context()


def gpu_prove(bfd, ctx, lam, log2_c, evals, point):
    n = len(point)
    pcs = bfd.Basefold.setup(bfd.BasefoldConfig(bfd.BinaryReedSolomonCode(lam, log2_c, n)), ctx)
    cd = pcs.commit(evals)
    claim = pcs.eval(cd, point)
    ch = bfd.GenericChallenger.keccak256()
    ch.observe(cd.comm)
    ch.observe_ext_slice(claim)
    proof = pcs.open(cd, [bfd.PolyEvals(claim, point)], ch)
    return cd.comm, claim, proof


@pytest.mark.parametrize("log2_c", [0, 1, 2])
def test_basefold_small_bit_exact(bfd, oracle, ctx, log2_c):
    for n in range(0, 10):
        _say(f"case log2_c={log2_c} n={n}")
        evals, point = bfd.random_vec(100 + n, 1 << n), bfd.random_vec(200 + n, n)
        root, claim, proof = gpu_prove(bfd, ctx, 128, log2_c, evals, point)
        root_o, claim_o, proof_o = oracle.basefold_prove(128, log2_c, evals, point)
        assert root == root_o, (log2_c, n)
        assert (claim == claim_o).all(), (log2_c, n)
        assert proof == proof_o, (log2_c, n)
        assert oracle.basefold_verify(128, log2_c, root, point, claim, proof) == 0


@pytest.mark.parametrize("lam,log2_c,n", [(128, 3, 6), (128, 3, 11), (100, 4, 5), (128, 8, 2), (80, 1, 7),
                                          (128, 5, 9)])
def test_basefold_other_rates_and_security(bfd, oracle, ctx, lam, log2_c, n):
    """Blowups 8..256 and lambda 80/100 (num_queries, binary_reed_solomon.rs:55-59, differs from 128):
    the layer-0 slots, twiddle levels and query counts of these codes against the restatement."""
    _say(f"case lam={lam} log2_c={log2_c} n={n}")
    evals, point = bfd.random_vec(300 + n, 1 << n), bfd.random_vec(400 + n, n)
    root, claim, proof = gpu_prove(bfd, ctx, lam, log2_c, evals, point)
    root_o, claim_o, proof_o = oracle.basefold_prove(lam, log2_c, evals, point)
    assert root == root_o and (claim == claim_o).all() and proof == proof_o
    assert len(proof) == bfd.lib().bfd_proof_size(lam, log2_c, n)
    assert oracle.basefold_verify(lam, log2_c, root, point, claim, proof) == 0


def test_basefold_cfg1_bit_exact(bfd, oracle, ctx):
    """2^16 evaluations, rate 1/2, seed 0: proof bytes identical (1,645,080 B)."""
    n, log2_c = 16, 1
    evals, point = bfd.random_vec(0, 1 << n), bfd.random_vec(1, n)
    root, claim, proof = gpu_prove(bfd, ctx, 128, log2_c, evals, point)
    root_o, claim_o, proof_o = oracle.basefold_prove(128, log2_c, evals, point)
    assert len(proof) == 1645080
    assert root == root_o and (claim == claim_o).all() and proof == proof_o


@pytest.mark.parametrize("n,log2_c", [(18, 2), (19, 1)])
def test_basefold_medium_bit_exact(bfd, oracle, ctx, n, log2_c):
    evals, point = bfd.random_vec(7, 1 << n), bfd.random_vec(8, n)
    root, claim, proof = gpu_prove(bfd, ctx, 128, log2_c, evals, point)
    root_o, claim_o, proof_o = oracle.basefold_prove(128, log2_c, evals, point)
    assert root == root_o and (claim == claim_o).all() and proof == proof_o


def gpu_prove_w(bfd, ctx, lam, log2_c, evals, width, point):
    n = len(point)
    pcs = bfd.Basefold.setup(bfd.BasefoldConfig(bfd.BinaryReedSolomonCode(lam, log2_c, n)), ctx)
    cd = pcs.commit(evals, width)
    claims = pcs.eval(cd, point)
    ch = bfd.GenericChallenger.keccak256()
    ch.observe(cd.comm)
    ch.observe_ext_slice(claims)
    proof = pcs.open(cd, [bfd.PolyEvals(claims, point)], ch)
    return cd.comm, claims, proof


def test_width_grid_bit_exact(bfd, oracle, ctx):
    """The reference harness grid (poly-commit/src/lib.rs:73-77: num_vars 0..9 x num_polys 1..9),
    rate 1/2: root, the width claims and the proof bytes equal the restatement's."""
    for n in range(0, 10):
        for w in range(1, 10):
            _say(f"case width n={n} w={w}")
            evals, point = bfd.random_vec(1000 + 10 * n + w, (1 << n) * w), bfd.random_vec(2000 + n, n)
            root, claims, proof = gpu_prove_w(bfd, ctx, 128, 1, evals, w, point)
            root_o, claims_o, proof_o = oracle.basefold_prove_w(128, 1, evals, w, point)
            assert root == root_o, (n, w)
            assert (claims == claims_o).all(), (n, w)
            assert proof == proof_o, (n, w)
            assert len(proof) == bfd.proof_size(128, 1, n, w)


@pytest.mark.parametrize("n,log2_c,w", [(16, 1, 256), (12, 2, 33)])
def test_width_reference_bench_shape_bit_exact(bfd, oracle, ctx, n, log2_c, w):
    """poly-commit/benches/poly-commit.rs:95-96 commits 2^16..2^19 rows x 256 columns at rate 1/2;
    the smallest of those shapes (and an odd width at rate 1/4) against the restatement."""
    _say(f"case width n={n} log2_c={log2_c} w={w}")
    evals, point = bfd.random_vec(31 + w, (1 << n) * w), bfd.random_vec(32, n)
    root, claims, proof = gpu_prove_w(bfd, ctx, 128, log2_c, evals, w, point)
    root_o, claims_o, proof_o = oracle.basefold_prove_w(128, log2_c, evals, w, point)
    assert root == root_o and (claims == claims_o).all() and proof == proof_o
    assert oracle.basefold_verify_w(128, log2_c, w, root, point, claims, proof) == 0


def test_width_commit_device_and_reopen(bfd, oracle, ctx):
    """Device-resident width-w matrix: the same root and proof as from host memory, twice."""
    import torch as torch_mod
    n, log2_c, w = 10, 1, 6
    evals, point = bfd.random_vec(77, (1 << n) * w), bfd.random_vec(78, n)
    dev = torch_mod.from_numpy(evals.view(np.int64).reshape(-1)).to("cuda:0")
    torch_mod.cuda.synchronize()
    pcs = bfd.Basefold.setup(bfd.BasefoldConfig(bfd.BinaryReedSolomonCode(128, log2_c, n)), ctx)
    cd = pcs.commit_device(dev.data_ptr(), 1 << n, w)
    claims = pcs.eval(cd, point)
    root_o, claims_o, proof_o = oracle.basefold_prove_w(128, log2_c, evals, w, point)
    assert cd.comm == root_o and (claims == claims_o).all()
    for _ in range(2):
        ch = bfd.GenericChallenger.keccak256()
        ch.observe(cd.comm)
        ch.observe_ext_slice(claims)
        assert pcs.open(cd, [bfd.PolyEvals(claims, point)], ch) == proof_o


def quad_case(bfd, n, seed):
    """quadratic.rs:150-171's shape: 5..9 tables, 5..9 pairs of random indices with random scalars."""
    rng = np.random.default_rng(seed)
    m, npairs = int(rng.integers(5, 10)), int(rng.integers(5, 10))
    polys = bfd.random_vec(seed, m << n).reshape(m, 1 << n, 2)
    scalars = bfd.random_vec(seed + 1, npairs)
    pairs = [(scalars[k], int(rng.integers(0, m)), int(rng.integers(0, m))) for k in range(npairs)]
    return bfd.Quadratic(polys, pairs)


def test_quadratic_sumcheck_bit_exact(bfd, oracle, ctx):
    """Quadratic sumcheck (quadratic.rs + prove_sumcheck): every output equals the restatement's."""
    for n in list(range(0, 11)) + [14]:
        _say(f"case quadratic n={n}")
        f = quad_case(bfd, n, 900 + n)
        got = bfd.prove_sumcheck_quadratic(f, bfd.GenericChallenger.keccak256(), ctx)
        want = oracle.quadratic_sumcheck(f.polys, [p[0] for p in f.pairs], [(p[1], p[2]) for p in f.pairs])
        for g, w in zip(got, want):
            assert (np.asarray(g).reshape(-1, 2) == np.asarray(w).reshape(-1, 2)).all(), n


def test_quadratic_sumcheck_many_pairs(bfd, oracle, ctx):
    """More pairs than the single-block rounds take (every round runs on the
    multi-block path): 130, 300 and 1000 pairs, which give 16, 8 and 4
    accumulation slots per pair and more pairs than one wave's lanes; and
    2^18 tables, whose first fixes multiply through the challenge's window
    table."""
    for n, m, npairs in [(6, 3, 130), (5, 6, 300), (3, 4, 1000), (18, 3, 3)]:
        _say(f"case quadratic n={n} m={m} pairs={npairs}")
        rng = np.random.default_rng(n * 1000 + npairs)
        polys = bfd.random_vec(n + npairs, m << n).reshape(m, 1 << n, 2)
        scalars = bfd.random_vec(n + npairs + 1, npairs)
        f = bfd.Quadratic(polys, [(scalars[k], int(rng.integers(0, m)), int(rng.integers(0, m)))
                                  for k in range(npairs)])
        got = bfd.prove_sumcheck_quadratic(f, bfd.GenericChallenger.keccak256(), ctx)
        want = oracle.quadratic_sumcheck(f.polys, [p[0] for p in f.pairs], [(p[1], p[2]) for p in f.pairs])
        for g, w in zip(got, want):
            assert (np.asarray(g).reshape(-1, 2) == np.asarray(w).reshape(-1, 2)).all(), (n, npairs)


def test_quadratic_prover_stepwise(bfd, oracle, ctx):
    """The step-wise QuadraticProver (bfd_sc_*) under a host prove_sumcheck driver
    gives the restatement's claim, round polys, challenges, evaluations and
    subclaim; evaluations() is None until every variable is fixed, and the
    sum at the final point equals the subclaim."""
    for n in (0, 1, 2, 9, 13):
        _say(f"case stepwise quadratic n={n}")
        f = quad_case(bfd, n, 970 + n)
        want = oracle.quadratic_sumcheck(f.polys, [p[0] for p in f.pairs], [(p[1], p[2]) for p in f.pairs])
        prover = bfd.QuadraticProver(f, ctx)
        assert prover.num_vars == n
        if n:
            assert prover.evaluations() is None
        claim = prover.compute_sum()
        ch = bfd.GenericChallenger.keccak256()
        sub, polys, chal = bfd.prove_sumcheck(prover, claim, ch)
        got = (claim, np.array(polys).reshape(-1, 2), np.array(chal).reshape(-1, 2), prover.evaluations(), sub)
        for g, w in zip(got, want):
            assert (np.asarray(g).reshape(-1, 2) == np.asarray(w).reshape(-1, 2)).all(), n
        assert (prover.compute_sum() == sub).all(), n
        ch2 = bfd.GenericChallenger.keccak256()
        bfd.prove_sumcheck_quadratic(f, ch2, ctx)
        assert ch.sample_bits(64) == ch2.sample_bits(64), n


def test_batch_quadratic_bit_exact(bfd, oracle, ctx):
    """Batch of three equal-size Quadratics (batch.rs) as one merged call, against the restatement."""
    n = 9
    fs = [quad_case(bfd, n, 950 + k) for k in range(3)]
    merged = bfd.batch_quadratics(fs, bfd.random_vec(960, 1))
    got = bfd.prove_sumcheck_quadratic(merged, bfd.GenericChallenger.keccak256(), ctx)
    want = oracle.quadratic_sumcheck(merged.polys, [p[0] for p in merged.pairs],
                                     [(p[1], p[2]) for p in merged.pairs])
    for g, w in zip(got, want):
        assert (np.asarray(g).reshape(-1, 2) == np.asarray(w).reshape(-1, 2)).all()


def test_quadratic_sumcheck_large_invariants(bfd, oracle, ctx):
    """2^22 tables: the verifier's final identity with independently evaluated finals
    (sum_k s_k P_a(r) P_b(r) = subclaim) and the transcript replay of every challenge."""
    n = 22
    _say(f"case quadratic n={n}")
    f = quad_case(bfd, n, 970)
    claim, rps, chal, fin, sub = bfd.prove_sumcheck_quadratic(f, bfd.GenericChallenger.keccak256(), ctx)
    ch = oracle.Challenger()
    ch.observe(claim.tobytes())
    for k in range(n):
        ch.observe(rps[k].tobytes())
        assert (ch.sample() == chal[k]).all()
    x = chal[::-1].copy()
    for p in [0, len(f.polys) - 1]:
        assert (oracle.mle_evaluate(f.polys[p], x) == fin[p:p + 1]).all()
    expect = np.zeros((1, 2), np.uint64)
    for (s, a, b) in f.pairs:
        expect ^= oracle.mul(oracle.mul(fin[a:a + 1], fin[b:b + 1]), np.asarray(s).reshape(1, 2))
    assert (sub == expect).all()


def test_quadratic_sumcheck_prefixes_and_device_inputs(bfd, oracle, ctx):
    """The Quadratic sumcheck's device transcript starts from any challenger
    state (caller prefixes of 5, 32 and 200 bytes, a pending sample): its
    challenges replay on the oracle's challenger and both end in the same
    state; tables in device memory give the host path's outputs and stay
    unmodified."""
    import torch
    # n 9 starts on the multi-block path, n 7 in the single-block rounds
    for n in (9, 7):
        _quadratic_prefixes_and_device_inputs(bfd, oracle, ctx, n, torch)


def _quadratic_prefixes_and_device_inputs(bfd, oracle, ctx, n, torch):
    f = quad_case(bfd, n, 990 + n)
    fresh = oracle.quadratic_sumcheck(f.polys, [p[0] for p in f.pairs], [(p[1], p[2]) for p in f.pairs])
    rng = np.random.default_rng(4)
    for plen, sample_after in [(5, False), (32, False), (200, False), (40, True)]:
        _say(f"quadratic n={n} prefix {plen} sample={sample_after}")
        prefix = bytes(rng.integers(0, 256, plen, dtype=np.uint8))
        ch, och = bfd.GenericChallenger.keccak256(), oracle.Challenger()
        ch.observe(prefix)
        och.observe(prefix)
        if sample_after:
            assert (ch.sample_ext_element() == och.sample()).all()
        claim, rps, chal, fin, sub = bfd.prove_sumcheck_quadratic(f, ch, ctx)
        assert (claim == np.asarray(fresh[0]).reshape(-1, 2)).all()
        och.observe(claim.tobytes())
        for k in range(n):
            och.observe(rps[k].tobytes())
            assert (och.sample() == chal[k]).all(), (plen, k)
        x = chal[::-1].copy()
        expect = np.zeros((1, 2), np.uint64)
        for (sc, a, b) in f.pairs:
            fa, fb = oracle.mle_evaluate(f.polys[a], x), oracle.mle_evaluate(f.polys[b], x)
            expect ^= oracle.mul(oracle.mul(fa, fb), np.asarray(sc).reshape(1, 2))
        assert (sub == expect).all()
        assert (ch.sample_ext_element() == och.sample()).all()
    P = np.ascontiguousarray(f.polys, dtype=np.uint64)
    dP = torch.from_numpy(P.view(np.int64).reshape(-1)).to("cuda:0")
    torch.cuda.synchronize()
    got = bfd.prove_sumcheck_quadratic(f, bfd.GenericChallenger.keccak256(), ctx, device_ptr=dP.data_ptr(),
                                       shape=(P.shape[0], P.shape[1]))
    want = bfd.prove_sumcheck_quadratic(f, bfd.GenericChallenger.keccak256(), ctx)
    for g, w in zip(got, want):
        assert (np.asarray(g) == np.asarray(w)).all()
    assert (dP.cpu().numpy() == P.view(np.int64).reshape(-1)).all()


def tamper_offsets(n, log2_c, w, q):
    """Byte offsets inside a bincode BasefoldProof (basefold.rs:49-66) for d = n rounds."""
    d, L = n, n + log2_c
    open_off = 72 * d + 48
    rec = 8 + 16 * w + 8 + 32 * L + 8 + sum(16 + 8 + 32 * (log2_c + d - 1 - k) for k in range(d))
    r0 = open_off
    return {
        "round_poly": 8 + 40 * min(1, d - 1) + 8,
        "final_poly": 8 + 40 * d + 8,
        "pi_root": 40 * d + 40,
        "row": r0 + 8,
        "main_sibling": r0 + 16 + 16 * w + 5,
        "pi_value": r0 + 24 + 16 * w + 32 * L,
        "pi_sibling": r0 + 24 + 16 * w + 32 * L + 24 + 3,
        "last_query_row": open_off + (q - 1) * rec + 8 + 16 * (w - 1),
        "query_count": open_off - 8,
    }


def verify_both(bfd, oracle, ctx, pcs, log2_c, w, root, point, claims, proof):
    ch = bfd.GenericChallenger.keccak256()
    ch.observe(root)
    ch.observe_ext_slice(claims)
    got = pcs.verify(root, [bfd.PolyEvals(claims, point)], proof, ch, w)
    want = oracle.basefold_verify_w(128, log2_c, w, root, point, claims, proof)
    return got, want


@pytest.mark.parametrize("n,log2_c,w", [(10, 1, 1), (9, 2, 3), (6, 0, 9), (1, 1, 2), (0, 1, 4)])
def test_verifier_matches_restatement(bfd, oracle, ctx, n, log2_c, w):
    """Basefold::verify on the device against the restatement's verifier: honest proofs accept,
    and every tampered part of the wire format gets the same error kind.  The query_count and
    truncation cases pin the restatement's stricter decoding (a proof must carry exactly
    num_queries openings and no trailing bytes, else malformed, 4); the reference zips the
    sampled indices with whatever openings the proof carries (basefold.rs:285) and would
    check only those."""
    _say(f"case verify n={n} log2_c={log2_c} w={w}")
    evals, point = bfd.random_vec(600 + n, (1 << n) * w), bfd.random_vec(601 + n, n)
    root, claims, proof = gpu_prove_w(bfd, ctx, 128, log2_c, evals, w, point)
    pcs = bfd.Basefold.setup(bfd.BasefoldConfig(bfd.BinaryReedSolomonCode(128, log2_c, n)), ctx)
    got, want = verify_both(bfd, oracle, ctx, pcs, log2_c, w, root, point, claims, proof)
    assert got == want == 0
    q = oracle.num_queries(128, log2_c)
    offs = tamper_offsets(n, log2_c, w, q)
    for name, off in offs.items():
        if (n == 0 and name in ("round_poly", "pi_root", "pi_value", "pi_sibling")) or q == 0:
            continue
        if n + log2_c == 0 and name == "main_sibling":
            continue
        bad = bytearray(proof)
        bad[off] ^= 0x04
        got, want = verify_both(bfd, oracle, ctx, pcs, log2_c, w, root, point, claims, bytes(bad))
        assert got == want != 0, (name, got, want)
    for cut in [1, 8, len(proof) // 2]:
        got, want = verify_both(bfd, oracle, ctx, pcs, log2_c, w, root, point, claims, proof[:-cut])
        assert got == want == 4, cut
    bad_claims = claims.copy()
    bad_claims[0, 0] ^= 2
    got, want = verify_both(bfd, oracle, ctx, pcs, log2_c, w, root, point, bad_claims, proof)
    assert got == want != 0


def test_verifier_batch(bfd, oracle, ctx):
    """Many proofs in one call: every result equals the restatement's for its proof."""
    n, log2_c, w = 8, 1, 3
    pcs = bfd.Basefold.setup(bfd.BasefoldConfig(bfd.BinaryReedSolomonCode(128, log2_c, n)), ctx)
    offs = tamper_offsets(n, log2_c, w, oracle.num_queries(128, log2_c))
    kinds = [None, "row", "final_poly", "pi_sibling", None, "last_query_row", None, "main_sibling"] * 3
    roots, evals, proofs, chs, want = [], [], [], [], []
    for k, kind in enumerate(kinds):
        e, pt = bfd.random_vec(700 + k, (1 << n) * w), bfd.random_vec(800 + k, n)
        root, claims, proof = gpu_prove_w(bfd, ctx, 128, log2_c, e, w, pt)
        if kind:
            b = bytearray(proof)
            b[offs[kind]] ^= 0x10
            proof = bytes(b)
        ch = bfd.GenericChallenger.keccak256()
        ch.observe(root)
        ch.observe_ext_slice(claims)
        roots.append(root)
        evals.append(bfd.PolyEvals(claims, pt))
        proofs.append(proof)
        chs.append(ch)
        want.append(oracle.basefold_verify_w(128, log2_c, w, root, pt, claims, proof))
    before = [c.export_state() for c in chs]
    with pytest.raises(bfd.BfdError) as e:
        pcs.verify_batch(roots[:2], evals[:2], proofs[:2], [chs[0], chs[0]], w)
    assert e.value.code == bfd.BFD_ERR_ARG
    assert [c.export_state() for c in chs] == before
    got = pcs.verify_batch(roots, evals, proofs, chs, w)
    assert got == want
    assert want.count(0) == 9


@pytest.mark.parametrize("n,log2_c,w", [(0, 1, 1), (0, 2, 5), (1, 1, 1), (13, 1, 1), (12, 2, 7), (17, 1, 3)])
def test_commit_eval_fused(bfd, oracle, ctx, n, log2_c, w):
    """commit + eval in one call (the claims from the one-pass reduction) equal the restatement's
    root and column MLEs, and bfd_eval (same reduction) agrees on the same commitment."""
    _say(f"case commit_eval n={n} log2_c={log2_c} w={w}")
    evals, point = bfd.random_vec(1300 + n + w, (1 << n) * w), bfd.random_vec(1400 + n, n)
    pcs = bfd.Basefold.setup(bfd.BasefoldConfig(bfd.BinaryReedSolomonCode(128, log2_c, n)), ctx)
    cd, vals = pcs.commit_eval(evals, point, w)
    root_o, claims_o, _ = oracle.basefold_prove_w(128, log2_c, evals, w, point)
    assert cd.comm == root_o and (vals == claims_o).all()
    assert (pcs.eval(cd, point) == claims_o).all()


def test_encode_matches_oracle_codeword_root(bfd, oracle, ctx):
    """The main-tree root pins the whole codeword (every leaf)."""
    for n, log2_c in [(5, 0), (6, 1), (7, 2), (12, 2)]:
        evals = bfd.random_vec(300 + n, 1 << n)
        pcs = bfd.Basefold.setup(bfd.BasefoldConfig(bfd.BinaryReedSolomonCode(128, log2_c, n)), ctx)
        cd = pcs.commit(evals)
        cw = oracle.rs_encode(log2_c, n, evals)
        assert cd.comm == oracle.merkle_root(cw.tobytes(), len(cw), 16)


def test_code_encode_fold_interpolate(bfd, oracle, ctx):
    """RandomFoldableCode on the device (bfd_code_encode / fold / interpolate):
    encode_batch at widths 1 and 5 (host and device buffers), fold at every
    level and interpolate at sampled positions equal the restatement's; and,
    independent of it, folding the codeword of m through levels d-1..0 while
    fixing m's last variable with the same challenges ends at c copies of the
    fixed value (encode0(final_poly) == pi, basefold.rs:198)."""
    import torch
    for log2_c, d in [(0, 4), (1, 6), (2, 9), (1, 14)]:
        _say(f"code log2_c={log2_c} d={d}")
        pcs = bfd.Basefold.setup(bfd.BasefoldConfig(bfd.BinaryReedSolomonCode(128, log2_c, d)), ctx)
        code = pcs.code
        m = bfd.random_vec(1300 + d, 1 << d)
        cw = code.encode(m)
        assert (cw == oracle.rs_encode(log2_c, d, m)).all()
        assert (code.encode0(m[:1]) == np.repeat(m[:1], 1 << log2_c, axis=0)).all()
        mw = bfd.random_vec(1400 + d, 5 << d)
        cw5 = code.encode_batch(mw, 5)
        assert (cw5 == oracle.rs_encode(log2_c, d, mw, 5)).all()
        dm = torch.from_numpy(mw.view(np.int64).reshape(-1)).to("cuda:0")
        dout = torch.zeros(2 * len(cw5), dtype=torch.int64, device="cuda:0")
        torch.cuda.synchronize()
        code.encode_batch_device(dm.data_ptr(), dout.data_ptr(), 5)
        assert (dout.cpu().numpy().view(np.uint64).reshape(-1, 2) == cw5).all()
        for i in range(d):
            w = bfd.random_vec(1500 + 31 * d + i, 2 << (log2_c + i))
            r = bfd.random_vec(1600 + 31 * d + i, 1)
            f = code.fold(i, w, r)
            assert (f == oracle.rs_fold(log2_c, d, i, w, r)).all(), (d, i)
            mid = len(w) // 2
            for j in sorted({0, 1 % mid, mid - 1, (7 * i + 3) % mid}):
                assert (code.interpolate(i, j, w[j], w[j + mid], r) == f[j]).all(), (d, i, j)
        if d <= 9:
            pi, poly = cw, m
            for i in reversed(range(d)):
                r = bfd.random_vec(1700 + 31 * d + i, 1)
                pi = code.fold(i, pi, r)
                half = len(poly) // 2
                poly = np.concatenate([poly[j:j + 1] ^ bfd._mul_host(r, poly[j + half:j + half + 1] ^ poly[j:j + 1])
                                       for j in range(half)])
            assert (pi == code.encode0(poly)).all(), d


def test_reopen_same_commitment(bfd, oracle, ctx):
    """open borrows the commitment immutably: two openings at different points."""
    n, log2_c = 8, 1
    evals = bfd.random_vec(9, 1 << n)
    pcs = bfd.Basefold.setup(bfd.BasefoldConfig(bfd.BinaryReedSolomonCode(128, log2_c, n)), ctx)
    cd = pcs.commit(evals)
    for seed in (10, 11):
        point = bfd.random_vec(seed, n)
        claim = pcs.eval(cd, point)
        ch = bfd.GenericChallenger.keccak256()
        ch.observe(cd.comm)
        ch.observe_ext_slice(claim)
        proof = pcs.open(cd, [bfd.PolyEvals(claim, point)], ch)
        _, _, proof_o = oracle.basefold_prove(128, log2_c, evals, point)
        assert proof == proof_o


def test_open_into_caller_buffer(bfd, oracle, ctx):
    """bfd_open_into: a buffer one byte short is refused before the
    challenger moves; a buffer of exactly bfd_proof_size gets the proof."""
    import ctypes
    n, log2_c = 9, 2
    evals, point = bfd.random_vec(12, 1 << n), bfd.random_vec(13, n)
    pcs = bfd.Basefold.setup(bfd.BasefoldConfig(bfd.BinaryReedSolomonCode(128, log2_c, n)), ctx)
    cd = pcs.commit(evals)
    claim = pcs.eval(cd, point)
    ch = bfd.GenericChallenger.keccak256()
    ch.observe(cd.comm)
    ch.observe_ext_slice(claim)
    before = ch.export_state()
    size = bfd.lib().bfd_proof_size(128, log2_c, n)
    p = np.ascontiguousarray(bfd._elems(point))
    v = np.ascontiguousarray(bfd._elems(claim))
    buf = (ctypes.c_uint8 * size)()
    plen = ctypes.c_size_t()
    rc = bfd.lib().bfd_open_into(ctx.h, pcs.code_h, cd.handle, bfd._p(p), bfd._p(v), ch.h, buf, size - 1,
                                 ctypes.byref(plen))
    assert rc == bfd.BFD_ERR_ARG
    assert ch.export_state() == before
    bfd._check(bfd.lib().bfd_open_into(ctx.h, pcs.code_h, cd.handle, bfd._p(p), bfd._p(v), ch.h, buf, size,
                                       ctypes.byref(plen)))
    assert plen.value == size
    _, _, proof_o = oracle.basefold_prove(128, log2_c, evals, point)
    assert bytes(buf) == proof_o


def test_product_sumcheck_bit_exact(bfd, oracle, ctx):
    # n 9, 10, 11 and 12 start the single-block tail at rounds 0, 1, 2 and 3;
    # from n 10 on the round kernels' blocks span 2^(n-8) entries of the eq
    # factor B, at n 17 two blocks share each entry; at n 21 round 1 folds
    # through the window table of its challenge
    for n in [1, 2, 3, 5, 8, 9, 10, 11, 12, 14, 17, 21]:
        _say(f"product sumcheck n={n}")
        f, g, r = bfd.random_vec(400 + n, 1 << n), bfd.random_vec(500 + n, 1 << n), bfd.random_vec(600 + n, n)
        got = bfd.prove_sumcheck_product(f, g, r, bfd.GenericChallenger.keccak256(), ctx)
        exp = oracle.product_sumcheck(f, g, r)
        for a, b in zip(got, exp):
            assert (np.asarray(a) == np.asarray(b)).all(), n


def test_product_sumcheck_full_size_invariants(bfd, oracle, ctx):
    """22 variables: every round satisfies p(0)+p(1) = claim, and the final
    subclaim equals f(x) g(x) eq(r, x) evaluated on the CPU."""
    n = 22
    f, g, r = bfd.random_vec(0, 1 << n), bfd.random_vec(1, 1 << n), bfd.random_vec(2, n)
    claim, polys, chal, fin, sub = bfd.prove_sumcheck_product(f, g, r, bfd.GenericChallenger.keccak256(), ctx)
    ch = oracle.Challenger()
    ch.observe(claim.tobytes())
    s = oracle.to_int(claim)
    one = oracle.from_canonical(oracle.from_int(1))
    for k in range(n):
        c0, c2, c3 = polys[k, 0:1], polys[k, 1:2], polys[k, 2:3]
        c1 = oracle.from_int(s) ^ c2 ^ c3
        ch.observe(c0.tobytes() + c2.tobytes() + c3.tobytes())
        x = ch.sample()
        assert (x == chal[k:k + 1]).all()
        acc = c3
        for c in (c2, c1, c0):
            acc = oracle.mul(acc, x) ^ c
        s = oracle.to_int(acc)
    assert s == oracle.to_int(sub)
    xs = chal[::-1].copy()
    fx, gx = oracle.mle_evaluate(f, xs), oracle.mle_evaluate(g, xs)
    assert (fin[0:1] == fx).all() and (fin[1:2] == gx).all()
    eq = one
    for i in range(n):
        eq = oracle.mul(eq, one ^ r[i:i + 1] ^ xs[i:i + 1])
    assert oracle.to_int(oracle.mul(oracle.mul(fx, gx), eq)) == s


def _replay_product(oracle, ch, n, claim, polys, chal):
    """Replays the product sumcheck's transcript on the oracle challenger ch
    (already in the caller's starting state): every challenge must equal the
    sampled one and every round polynomial satisfy p(0) + p(1) = subclaim.
    Returns the final subclaim as an integer."""
    ch.observe(claim.tobytes())
    s = oracle.to_int(claim)
    for k in range(n):
        c0, c2, c3 = polys[k, 0:1], polys[k, 1:2], polys[k, 2:3]
        c1 = oracle.from_int(s) ^ c2 ^ c3
        ch.observe(c0.tobytes() + c2.tobytes() + c3.tobytes())
        x = ch.sample()
        assert (x == chal[k:k + 1]).all(), k
        acc = c3
        for c in (c2, c1, c0):
            acc = oracle.mul(acc, x) ^ c
        s = oracle.to_int(acc)
    return s


def test_product_sumcheck_transcript_prefixes(bfd, oracle, ctx):
    """The product sumcheck's transcript runs on the device from any
    challenger state: after caller prefixes of 5, 32 (one digest's length),
    100 and 200 bytes (the first round then hashes two blocks), and after a
    prefix followed by a sample (unread output pending), the challenges replay
    on the oracle's challenger and both challengers end in the same state."""
    n = 6
    f, g, r = bfd.random_vec(700, 1 << n), bfd.random_vec(701, 1 << n), bfd.random_vec(702, n)
    fresh = oracle.product_sumcheck(f, g, r)
    rng = np.random.default_rng(3)
    for plen, sample_after in [(5, False), (32, False), (100, False), (200, False), (40, True)]:
        _say(f"product sumcheck prefix {plen} sample={sample_after}")
        prefix = bytes(rng.integers(0, 256, plen, dtype=np.uint8))
        ch, och = bfd.GenericChallenger.keccak256(), oracle.Challenger()
        ch.observe(prefix)
        och.observe(prefix)
        if sample_after:
            assert (ch.sample_ext_element() == och.sample()).all()
        claim, polys, chal, fin, sub = bfd.prove_sumcheck_product(f, g, r, ch, ctx)
        assert (claim == fresh[0]).all()
        s = _replay_product(oracle, och, n, claim, polys, chal)
        assert s == oracle.to_int(sub)
        assert (ch.sample_ext_element() == och.sample()).all()
        assert ch.sample_bits(20) == och.sample_bits(20)


def test_product_sumcheck_device_inputs(bfd, oracle, ctx):
    """Tables already in device memory give the same outputs as host tables,
    leave the caller's buffers unchanged, at 2^12 and at 2^22."""
    import torch
    for n in (12, 22):
        _say(f"product sumcheck device inputs n={n}")
        f, g, r = bfd.random_vec(800 + n, 1 << n), bfd.random_vec(900 + n, 1 << n), bfd.random_vec(1000 + n, n)
        df = torch.from_numpy(f.view(np.int64).reshape(-1)).to("cuda:0")
        dg = torch.from_numpy(g.view(np.int64).reshape(-1)).to("cuda:0")
        torch.cuda.synchronize()
        got = bfd.prove_sumcheck_product(None, None, r, bfd.GenericChallenger.keccak256(), ctx,
                                         device_ptrs=(df.data_ptr(), dg.data_ptr()))
        exp = bfd.prove_sumcheck_product(f, g, r, bfd.GenericChallenger.keccak256(), ctx)
        for a, b in zip(got, exp):
            assert (np.asarray(a) == np.asarray(b)).all(), n
        if n == 12:
            for a, b in zip(got, oracle.product_sumcheck(f, g, r)):
                assert (np.asarray(a) == np.asarray(b)).all()
        assert (df.cpu().numpy() == f.view(np.int64).reshape(-1)).all()
        assert (dg.cpu().numpy() == g.view(np.int64).reshape(-1)).all()


@pytest.mark.skipif(os.environ.get("BFD_SKIP_FULL") == "1", reason="full-size case disabled")
def test_basefold_full_size_verifies(bfd, oracle, ctx):
    """2^24 evaluations, rate 1/4: the claim agrees with the CPU MLE and the
    proof is accepted by the CPU verifier (a size-independent property)."""
    n, log2_c = 24, 2
    _say("full-size commit+open 2^24 rate 1/4")
    evals, point = bfd.random_vec(0, 1 << n), bfd.random_vec(1, n)
    root, claim, proof = gpu_prove(bfd, ctx, 128, log2_c, evals, point)
    assert (claim == oracle.mle_evaluate(evals, point)).all()
    assert len(proof) == 2235000
    assert oracle.basefold_verify(128, log2_c, root, point, claim, proof) == 0
    bad = bytearray(proof)
    bad[-40] ^= 1
    want = oracle.basefold_verify(128, log2_c, root, point, claim, bytes(bad))
    assert want != 0
    pcs = bfd.Basefold.setup(bfd.BasefoldConfig(bfd.BinaryReedSolomonCode(128, log2_c, n)), ctx)
    for pr, expect in [(proof, 0), (bytes(bad), want)]:
        ch = bfd.GenericChallenger.keccak256()
        ch.observe(root)
        ch.observe_ext_slice(claim)
        assert pcs.verify(root, [bfd.PolyEvals(claim, point)], pr, ch) == expect


def test_batch_equals_plain_and_oracle(bfd, oracle, ctx):
    """Uneven batch: every proof equals the plain path's and the oracle's."""
    rng = np.random.default_rng(7)
    sizes = [int(v) for v in rng.integers(1, 11, size=12)] + [12, 3, 1]
    # rate 1 (log2_c = 0) has no queries: header-only proofs
    for log2_c in (1, 2, 0, 3):
        _say(f"batch log2_c={log2_c} sizes={sizes}")
        evals = [bfd.random_vec(1000 + i, 1 << n) for i, n in enumerate(sizes)]
        points = [bfd.random_vec(2000 + i, n) for i, n in enumerate(sizes)]
        claims = [oracle.mle_evaluate(e, p) for e, p in zip(evals, points)]
        chs = [bfd.GenericChallenger.keccak256() for _ in sizes]
        roots, proofs = bfd.commit_open_batch(128, log2_c, evals, points, claims, chs, True, ctx=ctx)
        for i, n in enumerate(sizes):
            root_o, claim_o, proof_o = oracle.basefold_prove(128, log2_c, evals[i], points[i])
            assert roots[i] == root_o, (log2_c, i, n)
            assert proofs[i].tobytes() == proof_o, (log2_c, i, n)
            root_p, _, proof_p = gpu_prove(bfd, ctx, 128, log2_c, evals[i], points[i])
            assert proof_p == proofs[i].tobytes()


# bytes a caller's transcript observed before the harness prefix: the first
# round then hashes len + 48 + 64 bytes, which ends at an odd length, one
# byte short of the 136-byte rate, on it, one past it, and two or three
# blocks in
PREFIX_LENS = [5, 23, 24, 25, 160, 300]


def test_transcript_prefix_lengths(bfd, oracle, ctx):
    """The device transcript starts from any challenger state: plain and
    batched openings after a caller prefix equal the oracle's, and both
    leave the challenger in the same state."""
    n, log2_c = 6, 1
    pcs = bfd.Basefold.setup(bfd.BasefoldConfig(bfd.BinaryReedSolomonCode(128, log2_c, n)), ctx)
    for plen in PREFIX_LENS:
        _say(f"prefix {plen} bytes")
        pre = bytes((7 * t + plen) & 0xFF for t in range(plen))
        evals, point = bfd.random_vec(50 + plen, 1 << n), bfd.random_vec(60 + plen, n)
        root_o, claims_o, proof_o = oracle.basefold_prove_w(128, log2_c, evals, 1, point, pre=pre)
        assert oracle.basefold_verify_w(128, log2_c, 1, root_o, point, claims_o, proof_o, pre=pre) == 0
        cd = pcs.commit(evals)
        claim = pcs.eval(cd, point)
        assert cd.comm == root_o and (claim == claims_o).all()
        ch = bfd.GenericChallenger.keccak256()
        ch.observe(pre)
        ch.observe(cd.comm)
        ch.observe_ext_slice(claim)
        assert pcs.open(cd, [bfd.PolyEvals(claim, point)], ch) == proof_o, plen
        ch2 = bfd.GenericChallenger.keccak256()
        ch2.observe(pre)
        _, proofs = bfd.commit_open_batch(128, log2_c, [evals], [point], [claim], [ch2], True, ctx=ctx)
        assert proofs[0].tobytes() == proof_o, plen
        assert ch.sample_bits(64) == ch2.sample_bits(64), plen


def test_batch_rejects_shared_challenger(bfd, ctx):
    """Transcripts advance concurrently, so two polynomials may not share one challenger."""
    sizes = [4, 5]
    evals = [bfd.random_vec(1 + i, 1 << n) for i, n in enumerate(sizes)]
    points = [bfd.random_vec(3 + i, n) for i, n in enumerate(sizes)]
    ch = bfd.GenericChallenger.keccak256()
    with pytest.raises(bfd.BfdError) as e:
        claims = [np.zeros((1, 2), np.uint64)] * 2
        bfd.commit_open_batch(128, 1, evals, points, claims, [ch, ch], True, ctx=ctx)
    assert e.value.code == bfd.BFD_ERR_ARG


def test_batch_short_proof_buffer_refused(bfd, oracle, ctx):
    """A caller proof buffer one byte short is refused before any work: the
    challengers are left as they were, and the same call with the right sizes
    then gives the restatement's proofs."""
    import ctypes
    sizes = [6, 9]
    evals = [bfd.random_vec(21 + i, 1 << n) for i, n in enumerate(sizes)]
    points = [bfd.random_vec(23 + i, n) for i, n in enumerate(sizes)]
    chs = [bfd.GenericChallenger.keccak256() for _ in sizes]
    before = [c.export_state() for c in chs]
    L = bfd.lib()
    host = [np.ascontiguousarray(bfd._elems(e)) for e in evals]
    pts = [np.ascontiguousarray(bfd._elems(p)) for p in points]
    want = [int(L.bfd_proof_size(128, 1, n)) for n in sizes]

    def call(caps):
        bufs = [np.zeros(c, np.uint8) for c in caps]
        roots = np.zeros(32 * len(sizes), np.uint8)
        rc = L.bfd_commit_open_batch(
            ctx.h, 128, 1, len(sizes), (ctypes.c_uint32 * 2)(*sizes),
            (ctypes.c_void_p * 2)(*[h.ctypes.data for h in host]), 0, (bfd.u8p * 2)(*[bfd._p(p) for p in pts]),
            None, (ctypes.c_void_p * 2)(*[c.h.value if isinstance(c.h, ctypes.c_void_p) else c.h for c in chs]),
            1, bfd._p(roots), (bfd.u8p * 2)(*[bfd._p(b) for b in bufs]), (ctypes.c_size_t * 2)(*caps))
        return rc, bufs

    rc, _ = call([want[0], want[1] - 1])
    assert rc == bfd.BFD_ERR_ARG
    assert [c.export_state() for c in chs] == before
    rc, bufs = call(want)
    assert rc == 0
    for i, n in enumerate(sizes):
        assert bufs[i].tobytes() == oracle.basefold_prove(128, 1, evals[i], points[i])[2], n


def test_batch_uneven_mix_verifies(bfd, oracle, ctx):
    """The benchmark mix shape (sizes 12..22) at reduced count: roots equal the
    plain commitments and every proof verifies on the CPU."""
    rng = np.random.default_rng(0)
    sizes = [int(v) for v in rng.integers(12, 23, size=6)]
    _say(f"batch mix sizes={sizes}")
    evals = [bfd.random_vec(3000 + i, 1 << n) for i, n in enumerate(sizes)]
    points = [bfd.random_vec(4000 + i, n) for i, n in enumerate(sizes)]
    claims = [oracle.mle_evaluate(e, p) for e, p in zip(evals, points)]
    chs = [bfd.GenericChallenger.keccak256() for _ in sizes]
    roots, proofs = bfd.commit_open_batch(128, 1, evals, points, claims, chs, True, ctx=ctx)
    for i, n in enumerate(sizes):
        assert oracle.basefold_verify(128, 1, roots[i], points[i], claims[i], proofs[i].tobytes()) == 0, (i, n)
        root_p, _, proof_p = gpu_prove(bfd, ctx, 128, 1, evals[i], points[i])
        assert root_p == roots[i] and proof_p == proofs[i].tobytes(), (i, n)


def test_server_mix_equals_alone(bfd, oracle, ctx):
    """Continuous batching: requests of mixed sizes and rates submitted
    together (so their rounds share ticks) give the same root, claim and proof
    as each request alone on the plain path, and the oracle's."""
    import torch
    rng = np.random.default_rng(11)
    reqs = [(int(n), int(c)) for n, c in zip(rng.integers(1, 13, size=16), rng.choice([1, 2], size=16))]
    _say(f"server requests {reqs}")
    evals = [bfd.random_vec(5000 + i, 1 << n) for i, (n, c) in enumerate(reqs)]
    points = [bfd.random_vec(6000 + i, n) for i, (n, c) in enumerate(reqs)]
    dev = [torch.from_numpy(e.view(np.int64).reshape(-1)).to("cuda:0") for e in evals]
    torch.cuda.synchronize()
    srv = bfd.Server(hbm_budget=0, ctx=ctx)
    tickets = [srv.submit(d.data_ptr(), n, p, c, keep=d) for d, p, (n, c) in zip(dev, points, reqs)]
    for i, t in enumerate(tickets):
        n, c = reqs[i]
        root, claim, proof, lat = srv.wait(t, 120000)
        root_o, claim_o, proof_o = oracle.basefold_prove(128, c, evals[i], points[i])
        assert root == root_o and (claim == claim_o).all() and proof == proof_o, (i, n, c)
        root_p, claim_p, proof_p = gpu_prove(bfd, ctx, 128, c, evals[i], points[i])
        assert proof_p == proof and lat > 0
    srv.close()


def test_server_budget_admission(bfd, oracle, ctx):
    """A tight HBM budget serialises admission but every request completes;
    a request larger than the budget is refused at submit."""
    import torch
    n, c = 10, 1
    evals = [bfd.random_vec(7000 + i, 1 << n) for i in range(6)]
    points = [bfd.random_vec(7100 + i, n) for i in range(6)]
    dev = [torch.from_numpy(e.view(np.int64).reshape(-1)).to("cuda:0") for e in evals]
    torch.cuda.synchronize()
    one = 16 * (2 << n) + 32 * (2 * (2 << n)) + 16 * (1 << n) * 2 + 16 * (2 << n) + 32 * (2 << n) + 8192
    srv = bfd.Server(hbm_budget=int(1.5 * one), ctx=ctx)
    tickets = [srv.submit(d.data_ptr(), n, p, c, keep=d) for d, p in zip(dev, points)]
    for i, t in enumerate(tickets):
        root, claim, proof, lat = srv.wait(t, 120000)
        assert oracle.basefold_verify(128, c, root, points[i], claim, proof) == 0
    big = bfd.random_vec(1, 1 << 14)
    dbig = torch.from_numpy(big.view(np.int64).reshape(-1)).to("cuda:0")
    with pytest.raises(bfd.BfdError) as e:
        srv.submit(dbig.data_ptr(), 14, bfd.random_vec(2, 14), 1)
    assert e.value.code == 3
    srv.close()
