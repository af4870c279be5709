"""Pins the CPU restatement (oracle/bfo.c) before it is trusted as the checker.

Each check is independent of the oracle's own formulas: a reference-held
table, a library hash, a published RFC vector, a brute-force definition from
the reference's tests, or a mathematical invariant.
"""
import hashlib
import pathlib

import numpy as np
import pytest

GOLDEN = pathlib.Path(__file__).parent / "golden"


def rnd(O, n, seed):
    return O.random_vec(seed, n)


def test_polyval_nibble_pow_table_kat(oracle):
    """p3/src/polyval/portable.rs:169-376: T[i][k][j] = from_canonical(j<<4k)^(2^(2^(i+1)))."""
    O = oracle
    vals = [int(l, 16) for l in (GOLDEN / "polyval_nibble_pow_table.txt").read_text().split("\n")
            if l and not l.startswith("#")]
    assert len(vals) == 3072
    base = np.array([[(j << (4 * k)) & (2**64 - 1), (j << (4 * k)) >> 64]
                     for k in range(32) for j in range(16)], dtype=np.uint64)
    x = O.from_canonical(base)
    for i in range(6):
        for _ in range(2 ** (i + 1) - (2 ** i if i else 0)):
            x = O.mul(x, x)
        got = [O.to_int(e) for e in x]
        assert got == vals[i * 512:(i + 1) * 512], f"table block {i}"


def test_polyval_one_and_constants(oracle):
    O = oracle
    one = O.from_canonical(O.from_int(1))
    assert O.to_int(one) == 0xC2000000000000000000000000000001
    # basis(i) = from_canonical(1 << i) (polyval.rs:167-172) round-trips
    for i in range(128):
        assert O.to_int(O.to_canonical(O.from_canonical(O.from_int(1 << i)))) == 1 << i


@pytest.mark.parametrize("impl", [0, 1])
def test_polyval_mul_matches_textbook(oracle, impl):
    O = oracle
    a, b = rnd(O, 2000, 11), rnd(O, 2000, 12)
    assert (O.mul(a, b, impl) == O.mul(a, b, 2)).all()


def test_polyval_canonical_product_identity(oracle):
    """to_canonical(from_canonical(x) * from_canonical(y)) == x*y mod P."""
    O = oracle
    x, y = rnd(O, 500, 13), rnd(O, 500, 14)
    lhs = O.to_canonical(O.mul(O.from_canonical(x), O.from_canonical(y)))
    assert (lhs == O.polymul_mod(x, y)).all()


def test_polyval_field_axioms_and_inverse(oracle):
    """polyval.rs:265-295 identities in characteristic 2."""
    O = oracle
    x, y, z = rnd(O, 64, 15), rnd(O, 64, 16), rnd(O, 64, 17)
    one = np.repeat(O.from_canonical(O.from_int(1)), 64, axis=0)
    assert (O.mul(x, one) == x).all()
    assert (O.mul(x, y) == O.mul(y, x)).all()
    assert (O.mul(x, O.mul(y, z)) == O.mul(O.mul(x, y), z)).all()
    assert (O.mul(x, y ^ z) == (O.mul(x, y) ^ O.mul(x, z))).all()
    assert (O.mul(x, O.inv(x)) == one).all()


def test_keccak_kats(oracle):
    O = oracle
    assert O.keccak256(b"").hex() == "c5d2460186f7233c927e7db2dcc703c0e500b653ca82273b7bfad8045d85a470"
    rng = np.random.default_rng(3)
    for n in [0, 1, 16, 32, 64, 135, 136, 137, 272, 1000]:
        msg = bytes(rng.integers(0, 256, n, dtype=np.uint8))
        assert O.sha3_256(msg) == hashlib.sha3_256(msg).digest()


def test_chacha20_rfc8439_block(oracle):
    """RFC 8439 section 2.3.2 test vector pins the ChaCha core."""
    O = oracle
    key = [int.from_bytes(bytes(range(4 * i, 4 * i + 4)), "little") for i in range(8)]
    nonce = bytes.fromhex("000000090000004a00000000")
    cn = [1] + [int.from_bytes(nonce[4 * i:4 * i + 4], "little") for i in range(3)]
    out = O.chacha_block(20, key, cn)
    assert out[:4] == [0xE4E7F110, 0x15593BD1, 0x1FDD0F50, 0xC47120A3]
    assert out[12:] == [0xD19C12B5, 0xB94E16DE, 0xE883D0CB, 0x4E3C50A2]


@pytest.mark.parametrize("log_n", range(0, 9))
def test_ntt_matches_definition(oracle, log_n):
    """p3/src/field/binary.rs:194-203 (forward vs explicit novel-basis matrix)."""
    O = oracle
    d = rnd(O, 1 << log_n, 20 + log_n)
    assert (O.ntt_forward(log_n, d) == O.ntt_forward_definition(log_n, d)).all()


def test_ntt_round_trip(oracle):
    """p3/src/field/binary.rs:205-217."""
    O = oracle
    for log_n in range(0, 10):
        for width in (1, 3, 8):
            a = rnd(O, (1 << log_n) * width, log_n * 10 + width)
            b = O.ntt_forward(log_n, a, width)
            assert (O.ntt_backward(log_n, b, width) == a).all()


def test_ntt_last_layer_closed_form(oracle):
    """twiddles[log_n-1][j] = from_canonical(2j) since W_0 is the identity."""
    O = oracle
    L = 9
    tw = O.ntt_twiddles(L)
    last = tw[(1 << (L - 1)) - 1:]
    exp = O.from_canonical(np.array([[2 * j, 0] for j in range(1 << (L - 1))], dtype=np.uint64))
    assert (last == exp).all()


def _bitrev(v, bits):
    return int(format(v, f"0{bits}b")[::-1], 2) if bits else 0


@pytest.mark.parametrize("log2_c,d", [(0, 3), (1, 4), (2, 3), (1, 6)])
def test_rs_encode_closed_form(oracle, log2_c, d):
    """codeword[v] = P(point bitrev(v)), P = novel-basis poly of bitrev(data), via the brute-force matrix."""
    O = oracle
    N, L = 1 << d, log2_c + d
    data = rnd(O, N, 30 + d)
    coeffs = np.zeros(((1 << L), 2), dtype=np.uint64)
    coeffs[:N] = data[[_bitrev(j, d) for j in range(N)]]
    ev = O.ntt_forward_definition(L, coeffs)
    cw = O.rs_encode(log2_c, d, data)
    assert (cw == ev[[_bitrev(v, L) for v in range(1 << L)]]).all()


def test_num_queries(oracle):
    """binary_reed_solomon.rs:55-59: c=2 -> 309, c=4 -> 189, c=1 -> 0 (saturating cast)."""
    assert oracle.num_queries(128, 1) == 309
    assert oracle.num_queries(128, 2) == 189
    assert oracle.num_queries(128, 0) == 0


def test_mle_and_eq(oracle):
    """multilinear.rs:312-332 definitions restated: eval = sum eq(x, b) f(b)."""
    O = oracle
    for n in range(0, 7):
        f, x = rnd(O, 1 << n, 40 + n), rnd(O, n, 50 + n)
        eq = O.eq_table(x)
        acc = 0
        for k in range(1 << n):
            acc ^= O.to_int(O.mul(eq[k:k + 1], f[k:k + 1]))
        assert O.to_int(O.mle_evaluate(f, x)) == acc


@pytest.mark.parametrize("log2_c", [0, 1, 2])
def test_basefold_prove_verify_round_trip(oracle, log2_c):
    """poly-commit/src/basefold.rs:480-490 (num_vars 0..9), plus tampering."""
    O = oracle
    for n in range(0, 10):
        e, p = rnd(O, 1 << n, 60 + n), rnd(O, n, 70 + n)
        root, claim, proof = O.basefold_prove(128, log2_c, e, p)
        assert O.basefold_verify(128, log2_c, root, p, claim, proof) == 0
        assert (claim == O.mle_evaluate(e, p)).all()
        if n > 0:
            bad_claim = claim.copy()
            bad_claim[0, 0] ^= 1
            assert O.basefold_verify(128, log2_c, root, p, bad_claim, proof) != 0
        if O.num_queries(128, log2_c) and n > 0:
            bad = bytearray(proof)
            bad[-3] ^= 0x10
            assert O.basefold_verify(128, log2_c, root, p, claim, bytes(bad)) == 3


def test_transcript_prefix_binds_proof(oracle):
    """Bytes observed before the harness prefix change every challenge: the
    proof differs from the plain one, verifies under the same prefix, and
    fails under none or another one (the transcript is bound to its whole
    history)."""
    O = oracle
    n, log2_c = 5, 1
    e, p = rnd(O, 1 << n, 90), rnd(O, n, 91)
    root0, claims0, proof0 = O.basefold_prove_w(128, log2_c, e, 1, p)
    for plen in (1, 23, 137, 300):
        pre = bytes((3 * t + 1) & 0xFF for t in range(plen))
        root, claims, proof = O.basefold_prove_w(128, log2_c, e, 1, p, pre=pre)
        assert root == root0 and (claims == claims0).all()
        assert len(proof) == len(proof0) and proof != proof0
        assert O.basefold_verify_w(128, log2_c, 1, root, p, claims, proof, pre=pre) == 0
        assert O.basefold_verify_w(128, log2_c, 1, root, p, claims, proof) != 0
        assert O.basefold_verify_w(128, log2_c, 1, root, p, claims, proof, pre=pre[:-1] + b"\xff") != 0


def test_proof_size_formula(oracle):
    """bincode size of BasefoldProof at 2^n, rate 1/2 matches the wire-format arithmetic."""
    O = oracle
    for n, log2_c in [(6, 1), (8, 2)]:
        e, p = rnd(O, 1 << n, 1), rnd(O, n, 2)
        _, _, proof = O.basefold_prove(128, log2_c, e, p)
        q = O.num_queries(128, log2_c)
        L = n + log2_c
        per = 24 + 8 + 32 * L + 8 + sum(16 + 8 + 32 * (log2_c + i) for i in range(n))
        size = 8 + n * 40 + 24 + 8 + 32 * n + 8 + q * per
        assert len(proof) == size


def test_product_sumcheck_invariants(oracle):
    """Degree-3 product sumcheck: p(0)+p(1)=claim is asserted inside; the final
    subclaim must equal f(x) g(x) eq(r, x) computed independently."""
    O = oracle
    for n in range(1, 8):
        f, g, r = rnd(O, 1 << n, 80 + n), rnd(O, 1 << n, 90 + n), rnd(O, n, 95 + n)
        claim, polys, chal, fin, sub = O.product_sumcheck(f, g, r)
        x = chal[::-1].copy()  # challenge of round k fixes variable n-1-k
        fx, gx = O.mle_evaluate(f, x), O.mle_evaluate(g, x)
        assert (fin[0:1] == fx).all() and (fin[1:2] == gx).all()
        eq = O.mle_evaluate(O.eq_table(r), x)
        assert (O.mul(O.mul(fx, gx), eq) == sub).all()


def test_eval_sumcheck_invariants(oracle):
    """The Basefold sumcheck function (EvalProver, eq(point, .) * poly) driven by
    prove_sumcheck: every round polynomial satisfies p(0) + p(1) = subclaim (char 2:
    p(1) = c0 + c1 + c2), the challenges replay from the transcript (observe the
    claim, then [c0, c2] per round), the final table value is poly's MLE at the
    challenges and the final subclaim equals it times eq(point, x), each evaluated
    independently of the prover's formulas."""
    O = oracle
    for n in range(1, 10):
        poly, point = rnd(O, 1 << n, 500 + n), rnd(O, n, 600 + n)
        claim = O.mle_evaluate(poly, point)
        rps, chal, fin, sub = O.eval_sumcheck(poly, point, claim)
        ch = O.Challenger()
        ch.observe(claim.tobytes())
        s = claim
        for k in range(n):
            c0, c1, c2 = rps[k, 0:1], rps[k, 1:2], rps[k, 2:3]
            assert (c0 ^ (c0 ^ c1 ^ c2) == s).all(), (n, k)  # p(0) + p(1)
            ch.observe(c0.tobytes() + c2.tobytes())
            x = ch.sample()
            assert (x == chal[k:k + 1]).all(), (n, k)
            s = O.mul(O.mul(c2, x) ^ c1, x) ^ c0
        assert (s == sub).all()
        x = chal[::-1].copy()  # the challenge of round k fixes variable n-1-k
        px = O.mle_evaluate(poly, x)
        assert (fin == px).all()
        assert (O.mul(px, O.mle_evaluate(O.eq_table(point), x)) == sub).all(), n


# ---- width > 1 (one polynomial per column of a row-major matrix) ----
def one(O):
    return O.from_canonical(O.from_int(1))


def eq_by_product(O, r, count):
    """eq(r, j) = prod_i (r_i if bit i of j else 1 + r_i), straight from the definition."""
    o = one(O)
    out = []
    for j in range(count):
        acc = o
        for i in range(len(r)):
            f = r[i:i + 1] if (j >> i) & 1 else (r[i:i + 1] ^ o)
            acc = O.mul(acc, f)
        out.append(acc[0])
    return np.array(out, dtype=np.uint64).reshape(-1, 2)


def test_width_prove_verify_grid(oracle):
    """The reference's own harness grid (poly-commit/src/lib.rs:73-77: num_vars 0..9 x num_polys
    1..9) at rate 1/2: every proof verifies, claims are the column MLEs."""
    O = oracle
    for n in range(0, 10):
        for w in range(1, 10):
            e, p = rnd(O, (1 << n) * w, 300 + 10 * n + w), rnd(O, n, 400 + n)
            root, claims, proof = O.basefold_prove_w(128, 1, e, w, p)
            assert O.basefold_verify_w(128, 1, w, root, p, claims, proof) == 0, (n, w)
            cols = e.reshape(1 << n, w, 2)
            for t in range(w):
                assert (claims[t:t + 1] == O.mle_evaluate(np.ascontiguousarray(cols[:, t]), p)).all()


def test_width_encode_is_columnwise(oracle):
    """encode_batch of a w-column matrix equals the width-1 encoding of each column."""
    O = oracle
    for log2_c, d, w in [(1, 4, 3), (2, 5, 5), (0, 3, 2)]:
        m = rnd(O, (1 << d) * w, 500 + w)
        cw = O.rs_encode(log2_c, d, m, w).reshape(-1, w, 2)
        cols = m.reshape(1 << d, w, 2)
        for t in range(w):
            assert (cw[:, t] == O.rs_encode(log2_c, d, np.ascontiguousarray(cols[:, t]), 1)).all()


def parse_head(proof, d):
    """bincode head of BasefoldProof: round polys, final poly, pi roots."""
    pos, u64 = 0, lambda b, o: int.from_bytes(b[o:o + 8], "little")
    assert u64(proof, pos) == d
    pos += 8
    comps = []
    for _ in range(d):
        assert u64(proof, pos) == 2
        comps.append((proof[pos + 8:pos + 24], proof[pos + 24:pos + 40]))
        pos += 40
    assert u64(proof, pos) == 1
    final = proof[pos + 8:pos + 24]
    pos += 24
    assert u64(proof, pos) == d
    roots = [proof[pos + 8 + 32 * k:pos + 40 + 32 * k] for k in range(d)]
    return comps, final, roots


def test_width_scalars_pinned_by_final_poly(oracle):
    """basefold.rs:166-183, 398-408 restated independently: replay the verifier transcript
    (root, claims, log2(next_pow2(w)) scalar samples, then per round pi root, c0, c2 and the
    challenge), build the scalars by the eq product definition, combine the columns, and check
    that the proof's final poly is that combination's MLE at the round challenges."""
    O = oracle
    for n, w, log2_c in [(5, 3, 1), (6, 5, 2), (4, 8, 1), (3, 9, 0)]:
        e, p = rnd(O, (1 << n) * w, 600 + w), rnd(O, n, 700 + n)
        root, claims, proof = O.basefold_prove_w(128, log2_c, e, w, p)
        comps, final, roots = parse_head(proof, n)
        ch = O.Challenger()
        ch.observe(root)
        for t in range(w):
            ch.observe(claims[t].tobytes())
        lw = (w - 1).bit_length()
        r = np.concatenate([ch.sample() for _ in range(lw)]) if lw else np.zeros((0, 2), np.uint64)
        s = eq_by_product(O, r, w)
        chal = [None] * n
        for k in range(n):
            ch.observe(roots[k])
            ch.observe(comps[k][0])
            ch.observe(comps[k][1])
            chal[n - 1 - k] = ch.sample()[0]
        cols = e.reshape(1 << n, w, 2)
        comb = np.zeros((1 << n, 2), np.uint64)
        for t in range(w):
            comb ^= O.mul(np.ascontiguousarray(cols[:, t]), np.repeat(s[t:t + 1], 1 << n, axis=0))
        x = np.array(chal, dtype=np.uint64).reshape(-1, 2)
        assert O.mle_evaluate(comb, x).tobytes() == final, (n, w)


def test_width_tampering_and_size(oracle):
    """A flipped byte of an opened row fails the Merkle check, a wrong claim fails, and the
    proof grows by q * 16 * (w - 1) bytes over width 1 (the opened rows carry w elements)."""
    O = oracle
    n, log2_c = 6, 1
    q = O.num_queries(128, log2_c)
    p = rnd(O, n, 801)
    _, _, p1 = O.basefold_prove_w(128, log2_c, rnd(O, 1 << n, 800), 1, p)
    for w in [2, 4, 7]:
        e = rnd(O, (1 << n) * w, 810 + w)
        root, claims, proof = O.basefold_prove_w(128, log2_c, e, w, p)
        assert len(proof) == len(p1) + q * 16 * (w - 1)
        comps, final, roots = parse_head(proof, n)
        head = 8 + 40 * n + 24 + 8 + 32 * n + 8
        bad = bytearray(proof)
        bad[head + 8 + 3] ^= 1  # first opened row, first element
        assert O.basefold_verify_w(128, log2_c, w, root, p, claims, bytes(bad)) == 3
        bc = claims.copy()
        bc[w - 1, 1] ^= 4
        assert O.basefold_verify_w(128, log2_c, w, root, p, bc, proof) != 0


# ---- Quadratic sumcheck (sumcheck/src/function/quadratic.rs) ----
def quad_case(O, n, seed):
    """The reference test's shape (quadratic.rs:150-171): 5..9 polynomials, 5..9 pairs of random
    indices with random scalars."""
    rng = np.random.default_rng(seed)
    m, npairs = int(rng.integers(5, 10)), int(rng.integers(5, 10))
    polys = O.random_vec(seed, m << n).reshape(m, 1 << n, 2)
    scalars = O.random_vec(seed + 1, npairs)
    pairs = [(int(rng.integers(0, m)), int(rng.integers(0, m))) for _ in range(npairs)]
    return polys, scalars, pairs


def test_quadratic_sumcheck_pinned(oracle):
    """Claim by brute force; the challenges replay from the transcript (observe claim, then
    [c0, c2] per round, lib.rs:67,94-96); and the verifier's final identity: the last subclaim
    equals sum_k s_k P_a(r) P_b(r) with every P(r) evaluated independently as an MLE."""
    O = oracle
    for n in range(1, 9):
        polys, scalars, pairs = quad_case(O, n, 40 + n)
        claim, rps, chal, fin, sub = O.quadratic_sumcheck(polys, scalars, pairs)
        brute = np.zeros((1, 2), np.uint64)
        for (a, b), s in zip(pairs, scalars):
            prod = O.mul(polys[a], polys[b])
            acc = np.bitwise_xor.reduce(prod, axis=0).reshape(1, 2)
            brute ^= O.mul(acc, s.reshape(1, 2))
        assert (claim == brute).all(), n
        ch = O.Challenger()
        ch.observe(claim.tobytes())
        for k in range(n):
            ch.observe(rps[k, 0].tobytes())
            ch.observe(rps[k, 1].tobytes())
            assert (ch.sample() == chal[k]).all(), (n, k)
        x = chal[::-1].copy()
        evals = [O.mle_evaluate(polys[p], x) for p in range(len(polys))]
        for p in range(len(polys)):
            assert (fin[p:p + 1] == evals[p]).all()
        expect = np.zeros((1, 2), np.uint64)
        for (a, b), s in zip(pairs, scalars):
            expect ^= O.mul(O.mul(evals[a], evals[b]), s.reshape(1, 2))
        assert (sub == expect).all(), n


def test_verify_rejects_truncated_proofs(oracle):
    """A proof cut short anywhere is malformed (bincode cannot decode it), including the
    rate-1 case whose proof ends with an empty query list."""
    O = oracle
    for log2_c in [0, 1]:
        e, p = rnd(O, 1 << 5, 90), rnd(O, 5, 91)
        root, claim, proof = O.basefold_prove(128, log2_c, e, p)
        assert O.basefold_verify(128, log2_c, root, p, claim, proof) == 0
        for cut in [1, 8, 16, len(proof) // 2]:
            assert O.basefold_verify(128, log2_c, root, p, claim, proof[:-cut]) == 4, (log2_c, cut)
        assert O.basefold_verify(128, log2_c, root, p, claim, proof + b"\0") == 4
