"""GPU parity of the entry points a caller's own transcript or driver uses.

* Basefold::open round by round (bfd_open_begin .. bfd_open_finish) with the
  restatement's challenger (oracle.Challenger) as the transcript: the library
  never sees it, as with a plonky3 HashChallenger whose buffers are private.
  The harness prefix of poly-commit/src/lib.rs:77-100 is applied by the test.
* Basefold::eval on data without a commitment (bfd_eval_data).
* EvalProver step by step (bfd_sc_eval_create + bfd_sc_*) under a host
  prove_sumcheck loop.
* The device multiply at its carry bound: operands whose masked 4-bit classes
  are saturated, against the restatement's textbook multiply.
"""
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


class OracleTranscript:
    """The restatement's GenericChallenger behind the method names the
    library's step-wise open calls."""

    def __init__(self, oracle):
        self.c = oracle.Challenger()

    def observe(self, data: bytes):
        self.c.observe(bytes(data))

    def observe_ext_slice(self, elems):
        self.c.observe(np.ascontiguousarray(elems, dtype=np.uint64).tobytes())

    def sample_ext_element(self):
        return self.c.sample()

    def sample_bits(self, bits):
        return self.c.sample_bits(bits)


def _stepwise_prove(bfd, oracle, ctx, lam, log2_c, evals, width, point):
    n = len(point)
    pcs = bfd.Basefold.setup(bfd.BasefoldConfig(bfd.BinaryReedSolomonCode(lam, log2_c, n)), ctx)
    cd = pcs.commit(evals, width)
    claims = pcs.eval_data(evals, point, width)
    tr = OracleTranscript(oracle)
    tr.observe(cd.comm)
    tr.observe_ext_slice(claims)
    proof = pcs.open_stepwise(cd, [bfd.PolyEvals(claims, point)], tr)
    return cd.comm, claims, proof, tr


@pytest.mark.parametrize("log2_c", [0, 1, 2])
def test_stepwise_open_equals_oracle(bfd, oracle, ctx, log2_c):
    """n 0..9 at rates 1, 1/2 and 1/4 (basefold.rs:480-490's grid): the proof a
    foreign transcript drives equals the restatement's byte for byte."""
    for n in range(0, 10):
        _say(f"stepwise log2_c={log2_c} n={n}")
        evals, point = bfd.random_vec(100 + n, 1 << n), bfd.random_vec(200 + n, n)
        root, claims, proof, _ = _stepwise_prove(bfd, oracle, ctx, 128, log2_c, evals, 1, point)
        root_o, claim_o, proof_o = oracle.basefold_prove(128, log2_c, evals, point)
        assert root == root_o and (claims == claim_o).all(), (log2_c, n)
        assert proof == proof_o, (log2_c, n)


def test_stepwise_open_2e16_and_width(bfd, oracle, ctx):
    """2^16 at rate 1/2 (the small benchmark case), and widths 3 and 9, whose
    openings first sample the scalar challenges from the caller's transcript."""
    cases = [(16, 1, 1), (6, 1, 3), (9, 2, 9), (0, 1, 5)]
    for n, log2_c, w in cases:
        _say(f"stepwise n={n} log2_c={log2_c} w={w}")
        evals, point = bfd.random_vec(40 + n + w, (1 << n) * w), bfd.random_vec(50 + n, n)
        root, claims, proof, tr = _stepwise_prove(bfd, oracle, ctx, 128, log2_c, evals, w, point)
        root_o, claims_o, proof_o = oracle.basefold_prove_w(128, log2_c, evals, w, point)
        assert root == root_o and (claims == claims_o).all(), (n, w)
        assert proof == proof_o, (n, log2_c, w)


def test_stepwise_open_matches_device_transcript(bfd, oracle, ctx):
    """With a GenericChallenger the step-wise and the one-call opening give the
    same proof and leave the transcripts in the same state; empty evaluations
    give the default proof (basefold.rs:139-141)."""
    n, log2_c = 11, 2
    evals, point = bfd.random_vec(61, 1 << n), bfd.random_vec(62, n)
    pcs = bfd.Basefold.setup(bfd.BasefoldConfig(bfd.BinaryReedSolomonCode(128, log2_c, n)), ctx)
    cd = pcs.commit(evals)
    claim = pcs.eval(cd, point)
    a, b = bfd.GenericChallenger.keccak256(), bfd.GenericChallenger.keccak256()
    for ch in (a, b):
        ch.observe(b"prefix")
        ch.observe(cd.comm)
        ch.observe_ext_slice(claim)
    p1 = pcs.open(cd, [bfd.PolyEvals(claim, point)], a)
    p2 = pcs.open_stepwise(cd, [bfd.PolyEvals(claim, point)], b)
    assert p1 == p2
    assert a.export_state() == b.export_state()
    assert pcs.open(cd, [], a) == bfd.EMPTY_PROOF == pcs.open_stepwise(cd, [], b)


def test_stepwise_open_state_errors(bfd, ctx):
    """Calls out of order are refused with BFD_ERR_STATE and leave the opening usable."""
    import ctypes
    n = 3
    evals, point = bfd.random_vec(71, 1 << n), bfd.random_vec(72, n)
    pcs = bfd.Basefold.setup(bfd.BasefoldConfig(bfd.BinaryReedSolomonCode(128, 1, n)), ctx)
    cd = pcs.commit(evals)
    claim = pcs.eval(cd, point)
    L = bfd.lib()
    h = ctypes.c_void_p()
    scal = np.zeros((1, 2), np.uint64)
    bfd._check(L.bfd_open_begin(ctx.h, pcs.code_h, cd.handle, bfd._p(point), bfd._p(claim), bfd._p(scal), 0,
                                ctypes.byref(h)))
    try:
        x = np.ascontiguousarray(bfd.random_vec(73, 1))
        root = (ctypes.c_uint8 * 32)()
        rp = np.zeros((2, 2), np.uint64)
        fin = np.zeros((1, 2), np.uint64)
        assert L.bfd_open_challenge(h, bfd._p(x)) == bfd.BFD_ERR_STATE
        assert L.bfd_open_final_poly(h, bfd._p(fin)) == bfd.BFD_ERR_STATE
        bfd._check(L.bfd_open_round(h, root, bfd._p(rp)))
        assert L.bfd_open_round(h, root, bfd._p(rp)) == bfd.BFD_ERR_STATE
        bfd._check(L.bfd_open_challenge(h, bfd._p(x)))
        idx = (ctypes.c_uint64 * 1)(0)
        proof, plen = bfd.u8p(), ctypes.c_size_t()
        assert L.bfd_open_finish(h, idx, 1, ctypes.byref(proof), ctypes.byref(plen)) == bfd.BFD_ERR_ARG
    finally:
        L.bfd_opening_destroy(h)
    # a width-1 opening takes no scalar challenges
    h2 = ctypes.c_void_p()
    assert L.bfd_open_begin(ctx.h, pcs.code_h, cd.handle, bfd._p(point), bfd._p(claim), bfd._p(scal), 1,
                            ctypes.byref(h2)) == bfd.BFD_ERR_ARG


def test_eval_data_equals_oracle(bfd, oracle, ctx):
    """Basefold::eval on the data itself (basefold.rs:126-131), host and device,
    widths 1, 3 and 17, against the restatement's MLE of every column."""
    import torch
    for n, w in [(0, 1), (1, 3), (10, 1), (13, 3), (7, 17), (20, 1)]:
        _say(f"eval_data n={n} w={w}")
        data, point = bfd.random_vec(80 + n + w, (1 << n) * w), bfd.random_vec(90 + n, n)
        got = bfd.eval_data(data, point, w, ctx)
        mat = data.reshape(1 << n, w, 2)
        exp = np.concatenate([oracle.mle_evaluate(np.ascontiguousarray(mat[:, c]), point) for c in range(w)])
        assert (got == exp).all(), (n, w)
        dev = torch.from_numpy(data.view(np.int64).reshape(-1)).to("cuda:0")
        torch.cuda.synchronize()
        got_d = bfd.eval_data(None, point, w, ctx, device_ptr=dev.data_ptr(), height=1 << n)
        assert (got_d == exp).all(), (n, w)


def test_eval_prover_stepwise_equals_oracle(bfd, oracle, ctx):
    """EvalProver driven round by round by prove_sumcheck on the host (a
    GenericChallenger): the round polynomials, challenges, final evaluation
    and subclaim equal the restatement's for n 1..12; compute_sum agrees with
    the subclaim before every round (the reference's debug assertion)."""
    for n in range(1, 13):
        _say(f"eval prover n={n}")
        poly, point = bfd.random_vec(300 + n, 1 << n), bfd.random_vec(310 + n, n)
        claim = oracle.mle_evaluate(poly, point)
        rps_o, chal_o, fin_o, sub_o = oracle.eval_sumcheck(poly, point, claim)
        p = bfd.EvalProver(point, poly, ctx)
        assert (p.compute_sum() == claim).all(), n
        ch = bfd.GenericChallenger.keccak256()
        ch.observe_ext_slice(claim)
        sub = claim
        for k in range(n):
            if k > 0 and n <= 8:
                assert (p.compute_sum() == sub).all(), (n, k)
            rp = p.compute_round_poly(n - 1 - k, sub)
            assert (rp == rps_o[k]).all(), (n, k)
            ch.observe_ext_slice(p.compress_round_poly(rp))
            x = ch.sample_ext_element()
            assert (x == chal_o[k:k + 1]).all(), (n, k)
            p.fix_last_var(x)
            acc = np.zeros((1, 2), np.uint64)
            for c in rp[::-1]:
                acc = oracle.mul(acc, x) ^ c.reshape(1, 2)
            sub = acc
        assert p.num_vars == 0
        assert (p.evaluations() == fin_o).all() and (sub == sub_o).all(), n


def test_eval_prover_device_table_and_errors(bfd, oracle, ctx):
    """A device-resident table gives the same first round polynomial and is not
    modified; a coordinate of 0 is refused (r is inverted, eval_impr.rs:20-27)."""
    import torch
    n = 10
    poly, point = bfd.random_vec(401, 1 << n), bfd.random_vec(402, n)
    claim = oracle.mle_evaluate(poly, point)
    dev = torch.from_numpy(poly.view(np.int64).reshape(-1)).to("cuda:0")
    torch.cuda.synchronize()
    a = bfd.EvalProver(point, None, ctx, device_ptr=dev.data_ptr())
    b = bfd.EvalProver(point, poly, ctx)
    assert (a.compute_round_poly(n - 1, claim) == b.compute_round_poly(n - 1, claim)).all()
    a.fix_last_var(bfd.random_vec(403, 1))
    assert (dev.cpu().numpy() == poly.view(np.int64).reshape(-1)).all()
    bad = point.copy()
    bad[3] = 0
    with pytest.raises(bfd.BfdError):
        bfd.EvalProver(bad, poly, ctx)


def _weight_patterns():
    """Operands that saturate the multiply's masked 4-bit classes: all ones, each
    class alone (0x1111.., 0x2222.., 0x4444.., 0x8888..), three of four classes
    (0x7777.., 0xbbbb.., 0xdddd.., 0xeeee..), and 0 / 1 for the edges."""
    m = (1 << 128) - 1
    pats = [m, 0, 1, 1 << 127]
    for nib in [0x1, 0x2, 0x4, 0x8, 0x7, 0xB, 0xD, 0xE, 0x3, 0x5, 0x6, 0x9, 0xA, 0xC, 0xF]:
        pats.append(int(f"{nib:x}" * 32, 16))
    return [p & m for p in pats]


def test_device_multiply_carry_bound(bfd, oracle, ctx):
    """Every pair of the saturating patterns (both operand orders, general and
    shared-operand forms), and 2^20 random operands with bit density 15/16,
    against the textbook bit-serial multiply of the restatement."""
    pats = _weight_patterns()
    A = np.array([[p & (2**64 - 1), p >> 64] for p in pats for _ in pats], dtype=np.uint64)
    B = np.array([[p & (2**64 - 1), p >> 64] for _ in pats for p in pats], dtype=np.uint64)
    exp = oracle.mul(A, B, impl=2)
    assert (bfd.field_mul_batch(A, B, ctx=ctx) == exp).all()
    assert (bfd.field_mul_batch(A, B, ctx=ctx, wide=True) == exp).all()
    for i, p in enumerate(pats):
        b = np.array([[p & (2**64 - 1), p >> 64]], dtype=np.uint64)
        a = np.ascontiguousarray(A[:len(pats)])
        assert (bfd.field_mul_batch(a, b, shared_b=True, ctx=ctx) == oracle.mul(a, np.repeat(b, len(a), 0), impl=2)).all(), i
    rng = np.random.default_rng(17)
    n = 1 << 20
    clear_a = rng.random((n, 128)) < 1 / 16  # each bit cleared with probability 1/16
    clear_b = rng.random((n, 128)) < 1 / 16
    a = _dense(clear_a)
    b = _dense(clear_b)
    exp = oracle.mul(a, b, impl=2)
    assert (bfd.field_mul_batch(a, b, ctx=ctx) == exp).all()
    # the four-lane form on a ragged count (the last wave partly past n)
    assert (bfd.field_mul_batch(a[:n - 37], b[:n - 37], ctx=ctx, wide=True) == exp[:n - 37]).all()


def _dense(clear):
    """Elements with every bit set except where `clear` (n x 128 booleans) is true."""
    bits = (~clear).astype(np.uint64)
    w = np.uint64(1) << np.arange(64, dtype=np.uint64)
    lo = (bits[:, :64] * w).sum(axis=1, dtype=np.uint64)
    hi = (bits[:, 64:] * w).sum(axis=1, dtype=np.uint64)
    return np.ascontiguousarray(np.stack([lo, hi], axis=1))
