"""CPU-side checks of the HIP library: it loads, exports every symbol the
C header declares, and its host-only logic (input generator, Fiat-Shamir
challenger) agrees with the CPU restatement.  No GPU compute here."""
import ctypes

import numpy as np
import pytest


def test_library_exports_header_symbols(bfd):
    syms = bfd.header_symbols()
    assert len(syms) >= 25
    L = bfd.lib()
    missing = [s for s in syms if not hasattr(L, s)]
    assert not missing, missing


def test_random_vec_matches_oracle(bfd, oracle):
    for seed, n in [(0, 1), (0, 17), (1, 64), (12345, 1000)]:
        assert (bfd.random_vec(seed, n) == oracle.random_vec(seed, n)).all()


def test_challenger_matches_oracle(bfd, oracle):
    rng = np.random.default_rng(5)
    a, b = bfd.GenericChallenger.keccak256(), oracle.Challenger()
    for step in range(300):
        op = rng.integers(0, 4)
        if op == 0:
            msg = bytes(rng.integers(0, 256, rng.integers(0, 200), dtype=np.uint8))
            a.observe(msg)
            b.observe(msg)
        elif op == 1:
            assert (a.sample_ext_element() == b.sample()).all()
        elif op == 2:
            bits = int(rng.integers(1, 40))
            assert a.sample_bits(bits) == b.sample_bits(bits)
        else:
            c = a.clone()
            assert (c.sample_ext_element() == a.sample_ext_element()).all()
            b.sample()


def test_challenger_state_export_import(bfd, oracle):
    """The two byte buffers round-trip: a challenger rebuilt from another's
    exported state continues exactly as the original and as the oracle's; the
    input buffer is the observed bytes since the last digest, the output
    buffer at most one digest, consumed from the end."""
    rng = np.random.default_rng(11)
    a, b = bfd.GenericChallenger.keccak256(), oracle.Challenger()
    pending = b""
    for step in range(200):
        op = rng.integers(0, 3)
        if op == 0:
            msg = bytes(rng.integers(0, 256, rng.integers(0, 300), dtype=np.uint8))
            a.observe(msg)
            b.observe(msg)
            pending += msg
            inp, out = a.export_state()
            assert out == b"" and inp.endswith(pending) and len(inp) in (len(pending), len(pending) + 32)
        elif op == 1:
            assert (a.sample_ext_element() == b.sample()).all()
            inp, out = a.export_state()
            assert len(inp) == 32 and len(out) <= 32 and out == inp[:len(out)]
            pending = b""
        else:
            inp, out = a.export_state()
            a = bfd.GenericChallenger.import_state(inp, out)
            assert a.export_state() == (inp, out)
        bits = int(rng.integers(1, 60))
        c = bfd.GenericChallenger.import_state(*a.export_state())
        assert c.sample_bits(bits) == a.clone().sample_bits(bits)
    assert a.sample_bits(64) == b.sample_bits(64)
    with pytest.raises(bfd.BfdError):
        bfd.GenericChallenger.import_state(b"", bytes(33))


def test_errors_without_handles(bfd):
    L = bfd.lib()
    h = ctypes.c_void_p()
    rc = L.bfd_code_create(None, 128, 1, 4, ctypes.byref(h))
    assert rc == 1 and b"null" in L.bfd_last_error()


def test_proof_size_matches_oracle(bfd, oracle):
    for log2_c in (0, 1, 2):
        for n in (0, 1, 5, 9):
            e, p = oracle.random_vec(1, 1 << n), oracle.random_vec(2, n)
            _, _, proof = oracle.basefold_prove(128, log2_c, e, p)
            assert bfd.proof_size(128, log2_c, n) == len(proof)
    assert bfd.proof_size(128, 1, 16) == 1645080
    assert bfd.proof_size(128, 2, 24) == 2235000


def test_proof_size_width_matches_oracle(bfd, oracle):
    """bfd_proof_size_width (pure host arithmetic) against the restatement's bincode length."""
    for n, log2_c, w in [(0, 1, 3), (5, 1, 1), (6, 2, 4), (7, 0, 9)]:
        e, p = oracle.random_vec(5, (1 << n) * w), oracle.random_vec(6, n)
        _, _, proof = oracle.basefold_prove_w(128, log2_c, e, w, p)
        assert bfd.proof_size(128, log2_c, n, w) == len(proof)


def test_field_mul_inv_match_oracle(bfd, oracle):
    a, b = oracle.random_vec(11, 64), oracle.random_vec(12, 64)
    prod = oracle.mul(a, b)
    for i in range(64):
        assert (bfd._mul_host(a[i:i + 1], b[i:i + 1]) == prod[i:i + 1]).all()
    inv = oracle.inv(a)
    out = np.zeros((1, 2), dtype=np.uint64)
    for i in range(8):
        bfd.lib().bfd_field_inv(bfd._p(np.ascontiguousarray(a[i:i + 1])), bfd._p(out))
        assert (out == inv[i:i + 1]).all()


def test_batch_quadratics_is_the_alpha_combination(bfd, oracle):
    """Batch (sumcheck/src/function/batch.rs:15-44) of equal-size Quadratics: the merged
    function's claim is sum_k alpha^k claim_k and its final identity holds at its challenges."""
    n = 5
    fs = []
    for k in range(3):
        polys = oracle.random_vec(50 + k, 2 << n).reshape(2, 1 << n, 2)
        fs.append(bfd.Quadratic(polys, [(oracle.random_vec(60 + k, 1)[0], 0, 1), (oracle.random_vec(70 + k, 1)[0], 1, 1)]))
    alpha = oracle.random_vec(80, 1)
    merged = bfd.batch_quadratics(fs, alpha)
    claims = [oracle.quadratic_sumcheck(f.polys, [p[0] for p in f.pairs], [(p[1], p[2]) for p in f.pairs])[0]
              for f in fs]
    mc, rps, chal, fin, sub = oracle.quadratic_sumcheck(merged.polys, [p[0] for p in merged.pairs],
                                                        [(p[1], p[2]) for p in merged.pairs])
    expect, power = np.zeros((1, 2), np.uint64), oracle.from_canonical(oracle.from_int(1))
    for c in claims:
        expect ^= oracle.mul(c, power)
        power = oracle.mul(power, alpha)
    assert (mc == expect).all()


def test_tuning_struct_and_default(bfd):
    """bfd_tuning is 16 uint32 fields in the header's order; the built-in
    shape needs no device, and a null context is an argument error."""
    import re
    text = bfd.HEADER.read_text()
    body = re.search(r"typedef struct bfd_tuning \{(.*?)\} bfd_tuning;", text, re.S).group(1)
    names = re.findall(r"uint32_t (\w+)(?:\[\d+\])?;", body)
    assert tuple(names[:-1]) == bfd.Tuning.FIELDS and names[-1] == "reserved"
    assert ctypes.sizeof(bfd.Tuning) == 64
    t = bfd.Tuning.default()
    assert t.as_dict() == {"ntt_order": 0, "ntt_block": 256, "ntt_waves_per_cu": 32, "ntt_occupancy": 2,
                           "hash_block": 256, "hash_waves_per_cu": 64, "subtree_from_log2": 17,
                           "subtree_levels": 9, "seg_blocks_per_cu": 8}
    L = bfd.lib()
    assert L.bfd_tuning_set(None, 0, ctypes.byref(t)) == bfd.BFD_ERR_ARG
    assert L.bfd_tuning_get(None, 0, ctypes.byref(t)) == bfd.BFD_ERR_ARG
    assert L.bfd_tuning_default(None) == bfd.BFD_ERR_ARG
    assert L.bfd_tuning_load(None, b"/nonexistent") == bfd.BFD_ERR_ARG
