"""CPU-side checks of the HIP library: it loads, exports every symbol the
C header declares, and its host-only logic (input generator, Fiat-Shamir
challenger) agrees with the CPU restatement.  No GPU compute here."""
import ctypes

import numpy as np


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
