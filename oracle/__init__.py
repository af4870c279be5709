"""ctypes binding of the CPU restatement in ``oracle/bfo.c``.

Test infrastructure only: imported by ``tests/``, ``__graft_entry__.smoke()``
and the CPU-baseline leg of ``bench.py`` as the checker.  Field elements are
passed as numpy ``uint64`` arrays of shape ``(n, 2)`` (low word first), which
is the 16-byte little-endian layout of the reference's ``Polyval``.
"""
from __future__ import annotations

import ctypes
import os
import pathlib
import subprocess

import numpy as np

_DIR = pathlib.Path(__file__).resolve().parent
_LIB_PATH = _DIR / "liboracle.so"
_lib = None

u8p = ctypes.POINTER(ctypes.c_uint8)


def build(force: bool = False) -> None:
    """Compile the restatement with the committed Makefile."""
    if force or not _LIB_PATH.exists():
        subprocess.run(["make", "-C", str(_DIR), "-B" if force else "all"], check=True,
                       stdout=subprocess.DEVNULL)


def lib():
    global _lib
    if _lib is None:
        if not _LIB_PATH.exists():
            build()
        _lib = ctypes.CDLL(str(_LIB_PATH))
        _lib.bfo_num_queries.restype = ctypes.c_size_t
        _lib.bfo_challenger_new.restype = ctypes.c_void_p
        _lib.bfo_challenger_sample_bits.restype = ctypes.c_uint64
        _lib.bfo_challenger_sample_bits.argtypes = [ctypes.c_void_p, ctypes.c_int]
        _lib.bfo_challenger_observe.argtypes = [ctypes.c_void_p, u8p, ctypes.c_size_t]
        _lib.bfo_challenger_sample.argtypes = [ctypes.c_void_p, u8p]
        _lib.bfo_challenger_free.argtypes = [ctypes.c_void_p]
        _lib.bfo_random_vec.argtypes = [ctypes.c_uint64, ctypes.c_size_t, u8p]
        _lib.bfo_free.argtypes = [ctypes.c_void_p]
    return _lib


def _p(a: np.ndarray):
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(u8p)


def elems(n: int) -> np.ndarray:
    return np.zeros((n, 2), dtype=np.uint64)


def as_elems(a) -> np.ndarray:
    a = np.ascontiguousarray(a, dtype=np.uint64)
    return a.reshape(-1, 2)


def to_int(e) -> int:
    e = np.asarray(e, dtype=np.uint64).reshape(2)
    return int(e[0]) | (int(e[1]) << 64)


def from_int(v: int) -> np.ndarray:
    return np.array([[v & ((1 << 64) - 1), v >> 64]], dtype=np.uint64)


def set_threads(n: int) -> None:
    lib().bfo_set_threads(int(n))


# ---- field ----
def mul(a, b, impl: int = 1) -> np.ndarray:
    a, b = as_elems(a), as_elems(b)
    out = elems(len(a))
    lib().bfo_mul_batch(_p(a), _p(b), _p(out), ctypes.c_size_t(len(a)), impl)
    return out


def from_canonical(a) -> np.ndarray:
    a = as_elems(a)
    out = elems(len(a))
    lib().bfo_from_canonical_batch(_p(a), _p(out), ctypes.c_size_t(len(a)))
    return out


def to_canonical(a) -> np.ndarray:
    a = as_elems(a)
    out = elems(len(a))
    lib().bfo_to_canonical_batch(_p(a), _p(out), ctypes.c_size_t(len(a)))
    return out


def inv(a) -> np.ndarray:
    a = as_elems(a)
    out = elems(len(a))
    lib().bfo_inv_batch(_p(a), _p(out), ctypes.c_size_t(len(a)))
    return out


def polymul_mod(a, b) -> np.ndarray:
    a, b = as_elems(a), as_elems(b)
    out = elems(len(a))
    lib().bfo_polymul_mod_batch(_p(a), _p(b), _p(out), ctypes.c_size_t(len(a)))
    return out


# ---- hashing ----
def keccak256(msg: bytes) -> bytes:
    buf = (ctypes.c_uint8 * 32)()
    m = (ctypes.c_uint8 * max(1, len(msg))).from_buffer_copy(msg or b"\0")
    lib().bfo_keccak256(m, ctypes.c_size_t(len(msg)), buf)
    return bytes(buf)


def sha3_256(msg: bytes) -> bytes:
    buf = (ctypes.c_uint8 * 32)()
    m = (ctypes.c_uint8 * max(1, len(msg))).from_buffer_copy(msg or b"\0")
    lib().bfo_sha3_256(m, ctypes.c_size_t(len(msg)), buf)
    return bytes(buf)


def chacha_block(rounds: int, key_words, ctr_nonce_words) -> list:
    key = (ctypes.c_uint32 * 8)(*key_words)
    cn = (ctypes.c_uint32 * 4)(*ctr_nonce_words)
    out = (ctypes.c_uint32 * 16)()
    lib().bfo_chacha_block(rounds, key, cn, out)
    return list(out)


def random_vec(seed: int, n: int) -> np.ndarray:
    out = elems(n)
    lib().bfo_random_vec(ctypes.c_uint64(seed), ctypes.c_size_t(n), _p(out))
    return out


# ---- NTT / code ----
def ntt_twiddles(log_n: int) -> np.ndarray:
    out = elems((1 << log_n) - 1 if log_n else 0)
    if log_n:
        lib().bfo_ntt_twiddles(log_n, _p(out))
    return out


def ntt_forward(log_n: int, data, width: int = 1) -> np.ndarray:
    d = as_elems(data).copy()
    lib().bfo_ntt_forward(log_n, _p(d), ctypes.c_size_t(width))
    return d


def ntt_backward(log_n: int, data, width: int = 1) -> np.ndarray:
    d = as_elems(data).copy()
    lib().bfo_ntt_backward(log_n, _p(d), ctypes.c_size_t(width))
    return d


def ntt_forward_definition(log_n: int, data) -> np.ndarray:
    d = as_elems(data)
    out = elems(1 << log_n)
    lib().bfo_ntt_forward_definition(log_n, _p(d), _p(out))
    return out


def num_queries(lam: int, log2_c: int) -> int:
    return int(lib().bfo_num_queries(lam, log2_c))


def rs_encode(log2_c: int, d: int, data, width: int = 1) -> np.ndarray:
    x = as_elems(data)
    out = elems((1 << (d + log2_c)) * width)
    lib().bfo_rs_encode(log2_c, d, _p(x), ctypes.c_size_t(width), _p(out))
    return out


def rs_fold(log2_c: int, d: int, i: int, w, r) -> np.ndarray:
    w, r = as_elems(w), as_elems(r)
    out = elems(len(w) // 2)
    lib().bfo_rs_fold(log2_c, d, i, _p(w), _p(r), _p(out))
    return out


def mle_evaluate(evals, point) -> np.ndarray:
    e, p = as_elems(evals), as_elems(point)
    out = elems(1)
    lib().bfo_mle_evaluate(_p(e), len(p), _p(p), _p(out))
    return out


def eq_table(r) -> np.ndarray:
    r = as_elems(r)
    out = elems(1 << len(r))
    lib().bfo_eq_table(_p(r), len(r), _p(out))
    return out


def merkle_root(rows: bytes, height: int, row_bytes: int) -> bytes:
    buf = (ctypes.c_uint8 * len(rows)).from_buffer_copy(rows)
    out = (ctypes.c_uint8 * 32)()
    lib().bfo_merkle_root(buf, ctypes.c_size_t(height), ctypes.c_size_t(row_bytes), out)
    return bytes(out)


class Challenger:
    """GenericChallenger::keccak256() restated (p3/src/challenger.rs:45-49)."""

    def __init__(self):
        self._h = lib().bfo_challenger_new()

    def __del__(self):
        if getattr(self, "_h", None):
            lib().bfo_challenger_free(self._h)
            self._h = None

    def observe(self, data: bytes) -> None:
        buf = (ctypes.c_uint8 * max(1, len(data))).from_buffer_copy(data or b"\0")
        lib().bfo_challenger_observe(self._h, buf, len(data))

    def sample(self) -> np.ndarray:
        out = elems(1)
        lib().bfo_challenger_sample(self._h, _p(out))
        return out

    def sample_bits(self, bits: int) -> int:
        return int(lib().bfo_challenger_sample_bits(self._h, bits))


# ---- Basefold ----
def basefold_prove(lam: int, log2_c: int, evals, point):
    """Commit + harness transcript + open; returns (root, claim, proof bytes)."""
    e, p = as_elems(evals), as_elems(point)
    n = len(p)
    assert len(e) == 1 << n
    root = (ctypes.c_uint8 * 32)()
    claim = elems(1)
    proof = ctypes.POINTER(ctypes.c_uint8)()
    plen = ctypes.c_size_t()
    rc = lib().bfo_basefold_prove(lam, log2_c, n, _p(e), _p(p), root, _p(claim),
                                  ctypes.byref(proof), ctypes.byref(plen))
    if rc != 0:
        raise RuntimeError(f"oracle prove failed ({rc})")
    data = ctypes.string_at(proof, plen.value)
    lib().bfo_free(ctypes.cast(proof, ctypes.c_void_p))
    return bytes(root), claim, data


def basefold_verify(lam: int, log2_c: int, root: bytes, point, claim, proof: bytes) -> int:
    p, c = as_elems(point), as_elems(claim)
    rb = (ctypes.c_uint8 * 32).from_buffer_copy(root)
    pb = (ctypes.c_uint8 * max(1, len(proof))).from_buffer_copy(proof or b"\0")
    return int(lib().bfo_basefold_verify(lam, log2_c, len(p), rb, _p(p), _p(c), pb,
                                         ctypes.c_size_t(len(proof))))


def basefold_prove_w(lam: int, log2_c: int, evals, width: int, point, pre: bytes = b""):
    """Width-w form: evals is the row-major 2^n x width matrix (one polynomial
    per column). `pre`: bytes the transcript observed before the harness
    prefix. Returns (root, claims (width elements), proof bytes)."""
    e, p = as_elems(evals), as_elems(point)
    n = len(p)
    assert len(e) == (1 << n) * width
    root = (ctypes.c_uint8 * 32)()
    claims = elems(width)
    proof = ctypes.POINTER(ctypes.c_uint8)()
    plen = ctypes.c_size_t()
    pb = (ctypes.c_uint8 * max(1, len(pre))).from_buffer_copy(bytes(pre) or b"\0")
    rc = lib().bfo_basefold_prove_px(lam, log2_c, n, ctypes.c_size_t(width), _p(e), _p(p), pb,
                                     ctypes.c_size_t(len(pre)), root, _p(claims), ctypes.byref(proof),
                                     ctypes.byref(plen))
    if rc != 0:
        raise RuntimeError(f"oracle prove failed ({rc})")
    data = ctypes.string_at(proof, plen.value)
    lib().bfo_free(ctypes.cast(proof, ctypes.c_void_p))
    return bytes(root), claims, data


def basefold_verify_w(lam: int, log2_c: int, width: int, root: bytes, point, claims, proof: bytes,
                      pre: bytes = b"") -> int:
    p, c = as_elems(point), as_elems(claims)
    assert len(c) == width
    rb = (ctypes.c_uint8 * 32).from_buffer_copy(root)
    pb = (ctypes.c_uint8 * max(1, len(proof))).from_buffer_copy(proof or b"\0")
    qb = (ctypes.c_uint8 * max(1, len(pre))).from_buffer_copy(bytes(pre) or b"\0")
    return int(lib().bfo_basefold_verify_px(lam, log2_c, len(p), ctypes.c_size_t(width), rb, _p(p), _p(c), qb,
                                            ctypes.c_size_t(len(pre)), pb, ctypes.c_size_t(len(proof))))


def product_sumcheck(f, g, r):
    f, g, r = as_elems(f), as_elems(g), as_elems(r)
    n = len(r)
    claim = elems(1)
    polys = elems(3 * n)
    chal = elems(n)
    fin = elems(2)
    sub = elems(1)
    lib().bfo_product_sumcheck(n, _p(f), _p(g), _p(r), _p(claim), _p(polys), _p(chal),
                               _p(fin), _p(sub))
    return claim, polys.reshape(n, 3, 2), chal, fin, sub


def eval_sumcheck(poly, point, claim):
    """prove_sumcheck over EvalProver (eq(point, .) * poly) with a fresh
    challenger. Returns (round polys (n x 3 x 2, uncompressed), challenges,
    final table value, subclaim)."""
    P, p, c = as_elems(poly), as_elems(point), as_elems(claim)
    n = len(p)
    rp, ch, fin, sub = elems(3 * max(n, 1)), elems(max(n, 1)), elems(1), elems(1)
    pp = p if n else elems(1)
    rc = lib().bfo_sumcheck_eval(n, _p(P), _p(pp), _p(c), _p(rp), _p(ch), _p(fin), _p(sub))
    if rc != 0:
        raise RuntimeError(f"oracle eval sumcheck failed ({rc})")
    return rp[:3 * n].reshape(n, 3, 2), ch[:n], fin, sub


def quadratic_sumcheck(polys, scalars, pairs):
    """polys: m tables of 2^n elements (m x 2^n x 2), scalars: one per pair,
    pairs: [(a, b), ...]. Returns (claim, round polys (n x 2), challenges,
    final evaluations (m), subclaim)."""
    P = np.ascontiguousarray(polys, dtype=np.uint64)
    m = P.shape[0]
    n = (P.shape[1]).bit_length() - 1
    s = as_elems(scalars)
    idx = np.ascontiguousarray(np.array(pairs, dtype=np.uint32).reshape(-1))
    claim, rp, ch, fin, sub = elems(1), elems(2 * n), elems(n), elems(m), elems(1)
    rc = lib().bfo_sumcheck_quadratic(n, m, _p(P.reshape(-1, 2)), len(s), _p(s),
                                      idx.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)),
                                      _p(claim), _p(rp), _p(ch), _p(fin), _p(sub))
    if rc != 0:
        raise RuntimeError(f"oracle quadratic sumcheck failed ({rc})")
    return claim, rp.reshape(n, 2, 2), ch, fin, sub


if os.environ.get("BFO_THREADS"):
    set_threads(int(os.environ["BFO_THREADS"]))
