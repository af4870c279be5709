"""MI355X Basefold prover over POLYVAL GF(2^128): Python mirror of the
reference's ``PolyCommitScheme`` / sumcheck prover interface over the C ABI
in ``include/bfd.h`` (library ``libbfd.so`` built next to this file).

Names follow the reference (han0110/learn-zkp):

* ``BinaryReedSolomonCode(lambda_, log2_c, d)`` -- binary_reed_solomon.rs:19-27
* ``Basefold.setup(BasefoldConfig(code))``      -- poly-commit/src/basefold.rs:104-110
* ``Basefold.commit / data / eval / open``      -- basefold.rs:112-232
* ``GenericChallenger.keccak256()``             -- p3/src/challenger.rs:45-49
* ``prove_sumcheck_product``                    -- degree-3 product sumcheck
                                                   driven like sumcheck/src/lib.rs:56-79

Field elements are numpy ``uint64`` arrays of shape ``(n, 2)`` (low word
first): the 16-byte little-endian Montgomery value of ``Polyval``.

There is no CPU fallback: if the library is missing, or there is no GPU, the
calls raise.
"""
from __future__ import annotations

import ctypes
import pathlib
import subprocess
from dataclasses import dataclass, field
from typing import List, Optional, Sequence, Tuple

import numpy as np

_DIR = pathlib.Path(__file__).resolve().parent
LIB_PATH = _DIR / "libbfd.so"
HEADER = _DIR.parent / "include" / "bfd.h"

_lib = None
u8p = ctypes.POINTER(ctypes.c_uint8)


# status codes of include/bfd.h
BFD_OK, BFD_ERR_ARG, BFD_ERR_HIP, BFD_ERR_BUDGET, BFD_ERR_STATE = 0, 1, 2, 3, 4
# log2_m selectors of the launch-shape calls
BFD_TUNING_BATCH, BFD_TUNING_ALL = 0xFFFFFFFE, 0xFFFFFFFF


class Tuning(ctypes.Structure):
    """bfd_tuning: how the plain path (per codeword size) or the engine cuts
    its NTT and Merkle work into launches.  No reference counterpart; every
    shape gives the same results."""
    _fields_ = [("ntt_order", ctypes.c_uint32), ("ntt_block", ctypes.c_uint32),
                ("ntt_waves_per_cu", ctypes.c_uint32), ("ntt_occupancy", ctypes.c_uint32),
                ("hash_block", ctypes.c_uint32), ("hash_waves_per_cu", ctypes.c_uint32),
                ("subtree_from_log2", ctypes.c_uint32), ("subtree_levels", ctypes.c_uint32),
                ("seg_blocks_per_cu", ctypes.c_uint32), ("reserved", ctypes.c_uint32 * 7)]
    FIELDS = ("ntt_order", "ntt_block", "ntt_waves_per_cu", "ntt_occupancy", "hash_block", "hash_waves_per_cu",
              "subtree_from_log2", "subtree_levels", "seg_blocks_per_cu")

    @classmethod
    def default(cls) -> "Tuning":
        t = cls()
        _check(lib().bfd_tuning_default(ctypes.byref(t)))
        return t

    def as_dict(self) -> dict:
        return {k: int(getattr(self, k)) for k in self.FIELDS}

    def replace(self, **kw) -> "Tuning":
        t = Tuning()
        ctypes.pointer(t)[0] = self
        for k, v in kw.items():
            if k not in self.FIELDS:
                raise KeyError(k)
            setattr(t, k, v)
        return t


class BfdError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"bfd error {code}: {msg}")
        self.code = code


def build() -> None:
    """Compile libbfd.so for gfx950 with the package Makefile (always rebuilds)."""
    subprocess.run(["make", "-C", str(_DIR), "-B", "-j8"], check=True)


def lib():
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise BfdError(-1, f"{LIB_PATH} is not built; run __graft_entry__.build()")
        # PyTorch ships its own libamdhip64.so.7.  Load it first when it is
        # installed so this library binds to the same HIP runtime: two
        # runtimes in one process cannot both own the GPU.
        try:
            import torch  # noqa: F401
        except Exception:
            pass
        L = ctypes.CDLL(str(LIB_PATH))
        L.bfd_last_error.restype = ctypes.c_char_p
        L.bfd_ctx_resident_bytes.restype = ctypes.c_uint64
        L.bfd_ctx_resident_bytes.argtypes = [ctypes.c_void_p]
        L.bfd_code_num_queries.restype = ctypes.c_uint64
        L.bfd_code_num_queries.argtypes = [ctypes.c_void_p]
        for name in ("bfd_ctx_destroy", "bfd_ctx_synchronize", "bfd_code_destroy",
                     "bfd_commitment_destroy", "bfd_challenger_destroy"):
            getattr(L, name).argtypes = [ctypes.c_void_p]
        L.bfd_ctx_create.argtypes = [ctypes.c_int, ctypes.c_uint64, ctypes.POINTER(ctypes.c_void_p)]
        L.bfd_code_create.argtypes = [ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32,
                                      ctypes.POINTER(ctypes.c_void_p)]
        L.bfd_challenger_create.argtypes = [ctypes.POINTER(ctypes.c_void_p)]
        L.bfd_challenger_clone.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_void_p)]
        L.bfd_challenger_observe.argtypes = [ctypes.c_void_p, u8p, ctypes.c_size_t]
        L.bfd_challenger_sample.argtypes = [ctypes.c_void_p, u8p]
        L.bfd_challenger_sample_bits.argtypes = [ctypes.c_void_p, ctypes.c_uint32,
                                                 ctypes.POINTER(ctypes.c_uint64)]
        L.bfd_commit.argtypes = [ctypes.c_void_p, ctypes.c_void_p, u8p, ctypes.c_size_t, ctypes.c_size_t,
                                 ctypes.POINTER(ctypes.c_void_p), u8p]
        L.bfd_commit_device.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t,
                                        ctypes.c_size_t, ctypes.POINTER(ctypes.c_void_p), u8p]
        L.bfd_commitment_root.argtypes = [ctypes.c_void_p, u8p]
        L.bfd_eval.argtypes = [ctypes.c_void_p, ctypes.c_void_p, u8p, u8p]
        L.bfd_open.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, u8p, u8p, ctypes.c_void_p,
                               ctypes.POINTER(u8p), ctypes.POINTER(ctypes.c_size_t)]
        L.bfd_proof_free.argtypes = [u8p]
        L.bfd_proof_free.restype = None
        L.bfd_proof_size.argtypes = [ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32]
        L.bfd_proof_size.restype = ctypes.c_uint64
        L.bfd_proof_size_width.argtypes = [ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint64]
        L.bfd_proof_size_width.restype = ctypes.c_uint64
        L.bfd_open_into.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, u8p, u8p, ctypes.c_void_p,
                                    u8p, ctypes.c_size_t, ctypes.POINTER(ctypes.c_size_t)]
        L.bfd_sumcheck_product.argtypes = [ctypes.c_void_p, ctypes.c_uint32, u8p, u8p, u8p, ctypes.c_void_p,
                                           u8p, u8p, u8p, u8p, u8p]
        L.bfd_sumcheck_product_device.argtypes = [ctypes.c_void_p, ctypes.c_uint32, ctypes.c_void_p, ctypes.c_void_p,
                                                  u8p, ctypes.c_void_p, u8p, u8p, u8p, u8p, u8p]
        L.bfd_sumcheck_quadratic.argtypes = [ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint32, u8p, ctypes.c_uint32,
                                             u8p, ctypes.POINTER(ctypes.c_uint32), ctypes.c_void_p, u8p, u8p, u8p,
                                             u8p, u8p]
        L.bfd_sumcheck_quadratic_device.argtypes = [ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_void_p,
                                                    ctypes.c_uint32, u8p, ctypes.POINTER(ctypes.c_uint32),
                                                    ctypes.c_void_p, u8p, u8p, u8p, u8p, u8p]
        L.bfd_verify_batch.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_size_t, u8p,
                                       ctypes.POINTER(u8p), ctypes.POINTER(u8p), ctypes.POINTER(u8p),
                                       ctypes.POINTER(ctypes.c_size_t), ctypes.POINTER(ctypes.c_void_p),
                                       ctypes.POINTER(ctypes.c_int)]
        L.bfd_commit_eval.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_size_t,
                                      ctypes.c_size_t, u8p, ctypes.POINTER(ctypes.c_void_p), ctypes.c_void_p, u8p]
        L.bfd_eval_data.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_size_t, ctypes.c_size_t,
                                    u8p, u8p]
        L.bfd_eval_data_noalloc.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_size_t, u8p,
                                            u8p]
        L.bfd_open_begin.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, u8p, u8p, u8p,
                                     ctypes.c_size_t, ctypes.POINTER(ctypes.c_void_p)]
        L.bfd_open_rounds_left.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_uint32)]
        L.bfd_open_round.argtypes = [ctypes.c_void_p, u8p, u8p]
        L.bfd_open_challenge.argtypes = [ctypes.c_void_p, u8p]
        L.bfd_open_final_poly.argtypes = [ctypes.c_void_p, u8p]
        L.bfd_open_finish.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_uint64), ctypes.c_size_t,
                                      ctypes.POINTER(u8p), ctypes.POINTER(ctypes.c_size_t)]
        L.bfd_opening_destroy.argtypes = [ctypes.c_void_p]
        L.bfd_sc_eval_create.argtypes = [ctypes.c_void_p, ctypes.c_uint32, ctypes.c_void_p, ctypes.c_int, u8p,
                                         ctypes.POINTER(ctypes.c_void_p)]
        L.bfd_sc_destroy.argtypes = [ctypes.c_void_p]
        L.bfd_field_mul_batch.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int,
                                          ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        L.bfd_field_mul.argtypes = [u8p, u8p, u8p]
        L.bfd_field_inv.argtypes = [u8p, u8p]
        L.bfd_random_vec.argtypes = [ctypes.c_uint64, ctypes.c_size_t, u8p]
        L.bfd_commit_open_batch.argtypes = [ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_size_t,
                                            ctypes.POINTER(ctypes.c_uint32), ctypes.POINTER(ctypes.c_void_p),
                                            ctypes.c_int, ctypes.POINTER(u8p), ctypes.POINTER(u8p),
                                            ctypes.POINTER(ctypes.c_void_p), ctypes.c_int, u8p,
                                            ctypes.POINTER(u8p), ctypes.POINTER(ctypes.c_size_t)]
        L.bfd_server_create.argtypes = [ctypes.c_void_p, ctypes.c_uint64, ctypes.c_uint32,
                                        ctypes.POINTER(ctypes.c_void_p)]
        L.bfd_server_submit.argtypes = [ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32,
                                        ctypes.c_void_p, u8p, ctypes.POINTER(ctypes.c_uint64)]
        L.bfd_server_wait.argtypes = [ctypes.c_void_p, ctypes.c_uint64, ctypes.c_int64, u8p, u8p,
                                      ctypes.POINTER(u8p), ctypes.POINTER(ctypes.c_size_t),
                                      ctypes.POINTER(ctypes.c_double)]
        L.bfd_server_stats.argtypes = [ctypes.c_void_p] + [ctypes.POINTER(ctypes.c_uint64)] * 4
        L.bfd_server_destroy.argtypes = [ctypes.c_void_p]
        L.bfd_server_timing.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_double)]
        L.bfd_microbench.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_size_t, ctypes.c_int,
                                     ctypes.POINTER(ctypes.c_double)]
        L.bfd_ctx_kernel_times.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_char_p, ctypes.c_size_t,
                                           ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_uint32),
                                           ctypes.c_size_t, ctypes.POINTER(ctypes.c_size_t)]
        L.bfd_ctx_kernel_filter.argtypes = [ctypes.c_void_p, ctypes.c_char_p]
        tp = ctypes.c_void_p  # bfd_tuning *
        L.bfd_tuning_default.argtypes = [tp]
        L.bfd_tuning_get.argtypes = [ctypes.c_void_p, ctypes.c_uint32, tp]
        L.bfd_tuning_set.argtypes = [ctypes.c_void_p, ctypes.c_uint32, tp]
        L.bfd_tune.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint32, tp,
                               ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double)]
        L.bfd_tune_batch.argtypes = [ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_size_t,
                                     ctypes.POINTER(ctypes.c_uint32), ctypes.c_uint32, tp,
                                     ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double)]
        L.bfd_tuning_save.argtypes = [ctypes.c_void_p, ctypes.c_char_p]
        L.bfd_tuning_load.argtypes = [ctypes.c_void_p, ctypes.c_char_p]
        _lib = L
    return _lib


def _check(rc: int) -> None:
    if rc != 0:
        raise BfdError(rc, lib().bfd_last_error().decode(errors="replace"))


def _p(a: np.ndarray):
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(u8p)


def _elems(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.uint64).reshape(-1, 2)


def random_vec(seed: int, n: int) -> np.ndarray:
    """Polyval::random_vec(n, StdRng::seed_from_u64(seed)) (from_uniform_bytes.rs:5-27)."""
    out = np.zeros((n, 2), dtype=np.uint64)
    _check(lib().bfd_random_vec(ctypes.c_uint64(seed), n, _p(out)))
    return out


# ---------------------------------------------------------------------------
class Context:
    """One GPU, one stream, an HBM budget for resident commitments (0 = none)."""

    def __init__(self, device: int = 0, hbm_budget: int = 0):
        h = ctypes.c_void_p()
        _check(lib().bfd_ctx_create(device, hbm_budget, ctypes.byref(h)))
        self.h = h

    def close(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.bfd_ctx_destroy(self.h)
        self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def synchronize(self):
        _check(lib().bfd_ctx_synchronize(self.h))

    @property
    def resident_bytes(self) -> int:
        return int(lib().bfd_ctx_resident_bytes(self.h))

    def microbench(self, what: int, n: int, iters: int) -> float:
        ms = ctypes.c_double()
        _check(lib().bfd_microbench(self.h, what, n, iters, ctypes.byref(ms)))
        return ms.value

    def kernel_filter(self, only: Optional[str] = None) -> None:
        """Time only launches named `only` (None: every kernel)."""
        _check(lib().bfd_ctx_kernel_filter(self.h, only.encode() if only else None))

    def kernel_times(self, enable: bool = True) -> dict:
        """Returns {name: (total_ms, launches)} collected since the last call
        and (re)enables or disables collection."""
        cap, nl = 64, 64
        names = ctypes.create_string_buffer(cap * nl)
        ms = (ctypes.c_double * cap)()
        ln = (ctypes.c_uint32 * cap)()
        cnt = ctypes.c_size_t()
        _check(lib().bfd_ctx_kernel_times(self.h, int(enable), names, nl, ms, ln, cap, ctypes.byref(cnt)))
        out = {}
        for i in range(cnt.value):
            nm = names.raw[i * nl:(i + 1) * nl].split(b"\0", 1)[0].decode()
            out[nm] = (ms[i], ln[i])
        return out


    # -- launch shapes (bfd_tuning_*, bfd_tune*) --
    def tuning(self, log2_m: int) -> Tuning:
        """Shape of the plain path for codewords of 2^log2_m elements, or the
        engine's (log2_m = BFD_TUNING_BATCH)."""
        t = Tuning()
        _check(lib().bfd_tuning_get(self.h, log2_m, ctypes.byref(t)))
        return t

    def set_tuning(self, log2_m: int, t: Tuning) -> None:
        _check(lib().bfd_tuning_set(self.h, log2_m, ctypes.byref(t)))

    def tune(self, pcs: "Basefold", reps: int = 3) -> Tuple[Tuning, float, float]:
        """Tunes the plain path for pcs's code; returns (chosen shape, ms of
        commit+open under the built-in shape, ms under the chosen one)."""
        t, a, b = Tuning(), ctypes.c_double(), ctypes.c_double()
        _check(lib().bfd_tune(self.h, pcs.code_h, reps, ctypes.byref(t), ctypes.byref(a), ctypes.byref(b)))
        return t, a.value, b.value

    def tune_batch(self, log_sizes: Sequence[int], log2_c: int = 1, lambda_: int = 128,
                   reps: int = 3) -> Tuple[Tuning, float, float]:
        """Tunes the engine on a batched call of polynomials of these log2
        sizes; returns (chosen shape, ms built-in, ms chosen)."""
        ls = (ctypes.c_uint32 * len(log_sizes))(*log_sizes)
        t, a, b = Tuning(), ctypes.c_double(), ctypes.c_double()
        _check(lib().bfd_tune_batch(self.h, lambda_, log2_c, len(log_sizes), ls, reps, ctypes.byref(t),
                                    ctypes.byref(a), ctypes.byref(b)))
        return t, a.value, b.value

    def save_tuning(self, path: str) -> None:
        _check(lib().bfd_tuning_save(self.h, str(path).encode()))

    def load_tuning(self, path: str) -> None:
        _check(lib().bfd_tuning_load(self.h, str(path).encode()))


_default_ctx: Optional[Context] = None


def default_context() -> Context:
    global _default_ctx
    if _default_ctx is None:
        _default_ctx = Context(0)
    return _default_ctx


class GenericChallenger:
    """GenericChallenger<Polyval, Keccak256Hash> (p3/src/challenger.rs:24-90)."""

    def __init__(self, _h=None):
        if _h is None:
            _h = ctypes.c_void_p()
            _check(lib().bfd_challenger_create(ctypes.byref(_h)))
        self.h = _h

    @classmethod
    def keccak256(cls) -> "GenericChallenger":
        return cls()

    def clone(self) -> "GenericChallenger":
        h = ctypes.c_void_p()
        _check(lib().bfd_challenger_clone(self.h, ctypes.byref(h)))
        return GenericChallenger(h)

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.bfd_challenger_destroy(self.h)
        self.h = None

    def observe(self, data: bytes) -> None:
        buf = (ctypes.c_uint8 * max(1, len(data))).from_buffer_copy(bytes(data) or b"\0")
        _check(lib().bfd_challenger_observe(self.h, buf, len(data)))

    def observe_ext_slice(self, elems) -> None:
        self.observe(_elems(elems).tobytes())

    def sample_ext_element(self) -> np.ndarray:
        out = np.zeros((1, 2), dtype=np.uint64)
        _check(lib().bfd_challenger_sample(self.h, _p(out)))
        return out

    def sample_bits(self, bits: int) -> int:
        v = ctypes.c_uint64()
        _check(lib().bfd_challenger_sample_bits(self.h, bits, ctypes.byref(v)))
        return int(v.value)

    def export_state(self) -> Tuple[bytes, bytes]:
        """(input buffer, output buffer) of the underlying HashChallenger."""
        n_in, n_out = ctypes.c_size_t(), ctypes.c_size_t()
        out = (ctypes.c_uint8 * 32)()
        rc = lib().bfd_challenger_export(self.h, None, 0, ctypes.byref(n_in), out, ctypes.byref(n_out))
        if rc != 0 and n_in.value == 0:
            _check(rc)
        buf = (ctypes.c_uint8 * max(1, n_in.value))()
        _check(lib().bfd_challenger_export(self.h, buf, n_in.value, ctypes.byref(n_in), out, ctypes.byref(n_out)))
        return bytes(buf)[:n_in.value], bytes(out)[:n_out.value]

    @classmethod
    def import_state(cls, input_buf: bytes, output_buf: bytes) -> "GenericChallenger":
        h = ctypes.c_void_p()
        ib = (ctypes.c_uint8 * max(1, len(input_buf))).from_buffer_copy(bytes(input_buf) or b"\0")
        ob = (ctypes.c_uint8 * max(1, len(output_buf))).from_buffer_copy(bytes(output_buf) or b"\0")
        _check(lib().bfd_challenger_import(ib, len(input_buf), ob, len(output_buf), ctypes.byref(h)))
        return cls(h)


@dataclass
class BinaryReedSolomonCode:
    lambda_: int
    log2_c: int
    d: int


@dataclass
class BasefoldConfig:
    code: BinaryReedSolomonCode


@dataclass
class PolyEvals:
    values: np.ndarray
    point: np.ndarray


@dataclass
class BasefoldCommitmentData:
    comm: bytes
    handle: ctypes.c_void_p
    data: np.ndarray
    _owner: "Basefold" = field(repr=False, default=None)
    width: int = 1

    def __del__(self):
        if getattr(self, "handle", None) and _lib is not None:
            _lib.bfd_commitment_destroy(self.handle)
        self.handle = None


class DeviceCode:
    """RandomFoldableCode<Polyval> of BinaryReedSolomonCode
    (code/binary_reed_solomon.rs:38-111) on the code's device tables:
    encode_batch / encode / encode0 / fold / interpolate / num_queries."""

    def __init__(self, owner: "Basefold"):
        self._owner = owner  # keeps the code handle alive
        c = owner.config.code
        self.lambda_, self.log2_c, self.d = c.lambda_, c.log2_c, c.d

    def c(self) -> int:
        return 1 << self.log2_c

    def k_d(self) -> int:
        return 1 << self.d

    def n_d(self) -> int:
        return 1 << (self.log2_c + self.d)

    def num_queries(self) -> int:
        return self._owner.num_queries

    def encode0(self, m) -> np.ndarray:
        m = _elems(m)
        return np.ascontiguousarray(np.repeat(m[:1], self.c(), axis=0))

    def encode_batch(self, m, width: int = 1) -> np.ndarray:
        """k_d x width row-major message -> n_d x width codeword (flat (rows*width, 2))."""
        arr = _elems(m)
        if len(arr) != self.k_d() * width:
            raise BfdError(BFD_ERR_ARG, "message must hold k_d x width elements")
        out = np.zeros((self.n_d() * width, 2), dtype=np.uint64)
        o = self._owner
        _check(lib().bfd_code_encode(o.ctx.h, o.code_h, _p(arr), 0, self.k_d(), width, _p(out), 0))
        return out

    def encode(self, data) -> np.ndarray:
        return self.encode_batch(data, 1)

    def encode_batch_device(self, m_ptr: int, out_ptr: int, width: int = 1) -> None:
        o = self._owner
        _check(lib().bfd_code_encode(o.ctx.h, o.code_h, ctypes.c_void_p(m_ptr), 1, self.k_d(), width,
                                     ctypes.c_void_p(out_ptr), 1))

    def fold(self, i: int, w, r) -> np.ndarray:
        """fold(i, w, r): the first half of w after the inverse butterflies and interpolation."""
        w = _elems(w)
        if len(w) != 1 << (self.log2_c + i + 1):
            raise BfdError(BFD_ERR_ARG, "w must hold 2^(log2_c + i + 1) elements")
        out = np.zeros((len(w) // 2, 2), dtype=np.uint64)
        o = self._owner
        _check(lib().bfd_code_fold(o.ctx.h, o.code_h, i, _p(w), 0, _p(_elems(r)), _p(out)))
        return out

    def interpolate(self, i: int, j: int, lo, hi, r) -> np.ndarray:
        out = np.zeros((1, 2), dtype=np.uint64)
        _check(lib().bfd_code_interpolate(self._owner.code_h, i, ctypes.c_uint64(j), _p(_elems(lo)),
                                          _p(_elems(hi)), _p(_elems(r)), _p(out)))
        return out


class Basefold:
    """PolyCommitScheme<Polyval, Polyval> for Basefold over the binary RS code."""

    def __init__(self, config: BasefoldConfig, ctx: Optional[Context] = None):
        self.config = config
        self.ctx = ctx or default_context()
        h = ctypes.c_void_p()
        c = config.code
        _check(lib().bfd_code_create(self.ctx.h, c.lambda_, c.log2_c, c.d, ctypes.byref(h)))
        self.code_h = h

    @classmethod
    def setup(cls, config: BasefoldConfig, ctx: Optional[Context] = None) -> "Basefold":
        return cls(config, ctx)

    def __del__(self):
        if getattr(self, "code_h", None) and _lib is not None:
            _lib.bfd_code_destroy(self.code_h)
        self.code_h = None

    @property
    def num_queries(self) -> int:
        return int(lib().bfd_code_num_queries(self.code_h))

    @property
    def code(self) -> "DeviceCode":
        """The RandomFoldableCode (basefold/code.rs:16-84) behind this scheme."""
        return DeviceCode(self)

    def commit(self, data, width: Optional[int] = None) -> BasefoldCommitmentData:
        """RowMajorMatrix commit (basefold.rs:112-120).  ``data`` is the
        height x width matrix as an array of shape (height, width, 2) (one
        polynomial per column), or flat with ``width`` given (default 1)."""
        arr = np.ascontiguousarray(data, dtype=np.uint64)
        if width is None:
            width = arr.shape[1] if arr.ndim == 3 else 1
        flat = arr.reshape(-1, 2)
        if width < 1 or len(flat) % width:
            raise BfdError(BFD_ERR_ARG, "data does not hold whole rows of the given width")
        root = (ctypes.c_uint8 * 32)()
        h = ctypes.c_void_p()
        _check(lib().bfd_commit(self.ctx.h, self.code_h, _p(flat), len(flat) // width, width, ctypes.byref(h), root))
        return BasefoldCommitmentData(bytes(root), h, flat, self, width)

    def commit_device(self, ptr: int, height: int, width: int = 1) -> BasefoldCommitmentData:
        """Commit a height x width row-major matrix already resident in device memory at ``ptr``."""
        root = (ctypes.c_uint8 * 32)()
        h = ctypes.c_void_p()
        _check(lib().bfd_commit_device(self.ctx.h, self.code_h, ctypes.c_void_p(ptr), height, width,
                                       ctypes.byref(h), root))
        return BasefoldCommitmentData(bytes(root), h, None, self, width)

    def commit_eval(self, data, point, width: Optional[int] = None):
        """commit followed by eval at ``point`` in one call (the claims come from
        the same stream pass). Returns (commitment data, values (width, 2))."""
        arr = np.ascontiguousarray(data, dtype=np.uint64)
        if width is None:
            width = arr.shape[1] if arr.ndim == 3 else 1
        flat = arr.reshape(-1, 2)
        if width < 1 or len(flat) % width:
            raise BfdError(BFD_ERR_ARG, "data does not hold whole rows of the given width")
        pt = _elems(point) if len(point) else np.zeros((1, 2), dtype=np.uint64)
        root = (ctypes.c_uint8 * 32)()
        h = ctypes.c_void_p()
        vals = np.zeros((width, 2), dtype=np.uint64)
        _check(lib().bfd_commit_eval(self.ctx.h, self.code_h, _p(flat), 0, len(flat) // width, width, _p(pt),
                                     ctypes.byref(h), root, _p(vals)))
        return BasefoldCommitmentData(bytes(root), h, flat, self, width), vals

    def data(self, comm_data: BasefoldCommitmentData) -> np.ndarray:
        return comm_data.data

    def eval(self, comm_data: BasefoldCommitmentData, point) -> np.ndarray:
        """One value per column (basefold.rs:126-131): shape (width, 2)."""
        point = _elems(point)
        out = np.zeros((comm_data.width, 2), dtype=np.uint64)
        _check(lib().bfd_eval(self.ctx.h, comm_data.handle, _p(point), _p(out)))
        return out

    def verify(self, comm: bytes, evals: Sequence[PolyEvals], proof: bytes, challenger: GenericChallenger,
               width: Optional[int] = None) -> int:
        """Basefold::verify (basefold.rs:234-325); returns VERIFY_OK (0) or the
        error kind (INVALID_FINAL_POLY, INVALID_QUERY, MMCS_ERROR, MALFORMED_PROOF).
        ``width`` defaults to the number of values."""
        values = _elems(evals[0].values)
        return _verify_batch(self, width or len(values), [comm], [_elems(evals[0].point)], [values], [proof],
                             [challenger])[0]

    def verify_batch(self, comms: Sequence[bytes], evals: Sequence[PolyEvals], proofs: Sequence[bytes],
                     challengers: Sequence[GenericChallenger], width: int = 1) -> List[int]:
        """Many proofs of this code (each over `width` columns) in one device pass."""
        return _verify_batch(self, width, comms, [_elems(e.point) for e in evals], [_elems(e.values) for e in evals],
                             proofs, challengers)

    def eval_data(self, data, point, width: Optional[int] = None) -> np.ndarray:
        """Basefold::eval (basefold.rs:126-131) on the data itself, as the
        trait's eval(&Data, &Point) takes it: no commitment needed.  ``data``
        as for commit; one value per column, shape (width, 2)."""
        arr = np.ascontiguousarray(data, dtype=np.uint64)
        if width is None:
            width = arr.shape[1] if arr.ndim == 3 else 1
        flat = arr.reshape(-1, 2)
        return eval_data(flat, point, width, self.ctx)

    def open(self, comm_data: BasefoldCommitmentData, evals: Sequence[PolyEvals],
             challenger: GenericChallenger) -> bytes:
        if not evals:
            return EMPTY_PROOF  # basefold.rs:139-141
        point, values = _elems(evals[0].point), _elems(evals[0].values)
        if len(values) != comm_data.width:
            raise BfdError(BFD_ERR_ARG, "open needs one value per committed column")
        proof = u8p()
        plen = ctypes.c_size_t()
        _check(lib().bfd_open(self.ctx.h, self.code_h, comm_data.handle, _p(point), _p(values),
                              challenger.h, ctypes.byref(proof), ctypes.byref(plen)))
        try:
            return ctypes.string_at(proof, plen.value)
        finally:
            lib().bfd_proof_free(proof)


    def open_stepwise(self, comm_data: BasefoldCommitmentData, evals: Sequence[PolyEvals], challenger) -> bytes:
        """Basefold::open (basefold.rs:133-232) with every transcript step on
        the caller's challenger, which the library never sees: any object with
        observe(bytes), observe_ext_slice(elems), sample_ext_element() and
        sample_bits(bits), in the reference's order (bfd_open_begin ..
        bfd_open_finish)."""
        if not evals:
            return EMPTY_PROOF  # basefold.rs:139-141
        point, values = _elems(evals[0].point), _elems(evals[0].values)
        w = comm_data.width
        if len(values) != w:
            raise BfdError(BFD_ERR_ARG, "open needs one value per committed column")
        lw = (w - 1).bit_length() if w > 1 else 0
        scal = np.zeros((max(1, lw), 2), dtype=np.uint64)
        for i in range(lw):  # sample_random_scalars_by_eq (basefold.rs:398-408)
            scal[i] = _elems(challenger.sample_ext_element())[0]
        pt = point if len(point) else np.zeros((1, 2), dtype=np.uint64)
        h = ctypes.c_void_p()
        _check(lib().bfd_open_begin(self.ctx.h, self.code_h, comm_data.handle, _p(pt), _p(values), _p(scal), lw,
                                    ctypes.byref(h)))
        try:
            left = ctypes.c_uint32()
            _check(lib().bfd_open_rounds_left(h, ctypes.byref(left)))
            root = (ctypes.c_uint8 * 32)()
            rp = np.zeros((2, 2), dtype=np.uint64)
            for _ in range(left.value):
                _check(lib().bfd_open_round(h, root, _p(rp)))
                challenger.observe(bytes(root))               # basefold.rs:356
                challenger.observe_ext_slice(rp)              # sumcheck/src/lib.rs:94
                x = np.ascontiguousarray(_elems(challenger.sample_ext_element()))
                _check(lib().bfd_open_challenge(h, _p(x)))
            fin = np.zeros((1, 2), dtype=np.uint64)
            _check(lib().bfd_open_final_poly(h, _p(fin)))
            challenger.observe_ext_slice(fin)                 # basefold.rs:194
            q = self.num_queries
            L = self.config.code.log2_c + self.config.code.d
            idx = (ctypes.c_uint64 * max(1, q))(*[challenger.sample_bits(L) for _ in range(q)])
            proof = u8p()
            plen = ctypes.c_size_t()
            _check(lib().bfd_open_finish(h, idx, q, ctypes.byref(proof), ctypes.byref(plen)))
            try:
                return ctypes.string_at(proof, plen.value)
            finally:
                lib().bfd_proof_free(proof)
        finally:
            lib().bfd_opening_destroy(h)


# bincode of BasefoldProof::default(): four empty Vec length prefixes
# (basefold.rs:49-66), what open returns for no evaluations (:139-141)
EMPTY_PROOF = bytes(32)


def eval_data(flat, point, width: int = 1, ctx: Optional[Context] = None, device_ptr: Optional[int] = None,
              height: Optional[int] = None) -> np.ndarray:
    """MLE value of every column of a row-major height x width matrix at
    ``point`` (bfd_eval_data): host data as a flat (height*width, 2) array, or
    device data at ``device_ptr`` with ``height`` given."""
    ctx = ctx or default_context()
    point = _elems(point)
    pt = point if len(point) else np.zeros((1, 2), dtype=np.uint64)
    out = np.zeros((width, 2), dtype=np.uint64)
    if device_ptr is not None:
        _check(lib().bfd_eval_data(ctx.h, ctypes.c_void_p(device_ptr), 1, height, width, _p(pt), _p(out)))
    else:
        flat = _elems(flat)
        if width < 1 or len(flat) % width:
            raise BfdError(BFD_ERR_ARG, "data does not hold whole rows of the given width")
        _check(lib().bfd_eval_data(ctx.h, ctypes.c_void_p(flat.ctypes.data), 0, len(flat) // width, width, _p(pt),
                                   _p(out)))
    return out


def field_mul_batch(a, b, shared_b: bool = False, ctx: Optional[Context] = None, wide: bool = False) -> np.ndarray:
    """Elementwise POLYVAL product on the device (bfd_field_mul_batch): a[i] * b[i],
    or a[i] * b[0] with shared_b (through b[0]'s window table); wide takes the
    four-lanes-per-product form of the transcript steps."""
    ctx = ctx or default_context()
    a, b = np.ascontiguousarray(_elems(a)), np.ascontiguousarray(_elems(b))
    out = np.zeros_like(a)
    form = 1 if shared_b else (2 if wide else 0)
    _check(lib().bfd_field_mul_batch(ctx.h, ctypes.c_void_p(a.ctypes.data), ctypes.c_void_p(b.ctypes.data),
                                     form, ctypes.c_void_p(out.ctypes.data), len(a), 0))
    return out


# results of Basefold::verify (basefold.rs:234-325) as bfd_verify_batch reports them
VERIFY_OK, INVALID_FINAL_POLY, INVALID_QUERY, MMCS_ERROR, MALFORMED_PROOF = 0, 1, 2, 3, 4


def _verify_batch(pcs: "Basefold", width: int, roots, points, values, proofs, challengers) -> List[int]:
    count = len(proofs)
    rb = np.frombuffer(b"".join(bytes(r) for r in roots) or b"\0", dtype=np.uint8).copy()
    pts = [np.ascontiguousarray(_elems(p)) if len(p) else np.zeros((1, 2), np.uint64) for p in points]
    vals = [np.ascontiguousarray(_elems(v)) for v in values]
    bufs = [bytes(pr) or b"\0" for pr in proofs]  # no copy for bytes; the library reads them in place
    P = (u8p * max(1, count))(*[_p(a) for a in pts])
    V = (u8p * max(1, count))(*[_p(a) for a in vals])
    B = (u8p * max(1, count))(*[ctypes.cast(ctypes.c_char_p(b), u8p) for b in bufs])
    lens = (ctypes.c_size_t * max(1, count))(*[len(bytes(pr)) for pr in proofs])
    chs = (ctypes.c_void_p * max(1, count))(*[c.h for c in challengers])
    res = (ctypes.c_int * max(1, count))()
    _check(lib().bfd_verify_batch(pcs.ctx.h, pcs.code_h, width, count, _p(rb), P, V, B, lens, chs, res))
    return [int(res[i]) for i in range(count)]


def commit_open_batch(lambda_: int, log2_c: int, evals: Sequence, points: Sequence, claims: Sequence,
                      challengers: Sequence[GenericChallenger], observe_prefix: bool = True,
                      device_ptrs: Optional[Sequence[int]] = None, ctx: Optional[Context] = None):
    """Commit and open many independent polynomials in one batched call.

    evals[p]: numpy (2^n_p, 2) host evaluations (or pass device_ptrs and the
    sizes through evals as ints = n_p).  Each proof equals the plain
    commit+open proof of the same polynomial with the same challenger.
    Returns (roots: list[bytes], proofs: list[numpy uint8 arrays])."""
    ctx = ctx or default_context()
    count = len(points)
    keep = []
    if device_ptrs is None:
        host = [_elems(e) for e in evals]
        keep.extend(host)
        sizes = [int(len(e)).bit_length() - 1 for e in host]
        ptrs = (ctypes.c_void_p * count)(*[h.ctypes.data for h in host])
        on_dev = 0
    else:
        sizes = [int(n) for n in evals]
        ptrs = (ctypes.c_void_p * count)(*device_ptrs)
        on_dev = 1
    pts = [_elems(p) if len(p) else np.zeros((1, 2), dtype=np.uint64) for p in points]
    cls = [_elems(c) for c in claims]
    keep.extend(pts + cls)
    sz = (ctypes.c_uint32 * count)(*sizes)
    pp = (u8p * count)(*[_p(x) for x in pts])
    cp = (u8p * count)(*[_p(x) for x in cls])
    ch = (ctypes.c_void_p * count)(*[c.h.value if isinstance(c.h, ctypes.c_void_p) else c.h for c in challengers])
    roots = np.zeros(32 * count, dtype=np.uint8)
    bufs = [np.empty(int(lib().bfd_proof_size(lambda_, log2_c, n)), dtype=np.uint8) for n in sizes]
    proofs = (u8p * count)(*[_p(b) for b in bufs])
    lens = (ctypes.c_size_t * count)(*[len(b) for b in bufs])
    _check(lib().bfd_commit_open_batch(ctx.h, lambda_, log2_c, count, sz, ptrs, on_dev, pp, cp, ch,
                                       int(observe_prefix), _p(roots), proofs, lens))
    return [roots[32 * i:32 * (i + 1)].tobytes() for i in range(count)], [b[:lens[i]] for i, b in enumerate(bufs)]


def proof_size(lambda_: int, log2_c: int, d: int, width: int = 1) -> int:
    """Exact bincode size of a Basefold proof (openings carry whole rows of ``width``)."""
    return int(lib().bfd_proof_size_width(lambda_, log2_c, d, width))


class Server:
    """Continuous-batching commit+open service (bfd_server_*): submit requests
    whose evaluations live in device memory; collect root, claim, proof and
    the submit-to-done latency with wait()."""

    def __init__(self, hbm_budget: int = 0, max_active: int = 0, ctx: Optional[Context] = None):
        self.ctx = ctx or default_context()
        h = ctypes.c_void_p()
        _check(lib().bfd_server_create(self.ctx.h, hbm_budget, max_active, ctypes.byref(h)))
        self.h = h
        self._keep = {}

    def submit(self, device_ptr: int, n: int, point, log2_c: int, lambda_: int = 128, keep=None) -> int:
        point = _elems(point) if n else np.zeros((1, 2), dtype=np.uint64)
        t = ctypes.c_uint64()
        _check(lib().bfd_server_submit(self.h, lambda_, log2_c, n, ctypes.c_void_p(device_ptr), _p(point),
                                       ctypes.byref(t)))
        self._keep[t.value] = keep
        return t.value

    def wait(self, ticket: int, timeout_ms: int = -1):
        root = (ctypes.c_uint8 * 32)()
        claim = np.zeros((1, 2), dtype=np.uint64)
        proof = u8p()
        plen = ctypes.c_size_t()
        lat = ctypes.c_double()
        _check(lib().bfd_server_wait(self.h, ticket, timeout_ms, root, _p(claim), ctypes.byref(proof),
                                     ctypes.byref(plen), ctypes.byref(lat)))
        self._keep.pop(ticket, None)
        try:
            data = ctypes.string_at(proof, plen.value)
        finally:
            lib().bfd_proof_free(proof)
        return bytes(root), claim, data, lat.value

    def stats(self) -> dict:
        v = [ctypes.c_uint64() for _ in range(4)]
        _check(lib().bfd_server_stats(self.h, *[ctypes.byref(x) for x in v]))
        return dict(zip(["pending", "active", "resident_bytes", "completed"], [x.value for x in v]))

    def timing(self) -> dict:
        v = (ctypes.c_double * 8)()
        _check(lib().bfd_server_timing(self.h, v))
        return dict(zip(["commit_ticks", "open_ticks", "commit_busy_ms", "open_busy_ms", "finish_ms",
                         "enqueue_ms", "open_job_rounds"], list(v)[:7]))

    def close(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.bfd_server_destroy(self.h)
        self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def prove_sumcheck_product(f, g, r, challenger: GenericChallenger, ctx: Optional[Context] = None,
                           device_ptrs: Optional[Tuple[int, int]] = None):
    """Degree-3 product sumcheck sum_b f(b) g(b) eq(r, b).

    With device_ptrs = (f_ptr, g_ptr) the tables are read from device memory
    (2^n elements each) and f, g are ignored.  Returns (claim, compressed
    round polys (n, 3, 2), challenges (n, 2), final [f, g] (2, 2), subclaim)."""
    ctx = ctx or default_context()
    r = _elems(r)
    n = len(r)
    claim = np.zeros((1, 2), dtype=np.uint64)
    polys = np.zeros((max(1, 3 * n), 2), dtype=np.uint64)
    chal = np.zeros((max(1, n), 2), dtype=np.uint64)
    fin = np.zeros((2, 2), dtype=np.uint64)
    sub = np.zeros((1, 2), dtype=np.uint64)
    rr = r if n else np.zeros((1, 2), dtype=np.uint64)
    if device_ptrs is not None:
        _check(lib().bfd_sumcheck_product_device(ctx.h, n, ctypes.c_void_p(device_ptrs[0]),
                                                 ctypes.c_void_p(device_ptrs[1]), _p(rr), challenger.h, _p(claim),
                                                 _p(polys), _p(chal), _p(fin), _p(sub)))
    else:
        f, g = _elems(f), _elems(g)
        _check(lib().bfd_sumcheck_product(ctx.h, n, _p(f), _p(g), _p(rr), challenger.h, _p(claim), _p(polys),
                                          _p(chal), _p(fin), _p(sub)))
    return claim, polys[:3 * n].reshape(n, 3, 2), chal[:n], fin, sub


@dataclass
class Quadratic:
    """Quadratic<Polyval, Polyval> (sumcheck/src/function/quadratic.rs:13-28):
    sum over pairs (scalar, a, b) of scalar * P_a * P_b, with its tables."""
    polys: np.ndarray                       # (m, 2^n, 2)
    pairs: Sequence                         # [(scalar (2,) uint64, a, b), ...]


def prove_sumcheck_quadratic(f: Quadratic, challenger: GenericChallenger, ctx: Optional[Context] = None,
                             device_ptr: Optional[int] = None, shape: Optional[Tuple[int, int]] = None):
    """prove_sumcheck (sumcheck/src/lib.rs:56-103) of a Quadratic function; the
    claim is the function's sum (compute_sum).

    With device_ptr the tables are read from device memory (table-major,
    shape = (m, 2^n)) and f.polys is ignored; f.pairs still gives the pairs.
    Returns (claim, compressed round polys (n, 2, 2), challenges (n, 2),
    final evaluations of every table (m, 2), subclaim)."""
    ctx = ctx or default_context()
    if device_ptr is not None:
        m, size = shape
        P = None
    else:
        P = np.ascontiguousarray(f.polys, dtype=np.uint64)
        m, size = P.shape[0], P.shape[1]
    n = size.bit_length() - 1
    if size != 1 << n:
        raise BfdError(BFD_ERR_ARG, "tables must hold 2^n evaluations")
    sc = np.ascontiguousarray(np.array([np.asarray(p[0], dtype=np.uint64).reshape(2) for p in f.pairs]))
    idx = np.ascontiguousarray(np.array([[p[1], p[2]] for p in f.pairs], dtype=np.uint32).reshape(-1))
    claim = np.zeros((1, 2), dtype=np.uint64)
    polys = np.zeros((max(1, 2 * n), 2), dtype=np.uint64)
    chal = np.zeros((max(1, n), 2), dtype=np.uint64)
    fin = np.zeros((m, 2), dtype=np.uint64)
    sub = np.zeros((1, 2), dtype=np.uint64)
    ip = idx.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32))
    if device_ptr is not None:
        _check(lib().bfd_sumcheck_quadratic_device(ctx.h, n, m, ctypes.c_void_p(device_ptr), len(sc), _p(sc), ip,
                                                   challenger.h, _p(claim), _p(polys), _p(chal), _p(fin), _p(sub)))
    else:
        _check(lib().bfd_sumcheck_quadratic(ctx.h, n, m, _p(P.reshape(-1, 2)), len(sc), _p(sc), ip, challenger.h,
                                            _p(claim), _p(polys), _p(chal), _p(fin), _p(sub)))
    return claim, polys[:2 * n].reshape(n, 2, 2), chal[:n], fin, sub


class QuadraticProver:
    """QuadraticProver (sumcheck/src/function/quadratic.rs:31-136) with its
    tables resident in device memory (bfd_sc_*): the SumcheckFunctionProver a
    prove_sumcheck driver calls once per round.  compress / decompress touch no
    table and run here."""

    def __init__(self, f: Quadratic, ctx: Optional[Context] = None):
        self.ctx = ctx or default_context()
        P = np.ascontiguousarray(f.polys, dtype=np.uint64)
        m, size = P.shape[0], P.shape[1]
        n = size.bit_length() - 1
        if size != 1 << n:
            raise BfdError(BFD_ERR_ARG, "tables must hold 2^n evaluations")
        sc = np.ascontiguousarray(np.array([np.asarray(p[0], dtype=np.uint64).reshape(2) for p in f.pairs]))
        idx = np.ascontiguousarray(np.array([[p[1], p[2]] for p in f.pairs], dtype=np.uint32).reshape(-1))
        self.m = m
        h = ctypes.c_void_p()
        _check(lib().bfd_sc_quadratic_create(self.ctx.h, n, m, _p(P.reshape(-1, 2)), 0, len(sc), _p(sc),
                                             idx.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)), ctypes.byref(h)))
        self.h = h

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.bfd_sc_destroy(self.h)
        self.h = None

    @property
    def num_vars(self) -> int:
        v = ctypes.c_uint32()
        _check(lib().bfd_sc_num_vars(self.h, ctypes.byref(v)))
        return int(v.value)

    def compute_sum(self, round: int = 0) -> np.ndarray:
        out = np.zeros((1, 2), dtype=np.uint64)
        _check(lib().bfd_sc_compute_sum(self.h, _p(out)))
        return out

    def compute_round_poly(self, round: int, subclaim) -> np.ndarray:
        out = np.zeros((3, 2), dtype=np.uint64)
        _check(lib().bfd_sc_round_poly(self.h, _p(_elems(subclaim)), _p(out)))
        return out

    @staticmethod
    def compress_round_poly(round_poly) -> np.ndarray:
        rp = _elems(round_poly)
        return np.ascontiguousarray(rp[[0, 2]])

    @staticmethod
    def decompress_round_poly(subclaim, compressed) -> np.ndarray:
        c = _elems(compressed)
        return np.ascontiguousarray(np.stack([c[0], _elems(subclaim)[0] ^ c[1], c[1]]))

    def fix_last_var(self, x) -> None:
        _check(lib().bfd_sc_fix_last_var(self.h, _p(_elems(x))))

    def evaluations(self) -> Optional[np.ndarray]:
        if self.num_vars != 0:
            return None
        out = np.zeros((self.m, 2), dtype=np.uint64)
        _check(lib().bfd_sc_evaluations(self.h, _p(out)))
        return out


class EvalProver:
    """EvalProver (sumcheck/src/function/eval.rs:27-116 over eval_impr.rs) of
    eq(point, .) * poly with the table resident in device memory (bfd_sc_*):
    the Basefold sumcheck function, one SumcheckFunctionProver call per round.
    compress keeps [c0, c2]; decompress restores c1 = subclaim + c2
    (characteristic 2, eval.rs:68-79)."""

    def __init__(self, point, poly, ctx: Optional[Context] = None, device_ptr: Optional[int] = None):
        self.ctx = ctx or default_context()
        point = _elems(point)
        n = len(point)
        pt = point if n else np.zeros((1, 2), dtype=np.uint64)
        h = ctypes.c_void_p()
        if device_ptr is not None:
            _check(lib().bfd_sc_eval_create(self.ctx.h, n, ctypes.c_void_p(device_ptr), 1, _p(pt), ctypes.byref(h)))
        else:
            P = np.ascontiguousarray(_elems(poly))
            if len(P) != 1 << n:
                raise BfdError(BFD_ERR_ARG, "poly must hold 2^n evaluations")
            _check(lib().bfd_sc_eval_create(self.ctx.h, n, ctypes.c_void_p(P.ctypes.data), 0, _p(pt),
                                            ctypes.byref(h)))
        self.h = h

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.bfd_sc_destroy(self.h)
        self.h = None

    @property
    def num_vars(self) -> int:
        v = ctypes.c_uint32()
        _check(lib().bfd_sc_num_vars(self.h, ctypes.byref(v)))
        return int(v.value)

    def compute_sum(self, round: int = 0) -> np.ndarray:
        out = np.zeros((1, 2), dtype=np.uint64)
        _check(lib().bfd_sc_compute_sum(self.h, _p(out)))
        return out

    def compute_round_poly(self, round: int, subclaim) -> np.ndarray:
        out = np.zeros((3, 2), dtype=np.uint64)
        _check(lib().bfd_sc_round_poly(self.h, _p(np.ascontiguousarray(_elems(subclaim))), _p(out)))
        return out

    @staticmethod
    def compress_round_poly(round_poly) -> np.ndarray:
        rp = _elems(round_poly)
        return np.ascontiguousarray(rp[[0, 2]])

    @staticmethod
    def decompress_round_poly(subclaim, compressed) -> np.ndarray:
        c = _elems(compressed)
        return np.ascontiguousarray(np.stack([c[0], _elems(subclaim)[0] ^ c[1], c[1]]))

    def fix_last_var(self, x) -> None:
        _check(lib().bfd_sc_fix_last_var(self.h, _p(np.ascontiguousarray(_elems(x)))))

    def evaluations(self) -> Optional[np.ndarray]:
        if self.num_vars != 0:
            return None
        out = np.zeros((1, 2), dtype=np.uint64)
        _check(lib().bfd_sc_evaluations(self.h, _p(out)))
        return out


def _horner(coeffs: np.ndarray, x: np.ndarray) -> np.ndarray:
    acc = np.zeros((1, 2), dtype=np.uint64)
    for c in coeffs[::-1]:
        acc = _mul_host(acc, x) ^ c.reshape(1, 2)
    return acc


def prove_sumcheck(prover, claim, challenger: GenericChallenger):
    """prove_sumcheck (sumcheck/src/lib.rs:56-103) over any prover with the
    SumcheckFunctionProver methods: observe the claim, then per round observe
    the compressed round poly, sample r_i, fix the last variable and reduce
    the subclaim (horner, lib.rs:39-45).  Returns (subclaim, compressed round
    polys, challenges in round order)."""
    subclaim = np.array(_elems(claim), dtype=np.uint64).reshape(1, 2)
    n = prover.num_vars
    if n == 0:
        return subclaim, [], []
    challenger.observe_ext_slice(subclaim)
    polys, chal = [], []
    for rnd in reversed(range(n)):
        rp = prover.compute_round_poly(rnd, subclaim)
        comp = prover.compress_round_poly(rp)
        challenger.observe_ext_slice(comp)
        x = challenger.sample_ext_element()
        prover.fix_last_var(x)
        subclaim = _horner(rp, x)
        polys.append(comp)
        chal.append(x)
    return subclaim, polys, chal


def batch_quadratics(fs: Sequence[Quadratic], alpha) -> Quadratic:
    """Batch::new(fs, claims, alpha) (sumcheck/src/function/batch.rs:15-34) for
    Quadratics of equal num_vars, the case that needs neither E::TWO powers nor
    halve() (both degenerate in characteristic 2): the sum of alpha^k f_k is the
    Quadratic over the union of their tables with every pair's scalar times
    alpha^k.  Its round polys, challenges and final evaluations equal the
    Batch's (evaluations() concatenates the functions' tables in order)."""
    sizes = {np.asarray(f.polys).shape[1] for f in fs}
    if len(sizes) != 1:
        raise BfdError(BFD_ERR_ARG, "batched Quadratics must have equal num_vars")
    alpha = np.asarray(alpha, dtype=np.uint64).reshape(1, 2)
    power = np.array([[1, 0xC200000000000000]], dtype=np.uint64)  # alpha^0 = ONE (Montgomery form)
    polys, pairs, base = [], [], 0
    for f in fs:
        P = np.asarray(f.polys, dtype=np.uint64)
        for (s, a, b) in f.pairs:
            pairs.append((_mul_host(np.asarray(s, dtype=np.uint64).reshape(1, 2), power)[0], base + a, base + b))
        polys.append(P)
        base += P.shape[0]
        power = _mul_host(power, alpha)
    return Quadratic(np.concatenate(polys, axis=0), pairs)


def _mul_host(a: np.ndarray, b: np.ndarray) -> np.ndarray:
    """POLYVAL Montgomery product of single elements on the host (the library's
    host multiply, through bfd_field_mul)."""
    out = np.zeros((1, 2), dtype=np.uint64)
    _check(lib().bfd_field_mul(_p(np.ascontiguousarray(a)), _p(np.ascontiguousarray(b)), _p(out)))
    return out


def header_symbols() -> List[str]:
    """Names of the functions declared in include/bfd.h."""
    import re
    text = HEADER.read_text()
    return sorted(set(re.findall(r"^\s*(?:const\s+)?[a-z_0-9]+\s*\*?\s*(bfd_[a-z_0-9]+)\s*\(", text, re.M)))
