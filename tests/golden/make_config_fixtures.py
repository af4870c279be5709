"""Writes the full-size parity fixtures of the benchmark workloads.

The CPU restatement (oracle/) proves each workload at its own size here, in
the build container, and this script keeps what a GPU test needs to compare
byte for byte: roots, claims and SHA-256 digests of the bincode proofs
(whole proof and header, i.e. the round polynomials, final polynomial and
pi-tree roots), and every value of the product sumcheck.  The GPU tests read
``config_fixtures.json`` and never run the restatement at these sizes.

Workloads (inputs are ``Polyval::random_vec`` streams of StdRng seeds, the
same bytes on both sides):
  * plain 2^24 at rate 1/4 (log2_c 2): evaluations seed 0, point seed 1 (the
    benchmark's inputs), and evaluations seed 2, point seed 3;
  * product sumcheck f g eq(r, .) over 22 variables: f seed 0, g seed 1,
    r seed 2, fresh challenger (claim, 22 x [c0, c2, c3], challenges, final
    f and g, subclaim);
  * the 64-polynomial uneven mix: log2 sizes numpy default_rng(0).integers(
    12, 23, 64), rate 1/2, evaluations seed 10 + i, point seed 1000 + i.
Every proof follows the harness transcript (observe root, observe claim)
of poly-commit/src/lib.rs:77-100.

Run from the repository root (about three minutes on 8 cores):
    python tests/golden/make_config_fixtures.py
"""
import hashlib
import json
import os
import pathlib
import sys
import time

import numpy as np

ROOT = pathlib.Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))
OUT = pathlib.Path(__file__).with_name("config_fixtures.json")
LAMBDA = 128


def header_bytes(d: int) -> int:
    """bincode bytes of BasefoldProof before its query records: d compressed
    round polynomials of two elements, the final polynomial, d roots and the
    openings' length prefix (basefold.rs:49-66)."""
    return 8 + 40 * d + 8 + 16 + 8 + 32 * d + 8


def proof_entry(O, log2_c, evals, point, tag):
    t0 = time.time()
    root, claim, proof = O.basefold_prove(LAMBDA, log2_c, evals, point)
    d = len(point)
    print(f"  {tag}: {len(proof)} B in {time.time() - t0:.1f} s", flush=True)
    return {"root": root.hex(), "claim": claim.tobytes().hex(), "proof_len": len(proof),
            "proof_sha256": hashlib.sha256(proof).hexdigest(),
            "header_sha256": hashlib.sha256(proof[:header_bytes(d)]).hexdigest()}


def main() -> None:
    import oracle as O
    O.set_threads(int(os.environ.get("BFO_THREADS", os.cpu_count() or 1)))
    fx = {"lambda": LAMBDA}

    print("plain 2^24, rate 1/4", flush=True)
    fx["plain_2e24_rate4"] = []
    for es, ps in [(0, 1), (2, 3)]:
        n = 24
        e = proof_entry(O, 2, O.random_vec(es, 1 << n), O.random_vec(ps, n), f"seeds {es}/{ps}")
        e.update({"log_n": n, "log2_c": 2, "evals_seed": es, "point_seed": ps})
        fx["plain_2e24_rate4"].append(e)

    print("product sumcheck, 22 variables", flush=True)
    n = 22
    f, g, r = O.random_vec(0, 1 << n), O.random_vec(1, 1 << n), O.random_vec(2, n)
    claim, polys, chal, fin, sub = O.product_sumcheck(f, g, r)
    fx["product_sumcheck_22"] = {
        "n": n, "f_seed": 0, "g_seed": 1, "r_seed": 2, "claim": claim.tobytes().hex(),
        "round_polys": [polys[k].tobytes().hex() for k in range(n)],
        "challenges": [chal[k].tobytes().hex() for k in range(n)],
        "final_fg": fin.tobytes().hex(), "subclaim": sub.tobytes().hex()}

    print("64-polynomial mix, rate 1/2", flush=True)
    sizes = [int(v) for v in np.random.default_rng(0).integers(12, 23, size=64)]
    mix = []
    for i, n in enumerate(sizes):
        e = proof_entry(O, 1, O.random_vec(10 + i, 1 << n), O.random_vec(1000 + i, n), f"mix {i} 2^{n}")
        e.update({"log_n": n, "evals_seed": 10 + i, "point_seed": 1000 + i})
        mix.append(e)
    fx["mix64_rate2"] = {"log2_c": 1, "sizes": sizes, "proofs": mix}

    OUT.write_text(json.dumps(fx, indent=1) + "\n")
    print(f"wrote {OUT}")


if __name__ == "__main__":
    main()
