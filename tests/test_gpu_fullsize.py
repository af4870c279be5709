"""GPU parity at the benchmark workloads' own sizes.

The CPU restatement (oracle/) proved each workload in the build container and
tests/golden/make_config_fixtures.py kept the results in
tests/golden/config_fixtures.json: roots, claims and SHA-256 digests of every
bincode proof, and every value of the 22-variable product sumcheck.  These
tests run the HIP library on the same seeded inputs and compare byte for
byte, on the plain path, the batched engine and the server:

* plain commit+open of 2^24 evaluations at rate 1/4 (two input seeds);
* the degree-3 product sumcheck over 22 variables;
* all 64 proofs of the uneven 2^12..2^22 mix at rate 1/2, batched and plain;
* requests served while others run (2^14..2^19 at rates 1/2 and 1/4 against
  the restatement, and 2^22 and 2^24 requests against the fixtures).
"""
import hashlib
import json
import pathlib
import sys
import threading
import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

FIX = json.loads((pathlib.Path(__file__).parent / "golden" / "config_fixtures.json").read_text())
LAMBDA = FIX["lambda"]


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _header_bytes(d):
    return 8 + 40 * d + 8 + 16 + 8 + 32 * d + 8


def _check_proof(entry, root, claim, proof, what):
    d = entry["log_n"]
    assert root.hex() == entry["root"], f"{what}: root"
    assert np.asarray(claim).tobytes().hex() == entry["claim"], f"{what}: claim"
    assert len(proof) == entry["proof_len"], f"{what}: proof length"
    assert hashlib.sha256(proof[:_header_bytes(d)]).hexdigest() == entry["header_sha256"], \
        f"{what}: round polynomials, final polynomial or pi-tree roots"
    assert hashlib.sha256(proof).hexdigest() == entry["proof_sha256"], f"{what}: query openings"


@pytest.fixture(scope="module")
def ctx(bfd):
    return bfd.default_context()


def _to_device(torch, e):
    return torch.from_numpy(np.ascontiguousarray(e).view(np.int64).reshape(-1)).to("cuda:0")


@pytest.mark.parametrize("which", [0, 1])
def test_plain_2e24_rate4_matches_fixture(bfd, ctx, which):
    """2^24 evaluations at rate 1/4, harness transcript: root, claim and every
    proof byte (2,235,000) equal the restatement's."""
    import torch
    e = FIX["plain_2e24_rate4"][which]
    n, log2_c = e["log_n"], e["log2_c"]
    _say(f"plain 2^{n} rate 1/{1 << log2_c} seeds {e['evals_seed']}/{e['point_seed']}")
    evals, point = bfd.random_vec(e["evals_seed"], 1 << n), bfd.random_vec(e["point_seed"], n)
    dev = _to_device(torch, evals)
    torch.cuda.synchronize()
    pcs = bfd.Basefold.setup(bfd.BasefoldConfig(bfd.BinaryReedSolomonCode(LAMBDA, log2_c, n)), ctx)
    cd = pcs.commit_device(dev.data_ptr(), 1 << n)
    claim = pcs.eval(cd, point)
    ch = bfd.GenericChallenger.keccak256()
    ch.observe(cd.comm)
    ch.observe_ext_slice(claim)
    proof = pcs.open(cd, [bfd.PolyEvals(claim, point)], ch)
    _check_proof(e, cd.comm, claim, proof, "plain 2^24")


def test_product_sumcheck_22_matches_fixture(bfd, ctx):
    """22-variable product sumcheck f g eq(r, .): claim, every round polynomial
    [c0, c2, c3], every challenge, the final f and g values and the subclaim."""
    import torch
    e = FIX["product_sumcheck_22"]
    n = e["n"]
    f, g, r = bfd.random_vec(e["f_seed"], 1 << n), bfd.random_vec(e["g_seed"], 1 << n), bfd.random_vec(e["r_seed"], n)
    for mode in ("host", "device"):
        _say(f"product sumcheck n={n} {mode}")
        if mode == "host":
            out = bfd.prove_sumcheck_product(f, g, r, bfd.GenericChallenger.keccak256(), ctx)
        else:
            df, dg = _to_device(torch, f), _to_device(torch, g)
            torch.cuda.synchronize()
            out = bfd.prove_sumcheck_product(None, None, r, bfd.GenericChallenger.keccak256(), ctx,
                                             device_ptrs=(df.data_ptr(), dg.data_ptr()))
        claim, polys, chal, fin, sub = out
        assert claim.tobytes().hex() == e["claim"]
        for k in range(n):
            assert polys[k].tobytes().hex() == e["round_polys"][k], (mode, k)
            assert chal[k].tobytes().hex() == e["challenges"][k], (mode, k)
        assert fin.tobytes().hex() == e["final_fg"]
        assert sub.tobytes().hex() == e["subclaim"]


def _mix_inputs(bfd, torch, mix):
    sizes = mix["sizes"]
    evals = [bfd.random_vec(p["evals_seed"], 1 << p["log_n"]) for p in mix["proofs"]]
    points = [bfd.random_vec(p["point_seed"], p["log_n"]) for p in mix["proofs"]]
    dev = [_to_device(torch, e) for e in evals]
    torch.cuda.synchronize()
    return sizes, evals, points, dev


def test_mix64_batched_and_plain_match_fixture(bfd, ctx):
    """The 64-polynomial uneven mix (2^12..2^22, rate 1/2): every root and every
    proof of one batched call, and of the 64 plain calls, equal the restatement's."""
    import torch
    mix = FIX["mix64_rate2"]
    log2_c = mix["log2_c"]
    sizes, evals, points, dev = _mix_inputs(bfd, torch, mix)
    assert sizes == [p["log_n"] for p in mix["proofs"]]
    claims = [np.frombuffer(bytes.fromhex(p["claim"]), dtype=np.uint64).reshape(1, 2) for p in mix["proofs"]]
    _say("mix of 64: batched")
    chs = [bfd.GenericChallenger.keccak256() for _ in sizes]
    roots, proofs = bfd.commit_open_batch(LAMBDA, log2_c, sizes, points, claims, chs, True,
                                          device_ptrs=[d.data_ptr() for d in dev], ctx=ctx)
    for i, e in enumerate(mix["proofs"]):
        _check_proof(e, roots[i], claims[i], proofs[i].tobytes(), f"batched poly {i} (2^{e['log_n']})")
    _say("mix of 64: plain")
    pcs = {n: bfd.Basefold.setup(bfd.BasefoldConfig(bfd.BinaryReedSolomonCode(LAMBDA, log2_c, n)), ctx)
           for n in sorted(set(sizes))}
    for i, e in enumerate(mix["proofs"]):
        n = e["log_n"]
        cd = pcs[n].commit_device(dev[i].data_ptr(), 1 << n)
        claim = pcs[n].eval(cd, points[i])
        ch = bfd.GenericChallenger.keccak256()
        ch.observe(cd.comm)
        ch.observe_ext_slice(claim)
        _check_proof(e, cd.comm, claim, pcs[n].open(cd, [bfd.PolyEvals(claim, points[i])], ch),
                     f"plain poly {i} (2^{n})")


def test_server_staggered_arrivals(bfd, oracle, ctx):
    """Requests arrive in waves while earlier ones are in their open rounds, under
    an HBM budget that makes later ones wait for admission: every served proof
    equals the restatement's (2^14..2^19, rates 1/2 and 1/4), a 2^24 rate-1/4
    request and two 2^22 rate-1/2 requests equal the full-size fixtures, and a
    2^23 request is accepted by the restatement's verifier."""
    import torch
    rng = np.random.default_rng(21)
    small = [(int(n), int(c)) for n, c in zip(rng.integers(14, 20, size=18), rng.choice([1, 2], size=18))]
    e24 = FIX["plain_2e24_rate4"][0]
    mix = FIX["mix64_rate2"]
    big22 = [p for p in mix["proofs"] if p["log_n"] == 22][:2]
    reqs = []  # (n, log2_c, evals (host or None), device tensor, point, expected fixture or None)
    for i, (n, c) in enumerate(small):
        ev = bfd.random_vec(9000 + i, 1 << n)
        reqs.append((n, c, ev, _to_device(torch, ev), bfd.random_vec(9500 + i, n), None))
    ev24 = bfd.random_vec(e24["evals_seed"], 1 << 24)
    reqs.insert(5, (24, 2, None, _to_device(torch, ev24), bfd.random_vec(e24["point_seed"], 24), e24))
    del ev24
    for j, p in enumerate(big22):
        ev = bfd.random_vec(p["evals_seed"], 1 << 22)
        reqs.insert(9 + 4 * j, (22, 1, None, _to_device(torch, ev), bfd.random_vec(p["point_seed"], 22), p))
    ev23 = bfd.random_vec(9999, 1 << 23)
    reqs.insert(12, (23, 1, ev23, _to_device(torch, ev23), bfd.random_vec(9998, 23), "verify"))
    torch.cuda.synchronize()

    def footprint(n, c):
        # an upper bound of what a request holds: codeword 16M, main tree 64M,
        # pi vectors 16M, pi trees 64M, table and halved eq table 16N, and
        # the transcript region
        N = 1 << n
        M = N << c
        return 160 * M + 16 * N + (1 << 20)

    budget = footprint(24, 2) + footprint(22, 1)  # the 2^24 request leaves room for little else
    srv = bfd.Server(hbm_budget=budget, ctx=ctx)
    tickets = [None] * len(reqs)
    errors = []

    def client():
        try:
            for w0 in range(0, len(reqs), 5):  # waves of five, 30 ms apart
                for i in range(w0, min(w0 + 5, len(reqs))):
                    n, c, _, d, pt, _ = reqs[i]
                    tickets[i] = srv.submit(d.data_ptr(), n, pt, c, keep=d)
                time.sleep(0.03)
        except Exception as e:  # noqa: BLE001 - surfaced in the main thread
            errors.append(e)

    th = threading.Thread(target=client)
    th.start()
    got = []
    deadline = time.time() + 120
    for i in range(len(reqs)):
        while tickets[i] is None:
            assert not errors, f"submit failed: {errors[0]}"
            assert time.time() < deadline, "client thread stalled"
            time.sleep(0.001)
        got.append(srv.wait(tickets[i], 120000))
        _say(f"served {i}: 2^{reqs[i][0]} rate 1/{1 << reqs[i][1]} in {got[-1][3]:.1f} ms")
    th.join()
    srv.close()
    for i, ((n, c, ev, _, pt, fx), (root, claim, proof, lat)) in enumerate(zip(reqs, got)):
        _say(f"checking served {i}: 2^{n} rate 1/{1 << c}")
        assert lat > 0
        if fx == "verify":
            assert (claim == oracle.mle_evaluate(ev, pt)).all(), i
            assert oracle.basefold_verify(LAMBDA, c, root, pt, claim, proof) == 0, i
        elif fx is not None:
            _check_proof(fx, root, claim, proof, f"served 2^{n}")
        else:
            root_o, claim_o, proof_o = oracle.basefold_prove(LAMBDA, c, ev, pt)
            assert root == root_o and (claim == claim_o).all(), (i, n, c)
            assert proof == proof_o, (i, n, c)
