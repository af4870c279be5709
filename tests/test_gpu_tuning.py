"""Launch shapes and the autotuner: every shape the tuner can choose gives
the same commitments and proofs as the built-in shape and the CPU
restatement, on the plain path and in the batched engine; the tuner
installs a valid shape and the shapes survive a save/load round trip.

Sizes are chosen so the shapes change the launch plan: 2^15 evaluations at
rate 1/4 give a 2^17-leaf main tree, so Merkle thresholds 2^14..2^16 move
levels between per-level and LDS-subtree launches, and n = 15 or 16 covers
both remainders of the three-layer NTT passes (14 = 4*3 + 2, 15 = 5*3).
"""
import sys

import pytest

pytestmark = pytest.mark.gpu

# values exercised per field: the tuner's range and the extremes the ABI
# accepts
CANDIDATES = {
    "ntt_order": [0, 1, 2],
    "ntt_occupancy": [2, 3],
    "ntt_block": [64, 128, 256],
    "ntt_waves_per_cu": [4, 8, 64],
    "hash_block": [64, 128],
    "hash_waves_per_cu": [4, 16, 128],
    "subtree_from_log2": [1, 14, 15, 16, 18],
    "subtree_levels": [1, 6, 9],
}


# the values bfd_tune tries (any other field keeps its built-in value)
TUNER = {
    "ntt_order": [0, 1, 2],
    "ntt_occupancy": [2, 3],
    "ntt_block": [64, 128, 256],
    "ntt_waves_per_cu": [8, 16, 32, 64],
    "hash_block": [64, 128, 256],
    "hash_waves_per_cu": [16, 32, 64, 128],
    "subtree_from_log2": [14, 15, 16],
    "subtree_levels": [6, 7, 8, 9],
    "seg_blocks_per_cu": [],
}


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def tctx(bfd):
    # a context of its own: shapes are per context, the other tests keep
    # the built-in ones
    c = bfd.Context(0)
    yield c
    c.close()


def prove(bfd, pcs, evals, point):
    cd = pcs.commit(evals)
    claim = pcs.eval(cd, point)
    ch = bfd.GenericChallenger.keccak256()
    ch.observe(cd.comm)
    ch.observe_ext_slice(claim)
    return cd.comm, claim, pcs.open(cd, [bfd.PolyEvals(claim, point)], ch)


@pytest.mark.parametrize("n,log2_c", [(15, 2), (16, 1)])
def test_every_plain_shape_gives_the_same_proof(bfd, oracle, tctx, n, log2_c):
    evals, point = bfd.random_vec(300 + n, 1 << n), bfd.random_vec(400 + n, n)
    pcs = bfd.Basefold.setup(bfd.BasefoldConfig(bfd.BinaryReedSolomonCode(128, log2_c, n)), tctx)
    l = n + log2_c
    base = bfd.Tuning.default()
    tctx.set_tuning(l, base)
    ref = prove(bfd, pcs, evals, point)
    root_o, claim_o, proof_o = oracle.basefold_prove(128, log2_c, evals, point)
    assert ref[0] == root_o and (ref[1] == claim_o).all() and ref[2] == proof_o
    try:
        for field, values in CANDIDATES.items():
            for v in values:
                _say(f"shape n={n} log2_c={log2_c} {field}={v}")
                tctx.set_tuning(l, base.replace(**{field: v}))
                got = prove(bfd, pcs, evals, point)
                assert got[0] == ref[0] and (got[1] == ref[1]).all() and got[2] == ref[2], (field, v)
        # the codeword alone (RandomFoldableCode::encode_batch) under a
        # non-default plan, width 1 and 3
        tctx.set_tuning(l, base.replace(ntt_order=2, ntt_block=64, ntt_occupancy=3))
        for w in (1, 3):
            m = bfd.random_vec(500 + w, (1 << n) * w)
            got = pcs.code.encode_batch(m, w)
            tctx.set_tuning(l, base)
            want = pcs.code.encode_batch(m, w)
            assert (got == want).all(), w
            tctx.set_tuning(l, base.replace(ntt_order=2, ntt_block=64, ntt_occupancy=3))
    finally:
        tctx.set_tuning(l, base)


def test_invalid_shapes_rejected(bfd, tctx):
    base = bfd.Tuning.default()
    bad = [dict(ntt_order=3), dict(ntt_block=96), dict(ntt_block=512), dict(ntt_waves_per_cu=0),
           dict(ntt_occupancy=1), dict(hash_block=32), dict(hash_waves_per_cu=5000), dict(subtree_from_log2=0),
           dict(subtree_levels=0), dict(subtree_levels=10), dict(seg_blocks_per_cu=0)]
    for kw in bad:
        with pytest.raises(bfd.BfdError) as e:
            tctx.set_tuning(20, base.replace(**kw))
        assert e.value.code == bfd.BFD_ERR_ARG, kw
    t = base.replace()
    t.reserved[0] = 1
    with pytest.raises(bfd.BfdError):
        tctx.set_tuning(20, t)
    with pytest.raises(bfd.BfdError):
        tctx.set_tuning(48, base)
    assert tctx.tuning(20).as_dict() == base.as_dict()


def test_every_engine_shape_gives_the_same_proofs(bfd, oracle, tctx):
    sizes = [9, 14, 3, 15, 12, 1, 10]
    log2_c = 1
    evals = [bfd.random_vec(600 + i, 1 << n) for i, n in enumerate(sizes)]
    points = [bfd.random_vec(700 + i, n) for i, n in enumerate(sizes)]
    claims = [oracle.mle_evaluate(e, p) for e, p in zip(evals, points)]
    want = [oracle.basefold_prove(128, log2_c, e, p) for e, p in zip(evals, points)]
    base = bfd.Tuning.default()
    engine = {"ntt_order": [1, 2], "subtree_from_log2": [1, 12, 15, 19], "subtree_levels": [1, 7],
              "seg_blocks_per_cu": [1, 2, 32]}
    try:
        for field, values in engine.items():
            for v in values:
                _say(f"engine shape {field}={v}")
                tctx.set_tuning(bfd.BFD_TUNING_BATCH, base.replace(**{field: v}))
                chs = [bfd.GenericChallenger.keccak256() for _ in sizes]
                roots, proofs = bfd.commit_open_batch(128, log2_c, evals, points, claims, chs, True, ctx=tctx)
                for i in range(len(sizes)):
                    assert roots[i] == want[i][0], (field, v, i)
                    assert proofs[i].tobytes() == want[i][2], (field, v, i)
    finally:
        tctx.set_tuning(bfd.BFD_TUNING_BATCH, base)


def test_tuner_installs_a_valid_shape_and_proofs_hold(bfd, oracle, tctx, tmp_path):
    n, log2_c = 15, 2
    pcs = bfd.Basefold.setup(bfd.BasefoldConfig(bfd.BinaryReedSolomonCode(128, log2_c, n)), tctx)
    _say("tune plain n=15 log2_c=2")
    chosen, ms_def, ms_tuned = tctx.tune(pcs, reps=1)
    assert ms_def > 0 and ms_tuned > 0
    assert tctx.tuning(n + log2_c).as_dict() == chosen.as_dict()
    default = bfd.Tuning.default().as_dict()
    for k, v in chosen.as_dict().items():
        assert v == default[k] or v in TUNER[k], (k, v)
    evals, point = bfd.random_vec(800, 1 << n), bfd.random_vec(801, n)
    got = prove(bfd, pcs, evals, point)
    root_o, claim_o, proof_o = oracle.basefold_prove(128, log2_c, evals, point)
    assert got[0] == root_o and got[2] == proof_o
    _say("tune engine")
    bchosen, bdef, btuned = tctx.tune_batch([8, 12, 14, 10], log2_c=1, reps=1)
    assert bdef > 0 and btuned > 0
    assert tctx.tuning(bfd.BFD_TUNING_BATCH).as_dict() == bchosen.as_dict()
    # save / load round trip into a fresh context
    path = tmp_path / "shapes.txt"
    # make sure at least one entry differs from the built-in shape
    tctx.set_tuning(7, bfd.Tuning.default().replace(ntt_order=1, subtree_levels=6))
    tctx.save_tuning(path)
    text = path.read_text()
    assert "log2_m=7 " in text
    other = bfd.Context(0)
    try:
        other.load_tuning(path)
        for l in list(range(0, 48)):
            assert other.tuning(l).as_dict() == tctx.tuning(l).as_dict(), l
        assert other.tuning(bfd.BFD_TUNING_BATCH).as_dict() == tctx.tuning(bfd.BFD_TUNING_BATCH).as_dict()
        bad = tmp_path / "bad.txt"
        bad.write_text("log2_m=7 ntt_order=9\n")
        with pytest.raises(bfd.BfdError):
            other.load_tuning(bad)
        bad.write_text("log2_m=7 bogus=1\n")
        with pytest.raises(bfd.BfdError):
            other.load_tuning(bad)
        assert other.tuning(7).as_dict() == tctx.tuning(7).as_dict()
    finally:
        other.close()
        tctx.set_tuning(bfd.BFD_TUNING_ALL, bfd.Tuning.default())
        tctx.set_tuning(bfd.BFD_TUNING_BATCH, bfd.Tuning.default())
