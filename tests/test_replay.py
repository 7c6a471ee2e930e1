"""The oracle's proof orchestration pinned by an independent replay.

tests/replay.py re-derives every challenge and query set from a proof's bytes
with hashlib and checks roots, openings, running sums, triples, yr and the
final sumcheck against values it computes from the input (its own field
arithmetic, folds, basis polynomials, BLAKE3 and parser).  Here its pieces
are pinned first (textbook field products, the subspace-polynomial
definition, the published BLAKE3 vectors), then it is run over the oracle's
proofs, including explicit two-step configurations that exercise the glue
step with materialised basis polynomials, and shown to reject tampered
proofs.  The configuration shapes are checked against the literals of
config.rs."""
import json
import os
import random

import numpy as np
import pytest

import gfref
import ligerito_hip as lh
import replay as RP

# config.rs:6-67 (cfg20) and :69-134 (cfg24), typed from the reference
CFG_LITERALS = {
    20: {"recursive_steps": 1, "initial_dims": (1 << 14, 1 << 6), "initial_k": 6, "dims": [(1 << 10, 1 << 4)],
         "ks": [4]},
    24: {"recursive_steps": 2, "initial_dims": (1 << 18, 1 << 6), "initial_k": 6,
         "dims": [(1 << 14, 1 << 4), (1 << 10, 1 << 4)], "ks": [4, 4]},
}
VERIFIER_LITERALS = {20: (14, [10]), 24: (18, [14, 10])}  # (initial_dim, log_dims)

# two recursive steps at 2^16 (not a reference configuration): the glue
# step with basis polynomials small enough to materialise
TWO_STEP_16 = {"recursive_steps": 2, "initial_dims": (1 << 10, 1 << 6), "initial_k": 6,
               "dims": [(1 << 8, 1 << 2), (1 << 6, 1 << 2)], "ks": [2, 2]}
# three steps, other fold sizes
THREE_STEP_17 = {"recursive_steps": 3, "initial_dims": (1 << 12, 1 << 5), "initial_k": 5,
                 "dims": [(1 << 9, 1 << 3), (1 << 7, 1 << 2), (1 << 6, 1 << 1)], "ks": [3, 2, 1]}


def oracle_commit(O, bits):
    """commit callback for replay(): the oracle's ligero_commit."""

    def commit(v, m, n):
        arr = v.astype(np.uint32) if bits == 32 else v
        root, mat, _ = O.ligero_commit(arr, m, n, bits, want_matrix=True)
        M = 4 * m
        mm = mat.reshape(n, M) if bits == 32 else mat.reshape(n, M, 2)

        def rows(q):
            if bits == 32:
                return [int(x) for x in mm[:, q]]
            return [int(a) | (int(b) << 64) for a, b in mm[:, q]]

        return root, rows

    return commit


# ------------------------------------------------------- replay's pieces --
@pytest.mark.parametrize("bits", [32, 128])
def test_replay_field_is_the_textbook_field(bits):
    F = RP.Field(bits)
    rng = random.Random(bits)
    xs = [rng.getrandbits(bits) for _ in range(64)] + [0, 1, (1 << bits) - 1]
    ys = [rng.getrandbits(bits) for _ in range(len(xs))]
    want = [gfref.mul(x, y, bits) for x, y in zip(xs, ys)]
    assert [F.mul(x, y) for x, y in zip(xs, ys)] == want
    assert F.to_ints(F.vmul(F.vec(xs), F.vec(ys))) == want
    s = ys[3]
    assert F.to_ints(F.vmul_s(F.vec(xs), s)) == [gfref.mul(x, s, bits) for x in xs]
    a = xs[5]
    assert F.mul(a, F.inv(a)) == 1


@pytest.mark.parametrize("bits", [32, 128])
def test_replay_subspace_values_match_the_definition(bits):
    # s_t(x) = prod over u in span{1, 2, ..., 2^(t-1)} of (x + u)
    F = RP.Field(bits)
    sk = RP.sks_vks(F, 7)
    for t in range(7):
        assert sk[t] == gfref.subspace_poly(t, 1 << t, bits)
    for x in (5, 1234, 4095):
        s = RP.sks_at(F, sk, x, 7)
        assert s == [gfref.subspace_poly(t, x, bits) for t in range(7)]


def test_replay_blake3_matches_published_vectors(golden_dir):
    vec = json.load(open(os.path.join(golden_dir, "blake3_vectors.json")))["vectors"]
    for n, hexd in vec.items():
        data = bytes(i % 251 for i in range(int(n)))
        assert RP.blake3(data).hex() == hexd, n


def test_replay_merkle_matches_reference_kat(golden_dir):
    # the batched proof order of merkle_tree.rs:98-188 on the oracle's tree,
    # plus the reference's own two-query node list (merkle_tree.rs:214-235)
    import oracle as O

    rng = random.Random(3)
    leaves = [bytes(rng.getrandbits(8) for _ in range(40)) for _ in range(64)]
    layers = O.build_merkle_tree(leaves)
    for q in ([0, 2], [0, 1, 2], [5, 6, 7, 30, 63], list(range(0, 64, 3))):
        proof = O.merkle_prove(leaves, q)
        assert RP.merkle_verify(layers[-1], proof, 6, [leaves[i] for i in q], q)
        assert not RP.merkle_verify(layers[-1], proof[::-1], 6, [leaves[i] for i in q], q) or len(proof) < 2
        bad = [bytearray(x) for x in [leaves[i] for i in q]]
        bad[0][0] ^= 1
        assert not RP.merkle_verify(layers[-1], proof, 6, [bytes(b) for b in bad], q)


# ------------------------------------------------------------- configs --
@pytest.mark.parametrize("L", [20, 24])
def test_generated_configs_are_config_rs(O, L):
    assert lh.config_shape(L) == CFG_LITERALS[L]
    assert O.config_shape(L) == CFG_LITERALS[L]
    init_dim, log_dims = VERIFIER_LITERALS[L]
    sh = CFG_LITERALS[L]
    assert sh["initial_dims"][0] == 1 << init_dim
    assert [d[0] for d in sh["dims"]] == [1 << x for x in log_dims]


@pytest.mark.parametrize("L", range(16, 28))
def test_generated_configs_satisfy_the_prover_invariants(O, L):
    sh = lh.config_shape(L)
    assert sh == O.config_shape(L)
    m0, n0 = sh["initial_dims"]
    assert m0 * n0 == 1 << L
    assert n0 == 1 << sh["initial_k"]
    assert m0 == sh["dims"][0][0] * sh["dims"][0][1]
    for i, (m, n) in enumerate(sh["dims"]):
        assert n == 1 << sh["ks"][i]
        assert 4 * m >= 148
        if i + 1 < len(sh["dims"]):
            assert m == sh["dims"][i + 1][0] * sh["dims"][i + 1][1]
    assert sh["recursive_steps"] == len(sh["dims"]) >= 1


def test_config_shape_rejects_unsupported_sizes():
    for L in (15, 28):
        with pytest.raises(lh.LigeritoError) as e:
            lh.config_shape(L)
        assert e.value.code == -1


# ----------------------------------------------------- replay the oracle --
@pytest.mark.parametrize("L,bits", [(16, 32), (16, 128), (20, 32)])
def test_replay_accepts_oracle_proofs(O, L, bits):
    poly = O.splitmix(1, 1 << L, bits)
    pr = O.prove(poly, L, bits)
    log = RP.replay(pr, poly, O.config_shape(L), bits, oracle_commit(O, bits))
    assert len(log["queries"]) == 2 and all(len(q) == 148 for q in log["queries"])


@pytest.mark.parametrize("shape,bits", [(TWO_STEP_16, 32), (TWO_STEP_16, 128), (THREE_STEP_17, 32)])
def test_replay_accepts_oracle_multi_step_proofs(O, shape, bits):
    n = shape["initial_dims"][0] * shape["initial_dims"][1]
    poly = O.splitmix(8, n, bits)
    pr = O.prove_dims(poly, shape, bits, fs_seed=77)
    assert O.verify_dims(pr, shape, bits, fs_seed=77) == 1
    log = RP.replay(pr, poly, shape, bits, oracle_commit(O, bits), fs_seed=77)
    assert len(log["queries"]) == shape["recursive_steps"] + 1


def test_replay_oracle_cfg24_structure(O):
    # cfg24: two recursive roots, yr of 2^10 entries, 10 triples
    poly = O.splitmix(2, 1 << 24, 32)
    pr = O.prove(poly, 24, 32)
    p = RP.parse_proof(pr, 32)
    assert len(p["rec_roots"]) == 2 and len(p["rec_proofs"]) == 1
    assert len(p["yr"]) == 1 << 10
    assert len(p["tr"]) == 1 + 4 + 1 + 4
    assert len(p["initial"]["rows"]) == 148 and len(p["initial"]["rows"][0]) == 64
    assert len(p["final"]["rows"][0]) == 16
    RP.replay(pr, poly, O.config_shape(24), 32, oracle_commit(O, 32))


def test_replay_rejects_tampered_proofs(O):
    L, bits = 16, 32
    poly = O.splitmix(1, 1 << L, bits)
    pr = O.prove(poly, L, bits)
    sh = O.config_shape(L)
    commit = oracle_commit(O, bits)
    with pytest.raises(RP.ReplayError):
        RP.replay(pr, poly, sh, bits, commit, fs_seed=1235)
    other = O.splitmix(2, 1 << L, bits)
    with pytest.raises(RP.ReplayError):
        RP.replay(pr, other, sh, bits, commit)
    p = RP.parse_proof(pr, bits)
    # offsets: header 12, root 32, initial RLP (16 + rows + 8 + nodes), ...
    rows_at = 12 + 32 + 16
    nodes_at = rows_at + 148 * 64 * 4 + 8
    tr_at = len(pr) - 3 * 4 * len(p["tr"])
    for pos in (20, rows_at + 4 * 64 * 7 + 3, nodes_at + 32 * 5 + 1, tr_at + 5, tr_at + 3 * 4 * 4 + 1):
        bad = bytearray(pr)
        bad[pos] ^= 0x04
        with pytest.raises(RP.ReplayError):
            RP.replay(bytes(bad), poly, sh, bits, commit)
