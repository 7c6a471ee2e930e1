"""Pins the oracle's multilinear fold, basis induction, sumcheck and the
end-to-end prover/verifier (src/multilinear_poly.rs, src/sumcheck_polys.rs,
src/ligerito.rs, src/prover.rs, src/verifier.rs, src/config.rs:136-239)."""
import hashlib
import json
import os
import random

import numpy as np
import pytest

import gfref


def ints(a):
    return [int(x) for x in np.asarray(a).reshape(-1)]


def test_lagrange_basis_is_tensor(O):
    rng = random.Random(2)
    rs = [rng.getrandbits(32) for _ in range(4)]
    gr = ints(O.lagrange_basis(np.array(rs, dtype=np.uint32), 32))
    for idx in range(16):
        want = 1
        for i, r in enumerate(rs):  # rs[0] is the most significant bit
            bit = (idx >> (3 - i)) & 1
            want = gfref.mul(want, r if bit else r ^ 1, 32)
        assert gr[idx] == want


def test_partial_eval_is_row_combination(O):
    # partial_eval(rs) over the top k variables == sum_j gr[j] * column_j
    rng = random.Random(4)
    k, m = 3, 32
    e = [rng.getrandbits(32) for _ in range(m << k)]
    rs = [rng.getrandbits(32) for _ in range(k)]
    got = ints(O.partial_eval(np.array(e, dtype=np.uint32), np.array(rs, dtype=np.uint32), 32))
    gr = ints(O.lagrange_basis(np.array(rs, dtype=np.uint32), 32))
    for i in range(m):
        want = 0
        for j in range(1 << k):
            want ^= gfref.mul(gr[j], e[j * m + i], 32)
        assert got[i] == want


def test_eval_013_product(O):
    rng = random.Random(6)
    n = 16
    f = [rng.getrandbits(32) for _ in range(n)]
    g = [rng.getrandbits(32) for _ in range(n)]
    got = ints(O.eval_013_product(np.array(f, dtype=np.uint32), np.array(g, dtype=np.uint32), 32))
    h = n // 2
    dot = lambda a, b: __import__("functools").reduce(lambda x, y: x ^ y, [gfref.mul(p, q, 32) for p, q in zip(a, b)], 0)
    at3 = lambda lo, hi: [gfref.mul(3, y, 32) ^ gfref.mul(2, x, 32) for x, y in zip(lo, hi)]
    assert got[0] == dot(f[:h], g[:h])
    assert got[1] == dot(f[h:], g[h:])
    assert got[2] == dot(at3(f[:h], f[h:]), at3(g[:h], g[h:]))


def test_induce_matches_basis_brute_force(O):
    rng = random.Random(8)
    n, k, s = 6, 2, 5
    rows = [rng.getrandbits(32) for _ in range(s << k)]
    v = [rng.getrandbits(32) for _ in range(k)]
    q = sorted(rng.sample(range(64), s))
    alpha = rng.getrandbits(32)
    basis, enf = O.induce(n, np.array(rows, dtype=np.uint32), np.array(v, dtype=np.uint32), q, alpha, 32)
    basis = ints(basis)
    gr = ints(O.lagrange_basis(np.array(v, dtype=np.uint32), 32))
    want_b = [0] * (1 << n)
    want_e = 0
    ai = 1
    for i in range(s):
        W = [gfref.subspace_poly(t, q[i], 32) for t in range(n)]
        for j in range(1 << n):
            bj = ai
            for t in range(n):
                if (j >> t) & 1:
                    bj = gfref.mul(bj, W[t], 32)
            want_b[j] ^= bj
        dot = 0
        for j in range(1 << k):
            dot ^= gfref.mul(rows[i * (1 << k) + j], gr[j], 32)
        want_e ^= gfref.mul(dot, ai, 32)
        ai = gfref.mul(ai, alpha, 32)
    assert basis == want_b and enf == want_e


def test_ligero_commit_layout(O):
    # ligero.rs:11-68: column j = poly[j*m..], zero padded to 4m, encoded;
    # leaf i = row i bytes (LE), root = Merkle root.
    m, n = 16, 4
    poly = O.splitmix(3, m * n, 32)
    root, mat, layers = O.ligero_commit(poly, m, n, 32, want_matrix=True, want_layers=True)
    cols = []
    for j in range(n):
        buf = np.zeros(4 * m, dtype=np.uint32)
        buf[:m] = poly[j * m:(j + 1) * m]
        cols.append(O.rs_encode_non_systematic(buf, m, 32))
    assert np.array_equal(mat, np.concatenate(cols))
    leaves = [np.array([cols[j][i] for j in range(n)], dtype="<u4").tobytes() for i in range(4 * m)]
    assert layers == O.build_merkle_tree(leaves) and root == layers[-1]


@pytest.mark.parametrize("bits", [32, 128])
def test_prove_verify_L16(O, bits):
    poly = O.splitmix(1, 1 << 16, bits)
    pr = O.prove(poly, 16, bits)
    assert O.verify(pr, 16, bits) == 1
    assert O.prove(poly, 16, bits) == pr  # deterministic
    # a different transcript seed changes the proof and is rejected under 1234
    pr2 = O.prove(poly, 16, bits, fs_seed=99)
    assert pr2 != pr and O.verify(pr2, 16, bits, fs_seed=99) == 1
    assert O.verify(pr2, 16, bits) != 1


def test_proof_independent_of_thread_count(O):
    # the OpenMP loops (column encode, per-query induce) reduce by XOR, so the
    # bytes must not depend on how many threads ran them
    poly = O.splitmix(8, 1 << 16, 32)
    default = O.lib().oracle_get_threads()
    try:
        O.lib().oracle_set_threads(1)
        one = O.prove(poly, 16, 32)
        assert O.prove_timed(poly, 16, 32, threads=1) > 0
        O.lib().oracle_set_threads(3)
        assert O.prove(poly, 16, 32) == one
    finally:
        O.lib().oracle_set_threads(default)
    assert O.prove(poly, 16, 32) == one


def test_verify_rejects_tampering(O):
    poly = O.splitmix(1, 1 << 16, 32)
    pr = bytearray(O.prove(poly, 16, 32))
    for pos in (20, 200, len(pr) - 5, len(pr) // 2):
        bad = bytearray(pr)
        bad[pos] ^= 0x40
        assert O.verify(bytes(bad), 16, 32) != 1, pos
    assert O.verify(bytes(pr[:-1]), 16, 32) == -20


def test_merkle_with_ligero_data_style(O):
    # config.rs:143-203 uses poly[i] = i; here at 2^16 with a full prove
    poly = np.arange(1 << 16, dtype=np.uint32)
    pr = O.prove(poly, 16, 32)
    assert O.verify(pr, 16, 32) == 1


def test_prove_verify_config_20(O):
    # config.rs:205-238 (test_prove_verify_config_20) on seeded input
    poly = O.splitmix(1, 1 << 20, 32)
    pr = O.prove(poly, 20, 32)
    assert O.verify(pr, 20, 32) == 1
    golden = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "oracle_proofs.json")))
    assert hashlib.sha256(pr).hexdigest() == golden["L20_seed1_gf32"]["sha256"]
    assert len(pr) == golden["L20_seed1_gf32"]["bytes"]
