"""Parity of the HIP path with the oracle (bit-exact), through the C ABI.

Each case prints what it runs before it runs it, so a failing GPU step is
named in the log."""
import hashlib
import json
import os
import sys

import numpy as np
import pytest

import ligerito_hip as lh

pytestmark = pytest.mark.gpu


def say(*a):
    print("[gpu-case]", *a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def ctx():
    if lh.lib().lig_device_count() == 0:
        pytest.fail("no GPU visible: the HIP library cannot run (no CPU fallback exists)")
    c = lh.Context(0)
    yield c
    c.close()


def col(O, seed, n, bits):
    return O.splitmix(seed, n, bits)


@pytest.mark.parametrize("bits", [32, 128])
@pytest.mark.parametrize("logm", [0, 1, 3, 6, 9])
def test_rs_non_systematic_matches_oracle(O, ctx, bits, logm):
    m = 1 << logm
    M = 4 * m
    say("rs_non_systematic", bits, m)
    rs = lh.ReedSolomonEncoding(ctx, m, M, bits)
    assert np.array_equal(rs.twiddles(), O.compute_twiddles(logm + 2, bits))
    buf = O.empty(M, bits)
    buf[:m] = col(O, 11 + logm, m, bits)
    got = rs.encode_non_systematic(buf)
    want = O.rs_encode_non_systematic(buf, m, bits)
    assert np.array_equal(got, want)


@pytest.mark.parametrize("bits", [32, 128])
def test_rs_non_systematic_nonzero_tail(O, ctx, bits):
    # encode_non_systematic FFTs the whole block (reed_solomon.rs:104-113)
    m, M = 64, 256
    say("rs_non_systematic tail", bits)
    rs = lh.ReedSolomonEncoding(ctx, m, M, bits)
    buf = col(O, 5, M, bits)
    assert np.array_equal(rs.encode_non_systematic(buf), O.rs_encode_non_systematic(buf, m, bits))


@pytest.mark.parametrize("bits", [32, 128])
def test_rs_systematic_matches_oracle(O, ctx, bits):
    m, M = 256, 1024
    say("rs_systematic", bits)
    rs = lh.ReedSolomonEncoding(ctx, m, M, bits)
    msg = col(O, 3, m, bits)
    got = rs.encode(msg)
    assert np.array_equal(got, O.rs_encode(msg, M, bits))
    assert np.array_equal(got[:m], msg)  # reed_solomon.rs:121-131


@pytest.mark.parametrize("leaf_len", [4, 32, 64, 256, 1024, 1028, 16384])
def test_build_merkle_tree_matches_oracle(O, ctx, leaf_len):
    n = 64 if leaf_len < 16384 else 16
    say("merkle", leaf_len, n)
    rng = np.random.default_rng(leaf_len)
    leaves = rng.integers(0, 256, size=(n, leaf_len), dtype=np.uint8)
    got = lh.build_merkle_tree(ctx, leaves)
    want = O.build_merkle_tree([bytes(r) for r in leaves])
    assert got == want


@pytest.mark.parametrize("log_leaves", [17, 20])
def test_build_merkle_tree_wide_levels(O, ctx, log_leaves):
    # 2^20 leaves take the three-levels-per-thread kernel, then the
    # four-lane subtree kernels; 2^17 leaves only the latter
    n = 1 << log_leaves
    say("merkle wide", n)
    rng = np.random.default_rng(log_leaves)
    leaves = rng.integers(0, 256, size=(n, 16), dtype=np.uint8)
    got = lh.build_merkle_tree(ctx, leaves)
    want = O.build_merkle_tree([bytes(r) for r in leaves])
    assert got == want


# GF(2^32) shapes cover the transpose-and-scale tiles of 4, 8, 16, 32 and 64
# columns, two column tiles (128), the generic path (m < 64 or n < 4), and
# narrow matrices whose leading FFT passes run on the 64-column view (one
# such pass at 2^8 x 16 and 2^7 x 8, two at 2^14 x 16 and 2^13 x 4); rows
# of 64 bytes (16 GF(2^32) or 4 GF(2^128) columns) hash their leaves inside
# the tree's first top launch, from 64 leaves to 2^16
@pytest.mark.parametrize("bits,m,n", [(32, 1 << 10, 64), (32, 1 << 8, 16), (32, 64, 4), (32, 16, 1),
                                      (32, 1 << 7, 8), (32, 1 << 6, 32), (32, 32, 8), (32, 1 << 8, 2),
                                      (128, 1 << 8, 16), (128, 64, 64), (32, 1 << 9, 128),
                                      (32, 1 << 14, 16), (32, 1 << 13, 4), (32, 16, 16), (128, 64, 4)])
def test_ligero_commit_matches_oracle(O, ctx, bits, m, n):
    say("commit", bits, m, n)
    rs = lh.ReedSolomonEncoding(ctx, m, 4 * m, bits)
    poly = col(O, m + n, m * n, bits)
    w = lh.ligero_commit(poly, m, n, rs)
    root, mat, layers = O.ligero_commit(poly, m, n, bits, want_matrix=True, want_layers=True)
    assert w.root() == root
    assert np.array_equal(w.flat_mat(), mat)
    assert w.tree() == layers
    for r in (0, 1, 4 * m - 1):
        assert np.array_equal(w.extract_row(r), mat.reshape(n, 4 * m, -1)[:, r].reshape(w.extract_row(r).shape))


def test_merkle_with_ligero_data(O, ctx):
    # the reference's test_merkle_with_ligero_data (config.rs:143-203): the
    # input poly[i] = i, hardcoded_config_20's first commit (2^14 x 64 ->
    # 2^16 rows), the rows at queries [0, 1, 100, 1959, 50294, 50459] opened
    # and their batched Merkle proof verified against the root; here the root,
    # the tree, the rows and the proof also equal the oracle's
    m, n = 1 << 14, 64
    say("merkle with ligero data")
    poly = np.arange(m * n, dtype=np.uint32)
    rs = lh.ReedSolomonEncoding(ctx, m, 4 * m, 32)
    w = lh.ligero_commit(poly, m, n, rs)
    root, mat, layers = O.ligero_commit(poly, m, n, 32, want_matrix=True, want_layers=True)
    assert w.root() == root
    tree = w.tree()
    assert tree == layers
    queries = sorted([0, 1, 100, 50459, 50294, 1959])
    rows = [w.extract_row(q) for q in queries]
    for q, r in zip(queries, rows):
        assert np.array_equal(r, mat.reshape(n, 4 * m)[:, q])
    leaves = np.stack([np.ascontiguousarray(r).view(np.uint8) for r in rows])
    proof = lh.prove(tree, queries)
    depth = len(tree) - 1
    assert lh.verify(tree[-1], proof, depth, leaves, queries)
    rb = np.ascontiguousarray(mat.reshape(n, 4 * m).T).view(np.uint8)
    row_bytes = [rb[r].tobytes() for r in range(4 * m)]
    assert proof == O.merkle_prove(row_bytes, queries)
    assert O.merkle_verify(root, proof, depth, [bytes(x) for x in leaves], queries)
    # a changed row no longer verifies
    bad = leaves.copy()
    bad[2, 0] ^= 1
    assert not lh.verify(tree[-1], proof, depth, bad, queries)


def test_commit_tree_2e19_leaves(O, ctx):
    # 2^19 leaves: two single-level launches to 2^17 nodes, then subtree
    # launches of depth 7 and the one-workgroup top
    m, n = 1 << 17, 4
    say("commit tree top", m, n)
    rs = lh.ReedSolomonEncoding(ctx, m, 4 * m, 32)
    poly = col(O, 21, m * n, 32)
    w = lh.ligero_commit(poly, m, n, rs)
    root, _, layers = O.ligero_commit(poly, m, n, 32, want_layers=True)
    assert w.root() == root
    assert w.tree() == layers
    # a second commit on the same context
    w2 = lh.ligero_commit(poly[::-1].copy(), m, n, rs)
    root2, _, _ = O.ligero_commit(poly[::-1].copy(), m, n, 32)
    assert w2.root() == root2


def test_commit_encode_kernel_shape_gf32(O, ctx):
    # 4096 independent length-2^12 rows encoded to 2^14, 16 KiB leaves
    m, n = 1 << 12, 4096
    say("commit gf32 4096x2^12")
    rs = lh.ReedSolomonEncoding(ctx, m, 4 * m, 32)
    poly = col(O, 7, m * n, 32)
    w = lh.ligero_commit(poly, m, n, rs)
    root, _, _ = O.ligero_commit(poly, m, n, 32)
    assert w.root() == root


def test_commit_encode_kernel_shape_gf128(O, ctx):
    # 1024 independent length-2^12 GF(2^128) rows encoded to 2^14: column
    # chunks of the GF(2^128) passes beyond 64 columns, 16 KiB leaves
    m, n = 1 << 12, 1024
    say("commit gf128 1024x2^12")
    rs = lh.ReedSolomonEncoding(ctx, m, 4 * m, 128)
    poly = col(O, 7, m * n, 128)
    w = lh.ligero_commit(poly, m, n, rs)
    say("commit gf128 1024x2^12: oracle")
    root, mat, layers = O.ligero_commit(poly, m, n, 128, want_matrix=True, want_layers=True)
    assert w.root() == root
    assert w.tree() == layers
    for r in (0, 5, 4 * m - 1):
        assert np.array_equal(w.extract_row(r), mat.reshape(n, 4 * m, 2)[:, r])


# inductions run over 2^(L-6) and 2^(L-10) entries: L = 16..23 and 25 here
# and 24 below give every size from 2^6 to 2^19, hence every accumulate row
# length (2^6 to 2^10 entries)
@pytest.mark.parametrize("bits,L", [(32, 16), (128, 16), (32, 17), (32, 18), (32, 19), (32, 20), (32, 21), (32, 22),
                                    (32, 23), (32, 25)])
def test_prover_bytes_match_oracle(O, ctx, bits, L):
    say("prove", bits, L)
    cfg = lh.ProverConfig(ctx, L, bits)
    poly = col(O, 1, 1 << L, bits)
    p = lh.prover(cfg, poly)
    got = p.to_bytes()
    want = O.prove(poly, L, bits)
    assert got == want
    assert O.verify(got, L, bits) == 1
    assert lh.verifier(L, p) == 1
    if (bits, L) == (32, 20):
        g = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "oracle_proofs.json")))
        assert hashlib.sha256(got).hexdigest() == g["L20_seed1_gf32"]["sha256"]


@pytest.mark.parametrize("L", [26, 27])
def test_prover_largest_sizes(O, ctx, L):
    # the largest generated configurations (2^27 coefficients is the
    # library's limit): proof bytes equal the oracle's, both verifiers accept
    say("prove largest", L)
    cfg = lh.ProverConfig(ctx, L, 32)
    poly = col(O, 3, 1 << L, 32)
    p = lh.prover(cfg, poly)
    got = p.to_bytes()
    assert lh.verifier(L, p) == 1
    assert lh.verifier_dev(ctx, L, p) == 1
    say("prove largest", L, ": oracle")
    assert got == O.prove(poly, L, 32)


def test_prover_seeded_transcripts(O, ctx):
    cfg = lh.ProverConfig(ctx, 16, 32)
    poly = col(O, 9, 1 << 16, 32)
    for seed in (1, 99):
        say("prove seed", seed)
        got = lh.prover(cfg, poly, fs_seed=seed).to_bytes()
        assert got == O.prove(poly, 16, 32, fs_seed=seed)


def test_prover_2pow24_bit_exact(O, ctx):
    # the full-size workload: GPU proof bytes == oracle proof bytes, and both
    # verifiers accept
    say("prove 2^24")
    cfg = lh.ProverConfig(ctx, 24, 32)
    poly = col(O, 2, 1 << 24, 32)
    p = lh.prover(cfg, poly)
    got = p.to_bytes()
    assert lh.verifier(24, p) == 1
    want = O.prove(poly, 24, 32)
    assert got == want


def test_prover_gf128_2pow24_two_steps(O, ctx):
    # GF(2^128) at 2^24: two recursive commits, the glue step and the
    # generic fold path; the oracle takes minutes here, so its proof is
    # pinned by the digest committed under tests/golden
    say("prove gf128 2^24")
    g = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "oracle_proofs.json")))["L24_seed1_gf128"]
    cfg = lh.ProverConfig(ctx, 24, 128)
    poly = col(O, 1, 1 << 24, 128)
    p = lh.prover(cfg, poly)
    got = p.to_bytes()
    assert len(got) == g["bytes"]
    assert hashlib.sha256(got).hexdigest() == g["sha256"]
    assert lh.verifier(24, p) == 1
    assert lh.verifier_dev(ctx, 24, p) == 1


@pytest.mark.parametrize("bits,L", [(32, 16), (128, 16), (32, 20)])
def test_gpu_verifier_matches_host_verifier(O, ctx, bits, L):
    # verifier-side induce on the GPU: same verdicts as the host verifier on
    # valid proofs and on tampered ones
    say("verifier_dev", bits, L)
    poly = col(O, 3, 1 << L, bits)
    good = O.prove(poly, L, bits)
    p = lh.Proof.from_bytes(good, bits)
    assert lh.verifier_dev(ctx, L, p) == 1 == lh.verifier(L, p)
    assert lh.verifier_dev(ctx, L, p, fs_seed=7) == lh.verifier(L, p, fs_seed=7)
    rng = np.random.default_rng(L + bits)
    blob = bytearray(good)
    for _ in range(4):
        pos = int(rng.integers(44, len(blob)))
        t = bytearray(blob)
        t[pos] ^= 1 << int(rng.integers(0, 8))
        try:
            pt = lh.Proof.from_bytes(bytes(t), bits)
        except lh.LigeritoError:
            continue
        assert lh.verifier_dev(ctx, L, pt) == lh.verifier(L, pt)


def test_gpu_verifier_2pow24(O, ctx):
    say("verifier_dev 2^24")
    cfg = lh.ProverConfig(ctx, 24, 32)
    poly = col(O, 2, 1 << 24, 32)
    p = lh.prover(cfg, poly)
    assert lh.verifier_dev(ctx, 24, p) == 1


# ------------------------------------------------ replay of GPU proofs --
@pytest.mark.parametrize("bits,L", [(32, 16), (128, 16), (32, 20), (32, 24)])
def test_replay_accepts_gpu_proofs(O, ctx, bits, L):
    # the independent hashlib replay re-derives every challenge and query set
    # from the GPU proof's bytes and checks it against the input
    import replay as RP
    from test_replay import oracle_commit

    say("replay gpu proof", bits, L)
    cfg = lh.ProverConfig(ctx, L, bits)
    poly = col(O, 4, 1 << L, bits)
    got = lh.prover(cfg, poly, fs_seed=99).to_bytes()
    RP.replay(got, poly, lh.config_shape(L), bits, oracle_commit(O, bits), fs_seed=99)


@pytest.mark.parametrize("which,bits", [("two_step_16", 32), ("two_step_16", 128), ("three_step_17", 32),
                                        ("cfg24", 32)])
def test_prover_from_dims_matches_oracle(O, ctx, which, bits):
    # ProverConfig from explicit dimensions (data_structures.rs:9-21): the
    # proof equals the oracle's for the same shape, both verifiers accept it
    # with the matching VerifierConfig, and cfg24's literals give the same
    # bytes as the generated 2^24 configuration
    import replay as RP
    from test_replay import CFG_LITERALS, THREE_STEP_17, TWO_STEP_16, oracle_commit

    shape = {"two_step_16": TWO_STEP_16, "three_step_17": THREE_STEP_17, "cfg24": CFG_LITERALS[24]}[which]
    say("prove from dims", which, bits)
    cfg = lh.ProverConfigFromDims(ctx, shape, bits)
    assert cfg.shape() == shape
    poly = col(O, 8, cfg.num_coeffs, bits)
    p = lh.prover(cfg, poly, fs_seed=77)
    got = p.to_bytes()
    assert lh.verifier_with_config(shape, p, fs_seed=77) == 1
    assert lh.verifier_with_config(shape, p, fs_seed=77, ctx=ctx) == 1
    if which == "cfg24":
        assert got == lh.prover(lh.ProverConfig(ctx, 24, bits), poly, fs_seed=77).to_bytes()
        return
    assert got == O.prove_dims(poly, shape, bits, fs_seed=77)
    RP.replay(got, poly, shape, bits, oracle_commit(O, bits), fs_seed=77)


def test_prover_from_dims_rejects_broken_shapes(ctx):
    from test_replay import TWO_STEP_16

    say("prove from dims: invalid shapes")
    bad = []
    s = dict(TWO_STEP_16, initial_k=5)  # initial_k != log2 initial_dims.1
    bad.append(s)
    s = dict(TWO_STEP_16, dims=[(1 << 7, 1 << 2), (1 << 6, 1 << 2)])  # rows chain broken
    bad.append(s)
    s = dict(TWO_STEP_16, dims=[(1 << 8, 1 << 2), (1 << 5, 1 << 3)], ks=[2, 3])  # 4 * 32 < 148 rows
    bad.append(s)
    s = dict(TWO_STEP_16, recursive_steps=0, dims=[], ks=[])  # prover.rs:198
    bad.append(s)
    for sh in bad:
        with pytest.raises(lh.LigeritoError) as e:
            lh.ProverConfigFromDims(ctx, sh, 32)
        assert e.value.code == -1


@pytest.mark.parametrize("bits,logm,n", [(32, 0, 4), (32, 6, 2), (32, 8, 16), (32, 10, 64), (32, 7, 128),
                                         (128, 6, 4), (128, 8, 64), (128, 5, 128)])
def test_rs_encode_cols_dev_matches_oracle(O, ctx, bits, logm, n):
    # encode_cols (ligero.rs:29-38) on device-resident columns: column-major
    # full blocks (nonzero tails too), encoded in place on the device
    from hipmem import DeviceBuffer

    m = 1 << logm
    M = 4 * m
    say("rs encode cols dev", bits, m, n)
    rs = lh.ReedSolomonEncoding(ctx, m, M, bits)
    cols = col(O, 40 + logm + n, M * n, bits)
    d = DeviceBuffer.of(cols)
    rs.encode_non_systematic_cols_dev(d.value, n)
    got = d.download(cols)
    d.free()
    shape = (n, M) if bits == 32 else (n, M, 2)
    got, cols = got.reshape(shape), cols.reshape(shape)
    for j in range(n):
        assert np.array_equal(got[j], O.rs_encode_non_systematic(cols[j].copy(), m, bits)), f"column {j}"


@pytest.mark.parametrize("bits,logm,n", [(32, 12, 64), (32, 10, 16), (128, 9, 128), (128, 8, 64)])
def test_every_tuned_pass_shape_gives_the_same_commit(O, ctx, bits, logm, n):
    # the autotuner picks among launch shapes of the FFT passes; every
    # candidate must give the oracle's codeword and root
    m = 1 << logm
    say("tuning shapes", bits, m, n)
    rs = lh.ReedSolomonEncoding(ctx, m, 4 * m, bits)
    poly = col(O, 60 + logm, m * n, bits)
    root, mat, _ = O.ligero_commit(poly, m, n, bits, want_matrix=True)
    shapes = [(k, g) for k in (6, 5, 4, 3, 2) for g in (1, 2, 4, 8)] if bits == 32 else \
        [(k, g) for k in (3, 2, 1) for g in (1, 2, 4, 8, 16)]
    for k, g in shapes:
        rs.set_tuning(n, k, g)
        w = lh.ligero_commit(poly, m, n, rs)
        assert w.root() == root, (k, g)
        assert np.array_equal(w.flat_mat(), mat), (k, g)
    rs.set_tuning(n, 0, 0)
    assert lh.ligero_commit(poly, m, n, rs).root() == root
    # the autotuner's own choice is one of the candidates
    rs2 = lh.ReedSolomonEncoding(ctx, m, 4 * m, bits)
    k, g = rs2.autotune(n)
    assert (k, g) == (0, 0) or (k, g) in shapes
    assert lh.ligero_commit(poly, m, n, rs2).root() == root


def test_build_merkle_tree_variable_length_leaves(O, ctx):
    # build_merkle_tree<T: AsRef<[u8]>> (merkle_tree.rs:22-60): the crate's
    # own KAT leaves, then leaves of every length class (empty, partial and
    # whole blocks, one and several 1 KiB chunks, unaligned starts)
    say("merkle variable-length leaves")
    kat = [b"hello", b"world", b"rust", b"test"]
    assert lh.build_merkle_tree(ctx, kat) == O.build_merkle_tree(kat)
    rng = np.random.default_rng(11)
    lens = [0, 1, 3, 63, 64, 65, 127, 1023, 1024, 1025, 2048, 3001, 5, 70000, 16384, 9]
    leaves = [bytes(rng.integers(0, 256, size=L, dtype=np.uint8)) for L in lens]
    assert lh.build_merkle_tree(ctx, leaves) == O.build_merkle_tree(leaves)
    assert lh.build_merkle_tree(ctx, [b"only leaf"]) == O.build_merkle_tree([b"only leaf"])
    with pytest.raises(lh.LigeritoError) as e:
        lh.build_merkle_tree(ctx, [b"a", b"b", b"c"])
    assert e.value.code == -2  # merkle_tree.rs:26 asserts a power of two
