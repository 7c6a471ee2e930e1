"""C-ABI library checks that need no GPU: the library loads, exports every
symbol include/ligerito.h declares, and its host-side pieces (transcript,
Merkle proof logic, proof encoding, verifier) agree with the oracle."""
import hashlib
import random

import numpy as np
import pytest

import ligerito_hip as lh


@pytest.fixture(scope="module")
def L():
    return lh.lib()


def test_exports_every_header_symbol(L):
    syms = lh.header_symbols()
    assert len(syms) >= 40
    missing = [s for s in syms if not hasattr(L, s)]
    assert not missing, missing


def test_no_device_is_an_error_not_a_fallback(L):
    if L.lig_device_count() > 0:
        pytest.skip("a GPU is visible")
    with pytest.raises(lh.LigeritoError) as e:
        lh.Context(0)
    assert e.value.code == -5  # LIG_E_NO_DEVICE


def test_batch_and_server_need_a_device(L):
    if L.lig_device_count() > 0:
        pytest.skip("a GPU is visible")
    with pytest.raises(lh.LigeritoError) as e:
        lh.Batch(0, 32, 2)
    assert e.value.code == -5
    with pytest.raises(lh.LigeritoError) as e:
        lh.Server(0, 32, 2)
    assert e.value.code == -5


def test_transcript_matches_oracle(O):
    for seed in (1234, 0, 77):
        a, b = lh.FS(seed), O.FS(seed)
        for step in range(6):
            data = bytes([step]) * (step * 7)
            a.absorb(data)
            b.absorb(data)
            assert a.squeeze() == b.squeeze()
            assert a.get_field(32) == b.get_field(32)
            assert a.get_field(128) == b.get_field(128)
            assert a.get_query(1000 + step) == b.get_query(1000 + step)
        assert a.get_distinct_queries(4096, 148) == b.get_distinct_queries(4096, 148)
    with pytest.raises(lh.LigeritoError):
        lh.FS(1).get_distinct_queries(4, 5)


@pytest.mark.parametrize("n,queries", [(8, [0, 2]), (8, [0, 1, 2]), (16, [3, 4, 9, 15]), (1024, list(range(0, 1024, 7)))])
def test_merkle_prove_verify_host(O, n, queries):
    rng = random.Random(n)
    leaves = [bytes(rng.getrandbits(8) for _ in range(64)) for _ in range(n)]
    layers = O.build_merkle_tree(leaves)
    got = lh.prove(layers, queries)
    assert got == O.merkle_prove(leaves, queries)
    depth = len(layers) - 1
    sel = np.frombuffer(b"".join(leaves[q] for q in queries), dtype=np.uint8).reshape(len(queries), 64)
    assert lh.verify(layers[-1], got, depth, sel, queries)
    bad = sel.copy()
    bad[0, 0] ^= 1
    assert not lh.verify(layers[-1], got, depth, bad, queries)


@pytest.mark.parametrize("L_,bits", [(16, 32), (16, 128), (20, 32)])
def test_product_verifier_accepts_oracle_proofs(O, L_, bits):
    poly = O.splitmix(1, 1 << L_, bits)
    pr = O.prove(poly, L_, bits)
    p = lh.Proof.from_bytes(pr, bits)
    assert p.to_bytes() == pr  # canonical encoding round-trips
    assert lh.verifier(L_, p) == 1
    assert lh.verifier(L_, p, fs_seed=5) != 1
    bad = bytearray(pr)
    bad[len(bad) // 3] ^= 0x10
    assert lh.verifier(L_, lh.Proof.from_bytes(bytes(bad), bits)) != 1


def test_proof_size_matches_reference_rule(O):
    # ProofSize (data_structures.rs:105-194): elements x size_of::<F>() + 32 per node/root
    poly = O.splitmix(3, 1 << 16, 32)
    pr = O.prove(poly, 16, 32)
    p = lh.Proof.from_bytes(pr, 32)
    # encoding overhead = magic/version/field (12) + counts: RLP 3 x u64 each,
    # 4 list counts; subtract them from the byte length
    n_rlp = 2  # initial + final (L=16 has no recursive proofs)
    overhead = 12 + 8 * 3 * n_rlp + 8 * 4
    assert p.proof_size() == len(pr) - overhead


def test_malformed_proof_rejected():
    with pytest.raises(lh.LigeritoError) as e:
        lh.Proof.from_bytes(b"LGRT" + b"\0" * 10, 32)
    assert e.value.code == -6


def _bare(cls, **attrs):
    """A wrapper object with no library handle: calls that reach the C ABI
    would fail, so a test passing through one shows the check ran first."""
    o = cls.__new__(cls)
    for k, v in attrs.items():
        setattr(o, k, v)
    return o


def test_wrappers_reject_wrong_length_inputs():
    # the C entry points copy 2^log_n or m*n elements from the pointer; the
    # wrappers must refuse a short array before the call
    cfg = _bare(lh.ProverConfig, _h=None, bits=32, log_n=16, num_coeffs=1 << 16)
    with pytest.raises(lh.LigeritoError) as e:
        lh.prover(cfg, np.zeros((1 << 16) - 1, dtype=np.uint32))
    assert e.value.code == -1
    rs = _bare(lh.ReedSolomonEncoding, _h=None, bits=128)
    with pytest.raises(lh.LigeritoError) as e:
        lh.ligero_commit(np.zeros((64 * 4 - 1, 2), dtype=np.uint64), 64, 4, rs)
    assert e.value.code == -1
    b = _bare(lh.Batch, _h=None, bits=32)
    with pytest.raises(lh.LigeritoError) as e:
        b.prove([16, 17], [np.zeros(1 << 16, np.uint32), np.zeros(1 << 16, np.uint32)], [1, 2])
    assert e.value.code == -1
    with pytest.raises(lh.LigeritoError):
        b.prove([16], [np.zeros(1 << 16, np.uint32)], [1, 2])
    s = _bare(lh.Server, _h=None, bits=32, _keep={})
    with pytest.raises(lh.LigeritoError) as e:
        s.submit(17, np.zeros(1 << 16, np.uint32))
    assert e.value.code == -1


def test_merkle_verify_variable_length_leaves(O):
    # the crate's own leaves (merkle_tree.rs:195-255 use "hello", "world",
    # "rust", "test") and leaves of mixed lengths, as AsRef<[u8]> allows
    for leaves in ([b"hello", b"world", b"rust", b"test"],
                   [bytes(range(k % 251)) * (1 + k // 251) for k in (0, 1, 5, 64, 65, 1023, 1024, 1500)]):
        layers = O.build_merkle_tree(leaves)
        depth = len(layers) - 1
        for q in ([0, 2], [0, 1, 2], list(range(len(leaves)))[1::2]):
            proof = O.merkle_prove(leaves, q)
            assert lh.verify(layers[-1], proof, depth, [leaves[i] for i in q], q)
            bad = [bytearray(leaves[i]) + b"x" for i in q]
            assert not lh.verify(layers[-1], proof, depth, [bytes(b) for b in bad], q)
