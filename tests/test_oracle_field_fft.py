"""Pins the oracle's field arithmetic, FFT, twiddles and Reed–Solomon code.

Mirrors the reference tests in src/binary_field.rs:191-312,
src/binary_fft.rs:324-399 and src/reed_solomon.rs:116-183, and adds checks
that come from the mathematics (schoolbook GF products, subspace
polynomials, polynomial evaluation) so that a plausible mistake in the
restatement fails one of them.
"""
import json
import os
import random

import numpy as np
import pytest

import gfref

FIELDS = [16, 32, 128]


def rnd(bits, rng):
    return rng.getrandbits(bits)


def E(vals, bits, O):
    if bits == 128:
        return np.array([[v & (2**64 - 1), v >> 64] for v in vals], dtype=np.uint64)
    return np.array(vals, dtype={16: np.uint16, 32: np.uint32}[bits])


def ints(arr, bits):
    a = np.asarray(arr)
    if bits == 128:
        return [int(x[0]) | (int(x[1]) << 64) for x in a.reshape(-1, 2)]
    return [int(x) for x in a.reshape(-1)]


@pytest.mark.parametrize("bits", FIELDS)
def test_mul_matches_schoolbook(O, bits):
    rng = random.Random(bits)
    edge = [0, 1, 2, (1 << bits) - 1, 1 << (bits - 1), 3]
    a = edge + [rnd(bits, rng) for _ in range(300)]
    b = [rnd(bits, rng) for _ in range(len(edge))] + edge[::-1][:0] + [rnd(bits, rng) for _ in range(300)]
    b = b[: len(a)]
    got = ints(O.mul(E(a, bits, O), E(b, bits, O), bits), bits)
    want = [gfref.mul(x, y, bits) for x, y in zip(a, b)]
    assert got == want


@pytest.mark.parametrize("bits", FIELDS)
def test_field_axioms(O, bits):
    # binary_field.rs:195-305 with a fixed seed
    rng = random.Random(7 + bits)
    for _ in range(20):
        x, y, z = (rnd(bits, rng) for _ in range(3))
        m = lambda u, v: ints(O.mul(E([u], bits, O), E([v], bits, O), bits), bits)[0]
        assert m(x, 0) == 0 and m(x, 1) == x and m(1, 1) == 1
        assert m(x, y) == m(y, x)
        assert m(x, m(y, z)) == m(m(x, y), z)
        assert m(x, y ^ z) == m(x, y) ^ m(x, z)
        if x:
            assert m(x, O.inverse(x, bits)) == 1
    assert O.inverse(1, bits) == 1
    with pytest.raises(ZeroDivisionError):
        O.inverse(0, bits)
    # x + max = max - 1 style overflow check of binary_field.rs:213-219
    assert ((1 << bits) - 1) ^ 1 == (1 << bits) - 2


def test_twiddles_match_julia_kat(O, golden_dir):
    kat = json.load(open(os.path.join(golden_dir, "reference_kats.json")))["twiddles_gf16_log10"]
    tw = O.compute_twiddles(10, 16)
    assert len(tw) == kat["length"]
    assert [int(x) for x in tw[:5]] == kat["first5"]


@pytest.mark.parametrize("bits", [16, 32])
def test_sks_are_subspace_polynomials(O, bits):
    # s_i(v_i) = W_i(2^i): eval_sk_at_vks (binary_fft.rs:217-246) against the
    # product over the subspace span{1..2^(i-1)}.
    n = 1 << 8
    sks = ints(O.eval_sk_at_vks(n, bits), bits)
    assert len(sks) == 9
    for i in range(9):
        assert sks[i] == gfref.subspace_poly(i, 1 << i, bits)


@pytest.mark.parametrize("bits", [16, 32])
def test_evaluate_basis_is_subspace_product(O, bits):
    rng = random.Random(3)
    n = 6
    sks = O.eval_sk_at_vks(1 << n, bits)
    for _ in range(4):
        x = rnd(bits, rng)
        got = ints(O.evaluate_basis(1 << n, sks, x, bits), bits)
        W = [gfref.subspace_poly(t, x, bits) for t in range(n)]
        for j in range(1 << n):
            want = 1
            for t in range(n):
                if (j >> t) & 1:
                    want = gfref.mul(want, W[t], bits)
            assert got[j] == want


@pytest.mark.parametrize("bits", FIELDS)
def test_fft_ifft_roundtrip(O, bits):
    # binary_fft.rs:351-399 (values 1..16) plus a random vector
    tw = O.compute_twiddles(4, bits)
    v = E(list(range(1, 17)), bits, O)
    back = O.fft(O.fft(v, tw, bits), tw, bits, inverse=True)
    assert ints(back, bits) == list(range(1, 17))
    rng = random.Random(11)
    tw = O.compute_twiddles(9, bits)
    r = [rnd(bits, rng) for _ in range(512)]
    assert ints(O.fft(O.fft(E(r, bits, O), tw, bits), tw, bits, inverse=True), bits) == r


@pytest.mark.parametrize("bits", [16, 32])
def test_twiddles_linear_in_block_index(O, bits):
    # With beta = 0 every FFT stage's twiddles are a GF(2)-linear function of
    # the block index (layer_0 is i<<1 and next_s is linearised); the GPU
    # kernels rely on this to derive per-block tables.
    L = 10
    tw = ints(O.compute_twiddles(L, bits), bits)
    for s in range(L):
        base = (1 << s) - 1
        for b in range(1 << s):
            acc = 0
            for i in range(s):
                if (b >> i) & 1:
                    acc ^= tw[base + (1 << i)]
            assert tw[base + b] == acc


def test_reed_solomon_systematic(O):
    # reed_solomon.rs:121-131
    msg = E(list(range(1, 1025)), 16, O)
    enc = O.rs_encode(msg, 4096, 16)
    assert ints(enc[:1024], 16) == list(range(1, 1025))


def test_reed_solomon_non_systematic_reference_case(O):
    # reed_solomon.rs:133-182: linearity and the basis identity at x = 100
    a = list(range(1, 1025))
    b = list(range(1025, 2049))
    l = 42
    c = [x ^ gfref.mul(l, y, 16) for x, y in zip(a, b)]
    ap = E(a + [0] * 3072, 16, O)
    bp = E(b + [0] * 3072, 16, O)
    ea = ints(O.rs_encode_non_systematic(ap, 1024, 16), 16)
    eb = ints(O.rs_encode_non_systematic(bp, 1024, 16), 16)
    sks = O.eval_sk_at_vks(1024, 16)
    basis = ints(O.evaluate_basis(1024, sks, 100, 16), 16)
    want = 0
    for ci, bi in zip(c, basis):
        want ^= gfref.mul(ci, bi, 16)
    assert ea[100] ^ gfref.mul(l, eb[100], 16) == want


@pytest.mark.parametrize("bits", [16, 32, 128])
def test_non_systematic_codeword_is_polynomial_evaluation(O, bits):
    # Every codeword entry x equals sum_j m_j * prod_{t in bits(j)} W_t(x):
    # the FFT output is in natural order, evaluated at the field element x.
    m, M = 16, 64
    rng = random.Random(5)
    msg = [rnd(bits, rng) for _ in range(m)]
    enc = ints(O.rs_encode_non_systematic(E(msg + [0] * (M - m), bits, O), m, bits), bits)
    if bits == 128:
        xs = [0, 1, 17, 63]
    else:
        xs = range(M)
    for x in xs:
        W = [gfref.subspace_poly(t, x, bits) for t in range(4)]
        want = 0
        for j, mj in enumerate(msg):
            bj = 1
            for t in range(4):
                if (j >> t) & 1:
                    bj = gfref.mul(bj, W[t], bits)
            want ^= gfref.mul(mj, bj, bits)
        assert enc[x] == want, x


def test_pis_are_products_of_sks(O):
    sks = ints(O.eval_sk_at_vks(64, 32), 32)
    pis = ints(O.compute_pis(64, O.eval_sk_at_vks(64, 32), 32), 32)
    for j in range(64):
        want = 1
        for t in range(6):
            if (j >> t) & 1:
                want = gfref.mul(want, sks[t], 32)
        assert pis[j] == want
