"""The GF(2^32) carry-less product the kernels use (clmul32_holes in
ligerito-rust_amd/csrc/gf.hpp: 16 integer products over bit classes mod 4),
compiled for the host and checked against the textbook product of
tests/gfref.py (and, on random pairs, against the bit-serial product)."""
import os
import random
import shutil
import subprocess

import pytest

import gfref

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope="module")
def checker(tmp_path_factory):
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("no host C++ compiler")
    exe = str(tmp_path_factory.mktemp("gf") / "gf_host_check")
    r = subprocess.run([gxx, "-O2", "-std=c++17", os.path.join(HERE, "native", "gf_host_check.cpp"), "-o", exe],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    return exe


def test_clmul_holes_matches_textbook(checker):
    rng = random.Random(3)
    edge = [0, 1, 2, 3, 0xFFFFFFFF, 0x80000000, 0x11111111, 0x22222222, 0x44444444, 0x88888888, 0xF0F0F0F0,
            0x0F0F0F0F, 0xAAAAAAAA, 0x55555555, 0x8299]
    pairs = [(x, y) for x in edge for y in edge]
    pairs += [(rng.getrandbits(32), rng.getrandbits(32)) for _ in range(3000)]
    # dense residue classes: the most terms that can meet at one position
    pairs += [(rng.getrandbits(32) | 0x88888888, 0xFFFFFFFF) for _ in range(200)]
    inp = "".join(f"{x} {y}\n" for x, y in pairs)
    r = subprocess.run([checker], input=inp, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    lines = r.stdout.strip().split("\n")
    assert lines[-1] == "mismatches 0"
    for (x, y), line in zip(pairs, lines[:-1]):
        prod, red = (int(t) for t in line.split())
        assert prod == gfref.clmul(x, y), (hex(x), hex(y))
        assert red == gfref.mul(x, y, 32), (hex(x), hex(y))
