"""The library's transcript gives the oracle's squeezes and distinct queries
on every SHA-256 code path it can take on this host (SHA extensions,
eight-way AVX2 finalizations, portable), including rejection-heavy cases
(n close to the query count)."""
import os
import random
import subprocess
import sys

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)

SCRIPT = r'''
import random, sys
sys.path.insert(0, sys.argv[1]); sys.path.insert(0, sys.argv[2])
import ligerito_hip as lh, oracle as O
random.seed(11)
bad = 0
for trial in range(150):
    seed = random.randrange(1 << 32)
    a, b = lh.FS(seed), O.FS(seed)
    for step in range(random.randrange(1, 5)):
        data = bytes(random.randrange(256) for _ in range(4 * random.randrange(0, 20)))
        a.absorb(data); b.absorb(data)
        if random.random() < 0.5 and a.squeeze() != b.squeeze():
            bad += 1
        n = random.choice([1 << 20, 1 << 16, 4096, 300, 149, 148])
        if list(a.get_distinct_queries(n, 148)) != list(b.get_distinct_queries(n, 148)):
            bad += 1
        if a.squeeze() != b.squeeze():
            bad += 1
print(bad)
'''


@pytest.mark.parametrize("env", [{}, {"LIG_NO_SHA_NI": "1"}, {"LIG_NO_SHA_NI": "1", "LIG_NO_AVX2": "1"}])
def test_transcript_code_paths_match_oracle(O, env):
    e = dict(os.environ)
    e.update(env)
    r = subprocess.run([sys.executable, "-c", SCRIPT, os.path.join(ROOT, "ligerito-rust_amd"), os.path.join(ROOT, "oracle")],
                       env=e, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    assert r.stdout.strip() == "0"
