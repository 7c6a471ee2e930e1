"""Regenerates tests/golden/oracle_proofs.json: SHA-256 digests of proofs the
oracle produces (self-generated regression values, not reference outputs;
the reference is unseeded and has no proof encoding)."""
import hashlib
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "..", "..", "oracle"))
import oracle as O  # noqa: E402

out = {"_source": "self-generated by the oracle; SplitMix64 inputs, transcript seed 1234"}
# (24, 1, 128): the smallest GF(2^128) size with two recursive steps and a
# glue step; the oracle needs minutes for it, so GPU tests compare digests
for L, seed, bits in [(16, 1, 32), (20, 1, 32), (16, 1, 128), (24, 1, 128)]:
    pr = O.prove(O.splitmix(seed, 1 << L, bits), L, bits)
    assert O.verify(pr, L, bits) == 1
    out[f"L{L}_seed{seed}_gf{bits}"] = {"sha256": hashlib.sha256(pr).hexdigest(), "bytes": len(pr)}
json.dump(out, open(os.path.join(HERE, "oracle_proofs.json"), "w"), indent=1)
print(out)
