"""Batched proves and the prove server, through the C ABI.

Every proof made in a batch or in the server's mix must be byte-identical to
the plain single-proof path's proof of the same input and transcript seed
(and, for a sample, to the oracle's)."""
import ctypes as C
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


def plain_bytes(ctx, cfgs, L, bits, poly, seed):
    key = (L, bits)
    if key not in cfgs:
        cfgs[key] = lh.ProverConfig(ctx, L, bits)
    return lh.prover(cfgs[key], poly, fs_seed=seed).to_bytes()


def jobs(O, sizes, bits, seed0):
    polys = [O.splitmix(seed0 + i, 1 << L, bits) for i, L in enumerate(sizes)]
    seeds = [1234 + 17 * i for i in range(len(sizes))]
    return polys, seeds


def test_batch_equals_plain_and_oracle(O, ctx):
    sizes = [16, 17, 16, 18, 16, 17, 16]
    say("batch gf32", sizes)
    polys, seeds = jobs(O, sizes, 32, 100)
    b = lh.Batch(0, 32, n_streams=3)
    proofs = b.prove(sizes, polys, seeds)
    cfgs = {}
    for L, poly, seed, p in zip(sizes, polys, seeds, proofs):
        assert p.to_bytes() == plain_bytes(ctx, cfgs, L, 32, poly, seed)
        assert lh.verifier(L, p, seed) == 1
    assert proofs[0].to_bytes() == O.prove(polys[0], sizes[0], 32, fs_seed=seeds[0])
    b.close()


def test_batch_device_inputs(O, ctx):
    # device buffers from the HIP runtime the library itself loaded
    hip = C.CDLL("libamdhip64.so.7")
    sizes = [17, 16, 16]
    say("batch gf32 device inputs", sizes)
    polys, seeds = jobs(O, sizes, 32, 300)
    dptrs = []
    try:
        for p in polys:
            d = C.c_void_p()
            assert hip.hipMalloc(C.byref(d), C.c_size_t(p.nbytes)) == 0
            dptrs.append(d)
            assert hip.hipMemcpy(d, p.ctypes.data_as(C.c_void_p), C.c_size_t(p.nbytes), 1) == 0  # host to device
        b = lh.Batch(0, 32, n_streams=2)
        proofs = b.prove(sizes, [d.value for d in dptrs], seeds, on_device=True)
        cfgs = {}
        for L, poly, seed, p in zip(sizes, polys, seeds, proofs):
            assert p.to_bytes() == plain_bytes(ctx, cfgs, L, 32, poly, seed)
        b.close()
    finally:
        for d in dptrs:
            hip.hipFree(d)


def test_batch_gf128(O, ctx):
    sizes = [16, 16, 17]
    say("batch gf128", sizes)
    polys, seeds = jobs(O, sizes, 128, 500)
    b = lh.Batch(0, 128, n_streams=2)
    proofs = b.prove(sizes, polys, seeds)
    cfgs = {}
    for L, poly, seed, p in zip(sizes, polys, seeds, proofs):
        assert p.to_bytes() == plain_bytes(ctx, cfgs, L, 128, poly, seed)
    b.close()


def test_batch_rejects_bad_size(O, ctx):
    say("batch bad size")
    b = lh.Batch(0, 32, n_streams=2)
    poly = O.splitmix(1, 1 << 15, 32)
    with pytest.raises(lh.LigeritoError) as e:
        b.prove([15], [poly], [1])
    assert e.value.code == -1
    assert b.prove([], [], []) == []
    b.close()


def test_served_results_same_alone_and_in_mix(O, ctx):
    rng = np.random.default_rng(4)
    sizes = [int(x) for x in rng.integers(16, 19, size=10)]
    say("serve mix", sizes)
    polys, seeds = jobs(O, sizes, 32, 700)
    srv = lh.Server(0, 32, n_slots=3, hbm_budget=8 << 30)
    tickets = [srv.submit(L, p, s) for L, p, s in zip(sizes, polys, seeds)]
    results = [srv.wait(t) for t in tickets]
    srv.stop()
    cfgs = {}
    for L, poly, seed, (p, lat) in zip(sizes, polys, seeds, results):
        assert lat > 0
        assert p.to_bytes() == plain_bytes(ctx, cfgs, L, 32, poly, seed)


def test_served_tiny_budget_runs_one_at_a_time(O, ctx):
    sizes = [16, 17, 16]
    say("serve tiny budget", sizes)
    polys, seeds = jobs(O, sizes, 32, 900)
    srv = lh.Server(0, 32, n_slots=2, hbm_budget=1)
    tickets = [srv.submit(L, p, s) for L, p, s in zip(sizes, polys, seeds)]
    done = {}
    while len(done) < len(tickets):
        for t in tickets:
            if t not in done:
                r = srv.poll(t)
                if r is not None:
                    done[t] = r
    srv.stop()
    cfgs = {}
    for t, L, poly, seed in zip(tickets, sizes, polys, seeds):
        assert done[t][0].to_bytes() == plain_bytes(ctx, cfgs, L, 32, poly, seed)
