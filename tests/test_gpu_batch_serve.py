"""Batched proves and the prove server, through the C ABI.

Every proof made in a batch or in the server's mix must be byte-identical to
the plain single-proof path's proof of the same input and transcript seed
(and, for a sample, to the oracle's)."""
import ctypes as C
import sys
import threading
import time

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


def test_gang_proofs_equal_plain_and_oracle(O, ctx):
    # same-size jobs are proved in lockstep (one launch per kernel for the
    # gang, at most four jobs): 9 x 2^16 (gangs of 3), 5 x 2^19, 3 x 2^21,
    # 6 x 2^22 and 2 x 2^24 (two recursive steps: a second commit, and the
    # glue in the second section), every proof equal to the plain path's
    sizes = [16] * 9 + [19] * 5 + [21] * 3 + [22] * 6 + [24] * 2
    say("gang gf32", sizes)
    polys, seeds = jobs(O, sizes, 32, 700)
    b = lh.Batch(0, 32, n_streams=2)
    proofs = b.prove(sizes, polys, seeds)
    cfgs = {}
    for i, (L, poly, seed, p) in enumerate(zip(sizes, polys, seeds, proofs)):
        say("gang job", i, "log", L)
        assert p.to_bytes() == plain_bytes(ctx, cfgs, L, 32, poly, seed)
    for i in (0, 8, 9):
        assert proofs[i].to_bytes() == O.prove(polys[i], sizes[i], 32, fs_seed=seeds[i])
    assert lh.verifier(24, proofs[-1], seeds[-1]) == 1
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


def test_served_warm_keeps_results(O, ctx):
    # warm-up called while requests run and wait: it lets the running ones
    # finish, holds the queued ones back, prepares every slot for every size,
    # and what the slots prove afterwards is the plain prover's proof
    sizes = [18, 17, 18, 16, 17, 18]
    say("serve warm", sizes)
    polys, seeds = jobs(O, sizes, 32, 1100)
    srv = lh.Server(0, 32, n_slots=3, hbm_budget=8 << 30)
    early = [srv.submit(L, p, s) for L, p, s in zip(sizes, polys, seeds)]
    assert srv.warm([18, 16, 17, 16]) == 3 * 3
    late = [srv.submit(L, p, s) for L, p, s in zip(sizes, polys, seeds)]
    results = [srv.wait(t) for t in early + late]
    srv.stop()
    cfgs = {}
    for (L, poly, seed), (p, _) in zip(list(zip(sizes, polys, seeds)) * 2, results):
        assert p.to_bytes() == plain_bytes(ctx, cfgs, L, 32, poly, seed)
    # a budget that holds no prove warms nothing and still serves
    srv = lh.Server(0, 32, n_slots=2, hbm_budget=1)
    assert srv.warm([16]) == 0
    p, _ = srv.wait(srv.submit(16, polys[3], seeds[3]))
    srv.stop()
    assert p.to_bytes() == plain_bytes(ctx, cfgs, 16, 32, polys[3], seeds[3])
    with pytest.raises(lh.LigeritoError):
        lh.Server(0, 32, n_slots=1).warm([15])
    # GF(2^128) slots
    p128, s128 = jobs(O, [16], 128, 1200)
    srv = lh.Server(0, 128, n_slots=2, hbm_budget=8 << 30)
    assert srv.warm([16]) == 2
    p, _ = srv.wait(srv.submit(16, p128[0], s128[0]))
    srv.stop()
    assert p.to_bytes() == plain_bytes(ctx, cfgs, 16, 128, p128[0], s128[0])


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


def test_config_d_256_jobs_12_streams(O, ctx):
    # the batched workload at its stated size: 256 polynomials, log2 sizes
    # uniform in 16..22 (seed 3), each with its own transcript seed, on 12
    # worker streams; every proof equals the plain path's, one per size the
    # oracle's
    rng = np.random.default_rng(3)
    sizes = [int(x) for x in rng.integers(16, 23, size=256)]
    say("batch config D: 256 jobs, sizes", sorted(set(sizes)))
    seeds = [1234 + i for i in range(256)]
    polys = [O.splitmix(3000 + i, 1 << L, 32) for i, L in enumerate(sizes)]
    b = lh.Batch(0, 32, n_streams=12)
    proofs = b.prove(sizes, polys, seeds)
    b.close()
    say("batch config D: comparing with the plain path")
    cfgs = {}
    first = {}
    for i, (L, poly, seed, p) in enumerate(zip(sizes, polys, seeds, proofs)):
        got = p.to_bytes()
        assert got == plain_bytes(ctx, cfgs, L, 32, poly, seed), f"job {i} (2^{L})"
        first.setdefault(L, (i, got))
    for L, (i, got) in sorted(first.items()):
        say("batch config D: oracle check of job", i, "size", L)
        assert got == O.prove(polys[i], L, 32, fs_seed=seeds[i]), f"job {i} (2^{L})"


def test_config_e_poisson_8_slots(O, ctx):
    # the served workload: open-loop Poisson arrivals (seed 4), log2 sizes
    # uniform in 16..24, 8 slots; every result equals the plain path's
    rng = np.random.default_rng(4)
    n_req, rate = 48, 400.0
    sizes = [int(x) for x in rng.integers(16, 25, size=n_req)]
    gaps = rng.exponential(1.0 / rate, size=n_req)
    say("serve config E:", n_req, "requests, sizes", sorted(set(sizes)))
    polys = [O.splitmix(4000 + i, 1 << L, 32) for i, L in enumerate(sizes)]
    seeds = [4321 + i for i in range(n_req)]
    srv = lh.Server(0, 32, n_slots=8, hbm_budget=64 << 30)
    tickets = []
    t0 = time.perf_counter()
    due = 0.0
    for L, p, s, g in zip(sizes, polys, seeds, gaps):
        due += g
        while time.perf_counter() - t0 < due:
            pass
        tickets.append(srv.submit(L, p, s))
    results = [srv.wait(t) for t in tickets]
    srv.stop()
    say("serve config E: comparing with the plain path")
    cfgs = {}
    for i, (L, poly, seed, (p, lat)) in enumerate(zip(sizes, polys, seeds, results)):
        assert lat > 0
        assert p.to_bytes() == plain_bytes(ctx, cfgs, L, 32, poly, seed), f"request {i} (2^{L})"


def _hip_used_bytes(hip):
    free, total = C.c_size_t(0), C.c_size_t(0)
    assert hip.hipMemGetInfo(C.byref(free), C.byref(total)) == 0
    return total.value - free.value


def test_served_memory_stays_within_budget(O, ctx):
    # HBM-budget admission bounds the device memory the server really holds:
    # what each worker keeps cached between requests is charged too and
    # given back when a request would not fit otherwise
    hip = C.CDLL("libamdhip64.so.7")
    budget = 3 << 30
    overhead = 3 << 29  # contexts, streams and the runtime's per-queue scratch
    rng = np.random.default_rng(5)
    sizes = [int(x) for x in rng.integers(20, 25, size=40)]
    say("serve memory bound: budget", budget, "sizes", sizes)
    polys = {L: O.splitmix(5000 + L, 1 << L, 32) for L in set(sizes)}
    base = _hip_used_bytes(hip)
    peak = [0]
    stop = threading.Event()

    def sample():
        while not stop.is_set():
            peak[0] = max(peak[0], _hip_used_bytes(hip) - base)
            time.sleep(0.0005)

    th = threading.Thread(target=sample)
    th.start()
    try:
        srv = lh.Server(0, 32, n_slots=8, hbm_budget=budget)
        tickets = [srv.submit(L, polys[L], 1000 + i) for i, L in enumerate(sizes)]
        results = [srv.wait(t) for t in tickets]
        srv.stop()
    finally:
        stop.set()
        th.join()
    say("serve memory bound: peak above baseline", peak[0])
    assert peak[0] <= budget + overhead, f"device memory {peak[0]} above budget {budget} + {overhead}"
    cfgs = {}
    for i in (0, len(sizes) - 1):
        assert results[i][0].to_bytes() == plain_bytes(ctx, cfgs, sizes[i], 32, polys[sizes[i]], 1000 + i)
