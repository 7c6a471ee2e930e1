#!/usr/bin/env python3
"""Benchmark: plain Ligerito prove of 2^24 GF(2^32) coefficients on one GPU.

Workload: the reference's hardcoded_config_24 (initial matrix 2^18 x 2^6,
two recursive steps), input drawn from SplitMix64 seed 2 and resident in
HBM before the timed region, transcript seed 1234.  A step is one complete
prove (commit, fold, commit, openings, induce, sumcheck rounds, ...).

Prints one JSON line with throughput (coefficients proved per second), the
roofline of the dominant kernel group (the first commit's Reed–Solomon
encode) from HIP events on the library stream, and a CPU baseline: the
oracle (plain-C restatement of the reference algorithm) timing prover()
on the same input on all of this host's cores, and on one core over a 2^20
prefix.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "ligerito-rust_amd"))

LOG_N = 24
SEED = 2
FS_SEED = 1234
HBM_PEAK_GBS = 8000.0


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def _one_l3_cpus():
    """The CPUs of this thread's allowed set that share the last-level cache
    with the first of them (None when the topology is not readable)."""
    allowed = sorted(os.sched_getaffinity(0))
    try:
        with open(f"/sys/devices/system/cpu/cpu{allowed[0]}/cache/index3/shared_cpu_list") as fh:
            spec = fh.read().strip()
    except (OSError, IndexError):
        return None
    cpus = set()
    for part in spec.split(","):
        a, _, b = part.partition("-")
        cpus.update(range(int(a), int(b or a) + 1))
    cpus &= set(allowed)
    return sorted(cpus) or None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--log-n", type=int, default=LOG_N)
    ap.add_argument("--no-batch", action="store_true", help="skip the batched-prove measurement")
    ap.add_argument("--no-serve", action="store_true", help="skip the serving measurement")
    ap.add_argument("--no-commit", action="store_true", help="skip the isolated encode+commit measurement")
    ap.add_argument("--batch-jobs", type=int, default=256)
    ap.add_argument("--batch-streams", type=int, default=12)
    # 1000 arrivals (2.5 s at 400/s) put ten requests above the p99 mark
    ap.add_argument("--serve-requests", type=int, default=1000)
    ap.add_argument("--serve-rate", type=float, default=400.0, help="Poisson arrival rate, requests/s")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run 5 plain device copies of MIB MiB (scales hardware byte counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only: the Ligerito prove path runs on a single GPU here "
              f"(--gpus {args.gpus} is not supported)", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    if args.copy_calibration:
        src = torch.ones(args.copy_calibration << 18, dtype=torch.int32, device="cuda:0")
        dst = torch.empty_like(src)
        for _ in range(5):
            dst.copy_(src)
        torch.cuda.synchronize()
        log(f"copied {args.copy_calibration} MiB five times")
        return

    import ligerito_hip as lh

    L = args.log_n
    N = 1 << L
    ctx = lh.Context(0)
    log("plan setup (twiddle/pis tables), untimed like ReedSolomonEncoding::new in hardcoded_config_24")
    cfg = lh.ProverConfig(ctx, L, 32)

    # SplitMix64 input, generated on the host then placed in HBM
    poly = _splitmix_u32(SEED, N)
    d_poly = torch.from_numpy(poly.view(np.int32)).to("cuda:0")
    torch.cuda.synchronize()
    ptr = d_poly.data_ptr()

    # The calling thread (it drives the prove and polls for its results)
    # stays on one last-level-cache domain while the plain proves run.
    prev_cpus = os.sched_getaffinity(0)
    host_cpus = _one_l3_cpus()
    if host_cpus:
        os.sched_setaffinity(0, host_cpus)
        log(f"calling thread on CPUs {host_cpus[0]}-{host_cpus[-1]} (one L3) for the plain proves")
    for i in range(args.warmup):
        lh.prover_dev(cfg, ptr, FS_SEED)
    log(f"{args.warmup} warmup proves done")

    ctx.synchronize()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    proof = None
    for _ in range(args.steps):
        proof = lh.prover_dev(cfg, ptr, FS_SEED)
    ctx.synchronize()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    elapsed = t1 - t0
    log(f"timed {args.steps} proves in {elapsed * 1e3:.2f} ms")
    os.sched_setaffinity(0, prev_cpus)

    # Instrumented proves after the timed region: HIP events per phase and
    # around each FFT pass launch of the first commit (events add gaps between
    # launches, so they stay out of the timed region).
    n_prof = max(args.steps, 3)
    ctx.profile(True)
    for _ in range(n_prof):
        lh.prover_dev(cfg, ptr, FS_SEED)
    ctx.synchronize()
    phases = {p: ctx.profile_read(p) for p in ("encode0", "encode", "merkle", "fold", "induce", "sumcheck", "open",
                                               "pass")}
    ctx.profile(False)

    t_v = time.perf_counter()
    ok = lh.verifier(L, proof, FS_SEED)
    t_v = time.perf_counter() - t_v
    if ok != 1:
        log(f"verifier rejected the GPU proof ({ok})")
        sys.exit(1)
    lh.verifier_dev(ctx, L, proof, FS_SEED)
    t_vd = time.perf_counter()
    ok = lh.verifier_dev(ctx, L, proof, FS_SEED)
    t_vd = time.perf_counter() - t_vd
    if ok != 1:
        log(f"GPU-induce verifier rejected the GPU proof ({ok})")
        sys.exit(1)

    # host-resident input (PCIe-inclusive drop-in semantics), reported beside:
    # the median of three calls after one untimed call
    lh.prover(cfg, poly, FS_SEED)
    t_hs = []
    for _ in range(3):
        t2 = time.perf_counter()
        lh.prover(cfg, poly, FS_SEED)
        t_hs.append(time.perf_counter() - t2)
    t_host = sorted(t_hs)[1]

    # Dominant kernel group: the first commit's Reed-Solomon encode (the
    # transpose-and-scale and the three radix-2^6 FFT passes, 48% of a
    # prove's kernel time).  Its algorithmic bytes are the encode's: the
    # message read once (N*4) and the codeword written once (4N*4).
    enc_s = phases["encode0"][0] / n_prof / 1e3
    enc_alg = 5 * N * 4
    enc_gbps = enc_alg / enc_s / 1e9 if enc_s > 0 else 0.0
    # HBM bytes per encode from the committed PMC summary of the same
    # kernels on this workload (null when it was taken for another shape)
    traffic = None
    tf = os.path.join(ROOT, "profiles", "encode_traffic_prove_2e24.json")
    if os.path.exists(tf):
        with open(tf) as fh:
            tj = json.load(fh)
        if tj.get("algorithmic_bytes_per_encode") == enc_alg:
            traffic = tj.get("hbm_bytes_per_encode")
    # The radix-2^6 pass alone: one launch covers the whole 4m x n codeword
    # (four quarters of m*n = N elements), reading and writing 4N elements.
    p_ms, p_n = phases["pass"]
    p_avg_s = (p_ms / max(p_n, 1)) / 1e3
    pass_alg = 2 * (4 * N) * 4
    pass_gbps = pass_alg / p_avg_s / 1e9 if p_avg_s > 0 else 0.0
    # What binds the encode: GF(2^32) products by twiddle constants, 18
    # butterfly stages of N/2 products per quarter, four quarters (the
    # quarter FFT has log2(m) = L - 6 multiplying stages)
    ell = L - 6
    enc_products = 4 * ell * (N // 2)
    peaks = _peaks(ctx)
    # the launch shape the autotuner picks for this encode on this GPU (the
    # prove's plan made the same choice at its first commit)
    tune_rs = lh.ReedSolomonEncoding(ctx, N >> 6, N >> 4, 32)
    tuned = dict(zip(("kmax", "group"), tune_rs.autotune(64)))
    log(f"autotuned FFT passes of the first commit: {tuned} (0 = built-in default)")

    commit_c = None if args.no_commit else _commit_c(lh, torch, ctx, peaks)
    batched = None if args.no_batch else _batched(lh, torch, np, cfg, ptr, args.batch_jobs, args.batch_streams)
    served = None if args.no_serve else _served(lh, torch, np, args.serve_requests, args.serve_rate)

    cpu = None
    if not args.no_cpu_baseline:
        cpu = _cpu_baseline(poly, L)

    value = args.steps * N / elapsed
    line = {
        "metric": "Ligerito prove coeffs/s at 2^24, batched vs plain GPU; served p99 ms",
        "value": value,
        "unit": "coeffs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": elapsed / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32",
        "data": "synthetic",
        "config": {
            "workload": f"plain Ligerito prove (hardcoded_config_24 shape), 2^{L} GF(2^32) coefficients, "
                        f"device-resident input, SplitMix64 seed {SEED}, transcript seed {FS_SEED}",
            "log_n": L,
            "field": "GF(2^32)",
            "path": "plain single-proof GPU drop-in",
            "host_cpus": f"{host_cpus[0]}-{host_cpus[-1]}" if host_cpus else None,
        },
        "roofline": {
            "bound": "hbm",
            "kernel": "first-commit Reed-Solomon encode: k_transpose_scale32<6,5> and three k_fft_pass32<3,3> "
                      "launches (4 x 2^18 x 64 codeword)",
            "achieved": enc_gbps,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": enc_gbps / HBM_PEAK_GBS,
            "traffic": traffic,
            "algorithmic_bytes": enc_alg,
            "note": "5N*4 bytes per encode (message read once, codeword written once) over the encode's HIP-event "
                    "time on the library stream; traffic = FETCH_SIZE + WRITE_SIZE of the same kernels per encode",
            "avg_duration_us": enc_s * 1e6,
            "encodes_timed": n_prof,
            "binding": {
                "resource": "GF(2^32) products by twiddle constants (8 LDS nibble-table reads + 16 VALU each)",
                "products_per_encode": enc_products,
                "achieved_products_per_s": enc_products / enc_s if enc_s > 0 else 0.0,
                "peak_products_per_s": peaks["mul_tab"],
                "frac": (enc_products / enc_s) / peaks["mul_tab"] if enc_s > 0 and peaks["mul_tab"] else None,
            },
            "per_pass": {
                "kernel": "k_fft_pass32<3,3>",
                "algorithmic_bytes": pass_alg,
                "avg_duration_us": p_avg_s * 1e6,
                "launches_timed": p_n,
                "achieved": pass_gbps,
                "frac": pass_gbps / HBM_PEAK_GBS,
            },
        },
        "peaks": peaks,
        "autotune": {"first_commit_passes": tuned, "enabled": os.environ.get("LIG_AUTOTUNE", "1") != "0"},
        "cpu_baseline": cpu,
        "phases_ms_per_prove": {k: v[0] / n_prof for k, v in phases.items()},
        "host_input_coeffs_per_s": N / t_host,
        "verify_ms": {"host": t_v * 1e3, "gpu_induce": t_vd * 1e3},
        "commit_c": commit_c,
        "batched": batched,
        "served": served,
    }
    print(json.dumps(line), flush=True)


def _peaks(ctx):
    """Microbenchmarked peaks the rooflines are priced against (library
    kernels, best of five launches): GF(2^32) products/s with the FFT passes'
    LDS nibble tables, general GF(2^32) products/s, 32-bit VALU lane-ops/s."""
    out = {k: ctx.bench_rate(k) for k in ("mul_tab", "mul_clmul", "valu", "mul_tab5", "lds128", "mul_mask", "gf8_logexp")}
    log(f"peaks: GF(2^32) table products {out['mul_tab'] / 1e12:.2f} T/s (5-bit windows "
        f"{out['mul_tab5'] / 1e12:.2f}, register bit masks {out['mul_mask'] / 1e12:.2f}, GF(2^8) log/exp "
        f"{out['gf8_logexp'] / 1e12:.2f} T subfield products/s), general products {out['mul_clmul'] / 1e12:.2f} T/s, "
        f"VALU {out['valu'] / 1e12:.1f} T lane-ops/s, LDS 16-byte table reads {out['lds128'] / 1e12:.1f} TB/s")
    return out


def _gpu_poly(torch, L, seed):
    # uniform 32-bit elements from a seeded device generator
    g = torch.Generator(device="cuda:0")
    g.manual_seed(seed)
    return torch.randint(-(1 << 31), 1 << 31, (1 << L,), dtype=torch.int32, device="cuda:0", generator=g)


def _commit_c(lh, torch, ctx, peaks, reps=5):
    """Isolated encode+commit (ligero_commit): 4096 rows of 2^12 GF(2^32)
    elements (and 1024 of GF(2^128)) encoded at rate 1/4 to 2^14, one
    16 KiB leaf per encoded row, full Merkle tree.  Inputs in HBM."""
    out = {}
    for bits, m, n in ((32, 1 << 12, 4096), (128, 1 << 12, 1024)):
        words = m * n * (bits // 32)
        g = torch.Generator(device="cuda:0")
        g.manual_seed(77 + bits)
        d = torch.randint(-(1 << 31), 1 << 31, (words,), dtype=torch.int32, device="cuda:0", generator=g)
        torch.cuda.synchronize()
        rs = lh.ReedSolomonEncoding(ctx, m, 4 * m, bits)
        lh.ligero_commit_dev(d.data_ptr(), m, n, rs)
        ctx.profile(True)
        t0 = time.perf_counter()
        for _ in range(reps):
            lh.ligero_commit_dev(d.data_ptr(), m, n, rs)
        dt = (time.perf_counter() - t0) / reps
        enc_ms, _ = ctx.profile_read("encode")
        mk_ms, _ = ctx.profile_read("merkle")
        ctx.profile(False)
        enc_s = enc_ms / reps / 1e3
        mk_s = mk_ms / reps / 1e3
        fbytes = bits // 8
        alg = 5 * m * n * fbytes
        leaf_bytes = n * fbytes
        chunks = (leaf_bytes + 1023) // 1024
        comps = 4 * m * (leaf_bytes // 64 + (chunks - 1)) + (4 * m - 1)
        # products by twiddle constants: log2(m) stages of m/2 * n per quarter
        products = 4 * (m.bit_length() - 1) * (m // 2) * n
        out[f"gf{bits}"] = {
            "shape": f"{n} messages x 2^12 -> 2^14, GF(2^{bits})",
            "commit_ms": dt * 1e3,
            "encode_ms": enc_s * 1e3,
            "encode_GBps": alg / enc_s / 1e9,
            "encode_frac_of_hbm": alg / enc_s / 1e9 / HBM_PEAK_GBS,
            "encode_products": products,
            "merkle_ms": mk_s * 1e3,
            "merkle_compressions": comps,
            "merkle_Tops": comps * 792 / mk_s / 1e12,
            "merkle_frac_of_valu": comps * 792 / mk_s / peaks["valu"],
        }
        if bits == 32:
            out["gf32"]["encode_frac_of_mul_peak"] = products / enc_s / peaks["mul_tab"]
        else:
            # 32 nibble-table reads of 16 bytes per product
            out["gf128"]["encode_lds_bytes"] = products * 512
            out["gf128"]["encode_frac_of_lds_peak"] = products * 512 / enc_s / peaks["lds128"]
        log(f"commit C gf{bits}: {dt * 1e3:.3f} ms (encode {enc_s * 1e3:.3f} ms, merkle {mk_s * 1e3:.3f} ms)")
    return out


def _batched(lh, torch, np, cfg24, ptr24, n_jobs, streams=12):
    """Batched proves beside the plain path: (a) n_jobs polynomials with log2
    sizes uniform in 16..22 (seed 3), each with its own transcript seed;
    (b) eight 2^24 proves.  Inputs are in HBM; prover configurations are
    built before timing, as for the plain path."""
    rng = np.random.default_rng(3)
    sizes = [int(x) for x in rng.integers(16, 23, size=n_jobs)]
    seeds = [1234 + i for i in range(n_jobs)]
    polys = [_gpu_poly(torch, L, 1000 + i) for i, L in enumerate(sizes)]
    torch.cuda.synchronize()
    ptrs = [t.data_ptr() for t in polys]
    total = sum(1 << L for L in sizes)
    b = lh.Batch(0, 32, n_streams=streams)
    b.prove(sizes, ptrs, seeds, on_device=True)  # builds each worker's configurations
    # each figure is the median of `reps` timed runs (one run of the batch
    # is ~20 ms, so single runs vary with the host's scheduling)
    reps = 5

    def med(fn):
        ts = []
        for _ in range(reps):
            t0 = time.perf_counter()
            fn()
            ts.append(time.perf_counter() - t0)
        return sorted(ts)[reps // 2]

    t_batch = med(lambda: b.prove(sizes, ptrs, seeds, on_device=True))
    proofs = b.prove(sizes, ptrs, seeds, on_device=True)
    for i in (0, n_jobs // 2, n_jobs - 1):
        if lh.verifier(sizes[i], proofs[i], seeds[i]) != 1:
            raise SystemExit(f"verifier rejected batched proof {i}")
    # the same jobs one after another on the plain path
    ctx = lh.Context(0)
    cfgs = {L: lh.ProverConfig(ctx, L, 32) for L in set(sizes)}
    lh.prover_dev(cfgs[sizes[0]], ptrs[0], seeds[0])
    ctx.synchronize()

    def plain_all():
        for L, p, s in zip(sizes, ptrs, seeds):
            lh.prover_dev(cfgs[L], p, s)
        ctx.synchronize()

    t_plain = med(plain_all)
    log(f"batched {n_jobs} jobs: {t_batch * 1e3:.1f} ms batched, {t_plain * 1e3:.1f} ms plain (medians of {reps})")
    # eight 2^24 proves
    n24 = 8
    ptr_list = [ptr24] * n24
    seeds24 = [1234 + i for i in range(n24)]
    b.prove([24] * n24, ptr_list, seeds24, on_device=True)
    t_b24 = med(lambda: b.prove([24] * n24, ptr_list, seeds24, on_device=True))

    def plain24():
        for s in seeds24:
            lh.prover_dev(cfg24, ptr24, s)

    t_p24 = med(plain24)
    log(f"eight 2^24 proves: {t_b24 * 1e3:.1f} ms batched, {t_p24 * 1e3:.1f} ms plain (medians of {reps})")
    b.close()
    return {
        "workload": f"{n_jobs} GF(2^32) polynomials, log2 sizes uniform in 16..22 (seed 3), own transcript seeds, "
                    f"inputs in HBM, {streams} streams, same-size jobs proved in lockstep gangs of up to 4",
        "coeffs": total,
        "batched_coeffs_per_s": total / t_batch,
        "plain_coeffs_per_s": total / t_plain,
        "batched_over_plain": t_plain / t_batch,
        "batch_2pow24": {"jobs": n24, "batched_coeffs_per_s": n24 * (1 << 24) / t_b24,
                         "plain_coeffs_per_s": n24 * (1 << 24) / t_p24, "batched_over_plain": t_p24 / t_b24},
    }


def _served(lh, torch, np, n_req, rate, slots=8, budget=64 << 30):
    """Open-loop Poisson arrivals of prove requests (log2 sizes uniform in
    16..24, seed 4) against the prove server; inputs in pinned host memory
    (the upload is part of each request's latency)."""
    rng = np.random.default_rng(4)
    sizes = [int(x) for x in rng.integers(16, 25, size=n_req)]
    gaps = rng.exponential(1.0 / rate, size=n_req)
    host = {}
    for L in sorted(set(sizes)):
        g = torch.Generator()
        g.manual_seed(4000 + L)
        host[L] = torch.randint(-(1 << 31), 1 << 31, (1 << L,), dtype=torch.int32, generator=g).pin_memory()
    srv = lh.Server(0, 32, n_slots=slots, hbm_budget=budget)
    # untimed: every slot's configuration and device blocks for every size,
    # then one request from each input buffer (the runtime maps a pinned
    # buffer for the device on its first copy, a few ms at 2^24)
    srv.warm(sorted(host))
    for L in sorted(host):
        srv.wait(srv.submit(L, host[L].data_ptr(), 1))
    t_first = time.perf_counter()
    t_next = t_first
    tickets = []
    for i, L in enumerate(sizes):
        t_next += gaps[i]
        while time.perf_counter() < t_next:
            pass
        tickets.append(srv.submit(L, host[L].data_ptr(), 5000 + i))
    lats = []
    for t in tickets:
        _, lat = srv.wait(t)
        lats.append(lat)
    t_end = time.perf_counter()
    srv.stop()
    total = sum(1 << L for L in sizes)
    lats = np.array(lats)
    log(f"served {n_req} requests at {rate}/s: p50 {np.percentile(lats, 50):.2f} ms, p99 {np.percentile(lats, 99):.2f} ms")
    sz = np.array(sizes)
    by_size = {str(L): {"n": int((sz == L).sum()), "p50_ms": float(np.percentile(lats[sz == L], 50)),
                        "p90_ms": float(np.percentile(lats[sz == L], 90))} for L in sorted(set(sizes))}
    return {
        "workload": f"{n_req} requests, Poisson arrivals at {rate:g}/s (seed 4), log2 sizes uniform in 16..24, "
                    f"GF(2^32), inputs in pinned host memory, {slots} slots, HBM budget {budget >> 30} GiB",
        "throughput_coeffs_per_s": total / (t_end - t_first),
        "throughput_requests_per_s": n_req / (t_end - t_first),
        "p50_ms": float(np.percentile(lats, 50)),
        "p99_ms": float(np.percentile(lats, 99)),
        "mean_ms": float(lats.mean()),
        "by_log_n": by_size,
    }


def _splitmix_u32(seed, n):
    import numpy as np

    # SplitMix64, low 32 bits of each draw (the oracle's generator)
    idx = np.arange(1, n + 1, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = np.uint64(seed) + idx * np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return (z & np.uint64(0xFFFFFFFF)).astype(np.uint32)


def _cpu_baseline(poly, L):
    """The oracle (plain-C restatement of the reference prover) on this
    host: prover() alone, config built before the clock as src/main.rs:29-49
    times it.  All cores on the full 2^L input, then one core on the 2^20
    prefix of the same input (a whole 2^L prove on one core runs about a
    minute)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O

    O.build()
    cores = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or len(os.sched_getaffinity(0))
    log(f"cpu baseline: prover() on 2^{L} with the oracle on {cores} threads")
    dt = O.prove_timed(poly, L, 32, threads=cores)
    L1 = min(L, 20)
    log(f"cpu baseline: prover() on 2^{L1} with the oracle on 1 thread")
    dt1 = O.prove_timed(poly[: 1 << L1], L1, 32, threads=1)
    return {
        "value": (1 << L) / dt,
        "unit": "coeffs/s",
        "cores": cores,
        "kind": "port",
        "sample": f"prover() on the same 2^{L} GF(2^32) input, config prebuilt (oracle: plain-C restatement of "
                  f"the reference algorithm, bit-serial field multiply, OpenMP over columns/queries), {dt:.2f} s",
        "single_core": {"value": (1 << L1) / dt1, "unit": "coeffs/s", "cores": 1,
                        "sample": f"prover() on the first 2^{L1} coefficients of the same input, {dt1:.2f} s"},
    }


if __name__ == "__main__":
    main()
