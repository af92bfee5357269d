"""Benchmark: batched GCDs/s on 100k-bit pairs vs one pair per launch.

Workload (synthetic, seeded): NPAIRS pairs u = g a, v = g b with a 10k-bit
shared factor g and 90k-bit cofactors (~100k-bit operands), 32-bit moduli.
A step is one pass of the whole batch through the persistent kernel with
the inputs already resident in HBM (mgcd_batch_run: launch, digits back to
the host, conversion to positional limbs).  The one-pair-per-launch path
(mgcd_gcd_limbs: staging, launch, read-back per pair) is timed on a sample
for p50/p99 latency and its GCDs/s.  Results are checked bit-exact against
GMP mpz_gcd on the CPU-baseline sample and by divisibility on every pair.

Prints one JSON line (rank 0, N = 1 GPU only).
"""
import argparse
import json
import math
import os
import statistics
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import __graft_entry__ as ge  # noqa: E402

# MI355X peak VALU issue rate: 256 CUs x 4 SIMD x 16 lanes x 2.4 GHz lane-ops/s
# (= the FP64 vector peak 78.6 TFLOP/s counting an FMA as two flops).
VALU_PEAK_LANE_OPS = 256 * 4 * 16 * 2.4e9

# Algorithmic work per unit, in VALU issue slots of gfx950 (a quarter-rate
# v_rcp_f32 counts 4).  Units are the algorithm's own: active
# modulus-iterations, not the lanes the kernel keeps busy after their modulus
# retired.  FROZEN: the per-unit costs below were counted once from the
# compiled loops (6 / 40 + 12.73 x 11 / B / 19 / 15) and are not recounted
# when the kernel's instruction stream changes, so a leaner loop raises the
# roofline fraction instead of shrinking its numerator.
OPS_PER_RESIDUE_STEP = 6        # one Horner step per limb per modulus
OPS_PER_VALUE = 40              # per active modulus-iteration and pair: t' = p'/(t - b) with the
                                # batch-inversion share (4 mulmod), toSigned, key, select, reduce-scatter:
                                # the (512, 2, 16) loop's common path, 1,187 VALU per iteration for its
                                # 32 values (two slots of 16), plus ~100 per iteration, over 32 values
OPS_PER_EEA_STEP = 11           # one nearest-quotient extended-Euclid step (7 VALU + rcp)
EEA_STEPS = {32: 12.73, 27: 10.71, 22: 8.68}   # mean least-remainder Euclid steps, random y, m near 2^L
                                               # (CPU emulation, 2e6 each)
OPS_PER_REPLAY = 19             # per surviving modulus and recorded pivot: 3 mulmod + subtract
OPS_PER_RECOVERY = 15           # 2 mulmod + reduce per modulus-digit


# HBM bytes per 100k-bit pair (10k-bit shared factor, L = 32) of the batched kernel, measured with
# rocprofv3 FETCH_SIZE / WRITE_SIZE on the 4,096-pair launch and scaled by a copy kernel of known
# size (FETCH_SIZE counts half the bytes, WRITE_SIZE all): 2 x 3,385,532 + 3,332,201 KiB per launch.
HBM_BYTES_PER_HEADLINE_PAIR = (2 * 2047288 + 3332088) * 1024 / 4096


def elim_ops_per_unit(L, B):
    # one extended Euclid per modulus-iteration is shared by the B pairs of a group
    return OPS_PER_VALUE + EEA_STEPS[L] * OPS_PER_EEA_STEP / B


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--pairs", type=int, default=4096)
    ap.add_argument("--bits", type=int, default=100000)
    ap.add_argument("--gbits", type=int, default=10000)
    ap.add_argument("--L", type=int, default=32)
    ap.add_argument("--per-launch-sample", type=int, default=128)
    ap.add_argument("--workgroup", type=int, default=0)
    ap.add_argument("--slots", type=int, default=0)
    ap.add_argument("--group", type=int, default=0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--workload", choices=["headline", "mixed", "skewed", "serve", "single"], default="headline",
                    help="headline: the metric's 100k-bit batch (default); mixed: 4096 pairs of 1k-200k bits, "
                         "autotuned; skewed: 16384 pairs, 95%% 4k-bit and 5%% 500k-bit, 27-bit moduli, replayed "
                         "hipGraph; serve: Poisson arrivals into the serving loop at five offered rates; single: "
                         "one pair per call beside one-thread mpz_gcd (8k-bit pairs and edge pairs, 1M-bit pairs "
                         "at L = 32/27/22, 16-352 Kibit sweep)")
    ap.add_argument("--parts", choices=["all", "small", "large", "sweep"], default="all",
                    help="single: which part to run")
    ap.add_argument("--seeds", type=int, default=10, help="single: seeds per configuration")
    ap.add_argument("--Ls", default="32,27,22", help="single: moduli widths of the 1M-bit part")
    ap.add_argument("--relay", type=int, default=0, help="single: parts from which the exchange relays (0: default)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py runs on one GPU only: the modular GCD path is not sharded across GPUs; "
              "a batch of pairs would be split across GPUs pair by pair, which needs no exchange.",
              file=sys.stderr)
        sys.exit(2)

    m = ge.package()
    w = ge.workload()
    if args.workload != "headline":
        {"mixed": run_mixed, "skewed": run_skewed, "serve": run_serve,
         "single": run_single}[args.workload](args, m, w)
        return
    t0 = time.time()
    pairs = w.headline(args.pairs, seed=7, nbits=args.bits, gbits=args.gbits)
    log(f"generated {len(pairs)} pairs in {time.time() - t0:.1f}s")
    ctx = m.Context(args.L)
    batch = m.Batch.from_pairs(ctx, pairs, workgroup=args.workgroup, slots=args.slots, group=args.group)
    T, K, B = batch.tiling()
    log(f"batch resident; tiling T={T} K={K} B={B}")

    for i in range(args.warmup):
        batch.run()
        log(f"warmup {i} done")
    ctx.synchronize()
    dev_times = []
    t_start = time.perf_counter()
    for i in range(args.steps):
        batch.run()
        dev_times.append(batch.timing()[0])
    ctx.synchronize()
    elapsed = time.perf_counter() - t_start
    ms_per_step = elapsed / args.steps * 1e3
    value = args.pairs * args.steps / elapsed
    log(f"batched: {ms_per_step:.1f} ms/step, {value:.1f} GCD/s, kernel {statistics.mean(dev_times)*1e3:.1f} ms")

    # Correctness of every pair: g | u, g | v, g0 | g.
    results = batch.results()
    for i, (u, v) in enumerate(pairs):
        g = results[i]
        assert g > 0 and u % g == 0 and v % g == 0, f"pair {i}: result does not divide the operands"

    # Work counts for the roofline.
    res_steps = elim_units = rep_units = rec_units = 0
    for i in range(args.pairs):
        nq, R, k = batch.pair_stats(i)
        nu = (pairs[i][0].bit_length() + 31) // 32
        nv = (pairs[i][1].bit_length() + 31) // 32
        res_steps += nq * (nu + nv)
        elim_units += R * nq - R * (R - 1) // 2 + nq  # initial keys + iterations
        rep_units += (nq - R) * R                    # survivors x pivots
        rec_units += sum(nq - R - j for j in range(k))
    alg_ops = (res_steps * OPS_PER_RESIDUE_STEP + elim_units * elim_ops_per_unit(args.L, B)
               + rep_units * OPS_PER_REPLAY + rec_units * OPS_PER_RECOVERY)
    kern_s = statistics.mean(dev_times)
    achieved = alg_ops / kern_s / 1e12

    # One pair per launch.
    lat = []
    for i in range(args.per_launch_sample):
        u, v = pairs[i]
        t = time.perf_counter()
        g = ctx.gcd(u, v)
        lat.append(time.perf_counter() - t)
        assert g == results[i], f"pair {i}: one-per-launch result differs from batched"
    lat.sort()
    if lat:
        p50 = statistics.median(lat)
        p99 = lat[min(len(lat) - 1, int(math.ceil(0.99 * len(lat))) - 1)]
        single_rate = 1.0 / statistics.mean(lat)
        log(f"one-per-launch: p50 {p50*1e3:.2f} ms p99 {p99*1e3:.2f} ms -> {single_rate:.1f} GCD/s")
    else:
        p50 = p99 = float("nan")
        single_rate = float("nan")

    cpu = None
    exact_checked = 0
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        import oracle_lib as O
        # Every core this process may run on, up to the 16-core host share a
        # single-GPU box gives one GPU's jobs (its process guard counts threads).
        affinity = len(os.sched_getaffinity(0))
        threads = max(1, min(16, affinity))
        sample = min(args.pairs, 4096)
        limbs, u_off, nu, v_off, nv = m.pack_pairs(pairs[:sample])
        sec, gbits = O.time_mpz_gcd(limbs, u_off, nu, v_off, nv, threads)
        for i in range(sample):
            assert int(gbits[i]) == results[i].bit_length(), f"pair {i}: gcd size differs from mpz_gcd"
        for i in range(min(sample, 64)):
            assert results[i] == O.mpz_gcd(*pairs[i]), f"pair {i}: gcd differs from mpz_gcd"
            exact_checked += 1
        cpu = {"value": sample / sec, "unit": "GCD/s", "cores": threads, "kind": "reference",
               "affinity_cores": affinity, "nproc": os.cpu_count(), "cpu_model": _cpu_model(),
               "sample": f"GMP {O.gmp_version()} mpz_gcd (the reference's CPU path, testmodgcd.cpp:261-263) on the "
                         f"first {sample} pairs of the same workload, {threads} threads pulling pairs from a queue "
                         f"({affinity} cores in the affinity mask, {_cpu_model()}); {sec:.2f} s wall"}
        log(f"cpu baseline: {cpu['value']:.1f} GCD/s on {threads} threads")

    line = {
        "metric": "GCDs/sec batched vs one-per-launch, 100k-bit pairs; p50/p99 latency; bit-exact",
        "value": round(value, 3),
        "unit": "GCD/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "exact integers held in f64 (FP32 reciprocal for quotient estimates)",
        "data": "synthetic",
        "config": {"workload": f"batched {args.pairs} x {args.bits}-bit pairs, {args.gbits}-bit shared factor, "
                               f"{args.L}-bit moduli", "pairs": args.pairs, "bits": args.bits,
                   "gcd_bits": args.gbits, "L": args.L, "tiling": {"workgroup": T, "slots": K, "group": B}},
        "input_bits_per_s": round(value * 2 * args.bits, 1),
        "one_per_launch": {"gcds_per_s": round(single_rate, 3), "p50_ms": round(p50 * 1e3, 3),
                           "p99_ms": round(p99 * 1e3, 3), "sample": args.per_launch_sample},
        "batched_vs_one_per_launch": round(value / single_rate, 2),
        "bit_exact": {"mpz_gcd_equal": exact_checked, "divisibility_checked": args.pairs},
        "roofline": {"bound": "valu", "achieved": round(achieved, 3), "peak": round(VALU_PEAK_LANE_OPS / 1e12, 2),
                     "unit": "Tlane-op/s", "frac": round(achieved / (VALU_PEAK_LANE_OPS / 1e12), 4),
                     "traffic": (round(HBM_BYTES_PER_HEADLINE_PAIR * args.pairs)
                                 if (args.bits, args.gbits, args.L) == (100000, 10000, 32) else None),
                     "traffic_unit": "bytes per launch (rocprofv3 FETCH_SIZE + WRITE_SIZE, copy-calibrated)",
                     "kernel_ms": round(kern_s * 1e3, 3)},
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)
    batch.close()
    ctx.close()


def _one_per_launch(ctx, pairs, n, seed):
    """Mean seconds per pair of the one-pair path on a seeded sample."""
    import random
    rnd = random.Random(seed)
    sample = rnd.sample(range(len(pairs)), min(n, len(pairs)))
    lat = []
    for i in sample:
        u, v = pairs[i]
        t = time.perf_counter()
        g = ctx.gcd(u, v)
        lat.append(time.perf_counter() - t)
        assert g == math.gcd(u, v), f"pair {i}: one-per-launch result differs from math.gcd"
    lat.sort()
    return statistics.mean(lat), statistics.median(lat), lat[min(len(lat) - 1, int(math.ceil(0.99 * len(lat))) - 1)]


def _timed_runs(batch, ctx, steps, warmup):
    for _ in range(warmup):
        batch.run()
    ctx.synchronize()
    dev = []
    t0 = time.perf_counter()
    for _ in range(steps):
        batch.run()
        dev.append(batch.timing()[0])
    ctx.synchronize()
    return (time.perf_counter() - t0) / steps, statistics.mean(dev)


def run_mixed(args, m, w):
    """4096 pairs, log-uniform 1k..200k bits, g uniform in [0, half], 10% coprime, 1% v = 0, L = 32;
    size classes autotuned on first use; against one pair per launch."""
    pairs = w.mixed(4096, seed=3)
    ctx = m.Context(32)
    ctx.set_option(m.OPT_AUTOTUNE, 1)
    t0 = time.perf_counter()
    batch = m.Batch.from_pairs(ctx, pairs)
    tune_s = time.perf_counter() - t0
    step_s, kern_s = _timed_runs(batch, ctx, args.steps, args.warmup)
    res = batch.results()
    for i, (u, v) in enumerate(pairs):
        assert res[i] == math.gcd(u, v), f"pair {i}: result differs from math.gcd"
    launches = batch.launches()
    one = _one_per_launch(ctx, pairs, args.per_launch_sample * 4, 11) if args.per_launch_sample > 0 else None
    value = len(pairs) / step_s
    print(json.dumps({
        "metric": "GCDs/sec batched vs one-per-launch, 4096 mixed-size pairs; bit-exact",
        "value": round(value, 3), "unit": "GCD/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(step_s * 1e3, 3), "kernel_ms": round(kern_s * 1e3, 3),
        "config": {"workload": "4096 pairs, bit lengths log-uniform in [1k, 200k], g bits uniform in [0, len/2], "
                               "10% coprime, 1% v = 0, 32-bit moduli", "seed": 3},
        "launches": [{"tiling": x["tiling"], "pairs": x["pairs"], "nq": x["nq"],
                      "ms": round(x["device_s"] * 1e3, 3)} for x in launches],
        "autotune_s": round(tune_s, 2),
        "one_per_launch": ({"gcds_per_s": round(1 / one[0], 3), "p50_ms": round(one[1] * 1e3, 3),
                            "p99_ms": round(one[2] * 1e3, 3), "sample": args.per_launch_sample * 4}
                           if one else None),
        "batched_vs_one_per_launch": round(value * one[0], 2) if one else None,
        "bit_exact": {"math_gcd_equal": len(pairs)},
    }), flush=True)
    batch.close()
    ctx.close()


def run_skewed(args, m, w):
    """16384 pairs: 95% 4k-bit, 5% 500k-bit, g = 20% of the length, L = 27; the batch's launch
    sequence is captured once into a hipGraph and replayed."""
    pairs = w.skewed(16384, seed=4)
    ctx = m.Context(27)
    eager = m.Batch.from_pairs(ctx, pairs)
    eager.run()
    eager.run()
    launches = eager.launches()
    tot = sum(x["device_s"] for x in launches)
    res = eager.results()
    eager.close()
    ctx.set_option(m.OPT_BATCH_GRAPH, 1)
    batch = m.Batch.from_pairs(ctx, pairs)
    ctx.set_option(m.OPT_BATCH_GRAPH, 0)
    step_s, kern_s = _timed_runs(batch, ctx, args.steps, args.warmup)
    assert batch.results() == res, "graph replay differs from the eager run"
    for i, (u, v) in enumerate(pairs):
        assert res[i] == math.gcd(u, v), f"pair {i}: result differs from math.gcd"
    value = len(pairs) / step_s
    print(json.dumps({
        "metric": "GCDs/sec, skewed batch (95% 4k-bit, 5% 500k-bit) replayed as a hipGraph; bit-exact",
        "value": round(value, 3), "unit": "GCD/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(step_s * 1e3, 3), "kernel_ms": round(kern_s * 1e3, 3),
        "config": {"workload": "16384 pairs, 95% 4096-bit and 5% 500000-bit, g = 20% of the length, "
                               "27-bit moduli", "seed": 4},
        "launches_eager": [{"tiling": x["tiling"], "pairs": x["pairs"], "nq": x["nq"],
                            "ms": round(x["device_s"] * 1e3, 3),
                            "share": round(x["device_s"] / tot, 4) if tot else None} for x in launches],
        "bit_exact": {"math_gcd_equal": len(pairs)},
    }), flush=True)
    batch.close()
    ctx.close()


def run_serve(args, m, w):
    """Poisson arrivals at five offered rates; sizes 70% 2k, 25% 50k, 5% 1M bits, g = 10% of the
    length; 64 GB residue budget; three launch slots, one of them for the large requests.  Then the
    200/s arrivals again under a budget that binds (8 MiB: about three 1M-bit requests resident at
    once), so that admission holds requests back."""
    import random
    import threading
    rates = [25, 50, 100, 200, 400]
    n = 300
    ctx = m.Context(32)

    def serve(srv, rate, ri):
        sizes = w.serve_sizes(n, seed=5 + ri)
        reqs = [w.serve_pair(100 + ri, i, nb) for i, nb in enumerate(sizes)]
        rnd = random.Random(ri)
        gaps = [rnd.expovariate(rate) for _ in range(n)]
        tickets = [None] * n
        t_sub = [0.0] * n

        def feed():
            t = time.perf_counter()
            for i in range(n):
                t += gaps[i]
                d = t - time.perf_counter()
                if d > 0:
                    time.sleep(d)
                t_sub[i] = time.perf_counter()
                tickets[i] = srv.submit(*reqs[i])

        th = threading.Thread(target=feed)
        t0 = time.perf_counter()
        th.start()
        th.join()
        got, lat = [], []
        for i in range(n):
            g, l = srv.wait(tickets[i])
            got.append(g)
            lat.append(l)
        # Throughput over the serving span: first submit to last completion (each completion is
        # its submit time plus the latency the server measured), not including the checks below.
        span = max(ts + l for ts, l in zip(t_sub, lat)) - t0
        st = srv.stats()
        for i in range(n):
            assert got[i] == math.gcd(*reqs[i]), f"rate {rate} request {i}: result differs from math.gcd"

        def pct(xs, f):
            xs = sorted(xs)
            return round(xs[min(len(xs) - 1, int(math.ceil(f * len(xs))) - 1)] * 1e3, 3)

        per_size = {str(nb): {"n": sizes.count(nb), "p50_ms": pct([l for l, s2 in zip(lat, sizes) if s2 == nb], 0.5),
                              "p99_ms": pct([l for l, s2 in zip(lat, sizes) if s2 == nb], 0.99)}
                    for nb in sorted(set(sizes))}
        rec = {"offered_per_s": rate, "completed": n, "achieved_per_s": round(n / span, 2),
               "p50_ms": pct(lat, 0.5), "p99_ms": pct(lat, 0.99), "by_bits": per_size,
               "peak_resident_mb_so_far": round(st["peak_bytes"] / 2**20, 3)}
        log(f"rate {rate}/s: {rec}")
        return rec, st["peak_bytes"]

    out = []
    srv = m.Server(ctx, budget_bytes=64 << 30, slots=3)
    for ri, rate in enumerate(rates):
        out.append(serve(srv, rate, ri)[0])
    srv.close()
    tight = 8 << 20
    srv = m.Server(ctx, budget_bytes=tight, slots=3)
    rec, peak = serve(srv, 200, rates.index(200))
    srv.close()
    assert peak <= tight, f"resident bytes {peak} above the budget {tight}"
    rec["budget_mb"] = tight / 2**20
    ctx.close()
    print(json.dumps({
        "metric": "serving: p50/p99 latency and GCD/s at offered rates; bit-exact",
        "unit": "ms", "n_gpus": 1, "higher_is_better": False,
        "config": {"workload": f"Poisson arrivals, {n} requests per rate, sizes 70% 2k / 25% 50k / 5% 1M bits, "
                               "g = 10% of the length, 32-bit moduli, 64 GB budget, 3 launch slots (one for requests above 16,384 moduli); "
                               "then the 200/s arrivals under an 8 MiB budget"},
        "rates": out,
        "binding_budget": rec,
    }), flush=True)


# The reference's published one-pair times (ms; ModGCD-OneGPU.pdf p.6, Table 2), random pairs, L = 32.
PAPER_TABLE2 = {  # Kibit: (GTX 980 Ti, GTX 1080, Tesla P100, Tesla V100, GMP on a Xeon E5-2620)
    16: (9, 5, 8, 7, 0.6), 32: (16, 10, 15, 13, 1.7), 48: (22, 15, 20, 18, 2.6), 64: (33, 22, 27, 25, 3.8),
    80: (40, 27, 33, 31, 5.0), 96: (47, 32, 40, 35, 6.5), 112: (53, 38, 47, 40, 7.9), 128: (65, 45, 54, 46, 8.7),
    144: (78, 50, 61, 52, 10.9), 160: (88, 56, 85, 58, 12.9), 176: (98, 71, 98, 66, 14.9),
    192: (124, 80, 112, 73, 16.3), 208: (136, 87, 124, 81, 17.8), 224: (152, 103, 138, 94, 17.9),
    240: (170, 116, 152, 106, 20.1), 256: (199, 125, 164, 118, 22.5), 272: (220, 136, 177, 132, 24.1),
    288: (238, 158, 189, 143, 25.6), 304: (257, 174, 199, 154, 26.5), 320: (294, 186, 230, 165, 27.3),
    336: (321, 198, 261, 174, 30.6), 352: (342, 221, 276, 184, 32.9),
}


def _mpz_one_thread(O, m, u, v, reps):
    """Seconds per mpz_gcd(u, v) on one host thread (mean of reps calls) and the gcd's bit length."""
    limbs, u_off, nu, v_off, nv = m.pack_pairs([(u, v)] * reps)
    sec, gbits = O.time_mpz_gcd(limbs, u_off, nu, v_off, nv, 1)
    return sec / reps, int(gbits[0])


def _gpu_one_pair(ctx, u, v, reps):
    """Median wall seconds of the one-pair call (staging, launch, read-back, conversion), the last
    call's kernel seconds, its stats and its result."""
    wall = []
    g = None
    for _ in range(reps):
        t = time.perf_counter()
        g = ctx.gcd(u, v)
        wall.append(time.perf_counter() - t)
    return statistics.median(wall), ctx.last_kernel_seconds(), ctx.last_pair_stats(), g


def run_single(args, m, w):
    """One pair per call (GmpCudaDevice::gcd's path) beside mpz_gcd on one host thread, each pair timed
    separately, as the reference's harness does (testmodgcd.cpp:261-275):
      small: 8192-bit pairs (1024-bit odd g), seeds 0..S-1, plus coprime, u = v and v = 0; L = 32;
      large: ~1M-bit pairs (100k-bit g), seeds 0..S-1, at L = 32, 27, 22;
      sweep: 16..352 Kibit random pairs, 10 per size, L = 32, beside the reference's published table."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as O
    out = {"metric": "one-pair latency (ms) on one MI355X vs mpz_gcd on one host thread; bit-exact",
           "unit": "ms", "n_gpus": 1, "higher_is_better": False, "gmp": O.gmp_version(),
           "cpu_model": _cpu_model(), "timing": "GPU: median wall time of mgcd_gcd_limbs over the repeats "
           "(limbs staged, launch, digits converted, |G| read back); kernel_ms: the launch's HIP-event time; "
           "mpz: mean of repeated mpz_gcd calls on one thread, CLOCK_MONOTONIC_RAW"}
    ctxs = {}

    def ctx_of(L):
        if L not in ctxs:
            ctxs[L] = m.Context(L)
            if args.workgroup and args.slots:
                ctxs[L].set_option(m.OPT_SINGLE_WORKGROUP, args.workgroup)
                ctxs[L].set_option(m.OPT_SINGLE_SLOTS, args.slots)
            if args.relay:
                ctxs[L].set_option(m.OPT_RELAY_PARTS, args.relay)
        return ctxs[L]

    def one(L, name, u, v, reps_gpu, reps_cpu):
        ctx = ctx_of(L)
        gs, ks, st, g = _gpu_one_pair(ctx, u, v, reps_gpu)
        cs, gbits = _mpz_one_thread(O, m, u, v, reps_cpu)
        assert g.bit_length() == gbits and g == O.mpz_gcd(u, v), f"{name}: gcd differs from mpz_gcd"
        rec = {"pair": name, "bits": max(u.bit_length(), v.bit_length()), "gcd_bits": gbits,
               "gpu_ms": round(gs * 1e3, 3), "kernel_ms": round(ks * 1e3, 3), "mpz_ms": round(cs * 1e3, 4),
               "gpu_over_mpz": round(gs / cs, 3) if cs > 0 else None}
        if v != 0 and u != v:
            rec.update({"nq": st["nq"], "iters": st["iters"], "digits": st["digits"],
                        "tiling": list(st["tiling"]), "parts": st["parts"],
                        "us_per_iter": round(ks * 1e6 / max(1, st["iters"] + st["digits"]), 3)})
        log(f"L={L} {name}: {rec}")
        return rec

    def summary(recs):
        return {"median_gpu_ms": round(statistics.median(r["gpu_ms"] for r in recs), 3),
                "median_mpz_ms": round(statistics.median(r["mpz_ms"] for r in recs), 4),
                "gpu_faster_than_mpz": sum(r["gpu_ms"] < r["mpz_ms"] for r in recs), "pairs": recs}

    if args.parts in ("all", "small"):
        u, v = w.single_small(0)
        ctx_of(32).gcd(u, v)  # first call: buffers and code objects
        recs = [one(32, f"seed {s}", *w.single_small(s), 5, 200) for s in range(args.seeds)]
        edges = w.single_small_edges(0)
        recs += [one(32, name, u, v, 5, 200) for name, (u, v) in zip(("coprime", "u == v", "v == 0"), edges)]
        out["small_8192"] = {"L": 32, "config": "g 1024-bit odd, a, b 7168-bit, u = g a, v = g b; edge pairs",
                             **summary(recs)}
    if args.parts in ("all", "large"):
        res = {}
        for L in [int(x) for x in args.Ls.split(",")]:
            u, v = w.single_large(0)
            recs = [one(L, f"seed {s}", *w.single_large(s), 1, 1) for s in range(args.seeds)]
            res[str(L)] = summary(recs)
        out["large_1M"] = {"config": "g 100k-bit, a, b 900k-bit, u = g a, v = g b (~1M bits)", "by_L": res}
    if args.parts in ("all", "sweep"):
        rows = []
        ctx = ctx_of(32)
        for kib in sorted(PAPER_TABLE2):
            nbits = kib * 1024
            ps = [w.pair(w.Stream(8, kib * 100 + i), nbits, 0, "coprime") for i in range(10)]
            gms, kms, cms = [], [], []
            for u, v in ps:
                gs, ks, st, g = _gpu_one_pair(ctx, u, v, 1)
                cs, gbits = _mpz_one_thread(O, m, u, v, 3)
                assert g.bit_length() == gbits and g == math.gcd(u, v), f"{kib} Kibit: gcd differs"
                gms.append(gs)
                kms.append(ks)
                cms.append(cs)
            paper = PAPER_TABLE2[kib]
            rows.append({"kibit": kib, "gpu_ms": round(statistics.mean(gms) * 1e3, 3),
                         "kernel_ms": round(statistics.mean(kms) * 1e3, 3),
                         "mpz_ms": round(statistics.mean(cms) * 1e3, 3),
                         "paper_v100_ms": paper[3], "paper_gtx1080_ms": paper[1], "paper_gmp_ms": paper[4],
                         "nq": st["nq"], "parts": st["parts"]})
            log(f"sweep {kib} Kibit: {rows[-1]}")
        out["sweep_kibit"] = {"L": 32, "config": "random pairs (gcd ~ 1), 10 per size, mean; paper: PDF p.6 "
                                                 "Table 2 (mean of 10 pairs per size)", "rows": rows}
    for c in ctxs.values():
        c.close()
    print(json.dumps(out), flush=True)


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


if __name__ == "__main__":
    main()
