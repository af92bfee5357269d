"""Where a gcd launch's time goes, from the instrumented library.

Runs workloads through lib/libmgcd_phases.so (built with -DMGCD_PHASE_CLOCKS:
every work item of the gcd kernel sums, as its wave 0 sees them, the shader
cycles of each phase) and prints one JSON object: per workload the cycles of
each phase per work item and per elimination round, converted to microseconds
with the clock the real-time counter gives.

    python modgcd-onegpu_amd/tools/phase_breakdown.py [--pairs N] [--tilings T,K,B ...] [--no-single]

Workloads: 100k-bit pairs (10k-bit shared factor, L = 32) batched at each
tiling given (default (512, 2, 16)); one 100k-bit pair and one ~1M-bit pair
(100k-bit shared factor) on the one-pair path (skipped with --no-single).
"""
import argparse
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.dirname(HERE)
ROOT = os.path.dirname(PKG)
os.environ["MGCD_LIB"] = os.path.join(PKG, "lib", "libmgcd_phases.so")
sys.path.insert(0, ROOT)

import __graft_entry__ as ge  # noqa: E402


def summarize(c):
    ghz = c["memtime_cycles"] / (c["realtime_ticks"] * 10.0) if c["realtime_ticks"] else float("nan")
    us = lambda cyc: cyc / ghz / 1e3  # noqa: E731
    items, rounds = max(c["items"], 1), max(c["rounds"], 1)
    names = ["residues", "elim_compute", "elim_barrier", "elim_exchange", "elim_bookkeeping", "elim_barrier2",
             "elim_tail", "replay", "recovery"]
    tot = sum(c[n] for n in names)
    out = {"clock_ghz": round(ghz, 3), "items": c["items"], "rounds_per_item": round(c["rounds"] / items, 1),
           "item_us": round(us(tot / items), 2),
           "per_item_us": {n: round(us(c[n] / items), 2) for n in names},
           "share": {n: round(c[n] / tot, 4) for n in names},
           "per_round_us": {n: round(us(c[n] / rounds), 3) for n in names if n.startswith("elim_") and n != "elim_tail"}}
    out["per_round_us"]["total"] = round(sum(out["per_round_us"].values()), 3)
    waves = [i for i, x in enumerate(c.get("wave_compute", [])) if x]
    if waves:
        # every wave's own compute and barrier wait per round
        out["per_wave_round_us"] = {"compute": [round(us(c["wave_compute"][i] / rounds), 3) for i in waves],
                                    "barrier": [round(us(c["wave_barrier"][i] / rounds), 3) for i in waves]}
    if c.get("collects"):
        # one-pass collects: the slot loads' round trip, and how often they found every slot posted
        out["collect_first_load_us"] = round(us(c["first_load_cycles"] / c["collects"]), 3)
        out["collect_ready_at_first_load"] = round(c["collects_ready"] / c["collects"], 4)
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--pairs", type=int, default=2048)
    ap.add_argument("--tilings", nargs="*", default=["512,2,16"])
    ap.add_argument("--no-single", action="store_true")
    args = ap.parse_args()
    m = ge.package()
    w = ge.workload()
    res = {"library": os.path.relpath(m.LIB_PATH, ROOT)}
    ctx = m.Context(32)
    pairs = w.headline(args.pairs, seed=7)
    for tl in args.tilings:
        T, K, B = (int(x) for x in tl.split(","))
        b = m.Batch.from_pairs(ctx, pairs, workgroup=T, slots=K, group=B)
        b.run()
        ctx.phase_clocks(reset=True)
        b.run()
        dev = b.timing()[0]
        res[f"batched_100k_{T}x{K}x{B}"] = {"pairs": args.pairs, "launch_ms": round(dev * 1e3, 2),
                                            **summarize(ctx.phase_clocks(reset=True))}
        b.close()
        print(f"batched {tl} done", file=sys.stderr, flush=True)
    for name, (u, v) in (() if args.no_single else (("one_pair_100k", pairs[0]), ("one_pair_1M", w.single_large(0)))):
        ctx.gcd(u, v)
        ctx.phase_clocks(reset=True)
        ctx.gcd(u, v)
        st = ctx.last_pair_stats()
        res[name] = {"kernel_ms": round(ctx.last_kernel_seconds() * 1e3, 2), "tiling": list(st["tiling"]),
                     "parts": st["parts"], "iters": st["iters"], "digits": st["digits"],
                     **summarize(ctx.phase_clocks(reset=True))}
        print(f"{name} done", file=sys.stderr, flush=True)
    ctx.close()
    print(json.dumps(res, indent=1))


if __name__ == "__main__":
    main()
