"""Writes tests/golden/workload_digests.json: the SHA-256 of every seeded
workload that the benchmark and the GPU tests draw from
modgcd-onegpu_amd/python/workload.py, so that any change to the generator
(and with it to the inputs behind a published figure) shows up in the CPU
suite.

A workload's digest is SHA-256 over its pairs in order, each written as
u (little-endian bytes) | v (little-endian bytes) newline; zero is one 0 byte.

Run from the repository root: python tests/golden/make_workload_digests.py
"""
import hashlib
import json
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, os.path.join(ROOT, "modgcd-onegpu_amd", "python"))
import workload as w  # noqa: E402

SERVE_RATES = 5  # bench.py's serving legs draw seeds 5.. (sizes) and 100.. (pairs)
SERVE_N = 300


def digest(pairs):
    h = hashlib.sha256()
    for u, v in pairs:
        h.update(u.to_bytes((u.bit_length() + 7) // 8 or 1, "little"))
        h.update(b"|")
        h.update(v.to_bytes((v.bit_length() + 7) // 8 or 1, "little"))
        h.update(b"\n")
    return h.hexdigest()


def workloads():
    """name -> (description, function returning the pairs)."""
    def serve():
        out = []
        for ri in range(SERVE_RATES):
            sizes = w.serve_sizes(SERVE_N, seed=5 + ri)
            out += [w.serve_pair(100 + ri, i, nb) for i, nb in enumerate(sizes)]
        return out
    return {
        "single_small": ("8192-bit pairs, seeds 0-9, then the coprime, u == v and v == 0 edges",
                         lambda: [w.single_small(s) for s in range(10)] + w.single_small_edges(0)),
        "single_large": ("~1M-bit pairs (100k-bit g), seeds 0-9", lambda: [w.single_large(s) for s in range(10)]),
        "headline": ("4096 x 100k-bit pairs, 10k-bit g, seed 7", lambda: w.headline(4096)),
        "mixed": ("4096 pairs, 1k-200k bits log-uniform, seed 3", lambda: w.mixed()),
        "skewed": ("16384 pairs, 95% 4k / 5% 500k bits, g 20%, seed 4", lambda: w.skewed()),
        "serve": (f"{SERVE_RATES} x {SERVE_N} serving requests (70% 2k / 25% 50k / 5% 1M bits)", serve),
    }


def main():
    out = {}
    for name, (desc, f) in workloads().items():
        pairs = f()
        out[name] = {"pairs": len(pairs), "what": desc, "sha256": digest(pairs)}
        print(name, out[name]["pairs"], out[name]["sha256"])
    with open(os.path.join(ROOT, "tests", "golden", "workload_digests.json"), "w") as fh:
        json.dump(out, fh, indent=1)
        fh.write("\n")


if __name__ == "__main__":
    main()
