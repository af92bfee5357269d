"""Writes the committed golden fixtures of tests/golden/.

ref_moduli32.json   the output of the reference's own createModuli
                    (createModuli.cpp, L = 32, NUM_MODULI = 65536), built
                    from its sources by `make -C oracle ref`: count,
                    FNV-1a-64 of the array, the first 64 entries and every
                    1024th entry.
kernel_vectors.json seeded small pairs with the oracle's pivot sequence and
                    mixed-radix digits; each case is checked against the
                    exact-integer Weber run (tests/weber_bruteforce.py) and
                    GMP before it is written.

Run from the repository root: python tests/golden/make_fixtures.py
"""
import json
import math
import os
import random
import re
import subprocess
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, os.path.join(ROOT, "tests"))
import oracle_lib as O  # noqa: E402
import weber_bruteforce as wb  # noqa: E402

OUT = os.path.join(ROOT, "tests", "golden")


def fnv1a64(arr):
    h = 0xcbf29ce484222325
    for b in np.asarray(arr, np.uint32).tobytes():
        h = ((h ^ b) * 0x100000001b3) & 0xFFFFFFFFFFFFFFFF
    return h


def ref_moduli():
    exe = os.path.join(ROOT, "oracle", "_ref", "createModuli")
    if not os.path.exists(exe):
        subprocess.run(["make", "-C", os.path.join(ROOT, "oracle"), "ref"], check=True)
    out = subprocess.run([exe], check=True, capture_output=True, text=True).stdout
    mods = [int(x) for x in re.findall(r"^\s+(\d+),", out, re.M)]
    fx = {"source": "reference createModuli.cpp (L = 32, W = 64, NUM_MODULI = 65536), built by make -C oracle ref",
          "count": len(mods), "fnv1a64": "%016x" % fnv1a64(mods), "head": mods[:64],
          "samples": [[i, mods[i]] for i in range(0, len(mods), 1024)] + [[len(mods) - 1, mods[-1]]]}
    json.dump(fx, open(os.path.join(OUT, "ref_moduli32.json"), "w"), indent=1)
    print("ref_moduli32.json:", len(mods), "moduli")


def pair(rnd, bits, gbits):
    g = rnd.getrandbits(gbits) | (1 << (gbits - 1)) | 1 if gbits > 0 else 1
    ab = bits - max(gbits, 0)
    a = rnd.getrandbits(ab) | (1 << (ab - 1))
    b = rnd.getrandbits(ab) | (1 << (ab - 1))
    u, v = g * a, g * b
    return (u, v) if u >= v else (v, u)


def kernel_vectors():
    specs = [(32, 8192, 1024, 0), (32, 3000, 0, 1), (27, 2000, 500, 2), (22, 2000, 300, 3),
             (32, 1200, 1100, 4), (22, 900, 1, 5)]
    cases = []
    for L, bits, gbits, seed in specs:
        rnd = random.Random(1000 + seed)
        u, v = pair(rnd, bits, gbits)
        nq = O.num_moduli_needed(u.bit_length(), L)
        mods = O.create_moduli(L, nq)
        st, d, t = O.gcd_kernel(u, v, mods, L)
        bp, bd = wb.weber(u, v, [int(x) for x in mods])
        assert st == O.OK and t == bp and d == bd and O.horner(d) == math.gcd(u, v) == O.mpz_gcd(u, v)
        cases.append({"L": L, "bits": bits, "gcd_bits": gbits, "seed": seed, "nq": nq, "u": "%x" % u,
                      "v": "%x" % v, "status": st, "pivots": [list(x) for x in t], "digits": [list(x) for x in d],
                      "gcd": "%x" % math.gcd(u, v)})
        print(f"case L={L} bits={bits}: R={len(t)} k={len(d)}")
    # u == v
    rnd = random.Random(77)
    u = rnd.getrandbits(5000) | (1 << 4999)
    nq = O.num_moduli_needed(u.bit_length(), 32)
    st, d, t = O.gcd_kernel(u, u, O.create_moduli(32, nq), 32)
    assert st == O.OK and O.horner(d) == u
    cases.append({"L": 32, "bits": 5000, "gcd_bits": 5000, "seed": 77, "nq": nq, "u": "%x" % u, "v": "%x" % u,
                  "status": st, "pivots": [list(x) for x in t], "digits": [list(x) for x in d], "gcd": "%x" % u})
    json.dump({"generator": "tests/golden/make_fixtures.py (oracle/mgcd_oracle.c, cross-checked by "
                            "tests/weber_bruteforce.py and GMP)", "cases": cases},
              open(os.path.join(OUT, "kernel_vectors.json"), "w"))


if __name__ == "__main__":
    ref_moduli()
    kernel_vectors()
