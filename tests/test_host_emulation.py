"""CPU runs of the library's own device arithmetic, built for the host with g++.

tests/host/emu.h turns the device qualifiers into plain inline functions and
replaces v_rcp_f32 by a model with the error the device reciprocal is
certified to keep (mgcd_certify_quasi_quotient mode 2: |rcp(f) f - 1| <=
2^-23, rcp(-f) = -rcp(f)): the correctly rounded reciprocal moved by up to
one ulp, in four ways.  Under it:

* the modular inverses Arith32::modinv / Arith22::modinv (csrc/mgcd_arith.h,
  the role of modInv.h:141-188) are checked against pow(y, -1, m) on random
  inputs, Fibonacci-like worst cases and operands that round to the same
  float as the modulus (a floor-quotient form of this loop had a q = 0 fixed
  point there and hung on the GPU);
* the residue Horner step (horner_step, reduced lazily, each limb added one
  step late) and the reduction after the last step give the exact residue
  of random, all-ones and
  single-limb operands for the smallest, largest and random moduli of the
  L = 22, 27 and 32 tables, and every intermediate value stays inside the
  step's bound;
* the exchange-key decoders (csrc/mgcd_keys.h) never yield a modulus index
  outside the pair's N_Q, whatever payload a slot holds (0, ABORT, a
  recovery key read as an elimination key, random bits), and round-trip
  every well-formed key.
"""
import os
import random
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HOST = os.path.join(ROOT, "tests", "host")
CSRC = os.path.join(ROOT, "modgcd-onegpu_amd", "csrc")
PHI = (1 + 5 ** 0.5) / 2


def _build(tmp_path, name):
    exe = str(tmp_path / name)
    subprocess.run(["g++", "-O2", "-std=c++17", "-ffp-contract=off", "-include", os.path.join(HOST, "emu.h"),
                    "-I", CSRC, os.path.join(HOST, name + ".cpp"), "-o", exe], check=True)
    return exe


@pytest.fixture(scope="module")
def tools(tmp_path_factory):
    d = tmp_path_factory.mktemp("host")
    return {n: _build(d, n) for n in ("keys_check", "modinv_check", "horner_check")}, d


def test_key_decoders(tools):
    exes, _ = tools
    out = subprocess.run([exes["keys_check"]], check=True, capture_output=True, text=True, timeout=120).stdout
    rows = [line.split() for line in out.strip().splitlines()]
    assert [r[0] for r in rows] == ["foreign_payloads", "special_keys", "round_trips", "order"]
    for name, bad, cases in rows:
        assert int(cases) > 0
        assert int(bad) == 0, f"{name}: {bad} of {cases}"


def _moduli(L, mgcd):
    path = os.path.join(mgcd.DATA_DIR, f"moduli_L{L}.bin")
    if not os.path.exists(path):
        pytest.skip("moduli tables not built")
    raw = np.fromfile(path, dtype=np.uint32)
    # 48-byte header (magic[8], version, L, W, count, checksum, reserved), then the moduli
    assert raw[2:4].tolist() == [1, L]
    table = raw[12:12 + int(raw[5])]
    rnd = random.Random(L)
    pick = list(table[:16]) + list(table[-8:]) + [table[rnd.randrange(len(table))] for _ in range(40)]
    return [int(m) for m in pick]


def _cases(L, moduli):
    """(L, m, y) with 0 < |y| <= m/2 + 2^15 (the inverse's input range), y != 0 mod m."""
    rnd = random.Random(1000 + L)
    lim = 1 << 15 if L > 22 else 1 << 12
    out = []
    for m in moduli:
        ys = [rnd.randrange(-(m // 2 + lim), m // 2 + lim + 1) for _ in range(1500)]
        # Fibonacci-like worst cases: ratios near 1/phi, 1/phi^2, their neighbours
        for c in (1 / PHI, 1 / PHI ** 2, 2 ** 0.5 - 1, 1 / 3, 2 / 5):
            y0 = int(round(m * c))
            ys += [y0 + d for d in range(-40, 41)]
        # operands that round to the same float as m (and as m/2): |m - y| small
        ys += [m - d for d in range(1, 400)] + [d - m for d in range(1, 400)]
        ys += [(m + 1) // 2 + d for d in range(-300, 300)] + [-((m + 1) // 2) + d for d in range(-300, 300)]
        # powers of two and their neighbours, small values
        ys += [(1 << k) + d for k in range(1, L) for d in (-1, 0, 1)]
        ys += list(range(-50, 51))
        for y in ys:
            if y % m and abs(y) <= m // 2 + lim:
                out.append((L, m, y))
    return out


@pytest.mark.parametrize("L", [32, 27, 22])
@pytest.mark.parametrize("mode", [0, 1, 2, 3])
def test_modinv_matches_pow(tools, mgcd, L, mode):
    exes, d = tools
    cases = _cases(L, _moduli(L, mgcd))
    inp, outp = d / f"in_{L}_{mode}.bin", d / f"out_{L}_{mode}.bin"
    np.array(cases, dtype=np.int64).tofile(inp)
    res = subprocess.run([exes["modinv_check"], str(inp), str(outp), str(mode)], check=True, capture_output=True,
                         text=True, timeout=300).stdout.split()
    assert res[0] == "checked" and int(res[1]) == len(cases) and res[2] == "wrong"
    assert int(res[3]) == 0, f"{res[3]} wrong inverses of {len(cases)}"
    got = np.fromfile(outp, dtype=np.int64)
    for i in range(0, len(cases), 97):  # the 128-bit check above covers all; pow() a sample
        _, m, y = cases[i]
        assert int(got[i]) == pow(y, -1, m), (m, y)


@pytest.mark.parametrize("L", [22, 27, 32])
def test_residue_horner_steps(tools, mgcd, L):
    exes, d = tools
    mods = _moduli(L, mgcd)
    rnd = random.Random(90 + L)
    ms = [int(mods[0]), int(mods[-1])] + [int(mods[rnd.randrange(len(mods))]) for _ in range(30)]
    recs = []
    for m in ms:
        for limbs in ([rnd.getrandbits(32) for _ in range(300)], [0xFFFFFFFF] * 200, [0] * 5 + [0xFFFFFFFF],
                      [0xFFFFFFFF] + [0] * 50, [rnd.getrandbits(32)], [m - 1, m, m + 1, 0xFFFFFFFF, 1]):
            recs += [m, len(limbs)] + limbs
    path = d / f"horner_{L}.bin"
    np.array(recs, dtype="<i8").tofile(path)
    out = subprocess.run([exes["horner_check"], str(path)], check=True, capture_output=True, text=True,
                         timeout=120).stdout.split()
    assert out[0] == "checked" and int(out[1]) == len(ms) * 6
    assert int(out[3]) == 0, f"L={L}: {out[3]} residues differ from the exact ones"
    assert int(out[5]) == 0, f"L={L}: {out[5]} intermediate values outside the step's bound"
