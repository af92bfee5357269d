"""Pins the CPU oracle (no GPU): every function against something it does
not share code with - the reference's own createModuli (built from its
sources under oracle/_ref, or its committed fixture), Python integers,
GMP, closed forms, and a brute-force Weber run on exact integers."""
import json
import math
import os
import random

import numpy as np
import pytest

import weber_bruteforce as wb

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def is_prime32(n):
    """Deterministic Miller-Rabin for n < 2^32 (bases 2, 7, 61)."""
    if n < 2:
        return False
    for p in (2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37, 41, 43, 47, 53, 59, 61):
        if n % p == 0:
            return n == p
    d, s = n - 1, 0
    while d % 2 == 0:
        d //= 2
        s += 1
    for a in (2, 7, 61):
        x = pow(a, d, n)
        if x in (1, n - 1):
            continue
        for _ in range(s - 1):
            x = x * x % n
            if x == n - 1:
                break
        else:
            return False
    return True


def cw_usable(D, L, W=64):
    """Cavagnino-Werbrouck condition with Python integers."""
    FC = 1 << (W + L - 1)
    J = FC // D + 1
    Qcr = -(-J // (D * J - FC))
    return (Qcr * D - 1).bit_length() > W


def fnv1a64(arr):
    h = 0xcbf29ce484222325
    for b in np.asarray(arr, np.uint32).tobytes():
        h = ((h ^ b) * 0x100000001b3) & 0xFFFFFFFFFFFFFFFF
    return h


# ---------------------------------------------------------------- moduli ----
def test_ref_moduli_fixture_matches_oracle(oracle):
    fx = json.load(open(os.path.join(GOLDEN, "ref_moduli32.json")))
    ours = oracle.create_moduli(32, fx["count"])
    assert len(ours) == fx["count"] == 65536
    for i, m in fx["samples"]:
        assert int(ours[i]) == m
    assert ours[: len(fx["head"])].tolist() == fx["head"]
    assert fnv1a64(ours) == int(fx["fnv1a64"], 16)


@pytest.mark.parametrize("L,count", [(22, 97358), (27, 3000), (32, 3000)])
def test_oracle_moduli_prime_usable_descending_complete(oracle, L, count):
    mods = oracle.create_moduli(L, count + 10 if L != 22 else 200000)
    if L == 22:
        assert len(mods) == 97358  # every usable 22-bit prime
    mods = [int(x) for x in mods[:count]]
    assert all(a > b for a, b in zip(mods, mods[1:]))
    check = mods if L != 22 else mods[:: 7]
    for m in check:
        assert (1 << (L - 1)) < m < (1 << L) and is_prime32(m) and cw_usable(m, L), m
    # completeness: nothing usable was skipped above the 500th entry
    top = (1 << L) - 1
    head = set(mods[:501])
    for x in range(top, mods[min(500, len(mods) - 1)], -2):
        if x not in head:
            assert not (is_prime32(x) and cw_usable(x, L)), x


def test_cw_criterion_agrees(oracle):
    rnd = random.Random(1)
    for L in (22, 27, 32):
        for _ in range(300):
            D = rnd.randrange((1 << (L - 1)) + 1, 1 << L) | 1
            assert bool(oracle.lib().oracle_cw_usable(D, L)) == cw_usable(D, L)


# ---------------------------------------------------------------- sizing ----
SURVEY_NQ = {  # N_Q at L = 32 / 27 / 22 (numModuliNeededFor, GmpCudaDevice-gcd.cu:49-55)
    1024: (256, 256, 256), 4096: (768, 768, 768), 8192: (1024, 1280, 1280),
    50000: (5376, 5632, 5888), 100000: (9984, 10496, 11264), 200000: (18432, 19712, 20992),
    500000: (42752, 45568, 48640), 1000000: (81152, 86528, 92160),
}


@pytest.mark.parametrize("bits", sorted(SURVEY_NQ))
def test_num_moduli_needed_values(oracle, mgcd, bits):
    for L, want in zip((32, 27, 22), SURVEY_NQ[bits]):
        assert oracle.num_moduli_needed(bits, L) == want
        assert mgcd.num_moduli_needed(bits, L) == want


def test_num_moduli_needed_formula_range(oracle, mgcd):
    for bits in list(range(2, 5000, 37)) + list(range(5000, 3000000, 9973)):
        for L in (22, 27, 32):
            n = oracle.num_moduli_needed(bits, L)
            assert n == mgcd.num_moduli_needed(bits, L)
            exact = (1.6 - 0.015 * L) * bits / math.log(bits)
            assert n % 256 == 0 and n >= exact - 1 and n < exact + 257


# ------------------------------------------------------------ arithmetic ----
@pytest.mark.parametrize("L", [22, 27, 32])
def test_mod_and_mod_mul(oracle, L):
    rnd = random.Random(L)
    mods = oracle.create_moduli(L, 50)
    lib = oracle.lib()
    for m in mods[::10]:
        m = int(m)
        for x in [0, 1, m - 1, m, m + 1, (1 << 64) - 1, (m - 1) * (m - 1)] + [rnd.getrandbits(64) for _ in range(2000)]:
            assert lib.oracle_mod(x, m, L) == x % m
        for _ in range(2000):
            a, b = rnd.randrange(m), rnd.randrange(m)
            assert lib.oracle_mod_mul(a, b, m, L) == a * b % m


def test_cw_inverse_closed_form(oracle):
    for L in (22, 27, 32):
        for m in oracle.create_moduli(L, 20):
            m = int(m)
            assert oracle.lib().oracle_cw_inverse(m, L) == (1 << (64 + L - 1)) // m + 1


def test_mod_inv_exhaustive_small(oracle):
    for m in (3, 5, 7, 11, 13, 251, 257, 4093, 65521):
        for v in range(1, m):
            assert oracle.mod_inv(v, m) * v % m == 1


@pytest.mark.parametrize("L", [22, 27, 32])
def test_mod_inv_random(oracle, L):
    rnd = random.Random(100 + L)
    mods = oracle.create_moduli(L, 200)
    for _ in range(20000):
        m = int(mods[rnd.randrange(len(mods))])
        v = rnd.randrange(1, m)
        assert oracle.mod_inv(v, m) == pow(v, -1, m)
    # Fibonacci-like worst cases (longest Euclid chains)
    for m in mods[:5]:
        m = int(m)
        a, b = 1, 1
        while b < m:
            assert oracle.mod_inv(b, m) == pow(b, -1, m)
            a, b = b, a + b


def test_mod_mp(oracle):
    rnd = random.Random(7)
    for L in (22, 27, 32):
        mods = oracle.create_moduli(L, 8)
        for bits in (1, 31, 32, 33, 64, 100, 1000, 33 * 32, 40000):
            x = rnd.getrandbits(bits)
            for m in mods:
                assert oracle.mod_mp(x, int(m), L) == x % int(m)


# ---------------------------------------------------------------- kernel ----
def _small_pair(rnd, bits, gbits):
    g = rnd.getrandbits(gbits) | (1 << (gbits - 1)) | 1
    a = rnd.getrandbits(bits - gbits) | (1 << (bits - gbits - 1))
    b = rnd.getrandbits(bits - gbits) | (1 << (bits - gbits - 1))
    u, v = g * a, g * b
    return (u, v) if u >= v else (v, u)


@pytest.mark.parametrize("L,bits,gbits,seed", [(32, 600, 100, 1), (32, 2000, 300, 2), (27, 1500, 700, 3),
                                               (22, 1200, 64, 4), (32, 3000, 1, 5), (22, 2500, 2000, 6)])
def test_kernel_matches_bruteforce(oracle, L, bits, gbits, seed):
    rnd = random.Random(seed)
    u, v = _small_pair(rnd, bits, gbits)
    nq = oracle.num_moduli_needed(u.bit_length(), L)
    mods = [int(x) for x in oracle.create_moduli(L, nq)]
    st, digits, pivots = oracle.gcd_kernel(u, v, np.array(mods, np.uint32), L)
    bp, bd = wb.weber(u, v, mods)
    assert st == oracle.OK
    assert pivots == bp
    assert digits == bd
    assert oracle.horner(digits) == math.gcd(u, v) == oracle.mpz_gcd(u, v)


def _top_primes(L, n):
    out, x = [], (1 << L) - 1
    while len(out) < n:
        if is_prime32(x):
            out.append(x)
        x -= 2
    return out


@pytest.mark.parametrize("L,bits,gbits,seed", [(32, 2000, 300, 12), (27, 1500, 700, 13), (22, 1200, 64, 14)])
def test_exact_kernel_equals_reference_kernel_on_cw_tables(oracle, L, bits, gbits, seed):
    rnd = random.Random(seed)
    u, v = _small_pair(rnd, bits, gbits)
    nq = oracle.num_moduli_needed(u.bit_length(), L)
    mods = np.array(oracle.create_moduli(L, nq), np.uint32)
    assert oracle.gcd_kernel(u, v, mods, L, exact=True) == oracle.gcd_kernel(u, v, mods, L)


@pytest.mark.parametrize("L,bits,gbits,seed", [(32, 2000, 300, 15), (27, 1500, 700, 16), (22, 1200, 64, 17)])
def test_exact_kernel_matches_bruteforce_over_all_primes(oracle, L, bits, gbits, seed):
    # "All primes" moduli (no CW filter): the exact-reduction kernel must
    # still follow the brute-force Weber run step by step.
    rnd = random.Random(seed)
    u, v = _small_pair(rnd, bits, gbits)
    nq = oracle.num_moduli_needed(u.bit_length(), L)
    mods = _top_primes(L, nq)
    assert any(not cw_usable(m, L) for m in mods)
    st, digits, pivots = oracle.gcd_kernel(u, v, np.array(mods, np.uint32), L, exact=True)
    bp, bd = wb.weber(u, v, mods)
    assert st == oracle.OK
    assert pivots == bp
    assert digits == bd
    assert oracle.horner(digits) == math.gcd(u, v)


def test_kernel_degenerate_and_failures(oracle):
    L = 32
    mods = oracle.create_moduli(L, 4096)
    rnd = random.Random(9)
    u, v = _small_pair(rnd, 8192, 1024)
    nq = oracle.num_moduli_needed(u.bit_length(), L)
    # v == 0: the reference kernel runs out of moduli in recovery
    st, _, _ = oracle.gcd_kernel(u, 0, mods[:nq], L)
    assert st == oracle.RECOVERY_ERROR
    # u == v: one elimination step then recovery of u
    st, d, t = oracle.gcd_kernel(u, u, mods[:nq], L)
    assert st == oracle.OK and oracle.horner(d) == u and len(t) == 1
    # too few moduli: the reduction loop runs out
    st, _, _ = oracle.gcd_kernel(u, v, mods[:256], L)
    assert st == oracle.REDUX_ERROR


def test_golden_kernel_vectors(oracle):
    fx = json.load(open(os.path.join(GOLDEN, "kernel_vectors.json")))
    for case in fx["cases"]:
        L = case["L"]
        u, v = int(case["u"], 16), int(case["v"], 16)
        mods = oracle.create_moduli(L, case["nq"])
        st, d, t = oracle.gcd_kernel(u, v, mods, L)
        assert st == case["status"]
        assert [list(x) for x in t] == case["pivots"]
        assert [list(x) for x in d] == case["digits"]


def test_mpz_gcd_and_baseline_timer(oracle, mgcd):
    rnd = random.Random(3)
    pairs = [_small_pair(rnd, 5000, 700) for _ in range(20)] + [(12, 0), (0, 12), (1, 1)]
    for u, v in pairs:
        assert oracle.mpz_gcd(u, v) == math.gcd(u, v)
    sec, gbits = oracle.time_mpz_gcd(*mgcd.pack_pairs(pairs), threads=2)
    assert sec > 0
    assert [int(x) for x in gbits] == [math.gcd(u, v).bit_length() for u, v in pairs]
