"""Mixed-radix digits -> positional limbs on the device (the product tree that
replaces the host Horner loop, GmpCudaDevice-gcd.cu:113-129).

The expected value is the digits' defining sum, G = sum g_i prod_{j<i} m_j,
evaluated with Python integers: it pins the conversion independently of the
GCD kernel.  Digit counts straddle every power of two the tree splits at,
the one-thread-per-node / one-wave-per-chunk boundary (nodes of 4 and 8
digits), the first level of more than one 64-limb chunk per node (64 and
65 digits), the size of a 1M-bit pair's GCD (about 3,100 digits), and
nodes of more than 64 chunks, whose carries are chained 64 chunks at a
time (4,500 and 8,200 digits)."""
import math
import random

import pytest

pytestmark = pytest.mark.gpu

KS = [1, 2, 3, 4, 5, 7, 8, 9, 15, 16, 17, 31, 32, 33, 63, 64, 65, 127, 128, 129, 255, 256, 257, 1000, 1024, 1025, 3100,
      4500, 8200]


def _digits(moduli, k, rnd, mode):
    out = []
    for i in range(k):
        m = int(moduli[i % len(moduli)])
        lo, hi = -(m + 1) // 2, (m - 3) // 2  # the kernel's toSigned range
        if mode == "min":
            g = lo
        elif mode == "max":
            g = hi
        elif mode == "alt":
            g = lo if i % 2 else hi
        else:
            g = rnd.randint(lo, hi)
        out.append((m, g))
    return out


def _value(digits):
    v = 0
    for m, g in reversed(digits):
        v = v * m + g
    return abs(v)


@pytest.mark.parametrize("mode", ["random", "min", "max", "alt"])
def test_digit_sequences(ctx_for, mode):
    ctx = ctx_for(32)
    moduli = ctx.moduli()
    rnd = random.Random(71)
    for k in KS:
        d = _digits(moduli, k, rnd, mode)
        assert ctx.mixed_radix(d) == _value(d), (mode, k)


def test_zero_and_sparse_digits(ctx_for):
    ctx = ctx_for(32)
    moduli = ctx.moduli()
    rnd = random.Random(72)
    for k in (1, 17, 33, 300):
        d = [(int(moduli[i]), 0) for i in range(k)]
        assert ctx.mixed_radix(d) == 0
        # one nonzero digit anywhere, of either sign
        for pos in {0, k // 2, k - 1}:
            for g in (1, -1, (int(moduli[pos]) - 3) // 2, -(int(moduli[pos]) + 1) // 2):
                e = list(d)
                e[pos] = (e[pos][0], g)
                assert ctx.mixed_radix(e) == _value(e), (k, pos, g)
        # top digit zero, lower ones random: leading zero limbs are trimmed
        e = _digits(moduli, k, rnd, "random")
        e[-1] = (e[-1][0], 0)
        assert ctx.mixed_radix(e) == _value(e)


def test_carries_ripple_across_chunks(ctx_for):
    # g_0 = -1 and g_w = 1 make the node [0, 2w) equal P(0, w) - 1: every
    # column above P_lo's top limb reads V_lo's sign extension (all ones), so
    # the carry out of P_lo's top column ripples through all of them, across
    # every chunk (and, at 8,200 digits, across a batch of 64 chunks) of that
    # node; g_0 = 1, g_w = -1 runs the same with borrows.
    ctx = ctx_for(32)
    moduli = ctx.moduli()
    for k in (300, 3100, 8200):
        base = [(int(moduli[i]), 0) for i in range(k)]
        w = 1
        while w < k:
            for a, b in ((-1, 1), (1, -1)):
                e = list(base)
                e[0] = (e[0][0], a)
                e[w] = (e[w][0], b)
                assert ctx.mixed_radix(e) == _value(e), (k, w, a, b)
            w *= 2


def _signed_digits(n, moduli):
    # n in mixed radix over the moduli, digits in the kernel's toSigned range
    out = []
    i = 0
    while n != 0:
        m = int(moduli[i])
        r = n % m
        if r > (m - 3) // 2:
            r -= m
        out.append((m, r))
        n = (n - r) // m
        i += 1
    return out


def test_structured_values(ctx_for):
    # Values whose limbs are long runs of ones or zeros (2^n - 1, 2^n, ...):
    # a chunk's carry out, which holds a column's high-half sum and can pass
    # 2^32, lands on such runs and must ripple through them, across chunks
    # and batches of 64 chunks.  Either sign of G.
    ctx = ctx_for(32)
    moduli = ctx.moduli()
    for n in (2100, 6000, 140000):  # 140,000 bits: about 4,380 digits, 69 chunks at the root
        for x in (2**n - 1, 2**n, 2**n + 1, 2**n - 2 ** (n // 2), 3 * 2**n - 1):
            for sign in (1, -1):
                d = _signed_digits(sign * x, moduli)
                assert ctx.mixed_radix(d) == x, (n, sign, len(d))


@pytest.mark.parametrize("L", [22, 27])
def test_narrow_moduli(ctx_for, L):
    ctx = ctx_for(L)
    moduli = ctx.moduli()
    rnd = random.Random(73 + L)
    for k in (1, 16, 33, 700):
        d = _digits(moduli, k, rnd, "random")
        assert ctx.mixed_radix(d) == _value(d), (L, k)


def test_batches_device_and_host_conversion_agree(ctx_for, mgcd, wl):
    ctx = ctx_for(32)
    pairs = wl.mixed(n=96, seed=74, lo=1000, hi=200000)
    want = [math.gcd(u, v) for u, v in pairs]
    got = {}
    for dc in (1, 0):
        ctx.set_option(mgcd.OPT_DEVICE_CONVERT, dc)
        try:
            b = mgcd.Batch.from_pairs(ctx, pairs)
            b.run()
            got[dc] = b.results()
            b.close()
        finally:
            ctx.set_option(mgcd.OPT_DEVICE_CONVERT, 1)
    assert got[1] == want
    assert got[0] == want


def test_one_pair_with_many_digits(ctx_for, wl):
    # a 40k-bit shared factor: about 1,300 digits, converted on the device
    ctx = ctx_for(32)
    rnd = random.Random(75)
    g = rnd.getrandbits(40000) | (1 << 39999) | 1
    u = g * (rnd.getrandbits(60000) | (1 << 59999))
    v = g * (rnd.getrandbits(59000) | (1 << 58999))
    assert ctx.gcd(u, v) == math.gcd(u, v)
