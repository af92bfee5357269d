"""Mixed-radix digits -> positional limbs on the device (the product tree that
replaces the host Horner loop, GmpCudaDevice-gcd.cu:113-129).

The expected value is the digits' defining sum, G = sum g_i prod_{j<i} m_j,
evaluated with Python integers: it pins the conversion independently of the
GCD kernel.  Digit counts straddle every power of two the tree splits at,
the one-thread-per-node / one-workgroup-per-node boundary (nodes of 16 and
32 digits), and the size of a 1M-bit pair's GCD (about 3,100 digits)."""
import math
import random

import pytest

pytestmark = pytest.mark.gpu

KS = [1, 2, 3, 4, 5, 15, 16, 17, 31, 32, 33, 63, 64, 65, 127, 128, 129, 255, 256, 257, 1000, 1024, 1025, 3100]


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
