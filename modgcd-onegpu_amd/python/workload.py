"""Seeded synthetic GCD workloads.

Integers come from SplitMix64 streams: stream (seed, index) is the
SplitMix64 sequence started at state seed * 2^32 + index, so every pair is
reproducible on its own and on any machine.  A random n-bit operand is the
low n bits of consecutive outputs (little-endian words) with bit n-1 set.

Configurations (sizes in bits):
  single_small   g 1024-bit odd, a, b 7168-bit: u = g a, v = g b (~8192 bits),
                 plus edge pairs: coprime, u == v, v == 0
  single_large   g 100k-bit, a, b 900k-bit (~1M bits)
  mixed          lengths log-uniform in [1k, 200k], g bits uniform in
                 [0, len/2], 10% coprime, 1% v == 0
  skewed         95% 4k-bit, 5% 500k-bit, g = 20% of the length
  headline       100k-bit pairs with a 10k-bit shared factor
  serve          70% 2k, 25% 50k, 5% 1M bits, g = 10% of the length
"""
import math

import numpy as np

_GAMMA = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def splitmix64(state0, n):
    """First n outputs of SplitMix64 from state0 (numpy uint64 array)."""
    with np.errstate(over="ignore"):
        i = np.arange(1, n + 1, dtype=np.uint64)
        z = np.uint64(state0 & 0xFFFFFFFFFFFFFFFF) + i * _GAMMA
        z = (z ^ (z >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
        return z ^ (z >> np.uint64(31))


class Stream:
    """Sequential draws from one SplitMix64 stream."""

    def __init__(self, seed, index=0):
        self.state = ((seed & 0xFFFFFFFF) << 32) | (index & 0xFFFFFFFF)
        self.pos = 0

    def words(self, n):
        w = splitmix64(self.state + self.pos * 0x9E3779B97F4A7C15, n)
        self.pos += n
        return w

    def uniform(self):
        return float(self.words(1)[0] >> np.uint64(11)) / float(1 << 53)

    def bits(self, n, top=True, odd=False):
        """Uniform n-bit integer (bit n-1 set if top; bit 0 set if odd)."""
        if n <= 0:
            return 0
        w = self.words((n + 63) // 64)
        x = int.from_bytes(w.tobytes(), "little") & ((1 << n) - 1)
        if top:
            x |= 1 << (n - 1)
        if odd:
            x |= 1
        return x


def pair(st, nbits, gbits, kind="normal", g_odd=False):
    """One (u, v): u = g a, v = g b with g of gbits bits and a, b of
    nbits - gbits bits.  kind: normal | coprime (g = 1) | equal | vzero."""
    if kind == "coprime":
        gbits = 0
    gb = max(gbits, 0)
    g = st.bits(gb, odd=g_odd) if gb > 0 else 1
    ab = max(nbits - gb, 1)
    a = st.bits(ab)
    b = st.bits(ab)
    u, v = g * a, g * b
    if kind == "equal":
        v = u
    elif kind == "vzero":
        v = 0
    return u, v


def single_small(seed):
    st = Stream(seed, 0)
    return pair(st, 8192, 1024, g_odd=True)


def single_small_edges(seed):
    st = Stream(seed, 1)
    return [pair(st, 8192, 1024, "coprime"), pair(st, 8192, 1024, "equal", g_odd=True),
            pair(st, 8192, 1024, "vzero", g_odd=True)]


def single_large(seed, nbits=1000000, gbits=100000):
    st = Stream(seed, 0)
    return pair(st, nbits, gbits)


def headline(n, seed=7, nbits=100000, gbits=10000):
    return [pair(Stream(seed, i), nbits, gbits) for i in range(n)]


def mixed(n=4096, seed=3, lo=1000, hi=200000):
    out = []
    for i in range(n):
        st = Stream(seed, i)
        nb = int(round(math.exp(math.log(lo) + st.uniform() * (math.log(hi) - math.log(lo)))))
        r = st.uniform()
        kind = "coprime" if r < 0.10 else ("vzero" if r < 0.11 else "normal")
        gb = int(st.uniform() * 0.5 * nb)
        out.append(pair(st, nb, gb, kind))
    return out


def skewed(n=16384, seed=4, small=4096, large=500000, share=0.05, gfrac=0.20):
    out = []
    for i in range(n):
        st = Stream(seed, i)
        nb = large if st.uniform() < share else small
        out.append(pair(st, nb, int(gfrac * nb)))
    return out


def serve_sizes(n, seed=5):
    """Bit lengths of the serving mix: 70% 2k, 25% 50k, 5% 1M."""
    st = Stream(seed, 0xFFFFFFF0)
    sizes = []
    for _ in range(n):
        r = st.uniform()
        sizes.append(2000 if r < 0.70 else (50000 if r < 0.95 else 1000000))
    return sizes


def serve_pair(seed, i, nbits, gfrac=0.10):
    return pair(Stream(seed, i), nbits, int(gfrac * nbits))
