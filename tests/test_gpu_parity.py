"""GPU parity: the HIP path against the oracle and GMP, through the C ABI.

The trace entry point returns the kernel's pivot sequence (p, b) and its
signed mixed-radix digits; they must equal the oracle's restatement of the
reference kernel element by element, and the gcd must equal mpz_gcd.
"""
import math
import random

import pytest

pytestmark = pytest.mark.gpu


def _check_trace(ctx, oracle, mgcd, u, v, L):
    nq = mgcd.num_moduli_needed(max(u.bit_length(), 2), L)
    st, g, digits, pivots = ctx.gcd_trace(u, v)
    ost, od, ot = oracle.gcd_kernel(u, v, ctx.moduli()[:nq], L)
    status_map = {oracle.OK: mgcd.OK, oracle.REDUX_ERROR: mgcd.ERR_REDUX, oracle.RECOVERY_ERROR: mgcd.ERR_RECOVERY}
    assert st == status_map[ost], (st, ost)
    assert pivots == ot, f"pivots differ (first mismatch at {next(i for i, (a, b) in enumerate(zip(pivots, ot)) if a != b) if any(a != b for a, b in zip(pivots, ot)) else min(len(pivots), len(ot))})"
    if st == mgcd.OK:
        assert digits == od
        assert g == oracle.mpz_gcd(u, v) == math.gcd(u, v)
    return len(pivots), len(digits)


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_trace_small_L32(ctx_for, oracle, mgcd, wl, seed):
    u, v = wl.single_small(seed)
    R, k = _check_trace(ctx_for(32), oracle, mgcd, u, v, 32)
    assert R > 100 and k > 10


@pytest.mark.parametrize("L", [32, 22])
def test_trace_operands_divisible_by_table_moduli(ctx_for, oracle, mgcd, wl, L):
    # Table moduli dividing v only (t starts at infinity there), both (0/0)
    # and u only (t starts at 0, the first pivots): the sentinel states and
    # their transitions against the oracle, pivot by pivot.
    ctx = ctx_for(L)
    mods = ctx.moduli()
    rnd = random.Random(81 + L)
    u, v = wl.single_small(3)
    p = [int(mods[i]) for i in rnd.sample(range(200), 5)]
    v = v * p[0] * p[1] * p[2] * p[3]
    u = u * p[3] * p[4] * (rnd.getrandbits(160) | (1 << 159))
    assert u > v
    _check_trace(ctx, oracle, mgcd, u, v, L)


@pytest.mark.parametrize("L", [27, 22])
def test_trace_small_other_widths(ctx_for, oracle, mgcd, wl, L):
    u, v = wl.single_small(0)
    _check_trace(ctx_for(L), oracle, mgcd, u, v, L)


@pytest.mark.parametrize("L", [32, 27, 22])
def test_edge_pairs(ctx_for, oracle, mgcd, wl, L):
    ctx = ctx_for(L)
    coprime, equal, vzero = wl.single_small_edges(0)
    for u, v in (coprime, equal):
        _check_trace(ctx, oracle, mgcd, u, v, L)
    # v == 0: the reference kernel runs out of moduli (oracle agrees); the
    # product path returns |u| like mpz_gcd.
    u, v = vzero
    _check_trace(ctx, oracle, mgcd, u, v, L)
    assert ctx.gcd(u, 0) == u
    assert ctx.gcd(0, u) == u
    assert ctx.gcd(0, 0) == 0


def test_trace_multi_workgroup(ctx_for, oracle, mgcd):
    # N_Q = 2304 > 1024: the pair is split over 9 workgroups that exchange
    # their minima every iteration through the slot barrier.
    rnd = random.Random(11)
    g = rnd.getrandbits(3000) | (1 << 2999)
    u = g * (rnd.getrandbits(17000) | (1 << 16999))
    v = g * (rnd.getrandbits(17000) | (1 << 16999))
    if u < v:
        u, v = v, u
    _check_trace(ctx_for(32), oracle, mgcd, u, v, 32)


def test_trace_multi_workgroup_relay(mgcd, oracle, wl):
    # The same kind of pair with the exchange relayed through part 0 (forced
    # for every multi-part pair): pivots and digits still equal the oracle's.
    ctx = mgcd.Context(32)
    ctx.set_option(mgcd.OPT_RELAY_PARTS, 2)
    try:
        rnd = random.Random(12)
        g = rnd.getrandbits(3000) | (1 << 2999)
        u = g * (rnd.getrandbits(17000) | (1 << 16999))
        v = g * (rnd.getrandbits(17000) | (1 << 16999))
        if u < v:
            u, v = v, u
        _check_trace(ctx, oracle, mgcd, u, v, 32)
        pairs = wl.headline(40, seed=13, nbits=20000, gbits=3000)
        b = mgcd.Batch.from_pairs(ctx, pairs, workgroup=256, slots=2, group=16)
        b.run()
        assert b.results() == [math.gcd(x, y) for x, y in pairs]
        b.close()
        # 0 restores the default thresholds; 1 is not a threshold.
        ctx.set_option(mgcd.OPT_RELAY_PARTS, 0)
        with pytest.raises(mgcd.MgcdError):
            ctx.set_option(mgcd.OPT_RELAY_PARTS, 1)
        assert ctx.gcd(u, v) == math.gcd(u, v)
    finally:
        ctx.close()


@pytest.mark.parametrize("L", [32, 22])
def test_small_and_ragged_sizes(ctx_for, mgcd, L):
    ctx = ctx_for(L)
    rnd = random.Random(5 + L)
    cases = [(1, 1), (2, 1), (6, 4), (2**32 - 5, 2**31 + 11), (2**64 + 13, 2**63), (12, 18)]
    for bits in (33, 64, 65, 100, 500, 1000, 1500, 3000, 5000):
        gb = rnd.randrange(1, bits)
        g = rnd.getrandbits(gb) | 1
        cases.append((g * rnd.getrandbits(bits), g * rnd.getrandbits(bits) + g))
    for u, v in cases:
        assert ctx.gcd(u, v) == math.gcd(u, v), (u.bit_length(), v.bit_length())


# The ~1M-bit one-pair configuration (g 100k-bit, a, b 900k-bit): seeds 0-9 at
# L = 32, two seeds each at L = 27 and L = 22, every result against mpz_gcd
# (the reference's check, testmodgcd.cpp:283-295).
@pytest.mark.parametrize("L,seed", [(32, s) for s in range(10)] + [(27, 0), (27, 1), (22, 0), (22, 1)])
def test_single_large_1M(ctx_for, oracle, wl, L, seed):
    u, v = wl.single_large(seed)
    g = ctx_for(L).gcd(u, v)
    assert g == oracle.mpz_gcd(u, v)
    assert g.bit_length() >= 100000


def test_single_large_1M_L22_reference_sizing(ctx_for, mgcd, wl):
    # With the reference's N_Q (92,160) the 22-bit run stops in the reduction
    # loop: at L = 22 each iteration shrinks the pair by about L/2 bits and
    # the size bookkeeping (getGcdKernel.cu:390-399) runs out after ~80.7k
    # iterations.  The trace path keeps that behaviour; the product path
    # above retries with more moduli.
    u, v = wl.single_large(0)
    st, g, d, p = ctx_for(22).gcd_trace(u, v)
    assert st == mgcd.ERR_REDUX and len(p) > 80000


def test_forced_moduli_failure_parity(ctx_for, oracle, mgcd, wl):
    ctx = ctx_for(32)
    u, v = wl.single_small(3)
    ctx.set_option(mgcd.OPT_FORCE_MODULI, 256)
    try:
        st, g, d, p = ctx.gcd_trace(u, v)
        ost, od, ot = oracle.gcd_kernel(u, v, ctx.moduli()[:256], 32)
        assert ost == oracle.REDUX_ERROR and st == mgcd.ERR_REDUX
        assert p == ot
        with pytest.raises(mgcd.MgcdError):  # a forced N_Q is never retried
            ctx.gcd(u, v)
    finally:
        ctx.set_option(mgcd.OPT_FORCE_MODULI, 0)
    assert ctx.gcd(u, v) == math.gcd(u, v)


def test_batch_mixed_small(ctx_for, mgcd, wl):
    ctx = ctx_for(32)
    pairs = wl.mixed(n=160, seed=9, lo=64, hi=30000)
    pairs += [(0, 0), (5, 0), (0, 7), (2**100, 2**100)]
    b = mgcd.Batch.from_pairs(ctx, pairs)
    b.run()
    res = b.results()
    for i, (u, v) in enumerate(pairs):
        assert res[i] == math.gcd(u, v), i
    # the same pairs alone through the one-pair path give the same results
    for i in (0, 7, 50, 100, 150):
        assert ctx.gcd(*pairs[i]) == res[i]
    b.close()


def test_batch_alone_vs_mixed(ctx_for, mgcd, wl):
    ctx = ctx_for(27)
    mix = wl.mixed(n=64, seed=21, lo=1000, hi=40000)
    probe = mix[10]
    alone = mgcd.Batch.from_pairs(ctx, [probe])
    alone.run()
    together = mgcd.Batch.from_pairs(ctx, mix)
    together.run()
    assert alone.result(0) == together.result(10) == math.gcd(*probe)


@pytest.mark.parametrize("T,K,B", [(256, 1, 1), (256, 4, 1), (512, 1, 1), (1024, 2, 1), (256, 1, 8), (256, 2, 8),
                                   (256, 1, 16), (256, 2, 16), (256, 1, 32), (128, 1, 16)])
def test_batch_tilings_agree(ctx_for, mgcd, wl, T, K, B):
    # Groups of B pairs with equal N_Q share one modular inverse per modulus;
    # 27 pairs leave a partial group; the 700-bit pairs form their own groups.
    ctx = ctx_for(32)
    pairs = wl.headline(24, seed=1, nbits=12000, gbits=1200) + wl.headline(3, seed=2, nbits=700, gbits=100)
    b = mgcd.Batch.from_pairs(ctx, pairs, workgroup=T, slots=K, group=B)
    assert b.tiling() == (T, K, B)
    b.run()
    for i, (u, v) in enumerate(pairs):
        assert b.result(i) == math.gcd(u, v)


@pytest.mark.parametrize("T, K, B", [(512, 2, 16), (256, 1, 16)])
def test_batch_sentinel_states_match_single(ctx_for, mgcd, oracle, wl, T, K, B):
    # Grouped pairs whose operands table moduli divide (t starting at
    # infinity, 0/0 or 0 on those moduli), beside coprime pairs that finish
    # at infinity early and stay in the group: each pair's iteration and
    # digit counts and its gcd equal those of the pair run alone.
    ctx = ctx_for(32)
    mods = ctx.moduli()
    rnd = random.Random(83)
    pairs = []
    for i, (u, v) in enumerate(wl.headline(12, seed=4, nbits=9000, gbits=2000)):
        p = [int(mods[j]) for j in rnd.sample(range(600), 5)]
        pairs.append((u * p[3] * p[4] * (rnd.getrandbits(160) | 1 << 159), v * p[0] * p[1] * p[2] * p[3]))
    for i in range(4):
        pairs.append((rnd.getrandbits(9000) | 1 << 8999 | 1, rnd.getrandbits(8990) | 1))
    b = mgcd.Batch.from_pairs(ctx, pairs, workgroup=T, slots=K, group=B)
    b.run()
    for i, (u, v) in enumerate(pairs):
        st, g, d, p = ctx.gcd_trace(max(u, v), min(u, v))
        nq, R, k = b.pair_stats(i)
        assert (R, k) == (len(p), len(d)) and b.result(i) == g == math.gcd(u, v), i
        # pivot by pivot: the grouped pair's recorded pivots are the oracle's
        ost, od, ot = oracle.gcd_kernel(max(u, v), min(u, v), mods[:nq], 32)
        assert ost == oracle.OK and b.pivots(i) == ot == p, i


def test_batch_group_traces_match_single(ctx_for, mgcd, oracle, wl):
    # Every pair of a grouped batch follows its own pivot sequence: the
    # iteration and digit counts equal those of the pair run alone.
    ctx = ctx_for(32)
    pairs = wl.headline(8, seed=3, nbits=9000, gbits=2000)
    b = mgcd.Batch.from_pairs(ctx, pairs, workgroup=256, slots=1, group=16)
    b.run()
    mods = ctx.moduli()
    for i, (u, v) in enumerate(pairs):
        st, g, d, p = ctx.gcd_trace(max(u, v), min(u, v))
        nq, R, k = b.pair_stats(i)
        assert (R, k) == (len(p), len(d)) and b.result(i) == g
        ost, od, ot = oracle.gcd_kernel(max(u, v), min(u, v), mods[:nq], 32)
        assert b.pivots(i) == ot == p, i


def test_one_pair_512_thread_tiling(ctx_for, mgcd, oracle, wl):
    # The 512-thread one-pair tiling: pivot and digit parity on a pair split
    # over two of its workgroups, and the mid sizes that select it by default.
    ctx = ctx_for(32)
    u, v = wl.single_small(4)
    ctx.set_option(mgcd.OPT_SINGLE_WORKGROUP, 512)
    ctx.set_option(mgcd.OPT_SINGLE_SLOTS, 1)
    try:
        st, g, d, p = ctx.gcd_trace(u, v)
        assert ctx.last_pair_stats()["tiling"] == (512, 1) and ctx.last_pair_stats()["parts"] == 2
    finally:
        ctx.set_option(mgcd.OPT_SINGLE_WORKGROUP, 0)
        ctx.set_option(mgcd.OPT_SINGLE_SLOTS, 0)
    nq = mgcd.num_moduli_needed(u.bit_length(), 32)
    ost, od, ot = oracle.gcd_kernel(u, v, ctx.moduli()[:nq], 32)
    assert st == mgcd.OK and ost == oracle.OK and p == ot and d == od and g == oracle.mpz_gcd(u, v)
    for kib in (64, 160, 288):
        x, y = wl.pair(wl.Stream(71, kib), kib * 1024, kib * 100)
        assert ctx.gcd(x, y) == oracle.mpz_gcd(x, y)
        assert ctx.last_pair_stats()["tiling"] == (512, 1), kib



def _structured_pairs(rnd):
    # Structured and strongly unbalanced pairs, u >= v: powers of two,
    # 2^a - 1 and 2^b - 1 (gcd 2^gcd(a, b) - 1), consecutive Fibonacci
    # numbers (Euclid's worst case, coprime), v | u, v of 1 to 1000 bits
    # beside a 20k-bit u (coprime and with a shared factor), u = v + 1, u = 3v.
    fa, fb = 0, 1
    for _ in range(28000):
        fa, fb = fb, fa + fb
    pairs = [(2**20000, 2**7000), (2**18000 - 1, 2**12000 - 1), (fb, fa)]
    v = rnd.getrandbits(5000) | 1 << 4999
    pairs.append((v * (rnd.getrandbits(12000) | 1 << 11999), v))
    u = rnd.getrandbits(20000) | 1 << 19999
    for vb in (1, 2, 31, 32, 33, 64, 1000):
        vv = rnd.getrandbits(vb) | 1 << (vb - 1)
        g = rnd.getrandbits(max(vb // 2, 2)) | 1
        pairs += [(u, vv), (u * g, vv * g)]
    vv = rnd.getrandbits(15000) | 1 << 14999
    pairs += [(vv + 1, vv), (3 * vv, vv)]
    return pairs


@pytest.mark.parametrize("L", [32, 22])
def test_structured_and_unbalanced_pairs(ctx_for, oracle, mgcd, L):
    # Pivot and digit parity with the oracle for every pair alone, then the
    # same pairs as one batch (default and group tilings) against math.gcd.
    ctx = ctx_for(L)
    pairs = _structured_pairs(random.Random(7))
    for u, v in pairs:
        _check_trace(ctx, oracle, mgcd, u, v, L)
        assert ctx.gcd(u, v) == ctx.gcd(v, u) == math.gcd(u, v)
    want = [math.gcd(u, v) for u, v in pairs]
    for tiling in ({}, dict(workgroup=512, slots=2, group=16)):
        b = mgcd.Batch.from_pairs(ctx, pairs + [(v, u) for u, v in pairs], **tiling)
        b.run()
        assert b.results() == want + want
        b.close()
