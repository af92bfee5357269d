""""All primes" moduli tables (CW section 2.4) on the GPU: every L-bit prime
instead of the Cavagnino-Werbrouck-filtered ones.  The FP64 arithmetic is
exact for any odd modulus, so the kernel's pivots and digits over such a
table still equal the oracle's restatement of the reference kernel run over
the same table (in its exact-reduction mode), and the 22-bit ceiling of 97,358
moduli lifts to 140,336."""
import math

import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx_all(mgcd):
    made = {}

    def get(L):
        if L not in made:
            made[L] = mgcd.Context(L, all_primes=True)
        return made[L]
    yield get
    for c in made.values():
        c.close()


@pytest.mark.parametrize("L", [22, 27, 32])
def test_trace_parity_over_all_primes(ctx_all, oracle, mgcd, wl, L):
    ctx = ctx_all(L)
    for seed in (0, 1):
        u, v = wl.single_small(seed)
        nq = mgcd.num_moduli_needed(max(u.bit_length(), 2), L)
        st, g, digits, pivots = ctx.gcd_trace(u, v)
        # the oracle's exact-reduction mode: the reference's division by
        # multiply does not apply to primes outside the CW condition
        ost, od, ot = oracle.gcd_kernel(u, v, ctx.moduli()[:nq], L, exact=True)
        assert st == mgcd.OK and ost == oracle.OK
        assert pivots == ot and digits == od
        assert g == oracle.mpz_gcd(u, v) == math.gcd(u, v)


def test_all_primes_table_lifts_the_L22_ceiling(ctx_all, mgcd, oracle, wl):
    # 1.3 Mbit needs N_Q = 117,504 22-bit moduli, more than the 97,358 usable
    # under the CW filter.
    u, v = wl.single_large(3, nbits=1300000, gbits=130000)
    assert mgcd.num_moduli_needed(u.bit_length(), 22) > 97358
    cw = mgcd.Context(22)
    try:
        assert cw.num_moduli == 97358
        with pytest.raises(mgcd.MgcdError) as e:
            cw.gcd(u, v)
        assert e.value.status == mgcd.ERR_NOT_ENOUGH_MODULI
    finally:
        cw.close()
    ctx = ctx_all(22)
    assert ctx.num_moduli == 140336
    assert ctx.gcd(u, v) == oracle.mpz_gcd(u, v)


def test_batch_over_all_primes(ctx_all, mgcd, wl):
    ctx = ctx_all(32)
    pairs = wl.mixed(n=64, seed=76, lo=1000, hi=100000)
    b = mgcd.Batch.from_pairs(ctx, pairs)
    b.run()
    assert b.results() == [math.gcd(u, v) for u, v in pairs]
    b.close()
