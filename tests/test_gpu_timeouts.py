"""The timeout path, forced from the host: MGCD_OPT_SPIN_LIMIT bounds every
in-kernel wait (cross-workgroup exchange, residue jobs) to one poll, so pairs
split over several workgroups end with MGCD_ERR_TIMEOUT instead of a result,
and every kernel still exits normally.  The same context and batch then run
correctly once the limit is restored (the exchange slots are cleared per
launch)."""
import math

import pytest

pytestmark = pytest.mark.gpu


def test_one_pair_timeout_then_recovers(mgcd, wl):
    ctx = mgcd.Context(32)
    try:
        u, v = wl.headline(1, seed=9)[0]  # 100k bits: 39 workgroups in the one-pair path
        ctx.set_option(mgcd.OPT_SPIN_LIMIT, 1)
        with pytest.raises(mgcd.MgcdError) as e:
            ctx.gcd(u, v)
        assert e.value.status == mgcd.ERR_TIMEOUT
        ctx.set_option(mgcd.OPT_SPIN_LIMIT, 1 << 24)
        assert ctx.gcd(u, v) == math.gcd(u, v)
    finally:
        ctx.close()


def test_batch_timeouts_then_recovers(mgcd, wl):
    ctx = mgcd.Context(32)
    try:
        pairs = wl.headline(24, seed=10, nbits=30000, gbits=3000)  # multi-part groups
        want = [math.gcd(u, v) for u, v in pairs]
        b = mgcd.Batch.from_pairs(ctx, pairs)
        ctx.set_option(mgcd.OPT_SPIN_LIMIT, 1)
        b.run()
        st = [b.status(i) for i in range(len(pairs))]
        assert all(x in (mgcd.OK, mgcd.ERR_TIMEOUT) for x in st)
        assert st.count(mgcd.ERR_TIMEOUT) > 0
        for i, x in enumerate(st):
            if x == mgcd.OK:
                assert b.result(i) == want[i]
        ctx.set_option(mgcd.OPT_SPIN_LIMIT, 1 << 24)
        b.run()
        assert b.results() == want
        b.close()
    finally:
        ctx.close()


def test_spin_limit_argument_checked(mgcd):
    ctx = mgcd.Context(32)
    try:
        with pytest.raises(mgcd.MgcdError):
            ctx.set_option(mgcd.OPT_SPIN_LIMIT, 0)
    finally:
        ctx.close()
