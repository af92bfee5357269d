"""mgcd_gcd_limbs_noalloc (no device allocation, buffers in the library's fixed
scratch array) against mgcd_gcd_limbs and math.gcd, through the C ABI: two
calls in a row, the second with a smaller input, so the second runs over
what the first left in the shared array."""
import math

import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("L", [32, 22])
def test_noalloc_matches_gcd_limbs(ctx_for, wl, L):
    ctx = ctx_for(L)
    big = wl.pair(wl.Stream(41, L), 150000, 30000)        # N_Q 14,336 (L = 32), 16,128 (L = 22)
    small = wl.pair(wl.Stream(42, L), 8192, 1024, g_odd=True)
    for u, v in (big, small):
        want = ctx.gcd(u, v)
        got = ctx.gcd_noalloc(u, v)
        assert got == want == math.gcd(u, v)
        assert ctx.gcd_noalloc(v, u) == want  # either order


def test_noalloc_edges(ctx_for, wl, mgcd):
    ctx = ctx_for(32)
    u, v = wl.pair(wl.Stream(43, 0), 4096, 512)
    assert ctx.gcd_noalloc(u, 0) == u
    assert ctx.gcd_noalloc(u, u) == u
    assert ctx.gcd_noalloc(u, 1) == 1
    big = wl.pair(wl.Stream(44, 0), 300000, 1000)         # N_Q above the 16,384 limit
    with pytest.raises(mgcd.MgcdError) as e:
        ctx.gcd_noalloc(*big)
    assert e.value.status == mgcd.ERR_SCRATCH
    assert ctx.gcd_noalloc(u, v) == math.gcd(u, v)          # still usable afterwards
