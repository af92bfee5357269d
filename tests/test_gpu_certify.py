"""Exhaustive certification of the FP32-reciprocal quasi-quotient on the
device (the role of certifyQuoRemQuasi.cu): the quotient estimate used by
the modular inverse never exceeds floor(x/y) and falls short by at most one
whenever floor(x/y) < 2^19."""
import pytest

pytestmark = pytest.mark.gpu


def test_float_domain_point_by_point(ctx_for):
    # every x in [1, 2^23), y in [1, 2^22): 2^45 - ... evaluations
    fail, checked = ctx_for(22).certify(0)
    print(f"float domain: {checked} (x, y) checked, {fail} failures", flush=True)
    assert checked == ((1 << 22) - 1) * ((1 << 23) - 1)
    assert fail == 0


@pytest.mark.parametrize("L", [22, 27, 32])
def test_width_endpoints(ctx_for, L):
    fail, checked = ctx_for(22).certify(1, L)
    print(f"L={L}: {checked} endpoints checked, {fail} failures", flush=True)
    assert checked > (1 << L)
    assert fail == 0
