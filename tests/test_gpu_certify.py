"""Exhaustive certification on the device of the FP32-reciprocal quotient
estimate q = rint(fl32(x) * rcp(fl32(y))) used by the modular inverse (the
role of certifyQuoRemQuasi.cu): with r = x - q*y, |r| < y whenever
x < 2^20 y and |r| < x otherwise, so the inverse's Euclid always makes
progress; plus the v_rcp_f32 error bound and sign symmetry behind it."""
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


def test_rcp_error_bound_and_symmetry(ctx_for):
    fail, checked = ctx_for(22).certify(2)
    print(f"rcp: {checked} floats checked, {fail} failures", flush=True)
    assert checked == 34 << 23  # every float in [1, 2^34)
    assert fail == 0
