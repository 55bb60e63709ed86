"""GPU: the reciprocal the FP32 pair terms use (hardware reciprocal and one
FMA Newton step) equals the IEEE quotient 1.0f / x, bit for bit, for every
binary32 x in [2^-126, 2^126): an exhaustive comparison on the device.
Outside that range the reciprocals differ, but the pair term computed with
the fast one is then NaN or equal to the division's pair term, for every
non-negative pattern (also exhaustive); the pair loops take the fast sum and
recompute with the division only when it is NaN.  The pair terms, and with
them every FP32-pair walker, are therefore the same as with the division
the oracle uses."""
import pytest

import sadmc_amd as S

pytestmark = pytest.mark.gpu


def test_fast_reciprocal_equals_division_on_its_range():
    print("case reciprocal [2^-126, 2^126)", flush=True)
    assert S.check_recip_f32(0x00800000, 0x7E800000) == 0


def test_fast_reciprocal_differs_outside_its_range():
    print("case reciprocal outside its range", flush=True)
    assert S.check_recip_f32(0x00000001, 0x00800000) > 0  # denormal x
    assert S.check_recip_f32(0x7E800000, 0x7F800000) > 0  # 1/x denormal
    assert S.check_recip_f32(0x7F800000, 0x7F800001) == 1  # +inf (the Newton step makes NaN)
    assert S.check_recip_f32(0x7F800001, 0x80000000) == 0  # NaN stays NaN


def test_fast_pair_term_is_exact_or_nan_for_every_r2():
    print("case pair term over all non-negative r2", flush=True)
    assert S.check_recip_f32(0x00000000, 0x80000000, what=1) == 0
