"""The transport kernels' range-restricted correctly rounded square root
and division (xo_device.h sqrt_in_range, div_in_range) give the bits of the
general correctly rounded operations: every float of the square root's
range, and 2^30 random operand pairs of the division's, checked on the
device through xo_probe_cr_math."""
import pytest

from pyxopto_amd import _native as N

pytestmark = pytest.mark.gpu

# +0 and every float with biased exponent 31 (2^-96) .. 222 (below 2^96)
SQRT_RANGE_COUNT = 1 + (222 - 31 + 1) * (1 << 23)


def test_range_restricted_sqrt_and_division_are_correctly_rounded():
    ctx = N.Context(0, 0, 0x5EED000000000001)
    checked, bad_sqrt, bad_div = ctx.probe_cr_math(ndiv=1 << 30, seed=7)
    ctx.close()
    assert checked == SQRT_RANGE_COUNT
    assert bad_sqrt == 0
    assert bad_div == 0
