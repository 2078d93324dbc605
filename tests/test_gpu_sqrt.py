"""The SDF sphere distances' square root (sqrt_rn in rayna_amd/csrc/kernels.hip: sqrtf's sequence without its final
pass-through of zeros and infinities; sqrt_rn_wave, which scales small inputs only in waves that hold one) equals
the correctly rounded sqrtf on every float: all are evaluated on the GPU for all 2^32 bit patterns and compared bit
for bit (NaNs compare equal).  The smooth-min and sphere SDFs of the march use sqrt_rn_wave, and their bit-exact
agreement with the oracle rests on this."""
import pytest

from rayna_amd import RayError
from rayna_amd.renderer import debug_sqrt_mismatches

pytestmark = pytest.mark.gpu


def test_sqrt_rn_equals_sqrtf_on_every_bit_pattern(gpu):
    print("case sqrt_rn over 2^32 patterns", flush=True)
    assert debug_sqrt_mismatches(0, 1 << 32) == 0


def test_sqrt_check_range_is_bounded(gpu):
    assert debug_sqrt_mismatches(0x3F800000, 16) == 0  # [1, 1 + 15 ulp)
    with pytest.raises(RayError):
        debug_sqrt_mismatches(1, 1 << 32)  # past 2^32
