"""The VALU issue-rate calibration behind the bench's VALU roofline (rayna_measure_valu_peak): a rate, in wave64
instructions per second, that is physically plausible for 256 CUs and that a longer launch reproduces."""
import pytest

from rayna_amd import RayError
from rayna_amd.renderer import measure_valu_peak

pytestmark = pytest.mark.gpu


def test_valu_peak_is_plausible_and_reproducible(gpu):
    print("case valu peak", flush=True)
    r1, s1 = measure_valu_peak(1024)
    r2, s2 = measure_valu_peak(4096)
    # 256 CUs at up to 2.4 GHz retire at most a few wave64 VALU instructions per CU per cycle
    for r in (r1, r2):
        assert 0.1 * 256 * 2.4e9 < r < 4 * 256 * 2.4e9, r
    assert s2 > 3 * s1
    assert abs(r2 / r1 - 1) < 0.25, (r1, r2)


def test_valu_peak_rejects_bad_arguments(gpu):
    with pytest.raises(RayError):
        measure_valu_peak(0)
