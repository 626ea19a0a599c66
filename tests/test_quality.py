"""Quality figures (common/type.h:4960-5030 CheckImageQuality / NCC) against
their closed forms."""
import math

import numpy as np
import pytest

from fdk_amd.quality import check_image_quality, ncc


def test_identical_volumes():
    v = np.random.default_rng(0).standard_normal((4, 5, 6))
    q = check_image_quality(v, v)
    assert q["error_per_pixel"] == 0 and q["mse"] == 0 and q["psnr_db"] == math.inf
    assert q["qi"] == 1.0 and q["passed"]


@pytest.mark.parametrize("delta", [1e-3, 0.05, 0.5])
def test_constant_offset_closed_form(delta):
    v = np.random.default_rng(1).standard_normal((3, 4, 5))
    q = check_image_quality(v, v + delta)
    assert q["error_per_pixel"] == pytest.approx(delta, rel=1e-9)
    assert q["mse"] == pytest.approx(delta * delta, rel=1e-9)
    assert q["psnr_db"] == pytest.approx(20 * math.log10(2.0) - 10 * math.log10(delta * delta),
                                         rel=1e-9)
    assert q["qi"] == pytest.approx((2 - delta) / 2, rel=1e-12)
    # thresholds: error per pixel <= 0.03 and PSNR >= 26 dB
    assert q["passed"] == (delta <= 0.03 and q["psnr_db"] >= 26.0)


def test_ncc():
    rng = np.random.default_rng(2)
    a = rng.standard_normal(1000)
    assert ncc(a, 3.0 * a + 7.0) == pytest.approx(1.0, abs=1e-12)
    assert ncc(a, -a) == pytest.approx(-1.0, abs=1e-12)
    b = rng.standard_normal(1000)
    ref = np.corrcoef(a, b)[0, 1]
    assert ncc(a, b) == pytest.approx(ref, abs=1e-12)
    with pytest.raises(ValueError):
        ncc(a, b[:10])
