"""The texel unpacking's v / 255 (shading.h unorm8): the product with the
rounded reciprocal plus one FMA correction must equal the correctly rounded
IEEE quotient for every 8-bit v.  Checked here in exact rational arithmetic,
rounding each step to float32 (ties to even) as the GPU's v_mul_f32 and
v_fma_f32 do; and a near miss (dropping the correction) must fail."""
from fractions import Fraction as F

import numpy as np


def _rn32(t: F) -> np.float32:
    c = np.float32(float(t))
    cands = [np.nextafter(c, np.float32(-np.inf)), c, np.nextafter(c, np.float32(np.inf))]
    return min(cands, key=lambda x: (abs(F(float(x)) - t), int(np.float32(x).view(np.uint32)) & 1))


def _unorm8(v: int, corrected: bool = True) -> np.float32:
    x, r = np.float32(v), _rn32(F(1, 255))
    q = _rn32(F(float(x)) * F(float(r)))
    if not corrected:
        return q
    e = _rn32(F(float(x)) - F(float(q)) * 255)   # fma(-q, 255, x)
    return _rn32(F(float(e)) * F(float(r)) + F(float(q)))  # fma(e, r, q)


def test_unorm8_equals_ieee_quotient_for_every_byte():
    for v in range(256):
        assert _unorm8(v) == np.float32(v) / np.float32(255), v


def test_uncorrected_product_is_not_the_quotient():
    assert any(_unorm8(v, corrected=False) != np.float32(v) / np.float32(255) for v in range(256))
