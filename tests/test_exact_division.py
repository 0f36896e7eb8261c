"""The kernels replace x/12 and x/60 (flux operators, module_advect_em.F:199-213) by the
double product x*(1/c) rounded once to float.  Check it against IEEE float division on
every float mantissa in several binades (both signs) and on random floats."""
import numpy as np
import pytest


@pytest.mark.parametrize("c", [12.0, 60.0])
def test_double_product_equals_ieee_division(c):
    mant = np.arange(1 << 23, dtype=np.uint32)
    inv = np.float64(1.0) / np.float64(c)
    cf = np.float32(c)
    for exp in (1, 60, 126, 127, 140, 200, 253):
        bits = (np.uint32(exp) << np.uint32(23)) | mant
        x = bits.view(np.float32)
        for sx in (x, -x):
            ref = sx / cf
            got = (sx.astype(np.float64) * inv).astype(np.float32)
            assert np.array_equal(ref.view(np.uint32), got.view(np.uint32)), (c, exp)
    rng = np.random.default_rng(0)
    x = rng.standard_normal(1 << 22).astype(np.float32) * np.float32(1e4)
    assert np.array_equal((x / cf).view(np.uint32), (x.astype(np.float64) * inv).astype(np.float32).view(np.uint32))
