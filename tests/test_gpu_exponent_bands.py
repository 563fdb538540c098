"""GPU: accuracy by exponent band of A (the reference's exponent study,
test/exponent_test.cu:202-211, 249-252): A = 2^(u*(hi-lo)+lo), u ~ U[0,1),
for the bands [-15, 14], [-36, -15], [-100, -36]; B = FP16(U[0,1)).

Bars, against an FP64 product, relative to the library SGEMM's error on the
same inputs:
  * fp16 with per-row prescaling (the default) and tf32 (BF16 split): within
    2x in every band;
  * fp16 with prescaling off (the reference's hi + 2^-11 lo split): within 2x
    in [-15, 14]; in [-100, -36] every element is below FP16's smallest
    subnormal even after the 2^11 residual scaling, so the product is lost
    (relative error near 1) -- the range limit of that compute type, as in
    the reference.  [-36, -15] has no bar for it: the large elements keep
    hi/lo precision in FP16 subnormals, the small ones flush.
"""
from __future__ import annotations

import numpy as np
import pytest

from tests.gpu_util import dev_buffer, read_buffer, require_gpu

pytestmark = pytest.mark.gpu
BANDS = [(-15, 14), (-36, -15), (-100, -36)]


def band_inputs(rng, m, n, k, lo, hi):
    u = rng.random(m * k)
    A = np.exp2(u * (hi - lo) + lo).astype(np.float32)          # column-major m x k
    B16 = rng.random(k * n).astype(np.float16)                   # column-major k x n
    return A, B16


def rel_err(C, A, B16, m, n, k):
    ref = A.reshape(k, m).T.astype(np.float64) @ B16.astype(np.float64).reshape(n, k).T
    got = C.reshape(n, m).T.astype(np.float64)
    return float(np.linalg.norm(got - ref) / np.linalg.norm(ref))


@pytest.fixture(scope="module")
def h():
    require_gpu()
    import shgemm_amd as sg

    handle = sg.create()
    yield handle
    sg.destroy(handle)


@pytest.mark.parametrize("band", BANDS)
def test_exponent_band_errors(h, band):
    import shgemm_amd as sg

    m = n = k = 384
    rng = np.random.default_rng(abs(band[0]) * 100 + abs(band[1]))
    A, B16 = band_inputs(rng, m, n, k, *band)
    a_t, a_p = dev_buffer(A)
    b_t, b_p = dev_buffer(B16)
    b32_t, b32_p = dev_buffer(B16.astype(np.float32))
    c_t, c_p = dev_buffer(np.zeros(m * n, np.float32))
    sg.sgemm_fp32(h, 0, 0, m, n, k, 1.0, a_p, m, b32_p, k, 0.0, c_p, m)
    e_s = rel_err(read_buffer(c_t, 0, m * n), A, B16, m, n, k)
    errs = {}
    for name, ct, scaling in (("fp16", 0, True), ("fp16 classic", 0, False), ("tf32", 1, True)):
        sg.set_exponent_scaling(h, scaling)
        sg.shgemm(h, 0, 0, m, n, k, 1.0, a_p, m, b_p, k, 0.0, c_p, m, ct)
        errs[name] = rel_err(read_buffer(c_t, 0, m * n), A, B16, m, n, k)
    sg.set_exponent_scaling(h, True)
    assert errs["fp16"] <= 2.0 * e_s + 2.0 ** -22, (band, errs, e_s)
    assert errs["tf32"] <= 2.0 * e_s + 2.0 ** -22, (band, errs, e_s)
    print(band, {k: v / e_s for k, v in errs.items()}, e_s)
    if band == (-15, 14):
        assert errs["fp16 classic"] <= 2.0 * e_s + 2.0 ** -22, (band, errs, e_s)
    elif band == (-100, -36):
        assert errs["fp16 classic"] > 0.5, (band, errs, e_s)
