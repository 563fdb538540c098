"""GPU: operands at any leading dimension and base alignment run the same
unchecked full-tile loads as aligned ones (16- and 4-byte loads that need only
the element's alignment), with the same bits.

The reference picks its load width from the leading dimension alone
(src/dmem_accessor.hpp:25, 36, 87, 97, 107) and falls back to 2-byte loads for
FP16 rows at an odd ld; here an odd ld or an odd base offset must give exactly
the result of the same values stored aligned, for every plan whose kernel
reads the caller's operands, and stay within the FP64 bar."""
from __future__ import annotations

import numpy as np
import pytest

import oracle
from tests.gpu_util import FP16, FP16_CLASSIC, TF32, dev_buffer, read_buffer, require_gpu, run_sgemm

pytestmark = pytest.mark.gpu


@pytest.fixture()
def h():
    require_gpu()
    import shgemm_amd as sg

    handle = sg.create()
    yield handle
    sg.destroy(handle)


def _store(M, ld, rows, cols):
    """Column-major rows x cols matrix M into a flat buffer with leading dimension ld."""
    out = np.zeros(ld * cols, dtype=M.dtype)
    for j in range(cols):
        out[j * ld:j * ld + rows] = M[:, j]
    return out


def _call(h, op_a, op_b, m, n, k, A, B, compute_type, lda, ldb, off):
    import shgemm_amd as sg

    ra, ca = (m, k) if op_a == 0 else (k, m)
    rb, cb = (k, n) if op_b == 0 else (n, k)
    a_t, a_p = dev_buffer(_store(A, lda, ra, ca), off[0])
    b_t, b_p = dev_buffer(_store(B, ldb, rb, cb), off[1])
    c_t, c_p = dev_buffer(np.zeros(m * n, np.float32))
    ct = 0 if compute_type == FP16_CLASSIC else compute_type
    if compute_type == FP16_CLASSIC:
        sg.set_exponent_scaling(h, False)
    try:
        sg.shgemm(h, op_a, op_b, m, n, k, 1.0, a_p, lda, b_p, ldb, 0.0, c_p, m, ct)
    finally:
        if compute_type == FP16_CLASSIC:
            sg.set_exponent_scaling(h, True)
    return read_buffer(c_t, 0, m * n)


# plans whose kernels read A and B in the caller's layout, with the slices
# forced for the k-sliced ones (the ring and three-stage plans need aligned
# operands and run as other plans otherwise; plan 8 repacks both operands)
PLANS = [(0, 1), (1, 1), (2, 4), (3, 4), (4, 4), (5, 1), (9, 1), (10, 2), (11, 4), (8, 1), (-1, 1)]


@pytest.mark.parametrize("shape", [(512, 512, 2112), (389, 333, 1999)])
@pytest.mark.parametrize("op_a", [0, 1])
@pytest.mark.parametrize("op_b", [0, 1])
@pytest.mark.parametrize("compute_type", [FP16, TF32, FP16_CLASSIC])
def test_unaligned_operands_same_bits(h, op_a, op_b, compute_type, shape):
    """(512, 512, 2112): whole tiles of every class, k a multiple of 64 so plan
    8 applies; (389, 333, 1999): edge tiles in m and n and a k-tail, so the
    unchecked interior loads, the bounds-checked edge loads and the tail run
    on the same odd-ld operands."""
    import shgemm_amd as sg

    m, n, k = shape
    rng = np.random.default_rng(40 + 4 * op_a + 2 * op_b + compute_type + 100 * (k % 2))
    ra, ca = (m, k) if op_a == 0 else (k, m)
    rb, cb = (k, n) if op_b == 0 else (n, k)
    A = rng.uniform(-1, 1, (ra, ca)).astype(np.float32)
    B = rng.uniform(-1, 1, (rb, cb)).astype(np.float16)
    C64 = oracle.gemm_f64(op_a, op_b, 0, m, n, k, 1.0, _store(A, ra, ra, ca), ra,
                          _store(B, rb, rb, cb).astype(np.float32), rb, 0.0, None, m)
    Cs = run_sgemm(h, op_a, op_b, 0, m, n, k, 1.0, _store(A, ra, ra, ca), ra,
                   _store(B, rb, rb, cb).astype(np.float32), rb, 0.0, np.zeros(m * n, np.float32), m)
    err_s = oracle.rel_fro_error(Cs, C64)
    for plan, slices in PLANS:
        if plan in (9, 10, 11) and compute_type != FP16:
            continue
        sg.force_plan(h, plan, slices)
        try:
            # aligned reference: leading dimensions a multiple of 8 elements, aligned bases
            ref = _call(h, op_a, op_b, m, n, k, A, B, compute_type, (ra + 15) // 8 * 8, (rb + 15) // 8 * 8, (0, 0))
            # odd leading dimensions and bases one element past 16-byte alignment
            got = _call(h, op_a, op_b, m, n, k, A, B, compute_type, ra + 1, rb + 1, (1, 1))
            got2 = _call(h, op_a, op_b, m, n, k, A, B, compute_type, ra + 3, rb + 5, (3, 3))
        finally:
            sg.force_plan(h, -1, 1)
        assert np.array_equal(ref.view(np.uint32), got.view(np.uint32)), (plan, "odd ld, base + 1")
        assert np.array_equal(ref.view(np.uint32), got2.view(np.uint32)), (plan, "ld + 3 / + 5, base + 3")
        err = oracle.rel_fro_error(got, C64)
        assert err <= 2.0 * err_s + 2.0 ** -22, (plan, err, err_s)
