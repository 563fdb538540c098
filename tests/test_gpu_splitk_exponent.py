"""GPU: deterministic split-K (tall-skinny) and the wide-exponent sketch inputs.

Split-K bar: bit-identical results over 10 runs, and FP64 error within 2x of
the library's FP32 SGEMM on the same inputs.
Wide exponent (A = sign * U[1,2) * 2^e, e in [-15, 15]): the fp16 compute
type overflows FP16 without prescaling (as the reference's fp16 mode does);
with per-row prescaling, and in the tf32 (BF16 split) compute type, the error
is within 2x of SGEMM's.
"""
from __future__ import annotations

import numpy as np
import pytest

import oracle
from tests.gpu_util import dev_buffer, read_buffer, require_gpu, run_sgemm

pytestmark = pytest.mark.gpu
FP16, TF32 = 0, 1


@pytest.fixture(scope="module")
def h():
    require_gpu()
    import shgemm_amd as sg

    handle = sg.create()
    yield handle
    sg.destroy(handle)


def _fp64_nn(A, B32, m, n, k):
    Am = A.reshape(k, m).T.astype(np.float64)
    Bm = B32.reshape(n, k).T.astype(np.float64)
    return Am @ Bm


@pytest.mark.parametrize("scaling", [True, False])
@pytest.mark.parametrize("k", [1 << 16, 1 << 20])
def test_splitk_tall_skinny_bit_identical(h, k, scaling):
    """m = n = 128, k up to 2^20 (A U[-1,1], B N(0,1) FP16), op_n/op_n."""
    import torch

    import shgemm_amd as sg

    m = n = 128
    rng = np.random.default_rng(k)
    A = oracle.gen_uniform(rng, m * k)
    B16 = oracle.gen_f16_normal(rng, k * n)
    B32 = B16.astype(np.float32)
    a_t, a_p = dev_buffer(A)
    b_t, b_p = dev_buffer(B16)
    c_t, c_p = dev_buffer(np.zeros(m * n, np.float32))
    runs = []
    sg.set_exponent_scaling(h, scaling)
    try:
        for _ in range(10):
            c_t.fill_(float("nan"))
            sg.shgemm(h, 0, 0, m, n, k, 1.0, a_p, m, b_p, k, 0.0, c_p, m, FP16)
            runs.append(read_buffer(c_t, 0, m * n).copy())
    finally:
        sg.set_exponent_scaling(h, True)
    for r in runs[1:]:
        assert np.array_equal(r.view(np.uint32), runs[0].view(np.uint32))
    ref = _fp64_nn(A, B32, m, n, k)
    got = runs[0].reshape(n, m).T
    Cs = run_sgemm(h, 0, 0, 0, m, n, k, 1.0, A, m, B32, k, 0.0, np.zeros(m * n, np.float32), m).reshape(n, m).T
    err, err_s = oracle.rel_fro_error(got, ref), oracle.rel_fro_error(Cs, ref)
    print(f"split-K k={k}: err={err:.3e} sgemm_err={err_s:.3e}")
    assert err <= 2.0 * err_s
    del a_t, b_t, c_t
    torch.cuda.empty_cache()


def test_splitk_beta_and_transposed_c(h):
    """Split-K with beta != 0 and C stored transposed; alpha scaling."""
    import shgemm_amd as sg

    m, n, k = 96, 160, 40000
    rng = np.random.default_rng(77)
    A = oracle.gen_uniform(rng, k * m)          # op_t: stored k x m
    B16 = oracle.gen_f16_uniform(rng, n * k)    # op_t: stored n x k
    C0 = oracle.gen_uniform(rng, n * m)         # op_c = t: stored n x m
    a_t, a_p = dev_buffer(A)
    b_t, b_p = dev_buffer(B16)
    c_t, c_p = dev_buffer(C0)
    sg.shgemm(h, 1, 1, m, n, k, -2.0, a_p, k, b_p, n, 0.5, c_p, n, FP16, 1)
    got = read_buffer(c_t, 0, C0.size).reshape(m, n)
    Am = A.reshape(m, k).astype(np.float64)
    Bm = B16.astype(np.float64).reshape(k, n)
    ref = -2.0 * (Am @ Bm) + 0.5 * C0.reshape(m, n).astype(np.float64)
    s = -2.0 * (Am.astype(np.float32) @ Bm.astype(np.float32)) + np.float32(0.5) * C0.reshape(m, n)
    assert oracle.rel_fro_error(got, ref) <= 2.0 * oracle.rel_fro_error(s, ref) + 2.0 ** -22


def _wide_case(seed, m, n, k):
    rng = np.random.default_rng(seed)
    A = oracle.gen_wide_exponent(rng, m * k)
    B16 = oracle.gen_f16_normal(rng, k * n, sigma=1.0 / np.sqrt(n))
    return A, B16, B16.astype(np.float32)


def test_wide_exponent_fp16_needs_prescale(h):
    import shgemm_amd as sg

    m, n, k = 2048, 256, 2048
    A, B16, B32 = _wide_case(1, m, n, k)
    assert (np.abs(A) >= 65520).any()
    a_t, a_p = dev_buffer(A)
    b_t, b_p = dev_buffer(B16)
    c_t, c_p = dev_buffer(np.zeros(m * n, np.float32))
    ref = _fp64_nn(A, B32, m, n, k)
    Cs = run_sgemm(h, 0, 0, 0, m, n, k, 1.0, A, m, B32, k, 0.0, np.zeros(m * n, np.float32), m).reshape(n, m).T
    err_s = oracle.rel_fro_error(Cs, ref)
    # fp16 without prescaling: FP16 overflow of hi, as in the reference's fp16 mode
    sg.set_exponent_scaling(h, False)
    try:
        sg.shgemm(h, 0, 0, m, n, k, 1.0, a_p, m, b_p, k, 0.0, c_p, m, FP16)
        raw = read_buffer(c_t, 0, m * n)
    finally:
        sg.set_exponent_scaling(h, True)
    assert not np.isfinite(raw).all()
    # per-row prescaling (the default)
    sg.shgemm(h, 0, 0, m, n, k, 1.0, a_p, m, b_p, k, 0.0, c_p, m, FP16)
    got = read_buffer(c_t, 0, m * n).reshape(n, m).T
    err = oracle.rel_fro_error(got, ref)
    print(f"wide exponent fp16+prescale err={err:.3e} sgemm_err={err_s:.3e}")
    assert err <= 2.0 * err_s


@pytest.mark.parametrize("op_a", [0, 1])
def test_wide_exponent_tf32_bf16_split(h, op_a):
    import shgemm_amd as sg

    m, n, k = 2048, 256, 2048
    A, B16, B32 = _wide_case(2 + op_a, m, n, k)   # op_t: the same buffer read as k x m
    a_t, a_p = dev_buffer(A)
    b_t, b_p = dev_buffer(B16)
    c_t, c_p = dev_buffer(np.zeros(m * n, np.float32))
    lda = m if op_a == 0 else k
    sg.shgemm(h, op_a, 0, m, n, k, 1.0, a_p, lda, b_p, k, 0.0, c_p, m, TF32)
    got = read_buffer(c_t, 0, m * n).reshape(n, m).T
    Am = oracle.op_view(A.astype(np.float64), op_a, m, k, lda)
    ref = Am @ B32.reshape(n, k).T.astype(np.float64)
    Cs = run_sgemm(h, op_a, 0, 0, m, n, k, 1.0, A, lda, B32, k, 0.0, np.zeros(m * n, np.float32), m).reshape(n, m).T
    err, err_s = oracle.rel_fro_error(got, ref), oracle.rel_fro_error(Cs, ref)
    print(f"wide exponent tf32 op_a={op_a} err={err:.3e} sgemm_err={err_s:.3e}")
    assert err <= 2.0 * err_s


def test_wide_exponent_hsgemm_transposed_roles(h):
    """Sketch with transposed roles: Y^T (256 x m) = Omega^T (FP16) * A^T (FP32), via hsgemm."""
    import shgemm_amd as sg

    m, n, k = 2048, 256, 2048       # A is m x k, Omega is k x n
    A, B16, B32 = _wide_case(5, m, n, k)
    a_t, a_p = dev_buffer(A)        # A stored m x k column-major; A^T is op_t of it
    b_t, b_p = dev_buffer(B16)      # Omega stored k x n; Omega^T is op_t of it
    c_t, c_p = dev_buffer(np.zeros(n * m, np.float32))
    # hsgemm(op_a=t on Omega, op_b=t on A): C (n x m) = Omega^T A^T
    sg.hsgemm(h, 1, 1, n, m, k, 1.0, b_p, k, a_p, m, 0.0, c_p, n, FP16)
    got = read_buffer(c_t, 0, n * m).reshape(m, n).T      # n x m col-major
    ref = (A.reshape(k, m).T.astype(np.float64) @ B32.reshape(n, k).T.astype(np.float64)).T
    s = (A.reshape(k, m).T @ B32.reshape(n, k).T).T.astype(np.float64)
    err, err_s = oracle.rel_fro_error(got, ref), oracle.rel_fro_error(s, ref)
    print(f"hsgemm transposed roles err={err:.3e} fp32_err={err_s:.3e}")
    assert err <= 2.0 * err_s
