"""GPU parity of the SHGEMM / HSGEMM path against the FP64 oracle.

Bars (written into each test):
  * integer-valued inputs: bit-exact (every partial sum is an exact FP32
    integer and the split is exact);
  * real-valued inputs: relative Frobenius error vs FP64 within 2x the error
    of the library's plain FP32 SGEMM on the same inputs (plus 2^-22 slack on
    tiny shapes, where SGEMM's own error is a handful of roundings);
  * agreement with the FP64 evaluation of the split representation the kernel
    multiplies (emulate_shgemm) to FP32 accumulation accuracy.
"""
from __future__ import annotations

from pathlib import Path

import numpy as np
import pytest

import oracle
from tests.golden.make_golden import load_cases
from tests.gpu_util import FP16, FP16_CLASSIC, TF32, VARIANTS, lds, require_gpu, run_hsgemm, run_sgemm, run_shgemm, sizes

pytestmark = pytest.mark.gpu

GOLDEN = Path(__file__).resolve().parent / "golden"


@pytest.fixture(scope="module")
def h():
    require_gpu()
    import shgemm_amd as sg

    handle = sg.create()
    yield handle
    sg.destroy(handle)


def _err(C_flat, C64_flat, op_c, m, n, ldc):
    return oracle.rel_fro_error(oracle.c_matrix(C_flat, op_c, m, n, ldc), oracle.c_matrix(C64_flat, op_c, m, n, ldc))


def _sgemm_err(h, c):
    B32 = c["B_bits"].view(np.float16).astype(np.float32)
    Cs = run_sgemm(h, c["op_a"], c["op_b"], c["op_c"], c["m"], c["n"], c["k"], c["alpha"], c["A"], c["lda"], B32,
                   c["ldb"], c["beta"], c["C0"], c["ldc"])
    return _err(Cs, c["C64"], c["op_c"], c["m"], c["n"], c["ldc"])


@pytest.mark.parametrize("compute_type", VARIANTS)
def test_golden_integer_bit_exact(h, compute_type):
    for c in load_cases(GOLDEN / "gemm_int.npz"):
        C, _ = run_shgemm(h, c["op_a"], c["op_b"], c["op_c"], c["m"], c["n"], c["k"], c["alpha"], c["A"], c["lda"],
                          c["B_bits"], c["ldb"], c["beta"], c["C0"], c["ldc"], compute_type)
        view = oracle.c_matrix(C, c["op_c"], c["m"], c["n"], c["ldc"])
        ref = oracle.c_matrix(c["C64"], c["op_c"], c["m"], c["n"], c["ldc"])
        assert np.array_equal(view.astype(np.float64), ref), (c["op_a"], c["op_b"], c["op_c"], c["m"], c["n"], c["k"])


@pytest.mark.parametrize("compute_type", VARIANTS)
def test_golden_random_vs_fp64(h, compute_type):
    for c in load_cases(GOLDEN / "gemm_rand.npz"):
        C, _ = run_shgemm(h, c["op_a"], c["op_b"], c["op_c"], c["m"], c["n"], c["k"], c["alpha"], c["A"], c["lda"],
                          c["B_bits"], c["ldb"], c["beta"], c["C0"], c["ldc"], compute_type)
        err = _err(C, c["C64"], c["op_c"], c["m"], c["n"], c["ldc"])
        err_s = _sgemm_err(h, c)
        tag = (c["op_a"], c["op_b"], c["op_c"], c["m"], c["n"], c["k"], compute_type)
        assert err <= 2.0 * err_s + 2.0 ** -22, (tag, err, err_s)
        # padding rows of C (ld > rows) must be untouched
        rows = c["m"] if c["op_c"] == 0 else c["n"]
        pad = np.array(C).reshape(-1, c["ldc"])[:, rows:]
        pad0 = np.array(c["C0"]).reshape(-1, c["ldc"])[:, rows:]
        assert np.array_equal(pad, pad0), tag


def _random_case(rng, op_a, op_b, op_c, m, n, k, pad, alpha=1.0, beta=0.0):
    lda, ldb, ldc = lds(op_a, op_b, op_c, m, n, k, pad)
    na, nb, nc = sizes(op_a, op_b, op_c, m, n, k, lda, ldb, ldc)
    A = oracle.gen_uniform(rng, na)
    B16 = oracle.gen_f16_uniform(rng, nb)
    C0 = oracle.gen_uniform(rng, nc)
    return dict(op_a=op_a, op_b=op_b, op_c=op_c, m=m, n=n, k=k, lda=lda, ldb=ldb, ldc=ldc, alpha=alpha, beta=beta,
                A=A, B16=B16, B32=B16.astype(np.float32), C0=C0)


def _fp64(c):
    return oracle.gemm_f64(c["op_a"], c["op_b"], c["op_c"], c["m"], c["n"], c["k"], c["alpha"], c["A"], c["lda"],
                           c["B32"], c["ldb"], c["beta"], c["C0"], c["ldc"])


def _run(h, c, compute_type, off=(0, 0, 0)):
    return run_shgemm(h, c["op_a"], c["op_b"], c["op_c"], c["m"], c["n"], c["k"], c["alpha"], c["A"], c["lda"],
                      c["B16"], c["ldb"], c["beta"], c["C0"], c["ldc"], compute_type, off)


def _run_s(h, c):
    return run_sgemm(h, c["op_a"], c["op_b"], c["op_c"], c["m"], c["n"], c["k"], c["alpha"], c["A"], c["lda"],
                     c["B32"], c["ldb"], c["beta"], c["C0"], c["ldc"])


@pytest.mark.parametrize("compute_type", VARIANTS)
def test_cfg1_1024_nn_within_2x_sgemm(h, compute_type):
    """1024^3 op_n/op_n, A U[-1,1] FP32, B U[-1,1] FP16, alpha=1, beta=0."""
    rng = np.random.default_rng(1)
    c = _random_case(rng, 0, 0, 0, 1024, 1024, 1024, 0)
    C64 = _fp64(c)
    C, _ = _run(h, c, compute_type)
    Cs = _run_s(h, c)
    err = _err(C, C64, 0, 1024, 1024, 1024)
    err_s = _err(Cs, C64, 0, 1024, 1024, 1024)
    print(f"cfg1 compute_type={compute_type} err={err:.3e} sgemm_err={err_s:.3e} ratio={err / err_s:.3f}")
    assert err_s < 1e-6
    assert err <= 2.0 * err_s, (err, err_s)
    # the kernel multiplies exactly the split representation: FP32-accumulation agreement
    emu = oracle.emulate_shgemm(0, 0, 0, 1024, 1024, 1024, 1.0, c["A"], 1024, c["B32"], 1024, 0.0, None, 1024,
                                {FP16: 2, TF32: 1, FP16_CLASSIC: 0}[compute_type])
    assert _err(C, emu, 0, 1024, 1024, 1024) <= 2.0 * err_s


@pytest.mark.parametrize("compute_type", VARIANTS)
@pytest.mark.parametrize("op_a", [0, 1])
@pytest.mark.parametrize("op_b", [0, 1])
@pytest.mark.parametrize("op_c", [0, 1])
def test_all_ops_odd_shapes(h, op_a, op_b, op_c, compute_type):
    rng = np.random.default_rng(100 + 8 * op_a + 4 * op_b + 2 * op_c + compute_type)
    for (m, n, k, pad, alpha, beta, off) in [(131, 77, 301, 3, 1.0, 0.0, (0, 0, 0)),
                                             (257, 190, 515, 0, -1.5, 0.25, (1, 1, 1)),
                                             (64, 64, 64, 8, 1.0, 1.0, (4, 8, 4))]:
        c = _random_case(rng, op_a, op_b, op_c, m, n, k, pad, alpha, beta)
        C64 = _fp64(c)
        C, _ = _run(h, c, compute_type, off)
        Cs = _run_s(h, c)
        err = _err(C, C64, op_c, m, n, c["ldc"])
        err_s = _err(Cs, C64, op_c, m, n, c["ldc"])
        assert err <= 2.0 * err_s + 2.0 ** -22, ((m, n, k, off), err, err_s)


@pytest.mark.parametrize("level", [0, 1])
def test_fixed_levels(h, level):
    import shgemm_amd as sg

    rng = np.random.default_rng(5)
    c = _random_case(rng, 0, 0, 0, 300, 260, 400, 0)
    C64 = _fp64(c)
    sg.enable_kernel_level_fixing(h, level)
    try:
        C, lvl = _run(h, c, FP16)
    finally:
        sg.disable_kernel_level_fixing(h)
    assert int(lvl) == level
    assert _err(C, C64, 0, 300, 260, 300) <= 2.0 * _err(_run_s(h, c), C64, 0, 300, 260, 300) + 2.0 ** -22


def test_alpha_power_of_two_is_exact_scaling(h):
    rng = np.random.default_rng(9)
    c = _random_case(rng, 1, 0, 0, 200, 150, 333, 1)
    C1, _ = _run(h, c, FP16)
    C4, _ = _run(h, dict(c, alpha=4.0), FP16)
    v1 = oracle.c_matrix(C1, 0, 200, 150, c["ldc"])
    v4 = oracle.c_matrix(C4, 0, 200, 150, c["ldc"])
    assert np.array_equal(v4, v1 * np.float32(4.0))


def test_deterministic(h):
    rng = np.random.default_rng(3)
    c = _random_case(rng, 0, 0, 0, 513, 511, 777, 0, 1.0, 0.5)
    first, _ = _run(h, c, FP16)
    for _ in range(3):
        again, _ = _run(h, c, FP16)
        assert np.array_equal(first, again)


def test_k_zero_and_empty(h):
    rng = np.random.default_rng(4)
    c = _random_case(rng, 0, 0, 0, 40, 30, 0, 2, 1.0, 0.5)
    c["A"] = np.zeros(1, np.float32)
    c["B16"] = np.zeros(1, np.float16)
    c["lda"], c["ldb"] = 42, 1
    C, _ = _run(h, c, FP16)
    ref = oracle.c_matrix(c["C0"], 0, 40, 30, c["ldc"]) * np.float32(0.5)
    assert np.array_equal(oracle.c_matrix(C, 0, 40, 30, c["ldc"]), ref)
    # m == 0: nothing is written
    c2 = _random_case(rng, 0, 0, 0, 0, 30, 20, 1)
    C2, _ = _run(h, c2, FP16)
    assert np.array_equal(C2, c2["C0"])


@pytest.mark.parametrize("compute_type", [FP16, TF32])
@pytest.mark.parametrize("op_a", [0, 1])
@pytest.mark.parametrize("op_b", [0, 1])
def test_hsgemm_vs_fp64(h, op_a, op_b, compute_type):
    """hsgemm: C = op(A16) op(B32); computed by the library as C^T = op'(B) op'(A)."""
    rng = np.random.default_rng(50 + op_a * 2 + op_b)
    m, n, k = 190, 96, 260
    for op_c in (0, 1):
        lda = (m if op_a == 0 else k) + 1
        ldb = (k if op_b == 0 else n) + 2
        ldc = (m if op_c == 0 else n) + 3
        A16 = oracle.gen_f16_uniform(rng, lda * (k if op_a == 0 else m))
        B = oracle.gen_uniform(rng, ldb * (n if op_b == 0 else k))
        C0 = oracle.gen_uniform(rng, ldc * (n if op_c == 0 else m))
        C = run_hsgemm(h, op_a, op_b, op_c, m, n, k, 1.0, A16, lda, B, ldb, 0.0, C0, ldc, compute_type)
        Am = oracle.op_view(A16.astype(np.float64), op_a, m, k, lda)
        Bm = oracle.op_view(B.astype(np.float64), op_b, k, n, ldb)
        ref = Am @ Bm
        got = oracle.c_matrix(C, op_c, m, n, ldc)
        # FP32 reference of the same product for the 2x bar
        s = (Am.astype(np.float32) @ Bm.astype(np.float32)).astype(np.float64)
        err, err_s = oracle.rel_fro_error(got, ref), oracle.rel_fro_error(s, ref)
        assert err <= 2.0 * err_s + 2.0 ** -22, (op_c, err, err_s)


def test_invalid_arguments_raise(h):
    import shgemm_amd as sg

    rng = np.random.default_rng(6)
    c = _random_case(rng, 0, 0, 0, 16, 16, 16, 0)
    with pytest.raises(sg.ShgemmError) as ei:
        run_shgemm(h, 0, 0, 0, 16, 16, 16, 1.0, c["A"], 8, c["B16"], 16, 0.0, c["C0"], 16, FP16)
    assert ei.value.status == sg.STATUS_EINVAL and "lda" in str(ei.value)
    with pytest.raises(sg.ShgemmError):
        run_shgemm(h, 2, 0, 0, 16, 16, 16, 1.0, c["A"], 16, c["B16"], 16, 0.0, c["C0"], 16, FP16)
    with pytest.raises(sg.ShgemmError):
        run_shgemm(h, 0, 0, 0, 16, 16, 16, 1.0, c["A"], 16, c["B16"], 16, 0.0, c["C0"], 16, 7)


def test_stream_is_honoured(h):
    import torch

    import shgemm_amd as sg
    from tests.gpu_util import dev_buffer

    rng = np.random.default_rng(8)
    c = _random_case(rng, 0, 1, 0, 256, 256, 512, 0)
    C64 = _fp64(c)
    s = torch.cuda.Stream()
    sg.set_cuda_stream(h, s)
    try:
        a_t, a_p = dev_buffer(c["A"])
        b_t, b_p = dev_buffer(c["B16"])
        c_t, c_p = dev_buffer(c["C0"])
        torch.cuda.synchronize()
        sg.shgemm(h, 0, 1, 256, 256, 512, 1.0, a_p, c["lda"], b_p, c["ldb"], 0.0, c_p, c["ldc"], FP16, 0)
        s.synchronize()
        C = c_t[: c["C0"].size].cpu().numpy()
    finally:
        sg.set_cuda_stream(h, None)
    assert _err(C, C64, 0, 256, 256, 256) < 1e-6
