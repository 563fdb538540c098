"""GPU: the arithmetic premises the kernels rest on, checked on the device.

  * MFMA lane layout: v_mfma_f32_32x32x16_f16 / _bf16 and
    v_mfma_f32_32x32x2_f32 with the operand and accumulator layout the
    kernels assume give the exact product on integer-valued operands.
  * FP16 MFMA accumulation: faithful for one product onto the accumulator,
    but for sums of 16 products not one round-to-nearest-even: within 8 ulps
    of the largest operand and leaning toward zero on same-sign operands, so
    errors add up along a long MFMA chain.  The long-chain tests measure what
    that does to a whole GEMM: a single chain stays within 2x of the
    library's FP32 SGEMM (whose error equals the platform FP32 GEMM's), and
    k-slices summed in FP32 outside the chain (split-K with its fixed-order
    reduction; the reference likewise sums its hi product outside the MMA
    chain, src/wmmae_shgemm.hpp:86-98) bring it to the error of a
    round-to-nearest FP32 GEMM.
  * The paired split of the prescaled fp16 mode (split2_f16_scaled,
    v_fma_mix forms) equals the oracle's split_f16_same_scale of x * s bit
    for bit, over 2^22 values: wide exponents, FP16's top range, rounding
    ties and FP16 subnormals, at power-of-two scales 2^-10 .. 2^9.
  * The reference's remedy on gfx950: the hi product of every MFMA added
    outside the chain has 1.3-2x less error than two MFMA chains; the
    library's k-sliced plans for long k match or beat it.
  * The device's streaming HBM read rate (the attainable bandwidth beside
    the 8 TB/s peak).
"""
from __future__ import annotations

import ctypes
from pathlib import Path

import numpy as np
import pytest

import oracle
from tests.gpu_util import require_gpu

pytestmark = pytest.mark.gpu

LIB = Path(__file__).resolve().parent / "native" / "libshgemm_test_kernels.so"


@pytest.fixture(scope="module")
def tk():
    require_gpu()
    if not LIB.exists():
        pytest.fail(f"{LIB} is missing: run __graft_entry__.build() first")
    L = ctypes.CDLL(str(LIB))
    vp = ctypes.c_void_p
    L.tk_mfma.argtypes = [ctypes.c_int, ctypes.c_int, vp, vp, vp, vp]
    L.tk_mfma.restype = ctypes.c_int
    L.tk_split_pair.argtypes = [ctypes.c_int, vp, vp, vp, vp]
    L.tk_split_pair.restype = ctypes.c_int
    L.tk_chain_gemm.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, vp, vp, vp]
    L.tk_chain_gemm.restype = ctypes.c_int
    L.tk_read_bw.argtypes = [ctypes.c_int64, ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_float)]
    L.tk_read_bw.restype = ctypes.c_int
    return L


def _p(a: np.ndarray):
    return a.ctypes.data_as(ctypes.c_void_p)


def _mfma(tk, kind, A, B, C):
    """A: cases x 32 x K, B: cases x K x 32 (FP16/BF16 bits or FP32), C: cases x 32 x 32 FP32."""
    A, B = np.ascontiguousarray(A), np.ascontiguousarray(B)
    C = np.ascontiguousarray(C, dtype=np.float32)
    D = np.empty_like(C)
    rc = tk.tk_mfma(kind, C.shape[0], _p(A), _p(B), _p(C), _p(D))
    assert rc == 0, f"tk_mfma failed with HIP error {rc}"
    return D


def _bf16_bits(x: np.ndarray) -> np.ndarray:
    return (x.astype(np.float32).view(np.uint32) >> 16).astype(np.uint16)  # exact for the small integers used


@pytest.mark.parametrize("kind", [0, 1, 2])
def test_mfma_lane_layout_exact_on_integers(tk, kind):
    rng = np.random.default_rng(40 + kind)
    cases, K = 64, (2 if kind == 2 else 16)
    A = rng.integers(-4, 5, (cases, 32, K)).astype(np.float32)
    B = rng.integers(-4, 5, (cases, K, 32)).astype(np.float32)
    C = rng.integers(-4, 5, (cases, 32, 32)).astype(np.float32)
    if kind == 0:
        Ad, Bd = A.astype(np.float16).view(np.uint16), B.astype(np.float16).view(np.uint16)
    elif kind == 1:
        Ad, Bd = _bf16_bits(A), _bf16_bits(B)
    else:
        Ad, Bd = A, B
    D = _mfma(tk, kind, Ad, Bd, C)
    want = C.astype(np.float64) + A.astype(np.float64) @ B.astype(np.float64)
    assert np.array_equal(D.astype(np.float64), want)


def _f16_random(rng, shape, emin, emax):
    """FP16 values with exponents in [emin, emax], random 10-bit mantissas and signs."""
    e = rng.integers(emin, emax + 1, shape)
    mant = rng.integers(0, 1024, shape)
    v = (1.0 + mant / 1024.0) * np.exp2(e) * rng.choice([-1.0, 1.0], shape)
    return v.astype(np.float16)


def _ulp32(x: np.ndarray) -> np.ndarray:
    """Unit in the last place of |x| rounded to FP32 (x float64)."""
    a = np.abs(x).astype(np.float32)
    return (np.nextafter(a, np.float32(np.inf)) - a).astype(np.float64)


def _accumulate_cases(tk, A, B, C):
    """Device result, exact sum (FP64: every case below keeps all partial sums
    multiples of 2^-26 below 2^15, exact in any order), and the largest
    operand magnitude max(|c|, |a_i b_i|) per output."""
    exact = C.astype(np.float64) + A.astype(np.float64) @ B.astype(np.float64)
    D = _mfma(tk, 0, A.view(np.uint16), B.view(np.uint16), C).astype(np.float64)
    prods = np.abs(A.astype(np.float64)[:, :, :, None] * B.astype(np.float64)[:, None, :, :]).max(axis=2)
    return D, exact, np.maximum(np.abs(C).astype(np.float64), prods)


def test_f16_mfma_accumulation_bound_and_bias(tk):
    """What v_mfma_f32_32x32x16_f16 does with the sum of its 16 exact products
    and the accumulator (measured; it is not one round-to-nearest-even):
      * one product onto the accumulator, and two products onto zero, are
        rounded faithfully (within one ulp of the result);
      * sums of 16 products and the accumulator are within 8 ulps of the
        largest operand plus one ulp of the result, and differ from the
        correctly rounded sum in a fifth of the outputs or more;
      * on same-sign operands the error leans toward zero (mean below -0.05
        ulp of the result), so along a long MFMA chain it adds up instead of
        cancelling."""
    rng = np.random.default_rng(77)
    cases = 600
    # products: |a|, |b| in [2^-3, 2^4): exponents -6..7, lowest bit 2^-26;
    # C: exponents -3..12 with 23-bit mantissas (lowest bit >= 2^-26)
    A = _f16_random(rng, (cases, 32, 16), -3, 3)
    B = _f16_random(rng, (cases, 16, 32), -3, 3)
    ce = rng.integers(-3, 13, (cases, 32, 32))
    cm = rng.integers(0, 1 << 23, (cases, 32, 32))
    sign = rng.choice([-1.0, 1.0], (cases, 32, 32))
    C = ((1.0 + cm / float(1 << 23)) * np.exp2(ce) * sign).astype(np.float32)
    zero = np.zeros_like(C)

    A1 = A.copy()
    A1[:, :, 1:] = 0
    D, exact, _ = _accumulate_cases(tk, A1, B, C)
    assert np.all(np.abs(D - exact) <= _ulp32(exact))
    A2 = A.copy()
    A2[:, :, 2:] = 0
    D, exact, _ = _accumulate_cases(tk, A2, B, zero)
    assert np.all(np.abs(D - exact) <= _ulp32(exact))

    for name, a, b, c, same_sign in (("16 products + c, mixed signs", A, B, C, False),
                                      ("16 products + c, all positive", np.abs(A), np.abs(B), np.abs(C), True),
                                      ("16 positive products, c = 0", np.abs(A), np.abs(B), zero, True)):
        D, exact, big = _accumulate_cases(tk, a, b, c)
        err = D - exact
        rel_big, rel_res = err / _ulp32(big), err / _ulp32(exact)
        off_rne = np.mean(D != exact.astype(np.float32).astype(np.float64))
        nz = err != 0
        toward0 = np.mean(np.sign(err[nz]) == -np.sign(exact[nz])) if nz.any() else 0.0
        print(f"{name}: {D.size} outputs, differ from round-to-nearest-even {off_rne:.3f}, "
              f"max |err| {np.max(np.abs(rel_big)):.2f} ulp(largest operand), mean err {np.mean(rel_res):+.3f} "
              f"ulp(result), toward zero {toward0:.3f} of nonzero errors", flush=True)
        assert np.all(np.abs(err) <= 8.0 * _ulp32(big) + _ulp32(exact)), np.max(np.abs(rel_big))
        assert off_rne > 0.2
        if same_sign:
            assert np.mean(rel_res) < -0.05


def _split_inputs(rng, npairs):
    """The probe's distribution: wide exponents, FP16's top range, rounding
    ties and FP16 subnormals; power-of-two scales 2^-10 .. 2^9."""
    n = 2 * npairs
    u = rng.uniform(0.0, 1.0, (4, n))
    i = np.arange(n) % 4
    wide = (u[0] * 2 - 1) * np.exp2(np.floor(u[1] * 60) - 30)
    top = (u[0] * 2 - 1) * 65504.0 * 1.01
    ties = np.exp2(np.floor(u[1] * 40) - 30) * (1.0 + np.floor(u[2] * 2048) / 2048 + 2.0 ** -12)
    sub = (u[0] * 2 - 1) * np.exp2(-24 - np.floor(u[3] * 8))
    x = np.select([i == 0, i == 1, i == 2], [wide, top, ties], sub).astype(np.float32)
    x[:4] = [0.0, -0.0, 65504.0, -65504.0]
    s = np.exp2(np.floor(rng.uniform(0.0, 1.0, npairs) * 20) - 10).astype(np.float32)
    s[:2] = 1.0
    return x, s


def test_paired_split_equals_oracle(tk):
    npairs = 1 << 21
    x, s = _split_inputs(np.random.default_rng(7), npairs)
    hi = np.empty(npairs, np.uint32)
    lo = np.empty(npairs, np.uint32)
    rc = tk.tk_split_pair(npairs, _p(x), _p(s), _p(hi), _p(lo))
    assert rc == 0, f"tk_split_pair failed with HIP error {rc}"
    xs = x * np.repeat(s, 2)                  # exact: power-of-two scales, no FP32 subnormals here
    assert np.all((xs == 0) | (np.abs(xs) >= np.float32(2.0 ** -126)))
    ohi, olo = oracle.split_f16_same_scale(xs)
    got_hi = np.stack([hi & 0xFFFF, hi >> 16], axis=1).reshape(-1).astype(np.uint16)
    got_lo = np.stack([lo & 0xFFFF, lo >> 16], axis=1).reshape(-1).astype(np.uint16)
    # A zero input gives zero parts whose sign may differ from the oracle's
    # (the fma forms add a +0 term: -0 * s + 0 = +0); a zero of either sign
    # adds nothing to a product sum.  Every other input: bit for bit.
    zero = xs == 0
    assert zero.sum() >= 2
    assert np.all((got_hi[zero] & 0x7FFF) == 0) and np.all((got_lo[zero] & 0x7FFF) == 0)
    bad_hi = np.flatnonzero((got_hi != ohi.view(np.uint16)) & ~zero)
    bad_lo = np.flatnonzero((got_lo != olo.view(np.uint16)) & ~zero)
    assert bad_hi.size == 0 and bad_lo.size == 0, (bad_hi[:5], bad_lo[:5], x[bad_hi[:5]], x[bad_lo[:5]])


@pytest.mark.parametrize("k_log", [16, 18])
def test_long_mfma_chain_vs_round_to_nearest_fp32(tk, k_log):
    """m = n = 256, positive inputs (any bias of the accumulation adds up
    coherently).  One MFMA chain of k/16 steps per output (the unsplit large
    tile, plan 1) in the prescaled and the classic fp16 mode: within 2x of the
    library FP32 SGEMM's error (the metric's bar), but several times a
    round-to-nearest FP32 GEMM's on the host.  The same product as 64 k-slices
    summed in FP32 outside the MFMA chain (split-K, plan 3), and the plan the
    library picks itself: within 2x of the round-to-nearest FP32 GEMM."""
    import torch

    import shgemm_amd as sg

    m = n = 256
    k = 1 << k_log
    rng = np.random.default_rng(k_log)
    A = rng.uniform(0.0, 1.0, m * k).astype(np.float32)
    B16 = rng.uniform(0.0, 1.0, k * n).astype(np.float16)
    B32 = B16.astype(np.float32)
    ref = A.reshape(k, m).T.astype(np.float64) @ B32.reshape(n, k).T.astype(np.float64)
    host = oracle.sgemm_nn_packed(m, n, k, 1.0, A, m, B32, k, 0.0, None, m).reshape(n, m).T
    err_rn = oracle.rel_fro_error(host, ref)
    a, b, b32 = torch.from_numpy(A).cuda(), torch.from_numpy(B16).cuda(), torch.from_numpy(B32).cuda()
    c = torch.empty(m * n, device="cuda")
    h = sg.create()
    errs = {}
    try:
        sg.sgemm_fp32(h, 0, 0, m, n, k, 1.0, a, m, b32, k, 0.0, c, m)
        err_sg = oracle.rel_fro_error(c.cpu().numpy().reshape(n, m).T, ref)
        for name, plan, slices, scaling in (("chain prescaled", 1, 1, True), ("chain classic", 1, 1, False),
                                            ("64 slices", 3, 64, True), ("library plan", -1, 1, True)):
            sg.force_plan(h, plan, slices)
            sg.set_exponent_scaling(h, scaling)
            sg.shgemm(h, 0, 0, m, n, k, 1.0, a, m, b, k, 0.0, c, m, 0)
            errs[name] = oracle.rel_fro_error(c.cpu().numpy().reshape(n, m).T, ref)
    finally:
        sg.destroy(h)
    print(f"k=2^{k_log}: round-to-nearest FP32 {err_rn:.3e}, library SGEMM {err_sg:.3e}, "
          + ", ".join(f"{key} {v:.3e}" for key, v in errs.items()), flush=True)
    assert errs["chain prescaled"] <= 2.0 * err_sg and errs["chain classic"] <= 2.0 * err_sg, (errs, err_sg)
    assert errs["chain prescaled"] > 2.0 * err_rn          # the chain's accumulation loses bits
    assert errs["64 slices"] <= 2.0 * err_rn and errs["library plan"] <= 2.0 * err_rn, (errs, err_rn)


@pytest.mark.parametrize("k_log", [14, 16, 18])
def test_hi_outside_the_chain_vs_two_chains(tk, k_log):
    """The reference adds each MMA's hi product to its accumulator with an
    FP32 add outside the MMA chain (src/wmmae_shgemm.hpp:86-98).  The same
    split GEMM on gfx950 MFMA (m = n = 256, positive inputs), accumulated as
    two MFMA chains (the library's classic mode), with the hi product outside
    the chain, and with both outside: the chains lose bits against the
    reference's remedy (1.3x its error or more, growing with k), the
    residual's chain costs nothing (its products are 2^-11 of the hi ones),
    and the library's own plan for these shapes, k-slices summed in FP32
    outside the chain, is at least as accurate as the remedy from k = 2^16.
    Both stay above the blocked round-to-nearest host GEMM, whose sum over k
    is two-level."""
    import torch

    import shgemm_amd as sg

    m = n = 256
    k = 1 << k_log
    rng = np.random.default_rng(100 + k_log)
    A = rng.uniform(0.0, 1.0, (m, k)).astype(np.float32)        # row-major m x k
    B16 = rng.uniform(0.0, 1.0, (n, k)).astype(np.float16)      # column j of B in row j
    ref = A.astype(np.float64) @ B16.astype(np.float64).T
    A_cm = np.asfortranarray(A).ravel(order="F")                # column-major m x k
    host = oracle.sgemm_nn_packed(m, n, k, 1.0, A_cm, m, np.ascontiguousarray(B16.astype(np.float32)).ravel(), k,
                                  0.0, None, m)
    err_rn = oracle.rel_fro_error(host.reshape(n, m).T, ref)
    errs = {}
    for mode, name in ((0, "two chains"), (1, "hi outside"), (2, "both outside")):
        C = np.empty((m, n), np.float32)
        rc = tk.tk_chain_gemm(mode, m, n, k, _p(A), _p(B16.view(np.uint16)), _p(C))
        assert rc == 0, f"tk_chain_gemm failed with HIP error {rc}"
        errs[name] = oracle.rel_fro_error(C, ref)
    h = sg.create()
    try:
        a, b = torch.from_numpy(A_cm).cuda(), torch.from_numpy(B16.ravel()).cuda()
        c = torch.empty(m * n, device="cuda")
        plan = sg.query_plan(h, 0, 0, m, n, k, sg.fp16)[:2]
        sg.shgemm(h, 0, 0, m, n, k, 1.0, a, m, b, k, 0.0, c, m, sg.fp16)
        errs["library plan"] = oracle.rel_fro_error(c.cpu().numpy().reshape(n, m).T, ref)
    finally:
        sg.destroy(h)
    print(f"k=2^{k_log}: blocked round-to-nearest host FP32 {err_rn:.3e}, "
          + ", ".join(f"{key} {v:.3e}" for key, v in errs.items()) + f" (plan {plan})", flush=True)
    assert errs["two chains"] > 1.3 * errs["hi outside"], errs
    assert abs(errs["both outside"] - errs["hi outside"]) <= 0.05 * errs["hi outside"], errs
    if k_log >= 16:
        assert errs["library plan"] <= errs["hi outside"], errs


def test_hbm_read_ceiling(tk):
    """The streaming read rate of this device (1 GiB, 16-byte loads, one
    contiguous chunk per workgroup): the attainable figure beside the 8 TB/s
    peak for the bandwidth-bound kernels."""
    rates = {}
    for grid in (2048, 4096, 8192):
        ms = ctypes.c_float()
        rc = tk.tk_read_bw(1 << 30, grid, 20, ctypes.byref(ms))
        assert rc == 0, f"tk_read_bw failed with HIP error {rc}"
        rates[grid] = (1 << 30) / (ms.value * 1e-3) / 1e12
    print("HBM streaming read, 1 GiB: " + ", ".join(f"{g} workgroups {r:.2f} TB/s" for g, r in rates.items()),
          flush=True)
    assert 3.0 < max(rates.values()) < 8.0, rates
