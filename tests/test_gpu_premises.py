"""GPU: the arithmetic premises the kernels rest on, checked on the device.

  * MFMA lane layout: v_mfma_f32_32x32x16_f16 / _bf16 and
    v_mfma_f32_32x32x2_f32 with the operand and accumulator layout the
    kernels assume give the exact product on integer-valued operands.
  * FP16 MFMA accumulation is one round-to-nearest-even of the exact sum of
    the 16 products and the accumulator.  This is why the prescaled fp16 mode
    chains the hi and lo products in one MFMA accumulator: the reference sums
    the hi product outside the MMA chain (src/wmmae_shgemm.hpp:86-98) to work
    around accumulation that loses bits; where accumulation rounds once to
    nearest, nothing is lost by keeping it inside.  Checked on random cases
    whose exact sums are representable in FP64 (so numpy's FP64->FP32 cast is
    the correctly rounded result) and on crafted ties, sub-ulp sums and
    cancellations; the cases are chosen so that truncation would fail them.
  * The paired split of the prescaled fp16 mode (split2_f16_scaled,
    v_fma_mix forms) equals the oracle's split_f16_same_scale of x * s bit
    for bit, over 2^22 values: wide exponents, FP16's top range, rounding
    ties and FP16 subnormals, at power-of-two scales 2^-10 .. 2^9.
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


def _rtz_f32(x: np.ndarray) -> np.ndarray:
    r = x.astype(np.float32)
    over = np.abs(r.astype(np.float64)) > np.abs(x)
    r[over] = np.nextafter(r[over], np.float32(0.0))
    return r


def test_f16_mfma_accumulates_with_one_rne_rounding(tk):
    rng = np.random.default_rng(77)
    cases = 1500
    # products: exponents -6..7, lowest bit 2^-26; C: exponents -3..12 with
    # 23-bit mantissas (lowest bit >= 2^-26): every partial sum is a multiple
    # of 2^-26 below 2^15, exact in FP64 whatever the order
    A = _f16_random(rng, (cases, 32, 16), -3, 3)
    B = _f16_random(rng, (cases, 16, 32), -3, 3)
    ce = rng.integers(-3, 13, (cases, 32, 32))
    cm = rng.integers(0, 1 << 23, (cases, 32, 32))
    C = ((1.0 + cm / float(1 << 23)) * np.exp2(ce) * rng.choice([-1.0, 1.0], (cases, 32, 32))).astype(np.float32)
    exact = C.astype(np.float64) + A.astype(np.float64) @ B.astype(np.float64)
    # crafted cases in the last few: ties to even, sums below half an ulp,
    # cancellation, and accumulator-less sums (c = 0)
    t, u = 2.0 ** -13, 2.0 ** -12
    crafted = [
        (1.0, [(t, u)] * 16),                               # 16 x 2^-25 onto 1: exactly 2^-21
        (1.0, [(0.75 * 2.0 ** -11, 2.0 ** -12)]),           # 0.75 ulp: rounds up
        (1.0, [(0.5 * 2.0 ** -11, 2.0 ** -12)]),            # tie: to even (down)
        (1.0, [(1.5 * 2.0 ** -11, 2.0 ** -12)]),            # tie: to even (up)
        (1.0, [(-0.75 * 2.0 ** -12, 2.0 ** -12)]),          # below 1: rounds down by a quarter ulp
        (1.0, [(0.25 * 2.0 ** -11, 2.0 ** -12)]),           # quarter ulp: stays
        (1.0, [(0.3 * 2.0 ** -11, 2.0 ** -12)] * 2),        # two sub-half-ulp terms that sum above half
        (0.0, [(1.0, 1.0), (-1.0, 1.0), (t, 2.0 ** -11)]),  # cancellation leaves 2^-24
        (0.0, [(2048.0, 2048.0), (-2048.0, 2047.0), (t, u)]),
        (0.0, [(1.0, 1.0), (2.0 ** -12, 2.0 ** -12), (2.0 ** -12, 2.0 ** -12)]),
    ]
    ncr = len(crafted)
    A[-ncr:] = 0
    B[-ncr:] = 0
    for i, (c0, prods) in enumerate(crafted):
        j = cases - ncr + i
        C[j] = np.float32(c0)
        for kk, (x, y) in enumerate(prods):
            A[j, 0, kk] = np.float16(x)
            B[j, kk, 0] = np.float16(y)
        exact[j] = C[j].astype(np.float64) + A[j].astype(np.float64) @ B[j].astype(np.float64)
    D = _mfma(tk, 0, A.view(np.uint16), B.view(np.uint16), C)
    want = exact.astype(np.float32)          # correctly rounded (round to nearest, ties to even)
    rounded = want.astype(np.float64) != exact
    truncation_differs = _rtz_f32(exact) != want
    assert rounded.mean() > 0.5 and truncation_differs.mean() > 0.2   # the cases exercise the rounding
    bad = np.argwhere(D != want)
    assert bad.size == 0, (f"{len(bad)} of {D.size} outputs differ from one RNE rounding; first "
                           f"{[(tuple(b), float(D[tuple(b)]), float(exact[tuple(b)])) for b in bad[:4]]}")
    # the crafted cases in particular (output element (0, 0) of each)
    j0 = cases - ncr
    assert np.array_equal(D[j0:, 0, 0], want[j0:, 0, 0])


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
    bad_hi = np.flatnonzero(got_hi != ohi.view(np.uint16))
    bad_lo = np.flatnonzero(got_lo != olo.view(np.uint16))
    assert bad_hi.size == 0 and bad_lo.size == 0, (bad_hi[:5], bad_lo[:5], x[bad_hi[:5]], x[bad_lo[:5]])
