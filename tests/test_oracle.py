"""Pins the CPU oracle against independent implementations, the golden
fixtures, closed forms and brute force.  CPU only.

What pins what:
  * FP16 / BF16 conversions: numpy's IEEE binary16 cast and torch's bfloat16
    cast (independent round-to-nearest-even implementations), bit for bit,
    over special values, ties, subnormals and random bit patterns.
  * The FP16 split: the committed numpy fixture, and the invariants that fix
    the method (hi is the FP16 rounding of a; hi + 2^-11 lo reproduces a to
    2^-22; lo = 0 exactly when a is FP16-representable).
  * The BF16 splits: exact reconstruction a0 + a1 + a2 == a and b0 + b1 == b
    (exhaustive over all finite FP16 values).
  * FP64 GEMM: numpy float64 matmul (BLAS) on every op/ld/alpha/beta case,
    integer known answers exactly, brute-force loops on tiny shapes.
  * Planted mistakes (dropped lo term, wrong scale, transposed operand, sign)
    in an oracle-side emulation all exceed the error bar the GPU tests use.
"""
from __future__ import annotations

from pathlib import Path

import numpy as np
import pytest
import torch

import oracle
from tests.golden.make_golden import load_cases

GOLDEN = Path(__file__).resolve().parent / "golden"


@pytest.fixture(scope="module", autouse=True)
def _build():
    oracle.build()


def _special_floats():
    vals = [0.0, -0.0, 1.0, -1.0, 65504.0, 65519.996, 65520.0, 7e4, 2.0 ** -14, 2.0 ** -24, 2.0 ** -25,
            1.5 * 2.0 ** -25, 2.0 ** -26, 1e-40, -1e-45, np.inf, -np.inf, 3.4e38]
    ties = [(1.0 + (2 * i + 1) * 2.0 ** -11) for i in range(16)] + [2.0 ** -20 * (1 + 2 ** -10) for _ in range(1)]
    sub_ties = [(2 * i + 1) * 2.0 ** -25 for i in range(64)]
    return np.array(vals + ties + [-t for t in ties] + sub_ties, dtype=np.float32)


def _random_bits(n, seed=0):
    rng = np.random.default_rng(seed)
    bits = rng.integers(0, 2 ** 32, n, dtype=np.uint64).astype(np.uint32)
    f = bits.view(np.float32)
    return f[np.isfinite(f)]


def test_f16_conversion_matches_numpy():
    a = np.concatenate([_special_floats(), _random_bits(200000), np.random.default_rng(1).uniform(-70000, 70000, 100000)
                        .astype(np.float32), np.random.default_rng(2).uniform(-1e-4, 1e-4, 100000).astype(np.float32)])
    L = oracle.lib()
    ours = np.array([L.oracle_f32_to_f16(float(x)) for x in a[:4000]], dtype=np.uint16)
    with np.errstate(over="ignore"):
        ref = a[:4000].astype(np.float16).view(np.uint16)
    assert np.array_equal(ours, ref)
    hi, _ = oracle.split_f16(a)
    with np.errstate(over="ignore"):
        assert np.array_equal(hi, a.astype(np.float16).view(np.uint16))


def test_bf16_conversion_matches_torch():
    a = np.concatenate([_special_floats(), _random_bits(200000, seed=3)])
    p0, _, _ = oracle.split_bf16x3(a)
    ref = torch.from_numpy(a).to(torch.bfloat16).view(torch.int16).numpy().view(np.uint16)
    assert np.array_equal(p0, ref)


def test_split_f16_golden_fixture():
    z = np.load(GOLDEN / "split_f16.npz", allow_pickle=False)
    hi, lo = oracle.split_f16(z["a"])
    assert np.array_equal(hi, z["hi_bits"])
    finite = np.isfinite(z["hi_bits"].view(np.float16))
    assert np.array_equal(lo[finite], z["lo_bits"][finite])


def test_split_f16_invariants():
    rng = np.random.default_rng(4)
    # FP16 normal range, away from overflow: |a| in [2^-12, 2^15)
    a = (np.where(rng.random(100000) < 0.5, -1, 1) * np.ldexp(rng.uniform(1, 2, 100000),
                                                               rng.integers(-12, 15, 100000))).astype(np.float32)
    hi, lo = oracle.split_f16(a)
    h = oracle.f16_bits_to_f32(hi).astype(np.float64)
    l = oracle.f16_bits_to_f32(lo).astype(np.float64)
    rep = h + np.ldexp(l, -11)
    rel = np.abs(rep - a.astype(np.float64)) / np.abs(a.astype(np.float64))
    assert rel.max() <= 2.0 ** -22
    # hi carries the leading 11 bits: |a - hi| <= half an FP16 ulp
    assert (np.abs(a - h) <= np.ldexp(1.0, np.floor(np.log2(np.abs(a))).astype(int) - 11)).all()
    # lo vanishes exactly when a is an FP16 value
    exact = a.astype(np.float16).astype(np.float32)
    _, lo_e = oracle.split_f16(exact)
    assert (oracle.f16_bits_to_f32(lo_e) == 0).all()
    assert (l != 0).mean() > 0.99


def test_split_bf16x3_reconstructs_exactly():
    rng = np.random.default_rng(5)
    a = (np.where(rng.random(200000) < 0.5, -1, 1) * np.ldexp(rng.uniform(1, 2, 200000),
                                                              rng.integers(-100, 100, 200000))).astype(np.float32)
    p = [oracle.bf16_bits_to_f32(x).astype(np.float64) for x in oracle.split_bf16x3(a)]
    assert np.array_equal(p[0] + p[1] + p[2], a.astype(np.float64))


def test_split_bf16x2_exhaustive_fp16():
    bits = np.arange(65536, dtype=np.uint16)
    f = bits.view(np.float16).astype(np.float64)
    fin = np.isfinite(f)
    b0, b1 = (oracle.bf16_bits_to_f32(x).astype(np.float64) for x in oracle.split_bf16x2(bits))
    assert np.array_equal((b0 + b1)[fin], f[fin])


def _np_gemm(c):
    Am = oracle.op_view(c["A"].astype(np.float64), c["op_a"], c["m"], c["k"], c["lda"])
    Bm = oracle.op_view(c["B32"].astype(np.float64), c["op_b"], c["k"], c["n"], c["ldb"])
    out = c["C0"].astype(np.float64).copy()
    v = oracle._as_c_view(out, c["op_c"], c["m"], c["n"], c["ldc"])
    v[...] = c["alpha"] * (Am @ Bm) + c["beta"] * v
    return out


def _case(rng, op_a, op_b, op_c, m, n, k, pad, alpha, beta):
    lda = (m if op_a == 0 else k) + pad
    ldb = (k if op_b == 0 else n) + pad
    ldc = (m if op_c == 0 else n) + pad
    A = oracle.gen_uniform(rng, lda * (k if op_a == 0 else m))
    B16 = oracle.gen_f16_uniform(rng, ldb * (n if op_b == 0 else k))
    C0 = oracle.gen_uniform(rng, ldc * (n if op_c == 0 else m))
    return dict(op_a=op_a, op_b=op_b, op_c=op_c, m=m, n=n, k=k, lda=lda, ldb=ldb, ldc=ldc, alpha=alpha, beta=beta,
                A=A, B32=B16.astype(np.float32), C0=C0)


@pytest.mark.parametrize("op_a", [0, 1])
@pytest.mark.parametrize("op_b", [0, 1])
@pytest.mark.parametrize("op_c", [0, 1])
def test_gemm_f64_matches_numpy(op_a, op_b, op_c):
    rng = np.random.default_rng(10 + op_a + 2 * op_b + 4 * op_c)
    for (m, n, k, pad, alpha, beta) in [(37, 53, 300, 3, 1.0, 0.0), (300, 257, 70, 0, -0.5, 1.5), (1, 1, 1, 0, 2, 0)]:
        c = _case(rng, op_a, op_b, op_c, m, n, k, pad, alpha, beta)
        got = oracle.gemm_f64(op_a, op_b, op_c, m, n, k, alpha, c["A"], c["lda"], c["B32"], c["ldb"], beta, c["C0"],
                              c["ldc"])
        ref = _np_gemm(c)
        gv, rv = oracle.c_matrix(got, op_c, m, n, c["ldc"]), oracle.c_matrix(ref, op_c, m, n, c["ldc"])
        assert np.abs(gv - rv).max() <= 1e-12 * max(1.0, np.abs(rv).max())


def test_gemm_f64_golden_fixtures():
    for name, tol in (("gemm_int.npz", 0.0), ("gemm_rand.npz", 1e-12)):
        for c in load_cases(GOLDEN / name):
            B32 = c["B_bits"].view(np.float16).astype(np.float32)
            got = oracle.gemm_f64(c["op_a"], c["op_b"], c["op_c"], c["m"], c["n"], c["k"], c["alpha"], c["A"],
                                  c["lda"], B32, c["ldb"], c["beta"], c["C0"], c["ldc"])
            gv = oracle.c_matrix(got, c["op_c"], c["m"], c["n"], c["ldc"])
            rv = oracle.c_matrix(c["C64"], c["op_c"], c["m"], c["n"], c["ldc"])
            assert np.abs(gv - rv).max() <= tol * max(1.0, np.abs(rv).max()), name


def test_gemm_brute_force_tiny():
    rng = np.random.default_rng(12)
    for op_a in (0, 1):
        for op_b in (0, 1):
            for op_c in (0, 1):
                m, n, k = 3, 2, 5
                c = _case(rng, op_a, op_b, op_c, m, n, k, 1, 0.75, -0.5)
                got = oracle.gemm_f64(op_a, op_b, op_c, m, n, k, 0.75, c["A"], c["lda"], c["B32"], c["ldb"], -0.5,
                                      c["C0"], c["ldc"])
                for i in range(m):
                    for j in range(n):
                        s = 0.0
                        for p in range(k):
                            a = c["A"][i + p * c["lda"]] if op_a == 0 else c["A"][p + i * c["lda"]]
                            b = c["B32"][p + j * c["ldb"]] if op_b == 0 else c["B32"][j + p * c["ldb"]]
                            s += float(a) * float(b)
                        o = i + j * c["ldc"] if op_c == 0 else j + i * c["ldc"]
                        assert abs(got[o] - (0.75 * s - 0.5 * float(c["C0"][o]))) <= 1e-15


def test_gemm_f32_is_fp32_accurate():
    rng = np.random.default_rng(13)
    c = _case(rng, 0, 0, 0, 256, 256, 1024, 0, 1.0, 0.0)
    f64 = oracle.gemm_f64(0, 0, 0, 256, 256, 1024, 1.0, c["A"], 256, c["B32"], 1024, 0.0, None, 256)
    f32 = oracle.gemm_f32(0, 0, 0, 256, 256, 1024, 1.0, c["A"], 256, c["B32"], 1024, 0.0, None, 256)
    err = oracle.rel_fro_error(f32, f64)
    assert 1e-9 < err < 2e-6
    # the row-restricted form computes exactly the same rows
    part = oracle.gemm_f32_rows(0, 0, 0, 256, 256, 1024, 1.0, c["A"], 256, c["B32"], 1024, 0.0, None, 256, 64, 128)
    v, w = oracle.c_matrix(part, 0, 256, 256, 256), oracle.c_matrix(f32, 0, 256, 256, 256)
    assert np.array_equal(v[64:128], w[64:128])


@pytest.mark.parametrize("m,n,k", [(1, 1, 1), (15, 5, 7), (37, 29, 300), (200, 130, 513), (400, 4033, 260)])
def test_sgemm_nn_packed_exact_on_integers(m, n, k):
    """Integer inputs with |a|, |b| <= 8 and k <= 513: every partial sum is an
    integer below 2^24, so the packed FP32 GEMM must equal the FP64 product
    exactly, over ragged tiles (m % 16, n % 6, k % 256, n past one 4032-column
    panel), alpha/beta and a row range."""
    rng = np.random.default_rng(m * 7 + n)
    A = rng.integers(-8, 9, m * k).astype(np.float32)
    B = rng.integers(-8, 9, k * n).astype(np.float32)
    C0 = rng.integers(-8, 9, (m + 3) * n).astype(np.float32)
    ref = A.reshape(k, m).T.astype(np.float64) @ B.reshape(n, k).T.astype(np.float64)
    got = oracle.sgemm_nn_packed(m, n, k, 1.0, A, m, B, k, 0.0, None, m)
    assert np.array_equal(got.reshape(n, m).T.astype(np.float64), ref)
    got = oracle.sgemm_nn_packed(m, n, k, 2.0, A, m, B, k, -1.0, C0, m + 3)
    want = 2.0 * ref - C0.reshape(n, m + 3).T[:m].astype(np.float64)
    assert np.array_equal(got.reshape(n, m + 3).T[:m].astype(np.float64), want)
    r0, r1 = m // 3, m - m // 4
    part = oracle.sgemm_nn_packed(m, n, k, 1.0, A, m, B, k, 0.0, None, m, r0, r1)
    v = part.reshape(n, m).T
    assert np.array_equal(v[r0:r1].astype(np.float64), ref[r0:r1]) and not v[:r0].any() and not v[r1:].any()


def test_sgemm_nn_packed_is_fp32_accurate():
    rng = np.random.default_rng(14)
    m, n, k = 300, 200, 4096
    A = oracle.gen_uniform(rng, m * k)
    B = oracle.gen_uniform(rng, k * n)
    f64 = oracle.gemm_f64(0, 0, 0, m, n, k, 1.0, A, m, B, k, 0.0, None, m)
    err_p = oracle.rel_fro_error(oracle.sgemm_nn_packed(m, n, k, 1.0, A, m, B, k, 0.0, None, m), f64)
    err_f = oracle.rel_fro_error(oracle.gemm_f32(0, 0, 0, m, n, k, 1.0, A, m, B, k, 0.0, None, m), f64)
    assert 1e-9 < err_p < 2.0 * err_f


def test_emulation_and_planted_mistakes():
    """The 2x-SGEMM bar of the GPU tests rejects each plausible kernel mistake."""
    rng = np.random.default_rng(14)
    m = n = k = 256
    c = _case(rng, 0, 0, 0, m, n, k, 0, 1.0, 0.0)
    f64 = oracle.gemm_f64(0, 0, 0, m, n, k, 1.0, c["A"], m, c["B32"], k, 0.0, None, m)
    f32 = oracle.gemm_f32(0, 0, 0, m, n, k, 1.0, c["A"], m, c["B32"], k, 0.0, None, m)
    bar = 2.0 * oracle.rel_fro_error(f32, f64)
    for ct in (0, 1):
        emu = oracle.emulate_shgemm(0, 0, 0, m, n, k, 1.0, c["A"], m, c["B32"], k, 0.0, None, m, ct)
        assert oracle.rel_fro_error(emu, f64) <= bar / 4
    hi, lo = oracle.split_f16(c["A"])
    h = oracle.f16_bits_to_f32(hi).astype(np.float64)
    l = oracle.f16_bits_to_f32(lo).astype(np.float64)
    B = c["B32"].astype(np.float64).reshape(n, k).T
    good = (h + np.ldexp(l, -11)).reshape(k, m).T @ B
    assert oracle.rel_fro_error(good, oracle.c_matrix(f64, 0, m, n, m)) <= bar / 4
    mistakes = {
        "dropped lo": h.reshape(k, m).T @ B,
        "lo scale 2^-10": (h + np.ldexp(l, -10)).reshape(k, m).T @ B,
        "lo sign": (h - np.ldexp(l, -11)).reshape(k, m).T @ B,
        "A transposed": (h + np.ldexp(l, -11)).reshape(k, m) @ B,
    }
    ref = oracle.c_matrix(f64, 0, m, n, m)
    for name, wrong in mistakes.items():
        assert oracle.rel_fro_error(wrong, ref) > 8 * bar, name


def _running_case(rng, m, k):
    """Rows with their maximum in various k-tiles, subnormal and zero rows."""
    A = rng.uniform(-1.0, 1.0, (m, k)) * np.exp2(rng.integers(-40, 40, (m, 1)))
    grow = np.exp2(np.arange(k) // 32 * rng.integers(-3, 4, (m, 1)))
    A = (A * grow).astype(np.float32)
    A[3] = np.float32(1e-40) * rng.uniform(-1, 1, k).astype(np.float32)   # subnormal row
    A[5, :40] = 0.0                                                        # two zero tiles first
    A[7] = 0.0
    return A


@pytest.mark.parametrize("ktile,kslice", [(32, None), (64, None), (32, 96), (64, 128)])
def test_running_row_exponents_properties(ktile, kslice):
    """Running exponents: the last k-tile's equals the whole-row exponent of
    the C oracle (finite non-zero rows, unsplit); non-decreasing within a
    k-slice and restarting at each; every k-tile fits FP16 after scaling
    (max |a| 2^-e < 2^15), and where the exponent rose above the previous
    one the tile's maximum sets it (max |a| 2^-e >= 2^14)."""
    rng = np.random.default_rng(ktile + (kslice or 0))
    m, k = 48, 300
    A = _running_case(rng, m, k)
    e = oracle.running_row_exponents(A, ktile, kslice)
    T = e.shape[1]
    span = (kslice or k) // ktile if kslice else T
    if not kslice:
        live = np.abs(A).max(axis=1) > 0
        assert np.array_equal(e[live, -1], oracle.row_exponents(A)[live])
        assert np.all(e[~live] == -126)
    for t in range(T):
        tile = np.abs(A[:, t * ktile:(t + 1) * ktile]).max(axis=1).astype(np.float64)
        assert np.all(np.ldexp(tile, -e[:, t]) < 2.0 ** 15)
        first = t % span == 0
        prev = np.full(m, -126) if first else e[:, t - 1]
        assert np.all(e[:, t] >= prev)
        rose = e[:, t] > prev
        assert np.all(np.ldexp(tile[rose], -e[rose, t]) >= 2.0 ** 14)
    if kslice:  # each slice alone gives the same exponents
        for s0 in range(0, k, kslice):
            sub = oracle.running_row_exponents(A[:, s0:s0 + kslice], ktile)
            assert np.array_equal(sub, e[:, s0 // ktile:s0 // ktile + sub.shape[1]])


def test_running_split_reconstructs_and_matches_whole_row():
    """The running split reconstructs each element to 2^(e - 7) (the FP16 lo
    part's rounding at the k-tile's scale); rows whose maximum lies in their
    first k-tile split exactly as with the whole-row exponent."""
    rng = np.random.default_rng(31)
    m, k = 40, 256
    A = _running_case(rng, m, k)
    v = oracle.split_value_f16_running(A, 32)
    e = np.repeat(oracle.running_row_exponents(A, 32), 32, axis=1)[:, :k]
    assert np.all(np.abs(v - A.astype(np.float64)) <= np.ldexp(1.0, e - 7))
    first = np.abs(A[:, :32]).max(axis=1) == np.abs(A).max(axis=1)
    first &= np.abs(A).max(axis=1) > 0
    assert first.sum() >= 3
    w = oracle.split_value_f16_rowscaled(A[first])
    assert np.array_equal(v[first], w)


def test_growing_rows_separate_running_from_whole_row():
    """On balanced growing rows the running-exponent emulation stays within
    the FP32 bar of the FP64 product while the whole-row emulation does not
    (the case the GPU test uses to tell the two apart)."""
    rng = np.random.default_rng(8)
    m, n, k = 96, 64, 320
    A, B16 = oracle.gen_growing_rows(rng, m, n, k)
    B = B16.astype(np.float32)
    f64 = oracle.gemm_f64(0, 0, 0, m, n, k, 1.0, A, m, B, k, 0.0, None, m)
    f32 = oracle.gemm_f32(0, 0, 0, m, n, k, 1.0, A, m, B, k, 0.0, None, m)
    bar = 2.0 * oracle.rel_fro_error(f32, f64)
    run = oracle.emulate_shgemm(0, 0, 0, m, n, k, 1.0, A, m, B, k, 0.0, None, m, 3, ktile=32)
    whole = oracle.emulate_shgemm(0, 0, 0, m, n, k, 1.0, A, m, B, k, 0.0, None, m, 2)
    assert oracle.rel_fro_error(run, f64) <= bar / 4
    assert oracle.rel_fro_error(whole, f64) > 8 * bar


def test_wide_exponent_generator():
    rng = np.random.default_rng(15)
    a = oracle.gen_wide_exponent(rng, 100000)
    e = np.floor(np.log2(np.abs(a))).astype(int)
    assert e.min() == -15 and e.max() == 15
    assert (np.abs(a) >= 2.0 ** -15).all() and (np.abs(a) < 2.0 ** 16).all()
    assert 0.45 < (a > 0).mean() < 0.55


# Philox4x32-10 known-answer vectors published with the Random123 library
# (Salmon et al., SC'11): (counter, key) -> output words.
PHILOX_KAT = [
    ((0x00000000, 0x00000000, 0x00000000, 0x00000000), (0x00000000, 0x00000000),
     (0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8)),
    ((0xFFFFFFFF, 0xFFFFFFFF, 0xFFFFFFFF, 0xFFFFFFFF), (0xFFFFFFFF, 0xFFFFFFFF),
     (0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD)),
    ((0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344), (0xA4093822, 0x299F31D0),
     (0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1)),
]


@pytest.mark.parametrize("ctr,key,want", PHILOX_KAT)
def test_philox_known_answers(ctr, key, want):
    assert tuple(oracle.philox4x32_10(ctr, key)) == want


def test_philox_block_counter_and_key_layout():
    """Block b of seed s is Philox(counter (b lo, b hi, 0, 0), key (s lo, s hi))."""
    seed = (0x0123456789ABCDEF)
    w = oracle.philox_words(3, seed)
    for b in range(3):
        want = oracle.philox4x32_10((b, 0, 0, 0), (seed & 0xFFFFFFFF, seed >> 32))
        assert list(w[4 * b:4 * b + 4]) == want


def test_gaussian_box_muller_closed_form():
    """Normal 2p / 2p+1 of a block is the Box-Muller pair of words (2p, 2p+1)."""
    import math

    seed = 77
    w = oracle.philox_words(2, seed)
    z = oracle.gaussian(8, seed, 1.0)
    u = lambda x: ((int(x) >> 9) + 0.5) / 2.0 ** 23
    for b in range(2):
        for p in range(2):
            r = math.sqrt(-2.0 * math.log(u(w[4 * b + 2 * p])))
            t = 2.0 * math.pi * u(w[4 * b + 2 * p + 1])
            assert z[4 * b + 2 * p] == pytest.approx(r * math.cos(t), rel=1e-12, abs=1e-12)
            assert z[4 * b + 2 * p + 1] == pytest.approx(r * math.sin(t), rel=1e-12, abs=1e-12)
    big = oracle.gaussian(1 << 20, 5, 0.5)
    assert abs(big.mean()) < 3e-3 and abs(big.std() - 0.5) < 2e-3
