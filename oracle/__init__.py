"""CPU oracle for the SHGEMM hot path — test infrastructure only.

Only tests/, __graft_entry__.smoke() and the CPU-baseline leg of bench.py may
import this package, and only as the checker.  The product never calls it.

Contents
  gemm_f64 / gemm_f64_rows   FP64 reference product (C restatement in
                             oracle.c; replaces the un-vendored mateval
                             reference product used at test/shgemm_test.cu:91-100)
  gemm_f32 / gemm_f32_rows   host FP32 GEMM, the timed CPU baseline
  split_f16                  hi/lo FP16 split (reference src/wmmae_shgemm.hpp:29-31)
  split_bf16x3, split_bf16x2 BF16 parts of the tf32 compute type
  emulate_shgemm             FP64 evaluation of the split representation the
                             GPU multiplies: alpha*(op(hi) + 2^-11 op(lo)) op(B) + beta*C0,
                             and the prescaled forms (whole-row exponents,
                             running_row_exponents of the P0/P1/split-K kernels)
  rel_fro_error              ||C - C64||_F / ||C64||_F (the error metric)
  philox4x32_10, philox_words, gaussian
                             the sketch generator (Philox4x32-10 + Box-Muller),
                             restated for checking the device Omega
  gen_*                      seeded input generators shared by tests and bench
"""
from __future__ import annotations

import ctypes
import os
from pathlib import Path

import numpy as np

_HERE = Path(__file__).resolve().parent
_LIB_PATH = _HERE / "liboracle.so"
_lib = None

OP_N, OP_T = 0, 1


def build() -> Path:
    import subprocess

    r = subprocess.run(["make", "-s", "-C", str(_HERE)], capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("oracle build failed:\n" + r.stdout + r.stderr)
    return _LIB_PATH


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not _LIB_PATH.exists():
            build()
        L = ctypes.CDLL(str(_LIB_PATH))
        i64, i32, f64, f32 = ctypes.c_int64, ctypes.c_int, ctypes.c_double, ctypes.c_float
        vp = ctypes.c_void_p
        L.oracle_gemm_f64_rows.argtypes = [i32, i32, i32, i64, i64, i64, f64, vp, i64, vp, i64, f64, vp, i64, vp,
                                           i64, i64, i32]
        L.oracle_gemm_f32_rows.argtypes = [i32, i32, i32, i64, i64, i64, f32, vp, i64, vp, i64, f32, vp, i64, vp,
                                           i64, i64, i32]
        L.oracle_sgemm_nn_packed.argtypes = [i64, i64, i64, f32, vp, i64, vp, i64, f32, vp, i64, vp, i64, i64, i32]
        L.oracle_split_f16.argtypes = [vp, vp, vp, i64]
        L.oracle_split_bf16x3.argtypes = [vp, vp, vp, vp, i64]
        L.oracle_split_f16_same_scale.argtypes = [vp, vp, vp, i64]
        L.oracle_row_exponent.argtypes = [vp, i64, i64]
        L.oracle_row_exponent.restype = i32
        L.oracle_split_bf16x2.argtypes = [vp, vp, vp, i64]
        L.oracle_f32_to_f16.argtypes = [f32]
        L.oracle_f32_to_f16.restype = ctypes.c_uint16
        L.oracle_f32_to_bf16.argtypes = [f32]
        L.oracle_f32_to_bf16.restype = ctypes.c_uint16
        L.oracle_max_threads.restype = i32
        L.oracle_philox4x32_10.argtypes = [vp, vp, vp]
        L.oracle_philox_words.argtypes = [vp, i64, ctypes.c_uint64]
        L.oracle_gaussian.argtypes = [vp, i64, ctypes.c_uint64, f64]
        _lib = L
    return _lib


def max_threads() -> int:
    return int(lib().oracle_max_threads())


def _ptr(a: np.ndarray | None) -> int | None:
    return None if a is None else a.ctypes.data


def _check_f32(x: np.ndarray, name: str) -> np.ndarray:
    if x.dtype != np.float32:
        raise TypeError(f"{name} must be float32")
    return np.ascontiguousarray(x)


# ---------------------------------------------------------------------------
# GEMM.  All matrices are flat column-major buffers with explicit leading
# dimensions, exactly as the device API takes them.  B is passed as float32
# holding the FP16 values (the reference oracle convention).
# ---------------------------------------------------------------------------
def gemm_f64_rows(op_a, op_b, op_c, m, n, k, alpha, A, lda, B, ldb, beta, C0, ldc, r0, r1, out=None, threads=0):
    A = _check_f32(A, "A")
    B = _check_f32(B, "B")
    if C0 is not None:
        C0 = _check_f32(C0, "C0")
    size = (ldc * n) if op_c == OP_N else (ldc * m)
    if out is None:
        out = np.zeros(size, dtype=np.float64)
    lib().oracle_gemm_f64_rows(op_a, op_b, op_c, m, n, k, float(alpha), _ptr(A), lda, _ptr(B), ldb, float(beta),
                               _ptr(C0), ldc, _ptr(out), r0, r1, threads)
    return out


def gemm_f64(op_a, op_b, op_c, m, n, k, alpha, A, lda, B, ldb, beta, C0, ldc, threads=0):
    return gemm_f64_rows(op_a, op_b, op_c, m, n, k, alpha, A, lda, B, ldb, beta, C0, ldc, 0, m, threads=threads)


def gemm_f32_rows(op_a, op_b, op_c, m, n, k, alpha, A, lda, B, ldb, beta, C0, ldc, r0, r1, out=None, threads=0):
    A = _check_f32(A, "A")
    B = _check_f32(B, "B")
    if C0 is not None:
        C0 = _check_f32(C0, "C0")
    size = (ldc * n) if op_c == OP_N else (ldc * m)
    if out is None:
        out = np.zeros(size, dtype=np.float32)
    lib().oracle_gemm_f32_rows(op_a, op_b, op_c, m, n, k, float(alpha), _ptr(A), lda, _ptr(B), ldb, float(beta),
                               _ptr(C0), ldc, _ptr(out), r0, r1, threads)
    return out


def sgemm_nn_packed(m, n, k, alpha, A, lda, B, ldb, beta, C0, ldc, r0=0, r1=None, out=None, threads=0):
    """Host FP32 GEMM, op_n/op_n/op_n, rows [r0, r1) of C: packed panels and a
    16 x 6 AVX2 FMA register tile (the timed CPU baseline).  Each output is a
    round-to-nearest FP32 FMA chain over k in order."""
    A = _check_f32(A, "A")
    B = _check_f32(B, "B")
    r1 = m if r1 is None else r1
    if out is None:
        out = np.zeros(ldc * n, dtype=np.float32)
    if C0 is not None and beta != 0.0:
        C0 = _check_f32(C0, "C0")
    lib().oracle_sgemm_nn_packed(m, n, k, float(alpha), _ptr(A), lda, _ptr(B), ldb, float(beta),
                                 _ptr(C0) if (C0 is not None and beta != 0.0) else None, ldc, _ptr(out), r0, r1,
                                 threads)
    return out


def gemm_f32(op_a, op_b, op_c, m, n, k, alpha, A, lda, B, ldb, beta, C0, ldc, threads=0):
    return gemm_f32_rows(op_a, op_b, op_c, m, n, k, alpha, A, lda, B, ldb, beta, C0, ldc, 0, m, threads=threads)


# ---------------------------------------------------------------------------
# Splits (C restatement through ctypes; numpy versions for cross-checks)
# ---------------------------------------------------------------------------
def split_f16(a: np.ndarray) -> tuple[np.ndarray, np.ndarray]:
    """hi, lo as uint16 bit patterns of IEEE binary16."""
    a = _check_f32(a.ravel(), "a")
    hi = np.empty(a.size, np.uint16)
    lo = np.empty(a.size, np.uint16)
    lib().oracle_split_f16(_ptr(a), _ptr(hi), _ptr(lo), a.size)
    return hi, lo


def row_exponents(Am: np.ndarray) -> np.ndarray:
    """Per-row exponent e(row) of the prescaled fp16 split for an m x k matrix."""
    Am = np.ascontiguousarray(Am, dtype=np.float32)
    out = np.empty(Am.shape[0], np.int64)
    for r in range(Am.shape[0]):
        out[r] = lib().oracle_row_exponent(Am[r].ctypes.data, Am.shape[1], 1)
    return out


def split_f16_same_scale(x: np.ndarray):
    x = _check_f32(x.ravel(), "x")
    hi = np.empty(x.size, np.uint16)
    lo = np.empty(x.size, np.uint16)
    lib().oracle_split_f16_same_scale(_ptr(x), _ptr(hi), _ptr(lo), x.size)
    return hi, lo


def split_value_f16_rowscaled(Am: np.ndarray) -> np.ndarray:
    """FP64 value of the prescaled split (hi + lo) * 2^e(row) of an m x k matrix."""
    e = row_exponents(Am)
    xs = np.ldexp(Am.astype(np.float32), -e[:, None]).astype(np.float32)
    hi, lo = split_f16_same_scale(xs)
    v = f16_bits_to_f32(hi).astype(np.float64) + f16_bits_to_f32(lo).astype(np.float64)
    return np.ldexp(v.reshape(Am.shape), e[:, None])


def running_row_exponents(Am: np.ndarray, ktile: int, kslice: int | None = None) -> np.ndarray:
    """Running row exponents of the 64-row-wavefront kernels (the prescaled fp16 P0,
    P1, split-K and grouped tiles): an m x T array, T = number of ktile-wide
    k-tiles, whose entry (r, t) is the exponent k-tile t of row r is split
    with: max(previous entry, ilogb(max |tile|) - 14), clamped below at -126,
    an all-zero tile contributing -126 and a non-finite maximum 0 (as the
    whole-row row_exponents does).  Each k-slice of a split-K product (kslice
    columns, a multiple of ktile) starts again from -126."""
    Am = np.asarray(Am, dtype=np.float32)
    m, k = Am.shape
    span = kslice if kslice else k
    tiles = (k + ktile - 1) // ktile
    out = np.empty((m, tiles), np.int64)
    cur = np.full(m, -126, np.int64)
    for t in range(tiles):
        if (t * ktile) % span == 0:
            cur[:] = -126
        mx = np.abs(Am[:, t * ktile:(t + 1) * ktile]).max(axis=1).astype(np.float64)
        _, ex = np.frexp(mx)                      # mx = f * 2^ex, f in [0.5, 1)
        et = np.maximum(ex.astype(np.int64) - 1 - 14, -126)
        et = np.where(mx == 0.0, -126, np.where(np.isfinite(mx), et, 0))
        cur = np.maximum(cur, et)
        out[:, t] = cur
    return out


def split_value_f16_running(Am: np.ndarray, ktile: int, kslice: int | None = None) -> np.ndarray:
    """FP64 value of the running-exponent split (hi + lo) * 2^e(r, t) of an
    m x k matrix (e from running_row_exponents)."""
    Am = np.asarray(Am, dtype=np.float32)
    e = np.repeat(running_row_exponents(Am, ktile, kslice), ktile, axis=1)[:, :Am.shape[1]]
    xs = np.ldexp(Am, -e).astype(np.float32)
    hi, lo = split_f16_same_scale(xs)
    v = f16_bits_to_f32(hi).astype(np.float64) + f16_bits_to_f32(lo).astype(np.float64)
    return np.ldexp(v.reshape(Am.shape), e)


def split_bf16x3(a: np.ndarray):
    a = _check_f32(a.ravel(), "a")
    p = [np.empty(a.size, np.uint16) for _ in range(3)]
    lib().oracle_split_bf16x3(_ptr(a), _ptr(p[0]), _ptr(p[1]), _ptr(p[2]), a.size)
    return tuple(p)


def split_bf16x2(b16_bits: np.ndarray):
    b = np.ascontiguousarray(b16_bits.ravel().astype(np.uint16, copy=False))
    p = [np.empty(b.size, np.uint16) for _ in range(2)]
    lib().oracle_split_bf16x2(_ptr(b), _ptr(p[0]), _ptr(p[1]), b.size)
    return tuple(p)


def f16_bits_to_f32(bits: np.ndarray) -> np.ndarray:
    return bits.view(np.float16).astype(np.float32)


def bf16_bits_to_f32(bits: np.ndarray) -> np.ndarray:
    return (bits.astype(np.uint32) << 16).view(np.float32)


def split_value_f16(a: np.ndarray) -> np.ndarray:
    """FP64 value of the FP16 split representation hi + 2^-11 lo."""
    hi, lo = split_f16(a)
    return f16_bits_to_f32(hi).astype(np.float64) + np.ldexp(f16_bits_to_f32(lo).astype(np.float64), -11)


def emulate_shgemm(op_a, op_b, op_c, m, n, k, alpha, A, lda, B, ldb, beta, C0, ldc, compute_type=0,
                   ktile=32, kslice=None):
    """FP64 product of the operands exactly as the GPU kernel represents them.

    compute_type 0 (fp16, prescaling off): A -> hi + 2^-11 lo (FP16 each).
    compute_type 1 (tf32): A -> a0 + a1 + a2, B -> b0 + b1 (BF16 each); the
    kernel forms a0b0 + a0b1 + a1b0 + a1b1 + a2b0 (a2b1 is dropped).
    compute_type 2 (fp16, prescaled, the 256x256 / ring / three-stage /
    pre-split kernels): row r of op(A) -> (hi + lo) * 2^e(r).
    compute_type 3 (fp16, prescaled, the P0 / P1 / split-K kernels): k-tile t of row r
    -> (hi + lo) * 2^e(r, t), running exponents over ktile-wide k-tiles,
    restarting at every k-slice of kslice columns.
    The GPU result differs from this only by FP32 accumulation rounding.
    """
    A = _check_f32(A, "A")
    B = _check_f32(B, "B")
    if compute_type in (2, 3):
        Am = np.array(op_view(A, op_a, m, k, lda))
        Ae = split_value_f16_rowscaled(Am) if compute_type == 2 else split_value_f16_running(Am, ktile, kslice)
        Bm = op_view(B.astype(np.float64), op_b, k, n, ldb)
        out = np.zeros((ldc * n) if op_c == OP_N else (ldc * m), dtype=np.float64)
        view = _as_c_view(out, op_c, m, n, ldc)
        view[...] = alpha * (Ae @ Bm)
        if beta != 0.0 and C0 is not None:
            view += beta * _as_c_view(C0.astype(np.float64), op_c, m, n, ldc)
        return out
    if compute_type == 0:
        Ae = split_value_f16(A).astype(np.float64)
        Be = B.astype(np.float64)
        return _f64_gemm_np(op_a, op_b, op_c, m, n, k, alpha, Ae, lda, Be, ldb, beta, C0, ldc)
    a0, a1, a2 = (bf16_bits_to_f32(x).astype(np.float64) for x in split_bf16x3(A))
    b16 = B.astype(np.float16).view(np.uint16)
    b0, b1 = (bf16_bits_to_f32(x).astype(np.float64) for x in split_bf16x2(b16))
    terms = [(a0, b0), (a0, b1), (a1, b0), (a1, b1), (a2, b0)]
    out = None
    for ai, bi in terms:
        t = _f64_gemm_np(op_a, op_b, op_c, m, n, k, alpha, ai, lda, bi, ldb, 0.0, None, ldc)
        out = t if out is None else out + t
    if beta != 0.0:
        out = out + beta * _as_c_view(C0.astype(np.float64), op_c, m, n, ldc, full=True)
    return out


def _mat(buf, rows, cols, ld):
    """Column-major rows x cols view of a flat buffer with leading dimension ld."""
    return np.lib.stride_tricks.as_strided(buf, shape=(rows, cols), strides=(buf.itemsize, buf.itemsize * ld))


def op_view(buf, op, rows, cols, ld):
    """op(X) as a (rows x cols) array for a stored operand with leading dim ld."""
    if op == OP_N:
        return _mat(buf, rows, cols, ld)
    return _mat(buf, cols, rows, ld).T


def _as_c_view(buf, op_c, m, n, ldc, full=False):
    if op_c == OP_N:
        return _mat(buf, m, n, ldc)
    return _mat(buf, n, m, ldc).T


def _f64_gemm_np(op_a, op_b, op_c, m, n, k, alpha, A64, lda, B64, ldb, beta, C0, ldc):
    """numpy FP64 product in C's storage layout (flat, padding rows zero)."""
    Am = op_view(A64, op_a, m, k, lda)
    Bm = op_view(B64, op_b, k, n, ldb)
    prod = alpha * (Am @ Bm)
    size = (ldc * n) if op_c == OP_N else (ldc * m)
    out = np.zeros(size, dtype=np.float64)
    view = _as_c_view(out, op_c, m, n, ldc)
    view[...] = prod
    if beta != 0.0 and C0 is not None:
        view += beta * _as_c_view(C0.astype(np.float64), op_c, m, n, ldc)
    return out


def c_matrix(buf, op_c, m, n, ldc) -> np.ndarray:
    """The m x n matrix C held in a flat buffer (copy)."""
    return np.array(_as_c_view(buf, op_c, m, n, ldc))


def rel_fro_error(C, C64) -> float:
    C = np.asarray(C, dtype=np.float64)
    C64 = np.asarray(C64, dtype=np.float64)
    den = np.linalg.norm(C64)
    num = np.linalg.norm(C - C64)
    return float(num / den) if den != 0 else float(num)


# ---------------------------------------------------------------------------
# Seeded generators (numpy PCG64); the same streams feed tests and bench.
# ---------------------------------------------------------------------------
def gen_growing_rows(rng: np.random.Generator, m, n, k, step=3, width=32, b0=12):
    """A (m x k) and B (k x n, FP16 values), column-major, whose products are
    balanced while A's rows grow along k: odd rows of A and the k-rows of B
    carry 2^(step*j) and 2^(b0 - step*j) in column block j (width columns;
    B stays an FP16 normal for step*j <= b0 + 14), even rows of A are
    U[-1, 1]; one row of A is zero.  A split with each row's final exponent
    loses the early blocks' low bits; running exponents keep them."""
    j = np.arange(k) // width
    A = rng.uniform(1.0, 2.0, (m, k)) * np.where(rng.random((m, k)) < 0.5, -1.0, 1.0)
    A[1::2] *= np.exp2(step * j)[None, :]
    A[0::2] = rng.uniform(-1.0, 1.0, (A[0::2].shape))
    A[m // 2] = 0.0
    B = (rng.standard_normal((k, n)) * np.exp2(b0 - step * j)[:, None]).astype(np.float16)
    return (np.asfortranarray(A.astype(np.float32)).ravel(order="F"),
            np.asfortranarray(B).ravel(order="F"))


def gen_uniform(rng: np.random.Generator, size, lo=-1.0, hi=1.0) -> np.ndarray:
    return rng.uniform(lo, hi, size).astype(np.float32)


def gen_f16_uniform(rng, size, lo=-1.0, hi=1.0) -> np.ndarray:
    """FP16 values (as float16) drawn uniformly and rounded to nearest even."""
    return rng.uniform(lo, hi, size).astype(np.float32).astype(np.float16)


def gen_f16_normal(rng, size, sigma=1.0) -> np.ndarray:
    return (rng.standard_normal(size) * sigma).astype(np.float32).astype(np.float16)


def gen_wide_exponent(rng, size, emin=-15, emax=15) -> np.ndarray:
    """sign * U[1,2) * 2^e with integer e uniform in [emin, emax]."""
    sign = np.where(rng.random(size) < 0.5, -1.0, 1.0)
    mant = rng.uniform(1.0, 2.0, size)
    e = rng.integers(emin, emax + 1, size)
    return (sign * np.ldexp(mant, e)).astype(np.float32)


def philox4x32_10(ctr, key) -> list[int]:
    """Philox4x32-10 of one counter (4 words) and key (2 words)."""
    c = np.asarray(ctr, dtype=np.uint32)
    k = np.asarray(key, dtype=np.uint32)
    out = np.zeros(4, dtype=np.uint32)
    lib().oracle_philox4x32_10(_ptr(c), _ptr(k), _ptr(out))
    return [int(x) for x in out]


def philox_words(blocks: int, seed: int) -> np.ndarray:
    """Words of blocks 0..blocks-1 (counter (b, 0, 0) split in 32-bit halves, key = seed)."""
    out = np.zeros(4 * blocks, dtype=np.uint32)
    lib().oracle_philox_words(_ptr(out), blocks, seed)
    return out


def gaussian(count: int, seed: int, scale: float = 1.0) -> np.ndarray:
    """FP64 Box-Muller normals of the sketch generator, times scale."""
    out = np.zeros(count, dtype=np.float64)
    lib().oracle_gaussian(_ptr(out), count, seed, scale)
    return out
