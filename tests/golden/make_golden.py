"""Writes the golden fixtures under tests/golden/.

The reference repository holds no golden vectors, known-answer tests or
fixtures for this path (its tests print metrics only: test/shgemm_test.cu,
test/shgemm_k_slicing_test.cu, test/exponent_test.cu), and it cannot be built
here (CUDA only, submodules absent).  These fixtures are therefore computed
with numpy alone, independently of the C oracle under oracle/:

  split_f16.npz   inputs a (float32) incl. ties, subnormal and overflow edges,
                  with hi = float16(a) and lo = float16((a - hi) * 2^11) from
                  numpy's IEEE binary16 conversion (round to nearest even),
                  the split of reference src/wmmae_shgemm.hpp:29-31.
  gemm_int.npz    integer-valued GEMMs (|a|,|b| <= 32, k <= 1024) for all op
                  combinations: every partial sum is an exact integer below
                  2^24, so the FP16-split product is exact and C is known
                  bit-for-bit (lo = 0 because integers <= 2048 are FP16-exact).
  gemm_rand.npz   small random GEMMs (all op_a/op_b/op_c, odd sizes, padded
                  leading dimensions, alpha/beta) with the FP64 result computed
                  by numpy float64 matmul.

Run:  python tests/golden/make_golden.py
"""
from __future__ import annotations

from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent


def _mat(buf, rows, cols, ld):
    return np.lib.stride_tricks.as_strided(buf, shape=(rows, cols), strides=(buf.itemsize, buf.itemsize * ld))


def _op(buf, op, rows, cols, ld):
    return _mat(buf, rows, cols, ld) if op == 0 else _mat(buf, cols, rows, ld).T


def split_cases() -> dict:
    rng = np.random.default_rng(20240611)
    specials = [0.0, -0.0, 1.0, -1.0, 65504.0, 65519.99, 65520.0, 1e5, -1e5, 2.0 ** -14, 2.0 ** -24, 2.0 ** -25,
                2.0 ** -26, 1.5 * 2.0 ** -25, 3e-8, 6.1e-5, 1.0 + 2.0 ** -11, 1.0 + 3 * 2.0 ** -11,
                1.0 + 2.0 ** -12, 2048.5, 4097.0, np.pi, -np.e, 1.0 / 3.0, float(np.float32(0.1))]
    ties = [(1.0 + (2 * i + 1) * 2.0 ** -11) * s for i in range(8) for s in (1.0, -1.0)]
    rand_uniform = rng.uniform(-1, 1, 2000)
    rand_wide = np.ldexp(rng.uniform(1, 2, 2000) * np.where(rng.random(2000) < 0.5, -1, 1),
                         rng.integers(-30, 16, 2000))
    a = np.array(specials + ties, dtype=np.float32)
    a = np.concatenate([a, rand_uniform.astype(np.float32), rand_wide.astype(np.float32)])
    with np.errstate(invalid="ignore", over="ignore"):
        hi = a.astype(np.float16)
        r = a - hi.astype(np.float32)
        lo = (r * np.float32(2048.0)).astype(np.float16)
    return {"a": a, "hi_bits": hi.view(np.uint16), "lo_bits": lo.view(np.uint16)}


def _case(rng, op_a, op_b, op_c, m, n, k, pad, alpha, beta, integer):
    lda = (m if op_a == 0 else k) + pad
    ldb = (k if op_b == 0 else n) + pad
    ldc = (m if op_c == 0 else n) + pad
    a_cols = k if op_a == 0 else m
    b_cols = n if op_b == 0 else k
    c_cols = n if op_c == 0 else m
    if integer:
        A = rng.integers(-32, 33, lda * a_cols).astype(np.float32)
        B = rng.integers(-32, 33, ldb * b_cols).astype(np.float16)
        C0 = rng.integers(-64, 65, ldc * c_cols).astype(np.float32)
    else:
        A = rng.uniform(-1, 1, lda * a_cols).astype(np.float32)
        B = rng.uniform(-1, 1, ldb * b_cols).astype(np.float32).astype(np.float16)
        C0 = rng.uniform(-1, 1, ldc * c_cols).astype(np.float32)
    Am = _op(A.astype(np.float64), op_a, m, k, lda)
    Bm = _op(B.astype(np.float64), op_b, k, n, ldb)
    out = C0.astype(np.float64).copy()
    Cv = _mat(out, m, n, ldc) if op_c == 0 else _mat(out, n, m, ldc).T
    Cv[...] = alpha * (Am @ Bm) + beta * Cv
    return dict(op_a=op_a, op_b=op_b, op_c=op_c, m=m, n=n, k=k, lda=lda, ldb=ldb, ldc=ldc, alpha=alpha, beta=beta,
                A=A, B_bits=B.view(np.uint16), C0=C0, C64=out)


def gemm_cases(integer: bool) -> dict:
    rng = np.random.default_rng(7 if integer else 11)
    cases = []
    shapes = [(16, 16, 16), (37, 53, 91), (70, 33, 130)] if not integer else [(32, 48, 1024), (33, 20, 129)]
    for (m, n, k) in shapes:
        for op_a in (0, 1):
            for op_b in (0, 1):
                for op_c in (0, 1):
                    pad = (m + n + k + op_a + 2 * op_b + 4 * op_c) % 5
                    if integer:
                        alpha, beta = 1.0, (0.0 if (op_a + op_b) % 2 == 0 else 2.0)
                    else:
                        alpha, beta = (1.0, 0.0) if op_c == 0 else (-0.75, 0.5)
                    cases.append(_case(rng, op_a, op_b, op_c, m, n, k, pad, alpha, beta, integer))
    out = {}
    for i, c in enumerate(cases):
        for key, v in c.items():
            out[f"c{i}_{key}"] = np.asarray(v)
    out["count"] = np.asarray(len(cases))
    return out


def load_cases(path: Path) -> list[dict]:
    z = np.load(path, allow_pickle=False)
    cases = []
    for i in range(int(z["count"])):
        c = {}
        for key in ("op_a", "op_b", "op_c", "m", "n", "k", "lda", "ldb", "ldc"):
            c[key] = int(z[f"c{i}_{key}"])
        for key in ("alpha", "beta"):
            c[key] = float(z[f"c{i}_{key}"])
        for key in ("A", "B_bits", "C0", "C64"):
            c[key] = z[f"c{i}_{key}"]
        cases.append(c)
    return cases


def main():
    np.savez_compressed(HERE / "split_f16.npz", **split_cases())
    np.savez_compressed(HERE / "gemm_int.npz", **gemm_cases(True))
    np.savez_compressed(HERE / "gemm_rand.npz", **gemm_cases(False))
    for f in sorted(HERE.glob("*.npz")):
        print(f.name, f.stat().st_size)


if __name__ == "__main__":
    main()
