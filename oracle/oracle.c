/* Test infrastructure only: the CPU oracle for the SHGEMM hot path.
 *
 * Nothing in the product links, loads or calls this file.  Tests, the smoke
 * check and the CPU-baseline leg of the benchmark use it as the checker.
 *
 *   oracle_gemm_f64  C64 = alpha * op(A) * op(B) + beta * C0, FP64 accumulation
 *                    of FP32 inputs (B given as float of its FP16 value, as the
 *                    reference's tests do: test/shgemm_test.cu:214-224, 91-100).
 *                    Replaces mateval::get_error_AxB's reference product
 *                    (un-vendored submodule test/mateval, .gitmodules:7-9).
 *   oracle_gemm_f32  the same product with FP32 accumulation (any op layout).
 *   oracle_sgemm_nn_packed  the host FP32 GEMM timed as the CPU baseline
 *                    (op_n / op_n / op_n): packed panels and a 16 x 6 AVX2 FMA
 *                    register tile, the standard blocking of high-performance
 *                    CPU GEMM (Goto & van de Geijn, ACM TOMS 34(3), 2008).
 *   oracle_split_f16 hi = fp16(a), lo = fp16((a - hi) * 2^11), both
 *                    round-to-nearest-even (reference src/wmmae_shgemm.hpp:29-31;
 *                    method of Ootomo & Yokota, IJHPCA 2022).
 *   oracle_split_f16_same_scale / oracle_row_exponent  the prescaled fp16
 *                    split: each row of A scaled by 2^-e(row), e = ilogb(max|row|)
 *                    - 14, then hi = fp16(x), lo = fp16(x - hi).
 *   oracle_split_bf16x3 / oracle_split_bf16x2  the BF16 parts of the tf32
 *                    compute type.
 *   oracle_philox_words / oracle_gaussian  the sketch matrix generator:
 *                    Philox4x32-10 (Salmon et al., SC'11) words per counter
 *                    block, Box-Muller normals in FP64 from the same words
 *                    (the GPU computes them in FP32 and rounds to FP16).
 *
 * Conversions are written from the IEEE 754 binary16/bfloat16 definitions
 * (round to nearest, ties to even; overflow to infinity; gradual underflow).
 * Layout: column-major; op_t means the operand is stored transposed.
 */
#include <immintrin.h>
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ---------------------------------------------------------------- */
/* binary32 -> binary16, round to nearest even                       */
/* ---------------------------------------------------------------- */
uint16_t oracle_f32_to_f16(float f) {
  uint32_t x;
  memcpy(&x, &f, 4);
  const uint32_t sign = (x >> 16) & 0x8000u;
  const uint32_t ex = (x >> 23) & 0xffu;
  uint32_t man = x & 0x7fffffu;
  if (ex == 0xffu) { /* inf / nan */
    if (man == 0) return (uint16_t)(sign | 0x7c00u);
    return (uint16_t)(sign | 0x7e00u | (man >> 13));
  }
  const int e = (int)ex - 127; /* unbiased */
  if (e > 15) return (uint16_t)(sign | 0x7c00u);
  if (e >= -14) {
    /* normal half: 10 mantissa bits kept, 13 dropped */
    uint32_t h = ((uint32_t)(e + 15) << 10) | (man >> 13);
    const uint32_t rest = man & 0x1fffu;
    if (rest > 0x1000u || (rest == 0x1000u && (h & 1u))) h += 1; /* carries into exponent/inf correctly */
    return (uint16_t)(sign | h);
  }
  /* subnormal half (or zero): value = m * 2^-24 */
  if (ex == 0) return (uint16_t)sign; /* float subnormals are far below half's range */
  man |= 0x800000u;                  /* implicit bit; value = man * 2^(e-23) */
  const int shift = -e - 1;          /* bits to drop: man * 2^(e-23) / 2^-24 = man >> (-(e+1)) */
  if (shift > 24) return (uint16_t)sign;
  uint32_t h = man >> shift;
  const uint32_t rest = man & ((1u << shift) - 1u);
  const uint32_t half = 1u << (shift - 1);
  if (rest > half || (rest == half && (h & 1u))) h += 1;
  return (uint16_t)(sign | h);
}

float oracle_f16_to_f32(uint16_t h) {
  const uint32_t sign = (uint32_t)(h & 0x8000u) << 16;
  const uint32_t ex = (h >> 10) & 0x1fu;
  const uint32_t man = h & 0x3ffu;
  float f;
  if (ex == 0x1fu) {
    const uint32_t x = sign | 0x7f800000u | (man << 13);
    memcpy(&f, &x, 4);
    return f;
  }
  if (ex == 0) {
    f = ldexpf((float)man, -24);
    return sign ? -f : f;
  }
  const uint32_t x = sign | ((ex + 112u) << 23) | (man << 13);
  memcpy(&f, &x, 4);
  return f;
}

/* binary32 -> bfloat16, round to nearest even */
uint16_t oracle_f32_to_bf16(float f) {
  uint32_t x;
  memcpy(&x, &f, 4);
  if ((x & 0x7f800000u) == 0x7f800000u && (x & 0x7fffffu)) return (uint16_t)((x >> 16) | 0x40u); /* quiet nan */
  const uint32_t rest = x & 0xffffu;
  uint32_t hi = x >> 16;
  if (rest > 0x8000u || (rest == 0x8000u && (hi & 1u))) hi += 1;
  return (uint16_t)hi;
}

float oracle_bf16_to_f32(uint16_t b) {
  const uint32_t x = (uint32_t)b << 16;
  float f;
  memcpy(&f, &x, 4);
  return f;
}

void oracle_split_f16(const float* a, uint16_t* hi, uint16_t* lo, int64_t n) {
  for (int64_t i = 0; i < n; ++i) {
    const uint16_t h = oracle_f32_to_f16(a[i]);
    const float r = a[i] - oracle_f16_to_f32(h); /* exact */
    hi[i] = h;
    lo[i] = oracle_f32_to_f16(r * 2048.0f);
  }
}

/* Prescaled variant (row already multiplied by 2^-e): hi = fp16(x),
 * lo = fp16(x - hi), both at the same scale. */
void oracle_split_f16_same_scale(const float* x, uint16_t* hi, uint16_t* lo, int64_t n) {
  for (int64_t i = 0; i < n; ++i) {
    const uint16_t h = oracle_f32_to_f16(x[i]);
    hi[i] = h;
    lo[i] = oracle_f32_to_f16(x[i] - oracle_f16_to_f32(h));
  }
}

/* Row exponent of the prescaled fp16 split: ilogb(max |row|) - 14, clamped
 * to >= -126; 0 for an all-zero or non-finite row. */
int oracle_row_exponent(const float* row, int64_t n, int64_t stride) {
  float mx = 0.0f;
  for (int64_t i = 0; i < n; ++i) {
    const float v = fabsf(row[i * stride]);
    if (!(v <= mx)) mx = v; /* also picks up NaN */
  }
  if (mx == 0.0f || !isfinite(mx)) return 0;
  const int e = ilogbf(mx) - 14;
  return e < -126 ? -126 : e;
}

void oracle_split_bf16x3(const float* a, uint16_t* p0, uint16_t* p1, uint16_t* p2, int64_t n) {
  for (int64_t i = 0; i < n; ++i) {
    const uint16_t b0 = oracle_f32_to_bf16(a[i]);
    const float r1 = a[i] - oracle_bf16_to_f32(b0);
    const uint16_t b1 = oracle_f32_to_bf16(r1);
    const float r2 = r1 - oracle_bf16_to_f32(b1);
    p0[i] = b0;
    p1[i] = b1;
    p2[i] = oracle_f32_to_bf16(r2);
  }
}

void oracle_split_bf16x2(const uint16_t* b16, uint16_t* p0, uint16_t* p1, int64_t n) {
  for (int64_t i = 0; i < n; ++i) {
    const float f = oracle_f16_to_f32(b16[i]);
    const uint16_t b0 = oracle_f32_to_bf16(f);
    p0[i] = b0;
    p1[i] = oracle_f32_to_bf16(f - oracle_bf16_to_f32(b0));
  }
}

/* ---------------------------------------------------------------- */
/* GEMM                                                              */
/* ---------------------------------------------------------------- */
/* element (i, p) of op(A) (m x k) and (p, j) of op(B) (k x n) */
static inline float a_at(const float* A, int64_t lda, int op_a, int64_t i, int64_t p) {
  return op_a == 0 ? A[i + p * lda] : A[p + i * lda];
}
static inline float b_at(const float* B, int64_t ldb, int op_b, int64_t p, int64_t j) {
  return op_b == 0 ? B[p + j * ldb] : B[j + p * ldb];
}
static inline int64_t c_off(int op_c, int64_t i, int64_t j, int64_t ldc) {
  return op_c == 0 ? i + j * ldc : j + i * ldc;
}

#define KB 256
#define JB 256
#define IB 32

/* rows [r0, r1) of op(A) only: the row-sampled oracle for large shapes.
 * out (FP64) uses C's layout (op_c, ldc).  C0 may be NULL when beta == 0. */
void oracle_gemm_f64_rows(int op_a, int op_b, int op_c, int64_t m, int64_t n, int64_t k, double alpha,
                          const float* A, int64_t lda, const float* B, int64_t ldb, double beta, const float* C0,
                          int64_t ldc, double* out, int64_t r0, int64_t r1, int nthreads) {
  (void)m;
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#else
  (void)nthreads;
#endif
  const int64_t nrows = r1 - r0;
  const int64_t nib = (nrows + IB - 1) / IB;
#pragma omp parallel
  {
    double* acc = (double*)malloc(sizeof(double) * IB * JB);
    float* bp = (float*)malloc(sizeof(float) * KB * JB);
    float* ap = (float*)malloc(sizeof(float) * IB * KB);
#pragma omp for schedule(dynamic) collapse(2)
    for (int64_t jb = 0; jb < (n + JB - 1) / JB; ++jb) {
      for (int64_t ib = 0; ib < nib; ++ib) {
        const int64_t j0 = jb * JB, j1 = (j0 + JB < n) ? j0 + JB : n, nj = j1 - j0;
        const int64_t i0 = r0 + ib * IB, i1 = (i0 + IB < r1) ? i0 + IB : r1, ni = i1 - i0;
        memset(acc, 0, sizeof(double) * IB * JB);
        for (int64_t p0 = 0; p0 < k; p0 += KB) {
          const int64_t p1 = (p0 + KB < k) ? p0 + KB : k, np = p1 - p0;
          for (int64_t p = 0; p < np; ++p)
            for (int64_t j = 0; j < nj; ++j) bp[p * JB + j] = b_at(B, ldb, op_b, p0 + p, j0 + j);
          for (int64_t i = 0; i < ni; ++i)
            for (int64_t p = 0; p < np; ++p) ap[i * KB + p] = a_at(A, lda, op_a, i0 + i, p0 + p);
          for (int64_t i = 0; i < ni; ++i) {
            double* ci = acc + i * JB;
            for (int64_t p = 0; p < np; ++p) {
              const double a = (double)ap[i * KB + p];
              const float* bj = bp + p * JB;
              for (int64_t j = 0; j < nj; ++j) ci[j] += a * (double)bj[j];
            }
          }
        }
        for (int64_t i = 0; i < ni; ++i)
          for (int64_t j = 0; j < nj; ++j) {
            const int64_t o = c_off(op_c, i0 + i, j0 + j, ldc);
            double v = alpha * acc[i * JB + j];
            if (beta != 0.0) v += beta * (double)C0[o];
            out[o] = v;
          }
      }
    }
    free(acc);
    free(bp);
    free(ap);
  }
}

void oracle_gemm_f64(int op_a, int op_b, int op_c, int64_t m, int64_t n, int64_t k, double alpha, const float* A,
                     int64_t lda, const float* B, int64_t ldb, double beta, const float* C0, int64_t ldc,
                     double* out, int nthreads) {
  oracle_gemm_f64_rows(op_a, op_b, op_c, m, n, k, alpha, A, lda, B, ldb, beta, C0, ldc, out, 0, m, nthreads);
}

/* Host FP32 GEMM: same blocking, FP32 accumulation (each multiply-add rounded
 * to FP32; the compiler fuses them into FMAs), rows [r0, r1). */
void oracle_gemm_f32_rows(int op_a, int op_b, int op_c, int64_t m, int64_t n, int64_t k, float alpha,
                          const float* A, int64_t lda, const float* B, int64_t ldb, float beta, const float* C0,
                          int64_t ldc, float* out, int64_t r0, int64_t r1, int nthreads) {
  (void)m;
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#else
  (void)nthreads;
#endif
  const int64_t nrows = r1 - r0;
  const int64_t nib = (nrows + IB - 1) / IB;
#pragma omp parallel
  {
    float* acc = (float*)malloc(sizeof(float) * IB * JB);
    float* bp = (float*)malloc(sizeof(float) * KB * JB);
    float* ap = (float*)malloc(sizeof(float) * IB * KB);
#pragma omp for schedule(dynamic) collapse(2)
    for (int64_t jb = 0; jb < (n + JB - 1) / JB; ++jb) {
      for (int64_t ib = 0; ib < nib; ++ib) {
        const int64_t j0 = jb * JB, j1 = (j0 + JB < n) ? j0 + JB : n, nj = j1 - j0;
        const int64_t i0 = r0 + ib * IB, i1 = (i0 + IB < r1) ? i0 + IB : r1, ni = i1 - i0;
        memset(acc, 0, sizeof(float) * IB * JB);
        for (int64_t p0 = 0; p0 < k; p0 += KB) {
          const int64_t p1 = (p0 + KB < k) ? p0 + KB : k, np = p1 - p0;
          for (int64_t p = 0; p < np; ++p)
            for (int64_t j = 0; j < nj; ++j) bp[p * JB + j] = b_at(B, ldb, op_b, p0 + p, j0 + j);
          for (int64_t i = 0; i < ni; ++i)
            for (int64_t p = 0; p < np; ++p) ap[i * KB + p] = a_at(A, lda, op_a, i0 + i, p0 + p);
          for (int64_t i = 0; i < ni; ++i) {
            float* ci = acc + i * JB;
            for (int64_t p = 0; p < np; ++p) {
              const float a = ap[i * KB + p];
              const float* bj = bp + p * JB;
              for (int64_t j = 0; j < nj; ++j) ci[j] += a * bj[j];
            }
          }
        }
        for (int64_t i = 0; i < ni; ++i)
          for (int64_t j = 0; j < nj; ++j) {
            const int64_t o = c_off(op_c, i0 + i, j0 + j, ldc);
            float v = alpha * acc[i * JB + j];
            if (beta != 0.0f) v += beta * C0[o];
            out[o] = v;
          }
      }
    }
    free(acc);
    free(bp);
    free(ap);
  }
}

void oracle_gemm_f32(int op_a, int op_b, int op_c, int64_t m, int64_t n, int64_t k, float alpha, const float* A,
                     int64_t lda, const float* B, int64_t ldb, float beta, const float* C0, int64_t ldc, float* out,
                     int nthreads) {
  oracle_gemm_f32_rows(op_a, op_b, op_c, m, n, k, alpha, A, lda, B, ldb, beta, C0, ldc, out, 0, m, nthreads);
}

/* ---------------------------------------------------------------- */
/* Packed FP32 GEMM, op_n / op_n / op_n, rows [r0, r1) of C           */
/* ---------------------------------------------------------------- */
enum { PK_MR = 16, PK_NR = 6, PK_MC = 192, PK_KC = 256, PK_NC = 4032 };

/* acc(16 x 6) = sum over p < kc of ap[16p .. 16p+15] x bp[6p .. 6p+5]; then
 * C = alpha * acc + beta_eff * C on the mr x nr corner that exists. */
static void pk_tile(int64_t kc, const float* ap, const float* bp, float* c, int64_t ldc, int mr, int nr, float alpha,
                    float beta_eff) {
  /* twelve named accumulators so that they stay in registers */
  __m256 c00 = _mm256_setzero_ps(), c01 = c00, c02 = c00, c03 = c00, c04 = c00, c05 = c00;
  __m256 c10 = c00, c11 = c00, c12 = c00, c13 = c00, c14 = c00, c15 = c00;
  for (int64_t p = 0; p < kc; ++p) {
    const __m256 a0 = _mm256_load_ps(ap + PK_MR * p);
    const __m256 a1 = _mm256_load_ps(ap + PK_MR * p + 8);
    const float* b = bp + PK_NR * p;
    __m256 bj = _mm256_broadcast_ss(b + 0);
    c00 = _mm256_fmadd_ps(a0, bj, c00);
    c10 = _mm256_fmadd_ps(a1, bj, c10);
    bj = _mm256_broadcast_ss(b + 1);
    c01 = _mm256_fmadd_ps(a0, bj, c01);
    c11 = _mm256_fmadd_ps(a1, bj, c11);
    bj = _mm256_broadcast_ss(b + 2);
    c02 = _mm256_fmadd_ps(a0, bj, c02);
    c12 = _mm256_fmadd_ps(a1, bj, c12);
    bj = _mm256_broadcast_ss(b + 3);
    c03 = _mm256_fmadd_ps(a0, bj, c03);
    c13 = _mm256_fmadd_ps(a1, bj, c13);
    bj = _mm256_broadcast_ss(b + 4);
    c04 = _mm256_fmadd_ps(a0, bj, c04);
    c14 = _mm256_fmadd_ps(a1, bj, c14);
    bj = _mm256_broadcast_ss(b + 5);
    c05 = _mm256_fmadd_ps(a0, bj, c05);
    c15 = _mm256_fmadd_ps(a1, bj, c15);
  }
  float t[PK_NR][PK_MR] __attribute__((aligned(32)));
  _mm256_store_ps(t[0], c00);
  _mm256_store_ps(t[0] + 8, c10);
  _mm256_store_ps(t[1], c01);
  _mm256_store_ps(t[1] + 8, c11);
  _mm256_store_ps(t[2], c02);
  _mm256_store_ps(t[2] + 8, c12);
  _mm256_store_ps(t[3], c03);
  _mm256_store_ps(t[3] + 8, c13);
  _mm256_store_ps(t[4], c04);
  _mm256_store_ps(t[4] + 8, c14);
  _mm256_store_ps(t[5], c05);
  _mm256_store_ps(t[5] + 8, c15);
  for (int j = 0; j < nr; ++j) {
    float* cj = c + j * ldc;
    for (int i = 0; i < mr; ++i) cj[i] = alpha * t[j][i] + (beta_eff != 0.0f ? beta_eff * cj[i] : 0.0f);
  }
}

/* out = alpha * A * B + beta * C0 on rows [r0, r1) (out may be C0). */
void oracle_sgemm_nn_packed(int64_t m, int64_t n, int64_t k, float alpha, const float* A, int64_t lda,
                            const float* B, int64_t ldb, float beta, const float* C0, int64_t ldc, float* out,
                            int64_t r0, int64_t r1, int nthreads) {
  (void)m;
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#else
  (void)nthreads;
#endif
  if (r1 <= r0 || n <= 0) return;
  if (C0 != NULL && C0 != out && beta != 0.0f)
    for (int64_t j = 0; j < n; ++j) memcpy(out + j * ldc + r0, C0 + j * ldc + r0, sizeof(float) * (r1 - r0));
  if (k <= 0) {
    for (int64_t j = 0; j < n; ++j)
      for (int64_t i = r0; i < r1; ++i) out[i + j * ldc] = beta != 0.0f ? beta * out[i + j * ldc] : 0.0f;
    return;
  }
  float* bpk = (float*)aligned_alloc(64, sizeof(float) * PK_KC * PK_NC);
  for (int64_t jc = 0; jc < n; jc += PK_NC) {
    const int64_t nc = (n - jc < PK_NC) ? n - jc : PK_NC;
    const int64_t nsl = (nc + PK_NR - 1) / PK_NR;
    for (int64_t pc = 0; pc < k; pc += PK_KC) {
      const int64_t kc = (k - pc < PK_KC) ? k - pc : PK_KC;
      const float beta_eff = pc == 0 ? beta : 1.0f;
      /* B panel in slivers of 6 columns: bpk[s*6*kc + 6p + j] = B(pc + p, jc + 6s + j) */
#pragma omp parallel for schedule(static)
      for (int64_t sl = 0; sl < nsl; ++sl) {
        float* dst = bpk + sl * PK_NR * kc;
        for (int j = 0; j < PK_NR; ++j) {
          const int64_t col = jc + sl * PK_NR + j;
          if (col < n) {
            const float* src = B + col * ldb + pc;
            for (int64_t p = 0; p < kc; ++p) dst[PK_NR * p + j] = src[p];
          } else {
            for (int64_t p = 0; p < kc; ++p) dst[PK_NR * p + j] = 0.0f;
          }
        }
      }
#pragma omp parallel
      {
        float* apk = (float*)aligned_alloc(64, sizeof(float) * PK_MC * PK_KC);
#pragma omp for schedule(dynamic)
        for (int64_t ic = r0; ic < r1; ic += PK_MC) {
          const int64_t mc = (r1 - ic < PK_MC) ? r1 - ic : PK_MC;
          /* A block in slivers of 16 rows: apk[s*16*kc + 16p + i] = A(ic + 16s + i, pc + p) */
          for (int64_t s0 = 0; s0 < mc; s0 += PK_MR) {
            float* dst = apk + (s0 / PK_MR) * PK_MR * kc;
            const int64_t rows = (mc - s0 < PK_MR) ? mc - s0 : PK_MR;
            for (int64_t p = 0; p < kc; ++p) {
              const float* src = A + (pc + p) * lda + ic + s0;
              int64_t i = 0;
              for (; i < rows; ++i) dst[PK_MR * p + i] = src[i];
              for (; i < PK_MR; ++i) dst[PK_MR * p + i] = 0.0f;
            }
          }
          for (int64_t jr = 0; jr < nc; jr += PK_NR) {
            const int nr = (int)((nc - jr < PK_NR) ? nc - jr : PK_NR);
            for (int64_t ir = 0; ir < mc; ir += PK_MR) {
              const int mr = (int)((mc - ir < PK_MR) ? mc - ir : PK_MR);
              pk_tile(kc, apk + (ir / PK_MR) * PK_MR * kc, bpk + (jr / PK_NR) * PK_NR * kc,
                      out + (ic + ir) + (jc + jr) * ldc, ldc, mr, nr, pc == 0 ? alpha : alpha, beta_eff);
            }
          }
        }
        free(apk);
      }
    }
  }
  free(bpk);
}

int oracle_max_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

/* ---------------------------------------------------------------- */
/* Sketch generator: Philox4x32-10 + Box-Muller                      */
/* counter = (b mod 2^32, b / 2^32, 0, 0), key = (seed mod 2^32,     */
/* seed / 2^32); element 4b + j is normal j of block b.              */
/* ---------------------------------------------------------------- */
static void philox_block(uint64_t block, uint64_t seed, uint32_t w[4]) {
  const uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u, W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
  uint32_t c0 = (uint32_t)block, c1 = (uint32_t)(block >> 32), c2 = 0u, c3 = 0u;
  uint32_t k0 = (uint32_t)seed, k1 = (uint32_t)(seed >> 32);
  for (int r = 0; r < 10; ++r) {
    const uint64_t p0 = (uint64_t)M0 * c0, p1 = (uint64_t)M1 * c2;
    const uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
    const uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
    c1 = (uint32_t)p1;
    c3 = (uint32_t)p0;
    c0 = n0;
    c2 = n2;
    k0 += W0;
    k1 += W1;
  }
  w[0] = c0;
  w[1] = c1;
  w[2] = c2;
  w[3] = c3;
}

/* Philox4x32-10 with an explicit counter and key (known-answer tests). */
void oracle_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
  const uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u, W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
  uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3], k0 = key[0], k1 = key[1];
  for (int r = 0; r < 10; ++r) {
    const uint64_t p0 = (uint64_t)M0 * c0, p1 = (uint64_t)M1 * c2;
    const uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
    const uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
    c1 = (uint32_t)p1;
    c3 = (uint32_t)p0;
    c0 = n0;
    c2 = n2;
    k0 += W0;
    k1 += W1;
  }
  out[0] = c0;
  out[1] = c1;
  out[2] = c2;
  out[3] = c3;
}

void oracle_philox_words(uint32_t* out, int64_t blocks, uint64_t seed) {
  for (int64_t b = 0; b < blocks; ++b) philox_block((uint64_t)b, seed, out + 4 * b);
}

static double unit_open(uint32_t w) { return ((double)(w >> 9) + 0.5) * (1.0 / 8388608.0); }

void oracle_gaussian(double* out, int64_t count, uint64_t seed, double scale) {
  const double two_pi = 6.283185307179586476925286766559;
  for (int64_t b = 0; 4 * b < count; ++b) {
    uint32_t w[4];
    philox_block((uint64_t)b, seed, w);
    double z[4];
    for (int p = 0; p < 2; ++p) {
      const double r = sqrt(-2.0 * log(unit_open(w[2 * p])));
      const double t = two_pi * unit_open(w[2 * p + 1]);
      z[2 * p] = r * cos(t);
      z[2 * p + 1] = r * sin(t);
    }
    for (int j = 0; j < 4 && 4 * b + j < count; ++j) out[4 * b + j] = z[j] * scale;
  }
}
