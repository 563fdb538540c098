// Test-support kernels (not part of the library; the GPU tests load them to
// check premises of the library's arithmetic against the oracle on the host):
//   * tk_mfma runs one gfx950 MFMA instruction of the kinds the kernels use
//     (v_mfma_f32_32x32x16_f16 / _bf16 and v_mfma_f32_32x32x2_f32) per
//     wavefront, on operands given in row-major order;
//   * tk_split_pair runs the library's paired split of the prescaled fp16
//     mode (split2_f16_scaled in gemm_common.hpp) on given values;
//   * tk_chain_gemm runs a plain split GEMM (one wavefront per 32x32 block,
//     A split into FP16 hi + 2^11-scaled FP16 residual as the reference
//     does) with the products accumulated three ways: hi and residual each
//     in its own MFMA chain (the library's classic mode), the hi product of
//     every MFMA added to an FP32 accumulator outside the chain (the
//     reference's scheme, src/wmmae_shgemm.hpp:86-98), or both outside;
//   * tk_read_bw times a streaming read of a device buffer (16-byte loads,
//     8 per thread in flight, one contiguous chunk per workgroup): the HBM
//     read rate the bandwidth-bound kernels are judged against.
// Every entry point copies host arrays in and out (one wavefront per case;
// grids are sized exactly from the case count) and returns 0 or a HIP error
// code.
#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstring>
#include <utility>

#include "../../shgemm_amd/csrc/gemm_common.hpp"

namespace {

typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

// Output row of accumulator element j of lane l for the 32x32 MFMA blocks
// (column l % 32); the operand layout is A[l % 32][8 (l / 32) + e] and
// B[8 (l / 32) + e][l % 32] for the 16-deep kinds, A[l % 32][l / 32] and
// B[l / 32][l % 32] for the 2-deep FP32 kind.
__device__ __forceinline__ int acc_row(int l, int j) { return (j / 4) * 8 + (l / 32) * 4 + (j % 4); }

// a, b: FP16 bits (kind 0) or BF16 bits (kind 1), 32x16 and 16x32 per case.
template <int KIND>
__global__ void __launch_bounds__(64) mfma16_kernel(const uint16_t* __restrict__ a, const uint16_t* __restrict__ b,
                                                    const float* __restrict__ c, float* __restrict__ d) {
  const int l = threadIdx.x;
  const int64_t cs = blockIdx.x;
  const uint16_t* A = a + cs * 512;
  const uint16_t* B = b + cs * 512;
  const float* Cin = c + cs * 1024;
  float* D = d + cs * 1024;
  f32x16 acc;
  for (int j = 0; j < 16; ++j) acc[j] = Cin[acc_row(l, j) * 32 + (l % 32)];
  if constexpr (KIND == 0) {
    f16x8 x, y;
    for (int e = 0; e < 8; ++e) {
      x[e] = __builtin_bit_cast(_Float16, A[(l % 32) * 16 + 8 * (l / 32) + e]);
      y[e] = __builtin_bit_cast(_Float16, B[(8 * (l / 32) + e) * 32 + (l % 32)]);
    }
    acc = __builtin_amdgcn_mfma_f32_32x32x16_f16(x, y, acc, 0, 0, 0);
  } else {
    bf16x8 x, y;
    for (int e = 0; e < 8; ++e) {
      x[e] = __builtin_bit_cast(__bf16, A[(l % 32) * 16 + 8 * (l / 32) + e]);
      y[e] = __builtin_bit_cast(__bf16, B[(8 * (l / 32) + e) * 32 + (l % 32)]);
    }
    acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(x, y, acc, 0, 0, 0);
  }
  for (int j = 0; j < 16; ++j) D[acc_row(l, j) * 32 + (l % 32)] = acc[j];
}

// a: 32x2, b: 2x32 FP32 per case.
__global__ void __launch_bounds__(64) mfma_f32_kernel(const float* __restrict__ a, const float* __restrict__ b,
                                                      const float* __restrict__ c, float* __restrict__ d) {
  const int l = threadIdx.x;
  const int64_t cs = blockIdx.x;
  const float x = a[cs * 64 + (l % 32) * 2 + (l / 32)];
  const float y = b[cs * 64 + (l / 32) * 32 + (l % 32)];
  f32x16 acc;
  for (int j = 0; j < 16; ++j) acc[j] = c[cs * 1024 + acc_row(l, j) * 32 + (l % 32)];
  acc = __builtin_amdgcn_mfma_f32_32x32x2f32(x, y, acc, 0, 0, 0);
  for (int j = 0; j < 16; ++j) d[cs * 1024 + acc_row(l, j) * 32 + (l % 32)] = acc[j];
}

__global__ void __launch_bounds__(256) split_pair_kernel(const float* __restrict__ x, const float* __restrict__ s,
                                                          uint32_t* __restrict__ hi, uint32_t* __restrict__ lo,
                                                          int npairs) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= npairs) return;
  uint32_t h, o;
  shgemm_amd::split2_f16_scaled(x[2 * i], x[2 * i + 1], s[i], h, o);
  hi[i] = h;
  lo[i] = o;
}

// One wavefront per 32x32 block of C (row-major, m x n); A row-major m x k
// (FP32), B as n rows of k FP16 bits (column j of B contiguous); k % 16 == 0.
template <int MODE>
__global__ void __launch_bounds__(64) chain_gemm_kernel(const float* __restrict__ a, const uint16_t* __restrict__ b,
                                                        float* __restrict__ c, int m, int n, int k) {
  const int l = threadIdx.x;
  const int bm = blockIdx.x % (m / 32), bn = blockIdx.x / (m / 32);
  const float* A = a + (int64_t)(bm * 32 + l % 32) * k + 8 * (l / 32);
  const uint16_t* B = b + (int64_t)(bn * 32 + l % 32) * k + 8 * (l / 32);
  f32x16 hi_acc, lo_acc, zero;
  for (int j = 0; j < 16; ++j) hi_acc[j] = lo_acc[j] = zero[j] = 0.0f;
  for (int k0 = 0; k0 < k; k0 += 16) {
    f16x8 xh, xl, y;
    for (int e = 0; e < 8; ++e) {
      const float v = A[k0 + e];
      const _Float16 h = (_Float16)v;
      xh[e] = h;
      xl[e] = (_Float16)((v - (float)h) * 2048.0f);
      y[e] = __builtin_bit_cast(_Float16, B[k0 + e]);
    }
    if constexpr (MODE == 0) {
      hi_acc = __builtin_amdgcn_mfma_f32_32x32x16_f16(xh, y, hi_acc, 0, 0, 0);
    } else {
      const f32x16 t = __builtin_amdgcn_mfma_f32_32x32x16_f16(xh, y, zero, 0, 0, 0);
      for (int j = 0; j < 16; ++j) hi_acc[j] += t[j];
    }
    if constexpr (MODE == 2) {
      const f32x16 t = __builtin_amdgcn_mfma_f32_32x32x16_f16(xl, y, zero, 0, 0, 0);
      for (int j = 0; j < 16; ++j) lo_acc[j] += t[j];
    } else {
      lo_acc = __builtin_amdgcn_mfma_f32_32x32x16_f16(xl, y, lo_acc, 0, 0, 0);
    }
  }
  for (int j = 0; j < 16; ++j)
    c[(int64_t)(bm * 32 + acc_row(l, j)) * n + bn * 32 + l % 32] = __builtin_fmaf(lo_acc[j], 1.0f / 2048.0f, hi_acc[j]);
}

// Each workgroup reads one contiguous chunk of `per_wg` 16-byte words, 8
// loads per thread in flight; per_wg % 2048 == 0.
__global__ void __launch_bounds__(256) read_bw_kernel(const uint4* __restrict__ p, int64_t per_wg, uint32_t* out) {
  const uint4* q = p + (int64_t)blockIdx.x * per_wg;
  uint32_t acc = 0;
  for (int64_t i = threadIdx.x; i < per_wg; i += 256 * 8) {
    uint4 v[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) v[u] = q[i + u * 256];
#pragma unroll
    for (int u = 0; u < 8; ++u) acc ^= v[u].x ^ v[u].w;
  }
  if (acc == 0x9e3779b9u) out[0] = acc;  // keeps the loads; never true for the zero-filled buffer
}

struct DevBufs {
  void* p[4] = {nullptr, nullptr, nullptr, nullptr};
  ~DevBufs() {
    for (void* q : p)
      if (q) (void)hipFree(q);
  }
};

#define TK_CHECK(x)                     \
  do {                                  \
    const hipError_t e_ = (x);          \
    if (e_ != hipSuccess) return (int)e_; \
  } while (0)

}  // namespace

extern "C" {

// kind 0: f16 32x32x16, 1: bf16 32x32x16 (a, b: uint16 bits, 512 per case);
// kind 2: f32 32x32x2 (a, b: float, 64 per case).  c, d: 1024 floats per case.
int tk_mfma(int kind, int cases, const void* a, const void* b, const float* c, float* d) {
  if (cases <= 0 || kind < 0 || kind > 2) return -1;
  const size_t ab = kind == 2 ? 64 * sizeof(float) : 512 * sizeof(uint16_t);
  const size_t cd = 1024 * sizeof(float);
  DevBufs db;
  TK_CHECK(hipMalloc(&db.p[0], ab * cases));
  TK_CHECK(hipMalloc(&db.p[1], ab * cases));
  TK_CHECK(hipMalloc(&db.p[2], cd * cases));
  TK_CHECK(hipMalloc(&db.p[3], cd * cases));
  TK_CHECK(hipMemcpy(db.p[0], a, ab * cases, hipMemcpyHostToDevice));
  TK_CHECK(hipMemcpy(db.p[1], b, ab * cases, hipMemcpyHostToDevice));
  TK_CHECK(hipMemcpy(db.p[2], c, cd * cases, hipMemcpyHostToDevice));
  if (kind == 0)
    hipLaunchKernelGGL(mfma16_kernel<0>, dim3(cases), dim3(64), 0, 0, (const uint16_t*)db.p[0],
                       (const uint16_t*)db.p[1], (const float*)db.p[2], (float*)db.p[3]);
  else if (kind == 1)
    hipLaunchKernelGGL(mfma16_kernel<1>, dim3(cases), dim3(64), 0, 0, (const uint16_t*)db.p[0],
                       (const uint16_t*)db.p[1], (const float*)db.p[2], (float*)db.p[3]);
  else
    hipLaunchKernelGGL(mfma_f32_kernel, dim3(cases), dim3(64), 0, 0, (const float*)db.p[0], (const float*)db.p[1],
                       (const float*)db.p[2], (float*)db.p[3]);
  TK_CHECK(hipGetLastError());
  TK_CHECK(hipDeviceSynchronize());
  TK_CHECK(hipMemcpy(d, db.p[3], cd * cases, hipMemcpyDeviceToHost));
  return 0;
}

// hi[i] / lo[i]: the FP16 pair words of split2_f16_scaled(x[2i], x[2i+1], s[i])
// (element 2i in the low half, 2i+1 in the high half).
int tk_split_pair(int npairs, const float* x, const float* s, uint32_t* hi, uint32_t* lo) {
  if (npairs <= 0) return -1;
  DevBufs db;
  TK_CHECK(hipMalloc(&db.p[0], sizeof(float) * 2 * (size_t)npairs));
  TK_CHECK(hipMalloc(&db.p[1], sizeof(float) * (size_t)npairs));
  TK_CHECK(hipMalloc(&db.p[2], sizeof(uint32_t) * (size_t)npairs));
  TK_CHECK(hipMalloc(&db.p[3], sizeof(uint32_t) * (size_t)npairs));
  TK_CHECK(hipMemcpy(db.p[0], x, sizeof(float) * 2 * (size_t)npairs, hipMemcpyHostToDevice));
  TK_CHECK(hipMemcpy(db.p[1], s, sizeof(float) * (size_t)npairs, hipMemcpyHostToDevice));
  hipLaunchKernelGGL(split_pair_kernel, dim3((npairs + 255) / 256), dim3(256), 0, 0, (const float*)db.p[0],
                     (const float*)db.p[1], (uint32_t*)db.p[2], (uint32_t*)db.p[3], npairs);
  TK_CHECK(hipGetLastError());
  TK_CHECK(hipDeviceSynchronize());
  TK_CHECK(hipMemcpy(hi, db.p[2], sizeof(uint32_t) * (size_t)npairs, hipMemcpyDeviceToHost));
  TK_CHECK(hipMemcpy(lo, db.p[3], sizeof(uint32_t) * (size_t)npairs, hipMemcpyDeviceToHost));
  return 0;
}

// C (m x n, row-major) = A (m x k, row-major FP32) * B^T (B: n x k FP16 bits),
// accumulated as `mode` says (0: two MFMA chains, 1: hi outside the chain,
// 2: both outside); m, n multiples of 32, k of 16.
int tk_chain_gemm(int mode, int m, int n, int k, const float* a, const uint16_t* b, float* c) {
  if (mode < 0 || mode > 2 || m <= 0 || n <= 0 || k <= 0 || m % 32 || n % 32 || k % 16) return -1;
  DevBufs db;
  const size_t ab = sizeof(float) * (size_t)m * k, bb = sizeof(uint16_t) * (size_t)n * k, cb = sizeof(float) * (size_t)m * n;
  TK_CHECK(hipMalloc(&db.p[0], ab));
  TK_CHECK(hipMalloc(&db.p[1], bb));
  TK_CHECK(hipMalloc(&db.p[2], cb));
  TK_CHECK(hipMemcpy(db.p[0], a, ab, hipMemcpyHostToDevice));
  TK_CHECK(hipMemcpy(db.p[1], b, bb, hipMemcpyHostToDevice));
  const dim3 grid((unsigned)((m / 32) * (n / 32)));
  const float* A = (const float*)db.p[0];
  const uint16_t* B = (const uint16_t*)db.p[1];
  float* C = (float*)db.p[2];
  if (mode == 0) hipLaunchKernelGGL(chain_gemm_kernel<0>, grid, dim3(64), 0, 0, A, B, C, m, n, k);
  else if (mode == 1) hipLaunchKernelGGL(chain_gemm_kernel<1>, grid, dim3(64), 0, 0, A, B, C, m, n, k);
  else hipLaunchKernelGGL(chain_gemm_kernel<2>, grid, dim3(64), 0, 0, A, B, C, m, n, k);
  TK_CHECK(hipGetLastError());
  TK_CHECK(hipDeviceSynchronize());
  TK_CHECK(hipMemcpy(c, C, cb, hipMemcpyDeviceToHost));
  return 0;
}

// Milliseconds per read of `bytes` (a multiple of 2^20) with `grid`
// workgroups (dividing bytes / 2^15), median of `reps` timed passes after
// three untimed ones.
int tk_read_bw(int64_t bytes, int grid, int reps, float* ms) {
  if (bytes <= 0 || bytes % (1 << 20) || grid <= 0 || reps <= 0 || (bytes / 16) % ((int64_t)grid * 2048)) return -1;
  DevBufs db;
  TK_CHECK(hipMalloc(&db.p[0], (size_t)bytes));
  TK_CHECK(hipMalloc(&db.p[1], 4));
  TK_CHECK(hipMemset(db.p[0], 0, (size_t)bytes));
  const int64_t per = bytes / 16 / grid;
  hipEvent_t e0, e1;
  TK_CHECK(hipEventCreate(&e0));
  TK_CHECK(hipEventCreate(&e1));
  float best[64];
  const int r = reps < 64 ? reps : 64;
  for (int i = 0; i < 3 + r; ++i) {
    TK_CHECK(hipEventRecord(e0, 0));
    hipLaunchKernelGGL(read_bw_kernel, dim3((unsigned)grid), dim3(256), 0, 0, (const uint4*)db.p[0], per,
                       (uint32_t*)db.p[1]);
    TK_CHECK(hipEventRecord(e1, 0));
    TK_CHECK(hipEventSynchronize(e1));
    if (i >= 3) TK_CHECK(hipEventElapsedTime(&best[i - 3], e0, e1));
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  for (int i = 1; i < r; ++i)
    for (int j = i; j > 0 && best[j] < best[j - 1]; --j) std::swap(best[j], best[j - 1]);
  *ms = best[r / 2];
  return 0;
}

}  // extern "C"
