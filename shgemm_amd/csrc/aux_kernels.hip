// Auxiliary kernels: deterministic split-K reduction, beta-only update and the
// per-row exponent pass of the prescaled fp16 mode.
#include <algorithm>
#include "registry.hpp"

namespace shgemm_amd {

// ---------------------------------------------------------------------------
// Split-K reduction: C = alpha * sum_{s=0}^{S-1} ws[s] + beta * C, the slices
// summed in index order so the result does not depend on scheduling.
// ws element e is C element (e % inner, e / inner) in C's own storage order
// (inner = m for column-major C, n for transposed C).
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) splitk_reduce_kernel(const float* __restrict__ ws, int slices, int64_t elems,
                                                             int inner, float alpha, float beta,
                                                             float* __restrict__ c, int64_t ldc) {
  const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (e >= elems) return;
  float s = 0.0f;
  for (int i = 0; i < slices; ++i) s += ws[(int64_t)i * elems + e];
  const int64_t in = e % inner, out = e / inner;
  float* cp = c + in + out * ldc;
  *cp = (beta == 0.0f) ? alpha * s : alpha * s + beta * *cp;
}

// beta-only update (k == 0): C = beta * C, or zero when beta == 0.
__global__ void __launch_bounds__(256) scale_c_kernel(int64_t elems, int inner, float beta, float* __restrict__ c,
                                                       int64_t ldc) {
  const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (e >= elems) return;
  float* cp = c + (e % inner) + (e / inner) * ldc;
  *cp = (beta == 0.0f) ? 0.0f : beta * *cp;
}

// ---------------------------------------------------------------------------
// Per-row exponent of A for the prescaled fp16 mode.
//   row_max: |a| bit patterns, max-reduced with atomicMax (order-independent).
//   row_exp[r] = ilogb(max |A[r,:]|) - 14, so A[r,:] * 2^-row_exp lies below
//   2^15 (no FP16 overflow in hi) and its largest entries sit in [2^14, 2^15).
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) fill_u32_kernel(uint32_t* __restrict__ p, int n, uint32_t v) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i < n) p[i] = v;
}

// A stored with rows contiguous (op_n): element (r, k) at a[r + k*lda].
// grid.x covers rows in 256s, grid.y covers k in chunks of kchunk.
__global__ void __launch_bounds__(256) rowmax_rcontig_kernel(const float* __restrict__ a, int64_t lda, int m, int k,
                                                              int kchunk, uint32_t* __restrict__ row_max) {
  const int r = blockIdx.x * 256 + threadIdx.x;
  if (r >= m) return;
  const int k0 = blockIdx.y * kchunk, k1 = min(k, k0 + kchunk);
  uint32_t mx = 0;
  for (int kk = k0; kk < k1; ++kk) {
    const uint32_t b = __builtin_bit_cast(uint32_t, a[r + (int64_t)kk * lda]) & 0x7fffffffu;
    mx = b > mx ? b : mx;
  }
  atomicMax(&row_max[r], mx);
}

// A stored with k contiguous (op_t): element (r, k) at a[r*lda + k].
// One workgroup per (row, k chunk); a block reduction then one atomic.
__global__ void __launch_bounds__(256) rowmax_kcontig_kernel(const float* __restrict__ a, int64_t lda, int m, int k,
                                                              int kchunk, uint32_t* __restrict__ row_max) {
  const int r = blockIdx.x;
  const int k0 = blockIdx.y * kchunk, k1 = min(k, k0 + kchunk);
  uint32_t mx = 0;
  for (int kk = k0 + threadIdx.x; kk < k1; kk += 256) {
    const uint32_t b = __builtin_bit_cast(uint32_t, a[(int64_t)r * lda + kk]) & 0x7fffffffu;
    mx = b > mx ? b : mx;
  }
  __shared__ uint32_t red[256];
  red[threadIdx.x] = mx;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s) red[threadIdx.x] = max(red[threadIdx.x], red[threadIdx.x + s]);
    __syncthreads();
  }
  if (threadIdx.x == 0) atomicMax(&row_max[r], red[0]);
}

__global__ void __launch_bounds__(256) rowexp_kernel(const uint32_t* __restrict__ row_max, int m, int* __restrict__ row_exp) {
  const int r = blockIdx.x * 256 + threadIdx.x;
  if (r >= m) return;
  const uint32_t b = row_max[r];
  int e = 0;
  if (b != 0 && b < 0x7f800000u) {  // finite, non-zero
    const int biased = (int)(b >> 23);
    const int lg = biased != 0 ? biased - 127 : (31 - __builtin_clz(b)) - 149;  // ilogb(|a|max)
    e = lg - 14;
    e = e < -126 ? -126 : e;  // keep 2^-e a finite FP32 scale
  }
  row_exp[r] = e;
}


hipError_t launch_splitk_reduce(const float* ws, int slices, int64_t elems, int inner, float alpha, float beta,
                                float* c, int64_t ldc, hipStream_t s) {
  const int64_t blocks = (elems + 255) / 256;
  hipLaunchKernelGGL(splitk_reduce_kernel, dim3((unsigned)blocks), dim3(256), 0, s, ws, slices, elems, inner, alpha,
                     beta, c, ldc);
  return hipGetLastError();
}

hipError_t launch_scale_c(int64_t elems, int inner, float beta, float* c, int64_t ldc, hipStream_t s) {
  const int64_t blocks = (elems + 255) / 256;
  hipLaunchKernelGGL(scale_c_kernel, dim3((unsigned)blocks), dim3(256), 0, s, elems, inner, beta, c, ldc);
  return hipGetLastError();
}

hipError_t launch_row_exponents(const float* a, int64_t lda, int op_a, int m, int k, uint32_t* row_max,
                                int* row_exp, hipStream_t s) {
  const unsigned rb = (unsigned)((m + 255) / 256);
  hipLaunchKernelGGL(fill_u32_kernel, dim3(rb), dim3(256), 0, s, row_max, m, 0u);
  hipError_t err = hipGetLastError();
  if (err != hipSuccess) return err;
  if (k > 0) {
    if (op_a == OP_N) {
      const int kchunk = std::max(256, (int)((k + 65534) / 65535));
      const unsigned kb = (unsigned)((k + kchunk - 1) / kchunk);
      hipLaunchKernelGGL(rowmax_rcontig_kernel, dim3(rb, kb), dim3(256), 0, s, a, lda, m, k, kchunk, row_max);
    } else {
      const int kchunk = std::max(4096, (int)((k + 65534) / 65535));
      const unsigned kb = (unsigned)((k + kchunk - 1) / kchunk);
      hipLaunchKernelGGL(rowmax_kcontig_kernel, dim3((unsigned)m, kb), dim3(256), 0, s, a, lda, m, k, kchunk, row_max);
    }
    err = hipGetLastError();
    if (err != hipSuccess) return err;
  }
  hipLaunchKernelGGL(rowexp_kernel, dim3(rb), dim3(256), 0, s, row_max, m, row_exp);
  return hipGetLastError();
}

}  // namespace shgemm_amd
