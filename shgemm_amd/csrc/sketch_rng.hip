// On-device random projection matrices for sketches: FP16 Gaussian entries
// from a counter-based generator, so element i of a matrix depends only on
// (seed, i) -- the same Omega whatever the launch shape, stream or order.
//
// Generator: Philox4x32-10 (J. Salmon, M. Moraes, R. Dror, D. Shaw, "Parallel
// random numbers: as easy as 1, 2, 3", SC'11): 10 rounds of two 32x32->64-bit
// multiplies with the published multipliers and Weyl key increments.
// Counter of block b = (b mod 2^32, b / 2^32, 0, 0), key = (seed mod 2^32,
// seed / 2^32).  Each block gives four words w0..w3.
// Normals: Box-Muller on (w0, w1) and (w2, w3): u = ((w >> 9) + 1/2) * 2^-23
// lies in (0, 1) and is exact in FP32 (24 significant bits);
// r = sqrt(-2 ln u_a), z_a = r cos(2 pi u_b), z_b = r sin(2 pi u_b).
// Element 4b + j takes normal j of block b, times `scale`, rounded to FP16.
// The reference seeds its inputs with cuRAND Philox (test/shgemm_test.cu:247-252);
// this is the project's own generator, not a reproduction of that stream.
#include <hip/hip_runtime.h>
#include <cstdint>

#include "registry.hpp"

namespace shgemm_amd {

struct Philox4 {
  uint32_t w[4];
};

__host__ __device__ __forceinline__ Philox4 philox4x32_10(uint64_t block, uint64_t seed) {
  constexpr uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u, W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
  uint32_t c0 = (uint32_t)block, c1 = (uint32_t)(block >> 32), c2 = 0u, c3 = 0u;
  uint32_t k0 = (uint32_t)seed, k1 = (uint32_t)(seed >> 32);
#pragma unroll
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
  return Philox4{{c0, c1, c2, c3}};
}

__device__ __forceinline__ float unit_open(uint32_t w) { return ((float)(w >> 9) + 0.5f) * 1.1920928955078125e-7f; }

__device__ __forceinline__ void box_muller(uint32_t wa, uint32_t wb, float& za, float& zb) {
  const float r = sqrtf(-2.0f * logf(unit_open(wa)));
  float s, c;
  sincospif(2.0f * unit_open(wb), &s, &c);
  za = r * c;
  zb = r * s;
}

__global__ void __launch_bounds__(256) gaussian_f16_kernel(uint16_t* __restrict__ out, int64_t count, uint64_t seed,
                                                            float scale) {
  const int64_t b = (int64_t)blockIdx.x * 256 + threadIdx.x;
  const int64_t i0 = 4 * b;
  if (i0 >= count) return;
  const Philox4 x = philox4x32_10((uint64_t)b, seed);
  float z[4];
  box_muller(x.w[0], x.w[1], z[0], z[1]);
  box_muller(x.w[2], x.w[3], z[2], z[3]);
  uint16_t h[4];
#pragma unroll
  for (int j = 0; j < 4; ++j) h[j] = __builtin_bit_cast(uint16_t, (_Float16)(z[j] * scale));
  if (i0 + 4 <= count && (reinterpret_cast<uintptr_t>(out + i0) & 7) == 0) {
    uint2 v;
    v.x = (uint32_t)h[0] | ((uint32_t)h[1] << 16);
    v.y = (uint32_t)h[2] | ((uint32_t)h[3] << 16);
    *reinterpret_cast<uint2*>(out + i0) = v;
  } else {
    for (int j = 0; j < 4 && i0 + j < count; ++j) out[i0 + j] = h[j];
  }
}

__global__ void __launch_bounds__(256) philox_words_kernel(uint32_t* __restrict__ out, int64_t blocks, uint64_t seed) {
  const int64_t b = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (b >= blocks) return;
  const Philox4 x = philox4x32_10((uint64_t)b, seed);
  *reinterpret_cast<uint4*>(out + 4 * b) = make_uint4(x.w[0], x.w[1], x.w[2], x.w[3]);
}

hipError_t launch_gaussian_f16(uint16_t* out, int64_t count, uint64_t seed, float scale, hipStream_t s) {
  if (count <= 0) return hipSuccess;
  const int64_t blocks = (count + 3) / 4;
  const int64_t grid = (blocks + 255) / 256;
  if (grid > 0x7fffffff) return hipErrorInvalidValue;
  hipLaunchKernelGGL(gaussian_f16_kernel, dim3((unsigned)grid), dim3(256), 0, s, out, count, seed, scale);
  return hipGetLastError();
}

hipError_t launch_philox_words(uint32_t* out, int64_t blocks, uint64_t seed, hipStream_t s) {
  if (blocks <= 0) return hipSuccess;
  const int64_t grid = (blocks + 255) / 256;
  if (grid > 0x7fffffff) return hipErrorInvalidValue;
  hipLaunchKernelGGL(philox_words_kernel, dim3((unsigned)grid), dim3(256), 0, s, out, blocks, seed);
  return hipGetLastError();
}

}  // namespace shgemm_amd
