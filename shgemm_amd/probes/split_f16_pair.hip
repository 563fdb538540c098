// Measurement tool (not part of the library): checks bit for bit that the
// paired split of the prescaled fp16 mode (split2_f16_scaled, fma_mix
// instructions) equals the C form split_f16_same_scale(x * s) over 2^22
// inputs: wide exponents, the FP16 top range, rounding ties and the FP16
// subnormal range, each at a power-of-two scale in 2^-10..2^9.
// Thread i reads elements 2i, 2i+1 and scale i and writes words 2i, 2i+1 of
// each output; the grid covers npairs with a bounds check.
// Build and run:
//   hipcc -O2 -std=c++17 --offload-arch=gfx950 -Xclang -target-feature -Xclang -packed-fp32-ops \
//         shgemm_amd/probes/split_f16_pair.hip -o /tmp/split_probe && /tmp/split_probe
#include <hip/hip_runtime.h>
#include <cmath>
#include <cstdio>
#include <random>
#include <vector>
#include "../csrc/gemm_common.hpp"

using shgemm_amd::split2_f16_scaled;
using shgemm_amd::split_f16_same_scale;

__global__ void compare_split(const float* __restrict__ x, const float* __restrict__ s, uint32_t* __restrict__ ref,
                              uint32_t* __restrict__ alt, int npairs) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= npairs) return;
  const float x0 = x[2 * i], x1 = x[2 * i + 1], sc = s[i];
  _Float16 h0, l0, h1, l1;
  split_f16_same_scale(x0 * sc, h0, l0);
  split_f16_same_scale(x1 * sc, h1, l1);
  ref[2 * i] = (uint32_t)__builtin_bit_cast(uint16_t, h0) | ((uint32_t)__builtin_bit_cast(uint16_t, h1) << 16);
  ref[2 * i + 1] = (uint32_t)__builtin_bit_cast(uint16_t, l0) | ((uint32_t)__builtin_bit_cast(uint16_t, l1) << 16);
  uint32_t hh, ll;
  split2_f16_scaled(x0, x1, sc, hh, ll);
  alt[2 * i] = hh;
  alt[2 * i + 1] = ll;
}

int main() {
  const int npairs = 1 << 21, n = 2 * npairs;
  std::vector<float> x(n), s(npairs);
  std::mt19937_64 g(7);
  std::uniform_real_distribution<double> u(0.0, 1.0);
  for (int i = 0; i < n; ++i) {
    double v;
    switch (i % 4) {
      case 0: v = (u(g) * 2 - 1) * std::ldexp(1.0, (int)(u(g) * 60) - 30); break;       // wide exponents
      case 1: v = (u(g) * 2 - 1) * 65504.0 * 1.01; break;                                // FP16 top range
      case 2: v = std::ldexp(1.0 + std::floor(u(g) * 2048) / 2048 + std::ldexp(1.0, -12),  // rounding ties
                             (int)(u(g) * 40) - 30); break;
      default: v = (u(g) * 2 - 1) * std::ldexp(1.0, -24 - (int)(u(g) * 8)); break;     // FP16 subnormals
    }
    x[i] = (float)v;
  }
  for (int i = 0; i < npairs; ++i) s[i] = std::ldexp(1.0f, (int)(u(g) * 20) - 10);
  x[0] = 0.0f; x[1] = -0.0f; x[2] = 65504.0f; x[3] = -65504.0f; s[0] = s[1] = 1.0f;
  float *dx, *ds;
  uint32_t *dr, *da;
  if (hipMalloc(&dx, n * 4) || hipMalloc(&ds, npairs * 4) || hipMalloc(&dr, n * 4) || hipMalloc(&da, n * 4)) return 1;
  if (hipMemcpy(dx, x.data(), n * 4, hipMemcpyHostToDevice) || hipMemcpy(ds, s.data(), npairs * 4, hipMemcpyHostToDevice))
    return 1;
  hipLaunchKernelGGL(compare_split, dim3((npairs + 255) / 256), dim3(256), 0, 0, dx, ds, dr, da, npairs);
  if (hipDeviceSynchronize() != hipSuccess) {
    std::printf("launch failed\n");
    return 1;
  }
  std::vector<uint32_t> r(n), a(n);
  if (hipMemcpy(r.data(), dr, n * 4, hipMemcpyDeviceToHost) || hipMemcpy(a.data(), da, n * 4, hipMemcpyDeviceToHost))
    return 1;
  long bad = 0;
  for (int i = 0; i < n; ++i)
    if (r[i] != a[i] && bad++ < 10) std::printf("mismatch word %d: C form %08x paired %08x\n", i, r[i], a[i]);
  std::printf("pairs %d, mismatching words %ld\n", npairs, bad);
  (void)hipFree(dx);
  (void)hipFree(ds);
  (void)hipFree(dr);
  (void)hipFree(da);
  return bad ? 2 : 0;
}
