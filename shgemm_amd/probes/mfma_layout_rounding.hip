// Measurement tool (not part of the library): lane layout of the gfx950 MFMA
// instructions the kernels use (32x32x16 f16/bf16, 32x32x2 f32) and how the
// f16 MFMA rounds its accumulation (exact product sum + one round-to-nearest-
// even, the property that lets the hi and lo products share one accumulator).
// One wave of 64 lanes; every buffer is sized exactly for it.
// Build and run:
//   hipcc -O2 -std=c++17 --offload-arch=gfx950 -Xclang -target-feature -Xclang -packed-fp32-ops \
//         shgemm_amd/probes/mfma_layout_rounding.hip -o /tmp/mfma_probe && /tmp/mfma_probe
#include <hip/hip_runtime.h>
#include <cstdio>
#include <cstdlib>
#include <cmath>
#include <vector>
#include <cstring>

typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

#define CHK(x) do { hipError_t e = (x); if (e != hipSuccess) { printf("HIP error %s at %d\n", hipGetErrorString(e), __LINE__); exit(1);} } while (0)

// A: 32x16 row-major (float values), B: 16x32 row-major, C out 32x32 row-major.
// Assumed layout: lane l: A[l%32][8*(l/32)+e], B[8*(l/32)+e][l%32], D[(j/4)*8 + (l/32)*4 + j%4][l%32]
__global__ void k_f16(const float* A, const float* B, const float* Cin, float* D) {
  int l = threadIdx.x;
  f16x8 a, b;
  for (int e = 0; e < 8; ++e) { a[e] = (_Float16)A[(l % 32) * 16 + 8 * (l / 32) + e]; b[e] = (_Float16)B[(8 * (l / 32) + e) * 32 + (l % 32)]; }
  f32x16 c;
  for (int j = 0; j < 16; ++j) c[j] = Cin[((j / 4) * 8 + (l / 32) * 4 + (j % 4)) * 32 + (l % 32)];
  c = __builtin_amdgcn_mfma_f32_32x32x16_f16(a, b, c, 0, 0, 0);
  for (int j = 0; j < 16; ++j) D[((j / 4) * 8 + (l / 32) * 4 + (j % 4)) * 32 + (l % 32)] = c[j];
}
__global__ void k_bf16(const float* A, const float* B, const float* Cin, float* D) {
  int l = threadIdx.x;
  bf16x8 a, b;
  for (int e = 0; e < 8; ++e) { a[e] = (__bf16)A[(l % 32) * 16 + 8 * (l / 32) + e]; b[e] = (__bf16)B[(8 * (l / 32) + e) * 32 + (l % 32)]; }
  f32x16 c;
  for (int j = 0; j < 16; ++j) c[j] = Cin[((j / 4) * 8 + (l / 32) * 4 + (j % 4)) * 32 + (l % 32)];
  c = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, c, 0, 0, 0);
  for (int j = 0; j < 16; ++j) D[((j / 4) * 8 + (l / 32) * 4 + (j % 4)) * 32 + (l % 32)] = c[j];
}
// f32 32x32x2: A 32x2, B 2x32: lane l: A[l%32][l/32], B[l/32][l%32]
__global__ void k_f32(const float* A, const float* B, const float* Cin, float* D) {
  int l = threadIdx.x;
  float a = A[(l % 32) * 2 + (l / 32)], b = B[(l / 32) * 32 + (l % 32)];
  f32x16 c;
  for (int j = 0; j < 16; ++j) c[j] = Cin[((j / 4) * 8 + (l / 32) * 4 + (j % 4)) * 32 + (l % 32)];
  c = __builtin_amdgcn_mfma_f32_32x32x2f32(a, b, c, 0, 0, 0);
  for (int j = 0; j < 16; ++j) D[((j / 4) * 8 + (l / 32) * 4 + (j % 4)) * 32 + (l % 32)] = c[j];
}

static void ref(int K, const std::vector<float>& A, const std::vector<float>& B, const std::vector<float>& C, std::vector<double>& D) {
  for (int i = 0; i < 32; ++i) for (int j = 0; j < 32; ++j) {
    double s = C[i * 32 + j];
    for (int k = 0; k < K; ++k) s += (double)A[i * K + k] * (double)B[k * 32 + j];
    D[i * 32 + j] = s;
  }
}

template <class F>
static void run(F kern, int K, const std::vector<float>& A, const std::vector<float>& B, const std::vector<float>& C, std::vector<float>& D) {
  float *dA, *dB, *dC, *dD;
  CHK(hipMalloc(&dA, A.size() * 4)); CHK(hipMalloc(&dB, B.size() * 4)); CHK(hipMalloc(&dC, 1024 * 4)); CHK(hipMalloc(&dD, 1024 * 4));
  CHK(hipMemcpy(dA, A.data(), A.size() * 4, hipMemcpyHostToDevice));
  CHK(hipMemcpy(dB, B.data(), B.size() * 4, hipMemcpyHostToDevice));
  CHK(hipMemcpy(dC, C.data(), 1024 * 4, hipMemcpyHostToDevice));
  hipLaunchKernelGGL(kern, dim3(1), dim3(64), 0, 0, dA, dB, dC, dD);
  CHK(hipGetLastError());
  CHK(hipDeviceSynchronize());
  D.resize(1024);
  CHK(hipMemcpy(D.data(), dD, 1024 * 4, hipMemcpyDeviceToHost));
  CHK(hipFree(dA)); CHK(hipFree(dB)); CHK(hipFree(dC)); CHK(hipFree(dD));
}

static int layout_check(const char* name, decltype(&k_f16) kern, int K) {
  srand(1);
  std::vector<float> A(32 * K), B(K * 32), C(1024);
  for (auto& x : A) x = (float)(rand() % 9 - 4);
  for (auto& x : B) x = (float)(rand() % 9 - 4);
  for (auto& x : C) x = (float)(rand() % 9 - 4);
  std::vector<float> D; std::vector<double> R(1024);
  run(kern, K, A, B, C, D);
  ref(K, A, B, C, R);
  int bad = 0;
  for (int i = 0; i < 1024; ++i) if ((double)D[i] != R[i]) ++bad;
  printf("layout %-6s K=%2d mismatches=%d\n", name, K, bad);
  return bad;
}

// rounding cases on f16 MFMA: C = c0 everywhere, only row0/col0 product matters
static void round_case(const char* name, float c0, std::vector<std::pair<float, float>> prods) {
  std::vector<float> A(32 * 16, 0.f), B(16 * 32, 0.f), C(1024, c0);
  for (size_t k = 0; k < prods.size(); ++k) { A[0 * 16 + k] = prods[k].first; B[k * 32 + 0] = prods[k].second; }
  std::vector<float> D; std::vector<double> R(1024);
  run(k_f16, 16, A, B, C, D);
  ref(16, A, B, C, R);
  double rne = (double)(float)R[0];
  printf("round %-28s c0=%.9g exact=%.17g gpu=%.17g (fp32(exact)=%.17g) diff_ulps=%g\n", name, c0, R[0], (double)D[0], rne,
         ((double)D[0] - R[0]) / (double)(std::ldexp(1.0, std::ilogb(R[0]) - 23)));
}

int main() {
  int bad = 0;
  bad += layout_check("f16", k_f16, 16);
  bad += layout_check("bf16", k_bf16, 16);
  bad += layout_check("f32", k_f32, 2);
  const float t = std::ldexp(1.f, -13), u = std::ldexp(1.f, -12);  // t*u = 2^-25
  round_case("16 x 2^-25 onto 1", 1.f, std::vector<std::pair<float, float>>(16, {t, u}));
  round_case("1 x 0.75*2^-23 onto 1", 1.f, {{0.75f * std::ldexp(1.f, -11), std::ldexp(1.f, -12)}});
  round_case("1 x 0.5*2^-23 onto 1 (tie)", 1.f, {{0.5f * std::ldexp(1.f, -11), std::ldexp(1.f, -12)}});
  round_case("1 x 1.5*2^-23 onto 1 (tie)", 1.f, {{1.5f * std::ldexp(1.f, -11), std::ldexp(1.f, -12)}});
  round_case("1 x -0.75*2^-24 onto 1", 1.f, {{-0.75f * std::ldexp(1.f, -12), std::ldexp(1.f, -12)}});
  round_case("1 x 0.25*2^-23 onto 1", 1.f, {{0.25f * std::ldexp(1.f, -11), std::ldexp(1.f, -12)}});
  round_case("2 x 0.3*2^-23 onto 1", 1.f, std::vector<std::pair<float, float>>(2, {0.3f * std::ldexp(1.f, -11), std::ldexp(1.f, -12)}));
  round_case("1 + (-1) + 2^-24 onto 0", 0.f, {{1.f, 1.f}, {-1.f, 1.f}, {t, std::ldexp(1.f, -11)}});
  round_case("big cancel 2048*2048 - ...", 0.f, {{2048.f, 2048.f}, {-2048.f, 2047.f}, {t, u}});
  round_case("products 2^-30 x 16 onto 0", 0.f, std::vector<std::pair<float, float>>(16, {std::ldexp(1.f, -14), std::ldexp(1.f, -16)}));
  round_case("1 + 2^-24 + 2^-24 (c0=0)", 0.f, {{1.f, 1.f}, {std::ldexp(1.f, -12), std::ldexp(1.f, -12)}, {std::ldexp(1.f, -12), std::ldexp(1.f, -12)}});
  printf("probe done bad=%d\n", bad);
  return bad ? 1 : 0;
}
