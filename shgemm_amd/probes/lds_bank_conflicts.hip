// LDS bank-conflict probe for the SHGEMM operand layout on gfx950: 16-byte
// stores and loads in the three patterns of the kernels -- a wavefront storing
// one chunk of 64 consecutive rows (m-contiguous A), storing whole rows
// (k-contiguous operands), and reading MFMA operand fragments (lanes 0-31 rows
// of a 32-row block at chunk c, lanes 32-63 at chunk c+1) -- under three
// swizzle keys: 0 the window key r / (16 / KC), 1 ((r >> (S - 1)) ^ (r >> 3))
// with S = log2(16 / KC) (for KC = 8: r ^ (r >> 3)), 2 (r >> 3) for KC = 4.
// Count the conflicts with rocprofv3 --pmc SQ_LDS_BANK_CONFLICT SQ_INSTS_LDS.
// Every index stays below R * 8 slots (checked on the host for all patterns).
#include <hip/hip_runtime.h>
#include <cstdio>
#include <cstdlib>
constexpr int R = 128, NT = 512, ITERS = 2000;
template <int KEYT, int KC>
__device__ __forceinline__ int slot(int r, int c) {
  constexpr int S = KC == 4 ? 2 : 1;
  const int key = KEYT == 0 ? (r >> S) : KEYT == 1 ? ((r >> (S - 1)) ^ (r >> 3)) : (r >> 3);
  return r * KC + (c ^ (key & (KC - 1)));
}
template <int MODE, int KEYT, int KC>
__global__ void __launch_bounds__(NT) probe(float* out) {
  __shared__ uint4 lds[R * 8];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  uint4 acc = {0, 0, 0, 0};
  for (int i = tid; i < R * 8; i += NT) lds[i] = uint4{(unsigned)i, 1u, 2u, 3u};
  __syncthreads();
  for (int it = 0; it < ITERS; ++it) {
    if constexpr (MODE == 0) {
      const int r = tid % R, c = tid / R;  // c in 0..3
      lds[slot<KEYT, KC>(r, c)] = uint4{(unsigned)it, (unsigned)r, (unsigned)c, 0u};
    } else if constexpr (MODE == 1) {
      const int r = tid / KC, c = tid % KC;
      lds[slot<KEYT, KC>(r, c)] = uint4{(unsigned)it, (unsigned)r, (unsigned)c, 0u};
    } else {
      const int lrow = lane & 31, lh = lane >> 5;
      const int r = (wave & 3) * 32 + lrow, c = (it & 1) * 2 + lh;
      const uint4 v = lds[slot<KEYT, KC>(r, c)];
      acc.x += v.x; acc.y ^= v.y; acc.z += v.z; acc.w ^= v.w;
    }
    __builtin_amdgcn_s_waitcnt(0xc07f);
  }
  __syncthreads();
  out[blockIdx.x * NT + tid] = (float)(acc.x + acc.y + acc.z + acc.w) + (float)lds[tid].x;
}
int main() {
  float* out;
  if (hipMalloc(&out, 64 * NT * sizeof(float)) != hipSuccess) return 1;
  auto run = [&](auto k, const char* name) {
    hipLaunchKernelGGL(k, dim3(64), dim3(NT), 0, 0, out);
    if (hipDeviceSynchronize() != hipSuccess) { std::printf("%s failed\n", name); std::exit(1); }
    std::printf("%s done\n", name);
    std::fflush(stdout);
  };
  run(probe<0, 0, 4>, "m4o"); run(probe<0, 1, 4>, "m4n"); run(probe<1, 0, 4>, "k4o"); run(probe<1, 1, 4>, "k4n");
  run(probe<2, 0, 4>, "r4o"); run(probe<2, 1, 4>, "r4n");
  run(probe<0, 2, 4>, "m4k2"); run(probe<1, 2, 4>, "k4k2"); run(probe<2, 2, 4>, "r4k2");
  run(probe<0, 0, 8>, "m8o"); run(probe<0, 1, 8>, "m8n"); run(probe<1, 0, 8>, "k8o"); run(probe<1, 1, 8>, "k8n");
  run(probe<2, 0, 8>, "r8o"); run(probe<2, 1, 8>, "r8n");
  hipFree(out);
  return 0;
}
