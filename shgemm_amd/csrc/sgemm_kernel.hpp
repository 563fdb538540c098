// Plain FP32 SGEMM on gfx950 FP32 MFMA (v_mfma_f32_32x32x2_f32).
// The comparator the error-corrected SHGEMM is measured against: same inputs
// (B already widened to FP32), same tiling scheme, one accumulator.
//
// LDS: FP32 chunks of 4 values (16 bytes) in swizzled row-major slots (lds_slot_rm).
// MFMA k-order: inside each 16-k group, lane half h owns k = 8h..8h+7 and
// MFMA step s uses k = s (h=0) and 8+s (h=1); A and B use the same order, so
// every k is multiplied exactly once.
#pragma once
#include "gemm_common.hpp"

namespace shgemm_amd {

// FP32 tile, chunks of 4 k (16 bytes), row-major swizzled slots (lds_slot_rm).
// KCONT item = (row, chunk), chunk fastest: one 16-byte load.
// !KCONT item = (row, chunk), row fastest: four 4-byte loads down 4 columns,
// each instruction reading 64 consecutive values of one column.
template <bool KCONT, int R, int BK>
struct TileF32x4 {
  static constexpr int KC = BK / 4;
  static constexpr int NITEMS = R * KC;
  static constexpr int PER = NITEMS / 256;
  static_assert(NITEMS % 256 == 0 && PER >= 1, "tile must split evenly over 256 threads");
  static_assert(KCONT || R % 64 == 0, "row-fastest map needs whole wavefronts per chunk");
  f32x4 v[PER];

  __device__ __forceinline__ static void item(int i, int& r, int& c) {
    if constexpr (KCONT) {
      r = i / KC;
      c = i % KC;
    } else {
      r = i % R;
      c = i / R;
    }
  }

  __device__ __forceinline__ void load(const float* __restrict__ base, int64_t ld, int r0, int k0, int rlim,
                                       int klim, bool fast, int tid) {
#pragma unroll
    for (int p = 0; p < PER; ++p) {
      int r, c;
      item(tid + 256 * p, r, c);
      const int k = k0 + 4 * c, rr = r0 + r;
      if constexpr (KCONT) {
        if (fast) {
          v[p] = *reinterpret_cast<const f32x4*>(base + (int64_t)rr * ld + k);
        } else {
#pragma unroll
          for (int e = 0; e < 4; ++e) v[p][e] = (rr < rlim && k + e < klim) ? base[(int64_t)rr * ld + k + e] : 0.0f;
        }
      } else {
#pragma unroll
        for (int e = 0; e < 4; ++e)
          v[p][e] = (fast || (rr < rlim && k + e < klim)) ? base[rr + (int64_t)(k + e) * ld] : 0.0f;
      }
    }
  }

  __device__ __forceinline__ void store(uint4* __restrict__ dst, int tid) const {
#pragma unroll
    for (int p = 0; p < PER; ++p) {
      int r, c;
      item(tid + 256 * p, r, c);
      *reinterpret_cast<f32x4*>(&dst[lds_slot_rm<KC>(r, c)]) = v[p];
    }
  }
};

template <int BM_, int BN_, int BK_, int WM_, int WN_>
struct SgemmCfg {
  static constexpr int BM = BM_, BN = BN_, BK = BK_, WM = WM_, WN = WN_;
  static constexpr int KC = BK / 4;
  static_assert((BM / WM) * (BN / WN) == 4, "four wavefronts per workgroup");
  static_assert(BK % 16 == 0, "16-k groups");
  static constexpr int stage_chunks() { return (BM + BN) * KC; }
  static constexpr size_t lds_bytes() { return 2 * (size_t)stage_chunks() * 16; }
};
using SgemmCfgP1 = SgemmCfg<128, 128, 32, 64, 64>;

template <int OPA, int OPB, int OPC, class CFG>
__global__ void __launch_bounds__(256, 2) sgemm_kernel(const GemmParams p) {
  constexpr int BM = CFG::BM, BN = CFG::BN, BK = CFG::BK, WM = CFG::WM, WN = CFG::WN, KC = CFG::KC;
  constexpr int FM = WM / 32, FN = WN / 32, WAVES_M = BM / WM;
  constexpr int A_CH = BM * KC;
  constexpr int STAGE = CFG::stage_chunks();
  constexpr bool LANE_M = (OPC == OP_N);
  extern __shared__ uint4 lds[];

  TileF32x4<OPA == OP_T, BM, BK> ta;
  TileF32x4<OPB == OP_N, BN, BK> tb;

  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave % WAVES_M, wn = wave / WAVES_M;
  const int lrow = lane & 31, lh = lane >> 5;
  int tm, tn;
  tile_coords(p, blockIdx.x, tm, tn);
  const int m0 = tm * BM, n0 = tn * BN;
  const int ktiles = (p.k + BK - 1) / BK;
  const float* __restrict__ A = p.a;
  const float* __restrict__ B = static_cast<const float*>(p.b);
  const bool a_full = p.vec_a && (m0 + BM <= p.m);
  const bool b_full = p.vec_b && (n0 + BN <= p.n);

  f32x16 acc[FM][FN];
#pragma unroll
  for (int i = 0; i < FM; ++i)
#pragma unroll
    for (int j = 0; j < FN; ++j)
#pragma unroll
      for (int e = 0; e < 16; ++e) acc[i][j][e] = 0.0f;

  auto gload = [&](int kt) {
    const int k0 = kt * BK;
    const bool kf = (k0 + BK <= p.k);
    ta.load(A, p.lda, m0, k0, p.m, p.k, a_full && kf, tid);
    tb.load(B, p.ldb, n0, k0, p.n, p.k, b_full && kf, tid);
  };
  auto sstore = [&](int stage) {
    uint4* st = lds + stage * STAGE;
    ta.store(st, tid);
    tb.store(st + A_CH, tid);
  };
  auto compute = [&](const uint4* __restrict__ st) {
#pragma unroll
    for (int g = 0; g < BK / 16; ++g) {
      const int c = 4 * g + 2 * lh;
      f32x4 a[FM][2], b[FN][2];
#pragma unroll
      for (int i = 0; i < FM; ++i) {
        const int r = wm * WM + i * 32 + lrow;
        a[i][0] = *reinterpret_cast<const f32x4*>(&st[lds_slot_rm<KC>(r, c)]);
        a[i][1] = *reinterpret_cast<const f32x4*>(&st[lds_slot_rm<KC>(r, c + 1)]);
      }
#pragma unroll
      for (int j = 0; j < FN; ++j) {
        const int r = wn * WN + j * 32 + lrow;
        b[j][0] = *reinterpret_cast<const f32x4*>(&st[A_CH + lds_slot_rm<KC>(r, c)]);
        b[j][1] = *reinterpret_cast<const f32x4*>(&st[A_CH + lds_slot_rm<KC>(r, c + 1)]);
      }
#pragma unroll
      for (int s = 0; s < 8; ++s)
#pragma unroll
        for (int i = 0; i < FM; ++i)
#pragma unroll
          for (int j = 0; j < FN; ++j) {
            const float av = a[i][s >> 2][s & 3], bv = b[j][s >> 2][s & 3];
            if constexpr (LANE_M) acc[i][j] = mfma_f32(bv, av, acc[i][j]);
            else acc[i][j] = mfma_f32(av, bv, acc[i][j]);
          }
    }
  };

  if (ktiles > 0) {
    gload(0);
    sstore(0);
  }
  __syncthreads();
  for (int kt = 0; kt < ktiles; ++kt) {
    const int cur = kt & 1;
    if (kt + 1 < ktiles) gload(kt + 1);
    compute(lds + cur * STAGE);
    if (kt + 1 < ktiles) sstore(cur ^ 1);
    __syncthreads();
  }

  const float alpha = p.alpha, beta = p.beta;
#pragma unroll
  for (int i = 0; i < FM; ++i)
#pragma unroll
    for (int j = 0; j < FN; ++j)
#pragma unroll
      for (int e = 0; e < 16; ++e) {
        const int ridx = (e >> 2) * 8 + lh * 4 + (e & 3);
        int mm, nn;
        if constexpr (LANE_M) {
          mm = m0 + wm * WM + i * 32 + lrow;
          nn = n0 + wn * WN + j * 32 + ridx;
        } else {
          mm = m0 + wm * WM + i * 32 + ridx;
          nn = n0 + wn * WN + j * 32 + lrow;
        }
        if (mm < p.m && nn < p.n) {
          float* cp = p.c + (LANE_M ? (mm + nn * p.ldc) : (nn + mm * p.ldc));
          const float v = acc[i][j][e];
          *cp = (beta == 0.0f) ? alpha * v : alpha * v + beta * *cp;
        }
      }
}

}  // namespace shgemm_amd
