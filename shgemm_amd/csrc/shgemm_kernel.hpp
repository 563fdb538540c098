// Error-corrected SHGEMM kernel for gfx950 (CDNA4), plus the deterministic
// split-K reduction and the per-row exponent pass.
//
// Reference counterpart: src/shgemm.cu:36-119 (kernel entry),
// src/shgemm_pipeline_core.hpp:108-184 (k-loop), src/shgemm_core.hpp:48-75
// (per-warp MMA loop), src/wmmae_shgemm.hpp:11-125 (split + two products),
// src/dmem_accessor.hpp:187-356 (alpha/beta epilogue).
//
// Structure (one workgroup = 4 wavefronts of 64 lanes, one BM x BN tile of C):
//   * global -> registers: the FP32 operand and the FP16 operand of the next
//     k-tile are loaded while the current one is multiplied;
//   * registers -> LDS: the FP32 operand is split here, once per workgroup,
//     into FP16 hi and 2^11-scaled FP16 lo (or three BF16 parts); both are
//     written k-chunk-major (see lds_slot) so each MFMA operand is one
//     16-byte LDS read per lane whatever the global layout was;
//   * LDS -> MFMA: v_mfma_f32_32x32x16_{f16,bf16}, FP32 accumulation, the hi
//     and lo products in separate accumulators;
//   * epilogue: C = alpha * (hi + 2^-11 lo) + beta * C, or the slice partial
//     into the split-K workspace.
#pragma once
#include <type_traits>
#include "gemm_common.hpp"

namespace shgemm_amd {

// ---------------------------------------------------------------------------
// Operand tiles: R rows x BK k-values, NT threads, k cut into KC = BK/8
// chunks of 8 values (16 bytes once in FP16/BF16).  One item = (row r,
// chunk c): the thread loads those 8 values, converts them and writes one
// 16-byte LDS slot per part (lds_slot_h).
// Element (r, k) sits at base[r*ld + k] when KCONT (k contiguous) and at
// base[r + k*ld] otherwise.  The item -> thread maps are chosen for both
// coalesced global loads and conflict-free LDS stores:
//   KCONT: chunk fastest (KC lanes share one row segment; 16- or 32-byte loads);
//          a wavefront stores 64/KC whole rows.
//   !KCONT: row fastest (each load instruction reads 64 consecutive values of
//          one k column); a wavefront stores one chunk of 64 consecutive rows.
// ---------------------------------------------------------------------------
template <bool KCONT, int R, int BK, int NT>
struct ItemMap {
  static constexpr int KC = BK / 8;
  static constexpr int NITEMS = R * KC;
  static constexpr int PER = NITEMS / NT;
  static_assert(NITEMS % NT == 0 && PER >= 1, "tile must split evenly over the workgroup");
  static_assert(KCONT || R % 64 == 0, "row-fastest map needs whole wavefronts per chunk");
  // row-fastest with R dividing NT: every item of a thread lies in its row tid % R
  static constexpr bool ONE_ROW = !KCONT && NT % R == 0;
  __device__ __forceinline__ static void item(int i, int& r, int& c) {
    if constexpr (KCONT) {
      r = i / KC;
      c = i % KC;
    } else {
      r = i % R;
      c = i / R;
    }
  }
};

// FP32 operand (A of SHGEMM).
template <bool KCONT, int R, int BK, int NT>
struct TileF32 : ItemMap<KCONT, R, BK, NT> {
  using M = ItemMap<KCONT, R, BK, NT>;
  static constexpr int PER = M::PER, KC = M::KC;
  float v[PER][8];
  float scale[PER];  // per-row prescale 2^-e(row)

  __device__ __forceinline__ void init_scale(const int* row_exp, int r0, int rlim, int tid) {
#pragma unroll
    for (int p = 0; p < PER; ++p) {
      int r, c;
      M::item(tid + NT * p, r, c);
      const int rr = r0 + r;
      scale[p] = (row_exp != nullptr && rr < rlim) ? __builtin_ldexpf(1.0f, -row_exp[rr]) : 1.0f;
    }
  }

  // The largest |a| of item p's 8 values (a NaN is passed over unless all
  // are NaN): v_max3 with |.| operand modifiers, four instructions.
  __device__ __forceinline__ float item_absmax(int p) const {
    float mx = __builtin_fabsf(v[p][0]);
#pragma unroll
    for (int e = 1; e < 8; ++e) mx = __builtin_fmaxf(mx, __builtin_fabsf(v[p][e]));
    return mx;
  }

  __device__ __forceinline__ void load_fast(const float* __restrict__ base, int64_t ld, int r0, int k0, int tid) {
#pragma unroll
    for (int p = 0; p < PER; ++p) {
      int r, c;
      M::item(tid + NT * p, r, c);
      if constexpr (KCONT) {
        const float* ptr = base + (int64_t)(r0 + r) * ld + (k0 + 8 * c);
        const f32x4 x = load_unaligned<f32x4>(ptr);
        const f32x4 y = load_unaligned<f32x4>(ptr + 4);
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          v[p][e] = x[e];
          v[p][4 + e] = y[e];
        }
      } else {
        const float* ptr = base + (r0 + r) + (int64_t)(k0 + 8 * c) * ld;
#pragma unroll
        for (int e = 0; e < 8; ++e) v[p][e] = ptr[(int64_t)e * ld];
      }
    }
  }

  // Full-k-tile load with rows past rlim clamped to the last row: the values
  // only reach output rows that are never stored, and every access stays in
  // the operand.  Always issues the same instructions (PER*2 or PER*8 loads).
  // Addresses are a wave-uniform base (scalar registers) plus a 32-bit
  // per-lane offset, so the loads use the scalar-base addressing form; the
  // caller guarantees the tile's offsets fit in 32 bits.
  __device__ __forceinline__ void load_clamped(const float* __restrict__ base, int64_t ld, int r0, int k0, int rlim,
                                               int tid) {
    const int lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
#pragma unroll
    for (int p = 0; p < PER; ++p) {
      if constexpr (KCONT) {
        int r, c;
        M::item(tid + NT * p, r, c);
        const int rr = min(r0 + r, rlim - 1) - r0;
        const char* tile = as_global(uniform_ptr(reinterpret_cast<const char*>(base + (int64_t)r0 * ld + k0)));
        const uint32_t boff = 4u * ((uint32_t)rr * (uint32_t)ld + 8u * (uint32_t)c);
        const f32x4 x = *reinterpret_cast<const f32x4*>(tile + boff);
        const f32x4 y = *reinterpret_cast<const f32x4*>(tile + boff + 16);
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          v[p][e] = x[e];
          v[p][4 + e] = y[e];
        }
      } else {
        // item = 64 * wblk + lane with wblk wave-uniform: r = 64*(wblk % (R/64)) + lane, c = wblk / (R/64)
        const int wblk = wave + (NT / 64) * p;
        const int r = 64 * (wblk % (R / 64)) + lane;
        const int c = wblk / (R / 64);
        const uint32_t rbytes = 4u * (uint32_t)min(r0 + r, rlim - 1);
#pragma unroll
        for (int e = 0; e < 8; ++e) {
          const char* col = as_global(uniform_ptr(reinterpret_cast<const char*>(base + (int64_t)(k0 + 8 * c + e) * ld)));
          v[p][e] = *reinterpret_cast<const float*>(col + rbytes);
        }
      }
    }
  }
  static constexpr int kClampedLoads = KCONT ? 2 * PER : 8 * PER;

  __device__ __forceinline__ void load_safe(const float* __restrict__ base, int64_t ld, int r0, int k0,
                                            int rlim, int klim, int tid) {
#pragma unroll
    for (int p = 0; p < PER; ++p) {
      int r, c;
      M::item(tid + NT * p, r, c);
      const int rr = r0 + r;
#pragma unroll
      for (int e = 0; e < 8; ++e) {
        const int kk = k0 + 8 * c + e;
        const bool in = rr < rlim && kk < klim;
        v[p][e] = in ? (KCONT ? base[(int64_t)rr * ld + kk] : base[rr + (int64_t)kk * ld]) : 0.0f;
      }
    }
  }

  // FP16 hi / lo into two LDS arrays of R*KC slots.  SCALED: the row prescale
  // is applied first and lo is the unscaled residual (MODE_FP16S); otherwise
  // lo is the 2^11-scaled residual (MODE_FP16).
  template <bool SCALED>
  __device__ __forceinline__ void store_f16(uint4* __restrict__ hi, uint4* __restrict__ lo, int tid) const {
#pragma unroll
    for (int p = 0; p < PER; ++p) {
      int r, c;
      M::item(tid + NT * p, r, c);
      const int s = lds_slot_h<KC>(r, c);
      if constexpr (SCALED) {
        uint4 h, l;
        split2_f16_scaled(v[p][0], v[p][1], scale[p], h.x, l.x);
        split2_f16_scaled(v[p][2], v[p][3], scale[p], h.y, l.y);
        split2_f16_scaled(v[p][4], v[p][5], scale[p], h.z, l.z);
        split2_f16_scaled(v[p][6], v[p][7], scale[p], h.w, l.w);
        hi[s] = h;
        lo[s] = l;
      } else {
        f16x8 h, l;
#pragma unroll
        for (int e = 0; e < 8; ++e) {
          _Float16 hh, ll;
          split_f16(v[p][e], hh, ll);
          h[e] = hh;
          l[e] = ll;
        }
        *reinterpret_cast<f16x8*>(&hi[s]) = h;
        *reinterpret_cast<f16x8*>(&lo[s]) = l;
      }
    }
  }

  // Three-part BF16 split into three LDS arrays.
  __device__ __forceinline__ void store_bf16(uint4* __restrict__ d0, uint4* __restrict__ d1,
                                             uint4* __restrict__ d2, int tid) const {
#pragma unroll
    for (int p = 0; p < PER; ++p) {
      int r, c;
      M::item(tid + NT * p, r, c);
      bf16x8 x0, x1, x2;
#pragma unroll
      for (int e = 0; e < 8; ++e) {
        __bf16 a0, a1, a2;
        split_bf16x3(v[p][e], a0, a1, a2);
        x0[e] = a0;
        x1[e] = a1;
        x2[e] = a2;
      }
      const int s = lds_slot_h<KC>(r, c);
      *reinterpret_cast<bf16x8*>(&d0[s]) = x0;
      *reinterpret_cast<bf16x8*>(&d1[s]) = x1;
      *reinterpret_cast<bf16x8*>(&d2[s]) = x2;
    }
  }
};

// FP16 operand (B of SHGEMM), raw binary16 bits.
//   KCONT: one 16-byte load per item.
//   !KCONT: the item covers 2 rows; eight 4-byte loads (2 rows at each k),
//           unpacked into two 16-byte slots.
template <bool KCONT, int R, int BK, int NT>
struct TileF16 {
  static constexpr int KC = BK / 8;
  static constexpr int ROWS_PER_ITEM = KCONT ? 1 : 2;
  static constexpr int NITEMS = R * KC / ROWS_PER_ITEM;
  static constexpr int PER = NITEMS / NT;
  static_assert(NITEMS % NT == 0 && PER >= 1, "tile must split evenly over the workgroup");
  // KCONT: v[p] holds 8 halves of one row; !KCONT: v[p][e] holds rows r, r+1 at k0+e
  u32x4 vk[KCONT ? PER : 1];
  uint32_t vr[KCONT ? 1 : PER][8];

  __device__ __forceinline__ static void item(int i, int& r, int& c) {
    if constexpr (KCONT) {
      r = i / KC;
      c = i % KC;
    } else {
      r = 2 * (i % (R / 2));
      c = i / (R / 2);
    }
  }

  __device__ __forceinline__ void load_fast(const uint16_t* __restrict__ base, int64_t ld, int r0, int k0, int tid) {
#pragma unroll
    for (int p = 0; p < PER; ++p) {
      int r, c;
      item(tid + NT * p, r, c);
      if constexpr (KCONT) {
        vk[p] = load_unaligned<u32x4>(base + (int64_t)(r0 + r) * ld + (k0 + 8 * c));
      } else {
        const uint16_t* ptr = base + (r0 + r) + (int64_t)(k0 + 8 * c) * ld;
#pragma unroll
        for (int e = 0; e < 8; ++e) vr[p][e] = load_unaligned<uint32_t>(ptr + (int64_t)e * ld);
      }
    }
  }

  __device__ __forceinline__ void load_safe(const uint16_t* __restrict__ base, int64_t ld, int r0, int k0,
                                            int rlim, int klim, int tid) {
#pragma unroll
    for (int p = 0; p < PER; ++p) {
      int r, c;
      item(tid + NT * p, r, c);
      if constexpr (KCONT) {
        const int rr = r0 + r;
        uint32_t w[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const int kk = k0 + 8 * c + 2 * q;
          const uint32_t x0 = (rr < rlim && kk < klim) ? base[(int64_t)rr * ld + kk] : 0u;
          const uint32_t x1 = (rr < rlim && kk + 1 < klim) ? base[(int64_t)rr * ld + kk + 1] : 0u;
          w[q] = x0 | (x1 << 16);
        }
        vk[p] = u32x4{w[0], w[1], w[2], w[3]};
      } else {
#pragma unroll
        for (int e = 0; e < 8; ++e) {
          const int kk = k0 + 8 * c + e, rr = r0 + r;
          const uint32_t x0 = (rr < rlim && kk < klim) ? base[rr + (int64_t)kk * ld] : 0u;
          const uint32_t x1 = (rr + 1 < rlim && kk < klim) ? base[rr + 1 + (int64_t)kk * ld] : 0u;
          vr[p][e] = x0 | (x1 << 16);
        }
      }
    }
  }

  // 8 halves of row r + h (h = 0: low halves, 1: high halves) of a !KCONT item
  __device__ __forceinline__ u32x4 gather_row(int p, int h) const {
    uint32_t w[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const uint32_t a = (vr[p][2 * q] >> (16 * h)) & 0xffffu, b = (vr[p][2 * q + 1] >> (16 * h)) & 0xffffu;
      w[q] = a | (b << 16);
    }
    return u32x4{w[0], w[1], w[2], w[3]};
  }

  __device__ __forceinline__ void store_raw(uint4* __restrict__ dst, int tid) const {
#pragma unroll
    for (int p = 0; p < PER; ++p) {
      int r, c;
      item(tid + NT * p, r, c);
      if constexpr (KCONT) {
        *reinterpret_cast<u32x4*>(&dst[lds_slot_h<KC>(r, c)]) = vk[p];
      } else {
#pragma unroll
        for (int h = 0; h < 2; ++h) *reinterpret_cast<u32x4*>(&dst[lds_slot_h<KC>(r + h, c)]) = gather_row(p, h);
      }
    }
  }

  __device__ __forceinline__ static void to_bf16x2(const u32x4& w, bf16x8& x0, bf16x8& x1) {
#pragma unroll
    for (int e = 0; e < 8; ++e) {
      const uint32_t hbits = (w[e >> 1] >> (16 * (e & 1))) & 0xffffu;
      __bf16 b0, b1;
      split_bf16x2(bits_to_f16((uint16_t)hbits), b0, b1);
      x0[e] = b0;
      x1[e] = b1;
    }
  }

  __device__ __forceinline__ void store_bf16(uint4* __restrict__ d0, uint4* __restrict__ d1, int tid) const {
#pragma unroll
    for (int p = 0; p < PER; ++p) {
      int r, c;
      item(tid + NT * p, r, c);
#pragma unroll
      for (int h = 0; h < ROWS_PER_ITEM; ++h) {
        bf16x8 x0, x1;
        if constexpr (KCONT) to_bf16x2(vk[p], x0, x1);
        else to_bf16x2(gather_row(p, h), x0, x1);
        const int s = lds_slot_h<KC>(r + h, c);
        *reinterpret_cast<bf16x8*>(&d0[s]) = x0;
        *reinterpret_cast<bf16x8*>(&d1[s]) = x1;
      }
    }
  }
};

// ---------------------------------------------------------------------------
// k-contiguous FP16 operand loaded straight into LDS (global_load_lds_dwordx4:
// lane l of an instruction writes 16 bytes at the instruction's LDS base +
// 16*l, measured on gfx950).  Instruction q of wavefront w fills physical
// slots [(q*WAVES + w)*64, +64); the lane at physical slot (r, cp) fetches
// logical chunk cp ^ key(r), the inverse of lds_slot_h, so the LDS image
// equals what TileF16::store_raw writes.  Rows past rlim are clamped (their
// values only reach columns that are never stored); the caller guarantees a
// full k-tile, 16-byte-aligned rows and ld % 8 == 0.
// ---------------------------------------------------------------------------
template <int R, int BK, int NT>
struct DirectF16 {
  static constexpr int KC = BK / 8;
  static constexpr int WAVES = NT / 64;
  static constexpr int PER = R * KC / NT;  // instructions per thread per k-tile
  static_assert((R * KC) % NT == 0 && PER >= 1, "tile must split into whole wavefront instructions");
  __device__ __forceinline__ static void load(const uint16_t* __restrict__ base, int64_t ld, int r0, int k0, int rlim,
                                              uint4* __restrict__ dst, int tid) {
    const int lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const char* tile = reinterpret_cast<const char*>(base + (int64_t)r0 * ld + k0);  // uniform; offsets fit 32 bits
#pragma unroll
    for (int q = 0; q < PER; ++q) {
      const int blk = q * WAVES + wave;
      const int slot = blk * 64 + lane;
      const int r = slot / KC, cp = slot % KC;
      const int c = cp ^ lds_key_h<KC>(r);
      const uint32_t rr = (uint32_t)(min(r0 + r, rlim - 1) - r0);
      const uint32_t boff = 2u * (rr * (uint32_t)ld + 8u * (uint32_t)c);
      __builtin_amdgcn_global_load_lds(tile + boff, reinterpret_cast<uint32_t*>(dst + blk * 64), 16, 0, 0);
    }
  }
};

// s_waitcnt immediate for "at most n vector-memory operations outstanding",
// other counters left unconstrained (gfx9 encoding).
__device__ __forceinline__ constexpr int vmcnt_imm(int n) {
  return (n & 0xF) | ((n >> 4) << 14) | (0x7 << 4) | (0xF << 8);
}

// One wavefront instruction of 16-byte global -> LDS copies: lane l's 16
// bytes at src land at LDS address dst + 16 * l (dst wavefront-uniform).
__device__ __forceinline__ void lds_dma16(const char* src, uint4* dst) {
  __builtin_amdgcn_global_load_lds(src, reinterpret_cast<uint32_t*>(dst), 16, 0, 0);
}

// Workgroup barrier for an LDS ring filled by global_load_lds: wait until at
// most N of this thread's vector-memory operations are outstanding (the
// older ones, whose LDS writes have then landed) and its LDS reads have
// completed, then s_barrier.  No memory fence: a workgroup release fence
// would also wait for the N newest loads, the ones meant to stay in flight.
// The asm statement's memory clobber keeps the compiler from moving memory
// accesses across the barrier.
template <int N>
__device__ __forceinline__ void lds_ring_barrier() {
  static_assert(N >= 0 && N < 64, "vmcnt range");
  __builtin_amdgcn_s_waitcnt((N & 0xF) | ((N >> 4) << 14) | (0x7 << 4));  // vmcnt(N) lgkmcnt(0)
  asm volatile("s_barrier" ::: "memory");
}

// Running row exponents of one output tile (LDS, after the two stages):
// by k-tile parity the maxima |a| (bits, LDS atomic max) of the chunks that
// exceed their row's exponent and the index of the last k-tile with such a
// chunk; per row the current exponent and the accumulator factor
// 2^(old - new) of the last rise.
template <int BM>
struct RunState {
  uint32_t rmax[2][BM];
  int erow[BM];
  float fac[BM];
  int flag[2];
  int pad[14];
};

// ---------------------------------------------------------------------------
// Tile configuration
// ---------------------------------------------------------------------------
template <int BM_, int BN_, int BK_, int WM_, int WN_, int PF_, int OCC_, bool IL_ = false, bool RUN_ = true>
struct TileCfg {
  static constexpr int BM = BM_, BN = BN_, BK = BK_, WM = WM_, WN = WN_;
  // PF: k-tiles of global loads in flight ahead of the one being multiplied
  // (register sets); OCC: minimum wavefronts per SIMD the register budget
  // must allow (the second __launch_bounds__ argument in HIP).
  // IL (with PF == 2, fp16 one-accumulator mode): the split and LDS store of
  // tile kt+1 are interleaved with the MFMAs of tile kt, and the barrier is
  // skewed ahead of the last k-step's MFMAs.
  static constexpr int PF = PF_, OCC = OCC_;
  static constexpr bool IL = IL_;
  static_assert(!IL || PF == 2, "interleaving needs the next tile already in registers");
  static constexpr int KC = BK / 8;
  static constexpr int WAVES = (BM / WM) * (BN / WN);
  static constexpr int NT = 64 * WAVES;  // threads per workgroup
  static_assert(WAVES == 4 || WAVES == 8 || WAVES == 16, "4, 8 or 16 wavefronts per workgroup");
  static_assert(WM % 32 == 0 && WN % 32 == 0 && BK % 16 == 0, "MFMA 32x32x16 granularity");
  static_assert(PF == 1 || PF == 2, "one or two k-tiles of prefetch");
  template <int MODE>
  static constexpr int stage_chunks() {
    return (MODE == MODE_BF16 ? 3 : 2) * BM * KC + (MODE == MODE_BF16 ? 2 : 1) * BN * KC;
  }
  // Prescaled fp16 on wavefront tiles up to 64 rows, without the skewed
  // schedule: the tile keeps a running exponent per row itself (RunState in
  // LDS after the two stages) instead of reading exponents from a separate
  // pass over A.  (The 128-row wavefront tiles hold their accumulators in
  // AGPRs and have no registers left for it.)  RUN_ = false keeps whole-row
  // exponents on a tile that could run them.
  template <int MODE>
  static constexpr bool running() {
    return RUN_ && MODE == MODE_FP16S && !IL && WM <= 64;
  }
  template <int MODE>
  static constexpr size_t lds_bytes() {
    return 2 * (size_t)stage_chunks<MODE>() * 16 + (running<MODE>() ? sizeof(RunState<BM>) : 0);
  }
};

// Tile configurations per split mode (the reference's two kernel levels,
// include/shgemm/shgemm.hpp:30-34, plus the split-K tile):
//   P1: large tiles; P0: small tiles for shapes with too few large tiles;
//   SK: split-K (tall-skinny shapes have small m and n).
// The prescaled mode has one accumulator: its large tile runs 8 wavefronts of
// 64 x 64 with two workgroups per CU (4 wavefronts per SIMD to overlap the
// split/store phase of one with the MFMA phase of another).  The
// two-accumulator modes use 4 wavefronts of 64 x 64.  Chosen by the tile
// sweep at m = n = k = 8192 and 16384.
// P2 (one-accumulator modes): a 256 x 256 tile of four 128 x 128 wavefronts,
// one per SIMD, accumulators in AGPRs -- half the LDS operand traffic per
// MFMA of the 64 x 64 wavefront tiles; the tile sweep measured it fastest at
// m = n = k = 8192 and 16384.  The two-accumulator mode has no room for it.
template <int MODE>
struct ModeCfg {
  using P1 = TileCfg<128, 128, 32, 64, 64, 1, 2>;
  using P0 = TileCfg<64, 64, 64, 32, 32, 1, 2>;
  using P0SK = P0;  // the small tile of the k-sliced plan
  using P0W = P0;   // the small tile on whole-row exponents (prescaled mode only)
  using SK = TileCfg<128, 128, 32, 64, 64, 1, 2>;
  using P2 = TileCfg<256, 256, 32, 128, 128, 1, 1>;
  // The tf32 pre-split ring: eight 64 x 128 wavefronts, two per SIMD, as in
  // the prescaled fp16 mode (16384^3 per call 308 against 283 TFLOP/s for
  // the four 128 x 128 wavefronts of P2, 8192^3 301 against 276; one box,
  // interleaved).
  using PS = TileCfg<256, 256, 32, 64, 128, 2, 2, true>;
  // the same ring on 128 x 128 tiles (shapes with fewer 256 x 256 tiles than CUs)
  using PS128 = TileCfg<128, 128, 32, 64, 64, 2, 2, true>;
  static constexpr bool kHasP2 = (MODE == MODE_BF16);
};
// The prescaled P2 keeps two k-tiles of loads in registers, splits tile kt+1
// into LDS between the MFMAs of tile kt and places the barrier before the
// last k-step's MFMAs (tile sweep: +12% at 4096^3, +7% at 8192^3 and 16384^3
// against the one-deep form).
template <>
struct ModeCfg<MODE_FP16S> {
  using P1 = TileCfg<128, 256, 32, 64, 64, 1, 4>;  // 8 wavefronts, 2 workgroups (4 wavefronts) per SIMD
  // The small tile keeps two k-tiles of loads in flight (512^3 to 4096^3 on
  // this tile: 1-3% faster, level at 1024^3); k-sliced it keeps one (the
  // 128 x 128 x 2^20 slices ran 8% slower with two).  Same arithmetic, same
  // bits.
  using P0 = TileCfg<64, 64, 64, 32, 32, 2, 3>;
  using P0SK = TileCfg<64, 64, 64, 32, 32, 1, 4>;
  // The small tile reading whole-row exponents from the row pass (plan 9):
  // the running check costs about 0.2 us per k-tile when each SIMD holds one
  // wavefront, which outweighs the pass for long k.
  using P0W = TileCfg<64, 64, 64, 32, 32, 2, 3, false, false>;
  using SK = TileCfg<128, 128, 32, 64, 64, 1, 3>;
  // The deep k-sliced tile (plan 10): 128 x 128 x 64, eight wavefronts of
  // 64 x 32, two k-tiles of loads in flight, one workgroup per CU.  For
  // bandwidth-bound tall-skinny products each workgroup loads every byte of
  // its slice once (the whole C tile) and keeps 96 KiB of loads in flight;
  // 128 x 128 x 2^20 at 256 slices: 165 us against 216 us for the small
  // tile's k-sliced plan at 128 slices.
  using SKD = TileCfg<128, 128, 64, 64, 32, 2, 2>;
  // The large tile with four 64 x 128 wavefronts (plan 11, k-sliced): twice
  // the MFMA work per fragment read of the eight 64 x 64 wavefronts, two
  // workgroups per CU.  Same per-element k order and exponents as P1, so
  // the same bits; the sketch shape 16384 x 256 x 16384 in 4 slices ran
  // 335 us against 412 us for P1-SK on one box.
  using P1W = TileCfg<128, 256, 32, 64, 128, 1, 2>;
  using P2 = TileCfg<256, 256, 32, 128, 128, 2, 1, true>;
  // The pre-split plan's tile (plan 8): 256 x 256 as eight 64 x 128
  // wavefronts, two per SIMD, so one wavefront's barrier and LDS waits are
  // covered by the other's MFMAs (the four 128 x 128 wavefronts of P2 run one
  // per SIMD); the same fragments per k-step and the same bits as P2.  (PF
  // and IL describe the register-split P2 form; the pre-split kernel stages
  // its loads in an LDS ring.)
  using PS = TileCfg<256, 256, 32, 64, 128, 2, 2, true>;
  // the same ring on 128 x 128 tiles (shapes with fewer 256 x 256 tiles than CUs)
  using PS128 = TileCfg<128, 128, 32, 64, 64, 2, 2, true>;
  static constexpr bool kHasP2 = true;
};

// One 16-deep k-step (s) of MFMA work for one wavefront: fragments from the
// A parts at ast (+ q*BM*KC slots per part) and the B parts at bst (+ q*BN*KC).
template <int MODE, bool LANE_M, int NACC, int FM, int FN, int KC, int BM, int BN, int WM, int WN>
__device__ __forceinline__ void mma_kstep(f32x16 (&acc)[NACC][FM][FN], const uint4* __restrict__ ast,
                                          const uint4* __restrict__ bst, int s, int wm, int wn, int lrow, int lh) {
  constexpr int A_CH = BM * KC, B_CH = BN * KC;
  {
    const int c = 2 * s + lh;
    if constexpr (MODE != MODE_BF16) {
      f16x8 ah[FM], al[FM], bb[FN];
#pragma unroll
      for (int i = 0; i < FM; ++i) {
        const int sl = lds_slot_h<KC>(wm * WM + i * 32 + lrow, c);
        ah[i] = *reinterpret_cast<const f16x8*>(&ast[sl]);
        al[i] = *reinterpret_cast<const f16x8*>(&ast[A_CH + sl]);
      }
#pragma unroll
      for (int j = 0; j < FN; ++j)
        bb[j] = *reinterpret_cast<const f16x8*>(&bst[lds_slot_h<KC>(wn * WN + j * 32 + lrow, c)]);
#pragma unroll
      for (int i = 0; i < FM; ++i)
#pragma unroll
        for (int j = 0; j < FN; ++j) {
          if constexpr (LANE_M) {
            acc[0][i][j] = mfma_f16(bb[j], ah[i], acc[0][i][j]);
            acc[NACC - 1][i][j] = mfma_f16(bb[j], al[i], acc[NACC - 1][i][j]);
          } else {
            acc[0][i][j] = mfma_f16(ah[i], bb[j], acc[0][i][j]);
            acc[NACC - 1][i][j] = mfma_f16(al[i], bb[j], acc[NACC - 1][i][j]);
          }
        }
    } else {
      bf16x8 a0[FM], a1[FM], a2[FM], b0[FN], b1[FN];
#pragma unroll
      for (int i = 0; i < FM; ++i) {
        const int sl = lds_slot_h<KC>(wm * WM + i * 32 + lrow, c);
        a0[i] = *reinterpret_cast<const bf16x8*>(&ast[sl]);
        a1[i] = *reinterpret_cast<const bf16x8*>(&ast[A_CH + sl]);
        a2[i] = *reinterpret_cast<const bf16x8*>(&ast[2 * A_CH + sl]);
      }
#pragma unroll
      for (int j = 0; j < FN; ++j) {
        const int sl = lds_slot_h<KC>(wn * WN + j * 32 + lrow, c);
        b0[j] = *reinterpret_cast<const bf16x8*>(&bst[sl]);
        b1[j] = *reinterpret_cast<const bf16x8*>(&bst[B_CH + sl]);
      }
#pragma unroll
      for (int i = 0; i < FM; ++i)
#pragma unroll
        for (int j = 0; j < FN; ++j) {
          // a0*b0 + a0*b1 + a1*b0 + a1*b1 + a2*b0 (a2*b1 ~ 2^-24 dropped), all in
          // one accumulator: each MFMA adds its exact product sum with one
          // round-to-nearest-even.
          f32x16& c = acc[0][i][j];
          if constexpr (LANE_M) {
            c = mfma_bf16(b0[j], a0[i], c);
            c = mfma_bf16(b1[j], a0[i], c);
            c = mfma_bf16(b0[j], a1[i], c);
            c = mfma_bf16(b1[j], a1[i], c);
            c = mfma_bf16(b0[j], a2[i], c);
          } else {
            c = mfma_bf16(a0[i], b0[j], c);
            c = mfma_bf16(a0[i], b1[j], c);
            c = mfma_bf16(a1[i], b0[j], c);
            c = mfma_bf16(a1[i], b1[j], c);
            c = mfma_bf16(a2[i], b0[j], c);
          }
        }
    }
  }
}

// Fragments of one 16-deep k-step of the fp16 one-accumulator mode, held in
// registers so that a k-step's reads can be issued ahead of its MFMAs.
template <int FM, int FN>
struct FragF16 {
  f16x8 ah[FM], al[FM], bb[FN];
};
template <int FM, int FN, int KC, int BM, int WM, int WN>
__device__ __forceinline__ void frag_load(FragF16<FM, FN>& f, const uint4* __restrict__ ast,
                                          const uint4* __restrict__ bst, int s, int wm, int wn, int lrow, int lh) {
  constexpr int A_CH = BM * KC;
  const int c = 2 * s + lh;
#pragma unroll
  for (int i = 0; i < FM; ++i) {
    const int sl = lds_slot_h<KC>(wm * WM + i * 32 + lrow, c);
    f.ah[i] = *reinterpret_cast<const f16x8*>(&ast[sl]);
    f.al[i] = *reinterpret_cast<const f16x8*>(&ast[A_CH + sl]);
  }
#pragma unroll
  for (int j = 0; j < FN; ++j) f.bb[j] = *reinterpret_cast<const f16x8*>(&bst[lds_slot_h<KC>(wn * WN + j * 32 + lrow, c)]);
}
// All hi products of the k-step, then all lo products: each accumulator
// still takes hi before lo (the same bits), and FM*FN independent MFMAs
// separate the two that write one accumulator.
template <bool LANE_M, int FM, int FN>
__device__ __forceinline__ void frag_mma(f32x16 (&acc)[1][FM][FN], const FragF16<FM, FN>& f) {
#pragma unroll
  for (int h = 0; h < 2; ++h)
#pragma unroll
    for (int i = 0; i < FM; ++i)
#pragma unroll
      for (int j = 0; j < FN; ++j) {
        const f16x8& a = h == 0 ? f.ah[i] : f.al[i];
        if constexpr (LANE_M) acc[0][i][j] = mfma_f16(f.bb[j], a, acc[0][i][j]);
        else acc[0][i][j] = mfma_f16(a, f.bb[j], acc[0][i][j]);
      }
}

// One BK-deep k-tile of MFMA work for one wavefront.
template <int MODE, bool LANE_M, int NACC, int FM, int FN, int KC, int BM, int BN, int WM, int WN>
__device__ __forceinline__ void mma_ktile(f32x16 (&acc)[NACC][FM][FN], const uint4* __restrict__ ast,
                                          const uint4* __restrict__ bst, int wm, int wn, int lrow, int lh) {
#pragma unroll
  for (int s = 0; s < KC / 2; ++s)
    mma_kstep<MODE, LANE_M, NACC, FM, FN, KC, BM, BN, WM, WN>(acc, ast, bst, s, wm, wn, lrow, lh);
}

// Epilogue.  Lane/register -> (m, n) of the 32x32 MFMA output block: the lane
// index gives the column, register e gives row (e/4)*8 + (lane/32)*4 + e%4.
// TO_WS: the tile is one k-slice's partial; it goes unscaled by alpha/beta to
// `part` (m*n floats in C's storage order, leading dimension m or n).
template <int MODE, bool LANE_M, bool TO_WS, int NACC, int FM, int FN, int WM, int WN, bool ROWEXP = true>
__device__ __forceinline__ void store_ctile(const f32x16 (&acc)[NACC][FM][FN], const GemmParams& p, int m0, int n0,
                                            int wm, int wn, int lrow, int lh, float* __restrict__ part) {
  // ROWEXP = false: the accumulators already carry their rows' exponents
  constexpr bool SCALED = (MODE == MODE_FP16S) && ROWEXP;
  constexpr float lo_w = 1.0f / 2048.0f;
  const float alpha = p.alpha, beta = p.beta;
#pragma unroll
  for (int i = 0; i < FM; ++i) {
#pragma unroll
    for (int j = 0; j < FN; ++j) {
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
          float v = acc[0][i][j][e];
          if constexpr (NACC == 2) v = __builtin_fmaf(acc[1][i][j][e], lo_w, v);
          if constexpr (SCALED) v = __builtin_ldexpf(v, p.row_exp[mm]);
          if constexpr (TO_WS) {
            part[LANE_M ? (mm + (int64_t)nn * p.m) : (nn + (int64_t)mm * p.n)] = v;
          } else {
            float* cp = p.c + (LANE_M ? (mm + nn * p.ldc) : (nn + mm * p.ldc));
            *cp = (beta == 0.0f) ? alpha * v : alpha * v + beta * *cp;
          }
        }
      }
    }
  }
}

// ---------------------------------------------------------------------------
// One output tile (tm, tn) over k in [kbeg, kend) by one workgroup.
//   OPA, OPB: operand storage (OP_N / OP_T), OPC: C storage.
//   MODE: split scheme (MODE_FP16S prescales A rows by 2^-row_exp).
//   SPLIT: 0 writes C (alpha, beta applied); 1 writes the k-slice partial to
//   `part`; 2 decides at run time (part != nullptr).
// Leaves LDS free for the next tile: the last k-tile ends with a barrier.
// ---------------------------------------------------------------------------
template <int OPA, int OPB, int OPC, int MODE, class CFG, int SPLIT>
__device__ __forceinline__ void shgemm_tile(const GemmParams& p, int tm, int tn, int kbeg, int kend,
                                            float* __restrict__ part, uint4* __restrict__ lds) {
  constexpr bool SCALED = (MODE == MODE_FP16S);
  constexpr int NACC = (MODE == MODE_FP16) ? 2 : 1;  // separate hi / lo accumulators
  constexpr int BM = CFG::BM, BN = CFG::BN, BK = CFG::BK, WM = CFG::WM, WN = CFG::WN, KC = CFG::KC;
  constexpr int FM = WM / 32, FN = WN / 32, WAVES_M = BM / WM;
  constexpr int NA = MODE == MODE_BF16 ? 3 : 2;
  constexpr int A_CH = BM * KC, B_CH = BN * KC;
  constexpr int STAGE = CFG::template stage_chunks<MODE>();
  // With C column-major the MFMA's lane index runs along m (coalesced C
  // stores); with C transposed it runs along n.
  constexpr bool LANE_M = (OPC == OP_N);

  constexpr int PF = CFG::PF;
  using S0 = std::integral_constant<int, 0>;
  using S1 = std::integral_constant<int, PF - 1>;
  constexpr int NT = CFG::NT;
  TileF32<OPA == OP_T, BM, BK, NT> ta[PF];
  TileF16<OPB == OP_N, BN, BK, NT> tb[PF];

  int tid = threadIdx.x;
  // In the persistent grouped kernel (SPLIT == 2) this body runs once per
  // work unit: an opaque copy of the thread index keeps the compiler from
  // hoisting every per-lane offset out of the unit loop, where they would
  // stay live (and spill) across all units instead of one tile.
  if constexpr (SPLIT == 2) asm volatile("" : "+v"(tid));
  // the wavefront index is uniform: scalar registers for it and what derives from it
  const int lane = tid & 63, wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int wm = wave % WAVES_M, wn = wave / WAVES_M;
  const int lrow = lane & 31, lh = lane >> 5;
  const int m0 = tm * BM, n0 = tn * BN;
  const int ktiles = kend > kbeg ? (kend - kbeg + BK - 1) / BK : 0;
  const float* __restrict__ A = p.a;
  const uint16_t* __restrict__ B = static_cast<const uint16_t*>(p.b);
  // unchecked loads: op_n A uses 4-byte loads, the other forms 16- or 4-byte
  // loads at any ld and base alignment (load_unaligned)
  const bool a_full = m0 + BM <= p.m;
  const bool b_full = n0 + BN <= p.n;

  // RUN: running row exponents kept by the tile (no separate pass over A).
  // Every thread keeps the current exponent e of its items' rows (the same
  // in every thread holding a chunk of the row) as the threshold
  // ethr = rise_threshold(e) and splits each k-tile with it.  A chunk whose
  // largest |a| reaches the threshold (one compare: its exponent would be
  // larger than e) merges that maximum into rmax[t&1] and records t in
  // flag[t&1] before the barrier that ends the k-tile's stores; a thread
  // whose items share one row tests their common maximum once.  After the
  // barrier the workgroup reads the flag (uniform) and, only when it is set,
  // gives each row the exponent max(e, that of rmax), splits the k-tile again from the registers into
  // the same stage, records the factor 2^(old - new) and waits at one more
  // barrier.  Before the k-tile's MFMAs the accumulators of a tile that rose
  // are scaled by their rows' factors (exact powers of two).  Rows start at
  // the floor -126 (all-zero rows keep it).  flag[t&1] is written again two
  // k-tiles later, after every read of it; rmax[t&1] is cleared after the
  // extra barrier, before the next barrier.
  constexpr bool RUN = CFG::template running<MODE>();
  constexpr int kFloor = -126;
  using RS = RunState<BM>;
  using MI = ItemMap<OPA == OP_T, BM, BK, NT>;
  RS* rs = nullptr;
  uint32_t ethr[MI::PER];
  if constexpr (RUN) {
    rs = reinterpret_cast<RS*>(lds + 2 * STAGE);
    for (int r = tid; r < BM; r += NT) {
      rs->rmax[0][r] = 0u;
      rs->rmax[1][r] = 0u;
      rs->erow[r] = kFloor;
    }
    if (tid < 2) rs->flag[tid] = -1;
#pragma unroll
    for (int q = 0; q < MI::PER; ++q) {
      ethr[q] = rise_threshold(kFloor);
#pragma unroll
      for (int i = 0; i < PF; ++i) ta[i].scale[q] = __builtin_ldexpf(1.0f, -kFloor);
    }
    __syncthreads();
  } else if constexpr (SCALED) {
#pragma unroll
    for (int i = 0; i < PF; ++i) ta[i].init_scale(p.row_exp, m0, p.m, tid);
  }

  f32x16 acc[NACC][FM][FN];
#pragma unroll
  for (int h = 0; h < NACC; ++h)
#pragma unroll
    for (int i = 0; i < FM; ++i)
#pragma unroll
      for (int j = 0; j < FN; ++j)
#pragma unroll
        for (int e = 0; e < 16; ++e) acc[h][i][j][e] = 0.0f;

  // `set` is a std::integral_constant so the staging registers are never
  // indexed dynamically (which would force them into scratch memory).
  auto gload = [&](auto set_c, int kt) {
    constexpr int set = decltype(set_c)::value;
    const int k0 = kbeg + kt * BK;
    const bool kf = (k0 + BK <= kend);
    if (a_full && kf) ta[set].load_fast(A, p.lda, m0, k0, tid);
    else ta[set].load_safe(A, p.lda, m0, k0, p.m, kend, tid);
    if (b_full && kf) tb[set].load_fast(B, p.ldb, n0, k0, tid);
    else tb[set].load_safe(B, p.ldb, n0, k0, p.n, kend, tid);
  };
  auto sstore = [&](auto set_c, int stage) {
    constexpr int set = decltype(set_c)::value;
    uint4* st = lds + stage * STAGE;
    if constexpr (MODE != MODE_BF16) {
      ta[set].template store_f16<SCALED>(st, st + A_CH, tid);
      tb[set].store_raw(st + NA * A_CH, tid);
    } else {
      ta[set].store_bf16(st, st + A_CH, st + 2 * A_CH, tid);
      tb[set].store_bf16(st + NA * A_CH, st + NA * A_CH + B_CH, tid);
    }
  };
  // RUN, after k-tile t (in register set `set_c`) is stored at the current
  // exponents and before the barrier that ends the stores: flag the chunks
  // that need a larger exponent (par_c == t & 1, an int or an
  // integral_constant).  A k-tile that rises is split and stored again, so
  // the store need not wait for the check (nor the split of the first
  // loaded values for the last loads).
  auto run_check = [&](auto set_c, int t, auto par_c) {
    constexpr int set = decltype(set_c)::value;
    const int par = static_cast<int>(par_c);
    if constexpr (MI::ONE_ROW) {
      float mx = ta[set].item_absmax(0);
#pragma unroll
      for (int q = 1; q < MI::PER; ++q) mx = __builtin_fmaxf(mx, ta[set].item_absmax(q));
      const uint32_t b = __builtin_bit_cast(uint32_t, mx);
      if (b >= ethr[0]) {
        int r, c;
        MI::item(tid, r, c);
        atomicMax(&rs->rmax[par][r], b);
        rs->flag[par] = t;
      }
    } else {
#pragma unroll
      for (int q = 0; q < MI::PER; ++q) {
        const uint32_t b = __builtin_bit_cast(uint32_t, ta[set].item_absmax(q));
        if (b >= ethr[q]) {
          int r, c;
          MI::item(tid + NT * q, r, c);
          atomicMax(&rs->rmax[par][r], b);
          rs->flag[par] = t;
        }
      }
    }
  };
  // RUN, after the barrier that ends k-tile t's stores (from register set
  // `set_c`) into `stage`: whether some row's exponent rose at t (uniform);
  // if so the new exponents (every register set's scales), the factors and
  // the k-tile split again (contains a barrier)
  auto run_settle = [&](auto set_c, int t, auto par_c, int stage) -> bool {
    const int par = static_cast<int>(par_c);
    if (__builtin_amdgcn_readfirstlane(rs->flag[par]) != t) return false;
#pragma unroll
    for (int q = 0; q < MI::PER; ++q) {
      int r, c;
      MI::item(tid + NT * q, r, c);
      const uint32_t b = rs->rmax[par][r];
      const int e0 = threshold_exponent(ethr[q]);
      const int et = b == 0u ? kFloor : exponent_of_max(b);
      const int en = et > e0 ? et : e0;
      if (c == 0) {
        rs->fac[r] = __builtin_ldexpf(1.0f, e0 - en);
        rs->erow[r] = en;
      }
      ethr[q] = rise_threshold(en);
#pragma unroll
      for (int i = 0; i < PF; ++i) ta[i].scale[q] = __builtin_ldexpf(1.0f, -en);
    }
    sstore(set_c, stage);
    __syncthreads();
#pragma unroll
    for (int q = 0; q < MI::PER; ++q) {
      int r, c;
      MI::item(tid + NT * q, r, c);
      if (c == 0) rs->rmax[par][r] = 0u;
    }
    return true;
  };
  // RUN: before the MFMAs of a k-tile at which some row's exponent rose,
  // every accumulator times its row's factor 2^(old - new) (1 for most rows)
  auto run_rescale = [&]() {
#pragma unroll
    for (int i = 0; i < FM; ++i) {
      if constexpr (LANE_M) {
        const float f = rs->fac[wm * WM + i * 32 + lrow];  // the lane's row in fragment row block i
#pragma unroll
        for (int j = 0; j < FN; ++j)
#pragma unroll
          for (int e = 0; e < 16; ++e) acc[0][i][j][e] *= f;
      } else {
#pragma unroll
        for (int g = 0; g < 4; ++g) {  // registers 4g..4g+3 hold rows 8g + 4lh + 0..3
          const float4 f = *reinterpret_cast<const float4*>(&rs->fac[wm * WM + i * 32 + g * 8 + lh * 4]);
#pragma unroll
          for (int j = 0; j < FN; ++j) {
            acc[0][i][j][4 * g + 0] *= f.x;
            acc[0][i][j][4 * g + 1] *= f.y;
            acc[0][i][j][4 * g + 2] *= f.z;
            acc[0][i][j][4 * g + 3] *= f.w;
          }
        }
      }
    }
  };
  auto compute = [&](const uint4* __restrict__ st) {
    mma_ktile<MODE, LANE_M, NACC, FM, FN, KC, BM, BN, WM, WN>(acc, st, st + NA * A_CH, wm, wn, lrow, lh);
  };

  // LDS double buffer, one barrier per k-tile.  The store into stage cur^1 is
  // safe because the barrier closing the previous iteration ended every read
  // of that stage.  With PF == 2 the global loads of tile kt+2 are issued
  // before tile kt is multiplied and stored to LDS only after tile kt+1 is,
  // so each load has two k-tiles of MFMA work to land behind.
  if constexpr (PF == 1) {
    using P0 = std::integral_constant<int, 0>;
    using P1 = std::integral_constant<int, 1>;
    bool rose = false;  // RUN: some row's exponent rose at the next k-tile to multiply
    if (ktiles > 0) {
      gload(S0{}, 0);
      sstore(S0{}, 0);
      if constexpr (RUN) run_check(S0{}, 0, P0{});
    }
    __syncthreads();
    if constexpr (RUN) {
      if (ktiles > 0) run_settle(S0{}, 0, P0{}, 0);  // nothing accumulated yet: no rescale
    }
    int kt = 0;
    // Interior tiles while the next k-tile is full: only the unchecked loads,
    // with no branch between the A and B loads, so the compiler neither
    // merges the checked paths' register state into the loop (which made it
    // wait for the A loads before issuing the B loads) nor splits the loads
    // across blocks.  Two k-tiles per trip, so each one's LDS stage (and
    // running-exponent parity) is a constant: the LDS addresses are fixed
    // per-lane offsets plus immediates, with no per-tile address arithmetic.
    const int kfull = (kend - kbeg) / BK;
    auto fast_step = [&](auto cur_c, auto bload) {
      constexpr int cur = decltype(cur_c)::value;  // == kt & 1
      using NX = std::integral_constant<int, cur ^ 1>;
      const int k1 = kbeg + (kt + 1) * BK;
      ta[0].load_fast(A, p.lda, m0, k1, tid);
      bload(k1);
      __builtin_amdgcn_sched_barrier(0);  // loads go out before the MFMA work
      if constexpr (RUN) {
        if (rose) run_rescale();
      }
      compute(lds + cur * STAGE);
      if constexpr (RUN) __builtin_amdgcn_sched_barrier(0);  // the split waits for the loads after the MFMAs
      sstore(S0{}, cur ^ 1);
      if constexpr (RUN) run_check(S0{}, kt + 1, NX{});
      __syncthreads();
      if constexpr (RUN) rose = run_settle(S0{}, kt + 1, NX{}, cur ^ 1);
      ++kt;
    };
    auto fast_loop = [&](auto bload) {
      for (; kt + 2 < kfull;) {  // kt is even here
        fast_step(P0{}, bload);
        fast_step(P1{}, bload);
      }
    };
    if (a_full && b_full) fast_loop([&](int k1) { tb[0].load_fast(B, p.ldb, n0, k1, tid); });
    for (; kt < ktiles; ++kt) {
      const int cur = kt & 1;
      if (kt + 1 < ktiles) gload(S0{}, kt + 1);
      if constexpr (RUN) {
        if (rose) run_rescale();
      }
      compute(lds + cur * STAGE);
      if (kt + 1 < ktiles) {
        sstore(S0{}, cur ^ 1);
        if constexpr (RUN) run_check(S0{}, kt + 1, cur ^ 1);
      }
      __syncthreads();
      if constexpr (RUN) {
        if (kt + 1 < ktiles) rose = run_settle(S0{}, kt + 1, cur ^ 1, cur ^ 1);
      }
    }
  } else {
    using Q0 = std::integral_constant<int, 0>;
    using Q1 = std::integral_constant<int, 1>;
    bool rose = false;  // RUN: some row's exponent rose at the next k-tile to multiply
    if (ktiles > 0) gload(S0{}, 0);
    if (ktiles > 1) gload(S1{}, 1);
    if (ktiles > 0) {
      sstore(S0{}, 0);
      if constexpr (RUN) run_check(S0{}, 0, Q0{});
    }
    __syncthreads();
    if constexpr (RUN) {
      if (ktiles > 0) run_settle(S0{}, 0, Q0{}, 0);  // nothing accumulated yet: no rescale
    }
    // Register set s holds tile kt+1 (odd/even alternate), stage kt&1 holds tile kt.
    int kt = 0;
    // Interior pairs while tiles kt+2 and kt+3 are full: unchecked loads only
    // (as in the one-deep loop), then the checked loop from the same state.
    const int kfull = (kend - kbeg) / BK;
    if constexpr (CFG::IL) {
      // Skewed schedule, two k-steps per k-tile.  Per k-tile (stage cur holds
      // tile kt, register set `set` holds tile kt+1, its other set receives
      // tile kt+2):
      //   reads of step 1 || MFMAs of step 0 || split of tile kt+1 into
      //   stage cur^1;  barrier;  reads of step 0 of tile kt+1 || MFMAs of
      //   step 1 of tile kt.
      // The barrier sits before the last k-step's MFMAs, so the first
      // fragments of the next tile load while they run and no k-tile starts
      // with an exposed LDS wait.  Every read of stage cur is issued before
      // the barrier, so the split into it one k-tile later is safe, and the
      // reads of tile kt+1 follow the barrier that ends its stores.
      static_assert(MODE != MODE_BF16 && NACC == 1 && KC == 4, "fp16 one-accumulator mode, two k-steps");
      using Frag = FragF16<FM, FN>;
      constexpr int MF = FM * FN * 2;  // MFMAs per k-step
      constexpr int NR = 2 * FM + FN;  // fragment reads per k-step
      Frag f0, f1;
      auto fload = [&](Frag& f, int stage, int step) {
        const uint4* st = lds + stage * STAGE;
        frag_load<FM, FN, KC, BM, WM, WN>(f, st, st + NA * A_CH, step, wm, wn, lrow, lh);
      };
      auto half = [&](auto set_c, int cur) {
        fload(f1, cur, 1);
        frag_mma<LANE_M>(acc, f0);
        sstore(set_c, cur ^ 1);
#pragma unroll
        for (int q = 0; q < NR; ++q) {
          __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);
          __builtin_amdgcn_sched_group_barrier(0x100, 1, 0);
        }
#pragma unroll
        for (int q = NR; q < MF; ++q) {
          __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);
          __builtin_amdgcn_sched_group_barrier(0x002, 4, 0);
          __builtin_amdgcn_sched_group_barrier(0x200, 1, 0);
        }
        __builtin_amdgcn_sched_barrier(0);
        __syncthreads();
        fload(f0, cur ^ 1, 0);
        frag_mma<LANE_M>(acc, f1);
#pragma unroll
        for (int q = 0; q < NR; ++q) {
          __builtin_amdgcn_sched_group_barrier(0x100, 1, 0);
          __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);
        }
        __builtin_amdgcn_sched_group_barrier(0x008, MF - NR, 0);
        __builtin_amdgcn_sched_barrier(0);
      };
      if (a_full && b_full && kt + 3 < kfull) {
        fload(f0, 0, 0);
        for (; kt + 3 < kfull; kt += 2) {
          const int k2 = kbeg + (kt + 2) * BK;
          ta[0].load_fast(A, p.lda, m0, k2, tid);
          tb[0].load_fast(B, p.ldb, n0, k2, tid);
          __builtin_amdgcn_sched_barrier(0);
          half(S1{}, 0);
          ta[1].load_fast(A, p.lda, m0, k2 + BK, tid);
          tb[1].load_fast(B, p.ldb, n0, k2 + BK, tid);
          __builtin_amdgcn_sched_barrier(0);
          half(S0{}, 1);
        }
      }
    } else if (a_full && b_full) {
      // (running exponents: tile kt+1, in set 1, has parity 1; tile kt+2,
      // in set 0, parity 0)
      for (; kt + 3 < kfull; kt += 2) {
        const int k2 = kbeg + (kt + 2) * BK;
        ta[0].load_fast(A, p.lda, m0, k2, tid);
        tb[0].load_fast(B, p.ldb, n0, k2, tid);
        __builtin_amdgcn_sched_barrier(0);
        if constexpr (RUN) {
          if (rose) run_rescale();
        }
        compute(lds);
        if constexpr (RUN) __builtin_amdgcn_sched_barrier(0);
        sstore(S1{}, 1);
        if constexpr (RUN) run_check(S1{}, kt + 1, Q1{});
        __syncthreads();
        if constexpr (RUN) rose = run_settle(S1{}, kt + 1, Q1{}, 1);
        ta[1].load_fast(A, p.lda, m0, k2 + BK, tid);
        tb[1].load_fast(B, p.ldb, n0, k2 + BK, tid);
        __builtin_amdgcn_sched_barrier(0);
        if constexpr (RUN) {
          if (rose) run_rescale();
        }
        compute(lds + STAGE);
        if constexpr (RUN) __builtin_amdgcn_sched_barrier(0);
        sstore(S0{}, 0);
        if constexpr (RUN) run_check(S0{}, kt + 2, Q0{});
        __syncthreads();
        if constexpr (RUN) rose = run_settle(S0{}, kt + 2, Q0{}, 0);
      }
    }
    for (; kt < ktiles; kt += 2) {
      if (kt + 2 < ktiles) gload(S0{}, kt + 2);
      if constexpr (RUN) {
        if (rose) run_rescale();
      }
      compute(lds);
      if (kt + 1 < ktiles) {
        sstore(S1{}, 1);
        if constexpr (RUN) run_check(S1{}, kt + 1, Q1{});
      }
      __syncthreads();
      if constexpr (RUN) {
        if (kt + 1 < ktiles) rose = run_settle(S1{}, kt + 1, Q1{}, 1);
      }
      if (kt + 1 >= ktiles) break;
      if (kt + 3 < ktiles) gload(S1{}, kt + 3);
      if constexpr (RUN) {
        if (rose) run_rescale();
      }
      compute(lds + STAGE);
      if (kt + 2 < ktiles) {
        sstore(S0{}, 0);
        if constexpr (RUN) run_check(S0{}, kt + 2, Q0{});
      }
      __syncthreads();
      if constexpr (RUN) {
        if (kt + 2 < ktiles) rose = run_settle(S0{}, kt + 2, Q0{}, 0);
      }
    }
  }

  if constexpr (RUN) {
    // the rows' final exponents onto the accumulators (exact powers of two);
    // erow last changed before a barrier
    if (ktiles > 0) {
#pragma unroll
      for (int i = 0; i < FM; ++i) {
        if constexpr (LANE_M) {
          const int e = rs->erow[wm * WM + i * 32 + lrow];
#pragma unroll
          for (int j = 0; j < FN; ++j)
#pragma unroll
            for (int x = 0; x < 16; ++x) acc[0][i][j][x] = __builtin_ldexpf(acc[0][i][j][x], e);
        } else {
#pragma unroll
          for (int g = 0; g < 4; ++g) {  // registers 4g..4g+3 hold rows 8g + 4lh + 0..3
            const int4 e = *reinterpret_cast<const int4*>(&rs->erow[wm * WM + i * 32 + g * 8 + lh * 4]);
#pragma unroll
            for (int j = 0; j < FN; ++j) {
              acc[0][i][j][4 * g + 0] = __builtin_ldexpf(acc[0][i][j][4 * g + 0], e.x);
              acc[0][i][j][4 * g + 1] = __builtin_ldexpf(acc[0][i][j][4 * g + 1], e.y);
              acc[0][i][j][4 * g + 2] = __builtin_ldexpf(acc[0][i][j][4 * g + 2], e.z);
              acc[0][i][j][4 * g + 3] = __builtin_ldexpf(acc[0][i][j][4 * g + 3], e.w);
            }
          }
        }
      }
    }
    // the RunState is read above: no thread may reuse it (next grouped unit)
    // before every thread has finished here
    __syncthreads();
  }
  constexpr bool ROWEXP = !RUN;
  if constexpr (SPLIT == 2) {
    if (part != nullptr)
      store_ctile<MODE, LANE_M, true, NACC, FM, FN, WM, WN, ROWEXP>(acc, p, m0, n0, wm, wn, lrow, lh, part);
    else store_ctile<MODE, LANE_M, false, NACC, FM, FN, WM, WN, ROWEXP>(acc, p, m0, n0, wm, wn, lrow, lh, part);
  } else {
    store_ctile<MODE, LANE_M, SPLIT == 1, NACC, FM, FN, WM, WN, ROWEXP>(acc, p, m0, n0, wm, wn, lrow, lh, part);
  }
}

// The single-problem kernel: blockIdx.x -> output tile (XCD-aware order);
// with SPLITK the grid is tiles x slices (also XCD-aware, slice-major) and
// the partial goes to p.ws + slice * m * n.
template <int OPA, int OPB, int OPC, int MODE, class CFG, bool SPLITK>
__global__ void __launch_bounds__(CFG::NT, CFG::OCC) shgemm_kernel(const GemmParams p) {
  extern __shared__ uint4 lds[];
  int tm, tn;
  if constexpr (SPLITK) {
    int slice;
    tile_slice_coords(p, tm, tn, slice);
    const int kbeg = slice * p.k_slice;
    shgemm_tile<OPA, OPB, OPC, MODE, CFG, 1>(p, tm, tn, kbeg, min(p.k, kbeg + p.k_slice),
                                             p.ws + (int64_t)slice * p.m * p.n, lds);
  } else {
    tile_coords(p, blockIdx.x, tm, tn);
    shgemm_tile<OPA, OPB, OPC, MODE, CFG, 0>(p, tm, tn, 0, p.k, nullptr, lds);
  }
}

// ---------------------------------------------------------------------------
// m-contiguous FP32 tile (A with op_n) straight into LDS, raw: BK columns of
// BM = 256 rows, column c at slots [c*64, c*64 + 64) (element (r, c) at float
// c*BM + r).  One wavefront instruction moves one column: lane l fetches rows
// 4l..4l+3.  Wavefront w issues columns w, w + WAVES, ...  The tile base is
// wave-uniform; the caller guarantees whole columns in range, a 16-byte
// aligned base with lda % 4 == 0, and BK*lda*4 below 2^32.
// ---------------------------------------------------------------------------
template <int BM, int BK, int NT>
struct DirectRawF32 {
  static constexpr int QPC = BM / 4;
  static constexpr int WAVES = NT / 64;
  static constexpr int PER = BK / WAVES;  // instructions per thread per k-tile
  static_assert(QPC == 64, "one wavefront instruction per column");
  static_assert(BK % WAVES == 0, "columns split evenly over the wavefronts");
  __device__ __forceinline__ static void load(const float* __restrict__ base, int64_t ld, int m0, int k0,
                                              uint4* __restrict__ dst, int tid) {
    const int lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const char* tile = uniform_ptr(reinterpret_cast<const char*>(base + (int64_t)k0 * ld + m0));
#pragma unroll
    for (int q = 0; q < PER; ++q) {
      const int c = q * WAVES + wave;
      const uint32_t boff = 4u * ((uint32_t)c * (uint32_t)ld + 4u * (uint32_t)lane);
      __builtin_amdgcn_global_load_lds(tile + boff, reinterpret_cast<uint32_t*>(dst + c * QPC), 16, 0, 0);
    }
  }
};

// ---------------------------------------------------------------------------
// Ring pipeline of the prescaled fp16 P2 tile (256 x 256 x 32, four
// wavefronts of 128 x 128) for op_a = op_b = op_n, interior tiles, whole
// k-tiles.  Each k-tile's raw FP32 A and FP16 B go straight to LDS into a
// 3-slot ring, issued two k-tiles ahead of their MFMA work, so a load has
// about two k-tiles of MFMA time to land and holds no registers in flight.
// Each wavefront reads its own A fragments as FP32 and splits them in
// registers just before the MFMAs that use them (the same LDS bytes as
// reading a stored hi/lo pair; the split work interleaves with the MFMAs);
// one barrier per k-tile.  Every thread issues the same loads each k-tile,
// so the waits are exact vmcnt counts.  The k order of each C element is the
// register-staged kernel's, so the results are bit-identical to it.
// LDS per slot (16-byte slots): raw A (BK*BM/4) + B (BN*KC); 3 slots = 144 KiB.
// ---------------------------------------------------------------------------
template <class CFG>
struct RingLayout {
  static constexpr int RAW = CFG::BK * CFG::BM / 4;
  static constexpr int BST = CFG::BN * CFG::KC;
  static constexpr int SLOT = RAW + BST;
  static constexpr int NSLOT = 3;
  static constexpr size_t lds_bytes() { return (size_t)NSLOT * SLOT * 16; }
};

template <int OPC, class CFG>
__device__ __forceinline__ void shgemm_tile_ring(const GemmParams& p, int m0, int n0, uint4* __restrict__ lds) {
  constexpr int BM = CFG::BM, BN = CFG::BN, BK = CFG::BK, WM = CFG::WM, WN = CFG::WN, KC = CFG::KC, NT = CFG::NT;
  constexpr int FM = WM / 32, FN = WN / 32, WAVES_M = BM / WM;
  constexpr bool LANE_M = (OPC == OP_N);
  using L = RingLayout<CFG>;
  using DA = DirectRawF32<BM, BK, NT>;
  using DB = DirectF16<BN, BK, NT>;
  constexpr int PER = DA::PER + DB::PER;  // vector-memory instructions per thread per k-tile

  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave % WAVES_M, wn = wave / WAVES_M;
  const int lrow = lane & 31, lh = lane >> 5;
  const int ktiles = p.k / BK;
  const float* __restrict__ A = p.a;
  const uint16_t* __restrict__ B = static_cast<const uint16_t*>(p.b);
  float sc[FM];
#pragma unroll
  for (int i = 0; i < FM; ++i)
    sc[i] = p.row_exp != nullptr ? __builtin_ldexpf(1.0f, -p.row_exp[m0 + wm * WM + i * 32 + lrow]) : 1.0f;

  f32x16 acc[FM][FN];
#pragma unroll
  for (int i = 0; i < FM; ++i)
#pragma unroll
    for (int j = 0; j < FN; ++j)
#pragma unroll
      for (int e = 0; e < 16; ++e) acc[i][j][e] = 0.0f;

  auto issue = [&](int kt) {
    uint4* slot = lds + (kt % L::NSLOT) * L::SLOT;
    DB::load(B, p.ldb, n0, kt * BK, p.n, slot + L::RAW, tid);
    DA::load(A, p.lda, m0, kt * BK, slot, tid);
  };
  auto compute = [&](int kt) {
    const uint4* slot = lds + (kt % L::NSLOT) * L::SLOT;
    const float* rf = reinterpret_cast<const float*>(slot);  // element (r, k) at rf[k*BM + r]
    const uint4* bst = slot + L::RAW;
#pragma unroll
    for (int s = 0; s < BK / 16; ++s) {
      const int c = 2 * s + lh;
      f16x8 ah[FM], al[FM], bb[FN];
#pragma unroll
      for (int i = 0; i < FM; ++i) {
        const int r = wm * WM + i * 32 + lrow;
        uint4 h, l;
        split2_f16_scaled(rf[(8 * c + 0) * BM + r], rf[(8 * c + 1) * BM + r], sc[i], h.x, l.x);
        split2_f16_scaled(rf[(8 * c + 2) * BM + r], rf[(8 * c + 3) * BM + r], sc[i], h.y, l.y);
        split2_f16_scaled(rf[(8 * c + 4) * BM + r], rf[(8 * c + 5) * BM + r], sc[i], h.z, l.z);
        split2_f16_scaled(rf[(8 * c + 6) * BM + r], rf[(8 * c + 7) * BM + r], sc[i], h.w, l.w);
        ah[i] = __builtin_bit_cast(f16x8, h);
        al[i] = __builtin_bit_cast(f16x8, l);
      }
#pragma unroll
      for (int j = 0; j < FN; ++j)
        bb[j] = *reinterpret_cast<const f16x8*>(&bst[lds_slot_h<KC>(wn * WN + j * 32 + lrow, c)]);
#pragma unroll
      for (int i = 0; i < FM; ++i)
#pragma unroll
        for (int j = 0; j < FN; ++j) {
          if constexpr (LANE_M) {
            acc[i][j] = mfma_f16(bb[j], ah[i], acc[i][j]);
            acc[i][j] = mfma_f16(bb[j], al[i], acc[i][j]);
          } else {
            acc[i][j] = mfma_f16(ah[i], bb[j], acc[i][j]);
            acc[i][j] = mfma_f16(al[i], bb[j], acc[i][j]);
          }
        }
    }
  };

  issue(0);
  issue(1);
  lds_ring_barrier<PER>();  // tile 0 landed (tile 1 may be in flight)
  for (int kt = 0; kt < ktiles; ++kt) {
    // the slot of tile kt+2 held tile kt-1, whose reads ended at the last barrier
    if (kt + 2 < ktiles) issue(kt + 2);
    __builtin_amdgcn_sched_barrier(0);
    compute(kt);
    __builtin_amdgcn_sched_barrier(0);
    // tile kt+1 landed once at most tile kt+2's loads are outstanding
    if (kt + 2 < ktiles) lds_ring_barrier<PER>();
    else lds_ring_barrier<0>();
  }
  f32x16 out[1][FM][FN];
#pragma unroll
  for (int i = 0; i < FM; ++i)
#pragma unroll
    for (int j = 0; j < FN; ++j) out[0][i][j] = acc[i][j];
  store_ctile<MODE_FP16S, LANE_M, false, 1, FM, FN, WM, WN>(out, p, m0, n0, wm, wn, lrow, lh, nullptr);
}

// The ring kernel runs only whole tiles: the host launches it when m and n
// are multiples of the tile, k is a multiple of BK with at least two k-tiles,
// A is 16-byte aligned with lda % 4 == 0, B with ldb % 8 == 0, and the tile
// offsets fit the 32-bit forms (BK*lda*4 and BN*ldb*2 below 2^32).
template <int OPC, class CFG>
__global__ void __launch_bounds__(CFG::NT, CFG::OCC) shgemm_kernel_ring(const GemmParams p) {
  extern __shared__ uint4 lds[];
  int tm, tn;
  tile_coords(p, blockIdx.x, tm, tn);
  shgemm_tile_ring<OPC, CFG>(p, tm * CFG::BM, tn * CFG::BN, lds);
}

// ---------------------------------------------------------------------------
// Pre-split kernel of the prescaled fp16 mode (256 x 256 x 32 tile; the
// library uses eight wavefronts of 64 x 128, ModeCfg::PS).  A arrives already
// split and packed (launch_presplit_a: per 256-row block and k-tile, 32 KiB
// holding the hi then lo LDS images; p.a points at it), so a k-tile of A is
// one contiguous block and needs no VALU split.  Every operand byte goes from
// global memory straight to LDS (global_load_lds_dwordx4) into a 3-stage
// ring, issued two k-tiles ahead of the MFMAs that use it.  No register holds
// a load in flight, so the accumulators, both fragment sets and the addresses
// fit the register budget of two wavefronts per SIMD; staging the loads in
// registers instead spills, and each spill reload waits for every load in
// flight (16384^3: kernel 13.1-13.5 against 13.9 ms).  B is the packed image
// (launch_pack_b) or, with DIRECT_B, the caller's op_n B read in place: the
// lane at slot s of the image fetches column s / KC, k-chunk (s % KC) ^
// lds_key_h<KC>(s / KC), so four lanes read one column's 64 contiguous bytes
// per k-tile; columns past n read column n - 1 (they feed only C columns past
// n, which are never stored) and the host guarantees BN * ldb * 2 < 2^32.
// Per k-tile kt in stage kt % 3:
//   issue tile kt+2 into stage (kt+2) % 3 (it held tile kt-1, whose reads
//   completed before the last barrier) || reads of step 1 || MFMAs of
//   step 0;  wait until tile kt+1 has landed;  barrier;  reads of step 0 of
//   tile kt+1 || MFMAs of step 1.
// Each C element accumulates in the order of the register-split P2 tile, so
// the results are bit-identical to it.  The host guarantees k % 64 == 0.
// ---------------------------------------------------------------------------
// Rows (columns) of one packed block of the pre-split operands
// (launch_presplit_a, launch_pack_b).  A tile of BM <= 256 rows reads rows
// m0 % 256 .. + BM of block m0 / 256: the LDS-image slot of row r is
// r * KC + (chunk ^ key(r)) with a key that depends on r % 16 only, so those
// rows are already the image of a BM-row tile.
constexpr int kPsBlock = 256;

template <class CFG>
struct PsRingLayout {
  static constexpr int A_CH = CFG::BM * CFG::KC, B_CH = CFG::BN * CFG::KC;
  static constexpr int STAGE = 2 * A_CH + B_CH;  // 16-byte slots: A hi, A lo, B
  static constexpr int NSTAGE = 3;
  static constexpr size_t lds_bytes() { return (size_t)NSTAGE * STAGE * 16; }
};

template <int OPC, class CFG, bool DIRECT_B, int MODE = MODE_FP16S>
__global__ void __launch_bounds__(CFG::NT, CFG::OCC) shgemm_kernel_ps(const GemmParams p) {
  constexpr int BM = CFG::BM, BK = CFG::BK, WM = CFG::WM, WN = CFG::WN, KC = CFG::KC, NT = CFG::NT;
  constexpr int FM = WM / 32, FN = WN / 32, WAVES_M = BM / WM, WAVES = NT / 64;
  constexpr bool LANE_M = (OPC == OP_N);
  // MODE_FP16 (prescaling off): hi and 2^11-scaled lo products in two
  // accumulators, as the register-split kernels of that mode keep them
  static_assert(MODE == MODE_FP16S || MODE == MODE_FP16, "FP16 parts");
  constexpr int NACC = MODE == MODE_FP16 ? 2 : 1;
  static_assert(KC == 4, "two k-steps per k-tile");
  using L = PsRingLayout<CFG>;
  constexpr int A_CH = L::A_CH, B_CH = L::B_CH, STAGE = L::STAGE;
  constexpr int A_PL = kPsBlock * KC, B_PL = kPsBlock * KC;  // slots of one packed plane per k-tile
  static_assert(kPsBlock % CFG::BM == 0 && kPsBlock % CFG::BN == 0, "tiles inside the packed blocks");
  static_assert(A_CH % NT == 0 && B_CH % NT == 0, "whole wavefront instructions per thread");
  constexpr int QA = A_CH / NT, QB = B_CH / NT;  // instructions per thread: each A half, B
  constexpr int PER = 2 * QA + QB;               // vector-memory instructions per thread per k-tile
  using Frag = FragF16<FM, FN>;
  constexpr int MF = FM * FN * 2;  // MFMAs per k-step
  constexpr int NR = 2 * FM + FN;  // fragment reads per k-step
  static_assert(PER <= NR && NR <= MF, "every load and fragment read has an MFMA beside it");
  extern __shared__ uint4 lds[];

  int tm, tn;
  tile_coords(p, blockIdx.x, tm, tn);
  const int m0 = tm * CFG::BM, n0 = tn * CFG::BN;
  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int wm = wave % WAVES_M, wn = wave / WAVES_M;
  const int lrow = lane & 31, lh = lane >> 5;
  const int ktiles = p.k / BK;
  // A's packed k-tiles of this tile's rows: the tile is rows m0 % 256 .. + BM
  // of packed block m0 / 256 (2 * A_PL slots per k-tile: hi, then lo)
  const char* const apack = uniform_ptr(reinterpret_cast<const char*>(p.a) +
                                        ((int64_t)(m0 / kPsBlock) * ktiles * (2 * A_PL) + (m0 % kPsBlock) * KC) * 16);
  // B: this tile's columns of the packed k-tiles (B_PL slots each), or column n0 of the caller's B
  const char* const bbase =
      DIRECT_B ? uniform_ptr(reinterpret_cast<const char*>(static_cast<const uint16_t*>(p.b) + (int64_t)n0 * p.ldb))
               : uniform_ptr(reinterpret_cast<const char*>(p.b) +
                             ((int64_t)(n0 / kPsBlock) * ktiles * B_PL + (n0 % kPsBlock) * KC) * 16);
  constexpr int64_t B_STEP = DIRECT_B ? BK * 2 : B_PL * 16;  // bytes from one k-tile of B to the next
  uint32_t boff[QB];
#pragma unroll
  for (int q = 0; q < QB; ++q) {
    const int s = (q * WAVES + wave) * 64 + lane;
    if constexpr (DIRECT_B) {
      const int j = s / KC, c = (s % KC) ^ lds_key_h<KC>(j);
      boff[q] = (uint32_t)(((int64_t)(min(n0 + j, p.n - 1) - n0) * p.ldb + 8 * c) * 2);
    } else {
      boff[q] = (uint32_t)s * 16u;
    }
  }

  f32x16 acc[NACC][FM][FN];
#pragma unroll
  for (int h = 0; h < NACC; ++h)
#pragma unroll
    for (int i = 0; i < FM; ++i)
#pragma unroll
      for (int j = 0; j < FN; ++j)
#pragma unroll
        for (int e = 0; e < 16; ++e) acc[h][i][j][e] = 0.0f;
  // all hi products of a k-step, then all lo products (into acc[NACC - 1])
  auto fmma = [&](const FragF16<FM, FN>& f) {
#pragma unroll
    for (int h = 0; h < 2; ++h)
#pragma unroll
      for (int i = 0; i < FM; ++i)
#pragma unroll
        for (int j = 0; j < FN; ++j) {
          const f16x8& a = h == 0 ? f.ah[i] : f.al[i];
          f32x16& c = acc[h == 0 ? 0 : NACC - 1][i][j];
          if constexpr (LANE_M) c = mfma_f16(f.bb[j], a, c);
          else c = mfma_f16(a, f.bb[j], c);
        }
  };

  // tile kt into stage st: lane l of instruction (q, wave) fills slot (q * WAVES + wave) * 64 + l
  auto issue = [&](int kt, int st) {
    uint4* const dst = lds + st * STAGE;
    const char* const at = apack + (int64_t)kt * (2 * A_PL * 16);
#pragma unroll
    for (int q = 0; q < QA; ++q) {
      const int blk = q * WAVES + wave;
      const uint32_t off = (uint32_t)(blk * 64 + lane) * 16u;
      lds_dma16(at + off, dst + blk * 64);
      lds_dma16(at + A_PL * 16 + off, dst + A_CH + blk * 64);
    }
    const char* const bt = bbase + (int64_t)kt * B_STEP;
#pragma unroll
    for (int q = 0; q < QB; ++q) {
      const int blk = q * WAVES + wave;
      lds_dma16(bt + boff[q], dst + 2 * A_CH + blk * 64);
    }
  };
  Frag f0, f1;
  auto fload = [&](Frag& f, int st, int step) {
    const uint4* const src = lds + st * STAGE;
    frag_load<FM, FN, KC, BM, WM, WN>(f, src, src + 2 * A_CH, step, wm, wn, lrow, lh);
  };
  // k-tile kt in stage cur; MORE: tile kt+2 exists; NEXT: tile kt+1 does
  auto ktile = [&](int kt, int cur, auto more_c, auto next_c) {
    constexpr bool MORE = decltype(more_c)::value, NEXT = decltype(next_c)::value;
    const int nxt = cur == 2 ? 0 : cur + 1, far = cur == 0 ? 2 : cur - 1;
    if constexpr (MORE) issue(kt + 2, far);
    fload(f1, cur, 1);
    fmma(f0);
#pragma unroll
    for (int q = 0; q < NR; ++q) {
      __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);
      __builtin_amdgcn_sched_group_barrier(0x100, 1, 0);
      if (MORE && q < PER) __builtin_amdgcn_sched_group_barrier(0x020, 1, 0);
    }
    __builtin_amdgcn_sched_group_barrier(0x008, MF - NR, 0);
    __builtin_amdgcn_sched_barrier(0);
    // tile kt+1 has landed once at most tile kt+2's loads are outstanding
    if constexpr (MORE) lds_ring_barrier<PER>();
    else lds_ring_barrier<0>();
    if constexpr (NEXT) fload(f0, nxt, 0);
    fmma(f1);
#pragma unroll
    for (int q = 0; q < NR; ++q) {
      __builtin_amdgcn_sched_group_barrier(0x100, 1, 0);
      __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);
    }
    __builtin_amdgcn_sched_group_barrier(0x008, MF - NR, 0);
    __builtin_amdgcn_sched_barrier(0);
  };
  using T = std::true_type;
  using F = std::false_type;
  issue(0, 0);
  issue(1, 1);  // ktiles >= 2
  lds_ring_barrier<PER>();  // tile 0 landed
  fload(f0, 0, 0);
  int kt = 0, cur = 0;
  for (; kt + 2 < ktiles; ++kt) {
    ktile(kt, cur, T{}, T{});
    cur = cur == 2 ? 0 : cur + 1;
  }
  ktile(kt, cur, F{}, T{});
  ktile(kt + 1, cur == 2 ? 0 : cur + 1, F{}, F{});
  store_ctile<MODE, LANE_M, false, NACC, FM, FN, WM, WN>(acc, p, m0, n0, wm, wn, lrow, lh, nullptr);
}

// ---------------------------------------------------------------------------
// Pre-split kernel of the tf32 compute type: the ring of the fp16 pre-split
// kernel on A's three BF16 parts and B's two, both packed in tile order
// (launch_presplit_a / launch_pack_b with MODE_BF16: per 256-row block and
// k-tile 48 KiB holding the a0, a1, a2 images; per 256-column block and
// k-tile 32 KiB holding b0, b1).  A 256 x 256 x 32 stage is 80 KiB, so the
// ring has two stages: tile kt+1 is issued into the other stage (it held
// tile kt-1, whose reads completed before the last barrier) while tile kt is
// multiplied, and the barrier after tile kt waits for it.  Each accumulator
// takes, per 16-deep k-step in k order, the products a0*b0, a0*b1, a1*b0,
// a1*b1, a2*b0 in that order, as the register-split kernels do (mma_kstep):
// the results are bit-identical to them.  The host guarantees k % 64 == 0.
// ---------------------------------------------------------------------------
template <class CFG>
struct PsBf16Layout {
  static constexpr int A_CH = CFG::BM * CFG::KC, B_CH = CFG::BN * CFG::KC;
  static constexpr int STAGE = 3 * A_CH + 2 * B_CH;  // 16-byte slots: a0, a1, a2, b0, b1
  static constexpr int NSTAGE = 2;
  static constexpr size_t lds_bytes() { return (size_t)NSTAGE * STAGE * 16; }
};

template <int FM, int FN>
struct FragB16 {
  bf16x8 a[3][FM], b[2][FN];
};

template <int OPC, class CFG>
__global__ void __launch_bounds__(CFG::NT, CFG::OCC) shgemm_kernel_ps_bf16(const GemmParams p) {
  constexpr int BM = CFG::BM, BK = CFG::BK, WM = CFG::WM, WN = CFG::WN, KC = CFG::KC, NT = CFG::NT;
  constexpr int FM = WM / 32, FN = WN / 32, WAVES_M = BM / WM, WAVES = NT / 64;
  constexpr bool LANE_M = (OPC == OP_N);
  static_assert(KC == 4, "two k-steps per k-tile");
  using L = PsBf16Layout<CFG>;
  constexpr int A_CH = L::A_CH, B_CH = L::B_CH, STAGE = L::STAGE;
  constexpr int A_PL = kPsBlock * KC, B_PL = kPsBlock * KC;  // slots of one packed plane per k-tile
  static_assert(kPsBlock % CFG::BM == 0 && kPsBlock % CFG::BN == 0, "tiles inside the packed blocks");
  static_assert(A_CH % NT == 0 && B_CH % NT == 0, "whole wavefront instructions per thread");
  constexpr int QA = A_CH / NT, QB = B_CH / NT;  // instructions per thread per part
  constexpr int PER = 3 * QA + 2 * QB;           // vector-memory instructions per thread per k-tile
  using Frag = FragB16<FM, FN>;
  constexpr int MF = FM * FN * 5;  // MFMAs per k-step
  constexpr int NR = 3 * FM + 2 * FN;  // fragment reads per k-step
  static_assert(PER <= NR && NR <= MF, "every load and fragment read has an MFMA beside it");
  extern __shared__ uint4 lds[];

  int tm, tn;
  tile_coords(p, blockIdx.x, tm, tn);
  const int m0 = tm * CFG::BM, n0 = tn * CFG::BN;
  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int wm = wave % WAVES_M, wn = wave / WAVES_M;
  const int lrow = lane & 31, lh = lane >> 5;
  const int ktiles = p.k / BK;
  // this tile's rows / columns of the packed blocks (3 * A_PL, 2 * B_PL slots per k-tile)
  const char* const apack = uniform_ptr(reinterpret_cast<const char*>(p.a) +
                                        ((int64_t)(m0 / kPsBlock) * ktiles * (3 * A_PL) + (m0 % kPsBlock) * KC) * 16);
  const char* const bpack = uniform_ptr(reinterpret_cast<const char*>(p.b) +
                                        ((int64_t)(n0 / kPsBlock) * ktiles * (2 * B_PL) + (n0 % kPsBlock) * KC) * 16);

  f32x16 acc[1][FM][FN];
#pragma unroll
  for (int i = 0; i < FM; ++i)
#pragma unroll
    for (int j = 0; j < FN; ++j)
#pragma unroll
      for (int e = 0; e < 16; ++e) acc[0][i][j][e] = 0.0f;

  // tile kt into stage st: lane l of instruction (q, wave) fills slot (q * WAVES + wave) * 64 + l of each part
  auto issue = [&](int kt, int st) {
    uint4* const dst = lds + st * STAGE;
    const char* const at = apack + (int64_t)kt * (3 * A_PL * 16);
    const char* const bt = bpack + (int64_t)kt * (2 * B_PL * 16);
#pragma unroll
    for (int q = 0; q < QA; ++q) {
      const int blk = q * WAVES + wave;
      const uint32_t off = (uint32_t)(blk * 64 + lane) * 16u;
#pragma unroll
      for (int part = 0; part < 3; ++part) lds_dma16(at + part * A_PL * 16 + off, dst + part * A_CH + blk * 64);
    }
#pragma unroll
    for (int q = 0; q < QB; ++q) {
      const int blk = q * WAVES + wave;
      const uint32_t off = (uint32_t)(blk * 64 + lane) * 16u;
#pragma unroll
      for (int part = 0; part < 2; ++part)
        lds_dma16(bt + part * B_PL * 16 + off, dst + 3 * A_CH + part * B_CH + blk * 64);
    }
  };
  Frag f0, f1;
  auto fload = [&](Frag& f, int st, int step) {
    const uint4* const src = lds + st * STAGE;
    const int c = 2 * step + lh;
#pragma unroll
    for (int i = 0; i < FM; ++i) {
      const int sl = lds_slot_h<KC>(wm * WM + i * 32 + lrow, c);
#pragma unroll
      for (int part = 0; part < 3; ++part) f.a[part][i] = *reinterpret_cast<const bf16x8*>(&src[part * A_CH + sl]);
    }
#pragma unroll
    for (int j = 0; j < FN; ++j) {
      const int sl = lds_slot_h<KC>(wn * WN + j * 32 + lrow, c);
#pragma unroll
      for (int part = 0; part < 2; ++part)
        f.b[part][j] = *reinterpret_cast<const bf16x8*>(&src[3 * A_CH + part * B_CH + sl]);
    }
  };
  // the five products one at a time over all accumulators (each accumulator
  // still takes them in mma_kstep's order)
  auto fmma = [&](const Frag& f) {
    constexpr int PA[5] = {0, 0, 1, 1, 2}, PB[5] = {0, 1, 0, 1, 0};
#pragma unroll
    for (int t = 0; t < 5; ++t)
#pragma unroll
      for (int i = 0; i < FM; ++i)
#pragma unroll
        for (int j = 0; j < FN; ++j) {
          if constexpr (LANE_M) acc[0][i][j] = mfma_bf16(f.b[PB[t]][j], f.a[PA[t]][i], acc[0][i][j]);
          else acc[0][i][j] = mfma_bf16(f.a[PA[t]][i], f.b[PB[t]][j], acc[0][i][j]);
        }
  };
  // k-tile kt in stage cur; MORE: tile kt+1 exists
  auto ktile = [&](int kt, int cur, auto more_c) {
    constexpr bool MORE = decltype(more_c)::value;
    if constexpr (MORE) issue(kt + 1, cur ^ 1);
    fload(f1, cur, 1);
    fmma(f0);
#pragma unroll
    for (int q = 0; q < NR; ++q) {
      __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);
      __builtin_amdgcn_sched_group_barrier(0x100, 1, 0);
      if (MORE && q < PER) __builtin_amdgcn_sched_group_barrier(0x020, 1, 0);
    }
    __builtin_amdgcn_sched_group_barrier(0x008, MF - NR, 0);
    __builtin_amdgcn_sched_barrier(0);
    lds_ring_barrier<0>();  // tile kt+1 has landed; every read of stage cur is done
    if constexpr (MORE) fload(f0, cur ^ 1, 0);
    fmma(f1);
#pragma unroll
    for (int q = 0; q < NR; ++q) {
      __builtin_amdgcn_sched_group_barrier(0x100, 1, 0);
      __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);
    }
    __builtin_amdgcn_sched_group_barrier(0x008, MF - NR, 0);
    __builtin_amdgcn_sched_barrier(0);
  };
  issue(0, 0);
  lds_ring_barrier<0>();  // tile 0 landed
  fload(f0, 0, 0);
  int kt = 0;
  for (; kt + 1 < ktiles; ++kt) ktile(kt, kt & 1, std::true_type{});
  ktile(kt, kt & 1, std::false_type{});
  store_ctile<MODE_BF16, LANE_M, false, 1, FM, FN, WM, WN>(acc, p, m0, n0, wm, wn, lrow, lh, nullptr);
}

// ---------------------------------------------------------------------------
// Deep-prefetch variant of the prescaled fp16 kernel for op_b = op_n with
// full k-tiles (k % BK == 0) and aligned B: B goes straight to LDS into a
// 3-stage ring and A through two register sets, both issued two k-tiles
// ahead of the MFMA work, so each load has two k-tiles of MFMA time to land.
// Every thread issues the same number of vector-memory instructions per
// k-tile (no bounds-checked paths: edge rows are clamped), which lets the
// barrier wait with an exact vmcnt for the tile it publishes.
// LDS: A hi/lo x 2 stages, B x 3 stages.
// ---------------------------------------------------------------------------
template <class CFG>
struct BdLayout {
  static constexpr int KC = CFG::KC;
  static constexpr int A_STAGE = 2 * CFG::BM * KC;  // hi + lo slots
  static constexpr int B_STAGE = CFG::BN * KC;
  static constexpr size_t lds_bytes() { return (size_t)(2 * A_STAGE + 3 * B_STAGE) * 16; }
};

template <int OPA, int OPC, class CFG>
__global__ void __launch_bounds__(CFG::NT, CFG::OCC) shgemm_kernel_bd(const GemmParams p) {
  constexpr int MODE = MODE_FP16S;
  constexpr int BM = CFG::BM, BN = CFG::BN, BK = CFG::BK, WM = CFG::WM, WN = CFG::WN, KC = CFG::KC;
  constexpr int FM = WM / 32, FN = WN / 32, WAVES_M = BM / WM, NT = CFG::NT;
  constexpr int A_CH = BM * KC;
  constexpr int A_STAGE = BdLayout<CFG>::A_STAGE, B_STAGE = BdLayout<CFG>::B_STAGE;
  constexpr bool LANE_M = (OPC == OP_N);
  using TA = TileF32<OPA == OP_T, BM, BK, NT>;
  using DB = DirectF16<BN, BK, NT>;
  // vector-memory instructions one k-tile issues per thread
  constexpr int VM_PER_TILE = TA::kClampedLoads + DB::PER;
  extern __shared__ uint4 lds[];
  uint4* const a_lds = lds;                  // 2 stages
  uint4* const b_lds = lds + 2 * A_STAGE;    // 3 stages

  TA ta0, ta1;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave % WAVES_M, wn = wave / WAVES_M;
  const int lrow = lane & 31, lh = lane >> 5;
  int tm, tn;
  tile_coords(p, blockIdx.x, tm, tn);
  const int m0 = tm * BM, n0 = tn * BN;
  const int ktiles = p.k / BK;
  const float* __restrict__ A = as_global(p.a);  // global_* (not flat_*) loads: a flat load waits for the LDS DMA in flight
  const uint16_t* __restrict__ B = as_global(static_cast<const uint16_t*>(p.b));
  ta0.init_scale(p.row_exp, m0, p.m, tid);
  ta1.init_scale(p.row_exp, m0, p.m, tid);

  f32x16 acc[1][FM][FN];
#pragma unroll
  for (int i = 0; i < FM; ++i)
#pragma unroll
    for (int j = 0; j < FN; ++j)
#pragma unroll
      for (int e = 0; e < 16; ++e) acc[0][i][j][e] = 0.0f;

  // issue one k-tile: B into ring slot kt%3, A into the given register set
  auto issue = [&](TA& ta, int kt) {
    DB::load(B, p.ldb, n0, kt * BK, p.n, b_lds + (kt % 3) * B_STAGE, tid);
    ta.load_clamped(A, p.lda, m0, kt * BK, p.m, tid);
  };
  auto store_a = [&](const TA& ta, int kt) {
    uint4* st = a_lds + (kt & 1) * A_STAGE;
    ta.template store_f16<true>(st, st + A_CH, tid);
  };
  auto compute = [&](int kt) {
    mma_ktile<MODE, LANE_M, 1, FM, FN, KC, BM, BN, WM, WN>(acc, a_lds + (kt & 1) * A_STAGE,
                                                          b_lds + (kt % 3) * B_STAGE, wm, wn, lrow, lh);
  };
  // Publish k-tile kt+1 to LDS: its A registers are stored by this thread, and
  // its B load must have landed; at most one newer k-tile of loads may still
  // be in flight (`newer`).
  auto publish = [&](bool newer) {
    if (newer) lds_ring_barrier<VM_PER_TILE>();
    else lds_ring_barrier<0>();
  };

  if (ktiles > 0) {
    issue(ta0, 0);
    __builtin_amdgcn_sched_barrier(0);
    if (ktiles > 1) issue(ta1, 1);
    __builtin_amdgcn_sched_barrier(0);
    store_a(ta0, 0);
    publish(ktiles > 1);
  }
  for (int kt = 0; kt < ktiles; kt += 2) {
    // even step: ta0 is free (tile kt stored), ta1 holds tile kt+1
    if (kt + 2 < ktiles) issue(ta0, kt + 2);
    __builtin_amdgcn_sched_barrier(0);
    compute(kt);
    __builtin_amdgcn_sched_barrier(0);
    if (kt + 1 < ktiles) store_a(ta1, kt + 1);
    publish(kt + 2 < ktiles);
    if (kt + 1 >= ktiles) break;
    // odd step: ta1 is free, ta0 holds tile kt+2
    if (kt + 3 < ktiles) issue(ta1, kt + 3);
    __builtin_amdgcn_sched_barrier(0);
    compute(kt + 1);
    __builtin_amdgcn_sched_barrier(0);
    if (kt + 2 < ktiles) store_a(ta0, kt + 2);
    publish(kt + 3 < ktiles);
  }

  store_ctile<MODE, LANE_M, false, 1, FM, FN, WM, WN>(acc, p, m0, n0, wm, wn, lrow, lh, nullptr);
}

}  // namespace shgemm_amd
