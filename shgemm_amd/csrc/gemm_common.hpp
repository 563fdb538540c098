// Shared device-side definitions for the gfx950 GEMM kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

namespace shgemm_amd {

typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));
typedef _Float16 f16x4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
typedef uint32_t u32x2 __attribute__((ext_vector_type(2)));

enum : int { OP_N = 0, OP_T = 1 };
// MODE_FP16:  A = hi + 2^-11 lo, both FP16, hi and lo products in separate
//             FP32 accumulators (reference src/wmmae_shgemm.hpp:29-31, 73-125).
// MODE_BF16:  A = a0 + a1 + a2 and B = b0 + b1, all BF16 (the tf32 compute
//             type); the five products share one accumulator.
// MODE_FP16S: each row of A prescaled by 2^-e(row) (exact), then
//             x = hi + lo with both FP16 at the same scale, so the hi and lo
//             products share one FP32 accumulator; C's row is scaled back by
//             2^e(row).  Same split precision as MODE_FP16, FP32 exponent range.
enum : int { MODE_FP16 = 0, MODE_BF16 = 1, MODE_FP16S = 2 };
constexpr int kNumModes = 3;

// Everything a GEMM launch needs.  Sizes are 32-bit (the host rejects larger
// ones); leading dimensions and offsets are 64-bit.
struct GemmParams {
  int m, n, k;
  float alpha, beta;
  const float* a;      // FP32 operand
  int64_t lda;
  const void* b;       // FP16 bits (SHGEMM) or FP32 (SGEMM)
  int64_t ldb;
  float* c;
  int64_t ldc;
  float* ws;           // split-K partials, slice-major, each slice m*n floats
  const int* row_exp;  // per-row exponent of A (nullptr: no prescaling)
  int k_slice;         // k extent per slice (multiple of the tile depth)
  int tiles_m, tiles_n;
  int group_m;         // tile rows per L2 group
  int vec_a, vec_b;    // 1 when base pointers and ld allow aligned 16-byte loads (the direct-to-LDS plans)
};

// Workgroup id -> output tile.  Consecutive workgroup ids are dispatched
// round-robin over the 8 XCDs; remap so each XCD owns a contiguous run of
// tiles, then walk that run in groups of `group_m` tile rows so the tiles in
// flight on one XCD share A and B panels in its L2.
// Linear tile index t -> (tm, tn): groups of `group_m` tile rows, column by
// column inside a group.
__device__ __forceinline__ void grouped_order(const GemmParams& p, int t, int& tm, int& tn) {
  const int per_group = p.group_m * p.tiles_n;
  const int g = t / per_group;
  const int first = g * p.group_m;
  const int rows = min(p.tiles_m - first, p.group_m);
  const int local = t - g * per_group;
  tm = first + local % rows;
  tn = local / rows;
}

// A pointer read from memory (not a kernel argument) is generic to the
// compiler, so its accesses become flat_* instructions, and a flat load also
// counts against the LDS wait counter.  A round trip through the global
// address space tells the compiler it points to device memory, so they are
// global_* instructions.
template <class T>
__device__ __forceinline__ T* as_global(T* ptr) {
  using G = __attribute__((address_space(1))) T;
  G* g = (G*)(uintptr_t)ptr;
  return (T*)g;
}

// Dispatch index -> position in an order where each XCD owns one contiguous
// run of the nb indices (XCD x receives dispatch indices x, x + 8, ...).
__device__ __forceinline__ int xcd_run(int bid, int nb) {
  const int q = nb >> 3, r = nb & 7;
  const int x = bid & 7, i = bid >> 3;
  return x * q + (x < r ? x : r) + i;
}

__device__ __forceinline__ void tile_coords(const GemmParams& p, int bid, int& tm, int& tn) {
  grouped_order(p, xcd_run(bid, p.tiles_m * p.tiles_n), tm, tn);
}

// Split-K grid (tiles x slices): the whole grid in dispatch order goes
// through the XCD runs, slice-major, so one XCD works on few slices and
// keeps their B panel (and the A panels its tiles share) in its L2.
__device__ __forceinline__ void tile_slice_coords(const GemmParams& p, int& tm, int& tn, int& slice) {
  const int nt = gridDim.x;
  const int run = xcd_run(blockIdx.x + blockIdx.y * nt, nt * gridDim.y);
  slice = run / nt;
  grouped_order(p, run - slice * nt, tm, tn);
}

// LDS operand layout: row-major 16-byte slots, the chunk index of each row
// XOR-swizzled so that
//   * a wavefront reading one chunk of 32 (or 64) consecutive rows -- the MFMA
//     operand pattern -- touches every bank position once per 16 rows, and
//   * a wavefront storing whole rows, or one chunk of consecutive rows, stores
//     a permutation of contiguous 256-byte windows.
// Both were measured bank-conflict-free on gfx950 (SQ_LDS_BANK_CONFLICT = 0),
// where the unswizzled row-major read costs 12 conflict cycles per instruction.
// Rows per 256-byte window = 16 / KC; the swizzle key is the window index.
template <int KC>
__device__ __forceinline__ int lds_slot_rm(int r, int c) {
  static_assert(KC >= 2 && KC <= 16 && (KC & (KC - 1)) == 0, "2..16 chunks per row");
  return r * KC + (c ^ ((r / (16 / KC)) & (KC - 1)));
}

// The same layout for the FP16 operand tiles of the SHGEMM kernels, except
// that with 4 or 8 chunks per row (32- and 64-deep k-tiles) the key is
// r ^ (r >> 2) or r ^ (r >> 3): then a wavefront storing one chunk of 64
// consecutive rows (m-contiguous A) is conflict-free too.  Measured on gfx950
// with 16-byte accesses: that store costs 8 conflict cycles per instruction
// under the window key and 0 under these; whole-row stores and the MFMA
// operand reads stay at 0 under both.
template <int KC>
__device__ __forceinline__ int lds_key_h(int r) {
  static_assert(KC >= 2 && KC <= 16 && (KC & (KC - 1)) == 0, "2..16 chunks per row");
  if constexpr (KC == 8) return (r ^ (r >> 3)) & 7;
  else if constexpr (KC == 4) return (r ^ (r >> 2)) & 3;
  else return (r / (16 / KC)) & (KC - 1);
}
template <int KC>
__device__ __forceinline__ int lds_slot_h(int r, int c) {
  return r * KC + (c ^ lds_key_h<KC>(r));
}

// FP32 -> FP16 hi + 2^11-scaled FP16 lo, both round-to-nearest-even.
// a - float(hi) is exact in FP32 (hi is the FP16 rounding of a).
__device__ __forceinline__ void split_f16(float a, _Float16& hi, _Float16& lo) {
  hi = (_Float16)a;
  const float r = a - (float)hi;
  lo = (_Float16)(r * 2048.0f);
}

// Prescaled FP32 -> FP16 hi + FP16 residual at the same scale.
__device__ __forceinline__ void split_f16_same_scale(float x, _Float16& hi, _Float16& lo) {
  hi = (_Float16)x;
  lo = (_Float16)(x - (float)hi);
}

// Two elements at once, scale s = 2^-e applied inside: hi = FP16(x*s),
// lo = FP16(x*s - hi), packed (element 0 in bits 15:0).  Each mixed-precision
// FMA forms x*s (+ -hi) exactly and rounds once to FP16, so the bits equal
// split_f16_same_scale(x*s) (x*s is exact in FP32: s is a power of two that
// keeps the row's largest element below 2^15).  Four VALU instructions per
// pair, where the C form is compiled to seven (hi is formed twice, once
// packed and once for the residual).
__device__ __forceinline__ void split2_f16_scaled(float x0, float x1, float s, uint32_t& hi, uint32_t& lo) {
  asm("v_fma_mixlo_f16 %0, %2, %4, 0\n\t"
      "v_fma_mixhi_f16 %0, %3, %4, 0\n\t"
      "v_fma_mixlo_f16 %1, %2, %4, -%0 op_sel_hi:[0,0,1]\n\t"
      "v_fma_mixhi_f16 %1, %3, %4, -%0 op_sel:[0,0,1] op_sel_hi:[0,0,1]"
      : "=&v"(hi), "=&v"(lo)
      : "v"(x0), "v"(x1), "v"(s));
}

// row_exp = ilogb(max |A[r,:]|) - 14 from the |a| bit pattern of the maximum
// (0 for an all-zero or non-finite row), clamped so 2^-e stays a finite FP32.
__device__ __forceinline__ int exponent_of_max(uint32_t b) {
  int e = 0;
  if (b != 0 && b < 0x7f800000u) {  // finite, non-zero
    const int biased = (int)(b >> 23);
    const int lg = biased != 0 ? biased - 127 : (31 - __builtin_clz(b)) - 149;  // ilogb(|a|max)
    e = lg - 14;
    e = e < -126 ? -126 : e;
  }
  return e;
}

// The smallest |a| bit pattern b with exponent_of_max(b) > e, for
// -126 <= e <= 113: 2^(e + 15), or the infinity pattern for e = 113 (no
// finite value exceeds it).  For every finite b, b >= rise_threshold(e)
// exactly when exponent_of_max(b) > e; an infinity or NaN pattern is always
// at or above it.
__device__ __forceinline__ uint32_t rise_threshold(int e) {
  return e < 113 ? (uint32_t)(e + 142) << 23 : 0x7f800000u;
}
// The e of a rise_threshold(e).
__device__ __forceinline__ int threshold_exponent(uint32_t t) {
  return t < 0x7f800000u ? (int)(t >> 23) - 142 : 113;
}

// FP32 -> three BF16 parts a0 + a1 + a2 (each residual exact in FP32).
__device__ __forceinline__ void split_bf16x3(float a, __bf16& a0, __bf16& a1, __bf16& a2) {
  a0 = (__bf16)a;
  const float r1 = a - (float)a0;
  a1 = (__bf16)r1;
  const float r2 = r1 - (float)a1;
  a2 = (__bf16)r2;
}

// FP16 -> two BF16 parts b0 + b1, exact (b has 11 significant bits).
__device__ __forceinline__ void split_bf16x2(_Float16 b, __bf16& b0, __bf16& b1) {
  const float f = (float)b;
  b0 = (__bf16)f;
  b1 = (__bf16)(f - (float)b0);
}

// A pointer the compiler must treat as wave-uniform (held in scalar
// registers), so loads off it use the scalar-base + 32-bit-offset form.
template <class T>
__device__ __forceinline__ const T* uniform_ptr(const T* p) {
  const uint64_t a = reinterpret_cast<uint64_t>(p);
  const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)a);
  const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(a >> 32));
  return reinterpret_cast<const T*>(((uint64_t)hi << 32) | lo);
}

// A T-sized value from an address aligned only to its element type (a row of
// FP16 B at an odd ld, op_t A at an ld that is not a multiple of 4).  The
// copy compiles to one global load of sizeof(T) bytes: gfx9 global memory
// runs in unaligned mode under ROCm, so the hardware splits the access where
// it crosses a cache line, and the load costs about as much as an aligned one
// (eight 2-byte loads per 16 bytes cost twice the whole GEMM at k = 16001).
template <class T, class E>
__device__ __forceinline__ T load_unaligned(const E* p) {
  T v;
  __builtin_memcpy(&v, p, sizeof(T));
  return v;
}

__device__ __forceinline__ _Float16 bits_to_f16(uint16_t u) {
  return __builtin_bit_cast(_Float16, u);
}

__device__ __forceinline__ f32x16 mfma_f16(const f16x8& a, const f16x8& b, const f32x16& c) {
  return __builtin_amdgcn_mfma_f32_32x32x16_f16(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ f32x16 mfma_bf16(const bf16x8& a, const bf16x8& b, const f32x16& c) {
  return __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ f32x16 mfma_f32(float a, float b, const f32x16& c) {
  return __builtin_amdgcn_mfma_f32_32x32x2f32(a, b, c, 0, 0, 0);
}

}  // namespace shgemm_amd
