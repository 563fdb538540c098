// Kernel table: every compiled kernel variant, filled by the instantiation
// units (inst_*.hip) and read by the host dispatcher.
#pragma once
#include <hip/hip_runtime.h>
#include <cstddef>
#include "gemm_common.hpp"

namespace shgemm_amd {

// LDS of one gfx950 compute unit (the most one workgroup can allocate).
constexpr size_t kMaxLdsBytes = 160 * 1024;

struct KernelEntry {
  const void* fn = nullptr;
  size_t lds = 0;      // dynamic LDS bytes
  int bm = 0, bn = 0, bk = 0;
  int mode = -1;
  int threads = 256;
  int wgs_per_cu = 2;  // workgroups one CU holds at the configuration's register budget
  bool needs_row_exp = false;  // prescaled fp16 kernel that reads exponents from the row pass
};

enum : int { LEVEL_P0 = 0, LEVEL_P1 = 1 };

// Kernel variants per (op_a, op_b, op_c, mode): tile config x split-K.
enum : int {
  V_P0 = 0,    // small tile
  V_P1 = 1,    // large tile
  V_SK = 2,    // 128 x 128 tile, split-K
  V_P1SK = 3,  // large tile, split-K
  V_P0SK = 4,  // small tile, split-K
  V_P2 = 5,    // 256 x 256 tile of four 128 x 128 wavefronts (one per SIMD)
  V_P2R = 6,   // the same tile with the 3-slot direct-to-LDS ring (prescaled fp16, op_n/op_n, whole tiles)
  kNumVariants = 7
};
inline bool variant_is_split(int v) { return v == V_SK || v == V_P1SK || v == V_P0SK; }
inline int variant_level(int v) { return (v == V_P0 || v == V_P0SK) ? LEVEL_P0 : LEVEL_P1; }

// Tile classes of the grouped kernel (small, 128 x 128, the mode's large tile,
// and for the prescaled fp16 mode a 128 x 512 tile that covers n up to 512).
enum : int { kNumGroupClasses = 4 };

struct KernelTable {
  // [op_a][op_b][op_c][mode][variant]
  KernelEntry shgemm[2][2][2][kNumModes][kNumVariants];
  // [op_a][op_b][op_c]
  KernelEntry sgemm[2][2][2];
  // deep-prefetch prescaled fp16 kernel (op_b = n, full k-tiles): [op_a][op_c]
  KernelEntry shgemm_bd[2][2];
  // pre-split prescaled fp16 kernel (A split once per call): [B read as it
  // lies (op_b = n) rather than packed][op_c]
  KernelEntry shgemm_ps[2][2];
  // pre-split tf32 kernel (A's three BF16 parts and B's two packed once per call): [op_c]
  KernelEntry shgemm_ps_bf16[2];
  // the pre-split kernels on 128 x 128 tiles: [B in place][op_c] and tf32 [op_c]
  KernelEntry shgemm_ps128[2][2];
  KernelEntry shgemm_ps128_bf16[2];
  // the 128 x 128 pre-split kernel with prescaling off (two accumulators): [B in place][op_c]
  KernelEntry shgemm_ps128_fp16[2][2];
  // prescaled fp16 small tile on whole-row exponents: [op_a][op_b][op_c]
  KernelEntry shgemm_p0w[2][2][2];
  // prescaled fp16 deep k-sliced 128 x 128 x 64 tile (plan 10): [op_a][op_b][op_c]
  KernelEntry shgemm_skd[2][2][2];
  // prescaled fp16 large tile of four 64 x 128 wavefronts, k-sliced (plan 11): [op_a][op_b][op_c]
  KernelEntry shgemm_wsk[2][2][2];
  // persistent grouped kernels: [op_a][op_b][op_c][mode][tile class]
  KernelEntry grouped[2][2][2][kNumModes][kNumGroupClasses];
  // Candidate large-tile configurations of the fp16 op_n/op_n/C-op_n kernel,
  // timed by the tile sweep and selectable with MTK_SHGEMM_TUNE_P1=<index>.
  static constexpr int kMaxTune = 24;
  KernelEntry tune[kMaxTune];
  int ntune = 0;
};

void register_shgemm_fp16_nn(KernelTable&);
void register_shgemm_fp16_nt(KernelTable&);
void register_shgemm_fp16_tn(KernelTable&);
void register_shgemm_fp16_tt(KernelTable&);
void register_shgemm_bf16_nn(KernelTable&);
void register_shgemm_bf16_nt(KernelTable&);
void register_shgemm_bf16_tn(KernelTable&);
void register_shgemm_bf16_tt(KernelTable&);
void register_shgemm_fp16s_nn(KernelTable&);
void register_shgemm_fp16s_nt(KernelTable&);
void register_shgemm_fp16s_tn(KernelTable&);
void register_shgemm_fp16s_tt(KernelTable&);
void register_sgemm(KernelTable&);
void register_tune(KernelTable&);
void register_shgemm_bd(KernelTable&);
void register_shgemm_ps(KernelTable&);
void register_all(KernelTable&);

// One problem of a group as the device sees it.  p.ws is the problem's own
// partial buffer (slice-major, m*n floats per slice) when it is k-sliced,
// nullptr otherwise; p.row_exp points at the problem's rows of the group's
// exponent buffer when its class's kernel reads whole-row exponents, nullptr
// otherwise (classes with running exponents).
struct GroupProblem {
  GemmParams p;
  int unit_begin;  // first work unit of the problem
  int pad;
};

// Arguments of one persistent grouped launch (one tile class).
struct GroupLaunch {
  const GroupProblem* probs;
  const int* unit_begin;  // probs[i].unit_begin, packed for the search
  int nprobs;
  int total_units;
  int* head;  // queue head, zero before the launch
};

// Per-problem segment of a grouped row-exponent pass (blocks = bx * by).
struct RowmaxSeg {
  const float* a;
  int64_t lda;
  uint32_t* row_max;
  int m, k, kchunk, rq, vec, bx;
  int block_begin, pad;
};

// Per-problem segment of a grouped split-K reduction (slices = 0: C = beta*C).
struct ReduceSeg {
  const float* ws;
  float* c;
  int64_t ldc;
  int64_t elems;
  int slices, inner;
  float alpha, beta;
  int block_begin, pad;
};

// Launch geometry of the row-max pass of one matrix.
struct RowmaxGeometry {
  int rq, kchunk, bx, by;
};
RowmaxGeometry rowmax_geometry(int op_a, int m, int k);

// Auxiliary kernels (defined in aux_kernels.hip).
hipError_t launch_splitk_reduce(const float* ws, int slices, int64_t elems, int inner, float alpha, float beta,
                                float* c, int64_t ldc, hipStream_t s);
hipError_t launch_scale_c(int64_t elems, int inner, float beta, float* c, int64_t ldc, hipStream_t s);
// Row exponents of A for the prescaled mode; `partials` (partial_bytes) holds
// the k-block maxima, so calls on one stream must not share it concurrently.
hipError_t launch_row_exponents(const float* a, int64_t lda, int op_a, int m, int k, uint32_t* row_max,
                                int* row_exp, uint32_t* partials, size_t partial_bytes, hipStream_t s);
hipError_t launch_fill_u32(uint32_t* p, int64_t n, uint32_t v, hipStream_t s);
// A split in the pre-split plan's tile order (256-row blocks x 32-deep
// k-tiles): prescaled FP16 hi / residual lo (MODE_FP16S, 32 KiB per tile,
// row_exp from the row pass), FP16 hi / 2^11-scaled residual (MODE_FP16,
// row_exp unused) or three BF16 parts (MODE_BF16, 48 KiB per tile, row_exp
// unused); k % 32 == 0.  The output needs
// presplit_bytes(m, k, mode) bytes.
size_t presplit_bytes(int m, int k, int mode);
hipError_t launch_presplit_a(const float* a, int64_t lda, int op_a, int m, int k, int mode, const int* row_exp,
                             uint4* out, hipStream_t s);
// B (FP16, k x n after op_b) in the same tile order: 256-column blocks x
// 32-deep k-tiles, 16 KiB per tile (MODE_FP16S: the FP16 bits) or 32 KiB
// (MODE_BF16: two BF16 parts); k % 32 == 0; pack_b_bytes(n, k, mode) bytes.
size_t pack_b_bytes(int n, int k, int mode);
hipError_t launch_pack_b(const uint16_t* b, int64_t ldb, int op_b, int n, int k, int mode, uint4* out,
                         hipStream_t s);
hipError_t launch_row_exp(const uint32_t* row_max, int64_t rows, int* row_exp, hipStream_t s);
hipError_t launch_grouped_rowmax(const RowmaxSeg* segs, const int* begin, int nseg, int64_t blocks, int op_a,
                                 hipStream_t s);
hipError_t launch_grouped_reduce(const ReduceSeg* segs, const int* begin, int nseg, int64_t blocks, hipStream_t s);

// Sketch generator (sketch_rng.hip).
hipError_t launch_gaussian_f16(uint16_t* out, int64_t count, uint64_t seed, float scale, hipStream_t s);
hipError_t launch_philox_words(uint32_t* out, int64_t blocks, uint64_t seed, hipStream_t s);

}  // namespace shgemm_amd
