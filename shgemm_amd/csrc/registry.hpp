// Kernel table: every compiled kernel variant, filled by the instantiation
// units (inst_*.hip) and read by the host dispatcher.
#pragma once
#include <hip/hip_runtime.h>
#include <cstddef>
#include "gemm_common.hpp"

namespace shgemm_amd {

struct KernelEntry {
  const void* fn = nullptr;
  size_t lds = 0;      // dynamic LDS bytes
  int bm = 0, bn = 0, bk = 0;
  int mode = -1;
  int threads = 256;
};

enum : int { LEVEL_P0 = 0, LEVEL_P1 = 1 };

// Kernel variants per (op_a, op_b, op_c, mode): tile config x split-K.
enum : int {
  V_P0 = 0,    // small tile
  V_P1 = 1,    // large tile
  V_SK = 2,    // 128 x 128 tile, split-K
  V_P1SK = 3,  // large tile, split-K
  V_P0SK = 4,  // small tile, split-K
  kNumVariants = 5
};
inline bool variant_is_split(int v) { return v == V_SK || v == V_P1SK || v == V_P0SK; }
inline int variant_level(int v) { return (v == V_P0 || v == V_P0SK) ? LEVEL_P0 : LEVEL_P1; }

struct KernelTable {
  // [op_a][op_b][op_c][mode][variant]
  KernelEntry shgemm[2][2][2][kNumModes][kNumVariants];
  // [op_a][op_b][op_c]
  KernelEntry sgemm[2][2][2];
  // deep-prefetch prescaled fp16 kernel (op_b = n, full k-tiles): [op_a][op_c]
  KernelEntry shgemm_bd[2][2];
  // Candidate large-tile configurations of the fp16 op_n/op_n/C-op_n kernel,
  // timed by the tile sweep and selectable with MTK_SHGEMM_TUNE_P1=<index>.
  static constexpr int kMaxTune = 16;
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
void register_all(KernelTable&);

// Auxiliary kernels (defined in aux_kernels.hip).
hipError_t launch_splitk_reduce(const float* ws, int slices, int64_t elems, int inner, float alpha, float beta,
                                float* c, int64_t ldc, hipStream_t s);
hipError_t launch_scale_c(int64_t elems, int inner, float beta, float* c, int64_t ldc, hipStream_t s);
hipError_t launch_row_exponents(const float* a, int64_t lda, int op_a, int m, int k, uint32_t* row_max,
                                int* row_exp, hipStream_t s);

}  // namespace shgemm_amd
