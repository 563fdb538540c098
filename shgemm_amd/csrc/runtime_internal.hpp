// Host-side internals shared by the runtime units (runtime.cpp, plans.cpp,
// group.cpp): the state behind a handle and small launch helpers.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include <mutex>
#include <string>
#include <utility>
#include <vector>

#include "../../include/shgemm/shgemm.hpp"
#include "errors.hpp"
#include "plans.hpp"
#include "registry.hpp"

namespace mtk {
namespace shgemm {
namespace detail {

// Copies of a shgemmHandle_t share one handle_state, so calls through them may
// come from several host threads and on several streams.  `mu` serialises the
// host side of each call; the scratch buffers below are ordered on the device
// by `scratch_ev`: every call that uses them records it on its stream
// afterwards, and a call on another stream first waits for it, so the calls
// that touch the scratch run one after another whatever their streams.
struct handle_state {
  int device = 0;
  int num_cus = 0;
  shgemm_amd::KernelTable table;
  std::mutex mu;
  hipEvent_t scratch_ev = nullptr;   // recorded after the last call that used the scratch
  hipStream_t scratch_stream = nullptr;  // stream of that call
  bool scratch_used = false;
  float* ws = nullptr;  // split-K partials
  size_t ws_bytes = 0;
  uint32_t* row_part = nullptr;  // k-block partial maxima of the row-exponent pass
  size_t row_part_bytes = 0;
  uint32_t* row_max = nullptr;  // prescale scratch
  int* row_exp = nullptr;
  size_t row_cap = 0;
  bool exp_scaling = true;
  bool shared_device = false;  // set_shared_device: no whole-CU 256x256 plans
  int tune_p1 = -1;         // MTK_SHGEMM_TUNE_P1: tile-sweep candidate for fp16 nn
  bool use_bd = false;      // MTK_SHGEMM_BD=1: deep-prefetch kernel where the cost model picks P1
  int forced_variant = -1;  // force_plan / MTK_SHGEMM_PLAN=<variant>,<slices>
  int forced_slices = 1;
  std::vector<shgemm_amd::PlanRow> plans;  // measured plan table
  std::string plans_path;
  // pre-split plan: A packed as FP16 hi / lo tiles, then B packed (grown on demand)
  uint16_t* ps_buf = nullptr;
  size_t ps_bytes = 0;
  // enable_kernel_timing: event pairs recorded around the main GEMM kernel
  std::vector<std::pair<hipEvent_t, hipEvent_t>> timing;
  int timing_used = 0;
};

}  // namespace detail
}  // namespace shgemm
}  // namespace mtk

namespace shgemm_amd {

inline void check_hip(hipError_t e, const char* what) {
  if (e != hipSuccess) throw hip_failure(std::string(what) + ": " + hipGetErrorString(e));
}

inline bool aligned(const void* p, size_t bytes) { return (reinterpret_cast<uintptr_t>(p) % bytes) == 0; }

// Tile ordering: group this many tile rows so tiles in flight share panels.
inline int group_rows(int tiles_m) { return tiles_m < 8 ? tiles_m : 8; }

void set_lds_limit(const KernelEntry& e);

// Argument checks shared by shgemm, sgemm_fp32 and the grouped entry points
// (throws mtk::shgemm::error).
void validate_args(const mtk::shgemm::shgemmHandle_t& handle, mtk::shgemm::operation_t op_a,
                   mtk::shgemm::operation_t op_b, mtk::shgemm::operation_t op_c, size_t m, size_t n, size_t k,
                   const float* alpha, const void* a, size_t lda, const void* b, size_t ldb, const float* beta,
                   const float* c, size_t ldc);

// Vector-load flags of an SHGEMM operand pair (16-byte FP32 A; 16- or 8-byte FP16 B).
inline int vec_a_ok(const float* a, size_t lda) { return aligned(a, 16) && lda % 4 == 0; }
inline int vec_b_ok(const void* b, size_t ldb, int op_b) {
  return op_b == OP_N ? (aligned(b, 16) && ldb % 8 == 0) : (aligned(b, 8) && ldb % 4 == 0);
}

// Split-K partials of one call (the handle's workspace and the split-K part of
// the allocation-free scratch have this size, so both choose the same plans).
constexpr size_t kWorkspaceBytes = size_t(64) << 20;
// k-block partial maxima of the row-exponent pass (it sizes its k-blocks to fit).
constexpr size_t kRowPartialBytes = size_t(16) << 20;

// Device memory one call may use beyond its operands, when it does not come
// from the handle (which grows its own on demand): split-K partials of
// kWorkspaceBytes, row-pass partial maxima of kRowPartialBytes, and per-row
// maxima and exponents for row_cap rows.
struct CallScratch {
  float* ws = nullptr;
  uint32_t* row_part = nullptr;
  uint32_t* row_max = nullptr;
  int* row_exp = nullptr;
  size_t row_cap = 0;
};

// The allocation-free entry point's scratch on the current device: a
// file-scope device array (static_scratch.hip) whose address is looked up
// once per device with that device current.  init_static_scratch() does the
// lookup for the current device (create() calls it); static_scratch()
// returns the views, looking the address up on a device's first use when no
// handle was created there.
void init_static_scratch();
CallScratch static_scratch();
size_t static_scratch_bytes();

// shgemm on the static scratch: nothing is allocated, freed or synchronised.
// The pre-split plan (whose buffer grows with A) runs as the 256x256 tile it
// is bit-identical to; a call whose plan needs per-row exponents for more
// rows than the scratch holds throws mtk::shgemm::error.
mtk::shgemm::detail::kernel_level shgemm_no_alloc(
    const mtk::shgemm::shgemmHandle_t handle, mtk::shgemm::operation_t op_a, mtk::shgemm::operation_t op_b,
    std::size_t m, std::size_t n, std::size_t k, const float* alpha, const float* a, std::size_t lda, const half* b,
    std::size_t ldb, const float* beta, float* c, std::size_t ldc, mtk::shgemm::tc_t compute_type,
    mtk::shgemm::operation_t op_c);

}  // namespace shgemm_amd
