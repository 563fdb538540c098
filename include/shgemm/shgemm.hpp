// mtk::shgemm — SHGEMM / HSGEMM for AMD Instinct MI355X (gfx950).
//
// Drop-in C++ interface for the API declared in the reference's
// include/shgemm/shgemm.hpp:1-159.  Names, enum values, argument order and
// argument meaning are kept; `set_cuda_stream` takes a hipStream_t.
//
//   C = alpha * op(A) * op(B) + beta * C          (column-major, BLAS style)
//   shgemm: A is FP32, B is FP16                   (reference shgemm.hpp:105-119)
//   hsgemm: A is FP16, B is FP32                   (reference shgemm.hpp:121-155)
//
// op_t on an operand means it is stored transposed (A: k x m, lda >= k).
// op_c == op_t stores C transposed: C is n x m, ldc >= n.
// alpha and beta are host pointers read at call time (reference shgemm.cu:532,600).
// A, B and C are device pointers owned by the caller.  All work is enqueued on
// handle.cuda_stream; nothing synchronises.
//
// Error behaviour: the reference validates nothing.  This library checks its
// arguments and throws mtk::shgemm::error (a std::runtime_error) before any
// work is enqueued; launch failures are reported the same way.
#ifndef SHGEMM_AMD_SHGEMM_HPP
#define SHGEMM_AMD_SHGEMM_HPP
#include <cstddef>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <hip/hip_runtime.h>
#include <hip/hip_fp16.h>

namespace mtk {
namespace shgemm {

class error : public std::runtime_error {
 public:
  explicit error(const std::string& what) : std::runtime_error(what) {}
};

namespace detail {
// reference shgemm.hpp:30-34
enum kernel_level {
  P0 = 0,  // small tiles
  P1 = 1,  // large tiles
  num_levels
};
struct handle_state;  // kernel table, workspace, device properties

// reference shgemm.hpp:9-28: the function type and launch descriptor of the
// reference's kernels.  Declared so that code naming them still compiles;
// this library's kernels are described behind shgemmHandle_t::state and never
// take this form.
typedef void (*kernel_func_t)(const unsigned, const unsigned, const unsigned, const float, const float* const,
                              const unsigned, const __half* const, const unsigned, const float, float* const,
                              const unsigned, float* const, const unsigned);
struct kernel {
  unsigned smem_m, smem_n, smem_k;
  unsigned num_blocks_filling = 0xffffffffu;
  unsigned block_size;
  unsigned smem_size;
  kernel_func_t func;
};
}  // namespace detail

// reference shgemm.hpp:36-39
enum operation_t { op_n, op_t };

// reference shgemm.hpp:41-44.  gfx950 has no TF32 matrix instruction:
// tf32 selects a BF16 split (A in three BF16 parts, B in two) which keeps the
// FP32 exponent range; fp16 selects the FP16 hi + 2^-11-scaled lo split.
enum tc_t { fp16, tf32 };

// reference shgemm.hpp:46-92.  The reference keeps 48 kernel descriptors in
// the struct; here they live behind `state`.  The public fields a caller may
// touch keep their names (including the reference's spelling of
// fixed_lernel_level).
struct shgemmHandle_t {
  hipStream_t cuda_stream = nullptr;
  unsigned debug_mode = 0;
  detail::kernel_level fixed_lernel_level = detail::num_levels;
  detail::handle_state* state = nullptr;
};

// reference shgemm.hpp:94-102
void create(shgemmHandle_t& handle);
void destroy(shgemmHandle_t& handle);
void enable_kernel_level_fixing(shgemmHandle_t& handle, const detail::kernel_level kernel_level);
void disable_kernel_level_fixing(shgemmHandle_t& handle);
void set_debug_mode(shgemmHandle_t& handle, const unsigned on);
void set_cuda_stream(shgemmHandle_t& handle, hipStream_t const cuda_stream);

// reference shgemm.hpp:105-119
detail::kernel_level shgemm(
    const shgemmHandle_t handle,
    const operation_t op_a,
    const operation_t op_b,
    const std::size_t m,
    const std::size_t n,
    const std::size_t k,
    const float* const alpha_ptr,
    const float* const a_ptr, const std::size_t lda,
    const half* const b_ptr, const std::size_t ldb,
    const float* const beta_ptr,
    float* const c_ptr, const std::size_t ldc,
    const tc_t compute_type,
    const operation_t op_c = op_n);

// reference shgemm.hpp:121-155: C = op(A16) op(B32) is computed as
// C^T = op'(B32) op'(A16) by swapping the operands and inverting every op.
inline detail::kernel_level hsgemm(
    const shgemmHandle_t handle,
    const operation_t op_a,
    const operation_t op_b,
    const std::size_t m,
    const std::size_t n,
    const std::size_t k,
    const float* const alpha_ptr,
    const half* const a_ptr, const std::size_t lda,
    const float* const b_ptr, const std::size_t ldb,
    const float* const beta_ptr,
    float* const c_ptr, const std::size_t ldc,
    const tc_t compute_type,
    const operation_t op_c = op_n) {
  const auto flip = [](const operation_t op) { return op == op_n ? op_t : op_n; };
  return shgemm(handle, flip(op_b), flip(op_a), n, m, k, alpha_ptr, b_ptr, ldb, a_ptr, lda,
                beta_ptr, c_ptr, ldc, compute_type, flip(op_c));
}

// ---------------------------------------------------------------------------
// Additions with no counterpart in the reference.
// ---------------------------------------------------------------------------

// Plain FP32 SGEMM on FP32 MFMA (v_mfma_f32_32x32x2_f32): the comparator the
// error-corrected SHGEMM is measured against (same layout rules as shgemm).
void sgemm_fp32(
    const shgemmHandle_t handle,
    const operation_t op_a, const operation_t op_b,
    const std::size_t m, const std::size_t n, const std::size_t k,
    const float* const alpha_ptr,
    const float* const a_ptr, const std::size_t lda,
    const float* const b_ptr, const std::size_t ldb,
    const float* const beta_ptr,
    float* const c_ptr, const std::size_t ldc,
    const operation_t op_c = op_n);

// Per-row power-of-two prescaling of the FP32 operand for the fp16 compute
// type.  On by default (MTK_SHGEMM_EXP_SCALING=0 turns it off at create, which
// gives the reference's fp16 mode with FP16's exponent range).  When on, the
// split runs on A * 2^-e and the row of C is scaled back exactly: the small,
// large and split-K tiles (plans 0-4, grouped) keep a running exponent per
// row over their k-range (rescaling their accumulator rows by exact powers of
// two when it rises); the 256x256, ring, three-stage and pre-split kernels
// and the small tile of plan 9 (plans 5-9) take each row's whole-row
// exponent from one extra pass over A.
void set_exponent_scaling(shgemmHandle_t& handle, const bool on);

// Device sharing (no counterpart in the reference): on a handle whose calls
// run beside other kernels (several streams serving requests at once), the
// plans whose 256x256 workgroups take a whole CU (variants 5, 6, 8) run as
// the large 128x256 tile instead, which leaves room on each CU for the
// concurrent work.  Off by default (the plan table is measured alone).
void set_shared_device(shgemmHandle_t& handle, const bool on);

// ---- Kernel plans --------------------------------------------------------
// A plan is a kernel variant plus a k-slice count.  Variants:
//   0 small tile (level P0), 1 large tile (P1), 2 128x128 split-K,
//   3 large-tile split-K, 4 small-tile split-K, 5 256x256 tile of four
//   128x128 wavefronts (one-accumulator modes; otherwise it runs as 1),
//   6 the 256x256 tile with a 3-slot direct-to-LDS ring for A and B
//   (prescaled fp16, op_n/op_n, m and n multiples of 256, k of 32, aligned
//   operands; otherwise it runs as 5), 7 large tile with the three-stage
//   direct-to-LDS B pipeline (prescaled fp16, op_b = op_n, aligned B, k a
//   multiple of 32; otherwise it runs as 1), 8 the 256x256 tile on a
//   pre-split A: one pass writes A's FP16 hi / residual in the tile order
//   of the kernel and one packs B the same way (4 bytes per element of A
//   and 2 per element of B, in a handle buffer grown on first use), then
//   the 256x256 tile runs without a split (prescaled fp16, k a multiple of
//   64; otherwise it runs as 5), 9 the small tile on whole-row exponents
//   from a pass over A instead of running ones (prescaled fp16; otherwise
//   it runs as 0), 10 the deep k-sliced 128x128x64 tile (eight wavefronts,
//   two k-tiles of loads in flight; prescaled fp16, split over k like 2;
//   otherwise it runs as 2), 11 the large tile as four 64x128 wavefronts,
//   split over k like 3 (prescaled fp16; otherwise it runs as 3).
// shgemm picks, in this order: a forced plan, a fixed kernel level, the
// nearest entry of the measured plan table (written by the autotuner,
// shgemm_amd/tune.py), and otherwise a cost model.
enum {
  plan_variant_p2 = 5,
  plan_variant_p2r = 6,
  plan_variant_bd = 7,
  plan_variant_ps = 8,
  plan_variant_p0w = 9,
  plan_variant_skd = 10,
  plan_variant_wsk = 11,
  num_plan_variants = 12
};
enum plan_source { plan_cost_model = 0, plan_table = 1, plan_forced = 2, plan_fixed_level = 3 };
struct plan_t {
  int variant = 1;
  int slices = 1;
  plan_source source = plan_cost_model;
};

// Every later shgemm call on this handle runs `variant` with `slices`
// k-slices (clamped to what the workspace and k allow); variant < 0 clears.
void force_plan(shgemmHandle_t& handle, const int variant, const int slices);
// Replaces the handle's plan table with the one in `path`; nullptr or ""
// leaves the handle with no table (cost model only).  create() loads
// $MTK_SHGEMM_PLANS, else plans_gfx950.tsv beside the library, if present.
void load_plans(shgemmHandle_t& handle, const char* path);

// Kernel timing (no counterpart in the reference; for profiling and
// benchmarks): HIP events on the handle's stream bracket the main GEMM
// kernel of each of the next `max_launches` shgemm / hsgemm / sgemm_fp32
// calls (not the row-exponent pass or the split-K reduction); 0 stops.
void enable_kernel_timing(shgemmHandle_t& handle, const int max_launches);
// Waits for the bracketed launches since the last query and returns their
// count and summed kernel time; the event pairs are then free again.
void kernel_timing(shgemmHandle_t& handle, int& launches, float& total_ms);
// Scratch and concurrency (no counterpart in the reference, whose handle
// holds no scratch in use).  Copies of a handle share its scratch (split-K
// partials, row exponents, the pre-split buffer); calls that use it are
// ordered on the device after the previous such call, whatever stream each
// runs on, and host threads may call through copies concurrently.  Growing
// the scratch synchronises and allocates, so it never happens while the
// handle's stream captures a graph: there the pre-split plan runs as the
// 256x256 tile when its buffer is too small (bit-identical), and a call whose
// row scratch would have to grow throws.  reserve() grows the scratch for
// calls up to m x n x k beforehand.  Graphs captured with a handle must not be
// replayed while other calls through that handle run.
void reserve(shgemmHandle_t& handle, const std::size_t m, const std::size_t n, const std::size_t k);

// The plan shgemm would use for these arguments (nothing is launched).
plan_t query_plan(const shgemmHandle_t handle, const operation_t op_a, const operation_t op_b,
                  const std::size_t m, const std::size_t n, const std::size_t k, const tc_t compute_type,
                  const operation_t op_c = op_n);

// ---- Grouped SHGEMM ------------------------------------------------------
// Many independent problems C_i = alpha_i op(A_i) op(B_i) + beta_i C_i that
// share op_a, op_b, op_c and compute_type, run by persistent kernels that
// take (problem, k-slice, tile) work units from a queue, longest first.  Each
// problem gets a tile class (0 small, 1 128x128, 2 the large tile) and a
// k-slice count from the plan table's grouped entries (or the arguments);
// class 3, a 128x512 tile, exists for the prescaled fp16 mode only.
// create_group validates every problem, allocates the group's descriptors
// and split-K partials, and captures the whole sequence (one launch per tile
// class used, split-K reduction; the tile classes keep running row
// exponents) in one HIP graph; run_group replays it on `stream`.  Results are bit-identical
// to running each problem alone with the same tile depth and slice count.
struct group_problem {
  std::size_t m = 0, n = 0, k = 0;
  const float* A = nullptr;
  std::size_t lda = 0;
  const half* B = nullptr;
  std::size_t ldb = 0;
  float* C = nullptr;
  std::size_t ldc = 0;
  float alpha = 1.0f, beta = 0.0f;
};
struct group_t;
group_t* create_group(const shgemmHandle_t& handle, const operation_t op_a, const operation_t op_b,
                      const group_problem* problems, const std::size_t count, const tc_t compute_type,
                      const operation_t op_c = op_n, const int* tile_class = nullptr, const int* slices = nullptr);
// use_graph = false enqueues the same launches one by one (no graph).
void run_group(group_t* group, hipStream_t stream, const bool use_graph = true);
void destroy_group(group_t* group);
// Per-problem choices (tile class, slices) and the group's work-unit count.
void group_plan(const group_t* group, int* tile_class, int* slices, std::size_t* units);

// ---- Batches -----------------------------------------------------------------
// Problems that change from call to call (a serving loop's admitted
// requests) run as one grouped launch without allocating: create_batch
// allocates `slots` device regions of `bytes` each (descriptors and split-K
// partials of one batch) and pinned staging for up to max_problems
// descriptors; run_batch plans the problems exactly as create_group would
// (tile class and k-slices from the plan table's grouped rows), copies their
// descriptors into the next slot asynchronously and enqueues fill, the
// persistent launches and the split-K reduction on `stream`.  A slot is
// reused only after its previous batch has finished (run_batch waits on the
// host for that when all slots are in flight).  Results are bit-identical to
// the same problems in a group and to single calls with the matching plans.
// run_batch throws if the batch needs more than `bytes` of device scratch.
struct batch_t;
batch_t* create_batch(const shgemmHandle_t& handle, const operation_t op_a, const operation_t op_b,
                      const tc_t compute_type, const operation_t op_c, const std::size_t max_problems,
                      const std::size_t bytes, const int slots = 2);
void run_batch(batch_t* batch, const group_problem* problems, const std::size_t count, hipStream_t stream);
// Tile class and slices of every problem of the last run_batch, and its work units.
void batch_plan(const batch_t* batch, int* tile_class, int* slices, std::size_t* units);
// Waits for the batch's launches in flight, then frees it.
void destroy_batch(batch_t* batch);

}  // namespace shgemm
}  // namespace mtk
#endif
