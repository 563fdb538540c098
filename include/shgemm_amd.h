/* shgemm_amd.h — C ABI of the MI355X SHGEMM library (libshgemm_amd.so).
 *
 * Plain pointers and sizes only.  Every entry point returns 0 on success or a
 * negative status; shgemm_amd_last_error() returns the message of the last
 * failure on the calling thread.  Device pointers are HIP device pointers;
 * alpha/beta are host pointers read at call time, as in the reference.
 *
 * Each function names the reference interface it replaces
 * (paths relative to the reference repository).
 */
#ifndef SHGEMM_AMD_H
#define SHGEMM_AMD_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SHGEMM_AMD_OK 0
#define SHGEMM_AMD_EINVAL (-1)  /* invalid argument */
#define SHGEMM_AMD_EHIP (-2)    /* HIP runtime error (allocation, launch) */
#define SHGEMM_AMD_EINTERNAL (-3)

/* operation_t (include/shgemm/shgemm.hpp:36-39) */
#define SHGEMM_AMD_OP_N 0
#define SHGEMM_AMD_OP_T 1
/* tc_t (include/shgemm/shgemm.hpp:41-44) */
#define SHGEMM_AMD_FP16 0
#define SHGEMM_AMD_TF32 1

typedef struct shgemm_amd_handle shgemm_amd_handle;

/* mtk::shgemm::create (include/shgemm/shgemm.hpp:94; src/shgemm.cu:401-490) */
int shgemm_amd_create(shgemm_amd_handle** out);
/* mtk::shgemm::destroy (include/shgemm/shgemm.hpp:95; src/shgemm.cu:492-496) */
int shgemm_amd_destroy(shgemm_amd_handle* h);
/* mtk::shgemm::set_cuda_stream (include/shgemm/shgemm.hpp:102; src/shgemm.cu:498-503);
 * `stream` is a hipStream_t (NULL = default stream). */
int shgemm_amd_set_stream(shgemm_amd_handle* h, void* stream);
/* mtk::shgemm::enable_kernel_level_fixing (include/shgemm/shgemm.hpp:97; src/shgemm.cu:505-509) */
int shgemm_amd_enable_kernel_level_fixing(shgemm_amd_handle* h, int level);
/* mtk::shgemm::disable_kernel_level_fixing (include/shgemm/shgemm.hpp:98; src/shgemm.cu:511-515) */
int shgemm_amd_disable_kernel_level_fixing(shgemm_amd_handle* h);
/* mtk::shgemm::set_debug_mode (include/shgemm/shgemm.hpp:100; src/shgemm.cu:688-690) */
int shgemm_amd_set_debug_mode(shgemm_amd_handle* h, unsigned on);

/* mtk::shgemm::shgemm (include/shgemm/shgemm.hpp:105-119; src/shgemm.cu:517-686).
 * A FP32, B FP16 (IEEE binary16 bits).  *level_out (may be NULL) receives the
 * kernel_level the reference returns. */
int shgemm_amd_shgemm(shgemm_amd_handle* h, int op_a, int op_b,
                      size_t m, size_t n, size_t k,
                      const float* alpha, const float* a, size_t lda,
                      const uint16_t* b, size_t ldb,
                      const float* beta, float* c, size_t ldc,
                      int compute_type, int op_c, int* level_out);

/* No counterpart in the reference: shgemm that allocates nothing, for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  Same arguments,
 * semantics and results as shgemm_amd_shgemm, except that its scratch (split-K
 * partials and row exponents) is one static device array of 96 MiB that every
 * GPU the library runs on holds until the process ends, instead of h's
 * buffers: no hipMalloc, hipFree or synchronisation happens in the call.  The
 * pre-split plan (whose buffer grows with A) runs as the 256x256 tile, which
 * gives the same bits.  Limit: a call whose plan reads per-row exponents
 * (plans 5-9 of the prescaled fp16 compute type) with m above 2^21 rows
 * returns SHGEMM_AMD_EINVAL.  Every call on one GPU shares that array, so
 * calls that overlap on one GPU (other streams, other threads) race. */
int shgemm_amd_shgemm_no_alloc(shgemm_amd_handle* h, int op_a, int op_b,
                               size_t m, size_t n, size_t k,
                               const float* alpha, const float* a, size_t lda,
                               const uint16_t* b, size_t ldb,
                               const float* beta, float* c, size_t ldc,
                               int compute_type, int op_c, int* level_out);

/* mtk::shgemm::hsgemm (include/shgemm/shgemm.hpp:121-155). A FP16, B FP32. */
int shgemm_amd_hsgemm(shgemm_amd_handle* h, int op_a, int op_b,
                      size_t m, size_t n, size_t k,
                      const float* alpha, const uint16_t* a, size_t lda,
                      const float* b, size_t ldb,
                      const float* beta, float* c, size_t ldc,
                      int compute_type, int op_c, int* level_out);

/* No counterpart in the reference (its comparator is cuBLAS SGEMM,
 * test/shgemm_test.cu:150-166): the library's own plain FP32 SGEMM. */
int shgemm_amd_sgemm_fp32(shgemm_amd_handle* h, int op_a, int op_b,
                          size_t m, size_t n, size_t k,
                          const float* alpha, const float* a, size_t lda,
                          const float* b, size_t ldb,
                          const float* beta, float* c, size_t ldc, int op_c);

/* No counterpart in the reference: per-row power-of-two prescaling of the
 * FP32 operand for the fp16 compute type (widens its exponent range). */
int shgemm_amd_set_exponent_scaling(shgemm_amd_handle* h, int on);
/* Calls on h share the device with concurrent kernels: plans with
 * whole-CU 256x256 workgroups run as the 128x256 tile (no counterpart in
 * the reference; for servers running several streams at once). */
int shgemm_amd_set_shared_device(shgemm_amd_handle* h, int on);

/* ---- Kernel plans ---------------------------------------------------------
 * Replace the reference's static per-architecture tile tables and its
 * level/split heuristic (src/shgemm.cu:427-489, 532, 643-656) with a measured
 * plan table (written by shgemm_amd/tune.py) and a cost model.
 * Plan variants: 0 small tile (P0), 1 large tile (P1), 2 128x128 split-K,
 * 3 large-tile split-K, 4 small-tile split-K, 5 256x256 tile of four
 * 128x128 wavefronts, 6 the same tile with a 3-slot direct-to-LDS ring for
 * A and B, 7 large tile with the three-stage direct-to-LDS B pipeline,
 * 8 the 256x256 tile on a pre-split A (A split and B packed once per call
 * into a handle buffer of 4 bytes per element of A and 2 per element of B;
 * prescaled fp16, k % 64 == 0), 9 the small tile on whole-row exponents
 * from a pass over A (prescaled fp16)
 * (5-9 run as the nearest applicable variant where they do not apply).
 * Plan sources (query): 0 cost model, 1 plan table, 2 forced, 3 fixed level. */
#define SHGEMM_AMD_PLAN_P2 5
#define SHGEMM_AMD_PLAN_P2R 6
#define SHGEMM_AMD_PLAN_BD 7
#define SHGEMM_AMD_PLAN_PS 8
#define SHGEMM_AMD_PLAN_P0W 9
/* Force variant/slices for every later call on h (variant < 0 clears). */
int shgemm_amd_force_plan(shgemm_amd_handle* h, int variant, int slices);
/* Kernel timing (no counterpart in the reference; for profiling): HIP events
 * on h's stream around the main GEMM kernel of the next max_launches calls
 * (0 stops).  shgemm_amd_kernel_timing waits for them and returns the count
 * and the summed milliseconds since the last query. */
int shgemm_amd_enable_kernel_timing(shgemm_amd_handle* h, int max_launches);
int shgemm_amd_kernel_timing(shgemm_amd_handle* h, int* launches, float* total_ms);
/* Grow h's scratch (row exponents, pre-split buffer) for calls up to
 * m x n x k, so that such calls allocate nothing, e.g. while h's stream
 * captures a graph (no counterpart in the reference). */
int shgemm_amd_reserve(shgemm_amd_handle* h, size_t m, size_t n, size_t k);
/* Replace h's plan table with the file at path (NULL or "": no table). */
int shgemm_amd_load_plans(shgemm_amd_handle* h, const char* path);
/* The plan shgemm would run for these arguments, for aligned operands. */
int shgemm_amd_query_plan(shgemm_amd_handle* h, int op_a, int op_b, size_t m, size_t n, size_t k,
                          int compute_type, int op_c, int* variant, int* slices, int* source);
/* Parse a plan table file on the host (no GPU needed); *rows gets its row count. */
int shgemm_amd_check_plan_file(const char* path, int* rows);

/* ---- Grouped SHGEMM ---------------------------------------------------------
 * No counterpart in the reference (one problem per call there,
 * include/shgemm/shgemm.hpp:105-119).  Each problem follows shgemm's
 * semantics; all share op_a, op_b, op_c and compute_type.  create plans the
 * problems (tile class 0 small / 1 128x128 / 2 large / 3 128x512 for the
 * prescaled fp16 mode, and k-slices: from
 * tile_class/slices when not NULL, else the plan table), allocates the
 * group's device scratch and captures one HIP graph; run replays it (or, with
 * use_graph = 0, enqueues the same launches one by one) on `stream`.  Runs of
 * one group must not overlap (they share its scratch). */
typedef struct shgemm_amd_problem {
  size_t m, n, k;
  const float* a;
  size_t lda;
  const uint16_t* b;  /* FP16 bits */
  size_t ldb;
  float* c;
  size_t ldc;
  float alpha, beta;
} shgemm_amd_problem;
typedef struct shgemm_amd_group shgemm_amd_group;
int shgemm_amd_group_create(shgemm_amd_handle* h, int op_a, int op_b, int op_c, int compute_type,
                            const shgemm_amd_problem* problems, size_t count,
                            const int* tile_class, const int* slices, shgemm_amd_group** out);
int shgemm_amd_group_run(shgemm_amd_group* g, void* stream, int use_graph);
/* Per-problem tile class and slice count (-1 / 0 for problems with no MFMA
 * work), and the total number of work units. Arrays may be NULL. */
int shgemm_amd_group_plan(shgemm_amd_group* g, int* tile_class, int* slices, size_t* units);
int shgemm_amd_group_destroy(shgemm_amd_group* g);

/* ---- Batches -------------------------------------------------------------------
 * No counterpart in the reference: grouped launches of problems that change
 * from call to call (a serving loop's admitted requests), with no allocation
 * per call.  create allocates `slots` device regions of `bytes` each and
 * pinned staging for max_problems descriptors; run plans the problems as
 * group_create would (plan table's grouped rows), copies the descriptors
 * into the next slot asynchronously and enqueues the launches on `stream`
 * (waiting on the host only when that slot's previous batch is still in
 * flight); SHGEMM_AMD_EINVAL when a batch needs more than `bytes`.  plan
 * reports the last run's per-problem tile class and slices. */
typedef struct shgemm_amd_batch shgemm_amd_batch;
int shgemm_amd_batch_create(shgemm_amd_handle* h, int op_a, int op_b, int op_c, int compute_type,
                            size_t max_problems, size_t bytes, int slots, shgemm_amd_batch** out);
int shgemm_amd_batch_run(shgemm_amd_batch* b, const shgemm_amd_problem* problems, size_t count, void* stream);
int shgemm_amd_batch_plan(shgemm_amd_batch* b, int* tile_class, int* slices, size_t* units);
int shgemm_amd_batch_destroy(shgemm_amd_batch* b);

/* ---- Sketch matrices -----------------------------------------------------
 * No counterpart in the reference (its tests fill inputs with cuRAND Philox,
 * test/shgemm_test.cu:247-252).  Counter-based Philox4x32-10 (Salmon et al.,
 * SC'11) with key = seed and counter = (block mod 2^32, block / 2^32, 0, 0);
 * element i depends only on (seed, i).  Enqueued on `stream` (NULL = default)
 * on the current device. */
/* out[i] = FP16(scale * N(0,1)): Box-Muller on word pairs (0,1), (2,3) of
 * block i/4, with u = ((w >> 9) + 1/2) * 2^-23 in (0, 1). */
int shgemm_amd_gaussian_f16(uint16_t* out, size_t count, uint64_t seed, float scale, void* stream);
/* out[4b + j] = word j of block b, b < blocks (out 16-byte aligned). */
int shgemm_amd_philox_words(uint32_t* out, size_t blocks, uint64_t seed, void* stream);

/* Message of the last failure on this thread ("" if none). */
const char* shgemm_amd_last_error(void);

/* Library version string. */
const char* shgemm_amd_version(void);

#ifdef __cplusplus
}
#endif
#endif
