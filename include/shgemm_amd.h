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

/* Message of the last failure on this thread ("" if none). */
const char* shgemm_amd_last_error(void);

/* Library version string. */
const char* shgemm_amd_version(void);

#ifdef __cplusplus
}
#endif
#endif
