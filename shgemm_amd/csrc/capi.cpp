// C ABI over the mtk::shgemm C++ API (declared in include/shgemm_amd.h).
// Exceptions never cross this boundary: each entry point returns a status and
// records the message for shgemm_amd_last_error().
#include <new>
#include <string>
#include <vector>

#include "../../include/shgemm/shgemm.hpp"
#include "../../include/shgemm_amd.h"
#include "errors.hpp"
#include "plans.hpp"
#include "registry.hpp"
#include "runtime_internal.hpp"

struct shgemm_amd_handle {
  mtk::shgemm::shgemmHandle_t h;
};

struct shgemm_amd_group {
  mtk::shgemm::group_t* g = nullptr;
};

struct shgemm_amd_batch {
  mtk::shgemm::batch_t* b = nullptr;
  std::vector<mtk::shgemm::group_problem> probs;  // reused conversion buffer
};

namespace {
thread_local std::string g_last_error;

template <class F>
int guarded(F&& f) {
  try {
    f();
    g_last_error.clear();
    return SHGEMM_AMD_OK;
  } catch (const shgemm_amd::hip_failure& e) {
    g_last_error = e.what();
    return SHGEMM_AMD_EHIP;
  } catch (const mtk::shgemm::error& e) {
    g_last_error = e.what();
    return SHGEMM_AMD_EINVAL;
  } catch (const std::bad_alloc&) {
    g_last_error = "host allocation failed";
    return SHGEMM_AMD_EINTERNAL;
  } catch (const std::exception& e) {
    g_last_error = e.what();
    return SHGEMM_AMD_EINTERNAL;
  } catch (...) {
    g_last_error = "unknown error";
    return SHGEMM_AMD_EINTERNAL;
  }
}

mtk::shgemm::operation_t to_op(int op) {
  if (op == SHGEMM_AMD_OP_N) return mtk::shgemm::op_n;
  if (op == SHGEMM_AMD_OP_T) return mtk::shgemm::op_t;
  throw mtk::shgemm::error("invalid operation (expected 0 = N or 1 = T)");
}

mtk::shgemm::tc_t to_tc(int t) {
  if (t == SHGEMM_AMD_FP16) return mtk::shgemm::fp16;
  if (t == SHGEMM_AMD_TF32) return mtk::shgemm::tf32;
  throw mtk::shgemm::error("invalid compute type (expected 0 = fp16 or 1 = tf32)");
}

void need(const shgemm_amd_handle* h) {
  if (h == nullptr) throw mtk::shgemm::error("handle is null");
}
}  // namespace

extern "C" {

int shgemm_amd_create(shgemm_amd_handle** out) {
  return guarded([&] {
    if (out == nullptr) throw mtk::shgemm::error("out is null");
    auto* h = new shgemm_amd_handle();
    try {
      mtk::shgemm::create(h->h);
    } catch (...) {
      delete h;
      throw;
    }
    *out = h;
  });
}

int shgemm_amd_destroy(shgemm_amd_handle* h) {
  return guarded([&] {
    if (h == nullptr) return;
    mtk::shgemm::destroy(h->h);
    delete h;
  });
}

int shgemm_amd_set_stream(shgemm_amd_handle* h, void* stream) {
  return guarded([&] {
    need(h);
    mtk::shgemm::set_cuda_stream(h->h, static_cast<hipStream_t>(stream));
  });
}

int shgemm_amd_enable_kernel_level_fixing(shgemm_amd_handle* h, int level) {
  return guarded([&] {
    need(h);
    if (level != 0 && level != 1) throw mtk::shgemm::error("kernel level must be 0 (P0) or 1 (P1)");
    mtk::shgemm::enable_kernel_level_fixing(h->h, static_cast<mtk::shgemm::detail::kernel_level>(level));
  });
}

int shgemm_amd_disable_kernel_level_fixing(shgemm_amd_handle* h) {
  return guarded([&] {
    need(h);
    mtk::shgemm::disable_kernel_level_fixing(h->h);
  });
}

int shgemm_amd_set_debug_mode(shgemm_amd_handle* h, unsigned on) {
  return guarded([&] {
    need(h);
    mtk::shgemm::set_debug_mode(h->h, on);
  });
}

int shgemm_amd_shgemm(shgemm_amd_handle* h, int op_a, int op_b, size_t m, size_t n, size_t k, const float* alpha,
                      const float* a, size_t lda, const uint16_t* b, size_t ldb, const float* beta, float* c,
                      size_t ldc, int compute_type, int op_c, int* level_out) {
  return guarded([&] {
    need(h);
    const auto lvl = mtk::shgemm::shgemm(h->h, to_op(op_a), to_op(op_b), m, n, k, alpha, a, lda,
                                         reinterpret_cast<const half*>(b), ldb, beta, c, ldc, to_tc(compute_type),
                                         to_op(op_c));
    if (level_out) *level_out = static_cast<int>(lvl);
  });
}

int shgemm_amd_shgemm_no_alloc(shgemm_amd_handle* h, int op_a, int op_b, size_t m, size_t n, size_t k,
                               const float* alpha, const float* a, size_t lda, const uint16_t* b, size_t ldb,
                               const float* beta, float* c, size_t ldc, int compute_type, int op_c, int* level_out) {
  return guarded([&] {
    need(h);
    const auto lvl = shgemm_amd::shgemm_no_alloc(h->h, to_op(op_a), to_op(op_b), m, n, k, alpha, a, lda,
                                                 reinterpret_cast<const half*>(b), ldb, beta, c, ldc,
                                                 to_tc(compute_type), to_op(op_c));
    if (level_out) *level_out = static_cast<int>(lvl);
  });
}

int shgemm_amd_hsgemm(shgemm_amd_handle* h, int op_a, int op_b, size_t m, size_t n, size_t k, const float* alpha,
                      const uint16_t* a, size_t lda, const float* b, size_t ldb, const float* beta, float* c,
                      size_t ldc, int compute_type, int op_c, int* level_out) {
  return guarded([&] {
    need(h);
    const auto lvl = mtk::shgemm::hsgemm(h->h, to_op(op_a), to_op(op_b), m, n, k, alpha,
                                         reinterpret_cast<const half*>(a), lda, b, ldb, beta, c, ldc,
                                         to_tc(compute_type), to_op(op_c));
    if (level_out) *level_out = static_cast<int>(lvl);
  });
}

int shgemm_amd_sgemm_fp32(shgemm_amd_handle* h, int op_a, int op_b, size_t m, size_t n, size_t k,
                          const float* alpha, const float* a, size_t lda, const float* b, size_t ldb,
                          const float* beta, float* c, size_t ldc, int op_c) {
  return guarded([&] {
    need(h);
    mtk::shgemm::sgemm_fp32(h->h, to_op(op_a), to_op(op_b), m, n, k, alpha, a, lda, b, ldb, beta, c, ldc,
                            to_op(op_c));
  });
}

int shgemm_amd_set_exponent_scaling(shgemm_amd_handle* h, int on) {
  return guarded([&] {
    need(h);
    mtk::shgemm::set_exponent_scaling(h->h, on != 0);
  });
}

int shgemm_amd_set_shared_device(shgemm_amd_handle* h, int on) {
  return guarded([&] {
    need(h);
    mtk::shgemm::set_shared_device(h->h, on != 0);
  });
}

int shgemm_amd_force_plan(shgemm_amd_handle* h, int variant, int slices) {
  return guarded([&] {
    need(h);
    mtk::shgemm::force_plan(h->h, variant, slices);
  });
}

int shgemm_amd_enable_kernel_timing(shgemm_amd_handle* h, int max_launches) {
  return guarded([&] {
    need(h);
    mtk::shgemm::enable_kernel_timing(h->h, max_launches);
  });
}

int shgemm_amd_kernel_timing(shgemm_amd_handle* h, int* launches, float* total_ms) {
  return guarded([&] {
    need(h);
    int n = 0;
    float ms = 0.0f;
    mtk::shgemm::kernel_timing(h->h, n, ms);
    if (launches) *launches = n;
    if (total_ms) *total_ms = ms;
  });
}

int shgemm_amd_reserve(shgemm_amd_handle* h, size_t m, size_t n, size_t k) {
  return guarded([&] {
    need(h);
    mtk::shgemm::reserve(h->h, m, n, k);
  });
}

int shgemm_amd_load_plans(shgemm_amd_handle* h, const char* path) {
  return guarded([&] {
    need(h);
    mtk::shgemm::load_plans(h->h, path);
  });
}

int shgemm_amd_query_plan(shgemm_amd_handle* h, int op_a, int op_b, size_t m, size_t n, size_t k, int compute_type,
                          int op_c, int* variant, int* slices, int* source) {
  return guarded([&] {
    need(h);
    const auto pl =
        mtk::shgemm::query_plan(h->h, to_op(op_a), to_op(op_b), m, n, k, to_tc(compute_type), to_op(op_c));
    if (variant) *variant = pl.variant;
    if (slices) *slices = pl.slices;
    if (source) *source = static_cast<int>(pl.source);
  });
}

int shgemm_amd_check_plan_file(const char* path, int* rows) {
  return guarded([&] {
    if (path == nullptr) throw mtk::shgemm::error("path is null");
    std::vector<shgemm_amd::PlanRow> v;
    std::string err;
    if (!shgemm_amd::parse_plan_file(path, v, err)) throw mtk::shgemm::error(err);
    if (rows) *rows = (int)v.size();
  });
}

int shgemm_amd_group_create(shgemm_amd_handle* h, int op_a, int op_b, int op_c, int compute_type,
                            const shgemm_amd_problem* problems, size_t count, const int* tile_class,
                            const int* slices, shgemm_amd_group** out) {
  return guarded([&] {
    need(h);
    if (out == nullptr) throw mtk::shgemm::error("out is null");
    if (problems == nullptr && count > 0) throw mtk::shgemm::error("problems is null");
    std::vector<mtk::shgemm::group_problem> v(count);
    for (size_t i = 0; i < count; ++i) {
      const shgemm_amd_problem& q = problems[i];
      v[i].m = q.m;
      v[i].n = q.n;
      v[i].k = q.k;
      v[i].A = q.a;
      v[i].lda = q.lda;
      v[i].B = reinterpret_cast<const half*>(q.b);
      v[i].ldb = q.ldb;
      v[i].C = q.c;
      v[i].ldc = q.ldc;
      v[i].alpha = q.alpha;
      v[i].beta = q.beta;
    }
    auto* g = new shgemm_amd_group();
    try {
      g->g = mtk::shgemm::create_group(h->h, to_op(op_a), to_op(op_b), v.data(), count, to_tc(compute_type),
                                       to_op(op_c), tile_class, slices);
    } catch (...) {
      delete g;
      throw;
    }
    *out = g;
  });
}

namespace {
void to_group_problems(const shgemm_amd_problem* problems, size_t count, std::vector<mtk::shgemm::group_problem>& v) {
  v.resize(count);
  for (size_t i = 0; i < count; ++i) {
    const shgemm_amd_problem& q = problems[i];
    v[i].m = q.m;
    v[i].n = q.n;
    v[i].k = q.k;
    v[i].A = q.a;
    v[i].lda = q.lda;
    v[i].B = reinterpret_cast<const half*>(q.b);
    v[i].ldb = q.ldb;
    v[i].C = q.c;
    v[i].ldc = q.ldc;
    v[i].alpha = q.alpha;
    v[i].beta = q.beta;
  }
}
}  // namespace

int shgemm_amd_batch_create(shgemm_amd_handle* h, int op_a, int op_b, int op_c, int compute_type,
                            size_t max_problems, size_t bytes, int slots, shgemm_amd_batch** out) {
  return guarded([&] {
    need(h);
    if (out == nullptr) throw mtk::shgemm::error("out is null");
    auto* b = new shgemm_amd_batch();
    try {
      b->b = mtk::shgemm::create_batch(h->h, to_op(op_a), to_op(op_b), to_tc(compute_type), to_op(op_c),
                                       max_problems, bytes, slots);
    } catch (...) {
      delete b;
      throw;
    }
    *out = b;
  });
}

int shgemm_amd_batch_run(shgemm_amd_batch* b, const shgemm_amd_problem* problems, size_t count, void* stream) {
  return guarded([&] {
    if (b == nullptr) throw mtk::shgemm::error("batch is null");
    if (problems == nullptr && count > 0) throw mtk::shgemm::error("problems is null");
    to_group_problems(problems, count, b->probs);
    mtk::shgemm::run_batch(b->b, b->probs.data(), count, static_cast<hipStream_t>(stream));
  });
}

int shgemm_amd_batch_plan(shgemm_amd_batch* b, int* tile_class, int* slices, size_t* units) {
  return guarded([&] {
    if (b == nullptr) throw mtk::shgemm::error("batch is null");
    mtk::shgemm::batch_plan(b->b, tile_class, slices, units);
  });
}

int shgemm_amd_batch_destroy(shgemm_amd_batch* b) {
  return guarded([&] {
    if (b == nullptr) return;
    mtk::shgemm::destroy_batch(b->b);
    delete b;
  });
}

int shgemm_amd_group_run(shgemm_amd_group* g, void* stream, int use_graph) {
  return guarded([&] {
    if (g == nullptr) throw mtk::shgemm::error("group is null");
    mtk::shgemm::run_group(g->g, static_cast<hipStream_t>(stream), use_graph != 0);
  });
}

int shgemm_amd_group_plan(shgemm_amd_group* g, int* tile_class, int* slices, size_t* units) {
  return guarded([&] {
    if (g == nullptr) throw mtk::shgemm::error("group is null");
    mtk::shgemm::group_plan(g->g, tile_class, slices, units);
  });
}

int shgemm_amd_group_destroy(shgemm_amd_group* g) {
  return guarded([&] {
    if (g == nullptr) return;
    mtk::shgemm::destroy_group(g->g);
    delete g;
  });
}

int shgemm_amd_gaussian_f16(uint16_t* out, size_t count, uint64_t seed, float scale, void* stream) {
  return guarded([&] {
    if (count == 0) return;
    if (out == nullptr) throw mtk::shgemm::error("out is null");
    if (count > ((size_t)1 << 40)) throw mtk::shgemm::error("count too large");
    const hipError_t e =
        shgemm_amd::launch_gaussian_f16(out, (int64_t)count, seed, scale, static_cast<hipStream_t>(stream));
    if (e != hipSuccess) throw shgemm_amd::hip_failure(std::string("gaussian_f16: ") + hipGetErrorString(e));
  });
}

int shgemm_amd_philox_words(uint32_t* out, size_t blocks, uint64_t seed, void* stream) {
  return guarded([&] {
    if (blocks == 0) return;
    if (out == nullptr) throw mtk::shgemm::error("out is null");
    if (reinterpret_cast<uintptr_t>(out) % 16 != 0) throw mtk::shgemm::error("out must be 16-byte aligned");
    if (blocks > ((size_t)1 << 38)) throw mtk::shgemm::error("blocks too large");
    const hipError_t e = shgemm_amd::launch_philox_words(out, (int64_t)blocks, seed, static_cast<hipStream_t>(stream));
    if (e != hipSuccess) throw shgemm_amd::hip_failure(std::string("philox_words: ") + hipGetErrorString(e));
  });
}

const char* shgemm_amd_last_error(void) { return g_last_error.c_str(); }

const char* shgemm_amd_version(void) { return "shgemm_amd 0.1.0 (gfx950)"; }

}  // extern "C"
