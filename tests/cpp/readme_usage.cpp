// The usage example of the reference's README (README.md:15-56) as a program
// built against this library's drop-in header include/shgemm/shgemm.hpp and
// linked to libshgemm_amd.so: create, set_cuda_stream, shgemm (A float,
// B half) and hsgemm (A half, B float) with the compute type named on the
// command line, destroy.  Inputs come from a fixed linear congruential
// generator; the inputs and both results are written as raw little-endian
// arrays for the test to check against the oracle.
//
//   readme_usage <m> <n> <k> <fp16|tf32> <out_dir> [debug]
//
// Prints "levels <shgemm level> <hsgemm level>"; with "debug" it turns the
// handle's debug mode on (set_debug_mode), so the library prints its call line.
#include <shgemm/shgemm.hpp>

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

namespace {

// The reference header's kernel descriptor types are part of its interface.
static_assert(sizeof(mtk::shgemm::detail::kernel) >= sizeof(mtk::shgemm::detail::kernel_func_t),
              "detail::kernel holds a kernel_func_t");
static_assert(mtk::shgemm::detail::num_levels == 2, "two kernel levels");

#define CHECK_HIP(x)                                                              \
  do {                                                                            \
    const hipError_t e_ = (x);                                                    \
    if (e_ != hipSuccess) {                                                       \
      std::fprintf(stderr, "%s failed: %s\n", #x, hipGetErrorString(e_));        \
      std::exit(3);                                                               \
    }                                                                             \
  } while (0)

struct Lcg {
  unsigned long long s;
  float next() {  // uniform on [-1, 1)
    s = s * 6364136223846793005ull + 1442695040888963407ull;
    return (float)((s >> 40) & 0xFFFFFF) / 8388608.0f - 1.0f;
  }
};

template <class T>
void write_file(const std::string& path, const std::vector<T>& v) {
  FILE* f = std::fopen(path.c_str(), "wb");
  if (f == nullptr || std::fwrite(v.data(), sizeof(T), v.size(), f) != v.size()) {
    std::fprintf(stderr, "cannot write %s\n", path.c_str());
    std::exit(4);
  }
  std::fclose(f);
}

template <class T>
T* to_device(const std::vector<T>& v) {
  T* p = nullptr;
  CHECK_HIP(hipMalloc(&p, v.size() * sizeof(T)));
  CHECK_HIP(hipMemcpy(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
  return p;
}

}  // namespace

int main(int argc, char** argv) {
  if (argc < 6) {
    std::fprintf(stderr, "usage: %s m n k fp16|tf32 out_dir [debug]\n", argv[0]);
    return 2;
  }
  const std::size_t m = std::strtoull(argv[1], nullptr, 10);
  const std::size_t n = std::strtoull(argv[2], nullptr, 10);
  const std::size_t k = std::strtoull(argv[3], nullptr, 10);
  const std::string dir = argv[5];
  const bool debug = argc > 6 && std::strcmp(argv[6], "debug") == 0;
  const std::size_t lda = m, ldb = k, ldc = m;

  Lcg g{12345};
  std::vector<float> a_fp32(m * k), b_fp32(k * n);
  std::vector<half> a_fp16(m * k), b_fp16(k * n);
  for (auto& x : a_fp32) x = g.next();
  for (auto& x : b_fp16) x = __float2half(g.next());
  for (auto& x : a_fp16) x = __float2half(g.next());
  for (auto& x : b_fp32) x = g.next();
  float* a_fp32_ptr = to_device(a_fp32);
  half* b_fp16_ptr = to_device(b_fp16);
  half* a_fp16_ptr = to_device(a_fp16);
  float* b_fp32_ptr = to_device(b_fp32);
  float *c_fp32_ptr = nullptr, *c2_fp32_ptr = nullptr;
  CHECK_HIP(hipMalloc(&c_fp32_ptr, m * n * sizeof(float)));
  CHECK_HIP(hipMalloc(&c2_fp32_ptr, m * n * sizeof(float)));
  hipStream_t cuda_stream;
  CHECK_HIP(hipStreamCreate(&cuda_stream));
  const float alpha_fp32 = 1.0f, beta_fp32 = 0.0f;

  int level_sh = -1, level_hs = -1;
  try {
    mtk::shgemm::shgemmHandle_t shgemm_handle;
    mtk::shgemm::create(shgemm_handle);

    // Optional
    mtk::shgemm::set_cuda_stream(shgemm_handle, cuda_stream);
    if (debug) mtk::shgemm::set_debug_mode(shgemm_handle, 1);

    const auto compute_type = std::strcmp(argv[4], "tf32") == 0 ? mtk::shgemm::tf32 : mtk::shgemm::fp16;

    // SHGEMM (A=float, B=half)
    level_sh = mtk::shgemm::shgemm(
        shgemm_handle,
        mtk::shgemm::op_n, mtk::shgemm::op_n,
        m, n, k,
        &alpha_fp32,
        a_fp32_ptr, lda,
        b_fp16_ptr, ldb,
        &beta_fp32,
        c_fp32_ptr, ldc,
        compute_type);

    // HSGEMM (A=half, B=float) is also available
    level_hs = mtk::shgemm::hsgemm(
        shgemm_handle,
        mtk::shgemm::op_n, mtk::shgemm::op_n,
        m, n, k,
        &alpha_fp32,
        a_fp16_ptr, lda,
        b_fp32_ptr, ldb,
        &beta_fp32,
        c2_fp32_ptr, ldc,
        compute_type);

    CHECK_HIP(hipStreamSynchronize(cuda_stream));
    mtk::shgemm::destroy(shgemm_handle);
  } catch (const mtk::shgemm::error& e) {
    std::fprintf(stderr, "mtk::shgemm::error: %s\n", e.what());
    return 5;
  }
  std::vector<float> c(m * n), c2(m * n);
  CHECK_HIP(hipMemcpy(c.data(), c_fp32_ptr, m * n * sizeof(float), hipMemcpyDeviceToHost));
  CHECK_HIP(hipMemcpy(c2.data(), c2_fp32_ptr, m * n * sizeof(float), hipMemcpyDeviceToHost));
  write_file(dir + "/a_fp32.bin", a_fp32);
  write_file(dir + "/b_fp16.bin", b_fp16);
  write_file(dir + "/a_fp16.bin", a_fp16);
  write_file(dir + "/b_fp32.bin", b_fp32);
  write_file(dir + "/c_shgemm.bin", c);
  write_file(dir + "/c_hsgemm.bin", c2);
  std::printf("levels %d %d\n", level_sh, level_hs);
  for (void* p : {(void*)a_fp32_ptr, (void*)b_fp16_ptr, (void*)a_fp16_ptr, (void*)b_fp32_ptr, (void*)c_fp32_ptr,
                  (void*)c2_fp32_ptr})
    CHECK_HIP(hipFree(p));
  CHECK_HIP(hipStreamDestroy(cuda_stream));
  return 0;
}
