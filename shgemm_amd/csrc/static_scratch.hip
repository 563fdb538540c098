// Device scratch of the allocation-free entry point (shgemm_amd_shgemm_no_alloc):
// one file-scope device array, `scratch`, of which every GPU the library runs
// on holds its own copy until the process ends.  On the host the array's
// name stands for a host-side stand-in, not for device memory, so host code
// never hands it to a kernel or a copy: the device address is looked up once
// per device with hipGetSymbolAddress, with that device current, and only
// that address is used.
//
// Layout (kScratchBytes = 96 MiB): split-K partials (kWorkspaceBytes, 64 MiB,
// the size of a handle's workspace, so both entry points choose the same
// plans), the row-exponent pass's k-block maxima (kRowPartialBytes, 16 MiB),
// then row maxima and row exponents for kRowCap rows (8 MiB each).  Every
// kernel writes each scratch element it reads later in the same call (split-K
// slices write every element of their tiles, the row pass writes each row's
// maximum per k-block), so nothing has to start at zero; calls that overlap
// on one GPU share the array and race.
#include <hip/hip_runtime.h>

#include <mutex>

#include "runtime_internal.hpp"

namespace {
constexpr size_t kRowCap = size_t(1) << 21;
constexpr size_t kScratchBytes =
    shgemm_amd::kWorkspaceBytes + shgemm_amd::kRowPartialBytes + kRowCap * (sizeof(uint32_t) + sizeof(int));
}  // namespace

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace shgemm_amd {
namespace {

constexpr int kMaxDevices = 64;
std::mutex g_mu;
void* g_addr[kMaxDevices] = {};

// Device address of `scratch` on the current device, looked up on its first use.
void* scratch_address() {
  int dev = -1;
  check_hip(hipGetDevice(&dev), "hipGetDevice");
  if (dev < 0 || dev >= kMaxDevices) throw hip_failure("device index beyond the static scratch table");
  std::lock_guard<std::mutex> lock(g_mu);
  if (g_addr[dev] == nullptr) {
    void* p = nullptr;
    check_hip(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(static scratch)");
    if (p == nullptr) throw hip_failure("hipGetSymbolAddress(static scratch) returned null");
    g_addr[dev] = p;
  }
  return g_addr[dev];
}

}  // namespace

void init_static_scratch() { (void)scratch_address(); }

size_t static_scratch_bytes() { return kScratchBytes; }

CallScratch static_scratch() {
  char* base = static_cast<char*>(scratch_address());
  CallScratch s;
  s.ws = reinterpret_cast<float*>(base);
  s.row_part = reinterpret_cast<uint32_t*>(base + kWorkspaceBytes);
  s.row_max = reinterpret_cast<uint32_t*>(base + kWorkspaceBytes + kRowPartialBytes);
  s.row_exp = reinterpret_cast<int*>(base + kWorkspaceBytes + kRowPartialBytes + kRowCap * sizeof(uint32_t));
  s.row_cap = kRowCap;
  return s;
}

}  // namespace shgemm_amd
