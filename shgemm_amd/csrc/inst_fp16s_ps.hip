// Pre-split prescaled fp16 kernels (A split once per call, A and B packed in
// tile order): C op_n and op_t, on the 256 x 256 tile of eight wavefronts.
#include "register_shgemm.hpp"
namespace shgemm_amd {
using CfgPS = ModeCfg<MODE_FP16S>::PS;
template <int OPC>
static void put_ps(KernelTable& t) {
  static_assert(CfgPS::lds_bytes<MODE_FP16S>() <= kMaxLdsBytes, "tile exceeds the LDS of a CU");
  KernelEntry& e = t.shgemm_ps[OPC];
  e.fn = reinterpret_cast<const void*>(&shgemm_kernel_ps<OPC, CfgPS>);
  e.lds = CfgPS::lds_bytes<MODE_FP16S>();
  e.bm = CfgPS::BM;
  e.bn = CfgPS::BN;
  e.bk = CfgPS::BK;
  e.mode = MODE_FP16S;
  e.threads = CfgPS::NT;
  e.needs_row_exp = true;
  e.wgs_per_cu = 1;
}
void register_shgemm_ps(KernelTable& t) {
  put_ps<OP_N>(t);
  put_ps<OP_T>(t);
}
}  // namespace shgemm_amd
