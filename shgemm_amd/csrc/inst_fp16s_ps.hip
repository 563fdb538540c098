// Pre-split prescaled fp16 kernels (A split once per call and packed in tile
// order; B packed the same way, or for op_b = op_n read as it lies): C op_n
// and op_t, on the 256 x 256 tile of eight wavefronts.
#include "register_shgemm.hpp"
namespace shgemm_amd {
using CfgPS = ModeCfg<MODE_FP16S>::PS;
template <int OPC, bool DIRECT_B>
static void put_ps(KernelTable& t) {
  static_assert(PsRingLayout<CfgPS>::lds_bytes() <= kMaxLdsBytes, "ring exceeds the LDS of a CU");
  KernelEntry& e = t.shgemm_ps[DIRECT_B][OPC];
  e.fn = reinterpret_cast<const void*>(&shgemm_kernel_ps<OPC, CfgPS, DIRECT_B>);
  e.lds = PsRingLayout<CfgPS>::lds_bytes();
  e.bm = CfgPS::BM;
  e.bn = CfgPS::BN;
  e.bk = CfgPS::BK;
  e.mode = MODE_FP16S;
  e.threads = CfgPS::NT;
  e.needs_row_exp = true;
  e.wgs_per_cu = 1;
}
// tf32 (BF16 parts): the same 256 x 256 tile of eight 64 x 128 wavefronts,
// one workgroup per CU.
using CfgPSB = ModeCfg<MODE_BF16>::PS;
template <int OPC>
static void put_ps_bf16(KernelTable& t) {
  static_assert(PsBf16Layout<CfgPSB>::lds_bytes() <= kMaxLdsBytes, "ring exceeds the LDS of a CU");
  KernelEntry& e = t.shgemm_ps_bf16[OPC];
  e.fn = reinterpret_cast<const void*>(&shgemm_kernel_ps_bf16<OPC, CfgPSB>);
  e.lds = PsBf16Layout<CfgPSB>::lds_bytes();
  e.bm = CfgPSB::BM;
  e.bn = CfgPSB::BN;
  e.bk = CfgPSB::BK;
  e.mode = MODE_BF16;
  e.threads = CfgPSB::NT;
  e.needs_row_exp = false;
  e.wgs_per_cu = 1;
}
void register_shgemm_ps(KernelTable& t) {
  put_ps_bf16<OP_N>(t);
  put_ps_bf16<OP_T>(t);
  put_ps<OP_N, false>(t);
  put_ps<OP_T, false>(t);
  put_ps<OP_N, true>(t);
  put_ps<OP_T, true>(t);
}
}  // namespace shgemm_amd
