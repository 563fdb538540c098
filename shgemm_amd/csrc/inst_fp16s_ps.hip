// Pre-split prescaled fp16 kernels (A split once per call and packed in tile
// order; B packed the same way, or for op_b = op_n read as it lies): C op_n
// and op_t, on the 256 x 256 tile of eight wavefronts (one workgroup per CU)
// and on 128 x 128 tiles of four (two per CU).
#include "register_shgemm.hpp"
namespace shgemm_amd {
template <class CFG, int OPC, bool DIRECT_B, int MODE>
static void put_ps(KernelEntry& e, int wgs_per_cu) {
  static_assert(PsRingLayout<CFG>::lds_bytes() <= kMaxLdsBytes, "ring exceeds the LDS of a CU");
  e.fn = reinterpret_cast<const void*>(&shgemm_kernel_ps<OPC, CFG, DIRECT_B, MODE>);
  e.lds = PsRingLayout<CFG>::lds_bytes();
  e.bm = CFG::BM;
  e.bn = CFG::BN;
  e.bk = CFG::BK;
  e.mode = MODE;
  e.threads = CFG::NT;
  e.needs_row_exp = MODE == MODE_FP16S;
  e.wgs_per_cu = wgs_per_cu;
}
template <int OPC, bool DIRECT_B>
static void put_ps(KernelTable& t) {
  put_ps<ModeCfg<MODE_FP16S>::PS, OPC, DIRECT_B, MODE_FP16S>(t.shgemm_ps[DIRECT_B][OPC], 1);
  put_ps<ModeCfg<MODE_FP16S>::PS128, OPC, DIRECT_B, MODE_FP16S>(t.shgemm_ps128[DIRECT_B][OPC], 2);
  // prescaling off: the 128 x 128 form only (the 64 x 128 wavefronts would
  // need 256 accumulator registers)
  put_ps<ModeCfg<MODE_FP16S>::PS128, OPC, DIRECT_B, MODE_FP16>(t.shgemm_ps128_fp16[DIRECT_B][OPC], 2);
}
// tf32 (BF16 parts): the same tiles.
template <class CFG, int OPC>
static void put_ps_bf16(KernelEntry& e, int wgs_per_cu) {
  static_assert(PsBf16Layout<CFG>::lds_bytes() <= kMaxLdsBytes, "ring exceeds the LDS of a CU");
  e.fn = reinterpret_cast<const void*>(&shgemm_kernel_ps_bf16<OPC, CFG>);
  e.lds = PsBf16Layout<CFG>::lds_bytes();
  e.bm = CFG::BM;
  e.bn = CFG::BN;
  e.bk = CFG::BK;
  e.mode = MODE_BF16;
  e.threads = CFG::NT;
  e.needs_row_exp = false;
  e.wgs_per_cu = wgs_per_cu;
}
template <int OPC>
static void put_ps_bf16(KernelTable& t) {
  put_ps_bf16<ModeCfg<MODE_BF16>::PS, OPC>(t.shgemm_ps_bf16[OPC], 1);
  put_ps_bf16<ModeCfg<MODE_BF16>::PS128, OPC>(t.shgemm_ps128_bf16[OPC], 2);
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
