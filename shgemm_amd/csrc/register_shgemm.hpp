// Instantiates the SHGEMM kernel variants of one (op_a, op_b, mode) and
// records them in the kernel table.  Included by one .hip unit per
// combination so the variants compile in parallel.
#pragma once
#include "registry.hpp"
#include "grouped_kernel.hpp"

namespace shgemm_amd {

template <int OPA, int OPB, int OPC, int MODE, class CFG, bool SPLITK>
inline void fill_shgemm(KernelEntry& e) {
  static_assert(CFG::template lds_bytes<MODE>() <= kMaxLdsBytes, "tile exceeds the LDS of a CU");
  e.fn = reinterpret_cast<const void*>(&shgemm_kernel<OPA, OPB, OPC, MODE, CFG, SPLITK>);
  e.lds = CFG::template lds_bytes<MODE>();
  e.bm = CFG::BM;
  e.bn = CFG::BN;
  e.bk = CFG::BK;
  e.mode = MODE;
  e.threads = CFG::NT;
  e.wgs_per_cu = (4 * CFG::OCC) / CFG::WAVES > 0 ? (4 * CFG::OCC) / CFG::WAVES : 1;
  e.needs_row_exp = MODE == MODE_FP16S && !CFG::template running<MODE>();
}

template <int OPA, int OPB, int OPC, int MODE, class CFG, bool SPLITK>
inline void put_shgemm(KernelTable& t, int variant) {
  fill_shgemm<OPA, OPB, OPC, MODE, CFG, SPLITK>(t.shgemm[OPA][OPB][OPC][MODE][variant]);
}

template <int OPA, int OPB, int OPC, int MODE, class CFG>
inline void put_grouped(KernelTable& t, int cls) {
  static_assert(CFG::template lds_bytes<MODE>() + 256 <= kMaxLdsBytes, "tile exceeds the LDS of a CU");
  KernelEntry& e = t.grouped[OPA][OPB][OPC][MODE][cls];
  e.fn = reinterpret_cast<const void*>(&shgemm_grouped_kernel<OPA, OPB, OPC, MODE, CFG>);
  e.lds = CFG::template lds_bytes<MODE>();
  e.bm = CFG::BM;
  e.bn = CFG::BN;
  e.bk = CFG::BK;
  e.mode = MODE;
  e.threads = CFG::NT;
  e.needs_row_exp = MODE == MODE_FP16S && !CFG::template running<MODE>();
}

template <int OPC, class CFG>
inline void put_ring(KernelTable& t) {
  static_assert(RingLayout<CFG>::lds_bytes() <= kMaxLdsBytes, "tile exceeds the LDS of a CU");
  KernelEntry& e = t.shgemm[OP_N][OP_N][OPC][MODE_FP16S][V_P2R];
  e.fn = reinterpret_cast<const void*>(&shgemm_kernel_ring<OPC, CFG>);
  e.lds = RingLayout<CFG>::lds_bytes();
  e.bm = CFG::BM;
  e.bn = CFG::BN;
  e.bk = CFG::BK;
  e.mode = MODE_FP16S;
  e.threads = CFG::NT;
  e.wgs_per_cu = 1;
  e.needs_row_exp = true;
}

template <int OPA, int OPB, int OPC, int MODE>
inline void put_shgemm_opc(KernelTable& t) {
  using G = GroupCfg<MODE>;
  put_grouped<OPA, OPB, OPC, MODE, typename G::C0>(t, 0);
  put_grouped<OPA, OPB, OPC, MODE, typename G::C1>(t, 1);
  put_grouped<OPA, OPB, OPC, MODE, typename G::C2>(t, 2);
  if constexpr (G::kHasC3) put_grouped<OPA, OPB, OPC, MODE, typename G::C3>(t, 3);
  using C = ModeCfg<MODE>;
  put_shgemm<OPA, OPB, OPC, MODE, typename C::P1, false>(t, V_P1);
  put_shgemm<OPA, OPB, OPC, MODE, typename C::P0, false>(t, V_P0);
  put_shgemm<OPA, OPB, OPC, MODE, typename C::SK, true>(t, V_SK);
  put_shgemm<OPA, OPB, OPC, MODE, typename C::P1, true>(t, V_P1SK);
  put_shgemm<OPA, OPB, OPC, MODE, typename C::P0SK, true>(t, V_P0SK);
  if constexpr (C::kHasP2) put_shgemm<OPA, OPB, OPC, MODE, typename C::P2, false>(t, V_P2);
  if constexpr (MODE == MODE_FP16S) {
    fill_shgemm<OPA, OPB, OPC, MODE, typename C::P0W, false>(t.shgemm_p0w[OPA][OPB][OPC]);
    fill_shgemm<OPA, OPB, OPC, MODE, typename C::SKD, true>(t.shgemm_skd[OPA][OPB][OPC]);
    fill_shgemm<OPA, OPB, OPC, MODE, typename C::P1W, true>(t.shgemm_wsk[OPA][OPB][OPC]);
  }
  if constexpr (MODE == MODE_FP16S && OPA == OP_N && OPB == OP_N) put_ring<OPC, typename C::P2>(t);
}

template <int OPA, int OPB, int MODE>
inline void put_shgemm_ops(KernelTable& t) {
  put_shgemm_opc<OPA, OPB, OP_N, MODE>(t);
  put_shgemm_opc<OPA, OPB, OP_T, MODE>(t);
}

}  // namespace shgemm_amd
