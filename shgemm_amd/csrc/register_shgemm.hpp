// Instantiates the SHGEMM kernel variants of one (op_a, op_b, mode) and
// records them in the kernel table.  Included by one .hip unit per
// combination so the variants compile in parallel.
#pragma once
#include "registry.hpp"
#include "shgemm_kernel.hpp"

namespace shgemm_amd {

template <int OPA, int OPB, int OPC, int MODE, class CFG, bool SPLITK, bool SCALED>
inline void put_shgemm(KernelTable& t, int level) {
  KernelEntry& e = t.shgemm[OPA][OPB][OPC][MODE][level][SPLITK ? 1 : 0][SCALED ? 1 : 0];
  e.fn = reinterpret_cast<const void*>(&shgemm_kernel<OPA, OPB, OPC, MODE, CFG, SPLITK, SCALED>);
  e.lds = CFG::template lds_bytes<MODE>();
  e.bm = CFG::BM;
  e.bn = CFG::BN;
  e.bk = CFG::BK;
}

template <int OPA, int OPB, int OPC, int MODE>
inline void put_shgemm_opc(KernelTable& t) {
  put_shgemm<OPA, OPB, OPC, MODE, CfgP1, false, false>(t, LEVEL_P1);
  put_shgemm<OPA, OPB, OPC, MODE, CfgP0, false, false>(t, LEVEL_P0);
  put_shgemm<OPA, OPB, OPC, MODE, CfgP1, true, false>(t, LEVEL_P1);
  if constexpr (MODE == MODE_FP16) {
    put_shgemm<OPA, OPB, OPC, MODE, CfgP1, false, true>(t, LEVEL_P1);
    put_shgemm<OPA, OPB, OPC, MODE, CfgP0, false, true>(t, LEVEL_P0);
    put_shgemm<OPA, OPB, OPC, MODE, CfgP1, true, true>(t, LEVEL_P1);
  }
}

template <int OPA, int OPB, int MODE>
inline void put_shgemm_ops(KernelTable& t) {
  put_shgemm_opc<OPA, OPB, OP_N, MODE>(t);
  put_shgemm_opc<OPA, OPB, OP_T, MODE>(t);
}

}  // namespace shgemm_amd
