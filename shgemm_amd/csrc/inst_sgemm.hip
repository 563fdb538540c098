// Plain FP32 SGEMM kernel variants (all op_a / op_b / op_c).
#include "registry.hpp"
#include "sgemm_kernel.hpp"
namespace shgemm_amd {
template <int OPA, int OPB, int OPC>
static void put_sgemm(KernelTable& t) {
  KernelEntry& e = t.sgemm[OPA][OPB][OPC];
  e.fn = reinterpret_cast<const void*>(&sgemm_kernel<OPA, OPB, OPC, SgemmCfgP1>);
  e.lds = SgemmCfgP1::lds_bytes();
  e.bm = SgemmCfgP1::BM;
  e.bn = SgemmCfgP1::BN;
  e.bk = SgemmCfgP1::BK;
}
void register_sgemm(KernelTable& t) {
  put_sgemm<OP_N, OP_N, OP_N>(t); put_sgemm<OP_N, OP_N, OP_T>(t);
  put_sgemm<OP_N, OP_T, OP_N>(t); put_sgemm<OP_N, OP_T, OP_T>(t);
  put_sgemm<OP_T, OP_N, OP_N>(t); put_sgemm<OP_T, OP_N, OP_T>(t);
  put_sgemm<OP_T, OP_T, OP_N>(t); put_sgemm<OP_T, OP_T, OP_T>(t);
}
}  // namespace shgemm_amd
