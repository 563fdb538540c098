// Deep-prefetch prescaled fp16 kernels (op_b = n): every op_a x op_c.
#include "register_shgemm.hpp"
namespace shgemm_amd {
using CfgBD = TileCfg<128, 256, 32, 64, 64, 1, 2>;
template <int OPA, int OPC>
static void put_bd(KernelTable& t) {
  KernelEntry& e = t.shgemm_bd[OPA][OPC];
  e.fn = reinterpret_cast<const void*>(&shgemm_kernel_bd<OPA, OPC, CfgBD>);
  e.lds = BdLayout<CfgBD>::lds_bytes();
  e.bm = CfgBD::BM;
  e.bn = CfgBD::BN;
  e.bk = CfgBD::BK;
  e.mode = MODE_FP16S;
  e.threads = CfgBD::NT;
  e.needs_row_exp = true;
}
void register_shgemm_bd(KernelTable& t) {
  put_bd<OP_N, OP_N>(t);
  put_bd<OP_N, OP_T>(t);
  put_bd<OP_T, OP_N>(t);
  put_bd<OP_T, OP_T>(t);
}
}  // namespace shgemm_amd
