// Candidate tile configurations of the fp16 op_n/op_n kernel (C op_n), for
// the tile sweep (MTK_SHGEMM_TUNE_P1=<index>).
#include "register_shgemm.hpp"
namespace shgemm_amd {
template <class CFG, int MODE = MODE_FP16>
static void put_tune(KernelTable& t) {
  KernelEntry& e = t.tune[t.ntune++];
  e.fn = reinterpret_cast<const void*>(&shgemm_kernel<OP_N, OP_N, OP_N, MODE, CFG, false>);
  e.lds = CFG::template lds_bytes<MODE>();
  e.bm = CFG::BM;
  e.bn = CFG::BN;
  e.bk = CFG::BK;
  e.mode = MODE;
}
void register_tune(KernelTable& t) {
  t.ntune = 0;
  put_tune<TileCfg<128, 128, 32, 64, 64, 1, 2>, MODE_FP16>(t);
  put_tune<TileCfg<128, 128, 32, 64, 64, 1, 2>, MODE_FP16S>(t);
  put_tune<TileCfg<128, 256, 32, 64, 128, 1, 2>, MODE_FP16S>(t);
  put_tune<TileCfg<128, 256, 32, 64, 128, 1, 1>, MODE_FP16S>(t);
  put_tune<TileCfg<256, 128, 32, 128, 64, 1, 2>, MODE_FP16S>(t);
  put_tune<TileCfg<128, 128, 32, 64, 64, 2, 2>, MODE_FP16S>(t);
  put_tune<TileCfg<256, 256, 32, 128, 128, 1, 1>, MODE_FP16S>(t);
  put_tune<TileCfg<128, 256, 64, 64, 128, 1, 1>, MODE_FP16S>(t);
}
}  // namespace shgemm_amd
