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
  e.threads = CFG::NT;
}
void register_tune(KernelTable& t) {
  t.ntune = 0;
  put_tune<TileCfg<128, 128, 32, 64, 64, 1, 2>, MODE_FP16>(t);     // 0
  put_tune<TileCfg<128, 256, 32, 64, 128, 1, 2>, MODE_FP16S>(t);   // 1 (default P1)
  put_tune<TileCfg<128, 128, 32, 64, 64, 1, 3>, MODE_FP16S>(t);    // 2
  put_tune<TileCfg<256, 256, 32, 64, 128, 1, 1>, MODE_FP16S>(t);   // 3: 8 waves
  put_tune<TileCfg<128, 256, 32, 64, 64, 1, 2>, MODE_FP16S>(t);    // 4: 8 waves
  put_tune<TileCfg<256, 128, 32, 64, 64, 1, 2>, MODE_FP16S>(t);    // 5: 8 waves
  put_tune<TileCfg<256, 128, 32, 64, 64, 1, 1>, MODE_FP16S>(t);    // 6: 8 waves
  put_tune<TileCfg<128, 256, 64, 64, 64, 1, 1>, MODE_FP16S>(t);    // 7: 8 waves, BK 64
  put_tune<TileCfg<256, 256, 32, 64, 64, 1, 1>, MODE_FP16S>(t);    // 8: 16 waves
  put_tune<TileCfg<256, 256, 32, 64, 128, 1, 1>, MODE_FP16S>(t);   // 9: 8 waves, w64x128
  put_tune<TileCfg<128, 512, 32, 64, 128, 1, 1>, MODE_FP16S>(t);   // 10: 8 waves
}
}  // namespace shgemm_amd
