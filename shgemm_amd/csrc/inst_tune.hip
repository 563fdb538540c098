// Candidate tile configurations of the op_n/op_n (C op_n) kernel for the tile
// sweep (MTK_SHGEMM_TUNE_P1=<index> replaces the unsplit large tile).
#include "register_shgemm.hpp"
namespace shgemm_amd {
template <class CFG, int MODE = MODE_FP16>
static void put_tune(KernelTable& t) {
  static_assert(CFG::template lds_bytes<MODE>() <= kMaxLdsBytes, "tile exceeds the LDS of a CU");
  if (t.ntune >= KernelTable::kMaxTune) return;  // list full: candidate not registered
  KernelEntry& e = t.tune[t.ntune++];
  e.fn = reinterpret_cast<const void*>(&shgemm_kernel<OP_N, OP_N, OP_N, MODE, CFG, false>);
  e.lds = CFG::template lds_bytes<MODE>();
  e.bm = CFG::BM;
  e.bn = CFG::BN;
  e.bk = CFG::BK;
  e.mode = MODE;
  e.threads = CFG::NT;
  e.needs_row_exp = MODE == MODE_FP16S && !CFG::template running<MODE>();
}
// Deep-prefetch kernel (B direct to a 3-stage LDS ring, A in two register sets).
template <class CFG>
static void put_tune_bd(KernelTable& t) {
  static_assert(BdLayout<CFG>::lds_bytes() <= kMaxLdsBytes, "tile exceeds the LDS of a CU");
  if (t.ntune >= KernelTable::kMaxTune) return;  // list full: candidate not registered
  KernelEntry& e = t.tune[t.ntune++];
  e.fn = reinterpret_cast<const void*>(&shgemm_kernel_bd<OP_N, OP_N, CFG>);
  e.lds = BdLayout<CFG>::lds_bytes();
  e.bm = CFG::BM;
  e.bn = CFG::BN;
  e.bk = CFG::BK;
  e.mode = MODE_FP16S;
  e.threads = CFG::NT;
  e.needs_row_exp = true;
}
// Ring-pipeline kernel (raw A and B straight to LDS rings; interior tiles).
template <class CFG>
static void put_tune_ring(KernelTable& t) {
  static_assert(RingLayout<CFG>::lds_bytes() <= kMaxLdsBytes, "tile exceeds the LDS of a CU");
  if (t.ntune >= KernelTable::kMaxTune) return;  // list full: candidate not registered
  KernelEntry& e = t.tune[t.ntune++];
  e.fn = reinterpret_cast<const void*>(&shgemm_kernel_ring<OP_N, CFG>);
  e.lds = RingLayout<CFG>::lds_bytes();
  e.bm = CFG::BM;
  e.bn = CFG::BN;
  e.bk = CFG::BK;
  e.mode = MODE_FP16S;
  e.threads = CFG::NT;
  e.needs_row_exp = true;
}
void register_tune(KernelTable& t) {
  t.ntune = 0;
  put_tune<TileCfg<128, 128, 32, 64, 64, 1, 2>, MODE_FP16>(t);      // 0: two accumulators, 4 waves
  put_tune<TileCfg<128, 256, 32, 64, 64, 1, 2>, MODE_FP16S>(t);     // 1: the prescaled P1 (8 waves of 64x64)
  put_tune<TileCfg<256, 256, 32, 128, 128, 1, 1>, MODE_FP16S>(t);   // 2: 4 waves of 128x128, 1 per SIMD
  put_tune<TileCfg<128, 128, 32, 64, 64, 1, 2>, MODE_FP16S>(t);     // 3: 4 waves of 64x64
  put_tune<TileCfg<128, 256, 32, 64, 128, 1, 2>, MODE_FP16S>(t);    // 4: 4 waves of 64x128
  put_tune<TileCfg<256, 128, 32, 128, 64, 1, 2>, MODE_FP16S>(t);    // 5: 4 waves of 128x64
  put_tune<TileCfg<256, 256, 32, 64, 128, 1, 2>, MODE_FP16S>(t);    // 6: 8 waves of 64x128
  put_tune<TileCfg<256, 256, 32, 128, 64, 1, 2>, MODE_FP16S>(t);    // 7: 8 waves of 128x64
  put_tune<TileCfg<256, 128, 32, 64, 64, 1, 2>, MODE_FP16S>(t);     // 8: 8 waves of 64x64
  put_tune<TileCfg<128, 256, 64, 64, 64, 1, 2>, MODE_FP16S>(t);     // 9: 8 waves of 64x64, BK 64
  put_tune<TileCfg<256, 256, 16, 128, 128, 1, 1>, MODE_FP16S>(t);   // 10: 4 waves of 128x128, BK 16
  put_tune<TileCfg<128, 256, 32, 128, 64, 1, 2>, MODE_FP16S>(t);    // 11: 4 waves of 128x64 in a row
  put_tune<TileCfg<256, 128, 32, 64, 128, 1, 2>, MODE_FP16S>(t);    // 12: 4 waves of 64x128 in a column
  put_tune<TileCfg<192, 256, 32, 96, 128, 1, 1>, MODE_FP16S>(t);    // 13: 4 waves of 96x128
  put_tune<TileCfg<256, 192, 32, 128, 96, 1, 1>, MODE_FP16S>(t);    // 14: 4 waves of 128x96
  put_tune_bd<TileCfg<128, 256, 32, 64, 64, 1, 2>>(t);              // 15: deep prefetch, 8 waves of 64x64
  put_tune_ring<TileCfg<256, 256, 32, 128, 128, 1, 1>>(t);          // 16: ring pipeline on the P2 tile
  put_tune<TileCfg<256, 256, 32, 128, 128, 2, 1>, MODE_FP16S>(t);   // 17: P2, two k-tiles of register prefetch
  put_tune<TileCfg<256, 256, 32, 128, 128, 2, 1, true>, MODE_FP16S>(t);  // 18: 17 with the skewed barrier (the library's P2)
}
}  // namespace shgemm_amd
