// SHGEMM kernel variants: prescaled single-accumulator fp16 split, op_a/op_b = tt.
#include "register_shgemm.hpp"
namespace shgemm_amd {
void register_shgemm_fp16s_tt(KernelTable& t) { put_shgemm_ops<OP_T, OP_T, MODE_FP16S>(t); }
}  // namespace shgemm_amd
