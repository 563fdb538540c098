// SHGEMM kernel variants: prescaled single-accumulator fp16 split, op_a/op_b = tn.
#include "register_shgemm.hpp"
namespace shgemm_amd {
void register_shgemm_fp16s_tn(KernelTable& t) { put_shgemm_ops<OP_T, OP_N, MODE_FP16S>(t); }
}  // namespace shgemm_amd
