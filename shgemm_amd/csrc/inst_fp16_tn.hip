// SHGEMM kernel variants: compute type fp16, op_a/op_b = tn.
#include "register_shgemm.hpp"
namespace shgemm_amd {
void register_shgemm_fp16_tn(KernelTable& t) { put_shgemm_ops<OP_T, OP_N, MODE_FP16>(t); }
}  // namespace shgemm_amd
