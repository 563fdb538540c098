// SHGEMM kernel variants: compute type fp16, op_a/op_b = nt.
#include "register_shgemm.hpp"
namespace shgemm_amd {
void register_shgemm_fp16_nt(KernelTable& t) { put_shgemm_ops<OP_N, OP_T, MODE_FP16>(t); }
}  // namespace shgemm_amd
