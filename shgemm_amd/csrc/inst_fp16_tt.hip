// SHGEMM kernel variants: compute type fp16, op_a/op_b = tt.
#include "register_shgemm.hpp"
namespace shgemm_amd {
void register_shgemm_fp16_tt(KernelTable& t) { put_shgemm_ops<OP_T, OP_T, MODE_FP16>(t); }
}  // namespace shgemm_amd
