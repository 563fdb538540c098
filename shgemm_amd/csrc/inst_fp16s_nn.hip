// SHGEMM kernel variants: prescaled single-accumulator fp16 split, op_a/op_b = nn.
#include "register_shgemm.hpp"
namespace shgemm_amd {
void register_shgemm_fp16s_nn(KernelTable& t) { put_shgemm_ops<OP_N, OP_N, MODE_FP16S>(t); }
}  // namespace shgemm_amd
