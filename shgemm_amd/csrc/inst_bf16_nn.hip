// SHGEMM kernel variants: compute type bf16, op_a/op_b = nn.
#include "register_shgemm.hpp"
namespace shgemm_amd {
void register_shgemm_bf16_nn(KernelTable& t) { put_shgemm_ops<OP_N, OP_N, MODE_BF16>(t); }
}  // namespace shgemm_amd
