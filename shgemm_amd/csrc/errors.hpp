// Internal exception type for HIP runtime failures, so the C ABI can tell
// them apart from argument errors.
#pragma once
#include <string>
#include "../../include/shgemm/shgemm.hpp"

namespace shgemm_amd {
class hip_failure : public mtk::shgemm::error {
 public:
  explicit hip_failure(const std::string& what) : mtk::shgemm::error(what) {}
};
}  // namespace shgemm_amd
