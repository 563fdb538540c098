// Measured plan table: for shape points (log2 m, log2 n, log2 k) per split
// mode and operand layout, the kernel choice the autotuner (tune.py) found
// fastest.  One text row per point:
//   kind mode op_a op_b op_c log2m log2n log2k choice slices [time_us]
// kind   single  -- one shgemm call: choice = plan variant 0..11
//        grouped -- one problem inside a grouped launch: choice = tile class 0..3
// mode   fp16 (two accumulators), fp16s (prescaled), bf16 (the tf32 type)
// Lines starting with '#' are comments.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

namespace shgemm_amd {

enum : int { PLAN_SINGLE = 0, PLAN_GROUPED = 1 };

struct PlanRow {
  int kind, mode, oa, ob, oc;
  int lm, ln, lk;
  int choice, slices;
  float time_us;
};

// Parses `path`; returns false and sets `err` (file:line: reason) on the first
// bad line.  Rows are appended to `out` only when the whole file is valid.
bool parse_plan_file(const std::string& path, std::vector<PlanRow>& out, std::string& err);

// Nearest row for the arguments: rows of the same kind and mode, with the same
// operand layout if any exist, else with layout N/N/N; nearest in L1 distance
// over (log2 m, log2 n, log2 k) among rows within 1.5 in every coordinate.
// nullptr when there is none.
const PlanRow* find_plan(const std::vector<PlanRow>& rows, int kind, int mode, int oa, int ob, int oc, int64_t m,
                         int64_t n, int64_t k);

// The plan table shipped beside the library (plans_gfx950.tsv in the directory
// of libshgemm_amd.so), or "" when the directory cannot be found.
std::string default_plan_path();

}  // namespace shgemm_amd
