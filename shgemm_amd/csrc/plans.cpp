// Plan table parsing and lookup (format in plans.hpp).
#include "plans.hpp"

#include <dlfcn.h>
#include <cmath>
#include <cstdlib>
#include <fstream>
#include <sstream>

#include "gemm_common.hpp"

namespace shgemm_amd {
namespace {

int parse_mode(const std::string& s) {
  if (s == "fp16") return MODE_FP16;
  if (s == "fp16s") return MODE_FP16S;
  if (s == "bf16") return MODE_BF16;
  return -1;
}

int parse_op(const std::string& s) {
  if (s == "N") return OP_N;
  if (s == "T") return OP_T;
  return -1;
}

bool parse_int(const std::string& s, int lo, int hi, int& v) {
  char* end = nullptr;
  const long x = std::strtol(s.c_str(), &end, 10);
  if (end == s.c_str() || *end != '\0' || x < lo || x > hi) return false;
  v = (int)x;
  return true;
}

}  // namespace

bool parse_plan_file(const std::string& path, std::vector<PlanRow>& out, std::string& err) {
  std::ifstream f(path);
  if (!f) {
    err = path + ": cannot open";
    return false;
  }
  std::vector<PlanRow> rows;
  std::string line;
  int lineno = 0;
  while (std::getline(f, line)) {
    ++lineno;
    const size_t first = line.find_first_not_of(" \t\r");
    if (first == std::string::npos || line[first] == '#') continue;
    std::istringstream ss(line);
    std::vector<std::string> tok;
    std::string t;
    while (ss >> t) tok.push_back(t);
    const std::string where = path + ":" + std::to_string(lineno) + ": ";
    if (tok.size() != 10 && tok.size() != 11) {
      err = where + "expected 10 or 11 fields";
      return false;
    }
    PlanRow r{};
    if (tok[0] == "single") r.kind = PLAN_SINGLE;
    else if (tok[0] == "grouped") r.kind = PLAN_GROUPED;
    else {
      err = where + "kind must be single or grouped";
      return false;
    }
    r.mode = parse_mode(tok[1]);
    r.oa = parse_op(tok[2]);
    r.ob = parse_op(tok[3]);
    r.oc = parse_op(tok[4]);
    if (r.mode < 0 || r.oa < 0 || r.ob < 0 || r.oc < 0) {
      err = where + "bad mode or operation";
      return false;
    }
    const int max_choice = r.kind == PLAN_SINGLE ? 11 : 3;
    if (!parse_int(tok[5], 0, 31, r.lm) || !parse_int(tok[6], 0, 31, r.ln) || !parse_int(tok[7], 0, 31, r.lk) ||
        !parse_int(tok[8], 0, max_choice, r.choice) || !parse_int(tok[9], 1, 256, r.slices)) {
      err = where + "bad shape, choice or slices";
      return false;
    }
    r.time_us = tok.size() == 11 ? std::strtof(tok[10].c_str(), nullptr) : 0.0f;
    rows.push_back(r);
  }
  out.insert(out.end(), rows.begin(), rows.end());
  return true;
}

const PlanRow* find_plan(const std::vector<PlanRow>& rows, int kind, int mode, int oa, int ob, int oc, int64_t m,
                         int64_t n, int64_t k) {
  if (m <= 0 || n <= 0 || k <= 0) return nullptr;
  const double lm = std::log2((double)m), ln = std::log2((double)n), lk = std::log2((double)k);
  for (int pass = 0; pass < 2; ++pass) {
    const int want_a = pass == 0 ? oa : OP_N, want_b = pass == 0 ? ob : OP_N, want_c = pass == 0 ? oc : OP_N;
    const PlanRow* best = nullptr;
    double best_d = 1e30;
    bool any_layout = false;
    for (const PlanRow& r : rows) {
      if (r.kind != kind || r.mode != mode || r.oa != want_a || r.ob != want_b || r.oc != want_c) continue;
      any_layout = true;
      const double dm = std::fabs(lm - r.lm), dn = std::fabs(ln - r.ln), dk = std::fabs(lk - r.lk);
      if (dm > 1.5 || dn > 1.5 || dk > 1.5) continue;
      const double d = dm + dn + dk;
      if (d < best_d) {
        best_d = d;
        best = &r;
      }
    }
    if (best != nullptr || any_layout) return best;  // layout tuned: no fallback to N/N/N
  }
  return nullptr;
}

std::string default_plan_path() {
  Dl_info info;
  if (dladdr(reinterpret_cast<void*>(&default_plan_path), &info) == 0 || info.dli_fname == nullptr) return "";
  std::string lib = info.dli_fname;
  const size_t slash = lib.rfind('/');
  const std::string dir = slash == std::string::npos ? "." : lib.substr(0, slash);
  return dir + "/plans_gfx950.tsv";
}

}  // namespace shgemm_amd
