// Host runtime of mtk::shgemm on HIP: handle life cycle, argument checks,
// kernel choice (plan: tile variant and k-slices) and launch.
//
// Reference counterpart: src/shgemm.cu:401-690 (create, destroy, stream and
// level setters, the shgemm dispatcher).
#include <hip/hip_runtime.h>
#include <algorithm>
#include <climits>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <mutex>
#include <string>

#include "runtime_internal.hpp"

using shgemm_amd::check_hip;

namespace mtk {
namespace shgemm {

namespace {

// Split-K workspace: shgemm_amd::kWorkspaceBytes (64 MiB) of FP32 partials
// (reference: 3.54 MB, include/shgemm/shgemm.hpp:90, sized for one 108-SM GPU).
using shgemm_amd::aligned;
using shgemm_amd::kWorkspaceBytes;
using shgemm_amd::group_rows;
using shgemm_amd::set_lds_limit;

const shgemm_amd::KernelTable& kernel_table() {
  static shgemm_amd::KernelTable t;
  static std::once_flag once;
  std::call_once(once, [] { shgemm_amd::register_all(t); });
  return t;
}

void launch(const shgemm_amd::KernelEntry& e, dim3 grid, const shgemm_amd::GemmParams& p, hipStream_t s) {
  if (e.fn == nullptr) throw error("internal: kernel variant not compiled");
  void* args[] = {const_cast<shgemm_amd::GemmParams*>(&p)};
  check_hip(hipLaunchKernel(e.fn, grid, dim3(e.threads), args, e.lds, s), "kernel launch");
}

// The main GEMM kernel of a call, bracketed by the handle's timing events
// while enable_kernel_timing has free event pairs.
void launch_main(detail::handle_state* st, const shgemm_amd::KernelEntry& e, dim3 grid, const shgemm_amd::GemmParams& p,
                 hipStream_t s) {
  const bool timed = st->timing_used < (int)st->timing.size();
  if (timed) check_hip(hipEventRecord(st->timing[st->timing_used].first, s), "timing event");
  launch(e, grid, p, s);
  if (timed) {
    check_hip(hipEventRecord(st->timing[st->timing_used].second, s), "timing event");
    ++st->timing_used;
  }
}

void release_timing(detail::handle_state* st) {
  for (auto& ev : st->timing) {
    (void)hipEventDestroy(ev.first);
    (void)hipEventDestroy(ev.second);
  }
  st->timing.clear();
  st->timing_used = 0;
}

const char* op_name(operation_t op) { return op == op_n ? "N" : "T"; }

// A kernel choice for one call: table variant (or an explicit entry) and the
// number of k-slices.  `level` is what shgemm returns (reference
// shgemm.cu:609, 683: P0 for the k-sliced path).
struct Plan {
  int variant = shgemm_amd::V_P1;
  int slices = 1;
  int level = detail::P1;
  int requested = shgemm_amd::V_P1;  // plan variant as named by the caller / table (0..5)
  plan_source source = plan_cost_model;
  const shgemm_amd::KernelEntry* entry = nullptr;
  const char* note = "";
  bool presplit = false;  // A split once per call (plan variant 8)
  bool direct_b = false;  // pre-split plan reading op_n B as it lies (no B pack pass)
};

// Largest useful split: every slice at least 4 k-tiles, all partials inside
// the workspace.
int64_t max_slices(const shgemm_amd::KernelEntry& e, int64_t m, int64_t n, int64_t k, size_t ws_bytes) {
  int64_t s = 1;
  while (k / (s * 2) >= 4 * e.bk && (s * 2) * m * n * (int64_t)sizeof(float) <= (int64_t)ws_bytes && s < 256) s *= 2;
  return s;
}

// Estimated time (seconds) of one candidate on this device: the larger of an
// MFMA-rate term (workgroup waves x per-workgroup work at the rate measured
// for the tile) and an HBM term (operands once + C + split-K partials), plus
// a fixed cost per extra launch.
double plan_cost(const shgemm_amd::KernelEntry& e, int variant, int mode, int64_t slices, int64_t m, int64_t n,
                 int64_t k, int ncu) {
  const int64_t tiles = ((m + e.bm - 1) / e.bm) * ((n + e.bn - 1) / e.bn);
  const int level = shgemm_amd::variant_level(variant);
  // workgroups resident per CU and the effective MFMA rate one CU sustains on
  // this tile (TFLOP/s effective per CU), from the tile sweep
  const double per_cu_tflops = (mode == shgemm_amd::MODE_BF16 ? 0.8 : 1.8) * (level == 1 ? 1.0 : 0.55) *
                               (variant == shgemm_amd::V_P2 ? 1.03 : 1.0);
  const int wg_per_cu = level == 1 ? e.wgs_per_cu : 4;
  const int64_t wgs = tiles * slices;
  const int64_t waves = (wgs + (int64_t)ncu * wg_per_cu - 1) / ((int64_t)ncu * wg_per_cu);
  const double wg_flops = 2.0 * e.bm * e.bn * ((double)k / slices);
  const double t_mfma = waves * wg_flops / (per_cu_tflops * 1e12 / wg_per_cu);
  const double bytes = 4.0 * m * k + 2.0 * k * n + 4.0 * m * n + (slices > 1 ? 8.0 * slices * m * n : 0.0);
  const double t_hbm = bytes / 5.0e12;
  return std::max(t_mfma, t_hbm) + (slices > 1 ? 6e-6 : 0.0);
}

// The deep-prefetch kernel applies to the prescaled fp16 mode with op_b = N,
// 16-byte-aligned B and whole k-tiles, where its per-lane byte offsets fit 32
// bits: B rows (bn of them, ldb apart, FP16), A rows of a k-contiguous op_t
// tile (bm of them, lda apart, FP32) and, for op_n A, the row offset within a
// column (4 * (lda - 1) at most).
const shgemm_amd::KernelEntry* bd_entry(const detail::handle_state* st, int mode, int oa, int ob, int oc, int64_t k,
                                        bool vec_a, bool vec_b, int64_t lda, int64_t ldb) {
  const auto& bd = st->table.shgemm_bd[oa][oc];
  constexpr int64_t kLim = int64_t(1) << 32;
  const bool a_fits = oa == shgemm_amd::OP_N ? 4 * lda < kLim : (int64_t)bd.bm * lda * 4 < kLim;
  if (bd.fn != nullptr && mode == shgemm_amd::MODE_FP16S && ob == shgemm_amd::OP_N && vec_b && k % bd.bk == 0 &&
      (oa == shgemm_amd::OP_N || vec_a) && (int64_t)bd.bn * ldb * 2 < kLim && a_fits)
    return &bd;
  return nullptr;
}

// The pre-split kernel: every compute mode, an even number of whole k-tiles
// (A is packed by the plan's pass whatever its layout; B is packed too,
// except op_n B in the fp16 modes, read where it lies).
const shgemm_amd::KernelEntry* ps_entry(const detail::handle_state* st, int mode, int ob, int oc, int64_t m,
                                        int64_t n, int64_t k, int64_t ldb, bool& direct_b) {
  direct_b = false;
  // 128 x 128 tiles where fewer than 3/4 of the CUs would get a 256 x 256 one
  // (one box: 4096 x 2048 x 4096 128 tiles, 360 against 309 TFLOP/s for the
  // 256 x 256 form; 3072^3 144 tiles, level; 3072 x 4096 x 4096 192 tiles,
  // 421 against 437)
  const bool small = ((m + 255) / 256) * ((n + 255) / 256) * 4 < std::max(1, st->num_cus) * 3;
  const shgemm_amd::KernelEntry* e = nullptr;
  if (mode == shgemm_amd::MODE_FP16S) {
    // op_n B is read in place when the tile's column offsets fit 32 bits, else packed
    const auto& tab = small ? st->table.shgemm_ps128 : st->table.shgemm_ps;
    direct_b = ob == shgemm_amd::OP_N && (int64_t)tab[1][oc].bn * ldb * 2 < (int64_t(1) << 32);
    e = &tab[direct_b ? 1 : 0][oc];
  } else if (mode == shgemm_amd::MODE_BF16) {
    e = small ? &st->table.shgemm_ps128_bf16[oc] : &st->table.shgemm_ps_bf16[oc];  // B always packed
  } else if (mode == shgemm_amd::MODE_FP16) {
    // prescaling off: the 128 x 128 form at every size
    const auto& tab = st->table.shgemm_ps128_fp16;
    direct_b = ob == shgemm_amd::OP_N && (int64_t)tab[1][oc].bn * ldb * 2 < (int64_t(1) << 32);
    e = &tab[direct_b ? 1 : 0][oc];
  }
  if (e != nullptr && e->fn != nullptr && k >= 2 * e->bk && k % (2 * e->bk) == 0) return e;
  direct_b = false;
  return nullptr;
}

// The ring kernel runs whole 256 x 256 tiles and whole k-tiles (at least
// two), with vector-aligned operands whose tile offsets fit 32 bits.
bool ring_applies(const shgemm_amd::KernelEntry& e, int64_t m, int64_t n, int64_t k, bool vec_a, bool vec_b,
                  int64_t lda, int64_t ldb) {
  return e.fn != nullptr && vec_a && vec_b && m % e.bm == 0 && n % e.bn == 0 && k % e.bk == 0 && k >= 2 * e.bk &&
         (int64_t)e.bk * lda * 4 < (int64_t(1) << 32) && (int64_t)e.bn * ldb * 2 < (int64_t(1) << 32);
}

// A plan variant 0..8 with a slice count -> a launchable plan.  Split
// variants get 2..max_slices slices; when k or the workspace allows no split
// they run as the unsplit variant of the same tile.  Variants that do not
// apply to the arguments run as the nearest one that does.
Plan resolve_plan(const detail::handle_state* st, const shgemm_amd::KernelEntry (&tab)[shgemm_amd::kNumVariants],
                  int requested, int slices, int mode, int oa, int ob, int oc, int64_t m, int64_t n, int64_t k,
                  bool vec_a, bool vec_b, int64_t lda, int64_t ldb) {
  Plan pl;
  pl.requested = requested;
  // a handle sharing the device runs the whole-CU 256 x 256 plans as P1
  if (st->shared_device &&
      (requested == shgemm_amd::V_P2 || requested == shgemm_amd::V_P2R || requested == plan_variant_ps))
    requested = shgemm_amd::V_P1;
  if (requested == shgemm_amd::V_P2R &&
      !ring_applies(tab[shgemm_amd::V_P2R], m, n, k, vec_a, vec_b, lda, ldb))
    requested = tab[shgemm_amd::V_P2].fn != nullptr ? shgemm_amd::V_P2 : shgemm_amd::V_P1;
  if (requested == plan_variant_ps) {
    bool direct_b = false;
    pl.entry = ps_entry(st, mode, ob, oc, m, n, k, ldb, direct_b);
    if (pl.entry != nullptr) {
      pl.variant = shgemm_amd::V_P2;
      pl.presplit = true;
      pl.direct_b = direct_b;
      pl.note = direct_b ? " pre-split A, B in place" : " pre-split";
    } else {
      pl.variant = tab[shgemm_amd::V_P2].fn != nullptr ? shgemm_amd::V_P2 : shgemm_amd::V_P1;
    }
  } else if (requested == plan_variant_bd) {
    pl.variant = shgemm_amd::V_P1;
    pl.entry = bd_entry(st, mode, oa, ob, oc, k, vec_a, vec_b, lda, ldb);
    if (pl.entry != nullptr) pl.note = " deep-prefetch";
  } else if (requested == plan_variant_p0w) {
    pl.variant = shgemm_amd::V_P0;
    const auto& e = st->table.shgemm_p0w[oa][ob][oc];
    if (mode == shgemm_amd::MODE_FP16S && e.fn != nullptr) {
      pl.entry = &e;
      pl.note = " whole-row exponents";
    }
  } else if (requested == plan_variant_skd || requested == plan_variant_wsk) {
    // prescaled fp16 k-sliced tiles with their own kernel tables; in the
    // other modes plan 10 runs as 2 and plan 11 as 3
    const bool deep = requested == plan_variant_skd;
    const auto& e = deep ? st->table.shgemm_skd[oa][ob][oc] : st->table.shgemm_wsk[oa][ob][oc];
    const int base = deep ? shgemm_amd::V_SK : shgemm_amd::V_P1SK;
    if (mode == shgemm_amd::MODE_FP16S && e.fn != nullptr) {
      const int64_t smax = max_slices(e, m, n, k, st->ws_bytes);
      pl.slices = (int)std::max<int64_t>(1, std::min<int64_t>(std::max(2, slices), smax));
      if (pl.slices > 1) {
        pl.variant = base;
        pl.entry = &e;
        pl.note = deep ? " deep k-tile" : " 64x128 wavefronts";
      } else {
        pl.variant = shgemm_amd::V_P1;  // k too short to split: the unsplit large tile, as for plans 2 and 3
      }
    } else {
      return resolve_plan(st, tab, base, slices, mode, oa, ob, oc, m, n, k, vec_a, vec_b, lda, ldb);
    }
  } else if (requested >= 0 && requested < shgemm_amd::kNumVariants && tab[requested].fn != nullptr) {
    pl.variant = requested;
    if (shgemm_amd::variant_is_split(requested)) {
      const int64_t smax = max_slices(tab[requested], m, n, k, st->ws_bytes);
      pl.slices = (int)std::max<int64_t>(1, std::min<int64_t>(std::max(2, slices), smax));
      if (pl.slices == 1) pl.variant = requested == shgemm_amd::V_P0SK ? shgemm_amd::V_P0 : shgemm_amd::V_P1;
    }
  }
  pl.level = pl.slices > 1 ? detail::P0 : shgemm_amd::variant_level(pl.variant);
  return pl;
}

// MTK_SHGEMM_TUNE_P1 (tile sweep): an unsplit large-tile op_n/op_n/op_n plan
// runs the selected candidate configuration instead.
bool apply_tile_sweep(const detail::handle_state* st, Plan& pl, int mode, int oa, int ob, int oc) {
  if (st->tune_p1 < 0 || pl.slices != 1 || pl.variant != shgemm_amd::V_P1 || pl.entry != nullptr) return false;
  if (st->table.tune[st->tune_p1].mode != mode || oa != shgemm_amd::OP_N || ob != shgemm_amd::OP_N ||
      oc != shgemm_amd::OP_N)
    return false;
  pl.entry = &st->table.tune[st->tune_p1];
  pl.note = " tile-sweep candidate";
  return true;
}

Plan choose_plan(const shgemmHandle_t& handle, const detail::handle_state* st,
                 const shgemm_amd::KernelEntry (&tab)[shgemm_amd::kNumVariants], int mode, int oa, int ob, int oc,
                 int64_t m, int64_t n, int64_t k, bool vec_a, bool vec_b, int64_t lda, int64_t ldb) {
  // 1. explicit override (measurement and tuning)
  if (st->forced_variant >= 0) {
    Plan pl = resolve_plan(st, tab, st->forced_variant, st->forced_slices, mode, oa, ob, oc, m, n, k, vec_a, vec_b,
                           lda, ldb);
    pl.source = plan_forced;
    apply_tile_sweep(st, pl, mode, oa, ob, oc);
    return pl;
  }
  // 2. kernel level fixed by the caller (reference shgemm.cu:655): no split-K
  if (handle.fixed_lernel_level < detail::num_levels) {
    Plan pl;
    pl.variant = handle.fixed_lernel_level == detail::P1 ? shgemm_amd::V_P1 : shgemm_amd::V_P0;
    pl.requested = pl.variant;
    pl.level = handle.fixed_lernel_level;
    pl.source = plan_fixed_level;
    return pl;
  }
  // 3. measured plan table
  if (const shgemm_amd::PlanRow* r =
          shgemm_amd::find_plan(st->plans, shgemm_amd::PLAN_SINGLE, mode, oa, ob, oc, m, n, k)) {
    Plan pl = resolve_plan(st, tab, r->choice, r->slices, mode, oa, ob, oc, m, n, k, vec_a, vec_b, lda, ldb);
    pl.source = plan_table;
    apply_tile_sweep(st, pl, mode, oa, ob, oc);
    return pl;
  }
  // 4. cost model over every variant and split count
  Plan best;
  const int ncu = std::max(1, st->num_cus);
  double best_t = 1e30;
  for (int v = 0; v < shgemm_amd::kNumVariants; ++v) {
    const auto& e = tab[v];
    if (e.fn == nullptr || v == shgemm_amd::V_P2R) continue;  // the ring variant is chosen by measurement only
    if (v == shgemm_amd::V_P2 && st->shared_device) continue;
    const bool split = shgemm_amd::variant_is_split(v);
    const int64_t smax = split ? max_slices(e, m, n, k, st->ws_bytes) : 1;
    for (int64_t s = split ? 2 : 1; s <= smax; s *= 2) {
      const double t = plan_cost(e, v, mode, s, m, n, k, ncu);
      if (t < best_t) {
        best_t = t;
        best.variant = v;
        best.slices = (int)s;
      }
    }
  }
  best.requested = best.variant;
  best.level = best.slices > 1 ? detail::P0 : shgemm_amd::variant_level(best.variant);
  best.source = plan_cost_model;
  // tile-sweep candidate or deep-prefetch kernel for the large-tile fp16 path
  if (best.slices == 1 && best.variant == shgemm_amd::V_P1) {
    if (apply_tile_sweep(st, best, mode, oa, ob, oc)) {
    } else if (st->use_bd) {
      best.entry = bd_entry(st, mode, oa, ob, oc, k, vec_a, vec_b, lda, ldb);
      if (best.entry != nullptr) {
        best.requested = plan_variant_bd;
        best.note = " deep-prefetch";
      }
    }
  }
  return best;
}

// Whether `s` is capturing a graph.  A query that fails (the legacy stream
// while another stream captures) counts as capturing: nothing may then
// synchronise or allocate.
bool stream_capturing(hipStream_t s) {
  hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
  if (hipStreamIsCapturing(s, &cs) != hipSuccess) {
    (void)hipGetLastError();
    return true;
  }
  return cs != hipStreamCaptureStatusNone;
}

// Waits on the host until every earlier call that used the handle's scratch
// has finished (scratch_ev chains them), so that a scratch buffer may be freed.
void drain_scratch(detail::handle_state* st) {
  if (st->scratch_used) check_hip(hipEventSynchronize(st->scratch_ev), "hipEventSynchronize(scratch order)");
}

// Row-exponent scratch for m rows and the row pass's partial maxima.  Growing
// it synchronises and allocates, so it cannot happen during graph capture.
void ensure_rows(detail::handle_state* st, size_t m, bool capturing) {
  if (st->row_cap >= m && st->row_part != nullptr) return;
  if (capturing)
    throw error("this call needs the handle's row scratch to grow, which cannot happen while its stream captures "
                "a graph: call mtk::shgemm::reserve with the largest m first");
  drain_scratch(st);
  if (st->row_part == nullptr) {
    check_hip(hipMalloc(&st->row_part, shgemm_amd::kRowPartialBytes), "hipMalloc(row scratch)");
    st->row_part_bytes = shgemm_amd::kRowPartialBytes;
  }
  if (st->row_cap < m) {
    if (st->row_max) (void)hipFree(st->row_max);
    if (st->row_exp) (void)hipFree(st->row_exp);
    st->row_max = nullptr;
    st->row_exp = nullptr;
    st->row_cap = 0;
    check_hip(hipMalloc(&st->row_max, m * sizeof(uint32_t)), "hipMalloc(row scratch)");
    check_hip(hipMalloc(&st->row_exp, m * sizeof(int)), "hipMalloc(row scratch)");
    st->row_cap = m;
  }
}

// Pre-split buffer of at least `bytes`.  False when it would have to grow
// during graph capture or the device cannot hold it.
bool ensure_presplit(detail::handle_state* st, size_t bytes, bool capturing) {
  if (st->ps_bytes >= bytes) return true;
  if (capturing) return false;
  drain_scratch(st);
  if (st->ps_buf) (void)hipFree(st->ps_buf);
  st->ps_buf = nullptr;
  st->ps_bytes = 0;
  if (hipMalloc(&st->ps_buf, bytes) != hipSuccess) {
    (void)hipGetLastError();
    st->ps_buf = nullptr;
    return false;
  }
  st->ps_bytes = bytes;
  return true;
}

const char* source_name(plan_source s) {
  switch (s) {
    case plan_table: return "table";
    case plan_forced: return "forced";
    case plan_fixed_level: return "fixed level";
    default: return "cost model";
  }
}

int compute_mode(const detail::handle_state* st, tc_t compute_type) {
  return compute_type == tf32 ? shgemm_amd::MODE_BF16
         : st->exp_scaling    ? shgemm_amd::MODE_FP16S
                              : shgemm_amd::MODE_FP16;
}

}  // namespace

void create(shgemmHandle_t& handle) {
  handle.cuda_stream = nullptr;
  handle.fixed_lernel_level = detail::num_levels;
  handle.debug_mode = 0;
  const char* dbg = std::getenv("MTK_SHGEMM_DEBUG");
  if (dbg != nullptr && dbg[0] == '1') handle.debug_mode = 1;
  const char* lvl = std::getenv("MTK_SHGEMM_KERNEL_LEVEL");
  if (lvl != nullptr && (lvl[0] == '0' || lvl[0] == '1') && lvl[1] == '\0')
    handle.fixed_lernel_level = static_cast<detail::kernel_level>(lvl[0] - '0');

  auto* st = new detail::handle_state();
  try {
    // Measured plan table: $MTK_SHGEMM_PLANS ("none" for no table), else the
    // table shipped beside the library when it exists.
    const char* pe = std::getenv("MTK_SHGEMM_PLANS");
    std::string path = pe != nullptr ? std::string(pe) : shgemm_amd::default_plan_path();
    if (pe == nullptr && !path.empty() && !std::ifstream(path).good()) path.clear();
    if (!path.empty() && path != "none") {
      std::string err;
      if (!shgemm_amd::parse_plan_file(path, st->plans, err)) throw error("plan table: " + err);
      st->plans_path = path;
    }
    check_hip(hipGetDevice(&st->device), "hipGetDevice");
    check_hip(hipDeviceGetAttribute(&st->num_cus, hipDeviceAttributeMultiprocessorCount, st->device),
              "hipDeviceGetAttribute");
    st->table = kernel_table();
    for (auto& a : st->table.shgemm)
      for (auto& b : a)
        for (auto& c : b)
          for (auto& d : c)
            for (auto& e : d)
              if (e.fn) set_lds_limit(e);
    for (int i = 0; i < st->table.ntune; ++i) set_lds_limit(st->table.tune[i]);
    for (auto& a : st->table.shgemm_bd)
      for (auto& e : a)
        if (e.fn) set_lds_limit(e);
    for (auto& a : st->table.shgemm_ps)
      for (auto& e : a)
        if (e.fn) set_lds_limit(e);
    for (auto& e : st->table.shgemm_ps_bf16)
      if (e.fn) set_lds_limit(e);
    for (auto& a : st->table.shgemm_ps128)
      for (auto& e : a)
        if (e.fn) set_lds_limit(e);
    for (auto& e : st->table.shgemm_ps128_bf16)
      if (e.fn) set_lds_limit(e);
    for (auto& a : st->table.shgemm_ps128_fp16)
      for (auto& e : a)
        if (e.fn) set_lds_limit(e);
    check_hip(hipEventCreateWithFlags(&st->scratch_ev, hipEventDisableTiming), "hipEventCreate(scratch order)");
    shgemm_amd::init_static_scratch();
    check_hip(hipMalloc(&st->ws, kWorkspaceBytes), "hipMalloc(workspace)");
    st->ws_bytes = kWorkspaceBytes;
  } catch (...) {
    if (st->ws) (void)hipFree(st->ws);
    if (st->scratch_ev) (void)hipEventDestroy(st->scratch_ev);
    delete st;
    throw;
  }
  const char* tp = std::getenv("MTK_SHGEMM_TUNE_P1");
  if (tp != nullptr && tp[0] >= '0' && tp[0] <= '9') {
    const int i = std::atoi(tp);
    if (i < st->table.ntune) st->tune_p1 = i;
  }
  const char* fp = std::getenv("MTK_SHGEMM_PLAN");
  if (fp != nullptr) {
    int v = -1, sl = 1;
    if (std::sscanf(fp, "%d,%d", &v, &sl) >= 1 && v >= 0 && v < num_plan_variants) {
      st->forced_variant = v;
      st->forced_slices = std::max(1, sl);
    }
  }
  // Deep-prefetch kernel where the cost model picks the large tile: opt-in (at
  // 176 VGPRs it runs 2 wavefronts per SIMD; the tuner measures it per shape).
  const char* bd = std::getenv("MTK_SHGEMM_BD");
  st->use_bd = (bd != nullptr && bd[0] == '1');
  // Per-row exponent prescaling of the fp16 compute type is on unless
  // MTK_SHGEMM_EXP_SCALING=0 (then fp16 has FP16's exponent range, as in the reference).
  const char* sc = std::getenv("MTK_SHGEMM_EXP_SCALING");
  st->exp_scaling = !(sc != nullptr && sc[0] == '0');
  handle.state = st;
}

void destroy(shgemmHandle_t& handle) {
  if (handle.state == nullptr) return;
  auto* st = handle.state;
  // Work enqueued by this handle may still read the scratch buffers.
  (void)hipDeviceSynchronize();
  if (st->ws) (void)hipFree(st->ws);
  if (st->row_part) (void)hipFree(st->row_part);
  if (st->row_max) (void)hipFree(st->row_max);
  if (st->row_exp) (void)hipFree(st->row_exp);
  if (st->ps_buf) (void)hipFree(st->ps_buf);
  if (st->scratch_ev) (void)hipEventDestroy(st->scratch_ev);
  release_timing(st);
  delete st;
  handle.state = nullptr;
}

void enable_kernel_level_fixing(shgemmHandle_t& handle, const detail::kernel_level kernel_level) {
  if (kernel_level != detail::P0 && kernel_level != detail::P1) throw error("kernel level must be P0 or P1");
  handle.fixed_lernel_level = kernel_level;
}

void disable_kernel_level_fixing(shgemmHandle_t& handle) { handle.fixed_lernel_level = detail::num_levels; }

void set_debug_mode(shgemmHandle_t& handle, const unsigned on) { handle.debug_mode = on; }

void set_cuda_stream(shgemmHandle_t& handle, hipStream_t const cuda_stream) { handle.cuda_stream = cuda_stream; }

void set_exponent_scaling(shgemmHandle_t& handle, const bool on) {
  if (handle.state == nullptr) throw error("handle not created");
  std::lock_guard<std::mutex> lock(handle.state->mu);
  handle.state->exp_scaling = on;
}

void set_shared_device(shgemmHandle_t& handle, const bool on) {
  if (handle.state == nullptr) throw error("handle not created");
  std::lock_guard<std::mutex> lock(handle.state->mu);
  handle.state->shared_device = on;
}

void force_plan(shgemmHandle_t& handle, const int variant, const int slices) {
  if (handle.state == nullptr) throw error("handle not created");
  if (variant >= num_plan_variants) throw error("plan variant must be below " + std::to_string(num_plan_variants));
  if (variant >= 0 && (slices < 1 || slices > 256)) throw error("slices must be in 1..256");
  std::lock_guard<std::mutex> lock(handle.state->mu);
  handle.state->forced_variant = variant < 0 ? -1 : variant;
  handle.state->forced_slices = variant < 0 ? 1 : slices;
}

void load_plans(shgemmHandle_t& handle, const char* path) {
  if (handle.state == nullptr) throw error("handle not created");
  std::vector<shgemm_amd::PlanRow> rows;
  if (path != nullptr && path[0] != '\0') {
    std::string err;
    if (!shgemm_amd::parse_plan_file(path, rows, err)) throw error("plan table: " + err);
  }
  std::lock_guard<std::mutex> lock(handle.state->mu);
  handle.state->plans.swap(rows);
  handle.state->plans_path = path != nullptr ? path : "";
}

plan_t query_plan(const shgemmHandle_t handle, const operation_t op_a, const operation_t op_b, const std::size_t m,
                  const std::size_t n, const std::size_t k, const tc_t compute_type, const operation_t op_c) {
  if (handle.state == nullptr) throw error("handle not created");
  if ((op_a != op_n && op_a != op_t) || (op_b != op_n && op_b != op_t) || (op_c != op_n && op_c != op_t))
    throw error("invalid operation_t");
  if (compute_type != fp16 && compute_type != tf32) throw error("invalid tc_t");
  detail::handle_state* st = handle.state;
  std::lock_guard<std::mutex> lock(st->mu);
  const int mode = compute_mode(st, compute_type);
  const int oa = op_a == op_n ? shgemm_amd::OP_N : shgemm_amd::OP_T;
  const int ob = op_b == op_n ? shgemm_amd::OP_N : shgemm_amd::OP_T;
  const int oc = op_c == op_n ? shgemm_amd::OP_N : shgemm_amd::OP_T;
  // aligned operands with the smallest leading dimensions
  const int64_t lda = (int64_t)(op_a == op_n ? m : k), ldb = (int64_t)(op_b == op_n ? k : n);
  const Plan pl = choose_plan(handle, st, st->table.shgemm[oa][ob][oc][mode], mode, oa, ob, oc, (int64_t)m, (int64_t)n,
                              (int64_t)k, true, true, lda, ldb);
  plan_t out;
  out.variant = pl.presplit ? plan_variant_ps
                : (pl.entry != nullptr && (pl.requested == plan_variant_bd || pl.requested == plan_variant_p0w ||
                                          pl.requested == plan_variant_skd || pl.requested == plan_variant_wsk)
                       ? pl.requested
                       : pl.variant);
  out.slices = pl.slices;
  out.source = pl.source;
  return out;
}

namespace {

// One shgemm call.  `fixed` null: the handle's scratch, grown on demand and
// ordered across streams.  `fixed` set: that scratch, and nothing is
// allocated, freed or synchronised.
detail::kernel_level shgemm_call(const shgemmHandle_t handle, const operation_t op_a, const operation_t op_b,
                                 const std::size_t m, const std::size_t n, const std::size_t k,
                                 const float* const alpha_ptr, const float* const a_ptr, const std::size_t lda,
                                 const half* const b_ptr, const std::size_t ldb, const float* const beta_ptr,
                                 float* const c_ptr, const std::size_t ldc, const tc_t compute_type,
                                 const operation_t op_c, const shgemm_amd::CallScratch* fixed) {
  shgemm_amd::validate_args(handle, op_a, op_b, op_c, m, n, k, alpha_ptr, a_ptr, lda, b_ptr, ldb, beta_ptr, c_ptr, ldc);
  if (compute_type != fp16 && compute_type != tf32) throw error("invalid tc_t");
  detail::handle_state* st = handle.state;
  const hipStream_t stream = handle.cuda_stream;
  if (m == 0 || n == 0) return detail::P0;

  const float alpha = *alpha_ptr, beta = *beta_ptr;
  const int inner = (int)(op_c == op_n ? m : n);
  const int64_t elems = (int64_t)m * (int64_t)n;
  if (k == 0 || alpha == 0.0f) {
    // C = beta * C without touching A or B (BLAS semantics)
    check_hip(shgemm_amd::launch_scale_c(elems, inner, beta, c_ptr, (int64_t)ldc, stream), "scale kernel");
    return detail::P0;
  }

  const int mode = compute_mode(st, compute_type);
  const int oa = op_a == op_n ? shgemm_amd::OP_N : shgemm_amd::OP_T;
  const int ob = op_b == op_n ? shgemm_amd::OP_N : shgemm_amd::OP_T;
  const int oc = op_c == op_n ? shgemm_amd::OP_N : shgemm_amd::OP_T;
  const bool scaled = (mode == shgemm_amd::MODE_FP16S);

  shgemm_amd::GemmParams p{};
  p.m = (int)m;
  p.n = (int)n;
  p.k = (int)k;
  p.alpha = alpha;
  p.beta = beta;
  p.a = a_ptr;
  p.lda = (int64_t)lda;
  p.b = b_ptr;
  p.ldb = (int64_t)ldb;
  p.c = c_ptr;
  p.ldc = (int64_t)ldc;
  // Vector loads: 16 bytes of FP32 for A; 16 (k-contiguous) or 8 bytes of FP16 for B.
  p.vec_a = shgemm_amd::vec_a_ok(a_ptr, lda);
  p.vec_b = shgemm_amd::vec_b_ok(b_ptr, ldb, ob);

  std::lock_guard<std::mutex> lock(st->mu);
  const auto& tab = st->table.shgemm[oa][ob][oc][mode];
  Plan plan = choose_plan(handle, st, tab, mode, oa, ob, oc, (int64_t)m, (int64_t)n, (int64_t)k, p.vec_a != 0,
                          p.vec_b != 0, p.lda, p.ldb);
  // The capture state matters only to calls that use the handle's scratch.
  int cap_state = -1;
  const auto capturing = [&] {
    if (cap_state < 0) cap_state = stream_capturing(stream) ? 1 : 0;
    return cap_state == 1;
  };
  if (plan.presplit && fixed != nullptr) {
    // no buffer for the packed operands: the 256 x 256 tile (bit-identical)
    plan.presplit = false;
    plan.entry = nullptr;
    plan.variant = tab[shgemm_amd::V_P2].fn != nullptr ? shgemm_amd::V_P2 : shgemm_amd::V_P1;
    plan.note = " (pre-split plan as the 256x256 tile: no allocation)";
  }
  if (plan.presplit) {
    // A's packed parts (prescaled fp16 hi and lo: 4 bytes per element; tf32
    // three BF16 parts: 6; m rounded up to 256) and, unless op_n B is read in
    // place, B's packed tiles (2 or 4 bytes per element, n rounded up to 256),
    // in a handle buffer grown on demand; where the device cannot hold them,
    // or the buffer would have to grow during graph capture, the call runs
    // the register-split 256 x 256 tile (bit-identical).
    const size_t bytes = shgemm_amd::presplit_bytes(p.m, p.k, mode) +
                         (plan.direct_b ? 0 : shgemm_amd::pack_b_bytes(p.n, p.k, mode));
    if (!ensure_presplit(st, bytes, capturing())) {
      plan.presplit = false;
      plan.entry = nullptr;
      plan.variant = tab[shgemm_amd::V_P2].fn != nullptr ? shgemm_amd::V_P2 : shgemm_amd::V_P1;
      plan.note = " (pre-split buffer unavailable)";
    }
  }
  const shgemm_amd::KernelEntry& e = plan.entry != nullptr ? *plan.entry : tab[plan.variant];
  // Prescaled fp16: the kernels on wavefront tiles of up to 64 rows (P0, P1,
  // the split-K plans) keep running row exponents themselves; the 256 x 256
  // family and the three-stage kernel read them from a pass over A.
  const bool row_pass = scaled && e.needs_row_exp;
  shgemm_amd::CallScratch sc;
  if (fixed != nullptr) {
    if (row_pass && m > fixed->row_cap)
      throw error("m = " + std::to_string(m) + " exceeds the " + std::to_string(fixed->row_cap) +
                  " rows of row-exponent scratch the allocation-free entry point holds for this plan");
    sc = *fixed;
  } else {
    if (row_pass) ensure_rows(st, m, capturing());
    sc.ws = st->ws;
    sc.row_part = st->row_part;
    sc.row_max = st->row_max;
    sc.row_exp = st->row_exp;
    sc.row_cap = st->row_cap;
  }
  // calls on the handle's scratch run one after another, whatever their streams
  const bool ordered = fixed == nullptr && (row_pass || plan.presplit || plan.slices > 1) && !capturing();
  if (ordered && st->scratch_used && st->scratch_stream != stream)
    check_hip(hipStreamWaitEvent(stream, st->scratch_ev, 0), "hipStreamWaitEvent(scratch order)");

  if (row_pass) {
    check_hip(shgemm_amd::launch_row_exponents(a_ptr, p.lda, oa, p.m, p.k, sc.row_max, sc.row_exp, sc.row_part,
                                               shgemm_amd::kRowPartialBytes, stream),
              "row exponent pass");
    p.row_exp = sc.row_exp;
  }
  if (plan.presplit) {
    const size_t a_bytes = shgemm_amd::presplit_bytes(p.m, p.k, mode);
    uint4* apack = reinterpret_cast<uint4*>(st->ps_buf);
    uint4* bpack = reinterpret_cast<uint4*>(reinterpret_cast<char*>(st->ps_buf) + a_bytes);
    check_hip(shgemm_amd::launch_presplit_a(a_ptr, (int64_t)lda, oa, p.m, p.k, mode, scaled ? st->row_exp : nullptr,
                                            apack, stream),
              "pre-split pass");
    p.a = reinterpret_cast<const float*>(apack);  // the kernel reads the packed tiles
    if (!plan.direct_b) {
      check_hip(shgemm_amd::launch_pack_b(reinterpret_cast<const uint16_t*>(b_ptr), (int64_t)ldb, ob, p.n, p.k, mode,
                                          bpack, stream),
                "B pack pass");
      p.b = bpack;
    }
  }
  p.tiles_m = (int)((m + e.bm - 1) / e.bm);
  p.tiles_n = (int)((n + e.bn - 1) / e.bn);
  p.group_m = group_rows(p.tiles_m);
  const int64_t tiles = (int64_t)p.tiles_m * p.tiles_n;
  if (tiles > INT_MAX) throw error("too many output tiles");
  if (handle.debug_mode)
    std::printf(
        "[shape=(%zu,%zu,%zu),op_a=%s,op_b=%s,compute_type=%s] tile=%dx%dx%d slices=%d kernel_level=%d%s plan=%s%s\n",
        m, n, k, op_name(op_a), op_name(op_b), compute_type == tf32 ? "TF32" : "FP16", e.bm, e.bn, e.bk, plan.slices,
        plan.level, scaled ? " prescaled" : "", source_name(plan.source), plan.note);
  if (handle.debug_mode) std::fflush(stdout);
  if (plan.slices > 1) {
    const int64_t kslice = (((int64_t)k + plan.slices - 1) / plan.slices + e.bk - 1) / e.bk * e.bk;
    const int64_t used = ((int64_t)k + kslice - 1) / kslice;
    p.k_slice = (int)kslice;
    p.ws = sc.ws;
    launch_main(st, e, dim3((unsigned)tiles, (unsigned)used), p, stream);
    check_hip(shgemm_amd::launch_splitk_reduce(sc.ws, (int)used, elems, inner, alpha, beta, c_ptr, (int64_t)ldc,
                                               stream),
              "split-K reduction");
  } else {
    launch_main(st, e, dim3((unsigned)tiles), p, stream);
  }
  if (ordered) {
    check_hip(hipEventRecord(st->scratch_ev, stream), "hipEventRecord(scratch order)");
    st->scratch_stream = stream;
    st->scratch_used = true;
  }
  return static_cast<detail::kernel_level>(plan.level);
}

}  // namespace

detail::kernel_level shgemm(const shgemmHandle_t handle, const operation_t op_a, const operation_t op_b,
                            const std::size_t m, const std::size_t n, const std::size_t k, const float* const alpha_ptr,
                            const float* const a_ptr, const std::size_t lda, const half* const b_ptr,
                            const std::size_t ldb, const float* const beta_ptr, float* const c_ptr,
                            const std::size_t ldc, const tc_t compute_type, const operation_t op_c) {
  return shgemm_call(handle, op_a, op_b, m, n, k, alpha_ptr, a_ptr, lda, b_ptr, ldb, beta_ptr, c_ptr, ldc,
                     compute_type, op_c, nullptr);
}

void reserve(shgemmHandle_t& handle, const std::size_t m, const std::size_t n, const std::size_t k) {
  if (handle.state == nullptr) throw error("handle not created");
  if (m > (size_t)INT_MAX || n > (size_t)INT_MAX || k > (size_t)INT_MAX)
    throw error("m, n and k must each be below 2^31");
  auto* st = handle.state;
  std::lock_guard<std::mutex> lock(st->mu);
  if (stream_capturing(handle.cuda_stream)) throw error("reserve cannot run while the handle's stream captures a graph");
  if (m > 0) ensure_rows(st, m, false);
  // a handle sharing the device never runs the pre-split plan
  if (m > 0 && n > 0 && k > 0 && !st->shared_device) {
    // the pre-split plan needs k % 64 == 0; reserve for k rounded up to that
    const int64_t k64 = ((int64_t)k + 63) / 64 * 64;
    // sized for the larger (tf32) packing of A and B
    if (k64 <= INT_MAX) (void)ensure_presplit(st, shgemm_amd::presplit_bytes((int)m, (int)k64, shgemm_amd::MODE_BF16) +
                                                      shgemm_amd::pack_b_bytes((int)n, (int)k64, shgemm_amd::MODE_BF16),
                                              false);
  }
}

void sgemm_fp32(const shgemmHandle_t handle, const operation_t op_a, const operation_t op_b, const std::size_t m,
                const std::size_t n, const std::size_t k, const float* const alpha_ptr, const float* const a_ptr,
                const std::size_t lda, const float* const b_ptr, const std::size_t ldb, const float* const beta_ptr,
                float* const c_ptr, const std::size_t ldc, const operation_t op_c) {
  shgemm_amd::validate_args(handle, op_a, op_b, op_c, m, n, k, alpha_ptr, a_ptr, lda, b_ptr, ldb, beta_ptr, c_ptr, ldc);
  detail::handle_state* st = handle.state;
  const hipStream_t stream = handle.cuda_stream;
  if (m == 0 || n == 0) return;
  const float alpha = *alpha_ptr, beta = *beta_ptr;
  const int inner = (int)(op_c == op_n ? m : n);
  if (k == 0 || alpha == 0.0f) {
    check_hip(shgemm_amd::launch_scale_c((int64_t)m * (int64_t)n, inner, beta, c_ptr, (int64_t)ldc, stream),
              "scale kernel");
    return;
  }
  const int oa = op_a == op_n ? shgemm_amd::OP_N : shgemm_amd::OP_T;
  const int ob = op_b == op_n ? shgemm_amd::OP_N : shgemm_amd::OP_T;
  const int oc = op_c == op_n ? shgemm_amd::OP_N : shgemm_amd::OP_T;
  const auto& e = st->table.sgemm[oa][ob][oc];
  shgemm_amd::GemmParams p{};
  p.m = (int)m;
  p.n = (int)n;
  p.k = (int)k;
  p.alpha = alpha;
  p.beta = beta;
  p.a = a_ptr;
  p.lda = (int64_t)lda;
  p.b = b_ptr;
  p.ldb = (int64_t)ldb;
  p.c = c_ptr;
  p.ldc = (int64_t)ldc;
  p.vec_a = aligned(a_ptr, 16) && lda % 4 == 0;
  p.vec_b = aligned(b_ptr, 16) && ldb % 4 == 0;
  std::lock_guard<std::mutex> lock(st->mu);
  p.tiles_m = (int)((m + e.bm - 1) / e.bm);
  p.tiles_n = (int)((n + e.bn - 1) / e.bn);
  p.group_m = group_rows(p.tiles_m);
  const int64_t tiles = (int64_t)p.tiles_m * p.tiles_n;
  if (tiles > INT_MAX) throw error("too many output tiles");
  if (handle.debug_mode)
    std::printf("[shape=(%zu,%zu,%zu),op_a=%s,op_b=%s] sgemm_fp32 tile=%dx%dx%d\n", m, n, k, op_name(op_a),
                op_name(op_b), e.bm, e.bn, e.bk);
  launch_main(handle.state, e, dim3((unsigned)tiles), p, stream);
}

void enable_kernel_timing(shgemmHandle_t& handle, const int max_launches) {
  if (handle.state == nullptr) throw error("handle not created");
  if (max_launches < 0 || max_launches > 4096) throw error("max_launches must be in 0..4096");
  auto* st = handle.state;
  std::lock_guard<std::mutex> lock(st->mu);
  if (!st->timing.empty()) (void)hipStreamSynchronize(handle.cuda_stream);  // events may still be pending
  release_timing(st);
  for (int i = 0; i < max_launches; ++i) {
    hipEvent_t a = nullptr, b = nullptr;
    if (hipEventCreate(&a) != hipSuccess || hipEventCreate(&b) != hipSuccess) {
      if (a) (void)hipEventDestroy(a);
      release_timing(st);
      throw shgemm_amd::hip_failure("timing event creation failed");
    }
    st->timing.emplace_back(a, b);
  }
}

void kernel_timing(shgemmHandle_t& handle, int& launches, float& total_ms) {
  if (handle.state == nullptr) throw error("handle not created");
  auto* st = handle.state;
  std::lock_guard<std::mutex> lock(st->mu);
  launches = st->timing_used;
  total_ms = 0.0f;
  for (int i = 0; i < st->timing_used; ++i) {
    check_hip(hipEventSynchronize(st->timing[i].second), "timing event wait");
    float ms = 0.0f;
    check_hip(hipEventElapsedTime(&ms, st->timing[i].first, st->timing[i].second), "timing event");
    total_ms += ms;
  }
  st->timing_used = 0;
}

}  // namespace shgemm
}  // namespace mtk

namespace shgemm_amd {

void set_lds_limit(const KernelEntry& e) {
  if (e.fn != nullptr && e.lds > 65536)
    check_hip(hipFuncSetAttribute(e.fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)e.lds),
              "hipFuncSetAttribute");
}

void validate_args(const mtk::shgemm::shgemmHandle_t& handle, mtk::shgemm::operation_t op_a,
                   mtk::shgemm::operation_t op_b, mtk::shgemm::operation_t op_c, size_t m, size_t n, size_t k,
                   const float* alpha, const void* a, size_t lda, const void* b, size_t ldb, const float* beta,
                   const float* c, size_t ldc) {
  using mtk::shgemm::error;
  using mtk::shgemm::op_n;
  using mtk::shgemm::op_t;
  if (handle.state == nullptr) throw error("handle not created (call mtk::shgemm::create first)");
  if ((op_a != op_n && op_a != op_t) || (op_b != op_n && op_b != op_t) || (op_c != op_n && op_c != op_t))
    throw error("invalid operation_t");
  if (m > (size_t)INT_MAX || n > (size_t)INT_MAX || k > (size_t)INT_MAX)
    throw error("m, n and k must each be below 2^31");
  if (alpha == nullptr || beta == nullptr) throw error("alpha and beta must be host pointers, not null");
  const size_t a_rows = op_a == op_n ? m : k, b_rows = op_b == op_n ? k : n, c_rows = op_c == op_n ? m : n;
  if (lda < std::max<size_t>(1, a_rows)) throw error("lda < rows of the stored A");
  if (ldb < std::max<size_t>(1, b_rows)) throw error("ldb < rows of the stored B");
  if (ldc < std::max<size_t>(1, c_rows)) throw error("ldc < rows of the stored C");
  if (m > 0 && n > 0) {
    if (c == nullptr) throw error("C is null");
    if (k > 0 && (a == nullptr || b == nullptr)) throw error("A or B is null");
  }
}

mtk::shgemm::detail::kernel_level shgemm_no_alloc(
    const mtk::shgemm::shgemmHandle_t handle, mtk::shgemm::operation_t op_a, mtk::shgemm::operation_t op_b,
    std::size_t m, std::size_t n, std::size_t k, const float* alpha, const float* a, std::size_t lda, const half* b,
    std::size_t ldb, const float* beta, float* c, std::size_t ldc, mtk::shgemm::tc_t compute_type,
    mtk::shgemm::operation_t op_c) {
  if (handle.state == nullptr) throw mtk::shgemm::error("handle not created (call mtk::shgemm::create first)");
  const CallScratch sc = static_scratch();
  return mtk::shgemm::shgemm_call(handle, op_a, op_b, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc, compute_type, op_c,
                                  &sc);
}

void register_all(KernelTable& t) {
  register_shgemm_fp16_nn(t);
  register_shgemm_fp16_nt(t);
  register_shgemm_fp16_tn(t);
  register_shgemm_fp16_tt(t);
  register_shgemm_bf16_nn(t);
  register_shgemm_bf16_nt(t);
  register_shgemm_bf16_tn(t);
  register_shgemm_bf16_tt(t);
  register_shgemm_fp16s_nn(t);
  register_shgemm_fp16s_nt(t);
  register_shgemm_fp16s_tn(t);
  register_shgemm_fp16s_tt(t);
  register_sgemm(t);
  register_tune(t);
  register_shgemm_bd(t);
  register_shgemm_ps(t);
}
}  // namespace shgemm_amd
