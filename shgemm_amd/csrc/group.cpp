// Grouped SHGEMM, host side.  A set of problems is planned (tile class and
// k-slices per problem), its descriptors and split-K partials are laid out in
// one device region, each tile class's work units are ordered longest first,
// and the launches are
//   fill (queue heads) -> one persistent launch per tile class -> grouped
//   split-K reduction.
// Two owners of that sequence:
//   * a group (create_group / run_group): fixed problems, one device
//     allocation, the sequence captured once in a HIP graph whose tile
//     classes run one after another, each on the whole GPU (or, with
//     MTK_SHGEMM_GROUP_CONCURRENT=1, as parallel branches, each on a share of
//     it: measured slower on the 512-problem batch, 9.2-9.8 against 9.0 ms);
//   * a batch (create_batch / run_batch): preallocated slots, problems that
//     change from call to call (a serving step's admitted requests), planned
//     and launched eagerly on the caller's stream with no allocation.
// Problems of a tile class whose kernel reads whole-row exponents (prescaled
// fp16 without running exponents) add row-max segments: the fill also clears
// their row maxima and a grouped row-max and the row exponents run before the
// class launches.  The shipped classes keep running exponents.
// There is no grouped entry point in the reference; each problem follows the
// single-call semantics of mtk::shgemm::shgemm (src/shgemm.cu:517-686).
#include <hip/hip_runtime.h>
#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <vector>

#include "runtime_internal.hpp"

using shgemm_amd::check_hip;
using shgemm_amd::kNumGroupClasses;

namespace mtk {
namespace shgemm {

namespace {

// The launch sequence of one planned set of problems on one device region.
struct Launches {
  uint32_t* fill = nullptr;  // queue heads, then row maxima
  int64_t fill_n = 0;
  int op_a = 0;
  shgemm_amd::RowmaxSeg* rm = nullptr;
  int* rm_begin = nullptr;
  int rm_n = 0;
  int64_t rm_blocks = 0;
  uint32_t* row_max = nullptr;
  int* row_exp = nullptr;
  int64_t rows = 0;
  struct ClassLaunch {
    const shgemm_amd::KernelEntry* e;
    unsigned grid;      // workgroups when the classes run one after another (fills the GPU)
    unsigned grid_par;  // workgroups when they run side by side (a share of the GPU by work)
    shgemm_amd::GroupLaunch args;
  };
  std::vector<ClassLaunch> gemm;
  shgemm_amd::ReduceSeg* rd = nullptr;
  int* rd_begin = nullptr;
  int rd_n = 0;
  int64_t rd_blocks = 0;

  // The sequence on `s`.  With `side` streams and events (graph capture),
  // tile class i > 0 runs on side[i - 1] between a fork after the fill and a
  // join before the reduction, so the classes share the GPU.
  void enqueue(hipStream_t s, const hipStream_t* side = nullptr, const hipEvent_t* ev = nullptr) const {
    check_hip(shgemm_amd::launch_fill_u32(fill, fill_n, 0u, s), "group: fill");
    if (rm_n > 0) {
      check_hip(shgemm_amd::launch_grouped_rowmax(rm, rm_begin, rm_n, rm_blocks, op_a, s), "group: row max");
      check_hip(shgemm_amd::launch_row_exp(row_max, rows, row_exp, s), "group: row exponents");
    }
    const bool par = side != nullptr && gemm.size() > 1;
    if (par) check_hip(hipEventRecord(ev[0], s), "group: fork");
    for (size_t i = 0; i < gemm.size(); ++i) {
      const ClassLaunch& cl = gemm[i];
      hipStream_t q = s;
      if (par && i > 0) {
        q = side[i - 1];
        check_hip(hipStreamWaitEvent(q, ev[0], 0), "group: fork wait");
      }
      void* args[] = {const_cast<shgemm_amd::GroupLaunch*>(&cl.args)};
      check_hip(hipLaunchKernel(cl.e->fn, dim3(par ? cl.grid_par : cl.grid), dim3(cl.e->threads), args, cl.e->lds, q),
                "group: grouped kernel launch");
      if (par && i > 0) check_hip(hipEventRecord(ev[i], q), "group: join");
    }
    if (par)
      for (size_t i = 1; i < gemm.size(); ++i) check_hip(hipStreamWaitEvent(s, ev[i], 0), "group: join wait");
    if (rd_n > 0) check_hip(shgemm_amd::launch_grouped_reduce(rd, rd_begin, rd_n, rd_blocks, s), "group: reduce");
  }
};

// The grouped kernels of one (op_a, op_b, op_c, mode) with their resident
// workgroups per CU (the persistent grid) and the device's CU count.
struct ClassTable {
  const shgemm_amd::KernelEntry* e[kNumGroupClasses] = {};
  int per_cu[kNumGroupClasses] = {};
  int num_cus = 1;
};

ClassTable class_table(const detail::handle_state* st, int oa, int ob, int oc, int mode) {
  ClassTable t;
  t.num_cus = std::max(1, st->num_cus);
  for (int c = 0; c < kNumGroupClasses; ++c) {
    const shgemm_amd::KernelEntry& e = st->table.grouped[oa][ob][oc][mode][c];
    if (e.fn == nullptr) continue;
    // The grouped kernel adds a small static LDS word to the dynamic tile
    // buffers, so raise the dynamic limit explicitly for every class.
    check_hip(hipFuncSetAttribute(e.fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)e.lds),
              "hipFuncSetAttribute");
    int per_cu = 0;
    check_hip(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, e.fn, e.threads, e.lds), "occupancy query");
    if (per_cu < 1) throw error("internal: grouped kernel does not fit on a CU");
    t.e[c] = &e;
    t.per_cu[c] = per_cu;
  }
  return t;
}

// Default tile class without a plan table: the largest tile the problem fills.
int default_class(const ClassTable& ct, size_t m, size_t n) {
  for (int c = kNumGroupClasses - 1; c > 0; --c)
    if (ct.e[c] != nullptr && m >= (size_t)ct.e[c]->bm && n >= (size_t)ct.e[c]->bn) return c;
  return 0;
}

struct Bump {
  size_t off = 0;
  size_t take(size_t bytes) {
    const size_t o = (off + 255) & ~size_t(255);
    off = o + bytes;
    return o;
  }
};

// Plan and layout of one set of problems (no device pointers yet).
struct GroupPlan {
  struct Planned {
    size_t idx;
    int cls, used;
    int64_t kslice, units, row_off, ws_off;  // row_off: -1 no row exponents; ws_off in floats, -1: no partials
  };
  int mode = 0, oa = 0, ob = 0, oc = 0;
  std::vector<int> cls, slices;  // per problem in input order (-1 / 0: no MFMA work)
  size_t units = 0;
  std::vector<Planned> work;
  std::vector<size_t> scale_only;  // k == 0 or alpha == 0: C = beta * C
  std::vector<std::vector<size_t>> by_class;
  int64_t class_units[kNumGroupClasses] = {};
  double class_cost[kNumGroupClasses] = {}, total_cost = 0.0;
  int64_t rows = 0, ws_floats = 0;
  std::vector<shgemm_amd::RowmaxSeg> rm;
  std::vector<int> rm_begin;
  std::vector<int64_t> rm_row_off;
  int64_t rm_blocks = 0;
  std::vector<shgemm_amd::ReduceSeg> rd;
  std::vector<int> rd_begin;
  std::vector<std::pair<size_t, int64_t>> rd_ws;  // (segment, ws offset in floats)
  int64_t rd_blocks = 0;
  // device layout
  size_t off_probs[kNumGroupClasses] = {}, off_ub[kNumGroupClasses] = {};
  size_t off_rm = 0, off_rmb = 0, off_rd = 0, off_rdb = 0, off_fill = 0, off_rexp = 0, off_ws = 0, total = 0;
  int64_t fill_n = 0;
};

GroupPlan plan_group(const detail::handle_state* st, const ClassTable& ct, int mode, int oa, int ob, int oc,
                     const group_problem* problems, size_t count, const int* tile_class, const int* slices_in) {
  GroupPlan g;
  g.mode = mode;
  g.oa = oa;
  g.ob = ob;
  g.oc = oc;
  g.cls.assign(count, -1);
  g.slices.assign(count, 0);
  const bool scaled = mode == shgemm_amd::MODE_FP16S;
  for (size_t i = 0; i < count; ++i) {
    const group_problem& q = problems[i];
    if (q.m == 0 || q.n == 0) continue;
    if (q.k == 0 || q.alpha == 0.0f) {
      g.scale_only.push_back(i);
      continue;
    }
    int c, s;
    if (tile_class != nullptr) {
      c = tile_class[i];
      s = slices_in != nullptr ? slices_in[i] : 1;
    } else if (const shgemm_amd::PlanRow* r = shgemm_amd::find_plan(st->plans, shgemm_amd::PLAN_GROUPED, mode, oa,
                                                                    ob, oc, q.m, q.n, q.k)) {
      c = r->choice;
      s = slices_in != nullptr ? slices_in[i] : r->slices;
    } else {
      c = default_class(ct, q.m, q.n);
      s = slices_in != nullptr ? slices_in[i] : 1;
    }
    if (c < 0 || c >= kNumGroupClasses || ct.e[c] == nullptr) {
      if (tile_class != nullptr)
        throw error("tile_class[" + std::to_string(i) + "] = " + std::to_string(c) +
                    " is not compiled for this compute type (class 3 exists for the prescaled fp16 mode only)");
      c = default_class(ct, q.m, q.n);  // a table row naming a class this mode lacks
    }
    const shgemm_amd::KernelEntry& e = *ct.e[c];
    const int64_t k = (int64_t)q.k;
    const int64_t max_s = (k + e.bk - 1) / e.bk;
    s = (int)std::max<int64_t>(1, std::min<int64_t>(s, max_s));
    const int64_t kslice = s > 1 ? ((k + s - 1) / s + e.bk - 1) / e.bk * e.bk : k;
    const int used = (int)((k + kslice - 1) / kslice);
    const int64_t tiles = (int64_t)((q.m + e.bm - 1) / e.bm) * (int64_t)((q.n + e.bn - 1) / e.bn);
    // prescaled kernels that read row exponents (the others keep their own)
    const bool rexp = scaled && e.needs_row_exp;
    GroupPlan::Planned pl{i, c, used, used > 1 ? kslice : k, tiles * used, rexp ? g.rows : -1, -1};
    if (used > 1) {
      pl.ws_off = g.ws_floats;
      g.ws_floats += (int64_t)used * (int64_t)q.m * (int64_t)q.n;
    }
    if (rexp) g.rows += (int64_t)q.m;
    g.cls[i] = c;
    g.slices[i] = used;
    g.work.push_back(pl);
  }

  // per class: problems longest unit first, unit numbering
  g.by_class.assign(kNumGroupClasses, {});
  for (size_t w = 0; w < g.work.size(); ++w) g.by_class[g.work[w].cls].push_back(w);
  for (int c = 0; c < kNumGroupClasses; ++c) {
    auto& v = g.by_class[c];
    std::stable_sort(v.begin(), v.end(), [&](size_t x, size_t y) { return g.work[x].kslice > g.work[y].kslice; });
    for (size_t w : v) g.class_units[c] += g.work[w].units;
    if (g.class_units[c] > INT_MAX) throw error("too many work units in one tile class");
    g.units += (size_t)g.class_units[c];
  }
  // Estimated time of each class's work (for the side-by-side grid shares):
  // per problem the larger of its MFMA time and its HBM time at rates the
  // grouped kernels sustain (about 500 TFLOP/s and 5 TB/s effective).
  for (const auto& pl : g.work) {
    const group_problem& q = problems[pl.idx];
    const double fl = 2.0 * q.m * q.n * q.k, by = 4.0 * q.m * q.k + 2.0 * q.k * q.n + 4.0 * q.m * q.n;
    const double t = std::max(fl / 5e14, by / 5e12);
    g.class_cost[pl.cls] += t;
    g.total_cost += t;
  }

  // row-max segments
  if (scaled) {
    for (const auto& pl : g.work) {
      if (pl.row_off < 0) continue;
      const group_problem& q = problems[pl.idx];
      const shgemm_amd::RowmaxGeometry geo = shgemm_amd::rowmax_geometry(oa, (int)q.m, (int)q.k);
      shgemm_amd::RowmaxSeg sg{};
      sg.a = q.A;
      sg.lda = (int64_t)q.lda;
      sg.m = (int)q.m;
      sg.k = (int)q.k;
      sg.kchunk = geo.kchunk;
      sg.rq = geo.rq;
      sg.vec = shgemm_amd::vec_a_ok(q.A, q.lda);
      sg.bx = geo.bx;
      sg.block_begin = (int)g.rm_blocks;
      g.rm_begin.push_back((int)g.rm_blocks);
      g.rm_blocks += (int64_t)geo.bx * geo.by;
      if (g.rm_blocks > INT_MAX) throw error("group row-exponent pass too large");
      g.rm.push_back(sg);
      g.rm_row_off.push_back(pl.row_off);
    }
  }

  // reduce / scale segments
  auto add_reduce = [&](size_t i, int used) {
    const group_problem& q = problems[i];
    shgemm_amd::ReduceSeg sg{};
    sg.ws = nullptr;  // set when the device base is known
    sg.c = q.C;
    sg.ldc = (int64_t)q.ldc;
    sg.elems = (int64_t)q.m * (int64_t)q.n;
    sg.slices = used;
    sg.inner = (int)(oc == shgemm_amd::OP_N ? q.m : q.n);
    sg.alpha = q.alpha;
    sg.beta = q.beta;
    sg.block_begin = (int)g.rd_blocks;
    g.rd_begin.push_back((int)g.rd_blocks);
    g.rd_blocks += (sg.elems + 255) / 256;
    if (g.rd_blocks > INT_MAX) throw error("group reduction too large");
    g.rd.push_back(sg);
    return g.rd.size() - 1;
  };
  for (const auto& pl : g.work)
    if (pl.used > 1) g.rd_ws.emplace_back(add_reduce(pl.idx, pl.used), pl.ws_off);
  for (size_t i : g.scale_only) add_reduce(i, 0);

  // device layout: descriptors first (copied from the host), then the regions
  // the kernels write
  Bump b;
  for (int c = 0; c < kNumGroupClasses; ++c) {
    g.off_probs[c] = b.take(g.by_class[c].size() * sizeof(shgemm_amd::GroupProblem));
    g.off_ub[c] = b.take(g.by_class[c].size() * sizeof(int));
  }
  g.off_rm = b.take(g.rm.size() * sizeof(shgemm_amd::RowmaxSeg));
  g.off_rmb = b.take(g.rm_begin.size() * sizeof(int));
  g.off_rd = b.take(g.rd.size() * sizeof(shgemm_amd::ReduceSeg));
  g.off_rdb = b.take(g.rd_begin.size() * sizeof(int));
  g.fill_n = 4 + g.rows;  // 4 queue heads (16 bytes), then row maxima
  g.off_fill = b.take((size_t)g.fill_n * sizeof(uint32_t));
  g.off_rexp = b.take((size_t)g.rows * sizeof(int));
  g.off_ws = b.take((size_t)g.ws_floats * sizeof(float));
  g.total = std::max<size_t>(b.off, 256);
  return g;
}

// The descriptor image of plan `g` for the device region at `base` into
// `host` (g.off_fill bytes, which are copied to base), and the launches.
Launches build_image(const GroupPlan& g, const ClassTable& ct, const group_problem* problems, char* base, char* host) {
  Launches L;
  std::memset(host, 0, g.off_fill);
  L.fill = reinterpret_cast<uint32_t*>(base + g.off_fill);
  L.fill_n = g.fill_n;
  L.row_max = L.fill + 4;
  L.row_exp = reinterpret_cast<int*>(base + g.off_rexp);
  L.rows = g.rows;
  L.op_a = g.oa;
  float* ws = reinterpret_cast<float*>(base + g.off_ws);
  for (int c = 0; c < kNumGroupClasses; ++c) {
    if (g.by_class[c].empty()) continue;
    const shgemm_amd::KernelEntry& e = *ct.e[c];
    auto* hp = reinterpret_cast<shgemm_amd::GroupProblem*>(host + g.off_probs[c]);
    auto* hub = reinterpret_cast<int*>(host + g.off_ub[c]);
    int64_t ub = 0;
    for (size_t j = 0; j < g.by_class[c].size(); ++j) {
      const GroupPlan::Planned& pl = g.work[g.by_class[c][j]];
      const group_problem& q = problems[pl.idx];
      shgemm_amd::GemmParams p{};
      p.m = (int)q.m;
      p.n = (int)q.n;
      p.k = (int)q.k;
      p.alpha = q.alpha;
      p.beta = q.beta;
      p.a = q.A;
      p.lda = (int64_t)q.lda;
      p.b = q.B;
      p.ldb = (int64_t)q.ldb;
      p.c = q.C;
      p.ldc = (int64_t)q.ldc;
      p.ws = pl.used > 1 ? ws + pl.ws_off : nullptr;
      p.row_exp = pl.row_off >= 0 ? L.row_exp + pl.row_off : nullptr;
      p.k_slice = (int)pl.kslice;
      p.tiles_m = (int)((q.m + e.bm - 1) / e.bm);
      p.tiles_n = (int)((q.n + e.bn - 1) / e.bn);
      p.group_m = shgemm_amd::group_rows(p.tiles_m);
      p.vec_a = shgemm_amd::vec_a_ok(q.A, q.lda);
      p.vec_b = shgemm_amd::vec_b_ok(q.B, q.ldb, g.ob);
      hp[j].p = p;
      hp[j].unit_begin = (int)ub;
      hub[j] = (int)ub;
      ub += pl.units;
    }
    Launches::ClassLaunch cl;
    cl.e = &e;
    const int64_t full = (int64_t)ct.per_cu[c] * ct.num_cus;
    cl.grid = (unsigned)std::min<int64_t>(g.class_units[c], full);
    // side by side: this class's share of the GPU by its estimated time
    const int64_t share = (int64_t)std::ceil((double)full * g.class_cost[c] / std::max(1e-30, g.total_cost));
    cl.grid_par = (unsigned)std::max<int64_t>(1, std::min<int64_t>(g.class_units[c],
                                                                   std::max<int64_t>(share, ct.per_cu[c])));
    cl.args.probs = reinterpret_cast<const shgemm_amd::GroupProblem*>(base + g.off_probs[c]);
    cl.args.unit_begin = reinterpret_cast<const int*>(base + g.off_ub[c]);
    cl.args.nprobs = (int)g.by_class[c].size();
    cl.args.total_units = (int)g.class_units[c];
    cl.args.head = reinterpret_cast<int*>(L.fill) + c;
    L.gemm.push_back(cl);
  }
  if (!g.rm.empty()) {
    auto* hrm = reinterpret_cast<shgemm_amd::RowmaxSeg*>(host + g.off_rm);
    std::memcpy(hrm, g.rm.data(), g.rm.size() * sizeof(shgemm_amd::RowmaxSeg));
    std::memcpy(host + g.off_rmb, g.rm_begin.data(), g.rm_begin.size() * sizeof(int));
    // each problem's row maxima start at its row offset
    for (size_t j = 0; j < g.rm.size(); ++j) hrm[j].row_max = L.row_max + g.rm_row_off[j];
    L.rm = reinterpret_cast<shgemm_amd::RowmaxSeg*>(base + g.off_rm);
    L.rm_begin = reinterpret_cast<int*>(base + g.off_rmb);
    L.rm_n = (int)g.rm.size();
    L.rm_blocks = g.rm_blocks;
  }
  if (!g.rd.empty()) {
    auto* hrd = reinterpret_cast<shgemm_amd::ReduceSeg*>(host + g.off_rd);
    std::memcpy(hrd, g.rd.data(), g.rd.size() * sizeof(shgemm_amd::ReduceSeg));
    for (auto& [seg, wo] : g.rd_ws) hrd[seg].ws = ws + wo;
    std::memcpy(host + g.off_rdb, g.rd_begin.data(), g.rd_begin.size() * sizeof(int));
    L.rd = reinterpret_cast<shgemm_amd::ReduceSeg*>(base + g.off_rd);
    L.rd_begin = reinterpret_cast<int*>(base + g.off_rdb);
    L.rd_n = (int)g.rd.size();
    L.rd_blocks = g.rd_blocks;
  }
  return L;
}

int compute_mode_of(const detail::handle_state* st, tc_t compute_type) {
  return compute_type == tf32 ? shgemm_amd::MODE_BF16
         : st->exp_scaling    ? shgemm_amd::MODE_FP16S
                              : shgemm_amd::MODE_FP16;
}

void validate_group(const shgemmHandle_t& handle, operation_t op_a, operation_t op_b, operation_t op_c,
                    const group_problem* problems, size_t count, tc_t compute_type, const int* tile_class,
                    const int* slices_in) {
  if (handle.state == nullptr) throw error("handle not created (call mtk::shgemm::create first)");
  if (compute_type != fp16 && compute_type != tf32) throw error("invalid tc_t");
  if (count == 0 || problems == nullptr) throw error("a group needs at least one problem");
  if (count > (size_t)INT_MAX / 4) throw error("too many problems in one group");
  for (size_t i = 0; i < count; ++i) {
    const group_problem& q = problems[i];
    shgemm_amd::validate_args(handle, op_a, op_b, op_c, q.m, q.n, q.k, &q.alpha, q.A, q.lda, q.B, q.ldb, &q.beta,
                              q.C, q.ldc);
    if (tile_class != nullptr && (tile_class[i] < 0 || tile_class[i] >= kNumGroupClasses))
      throw error("tile_class[" + std::to_string(i) + "] must be in 0.." + std::to_string(kNumGroupClasses - 1));
    if (slices_in != nullptr && (slices_in[i] < 1 || slices_in[i] > 256))
      throw error("slices[" + std::to_string(i) + "] must be in 1..256");
  }
}

int op_index(operation_t op) { return op == op_n ? shgemm_amd::OP_N : shgemm_amd::OP_T; }

}  // namespace

struct group_t {
  int device = 0;
  size_t count = 0;
  std::vector<int> cls, slices;
  size_t units = 0;
  char* dmem = nullptr;
  Launches launches;
  bool concurrent = false;  // the graph runs the tile classes as parallel branches (else one after another)
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;

  ~group_t() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    if (dmem) (void)hipFree(dmem);
  }
};

group_t* create_group(const shgemmHandle_t& handle, const operation_t op_a, const operation_t op_b,
                      const group_problem* problems, const std::size_t count, const tc_t compute_type,
                      const operation_t op_c, const int* tile_class, const int* slices_in) {
  validate_group(handle, op_a, op_b, op_c, problems, count, compute_type, tile_class, slices_in);
  const detail::handle_state* st = handle.state;
  const int mode = compute_mode_of(st, compute_type);
  const int oa = op_index(op_a), ob = op_index(op_b), oc = op_index(op_c);
  const ClassTable ct = class_table(st, oa, ob, oc, mode);
  const GroupPlan plan = plan_group(st, ct, mode, oa, ob, oc, problems, count, tile_class, slices_in);

  auto g = std::make_unique<group_t>();
  g->device = st->device;
  g->count = count;
  g->cls = plan.cls;
  g->slices = plan.slices;
  g->units = plan.units;
  check_hip(hipMalloc(&g->dmem, plan.total), "hipMalloc(group)");
  std::vector<char> host(plan.off_fill + 1);
  g->launches = build_image(plan, ct, problems, g->dmem, host.data());
  // descriptors only (the scratch regions are written by the kernels)
  if (plan.off_fill > 0)
    check_hip(hipMemcpy(g->dmem, host.data(), plan.off_fill, hipMemcpyHostToDevice), "hipMemcpy(group descriptors)");

  const auto& gemm = g->launches.gemm;
  if (handle.debug_mode) {
    std::printf("[group] problems=%zu units=%zu rows=%lld partials=%.1f MiB", count, g->units,
                (long long)plan.rows, plan.ws_floats * 4.0 / (1 << 20));
    for (const auto& cl : gemm)
      std::printf(" | class %dx%dx%d: %d problems, %d units, grid %u (side by side %u)", cl.e->bm, cl.e->bn,
                  cl.e->bk, cl.args.nprobs, cl.args.total_units, cl.grid, cl.grid_par);
    std::printf("\n");
    std::fflush(stdout);
  }

  // ---- capture the launch sequence in one graph ----------------------------
  // The tile classes run one after another, each kernel filling the GPU
  // (MTK_SHGEMM_GROUP_CONCURRENT=1: as parallel branches on shares of it).
  const char* cc = std::getenv("MTK_SHGEMM_GROUP_CONCURRENT");
  g->concurrent = cc != nullptr && cc[0] == '1' && gemm.size() > 1;
  hipStream_t cs = nullptr;
  hipStream_t side[kNumGroupClasses] = {};
  hipEvent_t ev[kNumGroupClasses] = {};
  auto release = [&] {
    for (auto& q : side)
      if (q) (void)hipStreamDestroy(q);
    for (auto& e : ev)
      if (e) (void)hipEventDestroy(e);
    if (cs) (void)hipStreamDestroy(cs);
  };
  try {
    check_hip(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking), "hipStreamCreate");
    if (g->concurrent) {
      for (size_t i = 0; i + 1 < gemm.size(); ++i)
        check_hip(hipStreamCreateWithFlags(&side[i], hipStreamNonBlocking), "hipStreamCreate");
      for (size_t i = 0; i < gemm.size(); ++i)
        check_hip(hipEventCreateWithFlags(&ev[i], hipEventDisableTiming), "hipEventCreate");
    }
  } catch (...) {
    release();
    throw;
  }
  hipError_t err = hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal);
  if (err != hipSuccess) {
    release();
    check_hip(err, "hipStreamBeginCapture");
  }
  try {
    if (g->concurrent) g->launches.enqueue(cs, side, ev);
    else g->launches.enqueue(cs);
  } catch (...) {
    hipGraph_t tmp = nullptr;
    (void)hipStreamEndCapture(cs, &tmp);
    if (tmp) (void)hipGraphDestroy(tmp);
    release();
    throw;
  }
  err = hipStreamEndCapture(cs, &g->graph);
  release();
  check_hip(err, "hipStreamEndCapture");
  check_hip(hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0), "hipGraphInstantiate");
  return g.release();
}

void run_group(group_t* group, hipStream_t stream, const bool use_graph) {
  if (group == nullptr) throw error("group is null");
  if (use_graph) check_hip(hipGraphLaunch(group->exec, stream), "hipGraphLaunch");
  else group->launches.enqueue(stream);
}

void destroy_group(group_t* group) {
  if (group == nullptr) return;
  // Launches of this group may still read its descriptors.
  (void)hipDeviceSynchronize();
  delete group;
}

void group_plan(const group_t* group, int* tile_class, int* slices, std::size_t* units) {
  if (group == nullptr) throw error("group is null");
  if (tile_class) std::copy(group->cls.begin(), group->cls.end(), tile_class);
  if (slices) std::copy(group->slices.begin(), group->slices.end(), slices);
  if (units) *units = group->units;
}

// ---- batches ------------------------------------------------------------------
struct batch_t {
  shgemmHandle_t handle;
  operation_t op_a = op_n, op_b = op_n, op_c = op_n;
  tc_t compute_type = fp16;
  int mode = 0;
  ClassTable ct;
  size_t max_problems = 0, bytes = 0;
  struct Slot {
    char* dev = nullptr;      // descriptors and scratch of one batch in flight
    char* staging = nullptr;  // pinned host image of the descriptors
    size_t staging_bytes = 0;
    hipEvent_t done = nullptr;  // recorded after the batch's last launch
    bool used = false;
  };
  std::vector<Slot> slots;
  size_t next = 0;
  std::vector<int> last_cls, last_slices;
  size_t last_units = 0;

  ~batch_t() {
    for (Slot& s : slots) {
      if (s.done) (void)hipEventSynchronize(s.done);
      if (s.done) (void)hipEventDestroy(s.done);
      if (s.dev) (void)hipFree(s.dev);
      if (s.staging) (void)hipHostFree(s.staging);
    }
  }
};

batch_t* create_batch(const shgemmHandle_t& handle, const operation_t op_a, const operation_t op_b,
                      const tc_t compute_type, const operation_t op_c, const std::size_t max_problems,
                      const std::size_t bytes, const int slots) {
  if (handle.state == nullptr) throw error("handle not created (call mtk::shgemm::create first)");
  if (compute_type != fp16 && compute_type != tf32) throw error("invalid tc_t");
  if ((op_a != op_n && op_a != op_t) || (op_b != op_n && op_b != op_t) || (op_c != op_n && op_c != op_t))
    throw error("invalid operation_t");
  if (max_problems == 0 || max_problems > (size_t)INT_MAX / 4) throw error("max_problems must be in 1..2^29");
  if (slots < 1 || slots > 16) throw error("slots must be in 1..16");
  if (bytes < 4096) throw error("a batch needs at least 4096 bytes of device scratch per slot");
  const detail::handle_state* st = handle.state;
  auto b = std::make_unique<batch_t>();
  b->handle = handle;
  b->op_a = op_a;
  b->op_b = op_b;
  b->op_c = op_c;
  b->compute_type = compute_type;
  b->mode = compute_mode_of(st, compute_type);
  b->ct = class_table(st, op_index(op_a), op_index(op_b), op_index(op_c), b->mode);
  b->max_problems = max_problems;
  b->bytes = bytes;
  // descriptors of max_problems problems (per problem at most one descriptor,
  // unit index, row-max and reduce segment with their indices) plus alignment
  const size_t per_problem = sizeof(shgemm_amd::GroupProblem) + sizeof(shgemm_amd::RowmaxSeg) +
                             sizeof(shgemm_amd::ReduceSeg) + 3 * sizeof(int);
  const size_t staging = max_problems * per_problem + 16 * 256;
  b->slots.resize((size_t)slots);
  for (auto& s : b->slots) {
    check_hip(hipMalloc(&s.dev, bytes), "hipMalloc(batch slot)");
    check_hip(hipHostMalloc(&s.staging, staging, hipHostMallocDefault), "hipHostMalloc(batch staging)");
    s.staging_bytes = staging;
    check_hip(hipEventCreateWithFlags(&s.done, hipEventDisableTiming), "hipEventCreate(batch slot)");
  }
  return b.release();
}

void run_batch(batch_t* b, const group_problem* problems, const std::size_t count, hipStream_t stream) {
  if (b == nullptr) throw error("batch is null");
  if (count > b->max_problems)
    throw error("batch of " + std::to_string(count) + " problems exceeds its max_problems " +
                std::to_string(b->max_problems));
  validate_group(b->handle, b->op_a, b->op_b, b->op_c, problems, count, b->compute_type, nullptr, nullptr);
  const detail::handle_state* st = b->handle.state;
  GroupPlan plan = plan_group(st, b->ct, b->mode, op_index(b->op_a), op_index(b->op_b), op_index(b->op_c), problems,
                              count, nullptr, nullptr);
  if (plan.total > b->bytes)
    throw error("this batch needs " + std::to_string(plan.total) + " bytes of device scratch, more than the " +
                std::to_string(b->bytes) + " per slot the batch was created with");
  if (plan.off_fill > b->slots[0].staging_bytes) throw error("internal: batch descriptors exceed the staging buffer");
  batch_t::Slot& s = b->slots[b->next];
  b->next = (b->next + 1) % b->slots.size();
  // the slot's previous batch must have read its descriptors and partials
  if (s.used) check_hip(hipEventSynchronize(s.done), "hipEventSynchronize(batch slot)");
  const Launches L = build_image(plan, b->ct, problems, s.dev, s.staging);
  if (plan.off_fill > 0)
    check_hip(hipMemcpyAsync(s.dev, s.staging, plan.off_fill, hipMemcpyHostToDevice, stream),
              "hipMemcpyAsync(batch descriptors)");
  L.enqueue(stream);
  check_hip(hipEventRecord(s.done, stream), "hipEventRecord(batch slot)");
  s.used = true;
  b->last_cls = plan.cls;
  b->last_slices = plan.slices;
  b->last_units = plan.units;
}

void batch_plan(const batch_t* b, int* tile_class, int* slices, std::size_t* units) {
  if (b == nullptr) throw error("batch is null");
  if (tile_class) std::copy(b->last_cls.begin(), b->last_cls.end(), tile_class);
  if (slices) std::copy(b->last_slices.begin(), b->last_slices.end(), slices);
  if (units) *units = b->last_units;
}

void destroy_batch(batch_t* b) { delete b; }

}  // namespace shgemm
}  // namespace mtk
