// Grouped SHGEMM, host side: plans every problem of a group (tile class and
// k-slices), lays out the descriptors and split-K partials in one device
// allocation, orders each tile class's work units longest first, and
// captures the launch sequence in one HIP graph:
//   fill (queue heads) -> one persistent launch per tile class -> grouped
//   split-K reduction.
// Problems of a tile class whose kernel reads whole-row exponents (prescaled
// fp16 without running exponents) add row-max segments: the fill also clears
// their row maxima and a grouped row-max and the row exponents run before the
// class launches.  The shipped classes keep running exponents.
// There is no grouped entry point in the reference; each problem follows the
// single-call semantics of mtk::shgemm::shgemm (src/shgemm.cu:517-686).
#include <hip/hip_runtime.h>
#include <algorithm>
#include <climits>
#include <cstdio>
#include <cstring>
#include <memory>
#include <vector>

#include "runtime_internal.hpp"

using shgemm_amd::check_hip;

namespace mtk {
namespace shgemm {

struct group_t {
  int device = 0;
  unsigned debug = 0;
  size_t count = 0;
  std::vector<int> cls, slices;  // per problem in input order (-1 / 0: no MFMA work)
  size_t units = 0;
  char* dmem = nullptr;

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
    unsigned grid;
    shgemm_amd::GroupLaunch args;
  };
  std::vector<ClassLaunch> gemm;
  shgemm_amd::ReduceSeg* rd = nullptr;
  int* rd_begin = nullptr;
  int rd_n = 0;
  int64_t rd_blocks = 0;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;

  void enqueue(hipStream_t s) const {
    check_hip(shgemm_amd::launch_fill_u32(fill, fill_n, 0u, s), "group: fill");
    if (rm_n > 0) {
      check_hip(shgemm_amd::launch_grouped_rowmax(rm, rm_begin, rm_n, rm_blocks, op_a, s), "group: row max");
      check_hip(shgemm_amd::launch_row_exp(row_max, rows, row_exp, s), "group: row exponents");
    }
    for (const ClassLaunch& cl : gemm) {
      void* args[] = {const_cast<shgemm_amd::GroupLaunch*>(&cl.args)};
      check_hip(hipLaunchKernel(cl.e->fn, dim3(cl.grid), dim3(cl.e->threads), args, cl.e->lds, s),
                "group: grouped kernel launch");
    }
    if (rd_n > 0) check_hip(shgemm_amd::launch_grouped_reduce(rd, rd_begin, rd_n, rd_blocks, s), "group: reduce");
  }

  ~group_t() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    if (dmem) (void)hipFree(dmem);
  }
};

namespace {

// Default tile class without a plan table: the largest tile the problem fills.
int default_class(const shgemm_amd::KernelEntry (&cls)[shgemm_amd::kNumGroupClasses], size_t m, size_t n) {
  for (int c = shgemm_amd::kNumGroupClasses - 1; c > 0; --c)
    if (cls[c].fn != nullptr && m >= (size_t)cls[c].bm && n >= (size_t)cls[c].bn) return c;
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

}  // namespace

group_t* create_group(const shgemmHandle_t& handle, const operation_t op_a, const operation_t op_b,
                      const group_problem* problems, const std::size_t count, const tc_t compute_type,
                      const operation_t op_c, const int* tile_class, const int* slices_in) {
  if (handle.state == nullptr) throw error("handle not created (call mtk::shgemm::create first)");
  if (compute_type != fp16 && compute_type != tf32) throw error("invalid tc_t");
  if (count == 0 || problems == nullptr) throw error("a group needs at least one problem");
  if (count > (size_t)INT_MAX / 4) throw error("too many problems in one group");
  const detail::handle_state* st = handle.state;
  for (size_t i = 0; i < count; ++i) {
    const group_problem& q = problems[i];
    shgemm_amd::validate_args(handle, op_a, op_b, op_c, q.m, q.n, q.k, &q.alpha, q.A, q.lda, q.B, q.ldb, &q.beta,
                              q.C, q.ldc);
    if (tile_class != nullptr && (tile_class[i] < 0 || tile_class[i] >= shgemm_amd::kNumGroupClasses))
      throw error("tile_class[" + std::to_string(i) + "] must be 0, 1 or 2");
    if (slices_in != nullptr && (slices_in[i] < 1 || slices_in[i] > 256))
      throw error("slices[" + std::to_string(i) + "] must be in 1..256");
  }
  const int mode = compute_type == tf32 ? shgemm_amd::MODE_BF16
                   : st->exp_scaling    ? shgemm_amd::MODE_FP16S
                                        : shgemm_amd::MODE_FP16;
  const bool scaled = mode == shgemm_amd::MODE_FP16S;
  const int oa = op_a == op_n ? shgemm_amd::OP_N : shgemm_amd::OP_T;
  const int ob = op_b == op_n ? shgemm_amd::OP_N : shgemm_amd::OP_T;
  const int oc = op_c == op_n ? shgemm_amd::OP_N : shgemm_amd::OP_T;
  const auto& ctab = st->table.grouped[oa][ob][oc][mode];

  auto g = std::make_unique<group_t>();
  g->device = st->device;
  g->debug = handle.debug_mode;
  g->count = count;
  g->op_a = oa;
  g->cls.assign(count, -1);
  g->slices.assign(count, 0);

  // ---- plan every problem -------------------------------------------------
  struct Planned {
    size_t idx;
    int cls, used;
    int64_t kslice, units, row_off, ws_off;  // row_off: -1 no row exponents; ws_off in floats, -1: no partials
  };
  std::vector<Planned> work;        // problems with MFMA work
  std::vector<size_t> scale_only;   // k == 0 or alpha == 0: C = beta * C
  int64_t rows = 0, ws_floats = 0;
  for (size_t i = 0; i < count; ++i) {
    const group_problem& q = problems[i];
    if (q.m == 0 || q.n == 0) continue;
    if (q.k == 0 || q.alpha == 0.0f) {
      scale_only.push_back(i);
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
      c = default_class(ctab, q.m, q.n);
      s = slices_in != nullptr ? slices_in[i] : 1;
    }
    const shgemm_amd::KernelEntry& e = ctab[c];
    if (e.fn == nullptr) throw error("internal: grouped kernel class not compiled");
    const int64_t k = (int64_t)q.k;
    const int64_t max_s = (k + e.bk - 1) / e.bk;
    s = (int)std::max<int64_t>(1, std::min<int64_t>(s, max_s));
    const int64_t kslice = s > 1 ? ((k + s - 1) / s + e.bk - 1) / e.bk * e.bk : k;
    const int used = (int)((k + kslice - 1) / kslice);
    const int64_t tiles = (int64_t)((q.m + e.bm - 1) / e.bm) * (int64_t)((q.n + e.bn - 1) / e.bn);
    // prescaled kernels that read row exponents (the others keep their own)
    const bool rexp = scaled && e.needs_row_exp;
    Planned pl{i, c, used, used > 1 ? kslice : k, tiles * used, rexp ? rows : -1, -1};
    if (used > 1) {
      pl.ws_off = ws_floats;
      ws_floats += (int64_t)used * (int64_t)q.m * (int64_t)q.n;
    }
    if (rexp) rows += (int64_t)q.m;
    g->cls[i] = c;
    g->slices[i] = used;
    work.push_back(pl);
  }
  g->rows = rows;

  // ---- per class: problems longest unit first, unit numbering --------------
  std::vector<std::vector<size_t>> by_class(shgemm_amd::kNumGroupClasses);  // indices into `work`
  for (size_t w = 0; w < work.size(); ++w) by_class[work[w].cls].push_back(w);
  int64_t class_units[shgemm_amd::kNumGroupClasses] = {0, 0, 0};
  for (int c = 0; c < shgemm_amd::kNumGroupClasses; ++c) {
    auto& v = by_class[c];
    std::stable_sort(v.begin(), v.end(), [&](size_t x, size_t y) { return work[x].kslice > work[y].kslice; });
    for (size_t w : v) class_units[c] += work[w].units;
    if (class_units[c] > INT_MAX) throw error("too many work units in one tile class");
    g->units += (size_t)class_units[c];
  }

  // ---- row-max segments -----------------------------------------------------
  std::vector<shgemm_amd::RowmaxSeg> rm;
  std::vector<int> rm_begin;
  std::vector<int64_t> rm_row_off;
  int64_t rm_blocks = 0;
  if (scaled) {
    for (const Planned& pl : work) {
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
      sg.block_begin = (int)rm_blocks;
      rm_begin.push_back((int)rm_blocks);
      rm_blocks += (int64_t)geo.bx * geo.by;
      if (rm_blocks > INT_MAX) throw error("group row-exponent pass too large");
      rm.push_back(sg);
      rm_row_off.push_back(pl.row_off);
    }
  }

  // ---- reduce / scale segments ----------------------------------------------
  std::vector<shgemm_amd::ReduceSeg> rd;
  std::vector<int> rd_begin;
  int64_t rd_blocks = 0;
  auto add_reduce = [&](size_t i, int used) {
    const group_problem& q = problems[i];
    shgemm_amd::ReduceSeg sg{};
    sg.ws = nullptr;  // patched with the device base below
    sg.c = q.C;
    sg.ldc = (int64_t)q.ldc;
    sg.elems = (int64_t)q.m * (int64_t)q.n;
    sg.slices = used;
    sg.inner = (int)(op_c == op_n ? q.m : q.n);
    sg.alpha = q.alpha;
    sg.beta = q.beta;
    sg.block_begin = (int)rd_blocks;
    rd_begin.push_back((int)rd_blocks);
    rd_blocks += (sg.elems + 255) / 256;
    if (rd_blocks > INT_MAX) throw error("group reduction too large");
    rd.push_back(sg);
    return rd.size() - 1;
  };
  std::vector<std::pair<size_t, int64_t>> rd_ws;  // (segment, ws offset in floats)
  for (const Planned& pl : work)
    if (pl.used > 1) rd_ws.emplace_back(add_reduce(pl.idx, pl.used), pl.ws_off);
  for (size_t i : scale_only) add_reduce(i, 0);

  // ---- device layout ---------------------------------------------------------
  Bump b;
  size_t off_probs[shgemm_amd::kNumGroupClasses], off_ub[shgemm_amd::kNumGroupClasses];
  for (int c = 0; c < shgemm_amd::kNumGroupClasses; ++c) {
    off_probs[c] = b.take(by_class[c].size() * sizeof(shgemm_amd::GroupProblem));
    off_ub[c] = b.take(by_class[c].size() * sizeof(int));
  }
  const size_t off_rm = b.take(rm.size() * sizeof(shgemm_amd::RowmaxSeg));
  const size_t off_rmb = b.take(rm_begin.size() * sizeof(int));
  const size_t off_rd = b.take(rd.size() * sizeof(shgemm_amd::ReduceSeg));
  const size_t off_rdb = b.take(rd_begin.size() * sizeof(int));
  const int64_t fill_n = 4 + rows;  // 4 queue heads (16 bytes), then row maxima
  const size_t off_fill = b.take((size_t)fill_n * sizeof(uint32_t));
  const size_t off_rexp = b.take((size_t)rows * sizeof(int));
  const size_t off_ws = b.take((size_t)ws_floats * sizeof(float));
  const size_t total = std::max<size_t>(b.off, 256);
  check_hip(hipMalloc(&g->dmem, total), "hipMalloc(group)");
  char* base = g->dmem;
  g->fill = reinterpret_cast<uint32_t*>(base + off_fill);
  g->fill_n = fill_n;
  g->row_max = g->fill + 4;
  g->row_exp = reinterpret_cast<int*>(base + off_rexp);
  float* ws = reinterpret_cast<float*>(base + off_ws);

  // ---- host images of the descriptors ---------------------------------------
  std::vector<char> host(total, 0);
  for (int c = 0; c < shgemm_amd::kNumGroupClasses; ++c) {
    if (by_class[c].empty()) continue;
    const shgemm_amd::KernelEntry& e = ctab[c];
    auto* hp = reinterpret_cast<shgemm_amd::GroupProblem*>(host.data() + off_probs[c]);
    auto* hub = reinterpret_cast<int*>(host.data() + off_ub[c]);
    int64_t ub = 0;
    for (size_t j = 0; j < by_class[c].size(); ++j) {
      const Planned& pl = work[by_class[c][j]];
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
      p.row_exp = pl.row_off >= 0 ? g->row_exp + pl.row_off : nullptr;
      p.k_slice = (int)pl.kslice;
      p.tiles_m = (int)((q.m + e.bm - 1) / e.bm);
      p.tiles_n = (int)((q.n + e.bn - 1) / e.bn);
      p.group_m = shgemm_amd::group_rows(p.tiles_m);
      p.vec_a = shgemm_amd::vec_a_ok(q.A, q.lda);
      p.vec_b = shgemm_amd::vec_b_ok(q.B, q.ldb, ob);
      hp[j].p = p;
      hp[j].unit_begin = (int)ub;
      hub[j] = (int)ub;
      ub += pl.units;
    }
    // The grouped kernel adds a small static LDS word to the dynamic tile
    // buffers, so raise the dynamic limit explicitly for every class.
    check_hip(hipFuncSetAttribute(e.fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)e.lds),
              "hipFuncSetAttribute");
    int per_cu = 0;
    check_hip(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, e.fn, e.threads, e.lds), "occupancy query");
    if (per_cu < 1) throw error("internal: grouped kernel does not fit on a CU");
    group_t::ClassLaunch cl;
    cl.e = &e;
    cl.grid = (unsigned)std::min<int64_t>(class_units[c], (int64_t)per_cu * std::max(1, st->num_cus));
    cl.args.probs = reinterpret_cast<const shgemm_amd::GroupProblem*>(base + off_probs[c]);
    cl.args.unit_begin = reinterpret_cast<const int*>(base + off_ub[c]);
    cl.args.nprobs = (int)by_class[c].size();
    cl.args.total_units = (int)class_units[c];
    cl.args.head = reinterpret_cast<int*>(g->fill) + c;
    g->gemm.push_back(cl);
  }
  for (auto& [seg, wo] : rd_ws) rd[seg].ws = ws + wo;
  if (!rm.empty()) {
    std::memcpy(host.data() + off_rm, rm.data(), rm.size() * sizeof(shgemm_amd::RowmaxSeg));
    std::memcpy(host.data() + off_rmb, rm_begin.data(), rm_begin.size() * sizeof(int));
    // each problem's row maxima start at its row offset
    auto* hrm = reinterpret_cast<shgemm_amd::RowmaxSeg*>(host.data() + off_rm);
    for (size_t j = 0; j < rm.size(); ++j) hrm[j].row_max = g->row_max + rm_row_off[j];
    g->rm = reinterpret_cast<shgemm_amd::RowmaxSeg*>(base + off_rm);
    g->rm_begin = reinterpret_cast<int*>(base + off_rmb);
    g->rm_n = (int)rm.size();
    g->rm_blocks = rm_blocks;
  }
  if (!rd.empty()) {
    std::memcpy(host.data() + off_rd, rd.data(), rd.size() * sizeof(shgemm_amd::ReduceSeg));
    std::memcpy(host.data() + off_rdb, rd_begin.data(), rd_begin.size() * sizeof(int));
    g->rd = reinterpret_cast<shgemm_amd::ReduceSeg*>(base + off_rd);
    g->rd_begin = reinterpret_cast<int*>(base + off_rdb);
    g->rd_n = (int)rd.size();
    g->rd_blocks = rd_blocks;
  }
  // descriptors only (the scratch regions are written by the kernels)
  check_hip(hipMemcpy(base, host.data(), off_fill, hipMemcpyHostToDevice), "hipMemcpy(group descriptors)");

  if (g->debug) {
    std::printf("[group] problems=%zu units=%zu rows=%lld partials=%.1f MiB", count, g->units, (long long)rows,
                ws_floats * 4.0 / (1 << 20));
    for (const auto& cl : g->gemm)
      std::printf(" | class %dx%dx%d: %d problems, %d units, grid %u", cl.e->bm, cl.e->bn, cl.e->bk, cl.args.nprobs,
                  cl.args.total_units, cl.grid);
    std::printf("\n");
  }

  // ---- capture the launch sequence in one graph ----------------------------
  hipStream_t cs = nullptr;
  check_hip(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking), "hipStreamCreate");
  hipError_t err = hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal);
  if (err != hipSuccess) {
    (void)hipStreamDestroy(cs);
    check_hip(err, "hipStreamBeginCapture");
  }
  try {
    g->enqueue(cs);
  } catch (...) {
    hipGraph_t tmp = nullptr;
    (void)hipStreamEndCapture(cs, &tmp);
    if (tmp) (void)hipGraphDestroy(tmp);
    (void)hipStreamDestroy(cs);
    throw;
  }
  err = hipStreamEndCapture(cs, &g->graph);
  (void)hipStreamDestroy(cs);
  check_hip(err, "hipStreamEndCapture");
  check_hip(hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0), "hipGraphInstantiate");
  return g.release();
}

void run_group(group_t* group, hipStream_t stream, const bool use_graph) {
  if (group == nullptr) throw error("group is null");
  if (use_graph) check_hip(hipGraphLaunch(group->exec, stream), "hipGraphLaunch");
  else group->enqueue(stream);
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

}  // namespace shgemm
}  // namespace mtk
