// Persistent grouped SHGEMM: one launch runs the output tiles (and k-slices)
// of many independent problems of one tile class.  Each workgroup takes work
// units from a queue head in device memory (one atomicAdd per unit) until the
// list is exhausted; units are numbered problem by problem in the order the
// host chose (longest units first), so the queue hands out the long tiles
// early and the short ones fill the tail.  Every unit writes its own output
// (C tile, or its slice of the problem's split-K partials), so the result does
// not depend on which workgroup ran it.
#pragma once
#include "registry.hpp"
#include "shgemm_kernel.hpp"

namespace shgemm_amd {

// Copy of a wave-uniform struct from memory into scalar registers.
template <class T>
__device__ __forceinline__ T load_uniform(const T* src) {
  static_assert(sizeof(T) % 4 == 0, "whole dwords");
  T out;
  const uint32_t* s = reinterpret_cast<const uint32_t*>(src);
  uint32_t* d = reinterpret_cast<uint32_t*>(&out);
#pragma unroll
  for (int i = 0; i < (int)(sizeof(T) / 4); ++i) d[i] = __builtin_amdgcn_readfirstlane(s[i]);
  return out;
}

template <int OPA, int OPB, int OPC, int MODE, class CFG>
__global__ void __launch_bounds__(CFG::NT, CFG::OCC) shgemm_grouped_kernel(const GroupLaunch g) {
  extern __shared__ uint4 lds[];
  __shared__ int s_unit;
  for (;;) {
    if (threadIdx.x == 0) s_unit = atomicAdd(g.head, 1);
    __syncthreads();
    const int u = __builtin_amdgcn_readfirstlane(s_unit);
    __syncthreads();  // every wave has read s_unit before it is overwritten
    if (u >= g.total_units) break;
    // last problem whose first unit is <= u
    int lo = 0, hi = g.nprobs - 1;
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (__builtin_amdgcn_readfirstlane(g.unit_begin[mid]) <= u) lo = mid;
      else hi = mid - 1;
    }
    GemmParams p = load_uniform(&g.probs[lo].p);
    p.a = as_global(p.a);
    p.b = as_global(p.b);
    p.c = as_global(p.c);
    p.ws = as_global(p.ws);
    p.row_exp = as_global(p.row_exp);
    const int tiles = p.tiles_m * p.tiles_n;
    const int v = u - __builtin_amdgcn_readfirstlane(g.unit_begin[lo]);
    const int slice = v / tiles;
    int tm, tn;
    grouped_order(p, v - slice * tiles, tm, tn);
    int kbeg = 0, kend = p.k;
    float* part = nullptr;
    if (p.ws != nullptr) {
      kbeg = slice * p.k_slice;
      kend = min(p.k, kbeg + p.k_slice);
      part = p.ws + (int64_t)slice * p.m * p.n;
    }
    // one copy of the k-loop; the epilogue picks C or the partial buffer
    shgemm_tile<OPA, OPB, OPC, MODE, CFG, 2>(p, tm, tn, kbeg, kend, part, lds);
  }
}

// Tile classes of the grouped kernel: the small tile, the 128 x 128 tile and
// the mode's large tile (the single-call P0, SK and P1 configurations, so a
// problem run in a group gives the same bits as the same plan run alone; the
// k-sliced small-tile plan differs from P0 only in its prefetch depth, which
// does not change the arithmetic).  In the prescaled fp16 mode the large
// class is the 128 x 256 tile as four 64 x 128 wavefronts (plan 11's tile:
// P1's k order and exponents, so P1's bits), two workgroups per CU: the
// 512-problem batch ran 12.8 against 13.5 ms with the eight-wavefront form.
// Class 3 (prescaled fp16 only): 128 x 512 with eight 64 x 128 wavefronts, so
// a problem with n <= 512 reads each row panel of A once (the 128 x 256 tile
// reads it twice at n = 512); one workgroup per CU (96 KiB of LDS), two
// wavefronts per SIMD.  Same k-tile depth and running row exponents as P1, so
// its problems give P1's bits.  The other modes have no room for it (the BF16
// stage is 90 KiB, the two-accumulator mode needs 256 accumulator registers).
template <int MODE>
struct GroupCfg {
  using C0 = typename ModeCfg<MODE>::P0;
  using C1 = typename ModeCfg<MODE>::SK;
  using C2 = std::conditional_t<MODE == MODE_FP16S, typename ModeCfg<MODE_FP16S>::P1W, typename ModeCfg<MODE>::P1>;
  using C3 = TileCfg<128, 512, 32, 64, 128, 1, 2>;
  static constexpr bool kHasC3 = (MODE == MODE_FP16S);
};

}  // namespace shgemm_amd
