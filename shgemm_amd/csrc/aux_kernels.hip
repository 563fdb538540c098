// Auxiliary kernels: deterministic split-K reduction, beta-only update, the
// per-row exponent pass of the prescaled fp16 mode (single call: k-block
// partial maxima + a reduction; grouped: one pass with atomic maxima) and
// their grouped forms.
#include <algorithm>
#include "registry.hpp"

namespace shgemm_amd {

// ---------------------------------------------------------------------------
// Split-K reduction: C = alpha * sum_{s=0}^{S-1} ws[s] + beta * C, the slices
// summed in index order so the result does not depend on scheduling.
// ws element e is C element (e % inner, e / inner) in C's own storage order
// (inner = m for column-major C, n for transposed C).
// ---------------------------------------------------------------------------
// Sum of p[0], p[stride], ..., p[(slices-1)*stride], added one by one in
// index order.  The loads go out 16 at a time, the next 16 before the
// previous ones are added: with few elements and many slices (128 x 128 C
// over 128 slices) each thread's chain of loads is what bounds the pass.
__device__ __forceinline__ float sum_slices(const float* __restrict__ p, int64_t stride, int slices) {
  constexpr int D = 16;
  const int full = slices / D * D;
  float s = 0.0f;
  if (full > 0) {
    float v[D];
#pragma unroll
    for (int j = 0; j < D; ++j) v[j] = p[(int64_t)j * stride];
    for (int i = D; i < full; i += D) {
      float w[D];
#pragma unroll
      for (int j = 0; j < D; ++j) w[j] = p[(int64_t)(i + j) * stride];
#pragma unroll
      for (int j = 0; j < D; ++j) s += v[j];
#pragma unroll
      for (int j = 0; j < D; ++j) v[j] = w[j];
    }
#pragma unroll
    for (int j = 0; j < D; ++j) s += v[j];
  }
  for (int i = full; i < slices; ++i) s += p[(int64_t)i * stride];
  return s;
}

__device__ __forceinline__ void splitk_reduce_elem(const float* __restrict__ ws, int slices, int64_t elems, int inner,
                                                   float alpha, float beta, float* __restrict__ c, int64_t ldc,
                                                   int64_t e) {
  if (e >= elems) return;
  const float s = sum_slices(ws + e, elems, slices);
  const int64_t in = e % inner, out = e / inner;
  float* cp = c + in + out * ldc;
  *cp = (beta == 0.0f) ? alpha * s : alpha * s + beta * *cp;
}

__global__ void __launch_bounds__(256) splitk_reduce_kernel(const float* __restrict__ ws, int slices, int64_t elems,
                                                             int inner, float alpha, float beta,
                                                             float* __restrict__ c, int64_t ldc) {
  splitk_reduce_elem(ws, slices, elems, inner, alpha, beta, c, ldc, (int64_t)blockIdx.x * 256 + threadIdx.x);
}

// beta-only update (k == 0): C = beta * C, or zero when beta == 0.
__device__ __forceinline__ void scale_c_elem(int64_t elems, int inner, float beta, float* __restrict__ c, int64_t ldc,
                                             int64_t e) {
  if (e >= elems) return;
  float* cp = c + (e % inner) + (e / inner) * ldc;
  *cp = (beta == 0.0f) ? 0.0f : beta * *cp;
}

__global__ void __launch_bounds__(256) scale_c_kernel(int64_t elems, int inner, float beta, float* __restrict__ c,
                                                       int64_t ldc) {
  scale_c_elem(elems, inner, beta, c, ldc, (int64_t)blockIdx.x * 256 + threadIdx.x);
}

// ---------------------------------------------------------------------------
// Per-row exponent of A for the prescaled fp16 mode.
//   row_max: |a| bit patterns, max-reduced with atomicMax (order-independent).
//   row_exp[r] = ilogb(max |A[r,:]|) - 14, so A[r,:] * 2^-row_exp lies below
//   2^15 (no FP16 overflow in hi) and its largest entries sit in [2^14, 2^15).
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) fill_u32_kernel(uint32_t* __restrict__ p, int64_t n, uint32_t v) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i < n) p[i] = v;
}

__device__ __forceinline__ uint32_t abs_bits(float x) { return __builtin_bit_cast(uint32_t, x) & 0x7fffffffu; }
__device__ __forceinline__ uint32_t umax(uint32_t a, uint32_t b) { return a > b ? a : b; }

// A stored with rows contiguous (op_n): element (r, k) at a[r + k*lda].
// A workgroup covers `rq` row-quads (4 consecutive rows each) and 256/rq
// k-lanes, so narrow matrices (small m) still keep every thread loading.
// With `vec` a row-quad at one k is one 16-byte load (rows 4-aligned,
// lda % 4 == 0, 16-byte base); eight such loads are issued before they are
// reduced, so a thread keeps 128 bytes in flight.  grid.x covers row-quads in
// groups of rq, grid.y covers k in chunks of kchunk.  ATOMIC: the maxima go
// to dst[r] with atomicMax (order-independent); otherwise block y writes its
// partial maxima to dst[r] (dst already offset to its slice of the partials).
template <bool ATOMIC>
__device__ __forceinline__ void rowmax_rcontig_block(const float* __restrict__ a, int64_t lda, int m, int k,
                                                     int kchunk, int rq, int vec, uint32_t* __restrict__ dst,
                                                     int bx, int by) {
  const int q = threadIdx.x % rq, kl = threadIdx.x / rq, klanes = 256 / rq;
  // rows of this thread: a quad r4..r4+3 for 16-byte loads (vec), otherwise
  // rows q, q + rq, q + 2rq, q + 3rq of the block, so that each 4-byte load
  // instruction reads consecutive rows
  const int r4 = (bx * rq + q) * 4;
  const int rb = bx * rq * 4 + q;
  auto row = [&](int e) { return vec ? r4 + e : rb + e * rq; };
  uint32_t mx[4] = {0u, 0u, 0u, 0u};
  if (kl < klanes) {
    const int k0 = by * kchunk, k1 = min(k, k0 + kchunk);
    constexpr int U = 8;
    if (vec && r4 + 4 <= m) {
      for (int kk = k0 + kl; kk < k1; kk += U * klanes) {
        float4 v[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
          const int kx = kk + u * klanes;
          v[u] = kx < k1 ? *reinterpret_cast<const float4*>(a + r4 + (int64_t)kx * lda) : make_float4(0, 0, 0, 0);
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
          mx[0] = umax(mx[0], abs_bits(v[u].x));
          mx[1] = umax(mx[1], abs_bits(v[u].y));
          mx[2] = umax(mx[2], abs_bits(v[u].z));
          mx[3] = umax(mx[3], abs_bits(v[u].w));
        }
      }
    } else {
      // 4-byte loads, 8 columns in flight
      for (int kk = k0 + kl; kk < k1; kk += U * klanes) {
        float v[U][4];
#pragma unroll
        for (int u = 0; u < U; ++u) {
          const int kx = kk + u * klanes;
#pragma unroll
          for (int e = 0; e < 4; ++e) v[u][e] = (kx < k1 && row(e) < m) ? a[row(e) + (int64_t)kx * lda] : 0.0f;
        }
#pragma unroll
        for (int u = 0; u < U; ++u)
#pragma unroll
          for (int e = 0; e < 4; ++e) mx[e] = umax(mx[e], abs_bits(v[u][e]));
      }
    }
  }
  // combine the k-lanes of each thread's rows in LDS, then one store (or
  // atomic) per row
  __shared__ uint32_t red[4][256];
  for (int e = 0; e < 4; ++e) red[e][threadIdx.x] = mx[e];
  __syncthreads();
  if (kl == 0) {
    for (int e = 0; e < 4; ++e) {
      uint32_t v = 0;
      for (int j = 0; j < klanes; ++j) v = umax(v, red[e][q + j * rq]);
      const int r = row(e);
      if (r < m) {
        if constexpr (ATOMIC) atomicMax(&dst[r], v);
        else dst[r] = v;
      }
    }
  }
}


__global__ void __launch_bounds__(256) rowmax_rcontig_kernel(const float* __restrict__ a, int64_t lda, int m, int k,
                                                              int kchunk, int rq, int vec,
                                                              uint32_t* __restrict__ partial) {
  rowmax_rcontig_block<false>(a, lda, m, k, kchunk, rq, vec, partial + (int64_t)blockIdx.y * m, blockIdx.x,
                              blockIdx.y);
}

// A stored with k contiguous (op_t): element (r, k) at a[r*lda + k].
// One workgroup per (row, k chunk): 16-byte loads when `vec`, a block
// reduction, then one atomic.
template <bool ATOMIC>
__device__ __forceinline__ void rowmax_kcontig_block(const float* __restrict__ a, int64_t lda, int k, int kchunk,
                                                     int vec, uint32_t* __restrict__ dst, int r, int by) {
  const int k0 = by * kchunk, k1 = min(k, k0 + kchunk);
  const float* row = a + (int64_t)r * lda;
  uint32_t mx = 0;
  if (vec) {
    const int v0 = (k0 + 3) / 4 * 4, v1 = k1 / 4 * 4;  // 16-byte aligned middle
    for (int kk = k0 + threadIdx.x; kk < min(v0, k1); kk += 256) mx = umax(mx, abs_bits(row[kk]));
#pragma unroll 4
    for (int kk = v0 + 4 * threadIdx.x; kk < v1; kk += 1024) {
      const float4 v = *reinterpret_cast<const float4*>(row + kk);
      mx = umax(umax(mx, abs_bits(v.x)), umax(abs_bits(v.y), umax(abs_bits(v.z), abs_bits(v.w))));
    }
    for (int kk = max(v1, v0) + threadIdx.x; kk < k1; kk += 256) mx = umax(mx, abs_bits(row[kk]));
  } else {
#pragma unroll 8
    for (int kk = k0 + threadIdx.x; kk < k1; kk += 256) mx = umax(mx, abs_bits(row[kk]));
  }
  __shared__ uint32_t red[256];
  red[threadIdx.x] = mx;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s) red[threadIdx.x] = umax(red[threadIdx.x], red[threadIdx.x + s]);
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    if constexpr (ATOMIC) atomicMax(&dst[r], red[0]);
    else dst[r] = red[0];
  }
}

__global__ void __launch_bounds__(256) rowmax_kcontig_kernel(const float* __restrict__ a, int64_t lda, int m, int k,
                                                              int kchunk, int vec, uint32_t* __restrict__ partial) {
  rowmax_kcontig_block<false>(a, lda, k, kchunk, vec, partial + (int64_t)blockIdx.y * m, blockIdx.x, blockIdx.y);
}


__global__ void __launch_bounds__(256) rowexp_kernel(const uint32_t* __restrict__ row_max, int64_t m,
                                                      int* __restrict__ row_exp) {
  const int64_t r = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (r < m) row_exp[r] = exponent_of_max(row_max[r]);
}

// Reduction of kb partial maxima per row (in[b*m + r]) in groups of 256: a
// block covers 32 rows with 8 lanes over one group of partials and combines
// them in LDS.  With one group (gridDim.y == 1) it writes the row exponents;
// otherwise it writes one maximum per group to out[blockIdx.y*m + r] for the
// next level.
__global__ void __launch_bounds__(256) rowexp_reduce_kernel(const uint32_t* __restrict__ in, int kb, int m,
                                                             uint32_t* __restrict__ out, int* __restrict__ row_exp) {
  const int rl = threadIdx.x % 32, kl = threadIdx.x / 32;
  const int r = blockIdx.x * 32 + rl;
  const int b0 = blockIdx.y * 256, b1 = min(kb, b0 + 256);
  uint32_t mx = 0;
  if (r < m) {
#pragma unroll 8
    for (int b = b0 + kl; b < b1; b += 8) mx = umax(mx, in[(int64_t)b * m + r]);
  }
  __shared__ uint32_t red[8][32];
  red[kl][rl] = mx;
  __syncthreads();
  if (kl == 0 && r < m) {
    for (int j = 1; j < 8; ++j) mx = umax(mx, red[j][rl]);
    if (gridDim.y == 1) row_exp[r] = exponent_of_max(mx);
    else out[(int64_t)blockIdx.y * m + r] = mx;
  }
}


hipError_t launch_splitk_reduce(const float* ws, int slices, int64_t elems, int inner, float alpha, float beta,
                                float* c, int64_t ldc, hipStream_t s) {
  const int64_t blocks = (elems + 255) / 256;
  hipLaunchKernelGGL(splitk_reduce_kernel, dim3((unsigned)blocks), dim3(256), 0, s, ws, slices, elems, inner, alpha,
                     beta, c, ldc);
  return hipGetLastError();
}

hipError_t launch_scale_c(int64_t elems, int inner, float beta, float* c, int64_t ldc, hipStream_t s) {
  const int64_t blocks = (elems + 255) / 256;
  hipLaunchKernelGGL(scale_c_kernel, dim3((unsigned)blocks), dim3(256), 0, s, elems, inner, beta, c, ldc);
  return hipGetLastError();
}

RowmaxGeometry rowmax_geometry(int op_a, int m, int k) {
  RowmaxGeometry g{};
  if (op_a == OP_N) {
    // rq row-quads per workgroup (a power of two up to 256), the rest of the
    // 256 threads spread over k; about 64 loads per thread; grid.y <= 65535
    const int quads = (m + 3) / 4;
    g.rq = 1;
    while (g.rq < 256 && g.rq < quads) g.rq *= 2;
    const int klanes = 256 / g.rq;
    g.kchunk = std::max(64 * klanes, (int)((k + 65534) / 65535));
    g.bx = (quads + g.rq - 1) / g.rq;
  } else {
    g.rq = 0;
    g.kchunk = std::max(4096, (int)((k + 65534) / 65535));
    g.bx = m;
  }
  g.by = k > 0 ? (k + g.kchunk - 1) / g.kchunk : 0;
  return g;
}

// Single-call geometry: about 2048 workgroups (enough to spread over every
// CU), at most 4096 k-blocks (each writes one partial maximum per row), and
// the partials plus one reduction level inside `partial_cap` words.  A
// k-block count of 1 writes the maxima directly.
static RowmaxGeometry rowmax_partial_geometry(int op_a, int m, int k, int64_t partial_cap) {
  RowmaxGeometry g{};
  int64_t rb;
  int kb_min_chunk;
  if (op_a == OP_N) {
    const int quads = (m + 3) / 4;
    g.rq = 1;
    while (g.rq < 256 && g.rq < quads) g.rq *= 2;
    g.bx = (quads + g.rq - 1) / g.rq;
    rb = g.bx;
    // one load per k-lane at least, and 8192 elements per workgroup so
    // small matrices need few partials (one reduction level)
    kb_min_chunk = std::max(256 / g.rq, 8192 / (4 * g.rq));
  } else {
    g.rq = 0;
    g.bx = m;
    rb = m;
    kb_min_chunk = 4096;  // 256 threads x 16 values
  }
  int64_t kb = std::max<int64_t>(1, std::min<int64_t>(4096, (2048 + rb - 1) / rb));
  kb = std::min<int64_t>(kb, std::max<int64_t>(1, (partial_cap / std::max(1, m)) * 256 / 257));
  kb = std::min<int64_t>(kb, std::max<int64_t>(1, (k + kb_min_chunk - 1) / kb_min_chunk));
  g.kchunk = (int)((k + kb - 1) / kb);
  g.by = (k + g.kchunk - 1) / g.kchunk;
  return g;
}

hipError_t launch_row_exponents(const float* a, int64_t lda, int op_a, int m, int k, uint32_t* row_max,
                                int* row_exp, uint32_t* partials, size_t partial_bytes, hipStream_t s) {
  if (m <= 0) return hipSuccess;
  if (k <= 0) {
    hipError_t err = launch_fill_u32(row_max, m, 0u, s);
    return err != hipSuccess ? err : launch_row_exp(row_max, m, row_exp, s);
  }
  const bool aligned16 = (reinterpret_cast<uintptr_t>(a) % 16) == 0 && (lda % 4) == 0;
  const RowmaxGeometry g = rowmax_partial_geometry(op_a, m, k, (int64_t)(partial_bytes / sizeof(uint32_t)));
  uint32_t* dst = g.by == 1 ? row_max : partials;
  if (op_a == OP_N)
    hipLaunchKernelGGL(rowmax_rcontig_kernel, dim3((unsigned)g.bx, (unsigned)g.by), dim3(256), 0, s, a, lda, m, k,
                       g.kchunk, g.rq, aligned16 ? 1 : 0, dst);
  else
    hipLaunchKernelGGL(rowmax_kcontig_kernel, dim3((unsigned)g.bx, (unsigned)g.by), dim3(256), 0, s, a, lda, m, k,
                       g.kchunk, aligned16 ? 1 : 0, dst);
  hipError_t err = hipGetLastError();
  if (err != hipSuccess) return err;
  if (g.by == 1) return launch_row_exp(row_max, m, row_exp, s);
  // levels of 256-way reduction until one group remains; level outputs follow
  // the partials in the same scratch
  const unsigned rblocks = (unsigned)((m + 31) / 32);
  uint32_t* in = partials;
  int kb = g.by;
  while (true) {
    const int groups = (kb + 255) / 256;
    uint32_t* out = in + (int64_t)kb * m;
    hipLaunchKernelGGL(rowexp_reduce_kernel, dim3(rblocks, (unsigned)groups), dim3(256), 0, s, in, kb, m, out,
                       row_exp);
    err = hipGetLastError();
    if (err != hipSuccess || groups == 1) return err;
    in = out;
    kb = groups;
  }
}

hipError_t launch_fill_u32(uint32_t* p, int64_t n, uint32_t v, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(fill_u32_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, p, n, v);
  return hipGetLastError();
}

hipError_t launch_row_exp(const uint32_t* row_max, int64_t rows, int* row_exp, hipStream_t s) {
  if (rows <= 0) return hipSuccess;
  hipLaunchKernelGGL(rowexp_kernel, dim3((unsigned)((rows + 255) / 256)), dim3(256), 0, s, row_max, rows, row_exp);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Grouped forms: one launch over the blocks of every problem's segment.
// Block b belongs to the last segment whose first block is <= b.
// ---------------------------------------------------------------------------
__device__ __forceinline__ int find_segment(const int* __restrict__ begin, int n, int b) {
  int lo = 0, hi = n - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (begin[mid] <= b) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

template <int OPA>
__global__ void __launch_bounds__(256) grouped_rowmax_kernel(const RowmaxSeg* __restrict__ segs,
                                                              const int* __restrict__ begin, int nseg) {
  const int b = blockIdx.x;
  const int i = find_segment(begin, nseg, b);
  const RowmaxSeg sg = segs[i];
  const int local = b - sg.block_begin;
  const int bx = local % sg.bx, by = local / sg.bx;
  if constexpr (OPA == OP_N)
    rowmax_rcontig_block<true>(as_global(sg.a), sg.lda, sg.m, sg.k, sg.kchunk, sg.rq, sg.vec, as_global(sg.row_max),
                               bx, by);
  else rowmax_kcontig_block<true>(as_global(sg.a), sg.lda, sg.k, sg.kchunk, sg.vec, as_global(sg.row_max), bx, by);
}

__global__ void __launch_bounds__(256) grouped_reduce_kernel(const ReduceSeg* __restrict__ segs,
                                                              const int* __restrict__ begin, int nseg) {
  const int b = blockIdx.x;
  const int i = find_segment(begin, nseg, b);
  const ReduceSeg sg = segs[i];
  const int64_t e = (int64_t)(b - sg.block_begin) * 256 + threadIdx.x;
  float* c = as_global(sg.c);  // descriptor pointers: global, not flat, accesses
  if (sg.slices > 0) splitk_reduce_elem(as_global(sg.ws), sg.slices, sg.elems, sg.inner, sg.alpha, sg.beta, c, sg.ldc, e);
  else scale_c_elem(sg.elems, sg.inner, sg.beta, c, sg.ldc, e);  // k == 0 or alpha == 0
}

hipError_t launch_grouped_rowmax(const RowmaxSeg* segs, const int* begin, int nseg, int64_t blocks, int op_a,
                                 hipStream_t s) {
  if (nseg <= 0 || blocks <= 0) return hipSuccess;
  if (op_a == OP_N) hipLaunchKernelGGL(grouped_rowmax_kernel<OP_N>, dim3((unsigned)blocks), dim3(256), 0, s, segs, begin, nseg);
  else hipLaunchKernelGGL(grouped_rowmax_kernel<OP_T>, dim3((unsigned)blocks), dim3(256), 0, s, segs, begin, nseg);
  return hipGetLastError();
}

hipError_t launch_grouped_reduce(const ReduceSeg* segs, const int* begin, int nseg, int64_t blocks, hipStream_t s) {
  if (nseg <= 0 || blocks <= 0) return hipSuccess;
  hipLaunchKernelGGL(grouped_reduce_kernel, dim3((unsigned)blocks), dim3(256), 0, s, segs, begin, nseg);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Pre-split of A for the pre-split plan: the prescaled FP16 hi / residual lo
// of every element (the same values the GEMM kernels compute per tile),
// written once per call in the GEMM's tile order.  A is cut into 256-row
// blocks and 32-deep k-tiles; tile (bm, kt) is 32 KiB at
// out + (bm * (k / 32) + kt) * 2048 slots of 16 bytes: 1024 hi slots then
// 1024 lo slots, slot lds_slot_h<4>(r, c) holding row r's k-values
// 8c..8c+7 (the LDS image of the tile), rows past m zero.  A GEMM workgroup
// then streams one contiguous 32 KiB block per k-tile.
// Caller: k % 32 == 0.  One workgroup of 256 threads per tile.
// BF16 (the tf32 compute type): the three BF16 parts a0, a1, a2 of each
// element (split_bf16x3, as the register-split kernels make them), no row
// exponents; a tile is 48 KiB, the a0, a1 then a2 images.  MODE_FP16
// (prescaling off): FP16 hi and the 2^11-scaled residual (split_f16), no
// row exponents.
// ---------------------------------------------------------------------------
constexpr int kPackRows = 256, kPackK = 32, kPackSlots = kPackRows * kPackK / 8;

template <bool KCONT, int MODE>
__global__ void __launch_bounds__(256) presplit_pack_kernel(const float* __restrict__ a, int64_t lda, int m, int k,
                                                            int vec, const int* __restrict__ row_exp,
                                                            uint4* __restrict__ out) {
  constexpr bool BF16 = MODE == MODE_BF16;
  constexpr int PLANES = BF16 ? 3 : 2;
  __shared__ float t[KCONT ? 1 : kPackK][kPackRows + 1];
  const int mb = (m + kPackRows - 1) / kPackRows, kt_n = k / kPackK;
  const int bm = (int)(blockIdx.x % mb), kt = (int)(blockIdx.x / mb);
  const int r0 = bm * kPackRows, k0 = kt * kPackK, tid = threadIdx.x;
  if constexpr (!KCONT) {
    // m-contiguous A: lane group g = tid / 64 takes columns g, g + 4, ...;
    // each lane 4 consecutive rows (one 16-byte load where aligned and
    // inside m), so a wavefront instruction reads one whole 1 KiB column
    const int rr = 4 * (tid & 63), g = tid >> 6;
    const bool full = vec && (r0 + rr + 4 <= m);
#pragma unroll
    for (int j = 0; j < kPackK / 4; ++j) {
      const int col = g + 4 * j;
      const float* src = a + (int64_t)(r0 + rr) + (int64_t)(k0 + col) * lda;
      float x[4];
      if (full) {
        const float4 f = *reinterpret_cast<const float4*>(src);
        x[0] = f.x; x[1] = f.y; x[2] = f.z; x[3] = f.w;
      } else {
#pragma unroll
        for (int e = 0; e < 4; ++e) x[e] = r0 + rr + e < m ? src[e] : 0.0f;
      }
#pragma unroll
      for (int e = 0; e < 4; ++e) t[col][rr + e] = x[e];
    }
    __syncthreads();
  }
  uint4* tile = out + ((int64_t)bm * kt_n + kt) * (PLANES * kPackSlots);
#pragma unroll
  for (int q = 0; q < kPackSlots / 256; ++q) {
    const int s = tid + 256 * q;
    const int r = s / 4, c = (s % 4) ^ lds_key_h<4>(r);  // inverse of lds_slot_h<4>
    float v[8];
    if constexpr (KCONT) {
      const bool in = r0 + r < m;
      const float* src = a + (int64_t)(r0 + r) * lda + k0 + 8 * c;
#pragma unroll
      for (int e = 0; e < 8; ++e) v[e] = in ? src[e] : 0.0f;
    } else {
#pragma unroll
      for (int e = 0; e < 8; ++e) v[e] = t[8 * c + e][r];
    }
    if constexpr (BF16) {
      bf16x8 x0, x1, x2;
#pragma unroll
      for (int e = 0; e < 8; ++e) {
        __bf16 a0, a1, a2;
        split_bf16x3(v[e], a0, a1, a2);
        x0[e] = a0;
        x1[e] = a1;
        x2[e] = a2;
      }
      tile[s] = __builtin_bit_cast(uint4, x0);
      tile[kPackSlots + s] = __builtin_bit_cast(uint4, x1);
      tile[2 * kPackSlots + s] = __builtin_bit_cast(uint4, x2);
    } else if constexpr (MODE == MODE_FP16) {
      f16x8 h, l;
#pragma unroll
      for (int e = 0; e < 8; ++e) {
        _Float16 hh, ll;
        split_f16(v[e], hh, ll);
        h[e] = hh;
        l[e] = ll;
      }
      tile[s] = __builtin_bit_cast(uint4, h);
      tile[kPackSlots + s] = __builtin_bit_cast(uint4, l);
    } else {
      const float sc = r0 + r < m ? __builtin_ldexpf(1.0f, -row_exp[r0 + r]) : 1.0f;
      uint4 h, l;
      split2_f16_scaled(v[0], v[1], sc, h.x, l.x);
      split2_f16_scaled(v[2], v[3], sc, h.y, l.y);
      split2_f16_scaled(v[4], v[5], sc, h.z, l.z);
      split2_f16_scaled(v[6], v[7], sc, h.w, l.w);
      tile[s] = h;
      tile[kPackSlots + s] = l;
    }
  }
}

// B of the pre-split plan in the same tile order: per 256-column block of n
// and 32-deep k-tile, 16 KiB holding the tile's LDS image (slot
// lds_slot_h<4>(j, c) = column j's k-values 8c..8c+7, FP16 bits unchanged),
// columns past n zero.  B is k x n with op_b applied: op_n stores element
// (kk, j) at b[kk + j * ldb] (k contiguous), op_t at b[j + kk * ldb].  Loads
// are 16 bytes at any ldb and base alignment.  BF16 (the tf32 compute type):
// the two exact BF16 parts b0, b1 of each value (split_bf16x2), 32 KiB per
// tile, the b0 then b1 images.
template <bool KCONT, bool BF16>
__global__ void __launch_bounds__(256) pack_b_kernel(const uint16_t* __restrict__ b, int64_t ldb, int n, int k,
                                                     uint4* __restrict__ out) {
  constexpr int PLANES = BF16 ? 2 : 1;
  __shared__ uint16_t t[KCONT ? 1 : kPackK][kPackRows + 8];
  const int nb = (n + kPackRows - 1) / kPackRows, kt_n = k / kPackK;
  const int bn = (int)(blockIdx.x % nb), kt = (int)(blockIdx.x / nb);
  const int j0 = bn * kPackRows, k0 = kt * kPackK, tid = threadIdx.x;
  if constexpr (!KCONT) {
    // n-contiguous B: each thread loads 8 consecutive columns of one k-row
#pragma unroll
    for (int q = 0; q < kPackSlots / 256; ++q) {
      const int i = tid + 256 * q;
      const int kk = i / (kPackRows / 8), jj = 8 * (i % (kPackRows / 8));
      const uint16_t* src = b + (int64_t)(j0 + jj) + (int64_t)(k0 + kk) * ldb;
      if (j0 + jj + 8 <= n) {
        const uint4 w = load_unaligned<uint4>(src);
        *reinterpret_cast<uint4*>(&t[kk][jj]) = w;
      } else {
#pragma unroll
        for (int e = 0; e < 8; ++e) t[kk][jj + e] = j0 + jj + e < n ? src[e] : (uint16_t)0;
      }
    }
    __syncthreads();
  }
  uint4* tile = out + ((int64_t)bn * kt_n + kt) * (PLANES * kPackSlots);
#pragma unroll
  for (int q = 0; q < kPackSlots / 256; ++q) {
    const int s = tid + 256 * q;
    const int j = s / 4, c = (s % 4) ^ lds_key_h<4>(j);  // inverse of lds_slot_h<4>
    uint4 w;
    if constexpr (KCONT) {
      const uint16_t* src = b + (int64_t)(j0 + j) * ldb + k0 + 8 * c;
      w = j0 + j >= n ? uint4{0u, 0u, 0u, 0u} : load_unaligned<uint4>(src);
    } else {
      uint32_t h[4];
#pragma unroll
      for (int e = 0; e < 4; ++e) h[e] = (uint32_t)t[8 * c + 2 * e][j] | ((uint32_t)t[8 * c + 2 * e + 1][j] << 16);
      w = uint4{h[0], h[1], h[2], h[3]};
    }
    if constexpr (BF16) {
      bf16x8 x0, x1;
#pragma unroll
      for (int e = 0; e < 8; ++e) {
        __bf16 b0, b1;
        split_bf16x2(bits_to_f16((uint16_t)((w[e >> 1] >> (16 * (e & 1))) & 0xffffu)), b0, b1);
        x0[e] = b0;
        x1[e] = b1;
      }
      tile[s] = __builtin_bit_cast(uint4, x0);
      tile[kPackSlots + s] = __builtin_bit_cast(uint4, x1);
    } else {
      tile[s] = w;
    }
  }
}

size_t pack_b_bytes(int n, int k, int mode) {
  return (size_t)((n + kPackRows - 1) / kPackRows) * kPackRows * (size_t)k * (mode == MODE_BF16 ? 4 : 2);
}

hipError_t launch_pack_b(const uint16_t* b, int64_t ldb, int op_b, int n, int k, int mode, uint4* out,
                         hipStream_t s) {
  if (n <= 0 || k <= 0) return hipSuccess;
  if (k % kPackK != 0) return hipErrorInvalidValue;
  const int64_t blocks = (int64_t)((n + kPackRows - 1) / kPackRows) * (k / kPackK);
  if (blocks > INT32_MAX) return hipErrorInvalidValue;
  const bool bf16 = mode == MODE_BF16;
  auto fn = op_b == OP_N ? (bf16 ? pack_b_kernel<true, true> : pack_b_kernel<true, false>)
                         : (bf16 ? pack_b_kernel<false, true> : pack_b_kernel<false, false>);
  hipLaunchKernelGGL(fn, dim3((unsigned)blocks), dim3(256), 0, s, b, ldb, n, k, out);
  return hipGetLastError();
}

size_t presplit_bytes(int m, int k, int mode) {
  return (size_t)((m + kPackRows - 1) / kPackRows) * kPackRows * (size_t)k * (mode == MODE_BF16 ? 6 : 4);
}

hipError_t launch_presplit_a(const float* a, int64_t lda, int op_a, int m, int k, int mode, const int* row_exp,
                             uint4* out, hipStream_t s) {
  if (m <= 0 || k <= 0) return hipSuccess;
  if (k % kPackK != 0) return hipErrorInvalidValue;
  const int64_t blocks = (int64_t)((m + kPackRows - 1) / kPackRows) * (k / kPackK);
  if (blocks > INT32_MAX) return hipErrorInvalidValue;
  const int vec = (reinterpret_cast<uintptr_t>(a) % 16 == 0) && (lda % 4 == 0);
  auto pick = [&](auto kcont_c) {
    constexpr bool KC = decltype(kcont_c)::value;
    return mode == MODE_BF16 ? presplit_pack_kernel<KC, MODE_BF16>
           : mode == MODE_FP16 ? presplit_pack_kernel<KC, MODE_FP16>
                               : presplit_pack_kernel<KC, MODE_FP16S>;
  };
  auto fn = op_a == OP_T ? pick(std::true_type{}) : pick(std::false_type{});
  hipLaunchKernelGGL(fn, dim3((unsigned)blocks), dim3(256), 0, s, a, lda, m, k, vec, row_exp, out);
  return hipGetLastError();
}

}  // namespace shgemm_amd
