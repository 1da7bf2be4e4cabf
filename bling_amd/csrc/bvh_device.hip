// Device build of the triangle BVH: the binned-SAH builder of bvh_build.cpp
// run on the GPU, then collapsed to the traversal's wide node layout.
//
// The tree is the host builder's tree: the same 12 bins per axis, the same
// binary32 cost arithmetic in the same order, the same leaf rules, and
// children's extents taken from the bins.  Only the numbering of nodes and
// the order of primitives inside a leaf may differ.  Bin bounds are
// reduced with integer atomics on an order-preserving encoding of the
// floats, which gives exact minima and maxima whatever the order.
//
// Phases:
//   1. primitive records (box, centroid, id) and the root extent;
//   2. ranges of more than kSmall primitives, one level at a time: binning
//      by chunks of kChunk primitives (workgroup bins in LDS merged with
//      global atomics), one decision per range, a stable partition by a
//      scan over the range's chunks;
//   3. ranges of at most kSmall primitives: one workgroup builds the whole
//      subtree depth first (bins in LDS, a wave-ballot stable partition);
//      each partition moves a range between the two primitive buffers, and
//      leaves copy their primitives back to the first;
//   4. collapse to width W level by level (opening the largest inner child
//      first, as collapse_bvh does) and the exact traversal stack bound,
//      bottom up;
//   5. the triangle leaf records.
// A range that the host would split at the median (identical centroids,
// depth > 40) or a non-finite primitive box makes the build report
// "fallback": the caller then builds on the host.
#include <chrono>

#include "bgpu_internal.h"
#include "bvh_device.h"

namespace bgpu {
namespace {

constexpr int NB = 12;               // bins per axis (BuildParams::bins)
constexpr int kMaxLeafSah = 4;       // BuildParams::max_leaf
constexpr int kBinWords = 13;        // lo[3] hi[3] cmin[3] cmax[3] count
constexpr int kBinsWords = 3 * NB * kBinWords;
constexpr int kChunk = 4096;         // primitives per workgroup in the level passes
constexpr int kLevelBlock = 256;
constexpr int kTreeBlock = 64;       // one wave per small subtree
constexpr int kTreeStack = 48;       // > the 42 pending ranges a depth <= 41 subtree can hold
constexpr int kBalancedDepth = 40;   // deeper ranges use the host's median split

enum { kFlagFallback = 1, kFlagOverflow = 2, kFlagNonFinite = 4 };
enum { kDecLeaf = 0, kDecSplit = 1, kDecFallback = 2 };

struct DPrim {
  float lo[3];
  int32_t id;
  float hi[3];
  float pad0;
  float c[3];
  float pad1;
};
static_assert(sizeof(DPrim) == 48, "DPrim is three float4");

struct Ext {
  float lo[3], hi[3], cmin[3], cmax[3];
};
struct BTask {
  int32_t b, e, depth, parent, slot;
  int32_t buf;  // which primitive buffer holds the range (0: A, 1: B)
  int32_t pad[2];
  Ext x;
};
struct Dec {
  int32_t kind, axis, bin, nleft;
  float cmin, sc;
  int32_t node, pad;
  Ext le, re;
};
struct Counters {
  int32_t nodes, next, small, leaves, max_depth, flag, wide, small_used;
};

// -- order-preserving float encoding for integer atomics
__device__ __forceinline__ uint32_t ord(float f) {
  const uint32_t u = __float_as_uint(f);
  return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float unord(uint32_t u) {
  return __uint_as_float((u & 0x80000000u) ? (u & 0x7fffffffu) : ~u);
}
constexpr uint32_t kOrdPosInf = 0xff800000u;  // ord(+inf)
constexpr uint32_t kOrdNegInf = 0x007fffffu;  // ord(-inf)
__device__ __forceinline__ uint32_t empty_word(int w) {
  const int k = w % kBinWords;
  if (k == 12) return 0u;
  return (k < 3 || (k >= 6 && k < 9)) ? kOrdPosInf : kOrdNegInf;
}

// std::min / std::max and Haskell's min / max, as the host code uses them
__device__ __forceinline__ float smin(float a, float b) { return (b < a) ? b : a; }
__device__ __forceinline__ float smax(float a, float b) { return (a < b) ? b : a; }
__device__ __forceinline__ float hmin(float x, float y) { return (x <= y) ? x : y; }
__device__ __forceinline__ float hmax(float x, float y) { return (x <= y) ? y : x; }

__device__ __forceinline__ float area3(const float* lo, const float* hi) {  // bvh_build.cpp area()
  float d[3];
  for (int k = 0; k < 3; ++k) {
    const float x = hi[k] - lo[k];
    d[k] = (0.0f < x) ? x : 0.0f;
  }
  return 2.0f * (d[0] * d[1] + d[0] * d[2] + d[1] * d[2]);
}

__device__ __forceinline__ Box pad_box_d(const float* lo, const float* hi) {  // bvh_build.cpp pad_box()
  Box r;
  float ext = 0.0f;
  for (int k = 0; k < 3; ++k) ext = smax(ext, hi[k] - lo[k]);
  if (!isfinite(ext)) ext = 0.0f;
  for (int k = 0; k < 3; ++k) {
    r.lo[k] = lo[k];
    r.hi[k] = hi[k];
    const float m = smax(fabsf(lo[k]), fabsf(hi[k]));
    if (!isfinite(m)) continue;
    const float e = ext * 1e-5f + m * 4e-6f + 1e-30f;
    r.lo[k] = lo[k] - e;
    r.hi[k] = hi[k] + e;
  }
  return r;
}

__device__ __forceinline__ int bin_of(float c, float cmin, float sc) {
  const int bi = (int)((c - cmin) * sc);
  return min(NB - 1, max(0, bi));
}

// Extent of bins [i0, i1] of axis `ax`.
__device__ void bins_extent(const uint32_t* bins, int ax, int i0, int i1, Ext& x) {
  for (int k = 0; k < 3; ++k) {
    x.lo[k] = INFINITY;
    x.hi[k] = -INFINITY;
    x.cmin[k] = INFINITY;
    x.cmax[k] = -INFINITY;
  }
  for (int i = i0; i <= i1; ++i) {
    const uint32_t* w = bins + (ax * NB + i) * kBinWords;
    for (int k = 0; k < 3; ++k) {
      x.lo[k] = smin(x.lo[k], unord(w[k]));
      x.hi[k] = smax(x.hi[k], unord(w[3 + k]));
      x.cmin[k] = smin(x.cmin[k], unord(w[6 + k]));
      x.cmax[k] = smax(x.cmax[k], unord(w[9 + k]));
    }
  }
}

__device__ __forceinline__ bool binned(int depth, const Ext& x) {
  return !(depth > kBalancedDepth) && area3(x.lo, x.hi) > 0.0f;
}

// The decision of Builder::split (bvh_build.cpp) for a range of n >= 2
// primitives at `depth` with extent x, from its bins.
__device__ void decide(const uint32_t* bins, int n, int depth, const Ext& x, Dec& d) {
  const bool balanced = depth > kBalancedDepth;
  float best_cost = INFINITY;
  int best_axis = -1, best_bin = -1;
  const float parea = area3(x.lo, x.hi);
  if (!balanced && parea > 0.0f) {
    for (int ax = 0; ax < 3; ++ax) {
      const float ext = x.cmax[ax] - x.cmin[ax];
      if (!(ext > 0.0f)) continue;
      float ra[NB];
      int rc[NB];
      float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
      int cnt = 0;
      for (int i = NB - 1; i > 0; --i) {
        const uint32_t* w = bins + (ax * NB + i) * kBinWords;
        for (int k = 0; k < 3; ++k) {
          lo[k] = smin(lo[k], unord(w[k]));
          hi[k] = smax(hi[k], unord(w[3 + k]));
        }
        cnt += (int)w[12];
        ra[i] = area3(lo, hi);
        rc[i] = cnt;
      }
      for (int k = 0; k < 3; ++k) {
        lo[k] = INFINITY;
        hi[k] = -INFINITY;
      }
      cnt = 0;
      for (int i = 0; i < NB - 1; ++i) {
        const uint32_t* w = bins + (ax * NB + i) * kBinWords;
        for (int k = 0; k < 3; ++k) {
          lo[k] = smin(lo[k], unord(w[k]));
          hi[k] = smax(hi[k], unord(w[3 + k]));
        }
        cnt += (int)w[12];
        if (cnt == 0 || rc[i + 1] == 0) continue;
        const float c = 1.0f + (area3(lo, hi) * cnt + ra[i + 1] * rc[i + 1]) / parea;
        if (c < best_cost) {
          best_cost = c;
          best_axis = ax;
          best_bin = i;
        }
      }
    }
  }
  d.kind = kDecFallback;
  if (!balanced && n <= kMaxLeafSah && (best_axis < 0 || (float)n <= best_cost)) d.kind = kDecLeaf;
  else if (!balanced && best_axis < 0 && n <= kMaxLeaf) d.kind = kDecLeaf;
  else if (balanced && n <= kMaxLeaf) d.kind = kDecLeaf;
  else if (best_axis >= 0) {
    d.kind = kDecSplit;
    d.axis = best_axis;
    d.bin = best_bin;
    d.cmin = x.cmin[best_axis];
    d.sc = (float)NB / (x.cmax[best_axis] - x.cmin[best_axis]);
    int nl = 0;
    for (int i = 0; i <= best_bin; ++i) nl += (int)bins[(best_axis * NB + i) * kBinWords + 12];
    d.nleft = nl;
    bins_extent(bins, best_axis, 0, best_bin, d.le);
    bins_extent(bins, best_axis, best_bin + 1, NB - 1, d.re);
  }
}

__device__ __forceinline__ int leaf_link(int b, int n) { return ~((b << kLeafCountBits) | n); }

__device__ __forceinline__ void bin_prim(uint32_t* bins, const DPrim& p, const bool* use, const float* cmin,
                                         const float* sc) {
  for (int ax = 0; ax < 3; ++ax) {
    if (!use[ax]) continue;
    uint32_t* w = bins + (ax * NB + bin_of(p.c[ax], cmin[ax], sc[ax])) * kBinWords;
    for (int k = 0; k < 3; ++k) {
      atomicMin(&w[k], ord(p.lo[k]));
      atomicMax(&w[3 + k], ord(p.hi[k]));
      atomicMin(&w[6 + k], ord(p.c[k]));
      atomicMax(&w[9 + k], ord(p.c[k]));
    }
    atomicAdd(&w[12], 1u);
  }
}

// Bins the primitives [i0, i1) of one thread, consecutive ones: per axis
// the run of primitives that fall into the same bin is reduced in registers
// and merged into the LDS bins once (neighbouring primitives of a range are
// close in space, so runs are long and LDS atomics few).
__device__ void bin_run(uint32_t* bins, const DPrim* __restrict__ src, int i0, int i1, const bool* use,
                        const float* cmin, const float* sc) {
  int cur[3] = {-1, -1, -1};
  float acc[3][12];
  uint32_t cnt[3] = {0, 0, 0};
  auto flush = [&](int ax) {
    uint32_t* w = bins + (ax * NB + cur[ax]) * kBinWords;
    for (int k = 0; k < 3; ++k) {
      atomicMin(&w[k], ord(acc[ax][k]));
      atomicMax(&w[3 + k], ord(acc[ax][3 + k]));
      atomicMin(&w[6 + k], ord(acc[ax][6 + k]));
      atomicMax(&w[9 + k], ord(acc[ax][9 + k]));
    }
    atomicAdd(&w[12], cnt[ax]);
  };
  for (int i = i0; i < i1; ++i) {
    const DPrim p = src[i];
#pragma unroll
    for (int ax = 0; ax < 3; ++ax) {
      if (!use[ax]) continue;
      const int bi = bin_of(p.c[ax], cmin[ax], sc[ax]);
      if (bi != cur[ax]) {
        if (cur[ax] >= 0) flush(ax);
        cur[ax] = bi;
        cnt[ax] = 0;
        for (int k = 0; k < 3; ++k) {
          acc[ax][k] = p.lo[k];
          acc[ax][3 + k] = p.hi[k];
          acc[ax][6 + k] = p.c[k];
          acc[ax][9 + k] = p.c[k];
        }
      } else {
        for (int k = 0; k < 3; ++k) {
          acc[ax][k] = fminf(acc[ax][k], p.lo[k]);
          acc[ax][3 + k] = fmaxf(acc[ax][3 + k], p.hi[k]);
          acc[ax][6 + k] = fminf(acc[ax][6 + k], p.c[k]);
          acc[ax][9 + k] = fmaxf(acc[ax][9 + k], p.c[k]);
        }
      }
      cnt[ax]++;
    }
  }
#pragma unroll
  for (int ax = 0; ax < 3; ++ax)
    if (use[ax] && cur[ax] >= 0) flush(ax);
}

__device__ __forceinline__ void bin_setup(const Ext& x, bool* use, float* cmin, float* sc) {
  for (int ax = 0; ax < 3; ++ax) {
    const float ext = x.cmax[ax] - x.cmin[ax];
    use[ax] = ext > 0.0f;
    sc[ax] = use[ax] ? (float)NB / ext : 0.0f;
    cmin[ax] = x.cmin[ax];
  }
}

// ---- phase 1: primitive records and the root extent ----------------------

__global__ __launch_bounds__(kLevelBlock) void k_prims(const float* __restrict__ v, const int32_t* __restrict__ idx,
                                                      int n, DPrim* __restrict__ prims, uint32_t* ext_words,
                                                      Counters* ctr) {
  __shared__ uint32_t red[12];
  if (threadIdx.x < 12) red[threadIdx.x] = (threadIdx.x % 6 < 3) ? kOrdPosInf : kOrdNegInf;
  __syncthreads();
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) {
    DPrim p;
    for (int k = 0; k < 3; ++k) {
      p.lo[k] = INFINITY;
      p.hi[k] = -INFINITY;
    }
    bool finite = true;
    for (int j = 0; j < 3; ++j) {  // triangleBounds: extendAABBP over the three vertices
      const float* q = v + 3 * (size_t)idx[3 * (size_t)t + j];
      for (int k = 0; k < 3; ++k) {
        p.lo[k] = hmin(p.lo[k], q[k]);
        p.hi[k] = hmax(p.hi[k], q[k]);
      }
    }
    for (int k = 0; k < 3; ++k) {
      finite = finite && isfinite(p.lo[k]) && isfinite(p.hi[k]);
      p.c[k] = 0.5f * (p.lo[k] + p.hi[k]);
    }
    p.id = t;
    p.pad0 = 0.0f;
    p.pad1 = 0.0f;
    prims[t] = p;
    if (!finite) atomicOr(&ctr->flag, kFlagNonFinite);
    for (int k = 0; k < 3; ++k) {  // root extent: box (0..5), centroid bounds (6..11)
      atomicMin(&red[k], ord(p.lo[k]));
      atomicMax(&red[3 + k], ord(p.hi[k]));
      atomicMin(&red[6 + k], ord(p.c[k]));
      atomicMax(&red[9 + k], ord(p.c[k]));
    }
  }
  __syncthreads();
  if (threadIdx.x < 12) {
    const int k = threadIdx.x;
    if (k % 6 < 3) atomicMin(&ext_words[k], red[k]);
    else atomicMax(&ext_words[k], red[k]);
  }
}

__global__ void k_root_task(const uint32_t* ext_words, int n, BTask* task) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  BTask t;
  t.b = 0;
  t.e = n;
  t.depth = 0;
  t.parent = -1;
  t.slot = 0;
  t.buf = 0;
  t.pad[0] = t.pad[1] = 0;
  for (int k = 0; k < 3; ++k) {
    t.x.lo[k] = unord(ext_words[k]);
    t.x.hi[k] = unord(ext_words[3 + k]);
    t.x.cmin[k] = unord(ext_words[6 + k]);
    t.x.cmax[k] = unord(ext_words[9 + k]);
  }
  *task = t;
}

// ---- phase 2: level passes over the large ranges -------------------------

// blk_first[t] = first chunk of task t, blk_first[ntask] = number of chunks
__global__ __launch_bounds__(1024) void k_blockmap(const BTask* tasks, int ntask, int* blk_first) {
  __shared__ int part[1024];
  const int tid = threadIdx.x;
  const int per = (ntask + 1023) / 1024;
  const int t0 = min(ntask, tid * per), t1 = min(ntask, t0 + per);
  int s = 0;
  for (int t = t0; t < t1; ++t) s += (tasks[t].e - tasks[t].b + kChunk - 1) / kChunk;
  part[tid] = s;
  __syncthreads();
  if (tid == 0) {
    int acc = 0;
    for (int i = 0; i < 1024; ++i) {
      const int x = part[i];
      part[i] = acc;
      acc += x;
    }
    blk_first[ntask] = acc;
  }
  __syncthreads();
  int acc = part[tid];
  for (int t = t0; t < t1; ++t) {
    blk_first[t] = acc;
    acc += (tasks[t].e - tasks[t].b + kChunk - 1) / kChunk;
  }
}

__device__ __forceinline__ int task_of_block(const int* blk_first, int ntask, int blk) {
  int lo = 0, hi = ntask - 1;  // last t with blk_first[t] <= blk
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (blk_first[mid] <= blk) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

__global__ void k_clear_bins(uint32_t* bins, int nwords) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < nwords) bins[i] = empty_word(i % kBinsWords);
}

__global__ __launch_bounds__(kLevelBlock) void k_bin(const BTask* __restrict__ tasks, int ntask,
                                                    const int* __restrict__ blk_first, const DPrim* P0,
                                                    const DPrim* P1, uint32_t* gbins) {
  __shared__ uint32_t bins[kBinsWords];
  const int blk = blockIdx.x;
  if (blk >= blk_first[ntask]) return;
  const int t = task_of_block(blk_first, ntask, blk);
  const BTask task = tasks[t];
  if (!binned(task.depth, task.x)) return;
  for (int w = threadIdx.x; w < kBinsWords; w += blockDim.x) bins[w] = empty_word(w);
  __syncthreads();
  bool use[3];
  float cmin[3], sc[3];
  bin_setup(task.x, use, cmin, sc);
  const int cb = task.b + (blk - blk_first[t]) * kChunk, ce = min(task.e, cb + kChunk);
  constexpr int per = kChunk / kLevelBlock;
  const int i0 = cb + (int)threadIdx.x * per;
  bin_run(bins, task.buf ? P1 : P0, i0, min(ce, i0 + per), use, cmin, sc);
  __syncthreads();
  uint32_t* g = gbins + (size_t)t * kBinsWords;
  for (int w = threadIdx.x; w < kBinsWords; w += blockDim.x) {
    const uint32_t x = bins[w];
    if (x == empty_word(w)) continue;
    const int k = w % kBinWords;
    if (k == 12) atomicAdd(&g[w], x);
    else if (k < 3 || (k >= 6 && k < 9)) atomicMin(&g[w], x);
    else atomicMax(&g[w], x);
  }
}

__device__ void emit_child(int b, int e, int depth, int parent, int slot, int buf, const Ext& x, BTask* next,
                           int next_cap, BTask* small, int small_cap, Counters* ctr) {
  BTask c;
  c.b = b;
  c.e = e;
  c.depth = depth;
  c.parent = parent;
  c.slot = slot;
  c.buf = buf;
  c.pad[0] = c.pad[1] = 0;
  c.x = x;
  if (e - b > kSmall) {
    const int q = atomicAdd(&ctr->next, 1);
    if (q < next_cap) next[q] = c;
    else atomicOr(&ctr->flag, kFlagOverflow);
  } else {
    const int q = atomicAdd(&ctr->small, 1);
    if (q < small_cap) small[q] = c;
    else atomicOr(&ctr->flag, kFlagOverflow);
  }
}

__global__ void k_split(const BTask* __restrict__ tasks, int ntask, const uint32_t* __restrict__ gbins, Dec* decs,
                        BNode* nodes, int node_cap, BTask* next, int next_cap, BTask* small, int small_cap,
                        Counters* ctr) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= ntask) return;
  const BTask task = tasks[t];
  const int n = task.e - task.b;
  Dec d;
  d.kind = kDecLeaf;
  d.node = -1;
  if (task.parent >= 0) nodes[task.parent].child[task.slot] = pad_box_d(task.x.lo, task.x.hi);
  if (n > 1) decide(gbins + (size_t)t * kBinsWords, n, task.depth, task.x, d);
  if (d.kind == kDecFallback) atomicOr(&ctr->flag, kFlagFallback);
  if (d.kind == kDecLeaf) {
    if (task.parent >= 0) nodes[task.parent].link[task.slot] = leaf_link(task.b, n);
    atomicAdd(&ctr->leaves, 1);
  } else if (d.kind == kDecSplit) {
    const int node = atomicAdd(&ctr->nodes, 1);
    if (node >= node_cap) {
      atomicOr(&ctr->flag, kFlagOverflow);
      d.kind = kDecFallback;
    } else {
      d.node = node;
      if (task.parent >= 0) nodes[task.parent].link[task.slot] = node;
      atomicMax(&ctr->max_depth, task.depth);
      emit_child(task.b, task.b + d.nleft, task.depth + 1, node, 0, task.buf ^ 1, d.le, next, next_cap, small,
                 small_cap, ctr);
      emit_child(task.b + d.nleft, task.e, task.depth + 1, node, 1, task.buf ^ 1, d.re, next, next_cap, small,
                 small_cap, ctr);
    }
  }
  decs[t] = d;
}

// left-side primitives per chunk of the splitting tasks
__global__ __launch_bounds__(kLevelBlock) void k_count(const BTask* __restrict__ tasks, int ntask,
                                                      const int* __restrict__ blk_first, const Dec* __restrict__ decs,
                                                      const DPrim* P0, const DPrim* P1, int* cnt) {
  __shared__ int s;
  const int blk = blockIdx.x;
  if (blk >= blk_first[ntask]) return;
  const int t = task_of_block(blk_first, ntask, blk);
  const Dec& d = decs[t];
  if (d.kind != kDecSplit) {
    if (threadIdx.x == 0) cnt[blk] = 0;
    return;
  }
  if (threadIdx.x == 0) s = 0;
  __syncthreads();
  const BTask& task = tasks[t];
  const DPrim* A = task.buf ? P1 : P0;
  const int cb = task.b + (blk - blk_first[t]) * kChunk, ce = min(task.e, cb + kChunk);
  int c = 0;
  for (int i = cb + (int)threadIdx.x; i < ce; i += blockDim.x) c += bin_of(A[i].c[d.axis], d.cmin, d.sc) <= d.bin;
  atomicAdd(&s, c);
  __syncthreads();
  if (threadIdx.x == 0) cnt[blk] = s;
}

__device__ __forceinline__ int block_excl_scan(int v, int* sh, int* total) {  // 256 threads
  const int tid = threadIdx.x;
  sh[tid] = v;
  __syncthreads();
  for (int o = 1; o < kLevelBlock; o <<= 1) {
    const int x = tid >= o ? sh[tid - o] : 0;
    __syncthreads();
    sh[tid] += x;
    __syncthreads();
  }
  const int incl = sh[tid];
  *total = sh[kLevelBlock - 1];
  __syncthreads();
  return incl - v;
}

// off[blk] = left-side primitives in the task's chunks before blk
__global__ __launch_bounds__(kLevelBlock) void k_scan_chunks(int ntask, const int* __restrict__ blk_first,
                                                            const Dec* __restrict__ decs, const int* __restrict__ cnt,
                                                            int* off) {
  __shared__ int sh[kLevelBlock];
  const int t = blockIdx.x;
  if (t >= ntask || decs[t].kind != kDecSplit) return;
  const int f0 = blk_first[t], f1 = blk_first[t + 1];
  int carry = 0;
  for (int r = f0; r < f1; r += kLevelBlock) {
    const int i = r + (int)threadIdx.x;
    const int v = i < f1 ? cnt[i] : 0;
    int total;
    const int ex = block_excl_scan(v, sh, &total);
    if (i < f1) off[i] = carry + ex;
    carry += total;
  }
}

__global__ __launch_bounds__(kLevelBlock) void k_scatter(const BTask* __restrict__ tasks, int ntask,
                                                        const int* __restrict__ blk_first, const Dec* __restrict__ decs,
                                                        const int* __restrict__ off, DPrim* P0, DPrim* P1) {
  __shared__ int sh[kLevelBlock];
  const int blk = blockIdx.x;
  if (blk >= blk_first[ntask]) return;
  const int t = task_of_block(blk_first, ntask, blk);
  const Dec d = decs[t];
  if (d.kind != kDecSplit) return;
  const BTask& task = tasks[t];
  const DPrim* A = task.buf ? P1 : P0;  // the range moves to the other buffer
  DPrim* B = task.buf ? P0 : P1;
  const int cb = task.b + (blk - blk_first[t]) * kChunk, ce = min(task.e, cb + kChunk);
  int lrun = task.b + off[blk];
  int rrun = task.b + d.nleft + (cb - task.b - off[blk]);
  for (int r = cb; r < ce; r += kLevelBlock) {
    const int i = r + (int)threadIdx.x;
    const bool act = i < ce;
    DPrim p;
    int f = 0;
    if (act) {
      p = A[i];
      f = bin_of(p.c[d.axis], d.cmin, d.sc) <= d.bin;
    }
    int nl;
    const int ex = block_excl_scan(f, sh, &nl);
    const int nact = min(kLevelBlock, ce - r);
    if (act) B[f ? lrun + ex : rrun + ((int)threadIdx.x - ex)] = p;
    lrun += nl;
    rrun += nact - nl;
  }
}

// ---- phase 3: one workgroup per small subtree ------------------------------

// The split decision of `decide`, evaluated by the 64 lanes of one wave:
// lane l < 33 prices the split after bin l % 11 on axis l / 11, and a wave
// arg-min keeps the first cheapest candidate in (axis, bin) order, as the
// serial loop's strict comparison does.  Lane 0 writes the result to d.
__device__ void decide_wave(const uint32_t* bins, int n, int depth, const Ext& x, Dec& d, int lane) {
  const bool balanced = depth > kBalancedDepth;
  const float parea = area3(x.lo, x.hi);
  float cost = INFINITY;
  int key = 0x7fffffff;
  if (!balanced && parea > 0.0f && lane < 3 * (NB - 1)) {
    const int ax = lane / (NB - 1), sb = lane % (NB - 1);
    if (x.cmax[ax] - x.cmin[ax] > 0.0f) {
      float llo[3] = {INFINITY, INFINITY, INFINITY}, lhi[3] = {-INFINITY, -INFINITY, -INFINITY};
      float rlo[3] = {INFINITY, INFINITY, INFINITY}, rhi[3] = {-INFINITY, -INFINITY, -INFINITY};
      int cl = 0, cr = 0;
      for (int i = 0; i < NB; ++i) {
        const uint32_t* w = bins + (ax * NB + i) * kBinWords;
        if (i <= sb) {
          for (int k = 0; k < 3; ++k) {
            llo[k] = smin(llo[k], unord(w[k]));
            lhi[k] = smax(lhi[k], unord(w[3 + k]));
          }
          cl += (int)w[12];
        } else {
          for (int k = 0; k < 3; ++k) {
            rlo[k] = smin(rlo[k], unord(w[k]));
            rhi[k] = smax(rhi[k], unord(w[3 + k]));
          }
          cr += (int)w[12];
        }
      }
      if (cl != 0 && cr != 0) {
        const float c = 1.0f + (area3(llo, lhi) * cl + area3(rlo, rhi) * cr) / parea;
        if (c < INFINITY) {
          cost = c;
          key = lane;
        }
      }
    }
  }
  for (int o = 32; o > 0; o >>= 1) {
    const float oc = __shfl_xor(cost, o);
    const int ok = __shfl_xor(key, o);
    if (oc < cost || (oc == cost && ok < key)) {
      cost = oc;
      key = ok;
    }
  }
  if (lane != 0) return;
  const bool found = key != 0x7fffffff;
  const int best_axis = found ? key / (NB - 1) : -1, best_bin = found ? key % (NB - 1) : -1;
  const float best_cost = cost;
  d.kind = kDecFallback;
  if (!balanced && n <= kMaxLeafSah && (best_axis < 0 || (float)n <= best_cost)) d.kind = kDecLeaf;
  else if (!balanced && best_axis < 0 && n <= kMaxLeaf) d.kind = kDecLeaf;
  else if (balanced && n <= kMaxLeaf) d.kind = kDecLeaf;
  else if (best_axis >= 0) {
    d.kind = kDecSplit;
    d.axis = best_axis;
    d.bin = best_bin;
    d.cmin = x.cmin[best_axis];
    d.sc = (float)NB / (x.cmax[best_axis] - x.cmin[best_axis]);
    int nl = 0;
    for (int i = 0; i <= best_bin; ++i) nl += (int)bins[(best_axis * NB + i) * kBinWords + 12];
    d.nleft = nl;
    bins_extent(bins, best_axis, 0, best_bin, d.le);
    bins_extent(bins, best_axis, best_bin + 1, NB - 1, d.re);
  }
}

// A pending range of a subtree: its task and which buffer holds its
// primitives (0: A, 1: B).  Partitions alternate between the buffers; a
// leaf's primitives are copied to A, where the leaf records are read.
struct SubTask {
  BTask t;
  int32_t buf, pad[3];
};

__global__ __launch_bounds__(kTreeBlock) void k_subtree(const BTask* __restrict__ small, int nsmall, DPrim* A,
                                                       DPrim* B, BNode* nodes, int node_cap, Counters* ctr) {
  __shared__ uint32_t bins[kBinsWords];
  __shared__ SubTask stack[kTreeStack];
  __shared__ SubTask cur;
  __shared__ Dec dec;
  __shared__ int sp, base, used, cap, stop, iters, nleaves, maxd;  // counters merged once at the end
  const int tid = threadIdx.x;
  if (blockIdx.x >= nsmall) return;
  if (tid == 0) {
    const BTask r = small[blockIdx.x];
    stack[0].t = r;
    stack[0].buf = r.buf;
    sp = 1;
    stop = 0;
    used = 0;
    iters = 0;
    nleaves = 0;
    maxd = 0;
    const int n = r.e - r.b;
    cap = n > 1 ? n - 1 : 0;  // a binary tree over n primitives has at most n - 1 inner nodes
    base = cap ? atomicAdd(&ctr->nodes, cap) : 0;
    if (cap && base + cap > node_cap) {
      atomicOr(&ctr->flag, kFlagOverflow);
      sp = 0;
    }
  }
  for (;;) {
    __syncthreads();
    if (sp == 0 || stop) break;
    __syncthreads();
    if (tid == 0) {
      cur = stack[--sp];
      nodes[cur.t.parent].child[cur.t.slot] = pad_box_d(cur.t.x.lo, cur.t.x.hi);
      if (++iters > 2 * cap + 2) {  // a subtree over n primitives has at most 2n - 1 ranges
        atomicOr(&ctr->flag, kFlagOverflow);
        stop = 1;
      }
    }
    __syncthreads();
    const int b = cur.t.b, e = cur.t.e, n = e - b;
    const DPrim* src = cur.buf ? B : A;
    DPrim* dst = cur.buf ? A : B;
    bool leaf = n <= 1;
    if (!leaf) {
      if (binned(cur.t.depth, cur.t.x)) {
        for (int w = tid; w < kBinsWords; w += kTreeBlock) bins[w] = empty_word(w);
        __syncthreads();
        bool use[3];
        float cmin[3], sc[3];
        bin_setup(cur.t.x, use, cmin, sc);
        const int per = (n + kTreeBlock - 1) / kTreeBlock;
        const int i0 = b + tid * per;
        bin_run(bins, src, i0, min(e, i0 + per), use, cmin, sc);
      }
      __syncthreads();
      decide_wave(bins, n, cur.t.depth, cur.t.x, dec, tid);
      if (tid == 0) {
        if (dec.kind == kDecSplit) {
          if (used >= cap || sp + 2 > kTreeStack) {
            atomicOr(&ctr->flag, kFlagOverflow);
            dec.kind = kDecFallback;
          } else {
            dec.node = base + used++;
            nodes[cur.t.parent].link[cur.t.slot] = dec.node;
            maxd = max(maxd, cur.t.depth);
          }
        } else if (dec.kind == kDecFallback) {
          atomicOr(&ctr->flag, kFlagFallback);
        }
        if (dec.kind == kDecFallback) stop = 1;
      }
      __syncthreads();
      if (dec.kind == kDecFallback) continue;
      leaf = dec.kind == kDecLeaf;
    }
    if (leaf) {
      if (tid == 0) {
        nodes[cur.t.parent].link[cur.t.slot] = leaf_link(b, n);
        nleaves++;
      }
      if (cur.buf)
        for (int i = b + tid; i < e; i += kTreeBlock) A[i] = B[i];
      continue;
    }
    // stable partition of [b, e) from src into dst
    int lrun = b, rrun = b + dec.nleft;
    for (int r = b; r < e; r += kTreeBlock) {
      const int i = r + tid;
      const bool act = i < e;
      DPrim p;
      bool f = false;
      if (act) {
        p = src[i];
        const float cv = dec.axis == 0 ? p.c[0] : (dec.axis == 1 ? p.c[1] : p.c[2]);
        f = bin_of(cv, dec.cmin, dec.sc) <= dec.bin;
      }
      const unsigned long long m = __ballot(f);
      const int below = __popcll(m & ((1ull << tid) - 1ull));
      if (act) dst[f ? lrun + below : rrun + (tid - below)] = p;
      const int nl = __popcll(m);
      lrun += nl;
      rrun += min(kTreeBlock, e - r) - nl;
    }
    if (tid == 0) {
      SubTask c;
      c.buf = cur.buf ^ 1;
      c.pad[0] = c.pad[1] = c.pad[2] = 0;
      c.t.depth = cur.t.depth + 1;
      c.t.parent = dec.node;
      c.t.buf = 0;
      c.t.pad[0] = c.t.pad[1] = 0;
      c.t.b = b + dec.nleft;
      c.t.e = e;
      c.t.slot = 1;
      c.t.x = dec.re;
      stack[sp++] = c;
      c.t.b = b;
      c.t.e = b + dec.nleft;
      c.t.slot = 0;
      c.t.x = dec.le;
      stack[sp++] = c;
    }
  }
  __syncthreads();
  if (tid == 0) {
    atomicAdd(&ctr->small_used, used);
    atomicAdd(&ctr->leaves, nleaves);
    if (used) atomicMax(&ctr->max_depth, maxd);
  }
}

// ---- phase 4: collapse to width W and the traversal stack bound ----------

template <int W>
__global__ void k_collapse(const BNode* __restrict__ n2, int lo, int hi, int* wl, float* __restrict__ wide, int cap,
                           Counters* ctr) {
  const int j = lo + (int)(blockIdx.x * blockDim.x + threadIdx.x);
  if (j >= hi) return;
  const BNode& r = n2[wl[j]];
  Box cb[W];
  int lk[W];
  cb[0] = r.child[0];
  cb[1] = r.child[1];
  lk[0] = r.link[0];
  lk[1] = r.link[1];
  int n = 2;
  while (n < W) {  // open the largest inner child (collapse_bvh)
    int pick = -1;
    float pa = -1.0f;
    for (int i = 0; i < W; ++i)
      if (i < n && lk[i] >= 0) {
        const float a = area3(cb[i].lo, cb[i].hi);
        if (a > pa) {
          pa = a;
          pick = i;
        }
      }
    if (pick < 0) break;
    int gi = 0;
    for (int i = 0; i < W; ++i)
      if (i == pick) gi = lk[i];
    const BNode& g = n2[gi];
    for (int i = 0; i < W; ++i) {
      if (i == pick) {
        cb[i] = g.child[0];
        lk[i] = g.link[0];
      } else if (i == n) {
        cb[i] = g.child[1];
        lk[i] = g.link[1];
      }
    }
    n++;
  }
  int k = 0;
  for (int i = 0; i < W; ++i) k += (i < n && lk[i] >= 0);
  int base = k ? atomicAdd(&ctr->wide, k) : 0;
  if (base + k > cap) {
    atomicOr(&ctr->flag, kFlagOverflow);
    base = -1;
  }
  float* o = wide + (size_t)j * 8 * W;
  int q = 0;
  for (int c = 0; c < W; ++c) {
    const bool used = c < n;
    const Box& b = cb[used ? c : 0];
    for (int a = 0; a < 3; ++a) {
      o[(2 * a) * W + c] = b.lo[a];
      o[(2 * a + 1) * W + c] = b.hi[a];
    }
    int link = kNoChild;
    if (used) {
      link = lk[c];
      if (link >= 0) {
        if (base >= 0) {
          wl[base + q] = link;
          link = base + q;
        } else {
          link = kNoChild;
        }
        q++;
      }
    }
    o[6 * W + c] = __int_as_float(link);
    o[7 * W + c] = 0.0f;
  }
}

__global__ void k_bound(const float* __restrict__ wide, int W, int lo, int hi, int* bound) {
  const int j = lo + (int)(blockIdx.x * blockDim.x + threadIdx.x);
  if (j >= hi) return;
  const float* o = wide + (size_t)j * 8 * W + 6 * W;
  int n = 0, m = 0;
  for (int c = 0; c < W; ++c) {
    const int l = __float_as_int(o[c]);
    if (l == kNoChild) continue;
    n++;
    if (l >= 0) m = max(m, bound[l]);
  }
  bound[j] = n - 1 + m;
}

// ---- phase 5: triangle leaf records (p1, id | p2 - p1, 0 | p3 - p1, 0) ---

__global__ void k_leaves(const DPrim* __restrict__ A, int n, const float* __restrict__ v,
                         const int32_t* __restrict__ idx, float4* __restrict__ out) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const int id = A[k].id;
  const float* p1 = v + 3 * (size_t)idx[3 * (size_t)id];
  const float* p2 = v + 3 * (size_t)idx[3 * (size_t)id + 1];
  const float* p3 = v + 3 * (size_t)idx[3 * (size_t)id + 2];
  out[3 * (size_t)k] = make_float4(p1[0], p1[1], p1[2], __int_as_float(id));
  out[3 * (size_t)k + 1] = make_float4(p2[0] - p1[0], p2[1] - p1[1], p2[2] - p1[2], __int_as_float(0));
  out[3 * (size_t)k + 2] = make_float4(p3[0] - p1[0], p3[1] - p1[1], p3[2] - p1[2], 0.0f);
}

inline int grid_of(long long n, int block) { return (int)((n + block - 1) / block); }

}  // namespace

struct DeviceBvh::Impl {
  DevBuf prims_a, prims_b, nodes2, task[2], small, decs, bins, blk_first, cnt, off, ctr, ext, wl, bound, wide;
  int n = 0, ntask_cap = 0, small_cap = 0, node_cap = 0;
  std::vector<int> level_off;  // wide level offsets of the last collapse
};

DeviceBvh::DeviceBvh() : impl_(new Impl()) {}
DeviceBvh::~DeviceBvh() { delete impl_; }

int DeviceBvh::build(Ctx* ctx, const float* d_verts, const int32_t* d_idx, int ntri, bool* fallback) {
  auto t0 = std::chrono::steady_clock::now();
  *fallback = false;
  Impl& I = *impl_;
  hipStream_t st = ctx->stream;
  if (ntri <= kSmall) {
    *fallback = true;
    return BGPU_OK;
  }
  const int n = ntri;
  I.n = n;
  I.ntask_cap = 2 * (n / (kSmall + 1)) + 2;
  I.small_cap = 2 * (kBalancedDepth + 2) * I.ntask_cap;
  I.node_cap = n;
  const int max_chunks = n / kChunk + I.ntask_cap + 1;
  if (int rc = I.prims_a.alloc(ctx, sizeof(DPrim) * (size_t)n)) return rc;
  if (int rc = I.prims_b.alloc(ctx, sizeof(DPrim) * (size_t)n)) return rc;
  if (int rc = I.nodes2.alloc(ctx, sizeof(BNode) * (size_t)n)) return rc;
  for (int k = 0; k < 2; ++k)
    if (int rc = I.task[k].alloc(ctx, sizeof(BTask) * (size_t)I.ntask_cap)) return rc;
  if (int rc = I.small.alloc(ctx, sizeof(BTask) * (size_t)I.small_cap)) return rc;
  if (int rc = I.decs.alloc(ctx, sizeof(Dec) * (size_t)I.ntask_cap)) return rc;
  if (int rc = I.bins.alloc(ctx, 4ull * kBinsWords * I.ntask_cap)) return rc;
  if (int rc = I.blk_first.alloc(ctx, 4ull * (I.ntask_cap + 1))) return rc;
  if (int rc = I.cnt.alloc(ctx, 4ull * max_chunks)) return rc;
  if (int rc = I.off.alloc(ctx, 4ull * max_chunks)) return rc;
  if (int rc = I.ctr.alloc(ctx, sizeof(Counters))) return rc;
  if (int rc = I.ext.alloc(ctx, 4ull * 12)) return rc;
  Counters* dctr = I.ctr.as<Counters>();
  BGPU_HIP(hipMemsetAsync(dctr, 0, sizeof(Counters), st));
  {
    uint32_t e[12];
    for (int k = 0; k < 12; ++k) e[k] = (k % 6 < 3) ? kOrdPosInf : kOrdNegInf;
    BGPU_HIP(hipMemcpyAsync(I.ext.ptr, e, sizeof(e), hipMemcpyHostToDevice, st));
  }
  DPrim* A = I.prims_a.as<DPrim>();
  DPrim* B = I.prims_b.as<DPrim>();
  BNode* N2 = I.nodes2.as<BNode>();
  hipLaunchKernelGGL(k_prims, dim3(grid_of(n, kLevelBlock)), dim3(kLevelBlock), 0, st, d_verts, d_idx, n, A,
                     I.ext.as<uint32_t>(), dctr);
  hipLaunchKernelGGL(k_root_task, dim3(1), dim3(1), 0, st, I.ext.as<uint32_t>(), n, I.task[0].as<BTask>());
  BGPU_HIP(hipGetLastError());
  int cur = 0, ntask = 1, level = 0;
  Counters h{};
  while (ntask > 0) {
    if (++level > kBalancedDepth + 2) {  // every range this deep has fallen back already
      h.flag |= kFlagFallback;
      break;
    }
    BTask* tasks = I.task[cur].as<BTask>();
    BTask* next = I.task[cur ^ 1].as<BTask>();
    int* bf = I.blk_first.as<int>();
    const Dec* decs = I.decs.as<Dec>();
    // chunks of this level: at most n / kChunk + ntask; surplus workgroups exit
    const int grid = n / kChunk + ntask + 1;
    hipLaunchKernelGGL(k_blockmap, dim3(1), dim3(1024), 0, st, tasks, ntask, bf);
    hipLaunchKernelGGL(k_clear_bins, dim3(grid_of((long long)ntask * kBinsWords, 256)), dim3(256), 0, st,
                       I.bins.as<uint32_t>(), ntask * kBinsWords);
    hipLaunchKernelGGL(k_bin, dim3(grid), dim3(kLevelBlock), 0, st, tasks, ntask, bf, A, B, I.bins.as<uint32_t>());
    BGPU_HIP(hipMemsetAsync(&dctr->next, 0, sizeof(int), st));
    hipLaunchKernelGGL(k_split, dim3(grid_of(ntask, 64)), dim3(64), 0, st, tasks, ntask, I.bins.as<uint32_t>(),
                       I.decs.as<Dec>(), N2, I.node_cap, next, I.ntask_cap, I.small.as<BTask>(), I.small_cap, dctr);
    hipLaunchKernelGGL(k_count, dim3(grid), dim3(kLevelBlock), 0, st, tasks, ntask, bf, decs, A, B, I.cnt.as<int>());
    hipLaunchKernelGGL(k_scan_chunks, dim3(ntask), dim3(kLevelBlock), 0, st, ntask, bf, decs, I.cnt.as<int>(),
                       I.off.as<int>());
    hipLaunchKernelGGL(k_scatter, dim3(grid), dim3(kLevelBlock), 0, st, tasks, ntask, bf, decs, I.off.as<int>(), A, B);
    BGPU_HIP(hipGetLastError());
    BGPU_HIP(hipMemcpyAsync(&h, dctr, sizeof(h), hipMemcpyDeviceToHost, st));
    BGPU_HIP(hipStreamSynchronize(st));
    if (h.flag) break;
    ntask = std::min(h.next, I.ntask_cap);
    cur ^= 1;
  }
  const int large_nodes = h.nodes;
  if (!h.flag && h.small > 0) {
    const int ns = std::min(h.small, I.small_cap);
    hipLaunchKernelGGL(k_subtree, dim3(ns), dim3(kTreeBlock), 0, st, I.small.as<BTask>(), ns, A, B, N2, I.node_cap,
                       dctr);
    BGPU_HIP(hipGetLastError());
    BGPU_HIP(hipMemcpyAsync(&h, dctr, sizeof(h), hipMemcpyDeviceToHost, st));
    BGPU_HIP(hipStreamSynchronize(st));
  }
  if (h.flag) {
    *fallback = true;
    fallback_reason = h.flag;
    release();
    return BGPU_OK;
  }
  nodes2_used = large_nodes + h.small_used;  // inner nodes (the subtrees' reservations may leave holes)
  leaves = h.leaves;
  max_depth = h.max_depth;
  I.task[0].release();
  I.task[1].release();
  I.small.release();
  I.decs.release();
  I.bins.release();
  I.blk_first.release();
  I.cnt.release();
  I.off.release();
  I.prims_b.release();
  build_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  return BGPU_OK;
}

template <int W>
static void launch_collapse(hipStream_t st, int blocks, const BNode* n2, int lo, int hi, int* wl, float* wide, int cap,
                            Counters* ctr) {
  hipLaunchKernelGGL(k_collapse<W>, dim3(blocks), dim3(256), 0, st, n2, lo, hi, wl, wide, cap, ctr);
}

int DeviceBvh::collapse(Ctx* ctx, int width, int* count, int* stack_bound) {
  Impl& I = *impl_;
  hipStream_t st = ctx->stream;
  if (width != 2 && width != 4 && width != 8) return fail(BGPU_E_INVALID, "device BVH: width must be 2, 4 or 8");
  const int cap = std::max(1, nodes2_used);  // a wide node per distinct inner BVH2 node at most
  if (int rc = I.wide.alloc(ctx, 4ull * 8 * width * cap)) return rc;
  if (int rc = I.wl.alloc(ctx, 4ull * cap)) return rc;
  if (int rc = I.bound.alloc(ctx, 4ull * cap)) return rc;
  Counters h{};
  h.wide = 1;
  const int zero = 0;
  BGPU_HIP(hipMemcpyAsync(I.ctr.ptr, &h, sizeof(h), hipMemcpyHostToDevice, st));
  BGPU_HIP(hipMemcpyAsync(I.wl.ptr, &zero, sizeof(int), hipMemcpyHostToDevice, st));
  I.level_off.assign(1, 0);
  int lo = 0, hi = 1;
  while (lo < hi) {
    if ((int)I.level_off.size() > 2 * kBalancedDepth + 8)  // deeper than the binary tree
      return fail(BGPU_E_INVALID, "device BVH: collapse deeper than the binary tree");
    const int blocks = grid_of(hi - lo, 256);
    const BNode* n2 = I.nodes2.as<BNode>();
    if (width == 2) launch_collapse<2>(st, blocks, n2, lo, hi, I.wl.as<int>(), I.wide.as<float>(), cap, I.ctr.as<Counters>());
    else if (width == 4) launch_collapse<4>(st, blocks, n2, lo, hi, I.wl.as<int>(), I.wide.as<float>(), cap, I.ctr.as<Counters>());
    else launch_collapse<8>(st, blocks, n2, lo, hi, I.wl.as<int>(), I.wide.as<float>(), cap, I.ctr.as<Counters>());
    BGPU_HIP(hipGetLastError());
    BGPU_HIP(hipMemcpyAsync(&h, I.ctr.ptr, sizeof(h), hipMemcpyDeviceToHost, st));
    BGPU_HIP(hipStreamSynchronize(st));
    if (h.flag) return fail(BGPU_E_INVALID, "device BVH: collapse overflow");
    I.level_off.push_back(hi);
    lo = hi;
    hi = h.wide;
  }
  for (int l = (int)I.level_off.size() - 2; l >= 0; --l) {
    const int a = I.level_off[l], b = I.level_off[l + 1];
    hipLaunchKernelGGL(k_bound, dim3(grid_of(b - a, 256)), dim3(256), 0, st, I.wide.as<float>(), width, a, b,
                       I.bound.as<int>());
  }
  BGPU_HIP(hipGetLastError());
  int bound0 = 0;
  BGPU_HIP(hipMemcpyAsync(&bound0, I.bound.ptr, sizeof(int), hipMemcpyDeviceToHost, st));
  BGPU_HIP(hipStreamSynchronize(st));
  *count = hi;
  *stack_bound = bound0;
  return BGPU_OK;
}

const float* DeviceBvh::wide_nodes() const { return impl_->wide.as<float>(); }

int DeviceBvh::leaves_to(Ctx* ctx, const float* d_verts, const int32_t* d_idx, float4* dst) {
  Impl& I = *impl_;
  hipLaunchKernelGGL(k_leaves, dim3(grid_of(I.n, 256)), dim3(256), 0, ctx->stream, I.prims_a.as<DPrim>(), I.n,
                     d_verts, d_idx, dst);
  BGPU_HIP(hipGetLastError());
  return BGPU_OK;
}

void DeviceBvh::release() {
  delete impl_;
  impl_ = new Impl();
}

}  // namespace bgpu
