// Persistent traversal with lane refill, shared by the ray-batch kernels and
// the bounce-loop kernels.  One pass walks one tree (triangles: FAT false;
// shapes: FAT true) for work items 0..n-1.  Lanes take a new item as soon as
// theirs is done, so a wave does not idle on its slowest ray, and refill
// once `refill_min` lanes are idle (new items then arrive in groups of
// neighbours).  Items come either from slices of kSlice taken from a global
// counter (`work` non-null: waves balance each other), or from one static
// contiguous range per wave (`work` null: no atomics, for short queues).  A
// step first runs inner nodes until every lane of the wave is at a leaf or
// done (or, with `walk_min` > 0, until fewer than walk_min lanes still walk
// while others wait at leaves), then tests the leaves together.
//
// The item source `Src` supplies, per item k:
//   bool begin(int k, DRay& r, LaneScene& ls, Best& best)
//       loads the ray, its scene and the starting best hit; false when the
//       item needs no traversal (root box missed, already occluded, no tree);
//   void end(int k, const DRay& r, const LaneScene& ls, const Best& best)
//       commits the result; called exactly once per item.  Any-hit passes
//       report occlusion as best.prim >= 0.
#pragma once
#include "trav.h"

namespace bgpu {
namespace trav {

#ifndef BGPU_REFILL_MIN
#define BGPU_REFILL_MIN 32
#endif
#ifndef BGPU_SLICE
#define BGPU_SLICE 128
#endif
#ifndef BGPU_SPECULATIVE
#define BGPU_SPECULATIVE 0
#endif
constexpr unsigned int kSlice = BGPU_SLICE;
constexpr int kRefillMin = BGPU_REFILL_MIN;

// Launch-time scheduling of a persistent pass (results do not depend on it).
struct PersistTune {
  int refill_min = kRefillMin;  // idle lanes that trigger a refill (1..64)
  int walk_min = 0;             // 0: walk until every lane is at a leaf
};
// A lane that reaches a leaf while other lanes of its wave still walk inner
// nodes postpones the leaf and walks on (speculative traversal, Aila and
// Laine 2009), so that the leaf step runs with more lanes active.
constexpr bool kSpeculative = BGPU_SPECULATIVE != 0;

// The part of a scene a traversal pass reads, per lane.
struct LaneScene {
  const float4* nodes;
  const float4* leaf;
  const DShape* shapes;
  float pad_k, floor_k;
  int root;  // root of the pass's tree, -1 when absent
};

template <int W, bool ANY, bool FAT, bool STATS, class Src>
__device__ __forceinline__ void persist_trace(Src& src, int n, unsigned int* work, const LdsSlot& slot,
                                              unsigned long long* counters, PersistTune tune = {}) {
  const int lane = __lane_id();
  int spill_n[kSpill];
  float spill_t[kCull ? kSpill : 1];
  Stack st;
  st.init(slot, spill_n, spill_t);
  int item = -1, cur = kNoChild;
  int pend = kNoChild;  // postponed leaf, kNoChild when none
  DRay r;
  LaneScene ls{nullptr, nullptr, nullptr, 0.0f, 0.0f, -1};
  f3 inv = mk3(0.0f, 0.0f, 0.0f), oi = inv;
  Best best{0.0f, 0.0f, 0.0f, -1};
  unsigned long long nvisit = 0, ntest = 0;
  unsigned int s_next = 0, s_end = 0;  // the wave's slice (uniform)
  bool drained = false;
  if (!work) {  // one static range per wave
    const unsigned int waves = gridDim.x * (blockDim.x >> 6);
    const unsigned int wave = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    const unsigned int per = ((unsigned int)n + waves - 1) / waves;
    s_next = wave * per < (unsigned int)n ? wave * per : (unsigned int)n;
    s_end = s_next + per < (unsigned int)n ? s_next + per : (unsigned int)n;
    drained = true;  // no further slices
  }
  for (;;) {
    // refill idle lanes from the wave's slice, taking a new slice when empty
    const int n_idle = __popcll(__ballot(item < 0));
    if (n_idle > 0 && (n_idle >= tune.refill_min || n_idle == 64) && (!drained || s_next < s_end)) {
      if (!drained && s_next >= s_end) {
        unsigned int base = 0;
        if (lane == 0) base = atomicAdd(work, kSlice);
        base = __builtin_amdgcn_readfirstlane(__shfl(base, 0));
        if (base >= (unsigned int)n) {
          drained = true;
        } else {
          s_next = base;
          s_end = base + kSlice < (unsigned int)n ? base + kSlice : (unsigned int)n;
        }
      }
      if (s_next < s_end) {
        const unsigned long long m = __ballot(item < 0);
        const unsigned int pos = s_next + (unsigned int)__popcll(m & ((1ull << lane) - 1ull));
        if (item < 0 && pos < s_end) {
          item = (int)pos;
          const bool live = src.begin(item, r, ls, best);
          inv = mk3(safe_rcp(r.d.x), safe_rcp(r.d.y), safe_rcp(r.d.z));
          oi = mk3(r.o.x * inv.x, r.o.y * inv.y, r.o.z * inv.z);
          st.sp = 0;
          pend = kNoChild;
          cur = live && ls.root >= 0 ? ls.root : kNoChild;
        }
        const unsigned int took = (unsigned int)__popcll(m);
        s_next = s_next + took < s_end ? s_next + took : s_end;
      }
    }
    if (__ballot(item >= 0) == 0) {
      if (drained) break;
      continue;
    }
    if (kSpeculative && item >= 0 && cur < 0 && pend == kNoChild) {
      pend = cur;
      if (st.sp > 0) {
        float t;
        st.pop(&cur, &t);
      } else {
        cur = kNoChild;
      }
    }
    // inner nodes until every lane is at a leaf or done; with walk_min, the
    // walk also stops once fewer than walk_min lanes still walk while at
    // least one other lane waits at a leaf (the decision is wave-uniform, and
    // a waiting lane guarantees progress)
    const int n_live = tune.walk_min > 0 ? __popcll(__ballot(item >= 0)) : 0;
    while (item >= 0 && cur >= 0 && cur != kNoChild) {
      if (tune.walk_min > 0) {
        const int n_walk = __popcll(__ballot(true));
        if (n_walk < tune.walk_min && n_walk < n_live) break;
      }
      const float4* nd = ls.nodes + (size_t)cur * (2 * W);
      float v[8 * W];
#pragma unroll
      for (int q = 0; q < (FAT ? 2 * W : (7 * W + 3) / 4); ++q) {
        const float4 x = nd[q];
        v[4 * q] = x.x; v[4 * q + 1] = x.y; v[4 * q + 2] = x.z; v[4 * q + 3] = x.w;
      }
      if (STATS) ++nvisit;
      const float tlim = ANY ? r.tmax : best.t;
      float pad[W];
      if (FAT) node_margins<W>(v, r.o, ls.pad_k, ls.floor_k, pad);
      float key[W];
      int lk[W];
      int nhit = 0;
#pragma unroll
      for (int c = 0; c < W; ++c) {
        float lx = v[c], hx = v[W + c], ly = v[2 * W + c], hy = v[3 * W + c], lz = v[4 * W + c],
              hz = v[5 * W + c];
        lk[c] = __float_as_int(v[6 * W + c]);
        if (FAT) inflate(lx, hx, ly, hy, lz, hz, pad[c]);
        bool h;
        float tn;
        slab(lx, hx, ly, hy, lz, hz, inv, oi, r.tmin, tlim, &h, &tn);
        h = h && lk[c] != kNoChild;
        key[c] = h ? fminf(tn, 3.0e38f) : INFINITY;
        nhit += h ? 1 : 0;
      }
      if (nhit > 0) {
        sort_children<W>(key, lk);
#pragma unroll
        for (int c = W - 1; c >= 1; --c)
          if (c < nhit) st.push(lk[c], key[c]);
        cur = lk[0];
      } else if (st.sp > 0) {
        float t;
        st.pop(&cur, &t);
      } else {
        cur = kNoChild;
      }
      if (kSpeculative) {
        if (cur < 0 && pend == kNoChild) {  // postpone the first leaf, walk on
          pend = cur;
          if (st.sp > 0) {
            float t;
            st.pop(&cur, &t);
          } else {
            cur = kNoChild;
          }
        }
        // stop once no lane still looks for its first leaf
        if (__ballot(item >= 0 && pend == kNoChild && cur >= 0 && cur != kNoChild) == 0) break;
      }
    }
    // leaves: the postponed one, else the one the walk stopped at
    const bool at_leaf = kSpeculative ? (item >= 0 && pend != kNoChild) : (item >= 0 && cur < 0);
    if (at_leaf) {
      const int leaf = kSpeculative ? pend : cur;
      if (kSpeculative) pend = kNoChild;
      const int enc = ~leaf;
      const int first = enc >> 4, cnt = enc & 15;
      bool hit_any = false;
      for (int k = first; k < first + cnt; ++k) {
        const float4* pr = ls.leaf + 3 * k;
        if (STATS) ++ntest;
        const float4 r0 = pr[0], r1 = pr[1];
        const int id = __float_as_int(r0.w);
        float t, b1 = 0.0f, b2 = 0.0f;
        bool hit;
        if (!FAT) {
          const float4 r2 = pr[2];
          hit = tri_test(mk3(r0.x, r0.y, r0.z), mk3(r1.x, r1.y, r1.z), mk3(r2.x, r2.y, r2.z), r.o, r.d,
                         r.tmin, ANY ? r.tmax : best.t, &t, &b1, &b2);
        } else {
          const DShape& s = ls.shapes[~__float_as_int(r1.w)];
          DRay ro;
          ro.o = xform_point(s.w2o, r.o);
          ro.d = xform_vector(s.w2o, r.d);
          ro.tmin = r.tmin;
          ro.tmax = ANY ? r.tmax : best.t;
          if (ANY) {
            hit = shape_intersects(s, ro);
          } else {
            int ax;
            hit = shape_intersect(s, ro, &t, &ax);
          }
        }
        if (ANY) {
          if (hit) {
            hit_any = true;
            break;
          }
        } else if (hit && (best.prim < 0 || t < best.t || (t == best.t && id < best.prim))) {
          best.t = t;
          best.prim = id;
          best.b1 = b1;
          best.b2 = b2;
        }
      }
      if (ANY && hit_any) {
        best.prim = 0;  // occluded
        cur = kNoChild;
        st.sp = 0;
      } else if (!kSpeculative) {
        if (st.sp > 0) {
          float t;
          st.pop(&cur, &t);
        } else {
          cur = kNoChild;
        }
      }
    }
    // finished items commit their result and free the lane
    if (item >= 0 && cur == kNoChild && (!kSpeculative || pend == kNoChild)) {
      src.end(item, r, ls, best);
      item = -1;
      best.prim = -1;
    }
  }
  if (STATS) {
    unsigned long long a = nvisit, b = ntest;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
      a += __shfl_xor(a, off);
      b += __shfl_xor(b, off);
    }
    if (lane == 0) {  // counters: triangle nodes, tests, then shape nodes, tests
      atomicAdd(&counters[FAT ? 2 : 0], a);
      atomicAdd(&counters[FAT ? 3 : 1], b);
    }
  }
}

// Blocks that fill the device once for a persistent kernel launched with
// `lds_bytes` of dynamic LDS.
template <class K>
int persistent_grid(K kernel, int block, size_t lds_bytes) {
  int dev = 0, per_cu = 1;
  (void)hipGetDevice(&dev);
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, dev) != hipSuccess) return 1024;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, block, lds_bytes) != hipSuccess ||
      per_cu < 1)
    per_cu = 1;
  return per_cu * prop.multiProcessorCount;
}

// persistent_grid per LDS stack depth, computed on first use of each depth.
template <class K>
int persistent_grid_cached(int (&cache)[kMaxShort + 1], K kernel, int block, int nshort) {
  if (!cache[nshort]) cache[nshort] = persistent_grid(kernel, block, stack_lds_bytes(nshort, block));
  return cache[nshort];
}

}  // namespace trav
}  // namespace bgpu
