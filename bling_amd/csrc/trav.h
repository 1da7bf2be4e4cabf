// Device BVH traversal shared by the ray-batch kernels and the path tracer.
// One ray per lane over a BVH of width 2, 4 or 8 (8 * width floats per node:
// the children's slabs as six arrays of width floats, then the links).  Hit
// children are sorted near to far; the nearest is entered and the others are
// pushed with their entry distances, so that a closest-hit query skips popped
// subtrees that start beyond its current hit.  The stack keeps its top
// `nshort` entries (kMinShort..kMaxShort, chosen at launch) in dynamic LDS,
// laid out [depth][lane] so a wave's pushes hit 64 distinct banks, and spills
// deeper entries to a private array; scene creation bounds the worst-case
// depth by kStackCapacity.  Node boxes are tested with explicit
// FMAs (they are padded outward, so the test, and the culling by entry
// distance, only have to be conservative); primitives are tested with the
// reference's exact arithmetic.
#pragma once
#include "isect.h"

namespace bgpu {
namespace trav {

// Stack entries kept in LDS per lane (the range of the launch-time choice),
// and whether entries carry their entry distance (closest-hit queries then
// skip subtrees beyond the current hit).
#ifndef BGPU_TRAV_CULL
#define BGPU_TRAV_CULL 0
#endif
constexpr int kMinShort = kMinStackLds, kMaxShort = kMaxStackLds;
constexpr bool kCull = BGPU_TRAV_CULL != 0;
// private entries beyond the shallowest LDS part
constexpr int kSpill = kStackCapacity - kMinShort;
// LDS ints per lane for nshort entries: node links, then entry distances
__host__ __device__ constexpr int lds_per_lane(int nshort) { return kCull ? 2 * nshort : nshort; }
// Dynamic LDS bytes of a traversal kernel's block.
inline size_t stack_lds_bytes(int nshort, int block) {
  return (size_t)lds_per_lane(nshort) * (size_t)block * sizeof(int);
}

// Where a lane's LDS stack entries are (entry d at lds[d * stride],
// d < nshort), and the launch's walk threshold: a wave walks inner nodes
// until every lane is at a leaf or done, or (walk > 0) until fewer than
// `walk` lanes still walk while another waits at a leaf.
struct LdsSlot {
  int* lds;
  int stride;
  int nshort;
  int walk;
};
// This lane's slot in the block's dynamic LDS (stack_lds_bytes(nshort, BLOCK)
// bytes, given at launch).
template <int BLOCK>
__device__ __forceinline__ LdsSlot lane_slot(int nshort, int walk) {
  extern __shared__ int bgpu_stack_lds[];
  return LdsSlot{bgpu_stack_lds + threadIdx.x, BLOCK, nshort, walk};
}

// The counters live in registers; only the spill arrays, which are indexed
// by the stack depth, are private memory (they belong to the caller, so that
// the compiler keeps sp and stride out of scratch).
struct Stack {
  int* lds;    // lds_per_lane(nshort) entries per lane, `stride` apart
  int stride;
  int nshort;  // entries in LDS (uniform over the launch)
  int walk;    // walk threshold of the launch (LdsSlot)
  int sp;
  int* spill_n;    // kSpill entries
  float* spill_t;  // kSpill entries when kCull
  __device__ __forceinline__ void init(const LdsSlot& s, int* sn, float* stt) {
    lds = s.lds;
    stride = s.stride;
    nshort = s.nshort;
    walk = s.walk;
    sp = 0;
    spill_n = sn;
    spill_t = stt;
  }
  __device__ __forceinline__ void push(int v, float t) {
    if (sp < nshort) {
      lds[sp * stride] = v;
      if (kCull) lds[(nshort + sp) * stride] = __float_as_int(t);
    } else {
      spill_n[sp - nshort] = v;
      if (kCull) spill_t[sp - nshort] = t;
    }
    ++sp;
  }
  __device__ __forceinline__ void pop(int* v, float* t) {
    --sp;
    if (sp < nshort) {
      *v = lds[sp * stride];
      *t = kCull ? __int_as_float(lds[(nshort + sp) * stride]) : 0.0f;
    } else {
      *v = spill_n[sp - nshort];
      *t = kCull ? spill_t[sp - nshort] : 0.0f;
    }
  }
};

// Reciprocal direction for the (padded, conservative) node boxes.  Components
// below 1e-20 in magnitude are replaced by +-1e-20 so that lo*inv - o*inv
// never forms inf - inf: the box test then sees a ray that moves 1e-20 per
// unit t along that axis, far below the box padding.
__device__ __forceinline__ float safe_rcp(float d) {
  return 1.0f / (fabsf(d) < 1e-20f ? copysignf(1e-20f, d) : d);
}

__device__ __forceinline__ DRay load_ray(const DevRays& r, int i) {
  DRay ray;
  ray.o = mk3(r.ox[i], r.oy[i], r.oz[i]);
  ray.d = mk3(r.dx[i], r.dy[i], r.dz[i]);
  ray.tmin = r.tmin[i];
  ray.tmax = r.tmax[i];
  return ray;
}

__device__ __forceinline__ void slab(float lx, float hx, float ly, float hy, float lz, float hz,
                                     f3 inv, f3 oi, float tmin, float tmax, bool* h, float* tn) {
  float x0 = __builtin_fmaf(lx, inv.x, -oi.x), x1 = __builtin_fmaf(hx, inv.x, -oi.x);
  float y0 = __builtin_fmaf(ly, inv.y, -oi.y), y1 = __builtin_fmaf(hy, inv.y, -oi.y);
  float z0 = __builtin_fmaf(lz, inv.z, -oi.z), z1 = __builtin_fmaf(hz, inv.z, -oi.z);
  float n = fmaxf(fmaxf(fminf(x0, x1), fminf(y0, y1)), fmaxf(fminf(z0, z1), tmin));
  float f = fminf(fminf(fmaxf(x0, x1), fmaxf(y0, y1)), fminf(fmaxf(z0, z1), tmax));
  *h = n <= f;
  *tn = n;
}

// Margins of the children of a shape-tree node for a ray from o (see the
// false-hit bound in isect.h).  D is the distance from o to the farthest
// corner of the node's box (the union of its children; unused child slots
// repeat child 0's box, so the union is exact), an upper bound of |o - c| for
// the centre c of every shape below the node.  Child c gets
//   max(min(k * D, kSphereK0^2 * D^2 * h_c), floor_k * (|o|_1 + |box|_1))
// with h_c = the largest sigma(o2w) * sigma(w2o)^2 / (2 r) of the subtree's
// spheres (+inf when it holds another shape kind), stored in the node's
// last W floats, and |box|_1 the L1 norm of the box corner farthest from the
// world origin (rounding of the ray's transform and of t).
template <int W>
__device__ __forceinline__ void node_margins(const float* v, f3 o, float k, float floor_k, float* pad) {
  float lx = v[0], hx = v[W], ly = v[2 * W], hy = v[3 * W], lz = v[4 * W], hz = v[5 * W];
#pragma unroll
  for (int c = 1; c < W; ++c) {
    lx = fminf(lx, v[c]); hx = fmaxf(hx, v[W + c]);
    ly = fminf(ly, v[2 * W + c]); hy = fmaxf(hy, v[3 * W + c]);
    lz = fminf(lz, v[4 * W + c]); hz = fmaxf(hz, v[5 * W + c]);
  }
  const float fx = fmaxf(fabsf(o.x - lx), fabsf(o.x - hx));
  const float fy = fmaxf(fabsf(o.y - ly), fabsf(o.y - hy));
  const float fz = fmaxf(fabsf(o.z - lz), fabsf(o.z - hz));
  const float d2 = (fx * fx + fy * fy + fz * fz) * 1.0001f;  // rounded up
  const float cap = k * __builtin_sqrtf(d2) * 1.0001f;
  const float q = kSphereK0 * kSphereK0 * d2;
  const float ext = fmaxf(fabsf(lx), fabsf(hx)) + fmaxf(fabsf(ly), fabsf(hy)) + fmaxf(fabsf(lz), fabsf(hz));
  const float fl = floor_k * (fabsf(o.x) + fabsf(o.y) + fabsf(o.z) + ext);
#pragma unroll
  for (int c = 0; c < W; ++c) pad[c] = fmaxf(fminf(cap, q * v[7 * W + c]), fl);
}
__device__ __forceinline__ void inflate(float& lx, float& hx, float& ly, float& hy, float& lz,
                                        float& hz, float p) {
  lx -= p; ly -= p; lz -= p;
  hx += p; hy += p; hz += p;
}

struct Best {
  float t, b1, b2;
  int prim;
};

// Work counts of a STATS traversal: node fetches and primitive tests, per
// tree (index 0: triangles, 1: shapes).
struct Counts {
  unsigned long long nodes[2] = {0, 0}, tests[2] = {0, 0};
};

// Compare-exchange on (key, link) pairs, ascending by key.
__device__ __forceinline__ void cex(float& ka, int& la, float& kb, int& lb) {
  const bool s = kb < ka;
  const float k0 = s ? kb : ka, k1 = s ? ka : kb;
  const int l0 = s ? lb : la, l1 = s ? la : lb;
  ka = k0; kb = k1; la = l0; lb = l1;
}
template <int W>
__device__ __forceinline__ void sort_children(float (&k)[W], int (&l)[W]) {
  if constexpr (W == 2) {
    cex(k[0], l[0], k[1], l[1]);
  } else if constexpr (W == 4) {
    cex(k[0], l[0], k[1], l[1]); cex(k[2], l[2], k[3], l[3]);
    cex(k[0], l[0], k[2], l[2]); cex(k[1], l[1], k[3], l[3]);
    cex(k[1], l[1], k[2], l[2]);
  } else {  // Batcher's odd-even merge sort of 8
    cex(k[0], l[0], k[1], l[1]); cex(k[2], l[2], k[3], l[3]);
    cex(k[4], l[4], k[5], l[5]); cex(k[6], l[6], k[7], l[7]);
    cex(k[0], l[0], k[2], l[2]); cex(k[1], l[1], k[3], l[3]);
    cex(k[4], l[4], k[6], l[6]); cex(k[5], l[5], k[7], l[7]);
    cex(k[1], l[1], k[2], l[2]); cex(k[5], l[5], k[6], l[6]);
    cex(k[0], l[0], k[4], l[4]); cex(k[1], l[1], k[5], l[5]);
    cex(k[2], l[2], k[6], l[6]); cex(k[3], l[3], k[7], l[7]);
    cex(k[2], l[2], k[4], l[4]); cex(k[3], l[3], k[5], l[5]);
    cex(k[1], l[1], k[2], l[2]); cex(k[3], l[3], k[4], l[4]);
    cex(k[5], l[5], k[6], l[6]);
  }
}

// Traverses one tree.  CLOSEST: shrinks best.t; ANY: returns at the first hit.
// EXACT (with ANY): shapes are tested with `intersect` rather than
// `intersects`, so that "any hit" is exactly "a closest hit exists".  Each
// lane visits the same nodes in the same order whatever the other lanes do;
// the wave walks inner nodes together, then tests the leaves it reached
// together (st.walk sets when the walk stops early, see LdsSlot).
template <int W, bool ANY, bool FAT, bool STATS, bool EXACT = false>
__device__ __forceinline__ bool traverse(const DevScene& sc, int root, const DRay& r, f3 inv,
                                         f3 oi, Stack& st, Best& best, Counts& cn) {
  int cur = root;  // inner node (>= 0, not kNoChild), leaf (< 0) or kNoChild (done)
  st.sp = 0;
  // the next stack entry whose subtree may still hold a hit at or before
  // best.t (an equal t can still win the tie on the primitive id), or kNoChild
  auto next = [&]() -> int {
    while (st.sp > 0) {
      int n;
      float t;
      st.pop(&n, &t);
      if (ANY || !kCull || t <= best.t) return n;
    }
    return kNoChild;
  };
  for (;;) {
    const int n_live = st.walk > 0 ? __popcll(__ballot(true)) : 0;
    while (cur >= 0 && cur != kNoChild) {
      if (st.walk > 0) {
        const int n_walk = __popcll(__ballot(true));
        if (n_walk < st.walk && n_walk < n_live) break;
      }
      const float4* nd = sc.nodes + (size_t)cur * (2 * W);
      float v[8 * W];
#pragma unroll
      for (int q = 0; q < (FAT ? 2 * W : (7 * W + 3) / 4); ++q) {
        const float4 x = nd[q];
        v[4 * q] = x.x; v[4 * q + 1] = x.y; v[4 * q + 2] = x.z; v[4 * q + 3] = x.w;
      }
      if (STATS) ++cn.nodes[FAT ? 1 : 0];
      const float tlim = ANY ? r.tmax : best.t;
      float pad[W];
      if (FAT) node_margins<W>(v, r.o, sc.shape_pad_k, sc.shape_floor_k, pad);
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
        key[c] = h ? fminf(tn, 3.0e38f) : INFINITY;  // misses sort last
        nhit += h ? 1 : 0;
      }
      if (nhit > 0) {
        sort_children<W>(key, lk);
#pragma unroll
        for (int c = W - 1; c >= 1; --c)
          if (c < nhit) st.push(lk[c], key[c]);
        cur = lk[0];
      } else {
        cur = next();
      }
    }
    if (cur == kNoChild) break;
    if (cur < 0) {  // a leaf
      int enc = ~cur;
      int first = enc >> 4, cnt = enc & 15;
      for (int k = first; k < first + cnt; ++k) {
        const float4* pr = sc.leaf + 3 * k;
        if (STATS) ++cn.tests[FAT ? 1 : 0];
        float4 r0 = pr[0], r1 = pr[1];
        int id = __float_as_int(r0.w);
        float t, b1 = 0.0f, b2 = 0.0f;
        bool hit;
        if (!FAT) {
          float4 r2 = pr[2];
          hit = tri_test(mk3(r0.x, r0.y, r0.z), mk3(r1.x, r1.y, r1.z), mk3(r2.x, r2.y, r2.z), r.o,
                         r.d, r.tmin, ANY ? r.tmax : best.t, &t, &b1, &b2);
        } else {
          const DShape& s = sc.shapes[~__float_as_int(r1.w)];
          DRay ro;
          ro.o = xform_point(s.w2o, r.o);
          ro.d = xform_vector(s.w2o, r.d);
          ro.tmin = r.tmin;
          ro.tmax = ANY ? r.tmax : best.t;
          if (ANY && !EXACT) {
            hit = shape_intersects(s, ro);
          } else {
            int ax;
            hit = shape_intersect(s, ro, &t, &ax);
          }
        }
        if (ANY) {
          if (hit) return true;
        } else if (hit && (best.prim < 0 || t < best.t || (t == best.t && id < best.prim))) {
          best.t = t;
          best.prim = id;
          best.b1 = b1;
          best.b2 = b2;
        }
      }
      cur = next();
      if (cur == kNoChild) break;
    }
  }
  return false;
}

template <int W, bool STATS>
__device__ __forceinline__ void closest_w(const DevScene& sc, const DRay& r, f3 inv, f3 oi, Stack& st,
                                          Best& best, Counts& cn) {
  if (sc.tri_root >= 0) traverse<W, false, false, STATS>(sc, sc.tri_root, r, inv, oi, st, best, cn);
  if (sc.shape_root >= 0) traverse<W, false, true, STATS>(sc, sc.shape_root, r, inv, oi, st, best, cn);
}
template <int W, bool STATS, bool EXACT = false>
__device__ __forceinline__ bool any_w(const DevScene& sc, const DRay& r, f3 inv, f3 oi, Stack& st,
                                      Counts& cn) {
  Best best{r.tmax, 0.0f, 0.0f, -1};
  if (sc.tri_root >= 0 && traverse<W, true, false, STATS, EXACT>(sc, sc.tri_root, r, inv, oi, st, best, cn))
    return true;
  if (sc.shape_root >= 0 && traverse<W, true, true, STATS, EXACT>(sc, sc.shape_root, r, inv, oi, st, best, cn))
    return true;
  return false;
}

// Closest hit over both trees (triangles, then shapes), after the root-box
// test of the accelerator primitive (Primitive/KdTree.hs:240), for a scene of
// width W.
template <int W, bool STATS, bool ROOT_INF = false>
__device__ __forceinline__ void closest_hit_w(const DevScene& sc, const DRay& r, const LdsSlot& slot,
                                              Best& best, Counts& cn) {
  best.t = r.tmax;
  best.b1 = 0.0f;
  best.b2 = 0.0f;
  best.prim = -1;
  if (ROOT_INF) {  // the root box test of the unbounded ray; r.tmax only bounds the traversal
    DRay ru = r;
    ru.tmax = INFINITY;
    if (!slab_haskell(sc.root_lo, sc.root_hi, ru)) return;
  } else if (!slab_haskell(sc.root_lo, sc.root_hi, r)) {
    return;
  }
  f3 inv = mk3(safe_rcp(r.d.x), safe_rcp(r.d.y), safe_rcp(r.d.z));
  f3 oi = mk3(r.o.x * inv.x, r.o.y * inv.y, r.o.z * inv.z);
  int spill_n[kSpill];
  float spill_t[kCull ? kSpill : 1];
  Stack st;
  st.init(slot, spill_n, spill_t);
  closest_w<W, STATS>(sc, r, inv, oi, st, best, cn);
}

// Any hit (Primitive/KdTree.hs:244-246) for a scene of width W; EXACT:
// whether `intersect` (KdTree.hs:240-242) would find a hit.
template <int W, bool STATS, bool EXACT = false>
__device__ __forceinline__ bool any_hit_w(const DevScene& sc, const DRay& r, const LdsSlot& slot,
                                          Counts& cn) {
  if (!slab_haskell(sc.root_lo, sc.root_hi, r)) return false;
  f3 inv = mk3(safe_rcp(r.d.x), safe_rcp(r.d.y), safe_rcp(r.d.z));
  f3 oi = mk3(r.o.x * inv.x, r.o.y * inv.y, r.o.z * inv.z);
  int spill_n[kSpill];
  float spill_t[kCull ? kSpill : 1];
  Stack st;
  st.init(slot, spill_n, spill_t);
  return any_w<W, STATS, EXACT>(sc, r, inv, oi, st, cn);
}

// Width-dispatching forms for kernels whose rays may come from scenes of
// different widths; the width is uniform per scene, so the branch is uniform
// per wave where the scene is.
template <bool STATS, bool ROOT_INF = false>
__device__ __forceinline__ void closest_hit(const DevScene& sc, const DRay& r, const LdsSlot& slot,
                                            Best& best, Counts& cn) {
  if (sc.width == 8) closest_hit_w<8, STATS, ROOT_INF>(sc, r, slot, best, cn);
  else if (sc.width == 4) closest_hit_w<4, STATS, ROOT_INF>(sc, r, slot, best, cn);
  else closest_hit_w<2, STATS, ROOT_INF>(sc, r, slot, best, cn);
}
template <bool STATS, bool EXACT = false>
__device__ __forceinline__ bool any_hit(const DevScene& sc, const DRay& r, const LdsSlot& slot,
                                        Counts& cn) {
  if (sc.width == 8) return any_hit_w<8, STATS, EXACT>(sc, r, slot, cn);
  if (sc.width == 4) return any_hit_w<4, STATS, EXACT>(sc, r, slot, cn);
  return any_hit_w<2, STATS, EXACT>(sc, r, slot, cn);
}

}  // namespace trav
}  // namespace bgpu
