// Device BVH traversal shared by the ray-batch kernels and the path tracer.
// One ray per lane.  The traversal stack keeps its top kShort entries in LDS
// (laid out [depth][lane] so a wave's pushes hit 64 distinct banks) and spills
// deeper entries to a private array (scratch); the builder bounds the tree
// depth, hence the stack.  Node boxes are tested with explicit FMAs (the boxes
// are padded outward, so the test only has to be conservative); primitives
// are tested with the reference's exact arithmetic.
#pragma once
#include "isect.h"

namespace bgpu {
namespace trav {

constexpr int kShort = 16;
constexpr int kSpill = 48;  // kShort + kSpill >= tree depth bound (60)

struct Stack {
  int* lds;
  int stride;
  int sp;
  int spill[kSpill];
  __device__ __forceinline__ void push(int v) {
    if (sp < kShort) lds[sp * stride] = v;
    else spill[sp - kShort] = v;
    ++sp;
  }
  __device__ __forceinline__ int pop() {
    --sp;
    return sp < kShort ? lds[sp * stride] : spill[sp - kShort];
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

// Grows a box by k * (L1 distance from o to the box + L1 diagonal).
__device__ __forceinline__ void inflate(float& lx, float& hx, float& ly, float& hy, float& lz,
                                        float& hz, f3 o, float k) {
  float dist = fmaxf(fmaxf(lx - o.x, o.x - hx), 0.0f) + fmaxf(fmaxf(ly - o.y, o.y - hy), 0.0f) +
               fmaxf(fmaxf(lz - o.z, o.z - hz), 0.0f);
  float diag = (hx - lx) + (hy - ly) + (hz - lz);
  float p = k * (dist + diag);
  lx -= p; ly -= p; lz -= p;
  hx += p; hy += p; hz += p;
}

struct Best {
  float t, b1, b2;
  int prim;
};

// Traverses one tree.  CLOSEST: shrinks best.t; ANY: returns at the first hit.
template <bool ANY, bool FAT, bool STATS>
__device__ __forceinline__ bool traverse(const DevScene& sc, int root, const DRay& r, f3 inv,
                                         f3 oi, Stack& st, Best& best, unsigned long long& nvisit,
                                         unsigned long long& ntest) {
  int cur = root;
  st.sp = 0;
  for (;;) {
    if (cur >= 0) {
      const float4* nd = sc.nodes + 4 * cur;
      float4 a = nd[0], b = nd[1], c = nd[2];
      int4 l = reinterpret_cast<const int4*>(nd)[3];
      if (STATS) ++nvisit;
      float tlim = ANY ? r.tmax : best.t;
      bool h0, h1;
      float tn0, tn1;
      if (FAT) {
        float lx = a.x, hx = a.y, ly = a.z, hy = a.w, lz = b.x, hz = b.y;
        inflate(lx, hx, ly, hy, lz, hz, r.o, sc.shape_pad_k);
        slab(lx, hx, ly, hy, lz, hz, inv, oi, r.tmin, tlim, &h0, &tn0);
        lx = b.z; hx = b.w; ly = c.x; hy = c.y; lz = c.z; hz = c.w;
        inflate(lx, hx, ly, hy, lz, hz, r.o, sc.shape_pad_k);
        slab(lx, hx, ly, hy, lz, hz, inv, oi, r.tmin, tlim, &h1, &tn1);
      } else {
        slab(a.x, a.y, a.z, a.w, b.x, b.y, inv, oi, r.tmin, tlim, &h0, &tn0);
        slab(b.z, b.w, c.x, c.y, c.z, c.w, inv, oi, r.tmin, tlim, &h1, &tn1);
      }
      if (h0 && h1) {
        bool first0 = tn0 <= tn1;
        st.push(first0 ? l.y : l.x);
        cur = first0 ? l.x : l.y;
        continue;
      }
      if (h0) { cur = l.x; continue; }
      if (h1) { cur = l.y; continue; }
    } else {
      int enc = ~cur;
      int first = enc >> 4, cnt = enc & 15;
      for (int k = first; k < first + cnt; ++k) {
        const float4* pr = sc.leaf + 3 * k;
        if (STATS) ++ntest;
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
          if (ANY) {
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
    }
    if (st.sp == 0) break;
    cur = st.pop();
  }
  return false;
}


// Closest hit over both trees (triangles, then shapes), after the root-box
// test of the accelerator primitive (Primitive/KdTree.hs:240).
template <bool STATS>
__device__ __forceinline__ void closest_hit(const DevScene& sc, const DRay& r, int* lds, int stride,
                                            Best& best, unsigned long long& nvisit,
                                            unsigned long long& ntest) {
  best.t = r.tmax;
  best.b1 = 0.0f;
  best.b2 = 0.0f;
  best.prim = -1;
  if (!slab_haskell(sc.root_lo, sc.root_hi, r)) return;
  f3 inv = mk3(safe_rcp(r.d.x), safe_rcp(r.d.y), safe_rcp(r.d.z));
  f3 oi = mk3(r.o.x * inv.x, r.o.y * inv.y, r.o.z * inv.z);
  Stack st;
  st.lds = lds;
  st.stride = stride;
  if (sc.tri_root >= 0) traverse<false, false, STATS>(sc, sc.tri_root, r, inv, oi, st, best, nvisit, ntest);
  if (sc.shape_root >= 0) traverse<false, true, STATS>(sc, sc.shape_root, r, inv, oi, st, best, nvisit, ntest);
}

// Any hit (Primitive/KdTree.hs:244-246).
template <bool STATS>
__device__ __forceinline__ bool any_hit(const DevScene& sc, const DRay& r, int* lds, int stride,
                                        unsigned long long& nvisit, unsigned long long& ntest) {
  if (!slab_haskell(sc.root_lo, sc.root_hi, r)) return false;
  f3 inv = mk3(safe_rcp(r.d.x), safe_rcp(r.d.y), safe_rcp(r.d.z));
  f3 oi = mk3(r.o.x * inv.x, r.o.y * inv.y, r.o.z * inv.z);
  Stack st;
  st.lds = lds;
  st.stride = stride;
  Best best{r.tmax, 0.0f, 0.0f, -1};
  if (sc.tri_root >= 0 && traverse<true, false, STATS>(sc, sc.tri_root, r, inv, oi, st, best, nvisit, ntest))
    return true;
  if (sc.shape_root >= 0 && traverse<true, true, STATS>(sc, sc.shape_root, r, inv, oi, st, best, nvisit, ntest))
    return true;
  return false;
}

}  // namespace trav
}  // namespace bgpu
