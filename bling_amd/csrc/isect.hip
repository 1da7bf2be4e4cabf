// BVH traversal kernels for gfx950: closest hit (`intersect`,
// Primitive/KdTree.hs:240-242) and any hit (`intersects`, KdTree.hs:244-246).
//
// One ray per lane.  The traversal stack keeps its top kShort entries in LDS
// (laid out [depth][lane] so a wave's pushes hit 64 distinct banks) and spills
// deeper entries to a private array (scratch); the builder bounds the tree
// depth, hence the stack.  Node boxes are tested with explicit FMAs (the boxes
// are padded outward, so the test only has to be conservative); primitives
// are tested with the reference's exact arithmetic.  Triangles and shapes are
// in separate trees (see isect.h); the closest-hit result is the minimum t
// over both, ties going to the lowest primitive id.
#include "isect.h"

namespace bgpu {

namespace {

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

template <int BLOCK, bool STATS>
__global__ __launch_bounds__(BLOCK) void k_closest(DevScene sc, DevRays rays, int n, DevHits out,
                                                   unsigned long long* counters) {
  __shared__ int lds[kShort * BLOCK];
  const int tid = threadIdx.x;
  const int i = blockIdx.x * BLOCK + tid;
  if (i >= n) return;
  DRay r = load_ray(rays, i);
  Best best{r.tmax, 0.0f, 0.0f, -1};
  unsigned long long nvisit = 0, ntest = 0;
  if (slab_haskell(sc.root_lo, sc.root_hi, r)) {  // KdTree.hs:240
    f3 inv = mk3(safe_rcp(r.d.x), safe_rcp(r.d.y), safe_rcp(r.d.z));
    f3 oi = mk3(r.o.x * inv.x, r.o.y * inv.y, r.o.z * inv.z);
    Stack st;
    st.lds = lds + tid;
    st.stride = BLOCK;
    if (sc.tri_root >= 0)
      traverse<false, false, STATS>(sc, sc.tri_root, r, inv, oi, st, best, nvisit, ntest);
    if (sc.shape_root >= 0)
      traverse<false, true, STATS>(sc, sc.shape_root, r, inv, oi, st, best, nvisit, ntest);
  }
  if (best.prim >= 0) {
    out.t[i] = best.t;
    out.b1[i] = best.b1;
    out.b2[i] = best.b2;
    out.prim[i] = best.prim;
  } else {
    out.t[i] = 0.0f;
    out.b1[i] = 0.0f;
    out.b2[i] = 0.0f;
    out.prim[i] = -1;
  }
  if (STATS) {
    atomicAdd(&counters[0], nvisit);
    atomicAdd(&counters[1], ntest);
  }
}

template <int BLOCK, bool STATS>
__global__ __launch_bounds__(BLOCK) void k_any(DevScene sc, DevRays rays, int n, uint8_t* occ,
                                               unsigned long long* counters) {
  __shared__ int lds[kShort * BLOCK];
  const int tid = threadIdx.x;
  const int i = blockIdx.x * BLOCK + tid;
  if (i >= n) return;
  DRay r = load_ray(rays, i);
  bool found = false;
  unsigned long long nvisit = 0, ntest = 0;
  if (slab_haskell(sc.root_lo, sc.root_hi, r)) {  // KdTree.hs:244
    f3 inv = mk3(safe_rcp(r.d.x), safe_rcp(r.d.y), safe_rcp(r.d.z));
    f3 oi = mk3(r.o.x * inv.x, r.o.y * inv.y, r.o.z * inv.z);
    Stack st;
    st.lds = lds + tid;
    st.stride = BLOCK;
    Best best{r.tmax, 0.0f, 0.0f, -1};
    if (sc.tri_root >= 0)
      found = traverse<true, false, STATS>(sc, sc.tri_root, r, inv, oi, st, best, nvisit, ntest);
    if (!found && sc.shape_root >= 0)
      found = traverse<true, true, STATS>(sc, sc.shape_root, r, inv, oi, st, best, nvisit, ntest);
  }
  occ[i] = found ? 1 : 0;
  if (STATS) {
    atomicAdd(&counters[0], nvisit);
    atomicAdd(&counters[1], ntest);
  }
}

template <int BLOCK>
hipError_t run_closest(const DevScene& sc, const DevRays& r, int n, const DevHits& h, hipStream_t s,
                       unsigned long long* counters) {
  dim3 grid((n + BLOCK - 1) / BLOCK);
  if (counters)
    hipLaunchKernelGGL((k_closest<BLOCK, true>), grid, dim3(BLOCK), 0, s, sc, r, n, h, counters);
  else
    hipLaunchKernelGGL((k_closest<BLOCK, false>), grid, dim3(BLOCK), 0, s, sc, r, n, h, counters);
  return hipGetLastError();
}

template <int BLOCK>
hipError_t run_any(const DevScene& sc, const DevRays& r, int n, uint8_t* occ, hipStream_t s,
                   unsigned long long* counters) {
  dim3 grid((n + BLOCK - 1) / BLOCK);
  if (counters)
    hipLaunchKernelGGL((k_any<BLOCK, true>), grid, dim3(BLOCK), 0, s, sc, r, n, occ, counters);
  else
    hipLaunchKernelGGL((k_any<BLOCK, false>), grid, dim3(BLOCK), 0, s, sc, r, n, occ, counters);
  return hipGetLastError();
}

}  // namespace

hipError_t launch_closest(const DevScene& sc, const DevRays& r, int n, const DevHits& h,
                          const TraceConfig& cfg, hipStream_t s, unsigned long long* counters) {
  if (n <= 0) return hipSuccess;
  switch (cfg.block) {
    case 64: return run_closest<64>(sc, r, n, h, s, counters);
    case 256: return run_closest<256>(sc, r, n, h, s, counters);
    default: return run_closest<128>(sc, r, n, h, s, counters);
  }
}

hipError_t launch_any(const DevScene& sc, const DevRays& r, int n, uint8_t* occ,
                      const TraceConfig& cfg, hipStream_t s, unsigned long long* counters) {
  if (n <= 0) return hipSuccess;
  switch (cfg.block) {
    case 64: return run_any<64>(sc, r, n, occ, s, counters);
    case 256: return run_any<256>(sc, r, n, occ, s, counters);
    default: return run_any<128>(sc, r, n, occ, s, counters);
  }
}

}  // namespace bgpu
