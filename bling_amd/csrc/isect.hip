// Ray-batch kernels for gfx950: closest hit (`intersect`,
// Primitive/KdTree.hs:240-242) and any hit (`intersects`, KdTree.hs:244-246)
// over caller-supplied rays.  The traversal itself is in trav.h.
#include "trav.h"

namespace bgpu {

namespace {

__device__ __forceinline__ DRay load_ray(const DevRays& r, int i) {
  DRay ray;
  ray.o = mk3(r.ox[i], r.oy[i], r.oz[i]);
  ray.d = mk3(r.dx[i], r.dy[i], r.dz[i]);
  ray.tmin = r.tmin[i];
  ray.tmax = r.tmax[i];
  return ray;
}

template <int BLOCK, bool STATS, int W>
__global__ __launch_bounds__(BLOCK) void k_closest(DevScene sc, DevRays rays, int n, DevHits out,
                                                   unsigned long long* counters) {
  __shared__ int lds[trav::kLdsPerLane * BLOCK];
  const int tid = threadIdx.x;
  const int i = blockIdx.x * BLOCK + tid;
  if (i >= n) return;
  DRay r = load_ray(rays, i);
  trav::Best best;
  unsigned long long nvisit = 0, ntest = 0;
  trav::closest_hit_w<W, STATS>(sc, r, lds + tid, BLOCK, best, nvisit, ntest);
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

template <int BLOCK, bool STATS, int W>
__global__ __launch_bounds__(BLOCK) void k_any(DevScene sc, DevRays rays, int n, uint8_t* occ,
                                               unsigned long long* counters) {
  __shared__ int lds[trav::kLdsPerLane * BLOCK];
  const int tid = threadIdx.x;
  const int i = blockIdx.x * BLOCK + tid;
  if (i >= n) return;
  DRay r = load_ray(rays, i);
  unsigned long long nvisit = 0, ntest = 0;
  bool found = trav::any_hit_w<W, STATS>(sc, r, lds + tid, BLOCK, nvisit, ntest);
  occ[i] = found ? 1 : 0;
  if (STATS) {
    atomicAdd(&counters[0], nvisit);
    atomicAdd(&counters[1], ntest);
  }
}

template <int BLOCK, int W>
hipError_t run_closest_w(const DevScene& sc, const DevRays& r, int n, const DevHits& h, hipStream_t s,
                         unsigned long long* counters) {
  dim3 grid((n + BLOCK - 1) / BLOCK);
  if (counters)
    hipLaunchKernelGGL((k_closest<BLOCK, true, W>), grid, dim3(BLOCK), 0, s, sc, r, n, h, counters);
  else
    hipLaunchKernelGGL((k_closest<BLOCK, false, W>), grid, dim3(BLOCK), 0, s, sc, r, n, h, counters);
  return hipGetLastError();
}
template <int BLOCK>
hipError_t run_closest(const DevScene& sc, const DevRays& r, int n, const DevHits& h, hipStream_t s,
                       unsigned long long* counters) {
  if (sc.width == 8) return run_closest_w<BLOCK, 8>(sc, r, n, h, s, counters);
  if (sc.width == 4) return run_closest_w<BLOCK, 4>(sc, r, n, h, s, counters);
  return run_closest_w<BLOCK, 2>(sc, r, n, h, s, counters);
}

template <int BLOCK, int W>
hipError_t run_any_w(const DevScene& sc, const DevRays& r, int n, uint8_t* occ, hipStream_t s,
                     unsigned long long* counters) {
  dim3 grid((n + BLOCK - 1) / BLOCK);
  if (counters)
    hipLaunchKernelGGL((k_any<BLOCK, true, W>), grid, dim3(BLOCK), 0, s, sc, r, n, occ, counters);
  else
    hipLaunchKernelGGL((k_any<BLOCK, false, W>), grid, dim3(BLOCK), 0, s, sc, r, n, occ, counters);
  return hipGetLastError();
}
template <int BLOCK>
hipError_t run_any(const DevScene& sc, const DevRays& r, int n, uint8_t* occ, hipStream_t s,
                   unsigned long long* counters) {
  if (sc.width == 8) return run_any_w<BLOCK, 8>(sc, r, n, occ, s, counters);
  if (sc.width == 4) return run_any_w<BLOCK, 4>(sc, r, n, occ, s, counters);
  return run_any_w<BLOCK, 2>(sc, r, n, occ, s, counters);
}

// ---------------------------------------------------------------- persistent
// One pass over one tree (triangles: FAT false; shapes: FAT true) with lanes
// that take a new ray as soon as theirs is done, so a wave does not idle on
// its slowest ray.  Each wave takes slices of kSlice rays from a global
// counter.  A step first runs inner nodes until every lane of the wave is at
// a leaf or done, then tests the leaves together.
//   triangle pass: root-box test (KdTree.hs:240), then the triangle tree;
//                  writes every ray's result.
//   shape pass:    continues from the triangle pass's result (closest: its
//                  hit bounds the search; any: occluded rays are done).
#ifndef BGPU_REFILL_MIN
#define BGPU_REFILL_MIN 32
#endif
#ifndef BGPU_SLICE
#define BGPU_SLICE 128
#endif
constexpr unsigned int kSlice = BGPU_SLICE;
// idle lanes a wave waits for before taking new rays (all lanes idle always
// refills), so that new rays arrive in groups of neighbours
constexpr int kRefillMin = BGPU_REFILL_MIN;

template <int BLOCK, int W, bool ANY, bool FAT, bool STATS>
__global__ __launch_bounds__(BLOCK) void k_persist(DevScene sc, DevRays rays, int n, DevHits out,
                                                   uint8_t* occ, unsigned int* work,
                                                   unsigned long long* counters) {
  __shared__ int lds[trav::kLdsPerLane * BLOCK];
  const int lane = __lane_id();
  const int root = FAT ? sc.shape_root : sc.tri_root;
  trav::Stack st;
  st.lds = lds + threadIdx.x;
  st.stride = BLOCK;
  st.sp = 0;
  int ray = -1, cur = kNoChild;
  DRay r;
  f3 inv = mk3(0.0f, 0.0f, 0.0f), oi = inv;
  trav::Best best{0.0f, 0.0f, 0.0f, -1};
  unsigned long long nvisit = 0, ntest = 0;
  unsigned int s_next = 0, s_end = 0;  // the wave's slice (uniform)
  bool drained = false;
  for (;;) {
    // refill idle lanes from the wave's slice, taking a new slice when empty
    const int n_idle = __popcll(__ballot(ray < 0));
    if (!drained && n_idle > 0 && (n_idle >= kRefillMin || n_idle == 64)) {
      if (s_next >= s_end) {
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
        const unsigned long long m = __ballot(ray < 0);
        const unsigned int pos = s_next + (unsigned int)__popcll(m & ((1ull << lane) - 1ull));
        if (ray < 0 && pos < s_end) {
          ray = (int)pos;
          r = trav::load_ray(rays, ray);
          bool live = slab_haskell(sc.root_lo, sc.root_hi, r) && root >= 0;
          if (FAT) {
            if (ANY) {
              live = live && occ[ray] == 0;
            } else {
              best.prim = out.prim[ray];
              best.t = best.prim >= 0 ? out.t[ray] : r.tmax;
              best.b1 = best.prim >= 0 ? out.b1[ray] : 0.0f;
              best.b2 = best.prim >= 0 ? out.b2[ray] : 0.0f;
            }
          } else {
            best = trav::Best{r.tmax, 0.0f, 0.0f, -1};
          }
          inv = mk3(trav::safe_rcp(r.d.x), trav::safe_rcp(r.d.y), trav::safe_rcp(r.d.z));
          oi = mk3(r.o.x * inv.x, r.o.y * inv.y, r.o.z * inv.z);
          st.sp = 0;
          cur = live ? root : kNoChild;
        }
        const unsigned int took = (unsigned int)__popcll(m);
        s_next = s_next + took < s_end ? s_next + took : s_end;
      }
    }
    if (__ballot(ray >= 0) == 0) {
      if (drained) break;
      continue;
    }
    // inner nodes until every lane is at a leaf or done
    while (ray >= 0 && cur >= 0 && cur != kNoChild) {
      const float4* nd = sc.nodes + (size_t)cur * (2 * W);
      float v[8 * W];
#pragma unroll
      for (int q = 0; q < (7 * W + 3) / 4; ++q) {
        const float4 x = nd[q];
        v[4 * q] = x.x; v[4 * q + 1] = x.y; v[4 * q + 2] = x.z; v[4 * q + 3] = x.w;
      }
      if (STATS) ++nvisit;
      const float tlim = ANY ? r.tmax : best.t;
      const float pad = FAT ? trav::node_margin<W>(v, r.o, sc.shape_pad_k) : 0.0f;
      float key[W];
      int lk[W];
      int nhit = 0;
#pragma unroll
      for (int c = 0; c < W; ++c) {
        float lx = v[c], hx = v[W + c], ly = v[2 * W + c], hy = v[3 * W + c], lz = v[4 * W + c],
              hz = v[5 * W + c];
        lk[c] = __float_as_int(v[6 * W + c]);
        if (FAT) trav::inflate(lx, hx, ly, hy, lz, hz, pad);
        bool h;
        float tn;
        trav::slab(lx, hx, ly, hy, lz, hz, inv, oi, r.tmin, tlim, &h, &tn);
        h = h && lk[c] != kNoChild;
        key[c] = h ? fminf(tn, 3.0e38f) : INFINITY;
        nhit += h ? 1 : 0;
      }
      if (nhit > 0) {
        trav::sort_children<W>(key, lk);
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
    }
    // leaves
    if (ray >= 0 && cur < 0) {
      const int enc = ~cur;
      const int first = enc >> 4, cnt = enc & 15;
      bool hit_any = false;
      for (int k = first; k < first + cnt; ++k) {
        const float4* pr = sc.leaf + 3 * k;
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
      } else if (st.sp > 0) {
        float t;
        st.pop(&cur, &t);
      } else {
        cur = kNoChild;
      }
    }
    // finished rays write their result and free the lane
    if (ray >= 0 && cur == kNoChild) {
      if (ANY) {
        if (!FAT || best.prim >= 0) occ[ray] = best.prim >= 0 ? 1 : 0;
      } else if (!FAT || best.prim >= 0) {
        const bool h = best.prim >= 0;
        out.t[ray] = h ? best.t : 0.0f;
        out.b1[ray] = h ? best.b1 : 0.0f;
        out.b2[ray] = h ? best.b2 : 0.0f;
        out.prim[ray] = best.prim;
      }
      ray = -1;
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
    if (lane == 0) {
      atomicAdd(&counters[0], a);
      atomicAdd(&counters[1], b);
    }
  }
}

// Blocks that fill the device once for a persistent kernel.
template <class K>
int persistent_grid(K kernel, int block) {
  int dev = 0, per_cu = 1;
  (void)hipGetDevice(&dev);
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, dev) != hipSuccess) return 1024;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, block, 0) != hipSuccess || per_cu < 1)
    per_cu = 1;
  return per_cu * prop.multiProcessorCount;
}

template <int BLOCK, int W, bool ANY, bool STATS>
hipError_t run_persist(const DevScene& sc, const DevRays& r, int n, const DevHits& h, uint8_t* occ,
                       unsigned int* work, hipStream_t s, unsigned long long* counters) {
  // one counter per pass, zeroed together
  hipError_t e = hipMemsetAsync(work, 0, 2 * sizeof(unsigned int), s);
  if (e != hipSuccess) return e;
  static int g_tri = 0, g_shape = 0;
  if (!g_tri) g_tri = persistent_grid(k_persist<BLOCK, W, ANY, false, STATS>, BLOCK);
  if (!g_shape) g_shape = persistent_grid(k_persist<BLOCK, W, ANY, true, STATS>, BLOCK);
  hipLaunchKernelGGL((k_persist<BLOCK, W, ANY, false, STATS>), dim3(g_tri), dim3(BLOCK), 0, s, sc, r, n, h,
                     occ, work, counters);
  if (sc.shape_root >= 0)
    hipLaunchKernelGGL((k_persist<BLOCK, W, ANY, true, STATS>), dim3(g_shape), dim3(BLOCK), 0, s, sc, r, n, h,
                       occ, work + 1, counters);
  return hipGetLastError();
}

template <int BLOCK, bool ANY>
hipError_t run_persist_any_w(const DevScene& sc, const DevRays& r, int n, const DevHits& h, uint8_t* occ,
                             unsigned int* work, hipStream_t s, unsigned long long* counters) {
  if (counters) {
    if (sc.width == 8) return run_persist<BLOCK, 8, ANY, true>(sc, r, n, h, occ, work, s, counters);
    if (sc.width == 4) return run_persist<BLOCK, 4, ANY, true>(sc, r, n, h, occ, work, s, counters);
    return run_persist<BLOCK, 2, ANY, true>(sc, r, n, h, occ, work, s, counters);
  }
  if (sc.width == 8) return run_persist<BLOCK, 8, ANY, false>(sc, r, n, h, occ, work, s, counters);
  if (sc.width == 4) return run_persist<BLOCK, 4, ANY, false>(sc, r, n, h, occ, work, s, counters);
  return run_persist<BLOCK, 2, ANY, false>(sc, r, n, h, occ, work, s, counters);
}

}  // namespace

hipError_t launch_closest(const DevScene& sc, const DevRays& r, int n, const DevHits& h,
                          const TraceConfig& cfg, hipStream_t s, unsigned long long* counters,
                          unsigned int* work) {
  if (n <= 0) return hipSuccess;
  if (cfg.persistent && work) {
    switch (cfg.block) {
      case 64: return run_persist_any_w<64, false>(sc, r, n, h, nullptr, work, s, counters);
      case 256: return run_persist_any_w<256, false>(sc, r, n, h, nullptr, work, s, counters);
      default: return run_persist_any_w<128, false>(sc, r, n, h, nullptr, work, s, counters);
    }
  }
  switch (cfg.block) {
    case 64: return run_closest<64>(sc, r, n, h, s, counters);
    case 256: return run_closest<256>(sc, r, n, h, s, counters);
    default: return run_closest<128>(sc, r, n, h, s, counters);
  }
}

hipError_t launch_any(const DevScene& sc, const DevRays& r, int n, uint8_t* occ,
                      const TraceConfig& cfg, hipStream_t s, unsigned long long* counters,
                      unsigned int* work) {
  if (n <= 0) return hipSuccess;
  if (cfg.persistent && work) {
    const DevHits none{nullptr, nullptr, nullptr, nullptr};
    switch (cfg.block) {
      case 64: return run_persist_any_w<64, true>(sc, r, n, none, occ, work, s, counters);
      case 256: return run_persist_any_w<256, true>(sc, r, n, none, occ, work, s, counters);
      default: return run_persist_any_w<128, true>(sc, r, n, none, occ, work, s, counters);
    }
  }
  switch (cfg.block) {
    case 64: return run_any<64>(sc, r, n, occ, s, counters);
    case 256: return run_any<256>(sc, r, n, occ, s, counters);
    default: return run_any<128>(sc, r, n, occ, s, counters);
  }
}

}  // namespace bgpu
