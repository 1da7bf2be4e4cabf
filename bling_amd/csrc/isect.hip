// Ray-batch kernels for gfx950: closest hit (`intersect`,
// Primitive/KdTree.hs:240-242) and any hit (`intersects`, KdTree.hs:244-246)
// over caller-supplied rays.  The traversal itself is in trav.h.
#include "trav_persist.h"

namespace bgpu {

// Scratch of bgpu_intersect_closest_noalloc: the two pass counters of the
// persistent traversal (words 0-1) and, for BGPU_COUNT, the four 64-bit
// work counters (words 8-15).  One copy per GPU, for the
// life of the process; host code reaches it only through the address
// hipGetSymbolAddress returns on that device (scratch_address).
static __device__ __align__(256) unsigned int scratch[kScratchWords];

hipError_t scratch_address(void** ptr) { return hipGetSymbolAddress(ptr, HIP_SYMBOL(scratch)); }

namespace {

// Zeroes n words (a kernel rather than a memset node, so that a caller
// capturing the launches into a graph replays it correctly).
__global__ void k_zero_words(unsigned int* p, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0u;
}

__device__ __forceinline__ DRay load_ray(const DevRays& r, int i) {
  DRay ray;
  ray.o = mk3(r.ox[i], r.oy[i], r.oz[i]);
  ray.d = mk3(r.dx[i], r.dy[i], r.dz[i]);
  ray.tmin = r.tmin[i];
  ray.tmax = r.tmax[i];
  return ray;
}

// Adds a lane's work counts to counters[0..3]: triangle-tree node fetches
// and tests, then shape-tree node fetches and tests.
__device__ __forceinline__ void add_counts(unsigned long long* counters, const trav::Counts& cn) {
  atomicAdd(&counters[0], cn.nodes[0]);
  atomicAdd(&counters[1], cn.tests[0]);
  atomicAdd(&counters[2], cn.nodes[1]);
  atomicAdd(&counters[3], cn.tests[1]);
}

template <int BLOCK, bool STATS, int W>
__global__ __launch_bounds__(BLOCK) void k_closest(DevScene sc, DevRays rays, int n, DevHits out,
                                                   unsigned long long* counters, int nshort, int walk) {
  const int tid = threadIdx.x;
  const int i = blockIdx.x * BLOCK + tid;
  if (i >= n) return;
  DRay r = load_ray(rays, i);
  trav::Best best;
  trav::Counts cn;
  trav::closest_hit_w<W, STATS>(sc, r, trav::lane_slot<BLOCK>(nshort, walk), best, cn);
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
  if (STATS) add_counts(counters, cn);
}

template <int BLOCK, bool STATS, int W>
__global__ __launch_bounds__(BLOCK) void k_any(DevScene sc, DevRays rays, int n, uint8_t* occ,
                                               unsigned long long* counters, int nshort, int walk) {
  const int tid = threadIdx.x;
  const int i = blockIdx.x * BLOCK + tid;
  if (i >= n) return;
  DRay r = load_ray(rays, i);
  trav::Counts cn;
  bool found = trav::any_hit_w<W, STATS>(sc, r, trav::lane_slot<BLOCK>(nshort, walk), cn);
  occ[i] = found ? 1 : 0;
  if (STATS) add_counts(counters, cn);
}

template <int BLOCK, int W>
hipError_t run_closest_w(const DevScene& sc, const DevRays& r, int n, const DevHits& h, hipStream_t s,
                         unsigned long long* counters, int nshort, int walk) {
  dim3 grid((n + BLOCK - 1) / BLOCK);
  const size_t lds = trav::stack_lds_bytes(nshort, BLOCK);
  if (counters)
    hipLaunchKernelGGL((k_closest<BLOCK, true, W>), grid, dim3(BLOCK), lds, s, sc, r, n, h, counters, nshort,
                       walk);
  else
    hipLaunchKernelGGL((k_closest<BLOCK, false, W>), grid, dim3(BLOCK), lds, s, sc, r, n, h, counters, nshort,
                       walk);
  return hipGetLastError();
}
template <int BLOCK>
hipError_t run_closest(const DevScene& sc, const DevRays& r, int n, const DevHits& h, hipStream_t s,
                       unsigned long long* counters, int nshort, int walk) {
  if (sc.width == 8) return run_closest_w<BLOCK, 8>(sc, r, n, h, s, counters, nshort, walk);
  if (sc.width == 4) return run_closest_w<BLOCK, 4>(sc, r, n, h, s, counters, nshort, walk);
  return run_closest_w<BLOCK, 2>(sc, r, n, h, s, counters, nshort, walk);
}

template <int BLOCK, int W>
hipError_t run_any_w(const DevScene& sc, const DevRays& r, int n, uint8_t* occ, hipStream_t s,
                     unsigned long long* counters, int nshort, int walk) {
  dim3 grid((n + BLOCK - 1) / BLOCK);
  const size_t lds = trav::stack_lds_bytes(nshort, BLOCK);
  if (counters)
    hipLaunchKernelGGL((k_any<BLOCK, true, W>), grid, dim3(BLOCK), lds, s, sc, r, n, occ, counters, nshort,
                       walk);
  else
    hipLaunchKernelGGL((k_any<BLOCK, false, W>), grid, dim3(BLOCK), lds, s, sc, r, n, occ, counters, nshort,
                       walk);
  return hipGetLastError();
}
template <int BLOCK>
hipError_t run_any(const DevScene& sc, const DevRays& r, int n, uint8_t* occ, hipStream_t s,
                   unsigned long long* counters, int nshort, int walk) {
  if (sc.width == 8) return run_any_w<BLOCK, 8>(sc, r, n, occ, s, counters, nshort, walk);
  if (sc.width == 4) return run_any_w<BLOCK, 4>(sc, r, n, occ, s, counters, nshort, walk);
  return run_any_w<BLOCK, 2>(sc, r, n, occ, s, counters, nshort, walk);
}

// ---------------------------------------------------------------- persistent
// Ray batches through the persistent traversal (trav_persist.h), as two
// passes per launch:
//   triangle pass: root-box test (KdTree.hs:240), then the triangle tree;
//                  writes every ray's result;
//   shape pass:    continues from the triangle pass's result (closest: its
//                  hit bounds the search; any: occluded rays are done).
template <bool ANY, bool FAT>
struct BatchSrc {
  const DevScene& sc;
  DevRays rays;
  DevHits out;
  uint8_t* occ;
  __device__ __forceinline__ bool begin(int k, DRay& r, trav::LaneScene& ls, trav::Best& best) {
    r = trav::load_ray(rays, k);
    ls = trav::LaneScene{sc.nodes, sc.leaf, sc.shapes, sc.shape_pad_k, sc.shape_floor_k,
                         FAT ? sc.shape_root : sc.tri_root};
    bool live = slab_haskell(sc.root_lo, sc.root_hi, r);
    best = trav::Best{r.tmax, 0.0f, 0.0f, -1};
    if (FAT) {
      if (ANY) {
        live = live && occ[k] == 0;
      } else {
        best.prim = out.prim[k];
        if (best.prim >= 0) best = trav::Best{out.t[k], out.b1[k], out.b2[k], best.prim};
      }
    }
    return live;
  }
  __device__ __forceinline__ void end(int k, const DRay&, const trav::LaneScene&, const trav::Best& best) {
    const bool h = best.prim >= 0;
    if (ANY) {
      if (!FAT || h) occ[k] = h ? 1 : 0;
    } else if (!FAT || h) {
      out.t[k] = h ? best.t : 0.0f;
      out.b1[k] = h ? best.b1 : 0.0f;
      out.b2[k] = h ? best.b2 : 0.0f;
      out.prim[k] = best.prim;
    }
  }
};

template <int BLOCK, int W, bool ANY, bool FAT, bool STATS>
__global__ __launch_bounds__(BLOCK) void k_persist(DevScene sc, DevRays rays, int n, DevHits out,
                                                   uint8_t* occ, unsigned int* work,
                                                   unsigned long long* counters, int nshort,
                                                   trav::PersistTune tune) {
  BatchSrc<ANY, FAT> src{sc, rays, out, occ};
  trav::persist_trace<W, ANY, FAT, STATS>(src, n, work, trav::lane_slot<BLOCK>(nshort, tune.walk_min), counters,
                                          tune);
}

template <int BLOCK, int W, bool ANY, bool STATS>
hipError_t run_persist(const DevScene& sc, const DevRays& r, int n, const DevHits& h, uint8_t* occ,
                       unsigned int* work, hipStream_t s, unsigned long long* counters, int nshort,
                       trav::PersistTune tune) {
  // one counter per pass, zeroed together
  hipLaunchKernelGGL(k_zero_words, dim3(1), dim3(64), 0, s, work, 2);
  static int g_tri[trav::kMaxShort + 1] = {}, g_shape[trav::kMaxShort + 1] = {};
  const int gt = trav::persistent_grid_cached(g_tri, k_persist<BLOCK, W, ANY, false, STATS>, BLOCK, nshort);
  const int gs = trav::persistent_grid_cached(g_shape, k_persist<BLOCK, W, ANY, true, STATS>, BLOCK, nshort);
  const size_t lds = trav::stack_lds_bytes(nshort, BLOCK);
  hipLaunchKernelGGL((k_persist<BLOCK, W, ANY, false, STATS>), dim3(gt), dim3(BLOCK), lds, s, sc, r, n, h,
                     occ, work, counters, nshort, tune);
  if (sc.shape_root >= 0)
    hipLaunchKernelGGL((k_persist<BLOCK, W, ANY, true, STATS>), dim3(gs), dim3(BLOCK), lds, s, sc, r, n, h,
                       occ, work + 1, counters, nshort, tune);
  return hipGetLastError();
}

template <int BLOCK, bool ANY>
hipError_t run_persist_any_w(const DevScene& sc, const DevRays& r, int n, const DevHits& h, uint8_t* occ,
                             unsigned int* work, hipStream_t s, unsigned long long* counters, int nshort,
                             trav::PersistTune tune) {
  if (counters) {
    if (sc.width == 8) return run_persist<BLOCK, 8, ANY, true>(sc, r, n, h, occ, work, s, counters, nshort, tune);
    if (sc.width == 4) return run_persist<BLOCK, 4, ANY, true>(sc, r, n, h, occ, work, s, counters, nshort, tune);
    return run_persist<BLOCK, 2, ANY, true>(sc, r, n, h, occ, work, s, counters, nshort, tune);
  }
  if (sc.width == 8) return run_persist<BLOCK, 8, ANY, false>(sc, r, n, h, occ, work, s, counters, nshort, tune);
  if (sc.width == 4) return run_persist<BLOCK, 4, ANY, false>(sc, r, n, h, occ, work, s, counters, nshort, tune);
  return run_persist<BLOCK, 2, ANY, false>(sc, r, n, h, occ, work, s, counters, nshort, tune);
}

}  // namespace

hipError_t zero_words(unsigned int* p, int n, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_zero_words, dim3((n + 255) / 256), dim3(256), 0, s, p, n);
  return hipGetLastError();
}

hipError_t launch_closest(const DevScene& sc, const DevRays& r, int n, const DevHits& h,
                          const TraceConfig& cfg, hipStream_t s, unsigned long long* counters,
                          unsigned int* work) {
  if (n <= 0) return hipSuccess;
  const int ns = cfg.stack;
  if (ns < trav::kMinShort || ns > trav::kMaxShort) return hipErrorInvalidValue;
  if (cfg.refill < 1 || cfg.refill > 64 || cfg.walk < 0 || cfg.walk > 64) return hipErrorInvalidValue;
  trav::PersistTune tune;
  tune.refill_min = cfg.refill;
  tune.walk_min = cfg.walk;
  if (cfg.persistent && work) {
    switch (cfg.block) {
      case 64: return run_persist_any_w<64, false>(sc, r, n, h, nullptr, work, s, counters, ns, tune);
      case 256: return run_persist_any_w<256, false>(sc, r, n, h, nullptr, work, s, counters, ns, tune);
      default: return run_persist_any_w<128, false>(sc, r, n, h, nullptr, work, s, counters, ns, tune);
    }
  }
  switch (cfg.block) {
    case 64: return run_closest<64>(sc, r, n, h, s, counters, ns, cfg.walk);
    case 256: return run_closest<256>(sc, r, n, h, s, counters, ns, cfg.walk);
    default: return run_closest<128>(sc, r, n, h, s, counters, ns, cfg.walk);
  }
}

hipError_t launch_any(const DevScene& sc, const DevRays& r, int n, uint8_t* occ,
                      const TraceConfig& cfg, hipStream_t s, unsigned long long* counters,
                      unsigned int* work) {
  if (n <= 0) return hipSuccess;
  const int ns = cfg.stack;
  if (ns < trav::kMinShort || ns > trav::kMaxShort) return hipErrorInvalidValue;
  if (cfg.refill < 1 || cfg.refill > 64 || cfg.walk < 0 || cfg.walk > 64) return hipErrorInvalidValue;
  trav::PersistTune tune;
  tune.refill_min = cfg.refill;
  tune.walk_min = cfg.walk;
  if (cfg.persistent && work) {
    const DevHits none{nullptr, nullptr, nullptr, nullptr};
    switch (cfg.block) {
      case 64: return run_persist_any_w<64, true>(sc, r, n, none, occ, work, s, counters, ns, tune);
      case 256: return run_persist_any_w<256, true>(sc, r, n, none, occ, work, s, counters, ns, tune);
      default: return run_persist_any_w<128, true>(sc, r, n, none, occ, work, s, counters, ns, tune);
    }
  }
  switch (cfg.block) {
    case 64: return run_any<64>(sc, r, n, occ, s, counters, ns, cfg.walk);
    case 256: return run_any<256>(sc, r, n, occ, s, counters, ns, cfg.walk);
    default: return run_any<128>(sc, r, n, occ, s, counters, ns, cfg.walk);
  }
}

}  // namespace bgpu
