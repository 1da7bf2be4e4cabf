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

template <int BLOCK, bool STATS>
__global__ __launch_bounds__(BLOCK) void k_closest(DevScene sc, DevRays rays, int n, DevHits out,
                                                   unsigned long long* counters) {
  __shared__ int lds[trav::kShort * BLOCK];
  const int tid = threadIdx.x;
  const int i = blockIdx.x * BLOCK + tid;
  if (i >= n) return;
  DRay r = load_ray(rays, i);
  trav::Best best;
  unsigned long long nvisit = 0, ntest = 0;
  trav::closest_hit<STATS>(sc, r, lds + tid, BLOCK, best, nvisit, ntest);
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
  __shared__ int lds[trav::kShort * BLOCK];
  const int tid = threadIdx.x;
  const int i = blockIdx.x * BLOCK + tid;
  if (i >= n) return;
  DRay r = load_ray(rays, i);
  unsigned long long nvisit = 0, ntest = 0;
  bool found = trav::any_hit<STATS>(sc, r, lds + tid, BLOCK, nvisit, ntest);
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
