// Wavefront path tracer kernels for gfx950 (Integrator/Path.hs:41-87 as
// queues).  One bounce iteration:
//   reset     (1 thread)  totals += queue counts, queues := 0, retire consumed jobs
//   generate  (P slots)   free slots take the next work item (job, pass, pixel,
//                         sample) and fire a camera ray; every active slot is
//                         appended to the active queue (wave ballot + prefix)
//   extend    (active)    closest hit of the path ray
//   shade     (active)    emission, next-event estimation with MIS (one shadow
//                         ray + one BSDF-sampled MIS ray per vertex, Scene.hs:61-118),
//                         Russian roulette and the continuation sample
//   shadow    (queue)     any hit; unoccluded adds the light-sample contribution
//   mis       (queue)     closest hit; hitting the sampled light (or escaping to
//                         an infinite light) adds the BSDF-sample contribution
//   finish    (active)    terminated paths splat into their job's film
// Kernel grids cover the pool; queue lengths are read on the device, so the
// sequence replays from a captured hipGraph without host round trips.
#include "pt_device.h"
#include "pt_material.h"
#include "pt_texture.h"
#include "trav_persist.h"

namespace bgpu {

using namespace sh;

namespace {

// Same-address atomics serialise at the memory side (device-scope atomics on
// a chip of several L2 slices), so queue appends and counters are aggregated
// per block: one atomic per block and counter, positions from an LDS prefix.

// Block-wide exclusive prefix of `pred` plus one atomicAdd on `counter` for
// the block total; returns the lane's position.  Every thread of the block
// calls it (it synchronises the block); `s` holds BLOCK/64 + 1 entries.
template <int BLOCK, class T>
__device__ __forceinline__ T block_append(T* counter, bool pred, T* s) {
  constexpr int NW = BLOCK / 64;
  const int lane = __lane_id(), wave = threadIdx.x >> 6;
  const unsigned long long m = __ballot(pred);
  if (lane == 0) s[wave] = (T)__popcll(m);
  __syncthreads();
  if (threadIdx.x == 0) {
    T run = 0;
    for (int w = 0; w < NW; ++w) {
      T c = s[w];
      s[w] = run;
      run += c;
    }
    s[NW] = run ? atomicAdd(counter, run) : (T)0;
  }
  __syncthreads();
  T r = s[NW] + s[wave] + (T)__popcll(m & ((1ull << lane) - 1ull));
  __syncthreads();  // `s` may be reused by the caller
  return r;
}

// Block-wide count of `pred` added to `counter` with one atomic.
template <int BLOCK, class T>
__device__ __forceinline__ void block_count(T* counter, bool pred, T* s) {
  constexpr int NW = BLOCK / 64;
  const int lane = __lane_id(), wave = threadIdx.x >> 6;
  const unsigned long long m = __ballot(pred);
  if (lane == 0) s[wave] = (T)__popcll(m);
  __syncthreads();
  if (threadIdx.x == 0) {
    T run = 0;
    for (int w = 0; w < NW; ++w) run += s[w];
    if (run) atomicAdd(counter, run);
  }
  __syncthreads();
}

// A path's radiance is carried as unnormalised CIE XYZ sums
// (sum_b cie[b] * L[b]); the film divides by the Y integral.  XYZ is linear in
// the spectrum, so summing the XYZ of each contribution equals the XYZ of the
// summed spectrum up to float rounding.
struct Xyz {
  float x = 0.0f, y = 0.0f, z = 0.0f;
  __device__ __forceinline__ void add(int b, float c) {
    x = x + cCieX[b] * c;
    y = y + cCieY[b] * c;
    z = z + cCieZ[b] * c;
  }
};
__device__ __forceinline__ void xyz_store(float* a, int Pn, int i, const Xyz& v) {
  a[i] = v.x;
  a[Pn + i] = v.y;
  a[2 * Pn + i] = v.z;
}
__device__ __forceinline__ void xyz_accum(float* a, int Pn, int i, const Xyz& v) {
  a[i] += v.x;
  a[Pn + i] += v.y;
  a[2 * Pn + i] += v.z;
}
__device__ __forceinline__ void xyz_accum(float* a, int Pn, int i, const float* src) {
  a[i] += src[i];
  a[Pn + i] += src[Pn + i];
  a[2 * Pn + i] += src[2 * Pn + i];
}

// A ray stored in the pool as 7 (tmax given) or 8 SoA words.
__device__ __forceinline__ DRay load_pool_ray(const float* w, int Pn, int i, float tmax) {
  DRay r;
  r.o = mk3(w[i], w[Pn + i], w[2 * Pn + i]);
  r.d = mk3(w[3 * Pn + i], w[4 * Pn + i], w[5 * Pn + i]);
  r.tmin = w[6 * Pn + i];
  r.tmax = tmax;
  return r;
}
__device__ __forceinline__ DRay load_pool_ray(const float* w, int Pn, int i) {
  return load_pool_ray(w, Pn, i, w[7 * Pn + i]);
}

__device__ __forceinline__ unsigned int wave_sum(unsigned int v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
  return v;
}

// Adds 1 to *ctr for every lane with `pred`, merging equal counters: each
// wave groups its lanes by counter address (up to kGroups groups go to LDS,
// further ones straight to memory), then wave 0 merges the block's groups and
// issues one atomic per distinct counter.  Every thread of the block calls it.
constexpr int kGroups = 4;
template <int BLOCK>
__device__ __forceinline__ void block_increment(unsigned int* ctr, bool pred, unsigned long long* s_key,
                                                unsigned int* s_cnt, int* s_n) {
  constexpr int NW = BLOCK / 64;
  static_assert(NW * kGroups <= 64, "one merge entry per lane of wave 0");
  const int lane = __lane_id(), wave = threadIdx.x >> 6;
  const unsigned long long key = pred ? (unsigned long long)(uintptr_t)ctr : 0ull;
  unsigned long long pending = __ballot(pred);
  int ng = 0;
  while (pending) {
    const int leader = __ffsll((long long)pending) - 1;
    const unsigned long long lk = __shfl(key, leader);
    const unsigned long long m = __ballot(pred && key == lk) & pending;
    if (lane == leader) {
      if (ng < kGroups) {
        s_key[wave * kGroups + ng] = lk;
        s_cnt[wave * kGroups + ng] = (unsigned int)__popcll(m);
      } else {
        atomicAdd((unsigned int*)(uintptr_t)lk, (unsigned int)__popcll(m));
      }
    }
    ++ng;
    pending &= ~m;
  }
  if (lane == 0) s_n[wave] = ng < kGroups ? ng : kGroups;
  __syncthreads();
  if (wave == 0) {
    const int e = lane;
    const bool has = e < NW * kGroups && (e % kGroups) < s_n[e / kGroups];
    const unsigned long long k2 = has ? s_key[e] : 0ull;
    const unsigned int c2 = has ? s_cnt[e] : 0u;
    unsigned long long pend = __ballot(has);
    while (pend) {
      const int leader = __ffsll((long long)pend) - 1;
      const unsigned long long lk = __shfl(k2, leader);
      const unsigned long long m = __ballot(has && k2 == lk) & pend;
      const unsigned int sum = wave_sum(((m >> lane) & 1ull) ? c2 : 0u);
      if (lane == leader) atomicAdd((unsigned int*)(uintptr_t)lk, sum);
      pend &= ~m;
    }
  }
}

__global__ void k_init(PtKernelArgs a) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i == 0) {
    PtCounters z = {};
    *a.cnt = z;
  }
  if (i < kRing) a.ring[i] = RingChunk{0, 0, 0, 0, 0};
  if (i < a.pool.P) a.pool.state[i] = 0;
}

__global__ void k_admit(PtKernelArgs a, int n) {
  if (blockIdx.x != 0 || threadIdx.x != 0) return;
  PtCounters* c = a.cnt;
  for (int k = 0; k < n; ++k) {
    const AdmitReq r = a.admit[k];
    if (!r.valid || c->ring_tail - c->ring_head >= (unsigned)kRing) continue;
    const unsigned long long b = c->work_total > c->work_next ? c->work_total : c->work_next;
    a.ring[c->ring_tail % kRing] = RingChunk{b, b + r.size, r.local, r.slot, 0};
    c->work_total = b + r.size;
    c->ring_tail++;
  }
}

__global__ void k_reset(PtKernelArgs a) {
  if (blockIdx.x != 0 || threadIdx.x != 0) return;
  PtCounters* c = a.cnt;
  c->tot_extend += c->n_act;
  c->tot_shadow += c->n_shadow;
  c->tot_mis += c->n_mis;
  c->last_act = c->n_act;
  c->n_act = 0;
  c->n_shadow = 0;
  c->n_mis = 0;
  c->iter++;
  while (c->ring_head != c->ring_tail && a.ring[c->ring_head % kRing].wend <= c->work_next) c->ring_head++;
}

// Zeroes the shading-class counters (one thread per class).
__global__ void k_reset_classes(PtKernelArgs a) {
  if (threadIdx.x < kShadeClasses) {
    a.cnt->cls_count[threadIdx.x] = 0u;
    a.cnt->cls_cursor[threadIdx.x] = 0u;
  }
}

// Chunk holding work index w: the last ring entry whose wbegin <= w.
__device__ __forceinline__ const RingChunk& find_chunk(const PtKernelArgs& a, unsigned long long w,
                                                       unsigned int head, unsigned int tail) {
  unsigned int lo = 0, hi = tail - head;  // search [lo, hi)
  while (hi - lo > 1) {
    unsigned int mid = (lo + hi) >> 1;
    if (a.ring[(head + mid) % kRing].wbegin <= w) lo = mid;
    else hi = mid;
  }
  return a.ring[(head + lo) % kRing];
}

template <int BLOCK>
__global__ __launch_bounds__(BLOCK) void k_generate(PtKernelArgs a) {
  __shared__ unsigned long long s64[BLOCK / 64 + 1];
  __shared__ unsigned int s32[BLOCK / 64 + 1];
  const int i = blockIdx.x * BLOCK + threadIdx.x;
  const PathPool& P = a.pool;
  const bool inpool = i < P.P;
  const int st = inpool ? P.state[i] : 1;
  const bool want = inpool && st == 0;
  // work indices for the block's free slots, taken with one atomic while
  // work remains (a stale read only costs an idle slot this iteration)
  unsigned long long w = ~0ull;
  {
    constexpr int NW = BLOCK / 64;
    const int lane = __lane_id(), wave = threadIdx.x >> 6;
    const unsigned long long m = __ballot(want);
    if (lane == 0) s64[wave] = (unsigned long long)__popcll(m);
    __syncthreads();
    if (threadIdx.x == 0) {
      unsigned long long run = 0;
      for (int k = 0; k < NW; ++k) {
        unsigned long long c = s64[k];
        s64[k] = run;
        run += c;
      }
      const volatile PtCounters* vc = a.cnt;
      s64[NW] = (run && vc->work_next < vc->work_total) ? atomicAdd(&a.cnt->work_next, run) : ~0ull;
    }
    __syncthreads();
    if (want && s64[NW] != ~0ull) w = s64[NW] + s64[wave] + (unsigned long long)__popcll(m & ((1ull << lane) - 1ull));
    __syncthreads();
  }
  bool fresh = false;
  const unsigned int head = a.cnt->ring_head, tail = a.cnt->ring_tail;
  if (want && w < a.cnt->work_total && tail != head && tail - head <= (unsigned)kRing) {
    const RingChunk& ch = find_chunk(a, w, head, tail);
    const int j = ch.slot >= 0 && ch.slot < kMaxJobs ? ch.slot : 0;
    const DevJob& J = a.jobs[j];
    if (w >= ch.wbegin && w < ch.wend) {
      if (J.cancelled) {
        atomicAdd(J.skipped, 1ull);
      } else {
        unsigned long long local = ch.local + (w - ch.wbegin);
        int pass = (int)(local / J.per_pass);
        unsigned long long rem = local % J.per_pass;
        int pix = (int)(rem / (unsigned long long)J.N);
        int n = (int)(rem % (unsigned long long)J.N);
        int ix = J.view.ext_x0 + pix % J.view.ext_w, iy = J.view.ext_y0 + pix / J.view.ext_w;
        Sampler sm{pixel_key(J.seed, (uint32_t)pass, ix, iy), (uint32_t)J.nu, (uint32_t)J.nv,
                   (uint32_t)J.N, 4u * (uint32_t)J.sd, 3u * (uint32_t)J.sd};
        float ox, oy, lu = 0.0f, lv = 0.0f;
        sm.camera((uint32_t)n, &ox, &oy);
        if (J.view.lens_radius > 0.0f) sm.lens((uint32_t)n, &lu, &lv);  // a pinhole camera does not read it
        float sx = (float)ix + ox, sy = (float)iy + oy;
        DRay r = fire_ray(J.view, sx, sy, lu, lv);
        const int Pn = P.P;
        P.job[i] = j;
        P.pass[i] = pass;
        P.pix[i] = pix;
        P.samp[i] = n;
        P.dspec[i] = 1 << 16;  // depth 0, specular (camera ray)
        P.dl_sp[i] = 0;        // no pending direct-lighting branches
        P.sxy[i] = sx;
        P.sxy[Pn + i] = sy;
        P.ray[i] = r.o.x; P.ray[Pn + i] = r.o.y; P.ray[2 * Pn + i] = r.o.z;
        P.ray[3 * Pn + i] = r.d.x; P.ray[4 * Pn + i] = r.d.y; P.ray[5 * Pn + i] = r.d.z;
        P.ray[6 * Pn + i] = r.tmin;
        xyz_store(P.rad, Pn, i, Xyz{});  // throughput is 1 at depth 0, not stored
        P.state[i] = 1;
        fresh = true;
      }
    }
  }
  block_count<BLOCK>(&a.cnt->tot_camera, fresh, s64);
  const bool active = inpool && (st == 1 || fresh);
  const unsigned int q = block_append<BLOCK>(&a.cnt->n_act, active, s32);
  if (active) P.q_act[q] = (uint32_t)i;
}

__device__ __forceinline__ int wave_min(int v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    int o = __shfl_xor(v, off);
    v = o < v ? o : v;
  }
  return v;
}

// Runs body(scene) for every valid lane, one distinct scene at a time, with
// the scene index made wave-uniform so the scene descriptor loads are scalar.
template <class F>
__device__ __forceinline__ void for_uniform_scene(bool valid, int my, F&& body) {
  bool done = !valid;
  while (__ballot(!done)) {
    int s = wave_min(done ? 0x7fffffff : my);
    if (!done && my == s) {
      body(__builtin_amdgcn_readfirstlane(s));
      done = true;
    }
  }
}

// Traversal for a scene of width W (0: any width, chosen per scene).
template <int W, bool ROOT_INF = false>
__device__ __forceinline__ void pt_closest(const DevScene& sc, const DRay& r, const trav::LdsSlot& slot,
                                           trav::Best& best) {
  trav::Counts cn;  // unused (no STATS)
  if constexpr (W == 0) trav::closest_hit<false, ROOT_INF>(sc, r, slot, best, cn);
  else trav::closest_hit_w<W, false, ROOT_INF>(sc, r, slot, best, cn);
}
template <int W, bool EXACT = false>
__device__ __forceinline__ bool pt_any(const DevScene& sc, const DRay& r, const trav::LdsSlot& slot) {
  trav::Counts cn;  // unused (no STATS)
  if constexpr (W == 0) return trav::any_hit<false, EXACT>(sc, r, slot, cn);
  else return trav::any_hit_w<W, false, EXACT>(sc, r, slot, cn);
}

template <int BLOCK, int W>
__global__ __launch_bounds__(BLOCK) void k_extend(PtKernelArgs a) {
  const trav::LdsSlot slot = trav::lane_slot<BLOCK>(a.nshort, a.walk);
  const int q = blockIdx.x * BLOCK + threadIdx.x;
  const unsigned int n = a.cnt->n_act;
  if ((unsigned)(blockIdx.x * BLOCK) >= n) return;  // whole block idle
  const PathPool& P = a.pool;
  const int Pn = P.P;
  bool valid = (unsigned)q < n;
  int i = valid ? (int)P.q_act[q] : 0;
  // the ray is loaded before the scene lookup, so that the two latencies
  // overlap (a loop over scenes would keep the loads behind it)
  const DRay r = load_pool_ray(P.ray, Pn, i, INFINITY);
  int scene = valid ? a.jobs[P.job[i]].scene : 0;
  for_uniform_scene(valid, scene, [&](int scene_u) {
    {
      const DevScene sc = a.scenes[scene_u].geo;
      trav::Best best;
      pt_closest<W>(sc, r, slot, best);
      P.hit[i] = best.t;
      P.hit[Pn + i] = best.b1;
      P.hit[2 * Pn + i] = best.b2;
      P.hprim[i] = best.prim;
    }
  });
}

// Shading class of active path i: 0 escaped, 1 at the depth cutoff, 2 no
// BxDF, then one class per BxDF makeup of the hit's material, so that grouped
// waves run one shading code path.  The makeups of diffuse kinds only
// (Lambert 0, Oren-Nayar 1) come first: 3 + kind for one component,
// 5 + 2 kind0 + kind1 for two; classes below kDiffuseClasses are exactly the
// paths the diffuse shading variant may take.  Then 9 + (kind - 2) for one
// other component (kinds 2..5) and 13 + 5 kind0 + kind1 for two (kinds 0..4).
__device__ __forceinline__ int shade_class(const PtKernelArgs& a, int i) {
  const PathPool& P = a.pool;
  const int prim = P.hprim[i];
  if (prim < 0) return 0;
  const DevJob& J = a.jobs[P.job[i]];
  if ((P.dspec[i] & 0xffff) == J.md) return 1;
  const DevPtScene& S = a.scenes[J.scene];
  const int m = prim < S.ntri ? S.tri_info[prim].x : S.geo.shapes[prim - S.ntri].material;
  if (m < 0 || m >= S.nmats || S.mats[m].n == 0) return 2;
  const DMaterial& M = S.mats[m];
  const int k0 = M.c[0].kind, k1 = M.n > 1 ? M.c[1].kind : 0;
  const bool diffuse = k0 <= K_OREN && (M.n == 1 || k1 <= K_OREN);
  int c;
  if (diffuse) c = M.n == 1 ? 3 + k0 : 5 + 2 * k0 + k1;
  else c = M.n == 1 ? 9 + (k0 - K_MF) : 13 + 5 * k0 + k1;
  // a texture's alternate material has the same BxDF makeup (the same type)
  return c < kShadeClasses ? c : kShadeClasses - 1;
}

// Grouping the shade queue by class: counts per class (one atomic per class
// per block), then each block reserves its range of every class and writes
// its paths there (order within a class is arbitrary).
template <int BLOCK>
__global__ __launch_bounds__(BLOCK) void k_class_count(PtKernelArgs a) {
  __shared__ unsigned int h[kShadeClasses];
  const int q = blockIdx.x * BLOCK + threadIdx.x;
  const unsigned int n = a.cnt->n_act;
  if ((unsigned)(blockIdx.x * BLOCK) >= n) return;  // whole block idle
  if (threadIdx.x < kShadeClasses) h[threadIdx.x] = 0u;
  __syncthreads();
  if ((unsigned)q < n) {
    const int c = shade_class(a, (int)a.pool.q_act[q]);
    a.pool.q_cls[q] = (uint32_t)c;
    atomicAdd(&h[c], 1u);
  }
  __syncthreads();
  if (threadIdx.x < kShadeClasses && h[threadIdx.x] != 0u)
    atomicAdd(&a.cnt->cls_count[threadIdx.x], h[threadIdx.x]);
}

template <int BLOCK>
__global__ __launch_bounds__(BLOCK) void k_class_scatter(PtKernelArgs a) {
  static_assert(BLOCK >= 64, "class bases are scanned by the first wave");
  __shared__ unsigned int h[kShadeClasses], base[kShadeClasses];
  const int q = blockIdx.x * BLOCK + threadIdx.x;
  const unsigned int n = a.cnt->n_act;
  if ((unsigned)(blockIdx.x * BLOCK) >= n) return;  // whole block idle
  if (threadIdx.x < kShadeClasses) h[threadIdx.x] = 0u;
  __syncthreads();
  const bool valid = (unsigned)q < n;
  const int c = valid ? (int)a.pool.q_cls[q] : 0;
  const unsigned int pos = valid ? atomicAdd(&h[c], 1u) : 0u;
  __syncthreads();
  if (threadIdx.x < 64) {  // first wave: class bases by an exclusive scan of the counts
    const int t = threadIdx.x;
    const unsigned int cnt = t < kShadeClasses ? a.cnt->cls_count[t] : 0u;
    unsigned int incl = cnt;
#pragma unroll
    for (int off = 1; off < kShadeClasses; off <<= 1) {
      const unsigned int v = __shfl_up(incl, off);
      if (t >= off) incl += v;
    }
    if (t < kShadeClasses)
      base[t] = incl - cnt + (h[t] != 0u ? atomicAdd(&a.cnt->cls_cursor[t], h[t]) : 0u);
  }
  __syncthreads();
  if (valid) a.pool.q_sh[base[c] + pos] = a.pool.q_act[q];
}

// sampleOneLight (Scene.hs:110-118) with sampleLightMis / sampleBsdfMis
// (:61-82): the light sample's shadow ray and the BSDF sample's MIS ray are
// queued with the contributions they carry if unoccluded / if they reach the
// light, scaled by the path throughput thr_at(band).
template <int KM, int LK, class Thr>
__device__ __forceinline__ void sample_one_light(const DevPtScene& S, const PathPool& P, int Pn, int i,
                                                 const Bsdf& bsdf, f3 wo, f3 p, float eps, float lNumU, float ld1,
                                                 float ld2, float bComp, float bd1, float bd2, Thr thr_at,
                                                 bool& want_shadow, bool& want_mis) {
  int lc = S.nlights;
  if (lc > 0) {  // sampleOneLight (Scene.hs:110-118)
    int ln = 0;
    float lscale = 1.0f;
    if (lc > 1) {
      ln = (int)floorf(lNumU * (float)lc);
      ln = ln < lc - 1 ? ln : lc - 1;
      lscale = (float)lc;
    }
    {  // sampleLightMis (Scene.hs:61-69)
      LSample ls = sample_light<LK>(S, ln, p, eps, bsdf.cs.n, ld1, ld2);  // n: bsdfShadingNormal (Path.hs:61)
      if (ls.pdf != 0.0f && (!lrad_black<LK>(ls.li) || (has_sky_light<LK>() && ls.sky))) {
        SpecVal f = eval_bsdf<KM>(bsdf, wo, ls.wi);
        // a delta light's sample is not weighted (Scene.hs:65)
        float k = has_delta_light<LK>() && ls.delta ? 1.0f / ls.pdf
                                                    : power_heuristic(ls.pdf, bsdf_pdf<KM>(bsdf, wo, ls.wi)) / ls.pdf;
        const LRad lr = sample_radiance<LK>(S, ln, ls);
        Xyz c;
        bool liBlack = true, fBlack = true;
#pragma unroll 4
        for (int b = 0; b < 16; ++b) {
          float fv = spec_at<KM>(f, b), li = lrad_at<LK>(lr, b);
          if constexpr (has_delta_light<LK>()) li = li * ls.scale;  // sScale r s (1 for other lights: exact)
          liBlack = liBlack && li == 0.0f;
          fBlack = fBlack && fv == 0.0f;
          c.add(b, thr_at(b) * (((fv * li) * k) * lscale));
        }
        if (!(liBlack || fBlack)) {
          xyz_store(P.pend_s, Pn, i, c);
          P.sray[i] = ls.ray.o.x; P.sray[Pn + i] = ls.ray.o.y; P.sray[2 * Pn + i] = ls.ray.o.z;
          P.sray[3 * Pn + i] = ls.ray.d.x; P.sray[4 * Pn + i] = ls.ray.d.y; P.sray[5 * Pn + i] = ls.ray.d.z;
          P.sray[6 * Pn + i] = ls.ray.tmin; P.sray[7 * Pn + i] = ls.ray.tmax;
          want_shadow = true;
        }
      }
    }
    // sampleBsdfMis (Scene.hs:71-82): black for a delta light, whose le and
    // intLe are black and which no ray can hit
    if (!(has_delta_light<LK>() && (S.lights[ln].type == LT_POINT || S.lights[ln].type == LT_DIR))) {
      SpecVal fb;
      f3 wib;
      int tb;
      float bpdf = sample_bsdf<KM>(bsdf, wo, bComp, bd1, bd2, fb, &wib, &tb);
      if (bpdf != 0.0f) {
        float lpdf = light_pdf<LK>(S, ln, p, wib);
        float w = power_heuristic(bpdf, lpdf);
        // the emission a hit on the sampled light returns is known here:
        // the area light's radiance, or the infinite light's along wib
        const LRad e = has_inf_light<LK>() && S.lights[ln].type == LT_INF ? light_le<LK>(S, ln, wib)
                                                                           : lrad(S.lights[ln].L);
        Xyz c;
        bool fBlack = true;
#pragma unroll 4
        for (int b = 0; b < 16; ++b) {
          float fv = spec_at<KM>(fb, b);
          fBlack = fBlack && fv == 0.0f;
          c.add(b, (thr_at(b) * ((fv * w) * lscale)) * lrad_at<LK>(e, b));
        }
        // An area light is reached only if the ray hits its shape: the shape
        // test with tmax = inf (the traversal's own arithmetic, monotone in
        // tmax) culls rays that miss it, and its hit bounds the traversal,
        // since nothing beyond the light can be the closest hit.
        float mt = INFINITY;
        bool misses = false;
        if (has_area_light<LK>() && S.lights[ln].type == LT_AREA) {
          const DShape& lsh = S.lights[ln].sh;
          DRay ro;
          ro.o = xform_point(lsh.w2o, p);
          ro.d = xform_vector(lsh.w2o, wib);
          ro.tmin = eps;
          ro.tmax = INFINITY;
          float tl;
          int ax;
          if (shape_intersect_k<light_shapes<LK>()>(lsh, ro, &tl, &ax)) mt = nextafterf(tl, INFINITY);
          else misses = true;
        }
        if (!fBlack && !misses) {
          xyz_store(P.pend_m, Pn, i, c);
          P.mray[i] = p.x; P.mray[Pn + i] = p.y; P.mray[2 * Pn + i] = p.z;
          P.mray[3 * Pn + i] = wib.x; P.mray[4 * Pn + i] = wib.y; P.mray[5 * Pn + i] = wib.z;
          P.mray[6 * Pn + i] = eps;
          P.mray[7 * Pn + i] = mt;
          P.mlight[i] = ln;
          want_mis = true;
        }
      }
    }
  }
}

// Node textures (Texture.hs:59-414): for each active path whose hit's
// material takes inputs from the scene's texture node list, evaluates them
// at the shading DG (`tex dgs`) and writes the material's BxDF records, built
// from them as from constants (Material.hs:32-127), to the path's pmat
// entry, which make_bsdf takes for that material.  Runs only while such a
// scene is resident.
template <int BLOCK>
__global__ __launch_bounds__(BLOCK) void k_texture(PtKernelArgs a) {
  const int q = blockIdx.x * BLOCK + threadIdx.x;
  const unsigned int n = a.cnt->n_act;
  if ((unsigned)(blockIdx.x * BLOCK) >= n) return;
  const PathPool& P = a.pool;
  const int Pn = P.P;
  const bool valid = (unsigned)q < n;
  const int i = valid ? (int)P.q_act[q] : 0;
  const int prim = P.hprim[i];
  const bool hit = valid && prim >= 0;
  for_uniform_scene(hit, hit ? a.jobs[P.job[i]].scene : 0, [&](int scene_u) {
    const DevPtScene& S = a.scenes[scene_u];
    if (!S.mat_src) return;  // the scene has no node textures
    const int mat = prim < S.ntri ? S.tri_info[prim].x : S.geo.shapes[prim - S.ntri].material;
    if (mat < 0 || mat >= S.nmats || S.mats[mat].tex_kind != BGPU_TEX_NODES) return;
    const DMatSrc& src = S.mat_src[S.mats[mat].alt];
    const DRay ray = load_pool_ray(P.ray, Pn, i, INFINITY);
    DG dgg, dgs;
    float eps;
    int m2, light;
    hit_dg(S, ray, P.hit[i], P.hit[Pn + i], P.hit[2 * Pn + i], prim, &dgg, &dgs, &eps, &m2, &light);
    const tx::Tex T{S.tex_nodes, S.images, S.texels};
    const tx::TexPt t{dgs.p, dgs.u, dgs.v};
    MatIn in;
    in.type = src.type;
    for (int k = 0; k < 4; ++k)
      in.scalar[k] = ((src.slots >> (4 + k)) & 1) ? tx::eval_scalar(T, src.lo[4 + k], src.roots[4 + k], t)
                                                  : src.scalar[k];
    for (int k = 0; k < 3; ++k) {
      if ((src.slots >> k) & 1) {
        tx::eval_spectrum(T, src.lo[k], src.roots[k], t, in.spec[k]);
      } else {
        for (int b = 0; b < 16; ++b) in.spec[k][b] = src.spec[k][b];
      }
    }
    DMaterial dm;
    build_material(in, dm);
    if ((src.slots >> 8) & 1) {  // bump d dgg dgs (Reflection.hs:347-378)
      const float du = 0.01f, dv = 0.01f;
      const float uDisp =
          tx::eval_scalar(T, src.lo[8], src.roots[8], tx::TexPt{add3(dgs.p, scl3(dgs.dpdu, du)), dgs.u + du, dgs.v});
      const float vDisp =
          tx::eval_scalar(T, src.lo[8], src.roots[8], tx::TexPt{add3(dgs.p, scl3(dgs.dpdv, dv)), dgs.u, dgs.v + dv});
      const float disp = tx::eval_scalar(T, src.lo[8], src.roots[8], t);
      const float vscale = (vDisp - disp) / dv;
      const f3 dpdv = add3(dgs.dpdv, scl3(dgs.n, vscale));
      const float uscale = (uDisp - disp) / du;
      const f3 dpdu = add3(dgs.dpdu, scl3(dgs.n, uscale));
      f3 nb = norm3(cross3(dpdu, dpdv));
      if (dot3(nb, dgg.n) < 0.0f) nb = mk3(0.0f - nb.x, 0.0f - nb.y, 0.0f - nb.z);  // faceForward, Num's 0 - v
      dm.tex_map = 1;  // make_bsdf takes the bumped normal and dpdu from tex[0..5]
      dm.tex[0] = nb.x;
      dm.tex[1] = nb.y;
      dm.tex[2] = nb.z;
      dm.tex[3] = dpdu.x;
      dm.tex[4] = dpdu.y;
      dm.tex[5] = dpdu.z;
    }
    P.pmat[i] = dm;
  });
}

// Queue ranges of the shading variants: the whole queue, or (grouped queue)
// its diffuse classes [0, n_diff) and the rest [n_diff, n_act).
enum { SHADE_ALL = 0, SHADE_DIFFUSE_PART = 1, SHADE_OTHER_PART = 2 };

template <int PART>
__device__ __forceinline__ void shade_range(const PtKernelArgs& a, unsigned int* lo, unsigned int* hi) {
  const unsigned int nact = a.cnt->n_act;
  if constexpr (PART == SHADE_ALL) {
    *lo = 0u;
    *hi = nact;
  } else {
    unsigned int nd = 0u;
#pragma unroll
    for (int c = 0; c < kDiffuseClasses; ++c) nd += a.cnt->cls_count[c];
    nd = nd < nact ? nd : nact;
    *lo = PART == SHADE_DIFFUSE_PART ? 0u : nd;
    *hi = PART == SHADE_DIFFUSE_PART ? nd : nact;
  }
}

// Path shading (Path.hs:41-87).  KM: the BxDF kinds compiled in (kAllKinds,
// or kDiffuseKinds for the variant that shades only diffuse materials with
// far fewer registers); LK: the light kinds compiled in (kAllLights, or the
// quad-only or infinite-only sets); PART: the queue range it shades.
template <int BLOCK, int MINW, int KM, int PART, int LK = kAllLights>
__global__ __launch_bounds__(BLOCK, MINW) void k_shade(PtKernelArgs a) {
  unsigned int lo, hi;
  shade_range<PART>(a, &lo, &hi);
  if ((unsigned)(blockIdx.x * BLOCK) >= hi - lo) return;
  const int q = (int)lo + blockIdx.x * BLOCK + threadIdx.x;
  const PathPool& P = a.pool;
  const int Pn = P.P;
  bool valid = (unsigned)q < hi;
  int i = valid ? (int)(a.sort_shade ? P.q_sh[q] : P.q_act[q]) : 0;
  int jslot = valid ? P.job[i] : 0;
  if (a.direct && valid && a.jobs[jslot].integ == BGPU_INTEGRATOR_DIRECT) valid = false;  // k_shade_direct's
  bool want_shadow = false, want_mis = false;
  for_uniform_scene(valid, valid ? a.jobs[jslot].scene : 0, [&](int scene_u) {
    {
      const DevPtScene& S = a.scenes[scene_u];
      const DevJob& J = a.jobs[jslot];
      // loaded here rather than before the scene loop: values live across
      // the loop would stay in registers through the whole shading body
      const DRay ray = load_pool_ray(P.ray, Pn, i, INFINITY);
      int prim = P.hprim[i];
      int ds = P.dspec[i];
      int depth = ds & 0xffff;
      bool spec = (ds >> 16) != 0;
      // throughput band b: exactly 1 for camera rays (not stored at depth 0)
      const float* thr = P.thr + i;  // band b at thr[b * Pn]
      auto thr_at = [&](int b) { return depth == 0 ? 1.0f : thr[b * Pn]; };
      if (prim < 0) {  // Path.hs:44-47
        if (spec) {
          for (int li = 0; li < S.nlights; ++li) {
            const LRad le = light_le<LK>(S, li, ray.d);
            if (!lrad_black<LK>(le)) {
              Xyz c;
#pragma unroll
              for (int b = 0; b < 16; ++b) c.add(b, thr_at(b) * lrad_at<LK>(le, b));
              xyz_accum(P.rad, Pn, i, c);
            }
          }
        }
        P.state[i] = 2;
      } else if (depth == J.md) {  // Path.hs:51
        P.state[i] = 2;
      } else {
        float th = P.hit[i], hb1 = P.hit[Pn + i], hb2 = P.hit[2 * Pn + i];
        int pix = P.pix[i], n = P.samp[i];
        int ix = J.view.ext_x0 + pix % J.view.ext_w, iy = J.view.ext_y0 + pix / J.view.ext_w;
        Sampler sm{pixel_key(J.seed, (uint32_t)P.pass[i], ix, iy), (uint32_t)J.nu, (uint32_t)J.nv,
                   (uint32_t)J.N, 4u * (uint32_t)J.sd, 3u * (uint32_t)J.sd};
        DG dgg, dgs;
        float eps;
        int mat, light;
        hit_dg(S, ray, th, hb1, hb2, prim, &dgg, &dgs, &eps, &mat, &light);
        Bsdf bsdf = make_bsdf(S, mat, dgg, dgs, P.pmat ? P.pmat + i : nullptr);
        uint32_t d = (uint32_t)depth, un = (uint32_t)n;
        float lNumU = sm.get1(un, 1 + 4 * d);
        float ld1, ld2, bd1, bd2;
        sm.get2(un, 1 + 3 * d, &ld1, &ld2);
        float bComp = sm.get1(un, 2 + 4 * d);
        sm.get2(un, 2 + 3 * d, &bd1, &bd2);
        f3 wo = neg3(ray.d);
        f3 p = bsdf.p;
        // emission seen through a specular chain (Path.hs:60, intLe with rd)
        if (spec && light >= 0 && dot3(dgg.n, ray.d) > 0.0f) {
          const float* L = S.lights[light].L;
          Xyz c;
#pragma unroll
          for (int b = 0; b < 16; ++b) c.add(b, thr_at(b) * L[b]);
          xyz_accum(P.rad, Pn, i, c);
        }
        sample_one_light<KM, LK>(S, P, Pn, i, bsdf, wo, p, eps, lNumU, ld1, ld2, bComp, bd1, bd2, thr_at,
                                 want_shadow, want_mis);
        // Russian roulette (Path.hs:68-72)
        float pc = 1.0f;
        if (depth > J.rr) {
          float y = 0.0f;  // sY (Spectrum.hs:371-373)
#pragma unroll
          for (int b = 0; b < 16; ++b) y = y + thr_at(b) * cCieY[b];
          pc = hmin(0.75f, y / BGPU_CIE_Y_SUM);
        }
        float x = sm.get1(un, 3 + 4 * d);
        bool cont = !(x > pc);
        if (cont) {
          float uc = sm.get1(un, 0 + 4 * d);
          float u1, u2;
          sm.get2(un, 0 + 3 * d, &u1, &u2);
          SpecVal f;
          f3 wi;
          int tp;
          float spdf = sample_bsdf<KM>(bsdf, wo, uc, u1, u2, f, &wi, &tp);
          if (spdf == 0.0f) {
            cont = false;
          } else {
            // the new throughput is written band by band; a black sample ends
            // the path, and then the values written are never read
            float ipc = 1.0f / pc;
            bool fBlack = true;
            float* thw = P.thr + i;
#pragma unroll 4
            for (int b = 0; b < 16; ++b) {
              float fv = spec_at<KM>(f, b);
              fBlack = fBlack && fv == 0.0f;
              thw[b * Pn] = (fv * thr_at(b)) * ipc;
            }
            if (fBlack) {
              cont = false;
            } else {
              P.ray[i] = p.x; P.ray[Pn + i] = p.y; P.ray[2 * Pn + i] = p.z;
              P.ray[3 * Pn + i] = wi.x; P.ray[4 * Pn + i] = wi.y; P.ray[5 * Pn + i] = wi.z;
              P.ray[6 * Pn + i] = eps;
              P.dspec[i] = (depth + 1) | ((tp & BX_S) ? (1 << 16) : 0);
            }
          }
        }
        P.state[i] = cont ? 1 : 2;
      }
    }
  });
  __shared__ unsigned int s32[BLOCK / 64 + 1];
  const unsigned int qs = block_append<BLOCK>(&a.cnt->n_shadow, want_shadow, s32);
  if (want_shadow) P.q_shadow[qs] = (uint32_t)i;
  const unsigned int qm = block_append<BLOCK>(&a.cnt->n_mis, want_mis, s32);
  if (want_mis) P.q_mis[qm] = (uint32_t)i;
}

// Shading of direct-lighting paths (Integrator/DirectLighting.hs:23-59).
// A slot walks its sample's tree of specular branches depth first: at a hit
// it adds the emission towards wo and sampleOneLight (dimensions 2d, 2d + 1),
// then continues along the specular reflection and pushes the specular
// transmission (both sampled with uComp 0.5, uDir (0.5, 0.5)) while depth + 1
// < maxDepth, each branch weighted by its sampled f.  A path that escapes
// (black) or has no branch left pops the next pending one, or ends.
template <int BLOCK, int LK = kAllLights>
__global__ __launch_bounds__(BLOCK) void k_shade_direct(PtKernelArgs a) {
  const int q = blockIdx.x * BLOCK + threadIdx.x;
  const unsigned int nact = a.cnt->n_act;
  if ((unsigned)(blockIdx.x * BLOCK) >= nact) return;
  const PathPool& P = a.pool;
  const int Pn = P.P;
  bool valid = (unsigned)q < nact;
  int i = valid ? (int)(a.sort_shade ? P.q_sh[q] : P.q_act[q]) : 0;
  int jslot = valid ? P.job[i] : 0;
  valid = valid && a.jobs[jslot].integ == BGPU_INTEGRATOR_DIRECT;
  bool want_shadow = false, want_mis = false;
  for_uniform_scene(valid, valid ? a.jobs[jslot].scene : 0, [&](int scene_u) {
    const DevPtScene& S = a.scenes[scene_u];
    const DevJob& J = a.jobs[jslot];
    DRay ray;
    ray.o = mk3(P.ray[i], P.ray[Pn + i], P.ray[2 * Pn + i]);
    ray.d = mk3(P.ray[3 * Pn + i], P.ray[4 * Pn + i], P.ray[5 * Pn + i]);
    ray.tmin = P.ray[6 * Pn + i];
    ray.tmax = INFINITY;
    const int prim = P.hprim[i];
    const int depth = P.dspec[i] & 0xffff;
    const float* thr = P.thr + i;
    auto thr_at = [&](int b) { return depth == 0 ? 1.0f : thr[b * Pn]; };
    bool more = false;  // the slot continues with a branch of this vertex
    if (prim >= 0) {
      float th = P.hit[i], hb1 = P.hit[Pn + i], hb2 = P.hit[2 * Pn + i];
      int pix = P.pix[i], n = P.samp[i];
      int ix = J.view.ext_x0 + pix % J.view.ext_w, iy = J.view.ext_y0 + pix / J.view.ext_w;
      // SurfaceIntegrator li (2 md) (2 md): tables for every dimension used
      Sampler sm{pixel_key(J.seed, (uint32_t)P.pass[i], ix, iy), (uint32_t)J.nu, (uint32_t)J.nv, (uint32_t)J.N,
                 2u * (uint32_t)J.md, 2u * (uint32_t)J.md};
      DG dgg, dgs;
      float eps;
      int mat, light;
      hit_dg(S, ray, th, hb1, hb2, prim, &dgg, &dgs, &eps, &mat, &light);
      Bsdf bsdf = make_bsdf(S, mat, dgg, dgs, P.pmat ? P.pmat + i : nullptr);
      const uint32_t k = 2u * (uint32_t)depth, un = (uint32_t)n;
      const float uln = sm.get1(un, k), ubc = sm.get1(un, k + 1);
      float uld1, uld2, ubd1, ubd2;
      sm.get2(un, k, &uld1, &uld2);
      sm.get2(un, k + 1, &ubd1, &ubd2);
      const f3 wo = neg3(ray.d), p = bsdf.p;
      if (light >= 0 && dot3(dgg.n, wo) > 0.0f) {  // intLe int wo (Primitive.hs:68-76, Light.hs:85-93)
        const float* L = S.lights[light].L;
        Xyz c;
#pragma unroll
        for (int b = 0; b < 16; ++b) c.add(b, thr_at(b) * L[b]);
        xyz_accum(P.rad, Pn, i, c);
      }
      sample_one_light<kAllKinds, LK>(S, P, Pn, i, bsdf, wo, p, eps, uln, uld1, uld2, ubc, ubd1, ubd2,
                                              thr_at, want_shadow, want_mis);
      if (depth + 1 < J.md) {  // cont (DirectLighting.hs:45-59)
        SpecVal fr, ft;
        f3 wr, wt;
        const float pr = sample_bsdf_flt(bsdf, wo, BX_S | BX_R, fr, &wr);
        const float pt = sample_bsdf_flt(bsdf, wo, BX_S | BX_T, ft, &wt);
        const bool br = pr != 0.0f, bt = pt != 0.0f;
        if (br && bt) {  // reflection continues here, transmission waits
          const int e = P.dl_sp[i];
          // validate_job bounds maxDepth by dl_cap + 1 and at most maxDepth - 1
          // branches are pending at once, so a full stack is unreachable;
          // a dropped branch would bias the image, so it is counted
          if (e >= P.dl_cap) atomicAdd(&a.cnt->dl_dropped, 1ull);
          if (e < P.dl_cap) {
            float* st = P.dl_stk + (size_t)e * kDlWords * Pn + i;
            st[0] = p.x; st[Pn] = p.y; st[2 * Pn] = p.z;
            st[3 * Pn] = wt.x; st[4 * Pn] = wt.y; st[5 * Pn] = wt.z;
            st[6 * Pn] = eps;
#pragma unroll 4
            for (int b = 0; b < 16; ++b) st[(7 + b) * Pn] = spec_at(ft, b) * thr_at(b);
            st[23 * Pn] = __int_as_float(depth + 1);
            P.dl_sp[i] = e + 1;
          }
        }
        if (br || bt) {
          const SpecVal& f = br ? fr : ft;
          const f3 wi = br ? wr : wt;
          float* thw = P.thr + i;
#pragma unroll 4
          for (int b = 0; b < 16; ++b) thw[b * Pn] = spec_at(f, b) * thr_at(b);
          P.ray[i] = p.x; P.ray[Pn + i] = p.y; P.ray[2 * Pn + i] = p.z;
          P.ray[3 * Pn + i] = wi.x; P.ray[4 * Pn + i] = wi.y; P.ray[5 * Pn + i] = wi.z;
          P.ray[6 * Pn + i] = eps;
          P.dspec[i] = depth + 1;
          more = true;
        }
      }
    }
    if (!more) {
      const int e = P.dl_sp[i];
      if (e > 0) {  // the next pending branch
        const float* st = P.dl_stk + (size_t)(e - 1) * kDlWords * Pn + i;
        P.ray[i] = st[0]; P.ray[Pn + i] = st[Pn]; P.ray[2 * Pn + i] = st[2 * Pn];
        P.ray[3 * Pn + i] = st[3 * Pn]; P.ray[4 * Pn + i] = st[4 * Pn]; P.ray[5 * Pn + i] = st[5 * Pn];
        P.ray[6 * Pn + i] = st[6 * Pn];
        float* thw = P.thr + i;
#pragma unroll 4
        for (int b = 0; b < 16; ++b) thw[b * Pn] = st[(7 + b) * Pn];
        P.dspec[i] = __float_as_int(st[23 * Pn]);
        P.dl_sp[i] = e - 1;
        more = true;
      }
    }
    P.state[i] = more ? 1 : 2;
  });
  __shared__ unsigned int s32[BLOCK / 64 + 1];
  const unsigned int qs = block_append<BLOCK>(&a.cnt->n_shadow, want_shadow, s32);
  if (want_shadow) P.q_shadow[qs] = (uint32_t)i;
  const unsigned int qm = block_append<BLOCK>(&a.cnt->n_mis, want_mis, s32);
  if (want_mis) P.q_mis[qm] = (uint32_t)i;
}

template <int BLOCK, int W>
__global__ __launch_bounds__(BLOCK) void k_shadow(PtKernelArgs a) {
  const trav::LdsSlot slot = trav::lane_slot<BLOCK>(a.nshort, a.walk);
  const int q = blockIdx.x * BLOCK + threadIdx.x;
  const unsigned int n = a.cnt->n_shadow;
  if ((unsigned)(blockIdx.x * BLOCK) >= n) return;
  const PathPool& P = a.pool;
  const int Pn = P.P;
  bool valid = (unsigned)q < n;
  int i = valid ? (int)P.q_shadow[q] : 0;
  const DRay r = load_pool_ray(P.sray, Pn, i);  // before the scene lookup (k_extend)
  for_uniform_scene(valid, valid ? a.jobs[P.job[i]].scene : 0, [&](int scene_u) {
    {
      const DevScene sc = a.scenes[scene_u].geo;
      if (!pt_any<W>(sc, r, slot)) {  // occluded (Scene.hs:45-47)
        xyz_accum(P.rad, Pn, i, P.pend_s);
      }
    }
  });
}

template <int BLOCK, int W>
__global__ __launch_bounds__(BLOCK) void k_mis(PtKernelArgs a) {
  const trav::LdsSlot slot = trav::lane_slot<BLOCK>(a.nshort, a.walk);
  const int q = blockIdx.x * BLOCK + threadIdx.x;
  const unsigned int n = a.cnt->n_mis;
  if ((unsigned)(blockIdx.x * BLOCK) >= n) return;
  const PathPool& P = a.pool;
  const int Pn = P.P;
  bool valid = (unsigned)q < n;
  int i = valid ? (int)P.q_mis[q] : 0;
  // before the scene lookup (k_extend); tmax is bound by the sampled light's
  // hit (the root box sees tmax = inf)
  const DRay r = load_pool_ray(P.mray, Pn, i);
  const int ln = P.mlight[i];
  for_uniform_scene(valid, valid ? a.jobs[P.job[i]].scene : 0, [&](int scene_u) {
    {
      const DevPtScene& S = a.scenes[scene_u];
      const DevScene sc = S.geo;
      const DLight& L = S.lights[ln];
      trav::Best best;
      best.prim = -1;
      if (L.type == LT_INF) {
        // an infinite light is reached iff nothing is hit: any hit, with the
        // closest-hit query's shape tests, decides it (tmax is inf here)
        if (pt_any<W, true>(sc, r, slot)) best.prim = 0;
      } else {
        pt_closest<W, true>(sc, r, slot, best);
      }
      // pend_m already holds the contribution times the sampled light's
      // emission; here only whether the ray reaches that light is decided
      bool reaches = false;
      if (best.prim >= 0) {
        // the hit primitive's light must be the sampled AreaLight (Light.hs:47-50)
        if (best.prim >= S.ntri && L.type == LT_AREA) {
          int hl = sc.shapes[best.prim - S.ntri].light;
          if (hl >= 0 && S.lights[hl].id == L.id) {
            DG dgg, dgs;
            float eps;
            int mat, light;
            hit_dg(S, r, best.t, best.b1, best.b2, best.prim, &dgg, &dgs, &eps, &mat, &light);
            reaches = dot3(dgg.n, neg3(r.d)) > 0.0f;  // lEmit with -wi
          }
        }
      } else {
        reaches = L.type == LT_INF;  // le of an infinite light (Light.hs:98-106)
      }
      if (reaches) xyz_accum(P.rad, Pn, i, P.pend_m);
    }
  });
}

// ---- persistent traversal passes (scenes of one width W) -----------------
// Each traversal is two passes over its queue, triangles then shapes, with
// lanes refilled from one static range of queue entries per wave
// (trav_persist.h).  Items carry their own scene, so jobs of different
// scenes share a wave.
__device__ __forceinline__ const DevScene& scene_of(const PtKernelArgs& a, int i) {
  return a.scenes[a.jobs[a.pool.job[i]].scene].geo;
}
__device__ __forceinline__ trav::LaneScene lane_scene(const DevScene& g, bool fat) {
  return trav::LaneScene{g.nodes, g.leaf, g.shapes, g.shape_pad_k, g.shape_floor_k,
                         fat ? g.shape_root : g.tri_root};
}

template <bool FAT>
struct ExtendSrc {  // closest hit of the path ray; result in hit / hprim
  const PtKernelArgs& a;
  __device__ __forceinline__ bool begin(int k, DRay& r, trav::LaneScene& ls, trav::Best& best) {
    const PathPool& P = a.pool;
    const int Pn = P.P, i = (int)P.q_act[k];
    const DevScene& g = scene_of(a, i);
    r.o = mk3(P.ray[i], P.ray[Pn + i], P.ray[2 * Pn + i]);
    r.d = mk3(P.ray[3 * Pn + i], P.ray[4 * Pn + i], P.ray[5 * Pn + i]);
    r.tmin = P.ray[6 * Pn + i];
    r.tmax = INFINITY;
    ls = lane_scene(g, FAT);
    best = trav::Best{r.tmax, 0.0f, 0.0f, -1};
    if (FAT && P.hprim[i] >= 0) best = trav::Best{P.hit[i], P.hit[Pn + i], P.hit[2 * Pn + i], P.hprim[i]};
    return slab_haskell(g.root_lo, g.root_hi, r);
  }
  __device__ __forceinline__ void end(int k, const DRay&, const trav::LaneScene&, const trav::Best& best) {
    if (FAT && best.prim < 0) return;
    const PathPool& P = a.pool;
    const int Pn = P.P, i = (int)P.q_act[k];
    P.hit[i] = best.t;
    P.hit[Pn + i] = best.b1;
    P.hit[2 * Pn + i] = best.b2;
    P.hprim[i] = best.prim;
  }
};

template <bool FAT>
struct ShadowSrc {  // any hit of the shadow ray; unoccluded adds the light sample
  const PtKernelArgs& a;
  __device__ __forceinline__ bool begin(int k, DRay& r, trav::LaneScene& ls, trav::Best& best) {
    const PathPool& P = a.pool;
    const int Pn = P.P, i = (int)P.q_shadow[k];
    const DevScene& g = scene_of(a, i);
    r.o = mk3(P.sray[i], P.sray[Pn + i], P.sray[2 * Pn + i]);
    r.d = mk3(P.sray[3 * Pn + i], P.sray[4 * Pn + i], P.sray[5 * Pn + i]);
    r.tmin = P.sray[6 * Pn + i];
    r.tmax = P.sray[7 * Pn + i];
    ls = lane_scene(g, FAT);
    best = trav::Best{r.tmax, 0.0f, 0.0f, -1};
    return slab_haskell(g.root_lo, g.root_hi, r) && (!FAT || P.occ[i] == 0);
  }
  __device__ __forceinline__ void end(int k, const DRay&, const trav::LaneScene&, const trav::Best& best) {
    const PathPool& P = a.pool;
    const int i = (int)P.q_shadow[k];
    if (!FAT) {
      P.occ[i] = best.prim >= 0 ? 1 : 0;
    } else if (best.prim < 0 && P.occ[i] == 0) {  // unoccluded (Scene.hs:45-47, 61-69)
      xyz_accum(P.rad, P.P, i, P.pend_s);
    }
  }
};

template <bool FAT>
struct MisSrc {  // closest hit of the MIS ray into hit / hprim (resolved by k_mis_resolve)
  const PtKernelArgs& a;
  __device__ __forceinline__ bool begin(int k, DRay& r, trav::LaneScene& ls, trav::Best& best) {
    const PathPool& P = a.pool;
    const int Pn = P.P, i = (int)P.q_mis[k];
    const DevScene& g = scene_of(a, i);
    r.o = mk3(P.mray[i], P.mray[Pn + i], P.mray[2 * Pn + i]);
    r.d = mk3(P.mray[3 * Pn + i], P.mray[4 * Pn + i], P.mray[5 * Pn + i]);
    r.tmin = P.mray[6 * Pn + i];
    r.tmax = INFINITY;
    const bool root = slab_haskell(g.root_lo, g.root_hi, r);  // the reference's ray: tmax = inf
    r.tmax = P.mray[7 * Pn + i];  // bound by the sampled light's hit
    ls = lane_scene(g, FAT);
    best = trav::Best{r.tmax, 0.0f, 0.0f, -1};
    if (FAT && P.hprim[i] >= 0) best = trav::Best{P.hit[i], P.hit[Pn + i], P.hit[2 * Pn + i], P.hprim[i]};
    return root;
  }
  __device__ __forceinline__ void end(int k, const DRay&, const trav::LaneScene&, const trav::Best& best) {
    if (FAT && best.prim < 0) return;
    const PathPool& P = a.pool;
    const int Pn = P.P, i = (int)P.q_mis[k];
    P.hit[i] = best.t;
    P.hit[Pn + i] = best.b1;
    P.hit[2 * Pn + i] = best.b2;
    P.hprim[i] = best.prim;
  }
};

template <int BLOCK, int W, bool FAT>
__global__ __launch_bounds__(BLOCK) void k_extend_p(PtKernelArgs a) {
  const trav::LdsSlot slot = trav::lane_slot<BLOCK>(a.nshort, a.walk);
  ExtendSrc<FAT> src{a};
  trav::persist_trace<W, false, FAT, false>(src, (int)a.cnt->n_act, nullptr, slot, nullptr,
                                            trav::PersistTune{trav::kRefillMin, a.walk});
}
template <int BLOCK, int W, bool FAT>
__global__ __launch_bounds__(BLOCK) void k_shadow_p(PtKernelArgs a) {
  const trav::LdsSlot slot = trav::lane_slot<BLOCK>(a.nshort, a.walk);
  ShadowSrc<FAT> src{a};
  trav::persist_trace<W, true, FAT, false>(src, (int)a.cnt->n_shadow, nullptr, slot, nullptr,
                                            trav::PersistTune{trav::kRefillMin, a.walk});
}
template <int BLOCK, int W, bool FAT>
__global__ __launch_bounds__(BLOCK) void k_mis_p(PtKernelArgs a) {
  const trav::LdsSlot slot = trav::lane_slot<BLOCK>(a.nshort, a.walk);
  MisSrc<FAT> src{a};
  trav::persist_trace<W, false, FAT, false>(src, (int)a.cnt->n_mis, nullptr, slot, nullptr,
                                            trav::PersistTune{trav::kRefillMin, a.walk});
}

// The MIS ray reaches the sampled light: the hit primitive is that AreaLight
// seen from its emitting side, or an infinite light's ray escapes.
template <int BLOCK>
__global__ __launch_bounds__(BLOCK) void k_mis_resolve(PtKernelArgs a) {
  const int q = blockIdx.x * BLOCK + threadIdx.x;
  const unsigned int n = a.cnt->n_mis;
  if ((unsigned)(blockIdx.x * BLOCK) >= n) return;
  const PathPool& P = a.pool;
  const int Pn = P.P;
  bool valid = (unsigned)q < n;
  int i = valid ? (int)P.q_mis[q] : 0;
  for_uniform_scene(valid, valid ? a.jobs[P.job[i]].scene : 0, [&](int scene_u) {
    const DevPtScene& S = a.scenes[scene_u];
    const int prim = P.hprim[i];
    const DLight& L = S.lights[P.mlight[i]];
    bool reaches = false;
    if (prim >= 0) {
      // the hit primitive's light must be the sampled AreaLight (Light.hs:47-50)
      if (prim >= S.ntri && L.type == LT_AREA) {
        int hl = S.geo.shapes[prim - S.ntri].light;
        if (hl >= 0 && S.lights[hl].id == L.id) {
          DRay r;
          r.o = mk3(P.mray[i], P.mray[Pn + i], P.mray[2 * Pn + i]);
          r.d = mk3(P.mray[3 * Pn + i], P.mray[4 * Pn + i], P.mray[5 * Pn + i]);
          r.tmin = P.mray[6 * Pn + i];
          r.tmax = INFINITY;
          DG dgg, dgs;
          float eps;
          int mat, light;
          hit_dg(S, r, P.hit[i], P.hit[Pn + i], P.hit[2 * Pn + i], prim, &dgg, &dgs, &eps, &mat, &light);
          reaches = dot3(dgg.n, neg3(r.d)) > 0.0f;  // lEmit with -wi
        }
      }
    } else {
      reaches = L.type == LT_INF;  // le of an infinite light (Light.hs:98-106)
    }
    if (reaches) xyz_accum(P.rad, Pn, i, P.pend_m);
  });
}

// Lane of the (r+1)-th set bit of m (r < popcount(m)).
__device__ __forceinline__ int nth_set_lane(unsigned long long m, int r) {
  int lo = 0, hi = 63;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (__popcll(m & ((2ull << mid) - 1ull)) > r) hi = mid;
    else lo = mid + 1;
  }
  return lo;
}

// Film splats whose footprint is one pixel (the box filter's) are merged
// within the wave before the atomics: the wave's splats are gathered in lane
// order to lanes 0..cnt-1, runs of equal (job slot, pixel) keys are summed by
// a segmented scan, and the last lane of each run adds the sums.  Samples of
// a pixel are consecutive work indices, so they often finish in neighbouring
// lanes, whose same-address atomics would otherwise serialise.
__device__ __forceinline__ void merged_splat(const PtKernelArgs& a, bool single, unsigned long long key,
                                             float v[7]) {
  const unsigned long long part = __ballot(single);
  if (part == 0ull) return;
  const int lane = __lane_id();
  const int cnt = __popcll(part);
  const bool have = lane < cnt;
  const int src = have ? nth_set_lane(part, lane) : lane;
  const unsigned long long k = __shfl(key, src);
  float u[7];
#pragma unroll
  for (int c = 0; c < 7; ++c) u[c] = __shfl(v[c], src);
  const unsigned long long kp = __shfl_up(k, 1);
  const bool head = have && (lane == 0 || kp != k);
  const unsigned long long heads = __ballot(head);
  const int start = have ? 63 - __clzll(heads & ((2ull << lane) - 1ull)) : lane;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
#pragma unroll
    for (int c = 0; c < 7; ++c) {
      const float t = __shfl_up(u[c], off);
      if (lane - off >= start) u[c] += t;
    }
  }
  const bool tail = have && (lane == cnt - 1 || (lane < 63 && ((heads >> (lane + 1)) & 1ull)));
  if (tail) {
    const DevJob& J = a.jobs[(int)(k >> 32)];
    const size_t pix = (size_t)(k & 0xffffffffull);
    float* o = J.wxyz + 4 * pix;
    atomicAdd(o, u[0]);
    atomicAdd(o + 1, u[1]);
    atomicAdd(o + 2, u[2]);
    atomicAdd(o + 3, u[3]);
    if (J.sq) {
      float* s2 = J.sq + 3 * pix;
      atomicAdd(s2, u[4]);
      atomicAdd(s2 + 1, u[5]);
      atomicAdd(s2 + 2, u[6]);
    }
  }
}

template <int BLOCK>
__global__ __launch_bounds__(BLOCK) void k_finish(PtKernelArgs a) {
  __shared__ unsigned long long s_key[BLOCK / 64 * kGroups];
  __shared__ unsigned int s_cnt[BLOCK / 64 * kGroups];
  __shared__ int s_n[BLOCK / 64];
  const int q = blockIdx.x * BLOCK + threadIdx.x;
  const unsigned int n = a.cnt->n_act;
  if ((unsigned)(blockIdx.x * BLOCK) >= n) return;  // whole block idle
  const PathPool& P = a.pool;
  const int Pn = P.P;
  const int i = (unsigned)q < n ? (int)P.q_act[q] : 0;
  const bool term = (unsigned)q < n && P.state[i] == 2;
  unsigned int* ctr = nullptr;
  bool single = false;  // a one-pixel splat, merged below
  unsigned long long key = 0ull;
  float v[7] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
  if (term) {
    const int slot = P.job[i];
    const DevJob& J = a.jobs[slot];
    ctr = J.pass_done + P.pass[i];
    if (!J.cancelled) {
      // a NaN or infinite band makes its XYZ sums NaN or infinite, so the
      // sample is dropped as the spectral check of addSample would
      float X = P.rad[i], Y = P.rad[Pn + i], Z = P.rad[2 * Pn + i];
      const bool bad = isnan(X) || isinf(X) || isnan(Y) || isinf(Y) || isnan(Z) || isinf(Z);
      if (!bad) {  // addSample (Image.hs:250-299) into the sample's tile
        X = X / BGPU_CIE_Y_SUM;
        Y = Y / BGPU_CIE_Y_SUM;
        Z = Z / BGPU_CIE_Y_SUM;
        float sx = P.sxy[i], sy = P.sxy[Pn + i];
        float fw = J.view.fw, fh = J.view.fh, ifw = 1.0f / fw, ifh = 1.0f / fw;  // Image.hs:262-263 (sic)
        float dx = sx - 0.5f, dy = sy - 0.5f;
        int x0 = (int)ceilf(dx - fw), x1 = (int)floorf(dx + fw);
        int y0 = (int)ceilf(dy - fh), y1 = (int)floorf(dy + fh);
        {  // prender's tile of the sample's pixel: splitWindow's 16 x 16
           // window of the sample extent and the image mkImageTile makes for
           // it (Sampling.hs:55-58, Image.hs:108-120); addTile drops what
           // lies beyond the film (Image.hs:178-199)
          const DevView& V = J.view;
          const int pix = P.pix[i];
          const int ix = pix % V.ext_w, iy = pix / V.ext_w;  // relative to the extent
          const int xs = V.ext_x0 + (ix / 16) * 16, ys = V.ext_y0 + (iy / 16) * 16;
          const int xe = min(xs + 15, V.ext_x0 + V.ext_w - 1), ye = min(ys + 15, V.ext_y0 + V.ext_h - 1);
          const int lx = max(0, xs), ly = max(0, ys);
          const int rx = xe + (int)floorf(0.5f + fw) - 1, ry = ye + (int)floorf(0.5f + fh) - 1;
          x0 = max(x0, lx);
          y0 = max(y0, ly);
          x1 = min(min(x1, rx), V.film_w - 1);
          y1 = min(min(y1, ry), V.film_h - 1);
        }
        if (x0 == x1 && y0 == y1) {
          int fy = (int)floorf(fabsf(((float)y0 - dy) * ifh * 16.0f));
          fy = fy < 15 ? fy : 15;
          int fx = (int)floorf(fabsf(((float)x0 - dx) * ifw * 16.0f));
          fx = fx < 15 ? fx : 15;
          const float w = J.view.filter[fy * 16 + fx];
          single = true;
          key = ((unsigned long long)slot << 32) | ((unsigned)y0 * (unsigned)J.view.film_w + (unsigned)x0);
          v[0] = w;
          v[1] = X * w;
          v[2] = Y * w;
          v[3] = Z * w;
          v[4] = X * X * w;
          v[5] = Y * Y * w;
          v[6] = Z * Z * w;
        } else {
          for (int y = y0; y <= y1; ++y) {
            int fy = (int)floorf(fabsf(((float)y - dy) * ifh * 16.0f));
            fy = fy < 15 ? fy : 15;
            for (int x = x0; x <= x1; ++x) {
              int fx = (int)floorf(fabsf(((float)x - dx) * ifw * 16.0f));
              fx = fx < 15 ? fx : 15;
              float w = J.view.filter[fy * 16 + fx];
              float* o = J.wxyz + 4 * ((size_t)y * J.view.film_w + x);
              atomicAdd(o, w);
              atomicAdd(o + 1, X * w);
              atomicAdd(o + 2, Y * w);
              atomicAdd(o + 3, Z * w);
              if (J.sq) {
                float* s2 = J.sq + 3 * ((size_t)y * J.view.film_w + x);
                atomicAdd(s2, X * X * w);
                atomicAdd(s2 + 1, Y * Y * w);
                atomicAdd(s2 + 2, Z * Z * w);
              }
            }
          }
        }
      }
    }
    P.state[i] = 0;
  }
  merged_splat(a, single, key, v);
  block_increment<BLOCK>(ctr, term, s_key, s_cnt, s_n);
}

}  // namespace

hipError_t pt_enqueue_init(const PtKernelArgs& a, hipStream_t s) {
  static bool tables = false;
  if (!tables) {
    hipError_t e = hipMemcpyToSymbol(HIP_SYMBOL(cCieX), BGPU_CIE_X, sizeof(BGPU_CIE_X));
    if (e == hipSuccess) e = hipMemcpyToSymbol(HIP_SYMBOL(cCieY), BGPU_CIE_Y, sizeof(BGPU_CIE_Y));
    if (e == hipSuccess) e = hipMemcpyToSymbol(HIP_SYMBOL(cCieZ), BGPU_CIE_Z, sizeof(BGPU_CIE_Z));
    if (e != hipSuccess) return e;
    tables = true;
  }
  const int n = a.pool.P > kRing ? a.pool.P : kRing;
  hipLaunchKernelGGL(k_init, dim3((n + 255) / 256), dim3(256), 0, s, a);
  return hipGetLastError();
}

namespace {
template <bool SPECTRUM>
__global__ void k_tex_eval(DevPtScene sc, int lo, int root, int n, const float* xyz, const float* uv, float* out) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const tx::Tex T{sc.tex_nodes, sc.images, sc.texels};
  const tx::TexPt t{mk3(xyz[3 * k], xyz[3 * k + 1], xyz[3 * k + 2]), uv[2 * k], uv[2 * k + 1]};
  if (SPECTRUM) {
    float v[16];
    tx::eval_spectrum(T, lo, root, t, v);
    for (int b = 0; b < 16; ++b) out[16 * k + b] = v[b];
  } else {
    out[k] = tx::eval_scalar(T, lo, root, t);
  }
}
}  // namespace

hipError_t pt_texture_eval(const DevPtScene& sc, int lo, int root, bool spectrum, int n, const float* xyz,
                           const float* uv, float* out, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  const dim3 g((n + 255) / 256), b(256);
  if (spectrum) hipLaunchKernelGGL(k_tex_eval<true>, g, b, 0, s, sc, lo, root, n, xyz, uv, out);
  else hipLaunchKernelGGL(k_tex_eval<false>, g, b, 0, s, sc, lo, root, n, xyz, uv, out);
  return hipGetLastError();
}

hipError_t pt_enqueue_admit(const PtKernelArgs& a, int n, hipStream_t s) {
  hipLaunchKernelGGL(k_admit, dim3(1), dim3(64), 0, s, a, n);
  return hipGetLastError();
}

// generate and finish are light per thread and aggregate their counters per
// block, so they run in the largest blocks
constexpr int kWideBlock = 1024;

// The shading kernel for the diffuse-only kind set (PART: all of the queue,
// or its diffuse classes), compiled for 1, 2 or 4 waves per SIMD (a register
// budget of 512 / waves VGPRs; what does not fit spills to scratch).  At the
// default block and 2 waves it also comes in the light-kind variants of
// a.light_mode: quad lights only (the Cornell box), or quad, sphere and
// infinite lights.
template <int BLOCK, int PART>
static void launch_diffuse(const PtKernelArgs& a, int waves, dim3 g, dim3 b, hipStream_t s) {
  if constexpr (BLOCK == 128) {
    if (waves == 2 && a.light_mode == LIGHTS_QUAD) {
      hipLaunchKernelGGL((k_shade<BLOCK, 2, kDiffuseKinds, PART, kQuadLights>), g, b, 0, s, a);
      return;
    }
    if (waves == 2 && a.light_mode == LIGHTS_QSI) {
      hipLaunchKernelGGL((k_shade<BLOCK, 2, kDiffuseKinds, PART, kQSILights>), g, b, 0, s, a);
      return;
    }
  }
  if (waves == 4) hipLaunchKernelGGL((k_shade<BLOCK, 4, kDiffuseKinds, PART>), g, b, 0, s, a);
  else if (waves == 2) hipLaunchKernelGGL((k_shade<BLOCK, 2, kDiffuseKinds, PART>), g, b, 0, s, a);
  else hipLaunchKernelGGL((k_shade<BLOCK, 1, kDiffuseKinds, PART>), g, b, 0, s, a);
}

static inline void mark(hipEvent_t* marks, int k, hipStream_t s);

template <int BLOCK>
static void launch_shade(const PtKernelArgs& a, const PtConfig& cfg, dim3 g, dim3 b, hipStream_t s,
                         hipEvent_t* marks) {
  const int waves = cfg.shade_waves;
  if (a.sort_shade) {  // group the queue by shading class first
    constexpr int kSortBlock = 1024;  // fewer blocks, fewer atomics on the class counters
    dim3 gs((a.pool.P + kSortBlock - 1) / kSortBlock), bs(kSortBlock);
    hipLaunchKernelGGL(k_reset_classes, dim3(1), dim3(64), 0, s, a);
    hipLaunchKernelGGL(k_class_count<kSortBlock>, gs, bs, 0, s, a);
    hipLaunchKernelGGL(k_class_scatter<kSortBlock>, gs, bs, 0, s, a);
  }
  mark(marks, PH_CLASSIFY + 1, s);
  if (a.light_mode == LIGHTS_SKY) {  // a sun/sky map is resident: every class in the kernel that evaluates it
    if (waves == 2) hipLaunchKernelGGL((k_shade<BLOCK, 2, kAllKinds, SHADE_ALL, kAllLightsSky>), g, b, 0, s, a);
    else hipLaunchKernelGGL((k_shade<BLOCK, 1, kAllKinds, SHADE_ALL, kAllLightsSky>), g, b, 0, s, a);
    if (a.direct) hipLaunchKernelGGL((k_shade_direct<BLOCK, kAllLightsSky>), g, b, 0, s, a);
    mark(marks, PH_SHADE + 1, s);
    return;
  }
  if (a.shade_mode == SHADE_MODE_DIFFUSE) {
    launch_diffuse<BLOCK, SHADE_ALL>(a, cfg.diffuse_waves, g, b, s);
  } else if (a.shade_mode == SHADE_MODE_SPLIT) {
    launch_diffuse<BLOCK, SHADE_DIFFUSE_PART>(a, cfg.diffuse_waves, g, b, s);
    if (waves == 2) hipLaunchKernelGGL((k_shade<BLOCK, 2, kAllKinds, SHADE_OTHER_PART>), g, b, 0, s, a);
    else hipLaunchKernelGGL((k_shade<BLOCK, 1, kAllKinds, SHADE_OTHER_PART>), g, b, 0, s, a);
  } else if (waves == 2) {
    hipLaunchKernelGGL((k_shade<BLOCK, 2, kAllKinds, SHADE_ALL>), g, b, 0, s, a);
  } else {
    hipLaunchKernelGGL((k_shade<BLOCK, 1, kAllKinds, SHADE_ALL>), g, b, 0, s, a);
  }
  if (a.direct) hipLaunchKernelGGL((k_shade_direct<BLOCK>), g, b, 0, s, a);
  mark(marks, PH_SHADE + 1, s);
}

// Records the phase boundary event k when profiling (marks non-null).
static inline void mark(hipEvent_t* marks, int k, hipStream_t s) {
  if (marks) (void)hipEventRecord(marks[k], s);
}

template <int BLOCK, int W>
static hipError_t enqueue_iter(const PtKernelArgs& a, const PtConfig& cfg, bool persistent, hipStream_t s,
                               hipEvent_t* marks) {
  dim3 g((a.pool.P + BLOCK - 1) / BLOCK), b(BLOCK);
  dim3 gw((a.pool.P + kWideBlock - 1) / kWideBlock), bw(kWideBlock);
  const size_t lds = trav::stack_lds_bytes(a.nshort, BLOCK);  // the traversal kernels' stacks
  mark(marks, 0, s);
  hipLaunchKernelGGL(k_reset, dim3(1), dim3(64), 0, s, a);
  hipLaunchKernelGGL(k_generate<kWideBlock>, gw, bw, 0, s, a);
  mark(marks, PH_GENERATE + 1, s);
  if constexpr (W != 0) {
    if (persistent) {
      // grids that fill the device once, per LDS stack depth
      static int ge[trav::kMaxShort + 1] = {}, gs[trav::kMaxShort + 1] = {}, gm[trav::kMaxShort + 1] = {};
      const int e = trav::persistent_grid_cached(ge, k_extend_p<BLOCK, W, false>, BLOCK, a.nshort);
      const int sh = trav::persistent_grid_cached(gs, k_shadow_p<BLOCK, W, false>, BLOCK, a.nshort);
      const int m = trav::persistent_grid_cached(gm, k_mis_p<BLOCK, W, false>, BLOCK, a.nshort);
      hipLaunchKernelGGL((k_extend_p<BLOCK, W, false>), dim3(e), b, lds, s, a);
      hipLaunchKernelGGL((k_extend_p<BLOCK, W, true>), dim3(e), b, lds, s, a);
      mark(marks, PH_EXTEND + 1, s);
      if (a.textured) hipLaunchKernelGGL((k_texture<BLOCK>), g, b, 0, s, a);
      launch_shade<BLOCK>(a, cfg, g, b, s, marks);
      hipLaunchKernelGGL((k_shadow_p<BLOCK, W, false>), dim3(sh), b, lds, s, a);
      hipLaunchKernelGGL((k_shadow_p<BLOCK, W, true>), dim3(sh), b, lds, s, a);
      mark(marks, PH_SHADOW + 1, s);
      hipLaunchKernelGGL((k_mis_p<BLOCK, W, false>), dim3(m), b, lds, s, a);
      hipLaunchKernelGGL((k_mis_p<BLOCK, W, true>), dim3(m), b, lds, s, a);
      hipLaunchKernelGGL((k_mis_resolve<BLOCK>), g, b, 0, s, a);
      mark(marks, PH_MIS + 1, s);
      hipLaunchKernelGGL(k_finish<kWideBlock>, gw, bw, 0, s, a);
      mark(marks, PH_FINISH + 1, s);
      return hipGetLastError();
    }
  }
  hipLaunchKernelGGL((k_extend<BLOCK, W>), g, b, lds, s, a);
  mark(marks, PH_EXTEND + 1, s);
  if (a.textured) hipLaunchKernelGGL((k_texture<BLOCK>), g, b, 0, s, a);
  launch_shade<BLOCK>(a, cfg, g, b, s, marks);
  hipLaunchKernelGGL((k_shadow<BLOCK, W>), g, b, lds, s, a);
  mark(marks, PH_SHADOW + 1, s);
  hipLaunchKernelGGL((k_mis<BLOCK, W>), g, b, lds, s, a);
  mark(marks, PH_MIS + 1, s);
  hipLaunchKernelGGL(k_finish<kWideBlock>, gw, bw, 0, s, a);
  mark(marks, PH_FINISH + 1, s);
  return hipGetLastError();
}

template <int BLOCK>
static hipError_t enqueue_iter_w(const PtKernelArgs& a, int width, const PtConfig& cfg, hipStream_t s,
                                 hipEvent_t* marks) {
  const bool p = cfg.persistent != 0;
  switch (width) {
    case 2: return enqueue_iter<BLOCK, 2>(a, cfg, p, s, marks);
    case 4: return enqueue_iter<BLOCK, 4>(a, cfg, p, s, marks);
    case 8: return enqueue_iter<BLOCK, 8>(a, cfg, p, s, marks);
    default: return enqueue_iter<BLOCK, 0>(a, cfg, p, s, marks);
  }
}

hipError_t pt_enqueue_iteration(const PtKernelArgs& args, const PtConfig& cfg, int width, hipStream_t s,
                                hipEvent_t* marks) {
  if (cfg.stack < trav::kMinShort || cfg.stack > trav::kMaxShort) return hipErrorInvalidValue;
  if (cfg.walk < 0 || cfg.walk > 64) return hipErrorInvalidValue;
  PtKernelArgs a = args;
  a.nshort = cfg.stack;
  a.walk = cfg.walk;
  switch (cfg.block) {
    case 64: return enqueue_iter_w<64>(a, width, cfg, s, marks);
    case 256: return enqueue_iter_w<256>(a, width, cfg, s, marks);
    default: return enqueue_iter_w<128>(a, width, cfg, s, marks);
  }
}

}  // namespace bgpu
