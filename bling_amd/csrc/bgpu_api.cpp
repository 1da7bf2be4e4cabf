// C ABI entry points: context, scene and ray-batch queries.
#include <algorithm>
#include <cmath>
#include <cstring>

#include "bgpu_internal.h"

namespace bgpu {

static thread_local std::string g_err;

void set_error(const std::string& m) { g_err = m; }
int fail(int code, const std::string& m) {
  g_err = m;
  return code;
}
int hip_fail(hipError_t e, const char* what) {
  g_err = std::string(what) + ": " + hipGetErrorString(e);
  return e == hipErrorOutOfMemory ? BGPU_E_NOMEM : BGPU_E_HIP;
}

DevBuf::~DevBuf() { release(); }

int DevBuf::alloc(Ctx* c, size_t n) {
  release();
  if (n == 0) n = 16;
  size_t prev = c->used.fetch_add(n);
  if (c->budget && prev + n > c->budget) {
    c->used.fetch_sub(n);
    return fail(BGPU_E_BUDGET, "HBM budget exceeded: need " + std::to_string(prev + n) +
                                   " bytes, budget " + std::to_string(c->budget));
  }
  hipError_t e = hipMalloc(&ptr, n);
  if (e != hipSuccess) {
    c->used.fetch_sub(n);
    ptr = nullptr;
    return hip_fail(e, "hipMalloc");
  }
  ctx = c;
  bytes = n;
  return BGPU_OK;
}

int DevBuf::upload(Ctx* c, const void* src, size_t n, hipStream_t s) {
  int rc = alloc(c, n);
  if (rc) return rc;
  if (n) BGPU_HIP(hipMemcpyAsync(ptr, src, n, hipMemcpyHostToDevice, s));
  return BGPU_OK;
}

void DevBuf::release() {
  if (ptr) {
    (void)hipFree(ptr);
    ctx->used.fetch_sub(bytes);
  }
  ptr = nullptr;
  bytes = 0;
}

size_t Scene::device_bytes() const {
  size_t b = 0;
  const DevBuf* bufs[] = {&nodes, &leaf, &shapes, &tri_idx, &verts, &normals, &uvs, &tri_mat,
                          &tri_flags, &shape_x, &materials, &lights, &light_maps, &light_dist,
                          &filter_tbl};
  for (const DevBuf* d : bufs) b += d->bytes;
  return b;
}

int flatten_and_upload(Ctx* ctx, const bgpu_scene_desc* d, Scene* sc);

}  // namespace bgpu

using namespace bgpu;

// ---- entry points ----------------------------------------------------------

extern "C" {

const char* bgpu_last_error(void) { return bgpu::g_err.c_str(); }
int bgpu_version(void) { return BGPU_ABI_VERSION; }

// The static scratch array's address on each device, looked up once per
// device with that device current (each GPU holds its own copy).
static std::mutex g_scratch_mu;
static std::vector<void*> g_scratch;

static int device_scratch(int device, unsigned int** out) {
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if ((int)g_scratch.size() <= device) g_scratch.resize(device + 1, nullptr);
  if (!g_scratch[device]) {
    void* p = nullptr;
    hipError_t e = scratch_address(&p);
    if (e != hipSuccess || !p) return hip_fail(e != hipSuccess ? e : hipErrorInvalidSymbol, "hipGetSymbolAddress (scratch)");
    g_scratch[device] = p;
  }
  *out = static_cast<unsigned int*>(g_scratch[device]);
  return BGPU_OK;
}

int bgpu_init(int device, size_t hbm_budget_bytes, bgpu_ctx** out) {
  if (!out) return fail(BGPU_E_INVALID, "bgpu_init: out is NULL");
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0)
    return fail(BGPU_E_NODEV, "bgpu_init: no HIP device available");
  if (device < 0 || device >= ndev) return fail(BGPU_E_INVALID, "bgpu_init: bad device index");
  BGPU_HIP(hipSetDevice(device));
  hipDeviceProp_t prop;
  BGPU_HIP(hipGetDeviceProperties(&prop, device));
  std::unique_ptr<bgpu_ctx> c(new bgpu_ctx());
  c->device = device;
  c->num_cus = prop.multiProcessorCount;
  size_t free_b = 0, total_b = 0;
  BGPU_HIP(hipMemGetInfo(&free_b, &total_b));
  c->budget = hbm_budget_bytes ? std::min(hbm_budget_bytes, total_b) : (size_t)(0.9 * (double)free_b);
  BGPU_HIP(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
  BGPU_HIP(hipEventCreate(&c->ev0));
  BGPU_HIP(hipEventCreate(&c->ev1));
  if (int rc = device_scratch(device, &c->scratch)) return rc;
  *out = c.release();
  return BGPU_OK;
}

int bgpu_mem_info(bgpu_ctx* ctx, size_t* used, size_t* budget) {
  if (!ctx || !used || !budget) return fail(BGPU_E_INVALID, "bgpu_mem_info: NULL argument");
  *used = ctx->used.load();
  *budget = ctx->budget;
  return BGPU_OK;
}

int bgpu_shutdown(bgpu_ctx* ctx) {
  if (!ctx) return fail(BGPU_E_INVALID, "bgpu_shutdown: NULL context");
  (void)hipSetDevice(ctx->device);
  delete ctx;
  return BGPU_OK;
}

int bgpu_scene_create(bgpu_ctx* ctx, const bgpu_scene_desc* desc, bgpu_scene** out) {
  if (!ctx || !desc || !out) return fail(BGPU_E_INVALID, "bgpu_scene_create: NULL argument");
  *out = nullptr;
  BGPU_HIP(hipSetDevice(ctx->device));
  std::unique_ptr<bgpu_scene> sc(new bgpu_scene());
  sc->ctx = ctx;
  int rc = flatten_and_upload(ctx, desc, sc.get());
  if (rc) return rc;
  *out = sc.release();
  return BGPU_OK;
}

int bgpu_scene_destroy(bgpu_scene* scene) {
  if (!scene) return fail(BGPU_E_INVALID, "bgpu_scene_destroy: NULL scene");
  (void)hipSetDevice(scene->ctx->device);
  if (--scene->refs == 0) delete scene;
  return BGPU_OK;
}

int bgpu_scene_bvh_info(const bgpu_scene* scene, bgpu_bvh_info* out) {
  if (!scene || !out) return fail(BGPU_E_INVALID, "bgpu_scene_bvh_info: NULL argument");
  const BuildStats& b = scene->build;
  *out = bgpu_bvh_info{b.width, b.nodes, b.leaves, b.max_depth, b.stack_bound, b.device, b.build_ms};
  return BGPU_OK;
}

int bgpu_scene_texture_eval(bgpu_scene* scene, int node, size_t n, const float* xyz, const float* uv, float* out) {
  if (!scene || (n && (!xyz || !uv || !out))) return fail(BGPU_E_INVALID, "bgpu_scene_texture_eval: NULL argument");
  const PtScene& pt = scene->pt;
  if (node < 0 || (size_t)node >= pt.tex_op_host.size())
    return fail(BGPU_E_INVALID, "bgpu_scene_texture_eval: no such texture node");
  const int op = pt.tex_op_host[node];
  const bool spectrum = op >= BGPU_TN_CONST && op <= BGPU_TN_GRAPHPAPER;
  if (!spectrum && !(op >= BGPU_TN_SCONST && op <= BGPU_TN_SCALE))
    return fail(BGPU_E_INVALID, "bgpu_scene_texture_eval: the node is not a texture (a gradient step)");
  const int lo = pt.tex_lo_host[node];
  if (node - lo + 1 > 64) return fail(BGPU_E_INVALID, "bgpu_scene_texture_eval: texture of more than 64 nodes");
  if (n > (size_t)0x7fffffff) return fail(BGPU_E_INVALID, "bgpu_scene_texture_eval: too many points");
  if (n == 0) return BGPU_OK;
  Ctx* c = scene->ctx;
  std::lock_guard<std::mutex> lk(c->mu);
  BGPU_HIP(hipSetDevice(c->device));
  const size_t nout = n * (spectrum ? 16 : 1);
  DevBuf dx, du, dout;
  if (int rc = dx.upload(c, xyz, 12 * n, c->stream)) return rc;
  if (int rc = du.upload(c, uv, 8 * n, c->stream)) return rc;
  if (int rc = dout.alloc(c, 4 * nout)) return rc;
  BGPU_HIP(pt_texture_eval(pt.host, lo, node, spectrum, (int)n, dx.as<float>(), du.as<float>(), dout.as<float>(),
                           c->stream));
  BGPU_HIP(hipMemcpyAsync(out, dout.ptr, 4 * nout, hipMemcpyDeviceToHost, c->stream));
  BGPU_HIP(hipStreamSynchronize(c->stream));
  return BGPU_OK;
}

int bgpu_scene_bytes(const bgpu_scene* scene, size_t* bytes) {
  if (!scene || !bytes) return fail(BGPU_E_INVALID, "bgpu_scene_bytes: NULL argument");
  *bytes = scene->device_bytes();
  return BGPU_OK;
}

static int check_rays(const bgpu_rays* r) {
  if (!r || !r->ox || !r->oy || !r->oz || !r->dx || !r->dy || !r->dz || !r->tmin || !r->tmax)
    return fail(BGPU_E_INVALID, "ray batch has a NULL array");
  return BGPU_OK;
}

// Stages host rays into the context's buffer; returns device views.
static int stage_rays(Ctx* c, const bgpu_rays* r, size_t n, DevRays* dr) {
  size_t need = 8 * n * sizeof(float);
  if (c->stage_rays.bytes < need) {
    int rc = c->stage_rays.alloc(c, need);
    if (rc) return rc;
  }
  float* base = c->stage_rays.as<float>();
  const float* src[8] = {r->ox, r->oy, r->oz, r->dx, r->dy, r->dz, r->tmin, r->tmax};
  for (int k = 0; k < 8; ++k)
    BGPU_HIP(hipMemcpyAsync(base + k * n, src[k], n * sizeof(float), hipMemcpyHostToDevice, c->stream));
  *dr = DevRays{base, base + n, base + 2 * n, base + 3 * n, base + 4 * n, base + 5 * n, base + 6 * n, base + 7 * n};
  return BGPU_OK;
}

static int stage_out(Ctx* c, size_t need) {
  if (c->stage_hits.bytes < need) return c->stage_hits.alloc(c, need);
  return BGPU_OK;
}

// BGPU_COUNT: zeroed device counters for the counting kernel variant (4
// 64-bit: triangle-tree node fetches and tests, shape-tree node fetches and
// tests).
static int count_begin(Ctx* c, int flags, unsigned long long** cnt) {
  *cnt = nullptr;
  if (!(flags & BGPU_COUNT)) return BGPU_OK;
  if (!c->count_buf.ptr)
    if (int rc = c->count_buf.alloc(c, 4 * sizeof(unsigned long long))) return rc;
  BGPU_HIP(zero_words(c->count_buf.as<unsigned int>(), 8, c->stream));
  *cnt = c->count_buf.as<unsigned long long>();
  return BGPU_OK;
}

static int count_end(Ctx* c, unsigned long long* cnt) {
  if (!cnt) return BGPU_OK;
  unsigned long long h[4];
  BGPU_HIP(hipMemcpy(h, cnt, sizeof(h), hipMemcpyDeviceToHost));
  std::lock_guard<std::mutex> lk(c->stats_mu);
  c->stats.nodes_visited += h[0] + h[2];
  c->stats.prims_tested += h[1] + h[3];
  c->stats.shape_nodes_visited += h[2];
  c->stats.shapes_tested += h[3];
  return BGPU_OK;
}

int bgpu_intersect_closest(bgpu_scene* scene, const bgpu_rays* rays, size_t n, bgpu_hits* hits, int flags) {
  if (!scene || !hits || !hits->t || !hits->b1 || !hits->b2 || !hits->prim)
    return fail(BGPU_E_INVALID, "bgpu_intersect_closest: NULL argument");
  if (int rc = check_rays(rays)) return rc;
  if (n > (size_t)0x7fffffff) return fail(BGPU_E_INVALID, "ray batch larger than 2^31-1");
  if (n == 0) return BGPU_OK;
  Ctx* c = scene->ctx;
  std::lock_guard<std::mutex> lk(c->mu);
  BGPU_HIP(hipSetDevice(c->device));
  DevRays dr;
  DevHits dh;
  bool dev = flags & BGPU_DEVICE_PTRS;
  if (dev) {
    dr = DevRays{rays->ox, rays->oy, rays->oz, rays->dx, rays->dy, rays->dz, rays->tmin, rays->tmax};
    dh = DevHits{hits->t, hits->b1, hits->b2, hits->prim};
  } else {
    if (int rc = stage_rays(c, rays, n, &dr)) return rc;
    if (int rc = stage_out(c, 16 * n)) return rc;
    float* o = c->stage_hits.as<float>();
    dh = DevHits{o, o + n, o + 2 * n, reinterpret_cast<int32_t*>(o + 3 * n)};
  }
  unsigned long long* cnt;
  if (int rc = count_begin(c, flags, &cnt)) return rc;
  if (flags & BGPU_TIMED) BGPU_HIP(hipEventRecord(c->ev0, c->stream));
  if (!c->work_buf.ptr)
    if (int rc = c->work_buf.alloc(c, 256)) return rc;
  TraceConfig tc = c->trace_cfg;
  if (flags & BGPU_COHERENT) tc.persistent = 0;
  BGPU_HIP(launch_closest(scene->dev, dr, (int)n, dh, tc, c->stream, cnt, c->work_buf.as<unsigned int>()));
  if (flags & BGPU_TIMED) BGPU_HIP(hipEventRecord(c->ev1, c->stream));
  if (!dev) {
    BGPU_HIP(hipMemcpyAsync(hits->t, dh.t, n * 4, hipMemcpyDeviceToHost, c->stream));
    BGPU_HIP(hipMemcpyAsync(hits->b1, dh.b1, n * 4, hipMemcpyDeviceToHost, c->stream));
    BGPU_HIP(hipMemcpyAsync(hits->b2, dh.b2, n * 4, hipMemcpyDeviceToHost, c->stream));
    BGPU_HIP(hipMemcpyAsync(hits->prim, dh.prim, n * 4, hipMemcpyDeviceToHost, c->stream));
  }
  BGPU_HIP(hipStreamSynchronize(c->stream));
  if (int rc = count_end(c, cnt)) return rc;
  if (flags & BGPU_TIMED) {
    float ms = 0.0f;
    BGPU_HIP(hipEventElapsedTime(&ms, c->ev0, c->ev1));
    c->last_ms = ms;
  }
  return BGPU_OK;
}

int bgpu_intersect_closest_noalloc(bgpu_scene* scene, const bgpu_rays* rays, size_t n, bgpu_hits* hits,
                                   int flags) {
  if (!scene || !hits || !hits->t || !hits->b1 || !hits->b2 || !hits->prim)
    return fail(BGPU_E_INVALID, "bgpu_intersect_closest_noalloc: NULL argument");
  if (int rc = check_rays(rays)) return rc;
  if (!(flags & BGPU_DEVICE_PTRS))
    return fail(BGPU_E_INVALID, "bgpu_intersect_closest_noalloc: rays and hits must be device pointers "
                                "(BGPU_DEVICE_PTRS); host batches need staging buffers");
  if (n > (size_t)0x7fffffff)
    return fail(BGPU_E_INVALID, "bgpu_intersect_closest_noalloc: ray batch larger than 2^31-1");
  if (n == 0) return BGPU_OK;
  Ctx* c = scene->ctx;
  std::lock_guard<std::mutex> lk(c->mu);
  BGPU_HIP(hipSetDevice(c->device));
  const DevRays dr{rays->ox, rays->oy, rays->oz, rays->dx, rays->dy, rays->dz, rays->tmin, rays->tmax};
  const DevHits dh{hits->t, hits->b1, hits->b2, hits->prim};
  unsigned int* work = c->scratch;  // words 0-1: the passes' counters (zeroed by the launch)
  unsigned long long* cnt = nullptr;
  if (flags & BGPU_COUNT) {  // words 8-15: the four 64-bit work counters
    BGPU_HIP(zero_words(c->scratch + 8, 8, c->stream));
    cnt = reinterpret_cast<unsigned long long*>(c->scratch + 8);
  }
  if (flags & BGPU_TIMED) BGPU_HIP(hipEventRecord(c->ev0, c->stream));
  TraceConfig tc = c->trace_cfg;
  if (flags & BGPU_COHERENT) tc.persistent = 0;
  BGPU_HIP(launch_closest(scene->dev, dr, (int)n, dh, tc, c->stream, cnt, work));
  if (flags & BGPU_TIMED) BGPU_HIP(hipEventRecord(c->ev1, c->stream));
  BGPU_HIP(hipStreamSynchronize(c->stream));
  if (int rc = count_end(c, cnt)) return rc;
  if (flags & BGPU_TIMED) {
    float ms = 0.0f;
    BGPU_HIP(hipEventElapsedTime(&ms, c->ev0, c->ev1));
    c->last_ms = ms;
  }
  return BGPU_OK;
}

int bgpu_intersect_any(bgpu_scene* scene, const bgpu_rays* rays, size_t n, uint8_t* occluded, int flags) {
  if (!scene || !occluded) return fail(BGPU_E_INVALID, "bgpu_intersect_any: NULL argument");
  if (int rc = check_rays(rays)) return rc;
  if (n > (size_t)0x7fffffff) return fail(BGPU_E_INVALID, "ray batch larger than 2^31-1");
  if (n == 0) return BGPU_OK;
  Ctx* c = scene->ctx;
  std::lock_guard<std::mutex> lk(c->mu);
  BGPU_HIP(hipSetDevice(c->device));
  DevRays dr;
  uint8_t* docc;
  bool dev = flags & BGPU_DEVICE_PTRS;
  if (dev) {
    dr = DevRays{rays->ox, rays->oy, rays->oz, rays->dx, rays->dy, rays->dz, rays->tmin, rays->tmax};
    docc = occluded;
  } else {
    if (int rc = stage_rays(c, rays, n, &dr)) return rc;
    if (int rc = stage_out(c, n)) return rc;
    docc = c->stage_hits.as<uint8_t>();
  }
  unsigned long long* cnt;
  if (int rc = count_begin(c, flags, &cnt)) return rc;
  if (flags & BGPU_TIMED) BGPU_HIP(hipEventRecord(c->ev0, c->stream));
  if (!c->work_buf.ptr)
    if (int rc = c->work_buf.alloc(c, 256)) return rc;
  TraceConfig tc = c->trace_cfg;
  if (flags & BGPU_COHERENT) tc.persistent = 0;
  BGPU_HIP(launch_any(scene->dev, dr, (int)n, docc, tc, c->stream, cnt, c->work_buf.as<unsigned int>()));
  if (flags & BGPU_TIMED) BGPU_HIP(hipEventRecord(c->ev1, c->stream));
  if (!dev) BGPU_HIP(hipMemcpyAsync(occluded, docc, n, hipMemcpyDeviceToHost, c->stream));
  BGPU_HIP(hipStreamSynchronize(c->stream));
  if (int rc = count_end(c, cnt)) return rc;
  if (flags & BGPU_TIMED) {
    float ms = 0.0f;
    BGPU_HIP(hipEventElapsedTime(&ms, c->ev0, c->ev1));
    c->last_ms = ms;
  }
  return BGPU_OK;
}

int bgpu_last_kernel_ms(bgpu_ctx* ctx, double* ms) {
  if (!ctx || !ms) return fail(BGPU_E_INVALID, "bgpu_last_kernel_ms: NULL argument");
  *ms = ctx->last_ms;
  return BGPU_OK;
}

int bgpu_stats_get(bgpu_ctx* ctx, bgpu_stats* out) {
  if (!ctx || !out) return fail(BGPU_E_INVALID, "bgpu_stats_get: NULL argument");
  std::lock_guard<std::mutex> lk(ctx->stats_mu);
  *out = ctx->stats;
  return BGPU_OK;
}

int bgpu_stats_reset(bgpu_ctx* ctx) {
  if (!ctx) return fail(BGPU_E_INVALID, "bgpu_stats_reset: NULL context");
  std::lock_guard<std::mutex> lk(ctx->stats_mu);
  ctx->stats = bgpu_stats{};
  ctx->prof = bgpu_profile{};
  return BGPU_OK;
}

}  // extern "C"
