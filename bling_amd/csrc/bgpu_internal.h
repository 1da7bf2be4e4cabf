// Internal state behind the opaque bgpu handles.
#pragma once
#include <hip/hip_runtime.h>

#include <atomic>
#include <condition_variable>
#include <cstdint>
#include <deque>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/bgpu.h"
#include "bvh_build.h"
#include "bvh_device.h"
#include "isect.h"
#include "pt.h"

namespace bgpu {

void set_error(const std::string& msg);
int fail(int code, const std::string& msg);
int hip_fail(hipError_t e, const char* what);

#define BGPU_HIP(call)                                          \
  do {                                                          \
    hipError_t _e = (call);                                     \
    if (_e != hipSuccess) return ::bgpu::hip_fail(_e, #call);   \
  } while (0)

struct Ctx;

// Device allocation accounted against the context's HBM budget.
struct DevBuf {
  Ctx* ctx = nullptr;
  void* ptr = nullptr;
  size_t bytes = 0;
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  ~DevBuf();
  int alloc(Ctx* c, size_t n);  // 0 or BGPU_E_*
  int upload(Ctx* c, const void* src, size_t n, hipStream_t s);
  void release();
  template <class T>
  T* as() const { return static_cast<T*>(ptr); }
};

struct Scene;
struct Job;
struct Scheduler;
struct Engine;
void destroy_scheduler(Scheduler* s);
void destroy_engine(Engine* e);

struct Ctx {
  int device = 0;
  int num_cus = 0;
  size_t budget = 0;
  std::atomic<size_t> used{0};
  std::atomic<size_t> film_bytes{0};  // job films allocated (running or finished, not yet released)
  hipStream_t stream = nullptr;   // intersect / plain render stream
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  double last_ms = 0.0;
  std::mutex mu;                  // serialises submissions on `stream`
  // staging for host-pointer ray batches
  DevBuf stage_rays, stage_hits;
  DevBuf count_buf;               // BGPU_COUNT traversal counters
  DevBuf work_buf;                // ray counters of the persistent traversal kernels
  unsigned int* scratch = nullptr;  // this GPU's static scratch array (bgpu_intersect_closest_noalloc)
  std::mutex stats_mu;
  bgpu_stats stats{};
  bgpu_profile prof{};            // per-phase times while pt.profile is 1
  Scheduler* sched = nullptr;     // served path, created on first job_submit
  Engine* plain = nullptr;        // plain path engine, reused by bgpu_render calls
  std::mutex plain_mu;            // one plain render at a time
  std::atomic<int> live_jobs{0};  // served job handles not yet deleted
  TraceConfig trace_cfg;
  PtConfig pt_cfg;
  int bvh_width = 4;              // children per BVH node of scenes created next
  int bvh_device = 1;             // 1: build large triangle trees on the GPU
  ~Ctx();
};

struct Scene {
  Ctx* ctx = nullptr;
  uint32_t ntri = 0, nshape = 0;
  DevBuf nodes, leaf, shapes;
  DevScene dev{};
  BuildStats build{};
  // shading data for the path tracer
  PtScene pt{};
  DevBuf tri_idx, verts, normals, uvs, tri_mat, tri_flags, shape_x, materials, lights,
      light_maps, light_dist, filter_tbl, tex_nodes, tex_lo, images, texels, mat_src;
  bgpu_film film{};
  bgpu_camera camera{};
  std::atomic<int> refs{1};  // the user's handle plus engines referencing it
  size_t device_bytes() const;
  virtual ~Scene() = default;
};

// Host-side flattening of a scene description (shared by scene creation).
int flatten_and_upload(Ctx* ctx, const bgpu_scene_desc* d, Scene* sc);

}  // namespace bgpu

struct bgpu_ctx : bgpu::Ctx {};
struct bgpu_scene : bgpu::Scene {};
