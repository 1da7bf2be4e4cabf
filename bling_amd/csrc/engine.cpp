// Render engine: a pool of P path slots, a job table and one captured
// bounce iteration, replayed until the admitted work is done.
//
//   bgpu_render      the plain path: one job on a private engine, blocking
//                    (prender, Rendering.hs:207-236, all passes).
//   bgpu_job_*       the served path: a scheduler thread owns one engine;
//                    jobs arriving while others run are admitted as job-table
//                    slots and HBM (film buffers) free up, and their samples
//                    share the path pool with everything in flight
//                    (continuous batching).  Progress is polled per pass
//                    (PassDone, Rendering.hs:165-169).
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>

#include "bgpu_internal.h"

namespace bgpu {

static double now_ms() {
  return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch())
      .count();
}

struct HostJob {
  bgpu_job_desc desc{};
  Scene* scene = nullptr;  // holds a reference while the job exists
  DevView view{};          // the job's camera and film
  size_t px = 0;           // film pixels
  int slot = -1;
  unsigned long long total = 0, per_pass = 0;
  DevBuf film, sq, pass_done, skipped;
  std::vector<unsigned int> h_pass;  // host copy of per-pass counts
  unsigned long long h_skipped = 0;
  int state = BGPU_JOB_QUEUED;
  int passes_done = 0;
  double t_submit = 0.0, t_done = 0.0;
  bool cancel_req = false;
  bool released = false;  // user released the handle
  size_t bytes = 0;

  // Copies a validated description; the camera and film default to the
  // scene's (the caller's descriptors are not referenced afterwards).
  int prepare(Scene* s, const bgpu_job_desc& d) {
    if (int rc = make_view(d.camera ? *d.camera : s->camera, d.film ? *d.film : s->film, &view)) return rc;
    desc = d;
    desc.camera = nullptr;
    desc.film = nullptr;
    px = (size_t)view.film_w * view.film_h;
    per_pass = (unsigned long long)view.ext_w * view.ext_h * (unsigned long long)(d.spp_nu * d.spp_nv);
    if (per_pass > (1ull << 62) / (unsigned long long)d.passes)
      return fail(BGPU_E_INVALID, "job has too many samples");
    total = per_pass * (unsigned long long)d.passes;
    scene = s;
    s->refs++;
    return BGPU_OK;
  }
  virtual ~HostJob() {
    film.release();
    sq.release();
    pass_done.release();
    skipped.release();
    if (scene && --scene->refs == 0) {
      (void)hipSetDevice(scene->ctx->device);
      delete scene;
    }
  }
};

struct Engine {
  Ctx* ctx = nullptr;
  hipStream_t stream = nullptr;
  int P = 0, block = 128;
  DevBuf pool_buf, jobs_buf, ring_buf, scenes_buf, cnt_buf, admit_buf;
  PtKernelArgs args{};
  AdmitReq* h_admit = nullptr;
  PtCounters* h_cnt = nullptr;
  DevJob* h_job = nullptr;
  DevPtScene* h_scene = nullptr;
  unsigned int* h_pass = nullptr;  // pinned staging for progress reads
  size_t h_pass_cap = 0;
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;
  std::vector<Scene*> scenes;             // table index -> scene
  std::vector<HostJob*> slots;            // job slot -> job
  std::vector<int> free_slots;
  // Retired slots wait until a bounce iteration has run after retirement:
  // its reset step drops the finished job from the device's job ring, after
  // which the slot's table entry may be overwritten.
  std::vector<std::pair<int, unsigned int>> quarantine;  // (slot, iteration at retirement)

  ~Engine() {
    if (stream) (void)hipStreamSynchronize(stream);
    if (gexec) (void)hipGraphExecDestroy(gexec);
    if (graph) (void)hipGraphDestroy(graph);
    if (h_admit) (void)hipHostFree(h_admit);
    if (h_cnt) (void)hipHostFree(h_cnt);
    if (h_job) (void)hipHostFree(h_job);
    if (h_scene) (void)hipHostFree(h_scene);
    if (h_pass) (void)hipHostFree(h_pass);
    for (Scene* s : scenes)
      if (s && --s->refs == 0) delete s;
    pool_buf.release();
    jobs_buf.release();
    ring_buf.release();
    scenes_buf.release();
    cnt_buf.release();
    admit_buf.release();
    if (stream) (void)hipStreamDestroy(stream);
  }

  int init(Ctx* c, int pool, int blk) {
    ctx = c;
    P = std::max(1024, (pool + 255) / 256 * 256);
    block = blk;
    BGPU_HIP(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
    // pool layout: 11 int arrays + 91 float arrays of P, plus 3 queues
    size_t per = (size_t)P * 4;
    size_t total = per * (8 + 91 + 3);
    if (int rc = pool_buf.alloc(c, total)) return rc;
    char* b = pool_buf.as<char>();
    auto take = [&](size_t n) {
      char* r = b;
      b += n * per;
      return r;
    };
    PathPool& pp = args.pool;
    pp.P = P;
    pp.state = (int32_t*)take(1);
    pp.job = (int32_t*)take(1);
    pp.pass = (int32_t*)take(1);
    pp.pix = (int32_t*)take(1);
    pp.samp = (int32_t*)take(1);
    pp.dspec = (int32_t*)take(1);
    pp.hprim = (int32_t*)take(1);
    pp.mlight = (int32_t*)take(1);
    pp.sxy = (float*)take(2);
    pp.ray = (float*)take(7);
    pp.hit = (float*)take(3);
    pp.thr = (float*)take(16);
    pp.rad = (float*)take(16);
    pp.sray = (float*)take(8);
    pp.pend_s = (float*)take(16);
    pp.mray = (float*)take(7);
    pp.pend_m = (float*)take(16);
    pp.q_act = (uint32_t*)take(1);
    pp.q_shadow = (uint32_t*)take(1);
    pp.q_mis = (uint32_t*)take(1);
    if (int rc = jobs_buf.alloc(c, sizeof(DevJob) * kMaxJobs)) return rc;
    if (int rc = ring_buf.alloc(c, sizeof(uint32_t) * kMaxJobs)) return rc;
    if (int rc = scenes_buf.alloc(c, sizeof(DevPtScene) * kMaxScenes)) return rc;
    if (int rc = cnt_buf.alloc(c, sizeof(PtCounters))) return rc;
    if (int rc = admit_buf.alloc(c, sizeof(AdmitReq) * kMaxJobs)) return rc;
    args.scenes = scenes_buf.as<DevPtScene>();
    args.jobs = jobs_buf.as<DevJob>();
    args.ring = ring_buf.as<uint32_t>();
    args.cnt = cnt_buf.as<PtCounters>();
    args.admit = admit_buf.as<AdmitReq>();
    BGPU_HIP(hipHostMalloc((void**)&h_admit, sizeof(AdmitReq) * kMaxJobs, hipHostMallocDefault));
    BGPU_HIP(hipHostMalloc((void**)&h_cnt, sizeof(PtCounters), hipHostMallocDefault));
    BGPU_HIP(hipHostMalloc((void**)&h_job, sizeof(DevJob), hipHostMallocDefault));
    BGPU_HIP(hipHostMalloc((void**)&h_scene, sizeof(DevPtScene), hipHostMallocDefault));
    slots.assign(kMaxJobs, nullptr);
    for (int i = kMaxJobs - 1; i >= 0; --i) free_slots.push_back(i);
    BGPU_HIP(pt_enqueue_init(args, stream));
    // capture one bounce iteration
    BGPU_HIP(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
    hipError_t e = pt_enqueue_iteration(args, block, stream);
    hipError_t e2 = hipStreamEndCapture(stream, &graph);
    if (e != hipSuccess) return hip_fail(e, "pt_enqueue_iteration (capture)");
    if (e2 != hipSuccess) return hip_fail(e2, "hipStreamEndCapture");
    BGPU_HIP(hipGraphInstantiate(&gexec, graph, nullptr, nullptr, 0));
    BGPU_HIP(hipStreamSynchronize(stream));
    return BGPU_OK;
  }

  int scene_index(Scene* s) {
    for (size_t i = 0; i < scenes.size(); ++i)
      if (scenes[i] == s) return (int)i;
    int idx = -1;
    for (size_t i = 0; i < scenes.size(); ++i)
      if (!scenes[i]) { idx = (int)i; break; }
    if (idx < 0) {
      if ((int)scenes.size() >= kMaxScenes) return fail(BGPU_E_BUDGET, "too many scenes in flight");
      idx = (int)scenes.size();
      scenes.push_back(nullptr);
    }
    *h_scene = s->pt.host;
    h_scene->geo = s->dev;
    BGPU_HIP(hipMemcpyAsync((void*)(args.scenes + idx), h_scene, sizeof(DevPtScene), hipMemcpyHostToDevice, stream));
    BGPU_HIP(hipStreamSynchronize(stream));
    s->refs++;
    scenes[idx] = s;
    return idx;
  }

  // Drops scenes no admitted job references.
  void gc_scenes() {
    for (size_t i = 0; i < scenes.size(); ++i) {
      Scene* s = scenes[i];
      if (!s) continue;
      bool used = false;
      for (HostJob* j : slots)
        if (j && j->scene == s) used = true;
      if (!used) {
        scenes[i] = nullptr;
        if (--s->refs == 0) delete s;
      }
    }
  }

  static size_t job_bytes(const HostJob* j) {
    size_t b = j->px * 16 + (size_t)j->desc.passes * 4 + 8;
    if (j->desc.flags & BGPU_JOB_VARIANCE) b += j->px * 12;
    return b;
  }

  // Allocates the job's device state and queues its admission.
  int admit(HostJob* j) {
    if (free_slots.empty()) return fail(BGPU_E_BUDGET, "job table full");
    int sidx = scene_index(j->scene);
    if (sidx < 0) return sidx;
    const bgpu_job_desc& d = j->desc;
    const size_t px = j->px;
    if (int rc = j->film.alloc(ctx, px * 16)) return rc;
    if (d.flags & BGPU_JOB_VARIANCE)
      if (int rc = j->sq.alloc(ctx, px * 12)) return rc;
    if (int rc = j->pass_done.alloc(ctx, (size_t)d.passes * 4)) return rc;
    if (int rc = j->skipped.alloc(ctx, 8)) return rc;
    BGPU_HIP(hipMemsetAsync(j->film.ptr, 0, px * 16, stream));
    if (j->sq.ptr) BGPU_HIP(hipMemsetAsync(j->sq.ptr, 0, px * 12, stream));
    BGPU_HIP(hipMemsetAsync(j->pass_done.ptr, 0, (size_t)d.passes * 4, stream));
    BGPU_HIP(hipMemsetAsync(j->skipped.ptr, 0, 8, stream));
    int slot = free_slots.back();
    free_slots.pop_back();
    j->slot = slot;
    slots[slot] = j;
    j->h_pass.assign(d.passes, 0);
    DevJob& J = *h_job;
    std::memset(&J, 0, sizeof(J));
    J.per_pass = j->per_pass;
    J.seed = d.seed;
    J.scene = sidx;
    J.nu = d.spp_nu;
    J.nv = d.spp_nv;
    J.N = d.spp_nu * d.spp_nv;
    J.md = d.max_depth;
    J.sd = d.sample_depth;
    J.rr = d.rr_start_depth;
    J.passes = d.passes;
    J.wxyz = j->film.as<float>();
    J.sq = j->sq.ptr ? j->sq.as<float>() : nullptr;
    J.pass_done = j->pass_done.as<unsigned int>();
    J.skipped = j->skipped.as<unsigned long long>();
    J.cancelled = 0;
    J.view = j->view;
    BGPU_HIP(hipMemcpyAsync(args.jobs + slot, h_job, sizeof(DevJob), hipMemcpyHostToDevice, stream));
    h_admit[0] = AdmitReq{slot, 1, j->total};
    BGPU_HIP(hipMemcpyAsync(args.admit, h_admit, sizeof(AdmitReq), hipMemcpyHostToDevice, stream));
    BGPU_HIP(pt_enqueue_admit(args, 1, stream));
    BGPU_HIP(hipStreamSynchronize(stream));  // pinned staging is reused
    j->state = BGPU_JOB_RUNNING;
    return BGPU_OK;
  }

  int cancel(HostJob* j) {
    if (j->slot < 0) return BGPU_OK;
    int one = 1;
    BGPU_HIP(hipMemcpyAsync(&args.jobs[j->slot].cancelled, &one, sizeof(int), hipMemcpyHostToDevice, stream));
    BGPU_HIP(hipStreamSynchronize(stream));
    return BGPU_OK;
  }

  int run(int iters) {
    for (int k = 0; k < iters; ++k) BGPU_HIP(hipGraphLaunch(gexec, stream));
    return BGPU_OK;
  }

  int read_counters() {
    BGPU_HIP(hipMemcpyAsync(h_cnt, args.cnt, sizeof(PtCounters), hipMemcpyDeviceToHost, stream));
    BGPU_HIP(hipStreamSynchronize(stream));
    return BGPU_OK;
  }

  // Reads a job's per-pass completion; true when every sample is accounted for.
  int progress(HostJob* j, bool* complete) {
    // per-pass counts, padded to 8 bytes, then the 8-byte skipped counter
    size_t need = (size_t)(j->desc.passes + (j->desc.passes & 1)) * 4 + 8;
    if (h_pass_cap < need) {
      if (h_pass) (void)hipHostFree(h_pass);
      h_pass = nullptr;
      BGPU_HIP(hipHostMalloc((void**)&h_pass, need, hipHostMallocDefault));
      h_pass_cap = need;
    }
    BGPU_HIP(hipMemcpyAsync(h_pass, j->pass_done.ptr, (size_t)j->desc.passes * 4, hipMemcpyDeviceToHost, stream));
    BGPU_HIP(hipMemcpyAsync(h_pass + j->desc.passes + (j->desc.passes & 1), j->skipped.ptr, 8,
                            hipMemcpyDeviceToHost, stream));
    BGPU_HIP(hipStreamSynchronize(stream));
    std::memcpy(j->h_pass.data(), h_pass, (size_t)j->desc.passes * 4);
    std::memcpy(&j->h_skipped, h_pass + j->desc.passes + (j->desc.passes & 1), 8);
    unsigned long long done = j->h_skipped;
    int pd = 0;
    bool prefix = true;
    for (int p = 0; p < j->desc.passes; ++p) {
      done += j->h_pass[p];
      if (prefix && j->h_pass[p] == j->per_pass) pd++;
      else prefix = false;
    }
    j->passes_done = pd;
    *complete = done >= j->total;
    return BGPU_OK;
  }

  void retire(HostJob* j) {
    if (j->slot >= 0) {
      slots[j->slot] = nullptr;
      quarantine.emplace_back(j->slot, h_cnt->iter);
      j->slot = -1;
    }
  }

  // Frees quarantined slots once an iteration has run since their retirement
  // (h_cnt must be fresh).
  void release_quarantine() {
    for (auto it = quarantine.begin(); it != quarantine.end();) {
      if (h_cnt->iter > it->second) {
        free_slots.push_back(it->first);
        it = quarantine.erase(it);
      } else {
        ++it;
      }
    }
  }
};

static int validate_job(const bgpu_job_desc* d) {
  if (!d) return fail(BGPU_E_INVALID, "job description is NULL");
  if (d->spp_nu <= 0 || d->spp_nv <= 0)
    return fail(BGPU_E_INVALID, "mkStratifiedSampler: nx and ny must be > 0");
  if ((long long)d->spp_nu * d->spp_nv > (1 << 20)) return fail(BGPU_E_INVALID, "too many samples per pixel");
  if (d->max_depth < 0 || d->max_depth > 1000) return fail(BGPU_E_INVALID, "max_depth must be in [0, 1000]");
  if (d->sample_depth < 0 || d->sample_depth > 1000) return fail(BGPU_E_INVALID, "sample_depth must be in [0, 1000]");
  if (d->passes < 1 || d->passes > (1 << 20)) return fail(BGPU_E_INVALID, "passes must be >= 1");
  return BGPU_OK;
}

static void accumulate_stats(Ctx* ctx, const PtCounters& before, const PtCounters& after, double ms,
                             uint64_t samples) {
  std::lock_guard<std::mutex> lk(ctx->stats_mu);
  bgpu_stats& s = ctx->stats;
  s.rays_camera += after.tot_camera - before.tot_camera;
  s.rays_extend += (after.tot_extend + after.n_act) - (before.tot_extend + before.n_act);
  s.rays_shadow += (after.tot_shadow + after.n_shadow) - (before.tot_shadow + before.n_shadow);
  s.rays_mis += (after.tot_mis + after.n_mis) - (before.tot_mis + before.n_mis);
  s.samples += samples;
  s.gpu_ms += ms;
}

// ---------------------------------------------------------------- served
struct Scheduler {
  Ctx* ctx;
  Engine eng;
  std::thread th;
  std::mutex mu;
  std::condition_variable cv, done_cv;
  std::deque<HostJob*> pending;
  std::vector<HostJob*> running;
  bool stop = false;
  int error = 0;
  std::string error_msg;

  int start(Ctx* c) {
    ctx = c;
    if (int rc = eng.init(c, c->pt_cfg.pool, c->pt_cfg.block)) return rc;
    th = std::thread([this] { loop(); });
    return BGPU_OK;
  }

  void shutdown() {
    {
      std::lock_guard<std::mutex> lk(mu);
      stop = true;
    }
    cv.notify_all();
    if (th.joinable()) th.join();
    std::lock_guard<std::mutex> lk(mu);
    for (HostJob* j : running) {
      j->state = BGPU_JOB_CANCELLED;
      eng.retire(j);
    }
    for (HostJob* j : pending) j->state = BGPU_JOB_CANCELLED;
    for (HostJob* j : running)
      if (j->released) delete j;
    for (HostJob* j : pending)
      if (j->released) delete j;
    running.clear();
    pending.clear();
    done_cv.notify_all();
  }

  void set_error(int code, const std::string& m) {
    error = code;
    error_msg = m;
  }

  void loop() {
    (void)hipSetDevice(ctx->device);
    PtCounters prev{};
    bool have_prev = false;
    while (true) {
      std::vector<HostJob*> to_admit;
      {
        std::unique_lock<std::mutex> lk(mu);
        cv.wait(lk, [&] { return stop || !pending.empty() || !running.empty(); });
        if (stop) return;
        // cancellations of queued jobs never reach the device
        for (auto it = pending.begin(); it != pending.end();) {
          if ((*it)->cancel_req) {
            (*it)->state = BGPU_JOB_CANCELLED;
            (*it)->t_done = now_ms();
            HostJob* j = *it;
            it = pending.erase(it);
            if (j->released) delete j;
            done_cv.notify_all();
          } else {
            ++it;
          }
        }
        // admit by HBM budget, in arrival order
        while (!pending.empty()) {
          HostJob* j = pending.front();
          size_t need = Engine::job_bytes(j);
          if (ctx->used.load() + need > ctx->budget && !running.empty()) break;
          if (eng.free_slots.empty()) break;
          pending.pop_front();
          to_admit.push_back(j);
        }
      }
      for (HostJob* j : to_admit) {
        int rc = eng.admit(j);
        std::lock_guard<std::mutex> lk(mu);
        if (rc) {
          j->state = BGPU_JOB_FAILED;
          j->t_done = now_ms();
          set_error(rc, bgpu_last_error());
          eng.retire(j);
          done_cv.notify_all();
        } else {
          running.push_back(j);
        }
      }
      if (running.empty()) continue;
      double t0 = now_ms();
      if (eng.run(4) || eng.read_counters()) {
        std::lock_guard<std::mutex> lk(mu);
        for (HostJob* j : running) {
          j->state = BGPU_JOB_FAILED;
          j->t_done = now_ms();
        }
        set_error(BGPU_E_HIP, bgpu_last_error());
        running.clear();
        done_cv.notify_all();
        continue;
      }
      double t1 = now_ms();
      PtCounters cur = *eng.h_cnt;
      {
        std::lock_guard<std::mutex> lk(mu);
        eng.release_quarantine();
      }
      if (!have_prev) {
        prev = PtCounters{};
        have_prev = true;
      }
      uint64_t samples_before = 0, samples_after = 0;
      std::vector<HostJob*> finished;
      {
        std::lock_guard<std::mutex> lk(mu);
        for (HostJob* j : running) {
          if (j->cancel_req) eng.cancel(j);
          unsigned long long before = j->h_skipped;
          for (unsigned v : j->h_pass) before += v;
          bool complete = false;
          eng.progress(j, &complete);
          unsigned long long after = j->h_skipped;
          for (unsigned v : j->h_pass) after += v;
          samples_before += before;
          samples_after += after;
          if (complete) finished.push_back(j);
        }
      }
      accumulate_stats(ctx, prev, cur, t1 - t0, samples_after - samples_before);
      prev = cur;
      if (!finished.empty()) {
        std::lock_guard<std::mutex> lk(mu);
        for (HostJob* j : finished) {
          j->state = j->cancel_req ? BGPU_JOB_CANCELLED : BGPU_JOB_DONE;
          j->t_done = now_ms();
          eng.retire(j);
          running.erase(std::find(running.begin(), running.end(), j));
          {
            std::lock_guard<std::mutex> lk2(ctx->stats_mu);
            ctx->stats.jobs_done++;
          }
        }
        // device memory of finished jobs stays until the handle is released;
        // released ones are freed here
        for (HostJob* j : finished)
          if (j->released) delete j;
        eng.gc_scenes();
        done_cv.notify_all();
      }
    }
  }
};

void destroy_scheduler(Scheduler* s) {
  if (!s) return;
  s->shutdown();
  delete s;
}

}  // namespace bgpu

struct bgpu_job : bgpu::HostJob {};

using namespace bgpu;

extern "C" {

int bgpu_render(bgpu_scene* scene, const bgpu_job_desc* job, float* wxyz, float* sq) {
  if (!scene || !wxyz) return fail(BGPU_E_INVALID, "bgpu_render: NULL argument");
  if (int rc = validate_job(job)) return rc;
  Ctx* c = scene->ctx;
  BGPU_HIP(hipSetDevice(c->device));
  HostJob j;
  if (int rc = j.prepare(scene, *job)) return rc;
  int pool = (int)std::min<unsigned long long>((unsigned long long)c->pt_cfg.pool, j.total);
  Engine eng;
  if (int rc = eng.init(c, pool, c->pt_cfg.block)) return rc;
  if ((job->flags & BGPU_JOB_VARIANCE) == 0 && sq) j.desc.flags |= BGPU_JOB_VARIANCE;
  double t0 = now_ms();
  if (int rc = eng.admit(&j)) return rc;
  PtCounters before{};
  // every path ends within max_depth + 1 bounces, so the work drains in at
  // most (total / P + 1) * (max_depth + 2) iterations
  unsigned long long cap = (j.total / (unsigned long long)eng.P + 2) * (unsigned long long)(job->max_depth + 2) + 64;
  unsigned long long iters = 0;
  while (true) {
    if (int rc = eng.run(4)) return rc;
    if (int rc = eng.read_counters()) return rc;
    iters += 4;
    const PtCounters& cn = *eng.h_cnt;
    if (cn.work_next >= cn.work_total && cn.n_act == 0) break;
    if (iters > cap) return fail(BGPU_E_STATE, "render did not drain within the iteration bound");
  }
  bool complete = false;
  if (int rc = eng.progress(&j, &complete)) return rc;
  if (!complete) return fail(BGPU_E_STATE, "render finished with samples unaccounted for");
  double t1 = now_ms();
  accumulate_stats(c, before, *eng.h_cnt, t1 - t0, j.total);
  BGPU_HIP(hipMemcpy(wxyz, j.film.ptr, j.px * 16, hipMemcpyDeviceToHost));
  if (sq && j.sq.ptr) BGPU_HIP(hipMemcpy(sq, j.sq.ptr, j.px * 12, hipMemcpyDeviceToHost));
  eng.retire(&j);
  return BGPU_OK;
}

int bgpu_job_submit(bgpu_scene* scene, const bgpu_job_desc* job, bgpu_job** out) {
  if (!scene || !out) return fail(BGPU_E_INVALID, "bgpu_job_submit: NULL argument");
  *out = nullptr;
  if (int rc = validate_job(job)) return rc;
  Ctx* c = scene->ctx;
  bgpu_job* j = new bgpu_job();
  if (int rc = j->prepare(scene, *job)) {
    delete j;
    return rc;
  }
  if (Engine::job_bytes(j) > c->budget) {
    delete j;
    return fail(BGPU_E_BUDGET, "job film does not fit the HBM budget");
  }
  {
    std::lock_guard<std::mutex> lk(c->mu);
    if (!c->sched) {
      BGPU_HIP(hipSetDevice(c->device));
      Scheduler* s = new Scheduler();
      int rc = s->start(c);
      if (rc) {
        delete s;
        delete j;
        return rc;
      }
      c->sched = s;
    }
  }
  j->t_submit = now_ms();
  j->state = BGPU_JOB_QUEUED;
  Scheduler* s = c->sched;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->pending.push_back(j);
  }
  s->cv.notify_all();
  *out = j;
  return BGPU_OK;
}

int bgpu_job_poll(bgpu_job* job, float* wxyz, float* sq, int32_t* passes_done, int32_t* state) {
  if (!job) return fail(BGPU_E_INVALID, "bgpu_job_poll: NULL job");
  Scheduler* s = job->scene->ctx->sched;
  std::lock_guard<std::mutex> lk(s->mu);
  if (passes_done) *passes_done = job->passes_done;
  if (state) *state = job->state;
  if (wxyz || sq) {
    const size_t px = job->px;
    if (job->film.ptr && wxyz) BGPU_HIP(hipMemcpy(wxyz, job->film.ptr, px * 16, hipMemcpyDeviceToHost));
    else if (wxyz) std::memset(wxyz, 0, px * 16);
    if (sq && job->sq.ptr) BGPU_HIP(hipMemcpy(sq, job->sq.ptr, px * 12, hipMemcpyDeviceToHost));
    else if (sq) std::memset(sq, 0, px * 12);
  }
  return BGPU_OK;
}

int bgpu_job_wait(bgpu_job* job, double timeout_s) {
  if (!job) return fail(BGPU_E_INVALID, "bgpu_job_wait: NULL job");
  Scheduler* s = job->scene->ctx->sched;
  std::unique_lock<std::mutex> lk(s->mu);
  auto pred = [&] { return job->state >= BGPU_JOB_DONE; };
  if (timeout_s < 0) s->done_cv.wait(lk, pred);
  else if (!s->done_cv.wait_for(lk, std::chrono::duration<double>(timeout_s), pred))
    return fail(BGPU_E_TIMEOUT, "bgpu_job_wait: timed out");
  if (job->state == BGPU_JOB_FAILED) return fail(BGPU_E_HIP, "job failed: " + s->error_msg);
  if (job->state == BGPU_JOB_CANCELLED) return fail(BGPU_E_CANCELLED, "job was cancelled");
  return BGPU_OK;
}

int bgpu_job_cancel(bgpu_job* job) {
  if (!job) return fail(BGPU_E_INVALID, "bgpu_job_cancel: NULL job");
  Scheduler* s = job->scene->ctx->sched;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    if (job->state >= BGPU_JOB_DONE) return BGPU_OK;
    job->cancel_req = true;
  }
  s->cv.notify_all();
  return BGPU_OK;
}

int bgpu_job_release(bgpu_job* job) {
  if (!job) return fail(BGPU_E_INVALID, "bgpu_job_release: NULL job");
  Scheduler* s = job->scene->ctx->sched;
  std::lock_guard<std::mutex> lk(s->mu);
  if (job->state >= BGPU_JOB_DONE) {
    delete job;
  } else {
    job->released = true;
    job->cancel_req = true;
    s->cv.notify_all();
  }
  return BGPU_OK;
}

int bgpu_job_latency_ms(bgpu_job* job, double* ms) {
  if (!job || !ms) return fail(BGPU_E_INVALID, "bgpu_job_latency_ms: NULL argument");
  Scheduler* s = job->scene->ctx->sched;
  std::lock_guard<std::mutex> lk(s->mu);
  if (job->state < BGPU_JOB_DONE) return fail(BGPU_E_STATE, "job has not finished");
  *ms = job->t_done - job->t_submit;
  return BGPU_OK;
}

}  // extern "C"
