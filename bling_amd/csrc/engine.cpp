// Render engine: a pool of P path slots, a job table and one captured
// bounce iteration, replayed until the admitted work is done.
//
//   bgpu_render      the plain path: one job on a private engine, blocking
//                    (prender, Rendering.hs:111-140, all passes).
//   bgpu_job_*       the served path: a scheduler thread owns one engine;
//                    jobs arriving while others run are admitted as job-table
//                    slots and HBM (film buffers) free up, and their samples
//                    share the path pool with everything in flight
//                    (continuous batching).  Admitted jobs hand their
//                    samples to the device in chunks, earliest estimated
//                    finish first, so short jobs overtake long ones.
//                    Progress is polled per pass (PassDone,
//                    Rendering.hs:60-75).
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
  unsigned long long fed = 0;  // samples handed to the device work space
  unsigned long long seq = 0;  // arrival order
  DevBuf film, sq;
  std::vector<unsigned int> h_pass;  // host copy of per-pass counts
  unsigned long long h_skipped = 0;
  uint32_t prog_off = 0;  // the job's range in the engine's progress arena
  bool has_prog = false;
  bool cancel_sent = false;
  unsigned long long samples_seen = 0;  // finished + skipped at the last poll
  int passes_seen = 0;                  // complete passes at the last poll
  int state = BGPU_JOB_QUEUED;
  int passes_done = 0;
  double t_submit = 0.0, t_done = 0.0;
  bool cancel_req = false;
  bool released = false;  // user released the handle
  size_t bytes = 0;
  Ctx* counted = nullptr;  // served jobs count themselves in Ctx::live_jobs

  // Copies a validated description; the camera and film default to the
  // scene's (the caller's descriptors are not referenced afterwards).
  int prepare(Scene* s, const bgpu_job_desc& d) {
    if (int rc = make_view(d.camera ? *d.camera : s->camera, d.film ? *d.film : s->film, &view)) return rc;
    desc = d;
    desc.camera = nullptr;
    desc.film = nullptr;
    px = (size_t)view.film_w * view.film_h;
    per_pass = (unsigned long long)view.ext_w * view.ext_h * (unsigned long long)(d.spp_nu * d.spp_nv);
    // passes are counted by 32-bit device counters
    if (per_pass > 0xffffffffull) return fail(BGPU_E_INVALID, "job has more than 2^32-1 samples per pass");
    if (per_pass > (1ull << 62) / (unsigned long long)d.passes)
      return fail(BGPU_E_INVALID, "job has too many samples");
    total = per_pass * (unsigned long long)d.passes;
    scene = s;
    s->refs++;
    return BGPU_OK;
  }
  virtual ~HostJob() {
    if (counted) counted->live_jobs--;
    if (bytes && scene) scene->ctx->film_bytes -= bytes;
    film.release();
    sq.release();
    if (scene && --scene->refs == 0) {
      (void)hipSetDevice(scene->ctx->device);
      delete scene;
    }
  }
};

// Per-pass sample counters of every job in an engine live in one device
// arena, so that the progress of all jobs in flight is read with one copy:
// words [0, kSkipWords) hold each slot's 64-bit count of samples skipped after
// cancellation, the rest is handed out in per-job ranges of `passes` words.
constexpr uint32_t kProgWords = 1u << 20;
constexpr uint32_t kSkipWords = 2u * kMaxJobs;
constexpr int kMaxPasses = 1 << 16;

struct Engine {
  Ctx* ctx = nullptr;
  hipStream_t stream = nullptr;
  int P = 0;
  PtConfig cfg;
  int width = 0;  // BVH width the captured traversal kernels assume (0: any)
  DevBuf pool_buf, jobs_buf, ring_buf, scenes_buf, cnt_buf, admit_buf, prog_buf;
  DevBuf pmat_buf;  // per-slot BxDF records of node-textured hits (allocated on first need)
  PtKernelArgs args{};
  AdmitReq* h_admit = nullptr;
  PtCounters* h_cnt = nullptr;
  DevJob* h_job = nullptr;
  DevPtScene* h_scene = nullptr;
  uint32_t* h_prog = nullptr;  // pinned copy of the progress arena
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;
  std::vector<Scene*> scenes;             // table index -> scene
  std::vector<HostJob*> slots;            // job slot -> job
  std::vector<int> free_slots;
  std::vector<std::pair<uint32_t, uint32_t>> arena_used;  // sorted (offset, words)
  // Retired slots wait until a bounce iteration has run after retirement:
  // its reset step drops the finished job from the device's job ring, after
  // which the slot's table entry may be overwritten.
  std::vector<std::pair<int, unsigned int>> quarantine;  // (slot, iteration at retirement)
  unsigned long long n_admitted = 0;
  // device time of the iterations: events around each run() (read by poll)
  hipEvent_t ev_run0 = nullptr, ev_run1 = nullptr;
  bool run_timed = false;
  double dev_ms = 0.0;  // total, monotonic
  hipEvent_t marks[PH_COUNT + 1] = {};  // phase boundaries while profiling

  ~Engine() {
    if (stream) (void)hipStreamSynchronize(stream);
    if (ev_run0) (void)hipEventDestroy(ev_run0);
    if (ev_run1) (void)hipEventDestroy(ev_run1);
    for (hipEvent_t& e : marks)
      if (e) (void)hipEventDestroy(e);
    if (gexec) (void)hipGraphExecDestroy(gexec);
    if (graph) (void)hipGraphDestroy(graph);
    if (h_admit) (void)hipHostFree(h_admit);
    if (h_cnt) (void)hipHostFree(h_cnt);
    if (h_job) (void)hipHostFree(h_job);
    if (h_scene) (void)hipHostFree(h_scene);
    if (h_prog) (void)hipHostFree(h_prog);
    for (Scene* s : scenes)
      if (s && --s->refs == 0) delete s;
    pool_buf.release();
    jobs_buf.release();
    ring_buf.release();
    scenes_buf.release();
    cnt_buf.release();
    admit_buf.release();
    prog_buf.release();
    pmat_buf.release();
    if (stream) (void)hipStreamDestroy(stream);
  }

  int init(Ctx* c, const PtConfig& config) {
    ctx = c;
    cfg = config;
    P = std::max(1024, (cfg.pool + 1023) / 1024 * 1024);
    BGPU_HIP(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
    BGPU_HIP(hipEventCreate(&ev_run0));
    BGPU_HIP(hipEventCreate(&ev_run1));
    if (cfg.profile)
      for (hipEvent_t& e : marks) BGPU_HIP(hipEventCreate(&e));
    // pool layout: 9 int arrays + 53 float arrays of P, plus 5 queues, plus
    // the direct-lighting branch stacks (1 + kDlWords * dl_depth arrays)
    size_t per = (size_t)P * 4;
    size_t total = per * (9 + 53 + 5 + 1 + (size_t)kDlWords * cfg.dl_depth);
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
    pp.occ = (int32_t*)take(1);
    pp.sxy = (float*)take(2);
    pp.ray = (float*)take(7);
    pp.hit = (float*)take(3);
    pp.thr = (float*)take(16);
    pp.rad = (float*)take(3);
    pp.sray = (float*)take(8);
    pp.pend_s = (float*)take(3);
    pp.mray = (float*)take(8);
    pp.pend_m = (float*)take(3);
    pp.q_act = (uint32_t*)take(1);
    pp.q_shadow = (uint32_t*)take(1);
    pp.q_mis = (uint32_t*)take(1);
    pp.q_sh = (uint32_t*)take(1);
    pp.q_cls = (uint32_t*)take(1);
    pp.dl_sp = (int32_t*)take(1);
    pp.dl_stk = (float*)take((size_t)kDlWords * cfg.dl_depth);
    pp.dl_cap = cfg.dl_depth;
    if (int rc = jobs_buf.alloc(c, sizeof(DevJob) * kMaxJobs)) return rc;
    if (int rc = ring_buf.alloc(c, sizeof(RingChunk) * kRing)) return rc;
    if (int rc = scenes_buf.alloc(c, sizeof(DevPtScene) * kMaxScenes)) return rc;
    if (int rc = cnt_buf.alloc(c, sizeof(PtCounters))) return rc;
    if (int rc = admit_buf.alloc(c, sizeof(AdmitReq) * kMaxJobs)) return rc;
    if (int rc = prog_buf.alloc(c, sizeof(uint32_t) * kProgWords)) return rc;
    args.scenes = scenes_buf.as<DevPtScene>();
    args.jobs = jobs_buf.as<DevJob>();
    args.ring = ring_buf.as<RingChunk>();
    args.cnt = cnt_buf.as<PtCounters>();
    args.admit = admit_buf.as<AdmitReq>();
    BGPU_HIP(hipHostMalloc((void**)&h_admit, sizeof(AdmitReq) * kMaxJobs, hipHostMallocDefault));
    BGPU_HIP(hipHostMalloc((void**)&h_cnt, sizeof(PtCounters), hipHostMallocDefault));
    BGPU_HIP(hipHostMalloc((void**)&h_job, sizeof(DevJob), hipHostMallocDefault));
    BGPU_HIP(hipHostMalloc((void**)&h_scene, sizeof(DevPtScene), hipHostMallocDefault));
    BGPU_HIP(hipHostMalloc((void**)&h_prog, sizeof(uint32_t) * kProgWords, hipHostMallocDefault));
    std::memset(h_cnt, 0, sizeof(PtCounters));
    slots.assign(kMaxJobs, nullptr);
    for (int i = kMaxJobs - 1; i >= 0; --i) free_slots.push_back(i);
    BGPU_HIP(hipMemsetAsync(prog_buf.ptr, 0, sizeof(uint32_t) * kSkipWords, stream));
    BGPU_HIP(pt_enqueue_init(args, stream));
    if (int rc = capture(c->bvh_width, cfg.sort == 1, false, cfg.sort == 1 ? SHADE_MODE_SPLIT : SHADE_MODE_ALL))
      return rc;
    BGPU_HIP(hipStreamSynchronize(stream));
    return BGPU_OK;
  }

  // Captures one bounce iteration whose traversal kernels are specialised
  // for BVH width w (0: any width), shading in class-grouped order if `sort`
  // with the variants of `shade_mode` (SHADE_MODE_SPLIT needs `sort`), with
  // the texture kernel if `textured` (pmat_buf allocated).
  int capture(int w, bool sort, bool direct, int shade_mode, bool textured = false, int light_mode = LIGHTS_ALL) {
    args.sort_shade = sort ? 1 : 0;
    args.direct = direct ? 1 : 0;
    args.shade_mode = shade_mode;
    args.textured = textured ? 1 : 0;
    args.light_mode = light_mode;
    args.pool.pmat = pmat_buf.ptr ? pmat_buf.as<DMaterial>() : nullptr;
    if (gexec) BGPU_HIP(hipGraphExecDestroy(gexec));
    if (graph) BGPU_HIP(hipGraphDestroy(graph));
    gexec = nullptr;
    graph = nullptr;
    BGPU_HIP(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
    hipError_t e = pt_enqueue_iteration(args, cfg, w, stream);
    hipError_t e2 = hipStreamEndCapture(stream, &graph);
    if (e != hipSuccess) return hip_fail(e, "pt_enqueue_iteration (capture)");
    if (e2 != hipSuccess) return hip_fail(e2, "hipStreamEndCapture");
    BGPU_HIP(hipGraphInstantiate(&gexec, graph, nullptr, nullptr, 0));
    width = w;
    return BGPU_OK;
  }

  // First fit in the progress arena; false when no gap of `n` words is left.
  bool arena_alloc(uint32_t n, uint32_t* off) {
    uint32_t cur = kSkipWords;
    auto it = arena_used.begin();
    for (; it != arena_used.end(); ++it) {
      if (it->first - cur >= n) break;
      cur = it->first + it->second;
    }
    if (it == arena_used.end() && kProgWords - cur < n) return false;
    arena_used.insert(it, {cur, n});
    *off = cur;
    return true;
  }
  bool arena_fits(uint32_t n) const {
    uint32_t cur = kSkipWords;
    for (const auto& u : arena_used) {
      if (u.first - cur >= n) return true;
      cur = u.first + u.second;
    }
    return kProgWords - cur >= n;
  }
  void arena_free(uint32_t off) {
    for (auto it = arena_used.begin(); it != arena_used.end(); ++it)
      if (it->first == off) {
        arena_used.erase(it);
        return;
      }
  }
  uint32_t arena_end() const { return arena_used.empty() ? kSkipWords : arena_used.back().first + arena_used.back().second; }

  bool can_admit(const HostJob* j) const { return !free_slots.empty() && arena_fits((uint32_t)j->desc.passes); }

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
    size_t b = j->px * 16;
    if (j->desc.flags & BGPU_JOB_VARIANCE) b += j->px * 12;
    return b;
  }

  // Allocates the job's device state and queues its admission.  On failure
  // the caller retires the job, which returns what was taken.
  int admit(HostJob* j) {
    if (!can_admit(j)) return fail(BGPU_E_BUDGET, "job table or progress counters full");
    // a scene of another width than the captured kernels assume: fall back
    // to kernels that take any width (the stream is idle between iterations)
    const int want_w = (width != 0 && j->scene->dev.width != width) ? 0 : width;
    // shading-queue grouping (pt.sort 2): on while a resident scene mixes
    // BxDF makeups (the table holds only scenes with jobs in flight)
    bool want_sort = cfg.sort == 1;
    if (cfg.sort == 2) {
      want_sort = j->scene->pt.shade_kinds > 1;
      for (Scene* s : scenes)
        if (s && s->pt.shade_kinds > 1) want_sort = true;
    }
    // shading variants: the diffuse one alone while every resident scene is
    // diffuse only; split by class (diffuse, then the rest) on a grouped queue
    bool all_diffuse = j->scene->pt.diffuse_only;
    for (Scene* s : scenes)
      if (s && !s->pt.diffuse_only) all_diffuse = false;
    const int want_mode = all_diffuse ? SHADE_MODE_DIFFUSE : (want_sort ? SHADE_MODE_SPLIT : SHADE_MODE_ALL);
    // the diffuse variant's light kinds: those of every resident scene
    int kinds = j->scene->pt.light_kinds;
    for (Scene* s : scenes)
      if (s) kinds |= s->pt.light_kinds;
    // (a sun/sky map needs the kernels that evaluate it, whatever the knob)
    const int want_lights = (kinds >> kLightSky & 1)      ? LIGHTS_SKY
                            : cfg.light_variants == 0     ? LIGHTS_ALL
                            : (kinds & ~kQuadLights) == 0 ? LIGHTS_QUAD
                            : (kinds & ~kQSILights) == 0  ? LIGHTS_QSI
                                                          : LIGHTS_ALL;
    // the direct-lighting shading kernel runs while such a job is resident
    bool want_direct = j->desc.integrator == BGPU_INTEGRATOR_DIRECT;
    for (HostJob* o : slots)
      if (o && o->desc.integrator == BGPU_INTEGRATOR_DIRECT) want_direct = true;
    // the texture kernel runs while a resident scene has node textures; its
    // per-slot records are allocated on the first such job
    bool want_tex = j->scene->pt.node_tex;
    for (Scene* s : scenes)
      if (s && s->pt.node_tex) want_tex = true;
    if (want_tex && !pmat_buf.ptr) {
      BGPU_HIP(hipStreamSynchronize(stream));
      if (int rc = pmat_buf.alloc(ctx, sizeof(DMaterial) * (size_t)P)) return rc;
    }
    if (want_w != width || want_sort != (args.sort_shade != 0) || want_direct != (args.direct != 0) ||
        want_mode != args.shade_mode || want_tex != (args.textured != 0) || want_lights != args.light_mode) {
      BGPU_HIP(hipStreamSynchronize(stream));
      if (int rc = capture(want_w, want_sort, want_direct, want_mode, want_tex, want_lights)) return rc;
    }
    int sidx = scene_index(j->scene);
    if (sidx < 0) return sidx;
    const bgpu_job_desc& d = j->desc;
    const size_t px = j->px;
    if (int rc = j->film.alloc(ctx, px * 16)) return rc;
    if (d.flags & BGPU_JOB_VARIANCE)
      if (int rc = j->sq.alloc(ctx, px * 12)) return rc;
    j->bytes = job_bytes(j);
    ctx->film_bytes += j->bytes;
    int slot = free_slots.back();
    free_slots.pop_back();
    j->slot = slot;
    slots[slot] = j;
    uint32_t off = 0;
    (void)arena_alloc((uint32_t)d.passes, &off);  // can_admit checked the space
    j->prog_off = off;
    j->has_prog = true;
    uint32_t* prog = prog_buf.as<uint32_t>();
    BGPU_HIP(hipMemsetAsync(j->film.ptr, 0, px * 16, stream));
    if (j->sq.ptr) BGPU_HIP(hipMemsetAsync(j->sq.ptr, 0, px * 12, stream));
    BGPU_HIP(hipMemsetAsync(prog + off, 0, (size_t)d.passes * 4, stream));
    BGPU_HIP(hipMemsetAsync(prog + 2 * slot, 0, 8, stream));
    j->h_pass.assign(d.passes, 0);
    j->h_skipped = 0;
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
    J.integ = d.integrator;
    J.wxyz = j->film.as<float>();
    J.sq = j->sq.ptr ? j->sq.as<float>() : nullptr;
    J.pass_done = prog + off;
    J.skipped = reinterpret_cast<unsigned long long*>(prog + 2 * slot);
    J.cancelled = 0;
    J.view = j->view;
    BGPU_HIP(hipMemcpyAsync(args.jobs + slot, h_job, sizeof(DevJob), hipMemcpyHostToDevice, stream));
    BGPU_HIP(hipStreamSynchronize(stream));  // pinned staging is reused
    j->fed = 0;
    j->seq = ++n_admitted;
    j->state = BGPU_JOB_RUNNING;
    return BGPU_OK;
  }

  // Appends chunks of samples to the device work space with one admit
  // launch: reqs[k] = (job, samples).  Chunks of a job are handed over in
  // sample order.
  int feed(const std::vector<std::pair<HostJob*, unsigned long long>>& reqs) {
    if (reqs.empty()) return BGPU_OK;
    if (reqs.size() > (size_t)kMaxJobs) return fail(BGPU_E_STATE, "too many chunks in one admission");
    for (size_t k = 0; k < reqs.size(); ++k) {
      HostJob* j = reqs[k].first;
      h_admit[k] = AdmitReq{j->slot, 1, reqs[k].second, j->fed};
      j->fed += reqs[k].second;
    }
    BGPU_HIP(hipMemcpyAsync(args.admit, h_admit, sizeof(AdmitReq) * reqs.size(), hipMemcpyHostToDevice, stream));
    BGPU_HIP(pt_enqueue_admit(args, (int)reqs.size(), stream));
    BGPU_HIP(hipStreamSynchronize(stream));  // pinned staging is reused
    return BGPU_OK;
  }

  // Earliest estimated finish first: while fewer than `ahead` samples wait
  // in the device work space, hands chunks of at most `chunk` samples to the
  // job whose arrival time plus remaining work at `per_ms` samples per
  // millisecond is smallest (ties: arrival order).  Short jobs overtake
  // long ones that arrived less than their difference in run time earlier,
  // and no job waits behind work that arrived after its own estimated
  // finish, so long jobs are not starved.  The counters in h_cnt must be
  // fresh (read after the last iteration, no admission since).
  int feed_next(const std::vector<HostJob*>& js, unsigned long long ahead, unsigned long long chunk,
                double per_ms) {
    const PtCounters& c = *h_cnt;
    unsigned long long backlog = c.work_total > c.work_next ? c.work_total - c.work_next : 0;
    unsigned int ring_free = (unsigned)kRing - (c.ring_tail - c.ring_head);
    std::vector<std::pair<HostJob*, unsigned long long>> reqs;
    std::vector<HostJob*> open;
    for (HostJob* j : js)
      if (j->slot >= 0 && !j->cancel_sent && j->fed < j->total) open.push_back(j);
    const double inv = 1.0 / std::max(per_ms, 1.0);
    auto key = [inv](const HostJob* j) { return j->t_submit + (double)(j->total - j->fed) * inv; };
    std::sort(open.begin(), open.end(), [&](const HostJob* a, const HostJob* b) {
      const double ka = key(a), kb = key(b);
      return ka != kb ? ka < kb : a->seq < b->seq;
    });
    // chunks of one job are merged into one request per launch
    std::vector<unsigned long long> take(open.size(), 0);
    size_t k = 0;
    while (k < open.size() && backlog < ahead && reqs.size() + 1 < ring_free && reqs.size() < (size_t)kMaxJobs) {
      HostJob* j = open[k];
      const unsigned long long left = j->total - j->fed - take[k];
      const unsigned long long n = std::min(left, chunk);
      take[k] += n;
      backlog += n;
      if (take[k] == j->total - j->fed) {
        reqs.emplace_back(j, take[k]);
        ++k;
      }
    }
    if (k < open.size() && take[k] > 0) reqs.emplace_back(open[k], take[k]);
    return feed(reqs);
  }

  int cancel(HostJob* j) {
    if (j->slot < 0 || j->cancel_sent) return BGPU_OK;
    h_admit[0].slot = 1;  // staging: the flag value
    BGPU_HIP(hipMemcpyAsync(&args.jobs[j->slot].cancelled, &h_admit[0].slot, sizeof(int32_t),
                            hipMemcpyHostToDevice, stream));
    BGPU_HIP(hipStreamSynchronize(stream));
    j->cancel_sent = true;
    return BGPU_OK;
  }

  // Replays `iters` bounce iterations; their device time is read by poll.
  int run(int iters) {
    if (cfg.profile) return run_profiled(iters);
    BGPU_HIP(hipEventRecord(ev_run0, stream));
    for (int k = 0; k < iters; ++k) BGPU_HIP(hipGraphLaunch(gexec, stream));
    BGPU_HIP(hipEventRecord(ev_run1, stream));
    run_timed = true;
    return BGPU_OK;
  }

  // Profiling (pt.profile 1): the same kernels launched one iteration at a
  // time without the graph, with events between the phases; each
  // iteration's phase times and queue sizes go to the context's profile.
  int run_profiled(int iters) {
    for (int k = 0; k < iters; ++k) {
      BGPU_HIP(pt_enqueue_iteration(args, cfg, width, stream, marks));
      BGPU_HIP(hipMemcpyAsync(h_cnt, args.cnt, sizeof(PtCounters), hipMemcpyDeviceToHost, stream));
      BGPU_HIP(hipStreamSynchronize(stream));
      float ms[PH_COUNT];
      float total = 0.0f;
      for (int ph = 0; ph < PH_COUNT; ++ph) {
        BGPU_HIP(hipEventElapsedTime(&ms[ph], marks[ph], marks[ph + 1]));
        total += ms[ph];
      }
      const uint64_t act = h_cnt->n_act;
      const uint64_t items[PH_COUNT] = {act, act, act, act, h_cnt->n_shadow, h_cnt->n_mis, act};
      std::lock_guard<std::mutex> lk(ctx->stats_mu);
      for (int ph = 0; ph < PH_COUNT; ++ph) {
        ctx->prof.ms[ph] += ms[ph];
        ctx->prof.items[ph] += items[ph];
      }
      ctx->prof.iterations++;
      dev_ms += total;
    }
    return BGPU_OK;
  }

  // Reads the engine counters and the progress of `js` with one copy each
  // and one synchronisation; `done` receives the jobs whose every sample is
  // accounted for (finished, or skipped after cancellation).
  int poll(const std::vector<HostJob*>& js, std::vector<HostJob*>* done) {
    BGPU_HIP(hipMemcpyAsync(h_cnt, args.cnt, sizeof(PtCounters), hipMemcpyDeviceToHost, stream));
    if (!js.empty())
      BGPU_HIP(hipMemcpyAsync(h_prog, prog_buf.ptr, (size_t)arena_end() * 4, hipMemcpyDeviceToHost, stream));
    BGPU_HIP(hipStreamSynchronize(stream));
    if (run_timed) {
      float ms = 0.0f;
      BGPU_HIP(hipEventElapsedTime(&ms, ev_run0, ev_run1));
      dev_ms += ms;
      run_timed = false;
    }
    for (HostJob* j : js) {
      if (j->slot < 0 || !j->has_prog) continue;
      std::memcpy(j->h_pass.data(), h_prog + j->prog_off, (size_t)j->desc.passes * 4);
      std::memcpy(&j->h_skipped, h_prog + 2 * j->slot, 8);
      unsigned long long n = j->h_skipped;
      int pd = 0;
      bool prefix = true;
      for (int p = 0; p < j->desc.passes; ++p) {
        n += j->h_pass[p];
        if (prefix && j->h_pass[p] == j->per_pass) pd++;
        else prefix = false;
      }
      j->passes_seen = pd;
      j->samples_seen = n;
      // a cancelled job ends once what was handed over is accounted for
      if (done && n >= (j->cancel_sent ? j->fed : j->total)) done->push_back(j);
    }
    return BGPU_OK;
  }

  void retire(HostJob* j) {
    if (j->has_prog) {
      arena_free(j->prog_off);
      j->has_prog = false;
    }
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

void destroy_engine(Engine* e) { delete e; }

static int validate_job(const bgpu_job_desc* d, const PtConfig& cfg) {
  if (!d) return fail(BGPU_E_INVALID, "job description is NULL");
  if (d->integrator != BGPU_INTEGRATOR_PATH && d->integrator != BGPU_INTEGRATOR_DIRECT)
    return fail(BGPU_E_INVALID, "integrator must be BGPU_INTEGRATOR_PATH or BGPU_INTEGRATOR_DIRECT");
  if (d->integrator == BGPU_INTEGRATOR_DIRECT && (d->max_depth < 1 || d->max_depth - 1 > cfg.dl_depth))
    return fail(BGPU_E_INVALID, "directLighting maxDepth must be in [1, pt.dl_depth + 1] (pt.dl_depth = " +
                                    std::to_string(cfg.dl_depth) + ")");
  if (d->spp_nu <= 0 || d->spp_nv <= 0)
    return fail(BGPU_E_INVALID, "mkStratifiedSampler: nx and ny must be > 0");
  if ((long long)d->spp_nu * d->spp_nv > (1 << 20)) return fail(BGPU_E_INVALID, "too many samples per pixel");
  if (d->max_depth < 0 || d->max_depth > 1000) return fail(BGPU_E_INVALID, "max_depth must be in [0, 1000]");
  if (d->sample_depth < 0 || d->sample_depth > 1000) return fail(BGPU_E_INVALID, "sample_depth must be in [0, 1000]");
  if (d->passes < 1 || d->passes > kMaxPasses) return fail(BGPU_E_INVALID, "passes must be in [1, 65536]");
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
  s.branches_dropped += after.dl_dropped - before.dl_dropped;
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
    if (int rc = eng.init(c, c->pt_cfg)) return rc;
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
    double per_ms = 1e5;  // samples per millisecond
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
        // admit by HBM budget, in arrival order.  Films of jobs that are
        // running or finished but not yet released count against it; a job
        // that would not fit even once they are all released fails.
        size_t admitting = 0;
        bool blocked = false;
        while (!pending.empty()) {
          HostJob* j = pending.front();
          size_t need = Engine::job_bytes(j);
          const size_t used = ctx->used.load(), films = ctx->film_bytes.load();
          if (used - std::min(used, films) + need > ctx->budget) {
            pending.pop_front();
            j->state = BGPU_JOB_FAILED;
            j->t_done = now_ms();
            set_error(BGPU_E_BUDGET, "job film does not fit the HBM budget beside the resident scenes and engines");
            if (j->released) delete j;
            done_cv.notify_all();
            continue;
          }
          if (used + admitting + need > ctx->budget) {
            blocked = true;
            break;
          }
          if (!eng.can_admit(j) || to_admit.size() >= eng.free_slots.size()) break;
          pending.pop_front();
          to_admit.push_back(j);
          admitting += need;
        }
        // nothing to run and the next job waits for films to be released:
        // sleep until a release (or a new submission) changes that
        if (blocked && running.empty() && to_admit.empty()) {
          cv.wait_for(lk, std::chrono::milliseconds(50));
          continue;
        }
      }
      for (size_t k = 0; k < to_admit.size(); ++k) {
        HostJob* j = to_admit[k];
        if (!eng.can_admit(j)) {  // several jobs taken at once may not all fit
          std::lock_guard<std::mutex> lk(mu);
          pending.insert(pending.begin(), to_admit.begin() + k, to_admit.end());
          break;
        }
        int rc = eng.admit(j);
        std::lock_guard<std::mutex> lk(mu);
        if (rc) {
          j->state = BGPU_JOB_FAILED;
          j->t_done = now_ms();
          set_error(rc, bgpu_last_error());
          eng.retire(j);
          if (j->released) delete j;
          done_cv.notify_all();
        } else {
          running.push_back(j);
        }
      }
      std::vector<HostJob*> run_now;
      {
        std::lock_guard<std::mutex> lk(mu);
        if (running.empty()) continue;
        for (HostJob* j : running)
          if (j->cancel_req) eng.cancel(j);
        run_now = running;
      }
      double t0 = now_ms();
      const double dev0 = eng.dev_ms;
      uint64_t seen_before = 0;
      for (HostJob* j : run_now) seen_before += j->samples_seen;
      std::vector<HostJob*> finished;
      // jobs in `run_now` are deleted only by this thread, and poll and
      // feed write fields no other thread reads; results are published
      // under `mu`.  Four pools of samples stay queued ahead of the device:
      // enough for the iterations between polls, few enough that a short
      // job arriving now is reached within milliseconds.
      int rc = eng.feed_next(run_now, 4ull * (unsigned long long)eng.P, (unsigned long long)eng.P / 2, per_ms);
      if (!rc) rc = eng.run(4);
      if (!rc) rc = eng.poll(run_now, &finished);
      if (rc) {
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
      uint64_t seen_after = 0;
      for (HostJob* j : run_now) seen_after += j->samples_seen;
      accumulate_stats(ctx, prev, cur, eng.dev_ms - dev0, seen_after - seen_before);
      if (t1 > t0 && seen_after > seen_before)  // throughput estimate for feed_next
        per_ms = 0.9 * per_ms + 0.1 * (double)(seen_after - seen_before) / (t1 - t0);
      prev = cur;
      std::lock_guard<std::mutex> lk(mu);
      for (HostJob* j : run_now) j->passes_done = j->passes_seen;
      eng.release_quarantine();
      if (!finished.empty()) {
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
  PtConfig cfg_now;
  {
    std::lock_guard<std::mutex> lk(scene->ctx->mu);
    cfg_now = scene->ctx->pt_cfg;
  }
  if (int rc = validate_job(job, cfg_now)) return rc;
  Ctx* c = scene->ctx;
  BGPU_HIP(hipSetDevice(c->device));
  HostJob j;
  if (int rc = j.prepare(scene, *job)) return rc;
  if ((job->flags & BGPU_JOB_VARIANCE) == 0 && sq) j.desc.flags |= BGPU_JOB_VARIANCE;
  std::lock_guard<std::mutex> lk(c->plain_mu);
  if (!c->plain) {
    Engine* e = new Engine();
    if (int rc = e->init(c, c->pt_cfg)) {
      delete e;
      return rc;
    }
    c->plain = e;
  }
  Engine& eng = *c->plain;
  // a failure below leaves the engine in an unknown state: drop it
  auto bail = [&](int rc) {
    eng.retire(&j);
    destroy_engine(c->plain);
    c->plain = nullptr;
    return rc;
  };
  if (int rc = eng.poll({}, nullptr)) return bail(rc);
  eng.release_quarantine();
  const PtCounters before = *eng.h_cnt;
  const double dev0 = eng.dev_ms;
  if (int rc = eng.admit(&j)) return bail(rc);
  if (int rc = eng.feed({{&j, j.total}})) return bail(rc);
  // every path ends within max_depth + 1 bounces, so the job drains in at
  // most (total / P + 1) * (max_depth + 2) iterations
  // (a direct-lighting sample traces at most 2^max_depth - 1 branches, one
  // bounce each, in its slot)
  const unsigned long long per_path = job->integrator == BGPU_INTEGRATOR_DIRECT
                                          ? (1ull << std::min(job->max_depth, 20)) + 1
                                          : (unsigned long long)(job->max_depth + 2);
  unsigned long long cap = (j.total / (unsigned long long)eng.P + 2) * per_path + 64;
  unsigned long long iters = 0;
  std::vector<HostJob*> one{&j}, done;
  while (done.empty()) {
    if (int rc = eng.run(4)) return bail(rc);
    if (int rc = eng.poll(one, &done)) return bail(rc);
    iters += 4;
    if (iters > cap && done.empty()) return bail(fail(BGPU_E_STATE, "render did not drain within the iteration bound"));
  }
  j.passes_done = j.passes_seen;
  accumulate_stats(c, before, *eng.h_cnt, eng.dev_ms - dev0, j.total);
  {
    std::lock_guard<std::mutex> lk2(c->stats_mu);
    c->stats.jobs_done++;
  }
  hipError_t e1 = hipMemcpy(wxyz, j.film.ptr, j.px * 16, hipMemcpyDeviceToHost);
  hipError_t e2 = (sq && j.sq.ptr) ? hipMemcpy(sq, j.sq.ptr, j.px * 12, hipMemcpyDeviceToHost) : hipSuccess;
  eng.retire(&j);
  eng.gc_scenes();
  if (e1 != hipSuccess) return hip_fail(e1, "hipMemcpy (film)");
  if (e2 != hipSuccess) return hip_fail(e2, "hipMemcpy (sum of squares)");
  return BGPU_OK;
}

int bgpu_job_submit(bgpu_scene* scene, const bgpu_job_desc* job, bgpu_job** out) {
  if (!scene || !out) return fail(BGPU_E_INVALID, "bgpu_job_submit: NULL argument");
  *out = nullptr;
  PtConfig cfg_now;
  {
    std::lock_guard<std::mutex> lk(scene->ctx->mu);
    cfg_now = scene->ctx->pt_cfg;
  }
  if (int rc = validate_job(job, cfg_now)) return rc;
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
  j->t_submit = now_ms();
  j->state = BGPU_JOB_QUEUED;
  j->counted = c;
  c->live_jobs++;
  Scheduler* s = c->sched;
  {
    std::lock_guard<std::mutex> lk2(s->mu);
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
    s->cv.notify_all();  // its film may let a waiting job in
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

int bgpu_config_set(bgpu_ctx* ctx, const char* key, int64_t value) {
  if (!ctx || !key) return fail(BGPU_E_INVALID, "bgpu_config_set: NULL argument");
  const std::string k(key);
  auto block_ok = [](int64_t v) { return v == 64 || v == 128 || v == 256; };
  if (k == "trace.block") {
    if (!block_ok(value)) return fail(BGPU_E_INVALID, "trace.block must be 64, 128 or 256");
    std::lock_guard<std::mutex> lk(ctx->mu);
    ctx->trace_cfg.block = (int)value;
    return BGPU_OK;
  }
  if (k == "trace.persistent") {
    if (value != 0 && value != 1) return fail(BGPU_E_INVALID, "trace.persistent must be 0 or 1");
    std::lock_guard<std::mutex> lk(ctx->mu);
    ctx->trace_cfg.persistent = (int)value;
    return BGPU_OK;
  }
  if (k == "trace.stack") {
    if (value < kMinStackLds || value > kMaxStackLds)
      return fail(BGPU_E_INVALID, "trace.stack must be in [" + std::to_string(kMinStackLds) + ", " +
                                      std::to_string(kMaxStackLds) + "]");
    std::lock_guard<std::mutex> lk(ctx->mu);
    ctx->trace_cfg.stack = (int)value;
    return BGPU_OK;
  }
  if (k == "trace.refill" || k == "trace.walk") {
    const bool refill = k == "trace.refill";
    if (value < (refill ? 1 : 0) || value > 64)
      return fail(BGPU_E_INVALID, refill ? "trace.refill must be in [1, 64]" : "trace.walk must be in [0, 64]");
    std::lock_guard<std::mutex> lk(ctx->mu);
    (refill ? ctx->trace_cfg.refill : ctx->trace_cfg.walk) = (int)value;
    return BGPU_OK;
  }
  if (k == "bvh.width") {
    if (value != 2 && value != 4 && value != 8) return fail(BGPU_E_INVALID, "bvh.width must be 2, 4 or 8");
    std::lock_guard<std::mutex> lk(ctx->mu);
    ctx->bvh_width = (int)value;
    return BGPU_OK;
  }
  if (k == "bvh.device") {
    if (value != 0 && value != 1) return fail(BGPU_E_INVALID, "bvh.device must be 0 or 1");
    std::lock_guard<std::mutex> lk(ctx->mu);
    ctx->bvh_device = (int)value;
    return BGPU_OK;
  }
  // the copy, its validation and the assignment all happen under ctx->mu, so
  // that concurrent settings of different pt.* keys neither race nor get lost
  std::lock_guard<std::mutex> lk(ctx->mu);
  PtConfig next = ctx->pt_cfg;
  if (k == "pt.block") {
    if (!block_ok(value)) return fail(BGPU_E_INVALID, "pt.block must be 64, 128 or 256");
    next.block = (int)value;
  } else if (k == "pt.pool") {
    if (value < 1024 || value > (1 << 26)) return fail(BGPU_E_INVALID, "pt.pool must be in [1024, 2^26]");
    next.pool = (int)value;
  } else if (k == "pt.shade_waves") {
    if (value != 0 && value != 2) return fail(BGPU_E_INVALID, "pt.shade_waves must be 0 or 2");
    next.shade_waves = (int)value;
  } else if (k == "pt.stack") {
    if (value < kMinStackLds || value > kMaxStackLds)
      return fail(BGPU_E_INVALID, "pt.stack must be in [" + std::to_string(kMinStackLds) + ", " +
                                      std::to_string(kMaxStackLds) + "]");
    next.stack = (int)value;
  } else if (k == "pt.walk") {
    if (value < 0 || value > 64) return fail(BGPU_E_INVALID, "pt.walk must be in [0, 64]");
    next.walk = (int)value;
  } else if (k == "pt.diffuse_waves") {
    if (value != 0 && value != 2 && value != 4) return fail(BGPU_E_INVALID, "pt.diffuse_waves must be 0, 2 or 4");
    next.diffuse_waves = (int)value;
  } else if (k == "pt.light_variants") {
    if (value != 0 && value != 1) return fail(BGPU_E_INVALID, "pt.light_variants must be 0 or 1");
    next.light_variants = (int)value;
  } else if (k == "pt.sort") {
    if (value < 0 || value > 2) return fail(BGPU_E_INVALID, "pt.sort must be 0, 1 or 2");
    next.sort = (int)value;
  } else if (k == "pt.persistent") {
    if (value != 0 && value != 1) return fail(BGPU_E_INVALID, "pt.persistent must be 0 or 1");
    next.persistent = (int)value;
  } else if (k == "pt.dl_depth") {
    if (value < 0 || value > 15) return fail(BGPU_E_INVALID, "pt.dl_depth must be in [0, 15]");
    next.dl_depth = (int)value;
  } else if (k == "pt.profile") {
    if (value != 0 && value != 1) return fail(BGPU_E_INVALID, "pt.profile must be 0 or 1");
    next.profile = (int)value;
  } else {
    return fail(BGPU_E_INVALID, "unknown configuration key: " + k);
  }
  // engines capture their kernels and pool at creation: drop the idle ones
  if (ctx->live_jobs.load() > 0) return fail(BGPU_E_STATE, "served job handles are still live");
  if (ctx->sched) {
    destroy_scheduler(ctx->sched);
    ctx->sched = nullptr;
  }
  std::lock_guard<std::mutex> lk3(ctx->plain_mu);
  if (ctx->plain) {
    destroy_engine(ctx->plain);
    ctx->plain = nullptr;
  }
  ctx->pt_cfg = next;
  return BGPU_OK;
}

int bgpu_profile_get(bgpu_ctx* ctx, bgpu_profile* out) {
  if (!ctx || !out) return fail(BGPU_E_INVALID, "bgpu_profile_get: NULL argument");
  std::lock_guard<std::mutex> lk(ctx->stats_mu);
  *out = ctx->prof;
  return BGPU_OK;
}

int bgpu_config_get(bgpu_ctx* ctx, const char* key, int64_t* value) {
  if (!ctx || !key || !value) return fail(BGPU_E_INVALID, "bgpu_config_get: NULL argument");
  const std::string k(key);
  std::lock_guard<std::mutex> lk(ctx->mu);
  if (k == "trace.block") *value = ctx->trace_cfg.block;
  else if (k == "bvh.width") *value = ctx->bvh_width;
  else if (k == "bvh.device") *value = ctx->bvh_device;
  else if (k == "trace.persistent") *value = ctx->trace_cfg.persistent;
  else if (k == "trace.stack") *value = ctx->trace_cfg.stack;
  else if (k == "trace.refill") *value = ctx->trace_cfg.refill;
  else if (k == "trace.walk") *value = ctx->trace_cfg.walk;
  else if (k == "pt.stack") *value = ctx->pt_cfg.stack;
  else if (k == "pt.walk") *value = ctx->pt_cfg.walk;
  else if (k == "pt.block") *value = ctx->pt_cfg.block;
  else if (k == "pt.pool") *value = ctx->pt_cfg.pool;
  else if (k == "pt.shade_waves") *value = ctx->pt_cfg.shade_waves;
  else if (k == "pt.diffuse_waves") *value = ctx->pt_cfg.diffuse_waves;
  else if (k == "pt.light_variants") *value = ctx->pt_cfg.light_variants;
  else if (k == "pt.persistent") *value = ctx->pt_cfg.persistent;
  else if (k == "pt.sort") *value = ctx->pt_cfg.sort;
  else if (k == "pt.dl_depth") *value = ctx->pt_cfg.dl_depth;
  else if (k == "pt.profile") *value = ctx->pt_cfg.profile;
  else return fail(BGPU_E_INVALID, "unknown configuration key: " + k);
  return BGPU_OK;
}

}  // extern "C"
