// Wavefront path tracer: device scene data, job table, path pool and the
// bounce-loop kernels (Integrator/Path.hs:41-87 restated as queues).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <vector>

#include "isect.h"

namespace bgpu {

constexpr int kBands = 16;
constexpr int kMaxJobs = 1024;   // job table slots (reused)
constexpr int kRing = 2048;      // job chunks in the work queue (power of two)
constexpr int kMaxScenes = 64;   // scenes referenced by the engine

// BxDF flags and kinds (Reflection.hs:94-108, Reflection/*.hs)
enum { BX_R = 1, BX_T = 2, BX_D = 4, BX_G = 8, BX_S = 16 };
enum { K_LAMBERT = 0, K_OREN = 1, K_MF = 2, K_SPECR = 3, K_SPECT = 4, K_FB = 5 };
enum { FR_NOOP = 0, FR_DIEL = 1, FR_COND = 2 };

// K_FB (FresnelBlend over the anisotropic distribution) reuses the fields:
// r = rd (diffuse), eta = rs (specular), k = ra (absorption), e = ex,
// ei = ey, et = coating depth.  It is always a material's only component.
struct DBxdf {
  int32_t kind, type, fr;
  int32_t flip;  // brdfToBtdf: the diffuse lobe with wi mirrored to the other hemisphere
  float A, B, e, ei, et, pad2[3];
  float r[16], eta[16], k[16];
};
struct DMaterial {
  int32_t n;
  int32_t tex_kind, tex_map;  // selector texture (bgpu_texture), BGPU_TEX_NONE if constant
  int32_t alt;                // table index of the tex2 variant of this material
  float tex[12];
  DBxdf c[2];
};

enum { LT_AREA = 0, LT_INF = 1, LT_POINT = 2, LT_DIR = 3 };
struct DLight {
  int32_t type, shape, id, w, h, map_off, dist_off;
  int32_t sky;  // infinite: 1 for a sun/sky map (a kSkyFloats block at map_off), 0 for texels
  float L[16];  // area: radiance; point, directional: intensity
  float w2l[16], l2w[16];  // infinite: its transform; area: l2w is the shape's o2w
  float vec[4];  // point: position; directional: the unit direction toward the light
  DShape sh;     // area: a copy of the emitting shape's record, read with the light (one load less)
};

// Everything the bounce kernels need about one scene (device resident).
struct DevPtScene {
  DevScene geo;
  const int32_t* tri_idx;
  const float* verts;
  const float* normals;
  const float* uvs;
  const int2* tri_info;  // (material, mesh flags)
  const DShapeX* shapex;
  const DMaterial* mats;
  const DLight* lights;
  const float* maps;
  const float* dist;
  int32_t ntri, nlights, nmats, pad;
  // node textures (BGPU_TEX_NODES materials): the node list, each node's
  // range start (the lowest node of its subtree), the images as (width,
  // height, channels, offset into texels), and the inputs of each such
  // material (indexed by its DMaterial::alt)
  const bgpu_tex_node* tex_nodes;
  const int32_t* tex_lo;
  const int4* images;
  const float* texels;
  const struct DMatSrc* mat_src;
};

// The inputs of a material whose inputs are textures: its constant inputs,
// which slots are textured (bgpu_texture::slots) and their roots and range
// starts (index 8: the bump displacement).
struct DMatSrc {
  int32_t type, slots;
  int32_t roots[9], lo[9];
  float scalar[4];
  float spec[3][16];
};

// Camera and film of a job (ProjectiveCamera, Camera.hs:24-31; the film's
// 16x16 filter table, Image.hs:48-61; the sample extent, Image.hs:162-168).
struct DevView {
  float c2w[16], r2c[16];
  float lens_radius, focal_distance;
  int32_t cam_type, pad;
  float env_sx, env_sy;  // the environment camera's raster size (Camera.hs:70-76)
  int32_t film_w, film_h, ext_x0, ext_y0, ext_w, ext_h;
  float fw, fh;
  float filter[256];
};

// One job in the device job table.  Its samples (numbered pass-major:
// pass, pixel of the sample extent, sample n) enter the work space in
// chunks (RingChunk).
struct DevJob {
  unsigned long long per_pass;  // extent pixels * nu * nv
  uint64_t seed;
  int32_t scene, nu, nv, N, md, sd, rr, passes;
  float* wxyz;
  float* sq;
  unsigned int* pass_done;      // finished samples per pass
  unsigned long long* skipped;  // samples dropped after cancellation
  int32_t cancelled;
  int32_t integ;                // BGPU_INTEGRATOR_PATH or BGPU_INTEGRATOR_DIRECT
  DevView view;
};

// Device counters of the engine.
// Classes of the shading work a path needs this iteration (shade_class):
// the shade queue is grouped by class so that a wave runs one code path.
constexpr int kShadeClasses = 40;
// classes [0, kDiffuseClasses): escaped, depth cutoff, no BxDF, and the
// makeups of Lambert / Oren-Nayar components only
constexpr int kDiffuseClasses = 9;

struct PtCounters {
  unsigned long long work_next, work_total;
  unsigned int n_act, n_shadow, n_mis, last_act;
  unsigned long long tot_camera, tot_extend, tot_shadow, tot_mis, tot_samples;
  unsigned int ring_head, ring_tail;  // chunks not yet fully taken [head, tail)
  unsigned int iter, pad;
  unsigned long long dl_dropped;  // direct-lighting branches lost to a full branch stack (must stay 0)
  unsigned int cls_count[kShadeClasses];   // active paths per shading class
  unsigned int cls_cursor[kShadeClasses];  // slots of the grouped queue taken per class
};

// A chunk of a job's samples in the monotonic 64-bit work space: work
// indices [wbegin, wend) are the job's samples local, local + 1, ...
struct RingChunk {
  unsigned long long wbegin, wend, local;
  int32_t slot, pad;
};

// Chunk request consumed by the admit kernel: `size` samples of job `slot`
// starting at its sample `local`.
struct AdmitReq {
  int32_t slot, valid;
  unsigned long long size, local;
};

// Path pool, structure of arrays of P slots.
constexpr int kDlWords = 24;  // a pending direct-lighting branch: ray (7), throughput (16), depth

struct PathPool {
  int32_t P;
  int32_t* state;   // 0 free, 1 active, 2 done
  int32_t* job;
  int32_t* pass;
  int32_t* pix;
  int32_t* samp;
  int32_t* dspec;   // depth | spec << 16
  float* sxy;       // 2P: image position of the camera sample
  float* ray;       // 7P: ox oy oz dx dy dz tmin
  float* hit;       // 3P: t b1 b2
  int32_t* hprim;
  float* thr;       // 16P throughput
  float* rad;       // 3P radiance as unnormalised XYZ sums
  float* sray;      // 8P shadow ray
  float* pend_s;    // 3P XYZ of the light sample if the shadow ray is unoccluded
  float* mray;      // 8P MIS ray: ox oy oz dx dy dz tmin tmax
  float* pend_m;    // 3P XYZ of the BSDF sample if the MIS ray reaches the light
  int32_t* mlight;  // light index of the MIS ray
  int32_t* occ;     // shadow ray occluded by a triangle (persistent passes)
  // direct lighting: specular branches waiting to be traced, a stack of
  // dl_cap entries per slot; entry e, word k of slot i at dl_stk[(e * kDlWords + k) * P + i]
  // (ray ox oy oz dx dy dz tmin, 16 throughput bands, depth)
  int32_t* dl_sp;
  float* dl_stk;
  int32_t dl_cap, dl_pad;
  uint32_t* q_act;
  uint32_t* q_shadow;
  uint32_t* q_mis;
  uint32_t* q_sh;   // q_act grouped by shading class (when sorting)
  uint32_t* q_cls;  // shading class of q_act[q]
  // per slot, the BxDF records of a hit on a material whose inputs are
  // textures, evaluated at the hit (allocated while such a scene is resident)
  DMaterial* pmat;
};

struct PtConfig {
  int block = 128;
  int pool = 1 << 20;
  int shade_waves = 2;  // occupancy hint of the shading kernel (0 or 2 waves per SIMD)
  int diffuse_waves = 2;  // the same for its diffuse-only variant (0, 2 or 4)
  int light_variants = 1;  // the diffuse variant's light-kind variants (LIGHTS_*) on (1) or off (0)
  int stack = 12;         // traversal stack entries per lane kept in LDS (kMinStackLds..kMaxStackLds)
  int walk = 8;           // traversal: walk inner nodes while at least this many lanes do (0: all)
  int persistent = 0;   // traversal passes with lane refill (1) or one ray per thread (0)
  int sort = 2;         // shading queue grouped by material kind: 1 always, 0 never, 2 while a
                        // resident scene has more than one BxDF makeup
  int dl_depth = 4;     // pending specular branches per direct-lighting path
  int profile = 0;      // 1: iterations run as plain launches, timed per phase (no graph)
};

// Phases of one bounce iteration, as timed by profiling (bgpu_profile).
enum { PH_GENERATE = 0, PH_EXTEND = 1, PH_CLASSIFY = 2, PH_SHADE = 3, PH_SHADOW = 4, PH_MIS = 5, PH_FINISH = 6,
       PH_COUNT = 7 };

struct PtKernelArgs {
  PathPool pool;
  const DevPtScene* scenes;
  DevJob* jobs;
  RingChunk* ring;  // chunks in work order (kRing entries, modulo)
  PtCounters* cnt;
  AdmitReq* admit;
  int32_t sort_shade;  // shade reads q_sh (grouped by class) instead of q_act
  int32_t direct;      // direct-lighting jobs may be resident: run their shading kernel
  int32_t shade_mode;  // SHADE_MODE_*: which shading variants the iteration runs
  int32_t nshort;      // traversal stack entries per lane in LDS (PtConfig::stack)
  int32_t walk;        // traversal walk threshold (PtConfig::walk)
  int32_t textured;    // a resident scene has node-textured materials: run the texture kernel
  int32_t light_mode;  // LIGHTS_*: the light kinds the diffuse shading variant compiles in
};

// Light kinds of the diffuse shading variant: every kind; or, when every
// resident scene's lights are of these kinds, quads only, or quads, spheres
// and infinite lights (the other shapes' light code and its registers left
// out).
enum { LIGHTS_ALL = 0, LIGHTS_QUAD = 1, LIGHTS_QSI = 2, LIGHTS_SKY = 3 };
// Bits of a light-kind set: bit k < 5 an area light on bgpu shape type k,
// bit kLightInf an infinite light, kLightPoint a point light, kLightDir a
// directional light.
constexpr int kLightInf = 5;
constexpr int kLightPoint = 6;
constexpr int kLightDir = 7;
constexpr int kLightSky = 8;  // an infinite light with a sun/sky map (also sets kLightInf)
constexpr int kQuadLights = 1 << BGPU_SHAPE_QUAD;
constexpr int kQSILights = (1 << BGPU_SHAPE_QUAD) | (1 << BGPU_SHAPE_SPHERE) | (1 << kLightInf);

// Shading variants of an iteration: one kernel for every BxDF kind; the
// diffuse variant alone (every resident scene's materials are Lambert /
// Oren-Nayar); or, on the class-grouped queue, the diffuse variant over the
// diffuse classes and the full kernel over the rest.
enum { SHADE_MODE_ALL = 0, SHADE_MODE_DIFFUSE = 1, SHADE_MODE_SPLIT = 2 };

// Host-side shading data owned by a Scene.
struct PtScene {
  std::vector<int32_t> shape_light_host;  // device light index per shape, -1 if none
  DevPtScene host{};                      // filled by build_pt_data, pointers are device
  int shade_kinds = 0;                    // distinct BxDF makeups of the materials in use
  bool diffuse_only = false;              // every material in use has Lambert / Oren-Nayar components only
  bool node_tex = false;                  // a material's inputs are textures of the node list
  int light_kinds = 0;                    // bit k: an area light on shape type k; bit 5: an infinite light
  std::vector<int32_t> tex_lo_host;       // per texture node, the lowest node of its subtree
  std::vector<int32_t> tex_op_host;       // per texture node, its operation
};

// Fills a job view from a camera and film description.
int make_view(const bgpu_camera& cam, const bgpu_film& film, DevView* v);  // 0 or BGPU_E_INVALID

// Enqueues one bounce iteration (reset, generate, extend, shade, shadow,
// MIS, finish) on `s`.  Capturable into a hipGraph.  `width`: the BVH width
// of every scene the iteration will see (2, 4 or 8), or 0 for any width.
// `marks`: null, or PH_COUNT + 1 events recorded before the first phase and
// after each phase (for profiling; not for capture).
hipError_t pt_enqueue_iteration(const PtKernelArgs& a, const PtConfig& cfg, int width, hipStream_t s,
                                hipEvent_t* marks = nullptr);
// Enqueues the admission of jobs listed in `admit` (count n) on `s`.
hipError_t pt_enqueue_admit(const PtKernelArgs& a, int n, hipStream_t s);
// Evaluates the node texture [lo, root] of scene `sc` at n points (xyz, and
// the surface parameters uv): 16 bands per point for a spectrum texture,
// one value for a scalar one.  Device pointers.
hipError_t pt_texture_eval(const DevPtScene& sc, int lo, int root, bool spectrum, int n, const float* xyz,
                           const float* uv, float* out, hipStream_t s);
hipError_t pt_enqueue_init(const PtKernelArgs& a, hipStream_t s);

}  // namespace bgpu
