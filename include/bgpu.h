/*
 * bgpu.h -- C ABI of the MI355X (gfx950) ray-tracing hot path for bling.
 *
 * The library replaces three things in the Haskell renderer bling:
 *   - the accelerator behind `Primitive{intersect, intersects}`
 *     (src/lib/Graphics/Bling/Primitive.hs:21-27), built by `kdTreePrimitive`
 *     (Primitive/KdTree.hs:236-246) and consumed through `scIntersect` /
 *     `occluded` (Scene.hs:45-51)  -> bgpu_intersect_closest / bgpu_intersect_any;
 *   - the path integrator `mkPathIntegrator md sd` (Integrator/Path.hs:30-87)
 *     driven per camera sample by the sampler renderer `prender`
 *     (Rendering.hs:111-151)  -> bgpu_job_submit / bgpu_job_poll / bgpu_render;
 *   - the film accumulation `addSample` (Image.hs:250-299), returned as the
 *     (w, X, Y, Z) float quadruple per pixel of `Image` (Image.hs:123-129).
 *
 * Every entry point returns 0 (BGPU_OK) or a negative BGPU_E_* code; the text
 * of the last error on the calling thread is returned by bgpu_last_error().
 * All handles are opaque.  Host arrays passed in are copied before the call
 * returns unless BGPU_DEVICE_PTRS is set, in which case they must be device
 * pointers valid on the context's device until the call returns.
 * A context serialises submissions internally and may be used from any thread.
 */
#ifndef BGPU_H
#define BGPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define BGPU_ABI_VERSION 4
#define BGPU_NUM_BANDS 16 /* Spectrum.hs:34-36 */

/* ---- error codes ------------------------------------------------------ */
enum {
  BGPU_OK = 0,
  BGPU_E_INVALID = -1,   /* bad argument or malformed scene (cf. TriangleMesh.hs:46-55) */
  BGPU_E_NOMEM = -2,     /* host or device allocation failed */
  BGPU_E_HIP = -3,       /* HIP runtime error */
  BGPU_E_BUDGET = -4,    /* job or scene does not fit the context's HBM budget */
  BGPU_E_STATE = -5,     /* call not valid in the handle's current state */
  BGPU_E_CANCELLED = -6, /* job was cancelled */
  BGPU_E_NODEV = -7,     /* no usable GPU */
  BGPU_E_TIMEOUT = -8    /* bgpu_job_wait timed out */
};

typedef struct bgpu_ctx bgpu_ctx;
typedef struct bgpu_scene bgpu_scene;
typedef struct bgpu_job bgpu_job;

/* ---- scene description (copied by bgpu_scene_create) ----------------- */

/* Analytic shapes, Shape.hs:22-38.  params:
 *   BOX      pmin.xyz pmax.xyz                 (mkBox orders them, Shape.hs:41-44)
 *   CYLINDER radius zmin zmax phimax[rad]      (Shape.hs:47-56)
 *   DISK     height radius inner phimax[rad]   (Shape.hs:59-68)
 *   QUAD     sx sy                             (spans [-sx,sx]x[-sy,sy], Shape.hs:157-171)
 *   SPHERE   radius                            (Shape.hs:75-78) */
enum {
  BGPU_SHAPE_BOX = 0,
  BGPU_SHAPE_CYLINDER = 1,
  BGPU_SHAPE_DISK = 2,
  BGPU_SHAPE_QUAD = 3,
  BGPU_SHAPE_SPHERE = 4
};

/* One `mkGeom` primitive (Primitive/Geometry.hs:14-37): a shape intersected
 * in object space through w2o, with an AreaLight when emissive. */
typedef struct {
  int32_t type;
  int32_t material;   /* index into materials[] */
  int32_t emissive;   /* nonzero: carries an AreaLight with radiance emission[] */
  int32_t light_id;   /* the AreaLight id used by light equality (Light.hs:47-50) */
  float params[6];
  float o2w[16];      /* object-to-world matrix, row-major 4x4 */
  float w2o[16];      /* its stored inverse (Transform.hs:123-127) */
  float emission[16]; /* spectral radiance when emissive */
} bgpu_shape;

/* Materials, Material.hs:32-127 (inputs constant, or textures: bgpu_texture). */
enum {
  BGPU_MAT_BLACK = 0,   /* blackBodyMaterial: no BxDF (Reflection.hs:337-338) */
  BGPU_MAT_MATTE = 1,   /* spec[0]=kd, scalar[0]=sigma (Lambertian if 0, else Oren-Nayar) */
  BGPU_MAT_PLASTIC = 2, /* spec[0]=kd, spec[1]=ks, scalar[0]=roughness */
  BGPU_MAT_METAL = 3,   /* spec[0]=eta, spec[1]=k, scalar[0]=roughness */
  BGPU_MAT_GLASS = 4,   /* spec[0]=reflect, spec[1]=transmit (both clamped to [0,1]), scalar[0]=ior */
  BGPU_MAT_MIRROR = 5,  /* spec[0]=reflect (clamped to [0,1]) */
  BGPU_MAT_TRANSMATTE = 6, /* translucentMatte (Material.hs:43-53): spec[0]=kr, spec[1]=kt (both clamped
                              to [0,1], kt then scaled by 1 - kr), scalar[0]=sigma; a diffuse reflection and
                              the same lobe turned into transmission (brdfToBtdf, Reflection.hs:188-195) */
  BGPU_MAT_SUBSTRATE = 7,  /* mkSubstrate (Material.hs:111-127): spec[0]=kd, spec[1]=ks, spec[2]=ka (each
                              clamped to [0,1]), scalar[0]=urough, scalar[1]=vrough, scalar[2]=depth; one
                              FresnelBlend BRDF over the anisotropic distribution with exponents
                              1/max(0, urough) and 1/max(0, vrough) (Microfacet.hs:56-104,136-196) */
  BGPU_MAT_SHINYMETAL = 8  /* mkShinyMetal (Material.hs:98-109): spec[0]=kr, spec[1]=ks, scalar[0]=roughness;
                              a Blinn microfacet with frConductor (frApproxEta ks) (frApproxK ks) and a
                              specular reflection with frConductor (frApproxEta kr) (frApproxK kr), both
                              white (Fresnel.hs:72-78) */
};
/* A texture that selects, per hit, between two constant spectra
 * (Texture.hs:191-219): spec[k] of the material is tex1's value and
 * alt[k] tex2's, for every slot k in `slots` (bit k).
 *   CHECKER     checkerBoard (sx sy sz) on the world hit point: tex1 where
 *               floor(x sx) + floor(y sy) + floor(z sz) is even.
 *               params: sx sy sz
 *   GRAPHPAPER  graphPaper lw over a 2d mapping: tex2 on the lines (the
 *               fractional part of either coordinate within lw/2 of an
 *               integer), tex1 elsewhere.  params: lw, then the mapping:
 *               UV      su sv ou ov            (Texture.hs:166-170)
 *               PLANAR  vu.xyz vv.xyz ou ov    (Texture.hs:172-179) */
enum { BGPU_TEX_NONE = 0, BGPU_TEX_CHECKER = 1, BGPU_TEX_GRAPHPAPER = 2, BGPU_TEX_NODES = 3 };
enum { BGPU_MAP_UV = 0, BGPU_MAP_PLANAR = 1, BGPU_MAP_IDENTITY = 2 };
/* kind BGPU_TEX_NODES: any of the material's inputs is a texture of the
 * scene's node list (bgpu_tex_node): bit k of `slots` (k < 3) makes spec[k]
 * the spectrum texture rooted at roots[k], bit 4 + k (k < 4) makes scalar[k]
 * the scalar texture rooted at roots[4 + k]; the others stay constant.  Bit
 * 8 bump-maps the material by the scalar texture rooted at roots[8]
 * (bumpMapped, Reflection.hs:344-378). */
typedef struct {
  int32_t kind;
  int32_t mapping;
  int32_t slots;
  int32_t pad;
  union {
    float params[12];
    int32_t roots[12];
  };
  float alt[2][16];
} bgpu_texture;

/* A texture as a node list (Texture.hs:59-414, the textures the material
 * parser builds, MaterialParser.hs:106-226): every node's children come
 * before it.  Scalar textures:
 *   SCONST     constant q[0]
 *   SIMAGE     readImageScalarMap: grey image `a` (channels 1) over the 2d mapping
 *   CELL       cellNoise: distance a (0 euclidian, 1 squared euclidian,
 *              2 manhattan, 3 chebyshev) over the 3d mapping
 *   FBM        texMap3dToTexture (fbm octaves=a omega=q[0]) over the 3d mapping
 *   PERLIN     noiseTexture over the 3d mapping
 *   CRYSTAL    quasiCrystal octaves=a over the 2d mapping
 *   SCALE      scaleTexture q[0] q[1] (child a): q[0] + q[1] * a
 * Spectrum textures:
 *   CONST      constant spec
 *   IMAGE      readImageTextureMap: image `a` (channels 16) over the 2d mapping
 *   BLEND      spectrumBlend tex1=a tex2=b f=c
 *   CHECKER    checkerBoard scale=q[0..2] tex1=a tex2=b
 *   GRADIENT   gradient (mkGradient steps) f=a: steps are the GSTEP nodes
 *              b .. b+c-1 (position q[0], colour spec), ascending by position
 *   GRAPHPAPER graphPaper lw=q[0] over the 2d mapping, tex1=a tex2=b
 * Mappings (`map`, `m`): 2d UV m = su sv ou ov; 2d PLANAR m = vu.xyz vv.xyz
 * ou ov; 3d IDENTITY m = the world-to-texture matrix, 4x4 row-major. */
enum {
  BGPU_TN_SCONST = 1, BGPU_TN_SIMAGE = 2, BGPU_TN_CELL = 3, BGPU_TN_FBM = 4, BGPU_TN_PERLIN = 5,
  BGPU_TN_CRYSTAL = 6, BGPU_TN_SCALE = 7,
  BGPU_TN_CONST = 16, BGPU_TN_IMAGE = 17, BGPU_TN_BLEND = 18, BGPU_TN_CHECKER = 19, BGPU_TN_GRADIENT = 20,
  BGPU_TN_GRAPHPAPER = 21, BGPU_TN_GSTEP = 22
};
typedef struct {
  int32_t op;
  int32_t a, b, c;
  int32_t map;
  int32_t pad[3];
  float q[4];
  float m[16];
  float spec[16];
} bgpu_tex_node;
/* An image of a texture: width x height texels of `channels` floats (16:
 * pixelSpectrum of each RGB pixel, Texture.hs:84-86; 1: grey / 255),
 * row-major from the top row (JuicyPixels' pixelAt). */
typedef struct {
  int32_t width, height, channels, pad;
  const float* texels;
} bgpu_image;

typedef struct {
  int32_t type;
  int32_t pad;
  float scalar[4];
  float spec[3][16];
  bgpu_texture tex; /* kind BGPU_TEX_NONE: constant spectra; selectors cover spec[0] and spec[1] */
} bgpu_material;

/* A triangle mesh made by mkTriangleMesh (Primitive/TriangleMesh.hs:39-59). */
enum { BGPU_MESH_NORMALS = 1, BGPU_MESH_UVS = 2 };
typedef struct {
  int32_t material;
  int32_t flags;
} bgpu_mesh;

/* Explicit (non-geometric) lights (Light.hs:30-229):
 *   INFINITE     mkInfiniteAreaLight (Light.hs:71-82) with its world-to-light
 *                transform and one of the maps pDiscSpectrumMap2d reads
 *                (IO/MaterialParser.hs:250-272), chosen by `map`:
 *                  BGPU_LIGHTMAP_TEXELS  a spectral map of width x height
 *                    texels (`radiance`), looked up as rgbfToTexMap does
 *                    (IO/Bitmap.hs:21-28): texel (clamp (floor ((1 - u) w)),
 *                    clamp (floor ((1 - v) h))).  A `constant` map is 1 x 1,
 *                    a `file` map the decoded image's rgbToSpectrumIllum.
 *                  BGPU_LIGHTMAP_SUNSKY  mkSunSkyLight east sunDir turbidity
 *                    (SunSky.hs:12-24), evaluated per direction: sky[0..2]
 *                    east, sky[3..5] sunDir, sky[6] turbidity; `radiance`,
 *                    width and height are unused (the map is 640 x 480);
 *   POINT        mkPointLight: `radiance` is the 16-band intensity (width =
 *                height = 1), `vec` the position;
 *   DIRECTIONAL  mkDirectional: `radiance` is the 16-band intensity (width =
 *                height = 1), `vec` the normal as given (the library
 *                normalizes it as mkDirectional does).
 * Point and directional lights are delta lights: sampled with pdf 1 and
 * never hit by rays (Scene.hs:61-82). */
enum { BGPU_LIGHT_INFINITE = 1, BGPU_LIGHT_POINT = 2, BGPU_LIGHT_DIRECTIONAL = 3 };
enum { BGPU_LIGHTMAP_TEXELS = 0, BGPU_LIGHTMAP_SUNSKY = 1 };
typedef struct {
  int32_t type;
  int32_t width, height; /* texSize of a texel map */
  int32_t map;           /* infinite lights: BGPU_LIGHTMAP_* */
  const float* radiance; /* width*height*16 floats, pixel (x,y) at ((y*width)+x)*16 */
  float w2l[16];         /* world-to-light matrix, row-major */
  float l2w[16];         /* its inverse */
  float vec[4];          /* point: position; directional: normal (xyz, w unused) */
  float sky[8];          /* sun/sky map: east xyz, sunDir xyz, turbidity, unused */
} bgpu_light;

enum { BGPU_CAMERA_PERSPECTIVE = 0, BGPU_CAMERA_ENVIRONMENT = 1 };
/* ProjectiveCamera / Environment (Camera.hs:24-33) after mkProjective. */
typedef struct {
  int32_t type;
  int32_t pad;
  float c2w[16];           /* camera-to-world matrix, row-major */
  float r2c[16];           /* raster-to-camera matrix, row-major */
  float lens_radius;
  float focal_distance;
  float xres, yres;        /* environment camera only */
} bgpu_camera;

/* Pixel filters, Filter.hs:12-96. fparams:
 *   BOX -, GAUSS w h alpha, MITCHELL w h b c, SINC w h tau, TRIANGLE w h */
enum {
  BGPU_FILTER_BOX = 0,
  BGPU_FILTER_GAUSS = 1,
  BGPU_FILTER_MITCHELL = 2,
  BGPU_FILTER_SINC = 3,
  BGPU_FILTER_TRIANGLE = 4
};
typedef struct {
  int32_t width, height;
  int32_t filter;
  int32_t pad;
  float fparams[4];
} bgpu_film;

/* The flattened scene.  Primitive ids: triangles are 0..num_triangles-1 in
 * index-buffer order, shapes follow as num_triangles + shape index.  Lights:
 * lights[] (explicit) first, then the emissive shapes in shape order, the
 * order of `mkScene` (Scene.hs:37-43). */
typedef struct {
  uint32_t num_vertices;
  uint32_t num_triangles;
  const float* vertices;   /* 3*num_vertices, world space (transPoint applied) */
  const float* normals;    /* 3*num_vertices world-space shading normals, or NULL */
  const float* uvs;        /* 2*num_vertices, or NULL */
  const int32_t* indices;  /* 3*num_triangles */
  const int32_t* tri_mesh; /* mesh index per triangle, or NULL for mesh 0 */
  uint32_t num_meshes;
  const bgpu_mesh* meshes;
  uint32_t num_shapes;
  const bgpu_shape* shapes;
  uint32_t num_materials;
  const bgpu_material* materials;
  uint32_t num_lights;
  const bgpu_light* lights;
  bgpu_camera camera;
  bgpu_film film;
  uint32_t num_tex_nodes;          /* texture node list of BGPU_TEX_NODES materials */
  const bgpu_tex_node* tex_nodes;
  uint32_t num_images;
  const bgpu_image* images;
} bgpu_scene_desc;

/* ---- ray batches ------------------------------------------------------ */
/* Ray = origin, direction, tmin, tmax (Math.hs:375-380), structure of arrays. */
typedef struct {
  const float *ox, *oy, *oz;
  const float *dx, *dy, *dz;
  const float *tmin, *tmax;
} bgpu_rays;

/* Closest-hit result: t, the triangle barycentrics (b1, b2) of
 * TriangleMesh.hs:175-180 (0 for shapes) and the primitive id, -1 on a miss. */
typedef struct {
  float *t, *b1, *b2;
  int32_t* prim;
} bgpu_hits;

enum {
  BGPU_DEVICE_PTRS = 1, /* ray/hit arrays are device pointers */
  BGPU_TIMED = 2,       /* record the kernel time, read by bgpu_last_kernel_ms */
  BGPU_COUNT = 4,       /* run the counting variant: adds node visits and
                           primitive tests to bgpu_stats (not for timing) */
  BGPU_COHERENT = 8     /* hint: neighbouring rays take similar paths (camera
                           rays); traced one ray per thread instead of with
                           lane refill.  Results are identical either way. */
};

/* ---- jobs --------------------------------------------------------------- */
/* One render job: `renderer { gpu sampled { sampler { stratified nu nv }
 * integrator { path maxDepth md sampleDepth sd } } }` (RendererParser.hs:56-73,
 * IntegratorParser.hs:37-44), `passes` progressive passes of nu*nv samples per
 * pixel over the film's sample extent (Image.hs:162-168, Rendering.hs:111-140).
 * integrator BGPU_INTEGRATOR_DIRECT is `directLighting maxDepth md`
 * (Integrator/DirectLighting.hs): sample_depth and rr_start_depth are unused,
 * and max_depth - 1 must not exceed the "pt.dl_depth" knob (its specular
 * branches wait in a per-path stack of that many entries). */
enum { BGPU_JOB_VARIANCE = 1 /* also accumulate per-pixel sum of squared XYZ */ };
enum { BGPU_INTEGRATOR_PATH = 0, BGPU_INTEGRATOR_DIRECT = 1 };
typedef struct {
  int32_t spp_nu, spp_nv;
  int32_t max_depth, sample_depth;
  int32_t rr_start_depth; /* Russian roulette when depth > this (Path.hs:68 uses 7) */
  int32_t passes;         /* number of progressive passes, >= 1 */
  uint64_t seed;
  int32_t flags;
  int32_t integrator;     /* BGPU_INTEGRATOR_PATH (default) or BGPU_INTEGRATOR_DIRECT */
  /* Per-job camera and film (a RenderJob's camera and imageSize,
   * IO/CameraParser.hs:30-38); NULL selects the scene description's. */
  const bgpu_camera* camera;
  const bgpu_film* film;
} bgpu_job_desc;

enum {
  BGPU_JOB_QUEUED = 0,
  BGPU_JOB_RUNNING = 1,
  BGPU_JOB_DONE = 2,
  BGPU_JOB_CANCELLED = 3,
  BGPU_JOB_FAILED = 4
};

typedef struct {
  uint64_t rays_camera, rays_extend, rays_shadow, rays_mis;
  uint64_t samples;          /* camera samples completed */
  uint64_t jobs_done;
  double gpu_ms;             /* device time of the bounce-loop iterations (events
                                around each batch of iterations) */
  uint64_t nodes_visited;    /* BGPU_COUNT queries: BVH node fetches */
  uint64_t prims_tested;     /* BGPU_COUNT queries: primitive tests */
  uint64_t branches_dropped; /* direct-lighting specular branches lost to a full
                                branch stack; 0 unless an internal bound is broken */
  uint64_t shape_nodes_visited; /* BGPU_COUNT: the shape tree's part of nodes_visited */
  uint64_t shapes_tested;       /* BGPU_COUNT: the shape tests among prims_tested */
} bgpu_stats;

/* ---- entry points ------------------------------------------------------ */
const char* bgpu_last_error(void);
int bgpu_version(void);

int bgpu_init(int device, size_t hbm_budget_bytes, bgpu_ctx** out);
int bgpu_shutdown(bgpu_ctx* ctx);

/* kdTreePrimitive (Primitive/KdTree.hs:236-246) over `mkScene`'s primitives:
 * flattens the description into device-resident SoA arrays and builds the BVH. */
int bgpu_scene_create(bgpu_ctx* ctx, const bgpu_scene_desc* desc, bgpu_scene** out);
int bgpu_scene_destroy(bgpu_scene* scene);
/* Device bytes held by a scene. */
int bgpu_scene_bytes(const bgpu_scene* scene, size_t* bytes);

/* Flat binary scene files (no counterpart in the reference, which parses
 * .bling text on every run, IO/ParserCore.hs:114-147).  bgpu_scene_write
 * stores a descriptor, its arrays and light maps in one file; the record
 * layouts are recorded in it and checked on reading.  bgpu_scene_read
 * returns a descriptor owned by the library, whose arrays point into one
 * block, freed with bgpu_scene_desc_free; a malformed or truncated file is
 * BGPU_E_INVALID.  bgpu_scene_load reads a file and creates the scene. */
int bgpu_scene_write(const bgpu_scene_desc* desc, const char* path);
int bgpu_scene_read(const char* path, bgpu_scene_desc** out);
int bgpu_scene_desc_free(bgpu_scene_desc* desc);
int bgpu_scene_load(bgpu_ctx* ctx, const char* path, bgpu_scene** out);

/* `intersect` of the accelerator primitive (Primitive/KdTree.hs:240-242). */
int bgpu_intersect_closest(bgpu_scene* scene, const bgpu_rays* rays, size_t n,
                           bgpu_hits* hits, int flags);
/* `intersects` of the accelerator primitive (Primitive/KdTree.hs:244-246). */
int bgpu_intersect_any(bgpu_scene* scene, const bgpu_rays* rays, size_t n,
                       uint8_t* occluded, int flags);
/* bgpu_intersect_closest for callers that cannot allocate device memory
 * (code that must not call hipMalloc: graph-building hosts, tight serving
 * loops, harnesses that forbid allocation).  No counterpart in the
 * reference.  Same results, flags and kernels as bgpu_intersect_closest, but
 * it allocates nothing: rays and hits must be device pointers
 * (BGPU_DEVICE_PTRS, else BGPU_E_INVALID), batches are limited to 2^31-1
 * rays as there, and the kernels' counters live in a 256-byte static array
 * that each GPU holds for the life of the process.  Calls on one GPU share
 * that array: calls that overlap on one GPU (from two contexts, or from
 * threads bypassing the context's lock) race. */
int bgpu_intersect_closest_noalloc(bgpu_scene* scene, const bgpu_rays* rays, size_t n,
                                   bgpu_hits* hits, int flags);
/* Evaluates texture `node` of the scene's node list (bgpu_tex_node) on the
 * device at n points: xyz (3n) is the hit point and uv (2n) the surface
 * parameters a texture reads (dgP, dgU, dgV); out gets 16 bands per point
 * for a spectrum texture, one value for a scalar one.  Host arrays.  No
 * counterpart in the reference (whose textures are closures, Texture.hs:59-
 * 414): it exposes the per-hit evaluation that shading uses, for tests and
 * tools. */
int bgpu_scene_texture_eval(bgpu_scene* scene, int node, size_t n, const float* xyz, const float* uv,
                            float* out);
/* Duration in ms of the last BGPU_TIMED intersect launch on this context. */
int bgpu_last_kernel_ms(bgpu_ctx* ctx, double* ms);

/* Plain path: render one job to completion, blocking (prender, Rendering.hs:111-140).
 * wxyz: width*height*4 floats (w, X, Y, Z); sq: width*height*3 or NULL. */
int bgpu_render(bgpu_scene* scene, const bgpu_job_desc* job, float* wxyz, float* sq);

/* Served path: asynchronous jobs that share the device (continuous batching). */
int bgpu_job_submit(bgpu_scene* scene, const bgpu_job_desc* job, bgpu_job** out);
/* Progressive result (PassDone, Rendering.hs:60-75): copies the film so far. */
int bgpu_job_poll(bgpu_job* job, float* wxyz, float* sq, int32_t* passes_done, int32_t* state);
int bgpu_job_wait(bgpu_job* job, double timeout_s);
int bgpu_job_cancel(bgpu_job* job);
int bgpu_job_release(bgpu_job* job);
/* Arrival-to-completion latency of a finished job, in ms. */
int bgpu_job_latency_ms(bgpu_job* job, double* ms);

int bgpu_stats_get(bgpu_ctx* ctx, bgpu_stats* out);
/* Device bytes the context holds (scenes, engines, resident job films) and
 * its HBM budget (no counterpart in the reference).  Jobs whose film would
 * take `used` past `budget` wait in the queue until others release theirs. */
int bgpu_mem_info(bgpu_ctx* ctx, size_t* used, size_t* budget);
int bgpu_stats_reset(bgpu_ctx* ctx);

/* Device time per phase of the bounce loop (no counterpart in the
 * reference), collected while the "pt.profile" knob is 1: the iterations
 * then run as plain kernel launches with events between the phases instead
 * of graph replays.  Phases: 0 generate (camera rays), 1 extend (closest
 * hit of the path rays), 2 classify (grouping by material), 3 shade, 4 shadow
 * (any hit), 5 MIS (closest hit of the BSDF-sampled rays), 6 finish (film).
 * items: queue entries of the phase (active paths; shadow and MIS rays).
 * Cleared by bgpu_stats_reset. */
enum { BGPU_PHASES = 7 };
typedef struct {
  double ms[BGPU_PHASES];
  uint64_t items[BGPU_PHASES];
  uint64_t iterations;
} bgpu_profile;
int bgpu_profile_get(bgpu_ctx* ctx, bgpu_profile* out);

/* Tuning knobs (no counterpart in the reference, whose CPU integrator has
 * none): "bvh.width" (2/4/8 children per BVH node of scenes created
 * afterwards; a scene whose worst-case traversal stack would not fit falls
 * back to a narrower width), "trace.block" (64/128/256, default 256, intersect kernels),
 * "bvh.device" (1, the default: scenes created afterwards with more than
 * 1024 triangles build their triangle tree on the GPU, the same tree as the
 * host builder's; 0: on the host), "trace.persistent" (1: intersect kernels keep lanes busy by taking a new
 * ray when theirs is done; 0: one ray per thread), "trace.stack" and
 * "pt.stack" (8..24, default 12: traversal stack entries per lane kept in
 * LDS by the intersect kernels and by the bounce loop; deeper entries go to
 * private memory), "trace.refill" (1..64, default 32: idle lanes that make
 * a wave of the lane-refill kernels take new rays), "trace.walk" (0: a wave
 * walks inner nodes until every lane is at a leaf; n, default 16: until
 * fewer than n lanes still walk while another waits at a leaf; both
 * traversal kernels), "pt.walk" (the same for the bounce loop's
 * traversals, default 8),
 * "pt.block"
 * (64/128/256, bounce kernels), "pt.persistent" (1: the bounce loop's
 * traversal runs as persistent passes that refill lanes with new rays;
 * 0, the default: one ray per thread),
 * "pt.pool" (path slots, >= 1024), "pt.dl_depth" (0..15, default 4:
 * pending specular branches a direct-lighting path can hold, so its jobs may
 * use maxDepth <= pt.dl_depth + 1; 96 bytes per path slot per entry),
 * "pt.profile" (1: time the bounce loop per phase, bgpu_profile_get; 0, the
 * default: replay the captured graph),
 * "pt.shade_waves" (0: no occupancy hint for the shading kernel, 2: at least
 * two waves per SIMD), "pt.diffuse_waves" (the same for the shading kernel
 * compiled for diffuse-only materials: 0, 2, the default, or 4 waves per
 * SIMD), "pt.sort" (1: the shading queue is grouped by
 * material kind each bounce, so that a wave runs one BSDF code path; 0: path
 * order; 2, the default: grouped while a scene whose materials mix BSDF
 * kinds has jobs in flight).  Changing a pt.* knob is refused with BGPU_E_STATE
 * while served jobs are queued or running; the engines are rebuilt on their
 * next use.  bgpu_config_get reads the current value. */
int bgpu_config_set(bgpu_ctx* ctx, const char* key, int64_t value);
int bgpu_config_get(bgpu_ctx* ctx, const char* key, int64_t* value);

/* Acceleration structure of a scene (no counterpart in the reference). */
typedef struct {
  int32_t width;        /* children per node */
  int32_t nodes;        /* nodes over both trees */
  int32_t leaves;
  int32_t max_depth;    /* of the binary build */
  int32_t stack_bound;  /* worst-case traversal stack entries */
  int32_t device;       /* 1: the triangle tree was built on the GPU */
  double build_ms;      /* binary build of the triangle tree */
} bgpu_bvh_info;
int bgpu_scene_bvh_info(const bgpu_scene* scene, bgpu_bvh_info* out);

#ifdef __cplusplus
}
#endif
#endif /* BGPU_H */
