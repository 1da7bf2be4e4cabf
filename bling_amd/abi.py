"""ctypes mirror of include/bgpu.h and the loader of libbgpu.so.

The library is the product: every query and render goes through it.  Loading
fails loudly when the shared library has not been built.
"""
import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libbgpu.so")

BANDS = 16

# error codes
OK, E_INVALID, E_NOMEM, E_HIP, E_BUDGET, E_STATE, E_CANCELLED, E_NODEV, E_TIMEOUT = (
    0, -1, -2, -3, -4, -5, -6, -7, -8)

SHAPE_BOX, SHAPE_CYLINDER, SHAPE_DISK, SHAPE_QUAD, SHAPE_SPHERE = range(5)
MAT_BLACK, MAT_MATTE, MAT_PLASTIC, MAT_METAL, MAT_GLASS, MAT_MIRROR, MAT_TRANSMATTE, MAT_SUBSTRATE, \
    MAT_SHINYMETAL = range(9)
MESH_NORMALS, MESH_UVS = 1, 2
LIGHT_INFINITE, LIGHT_POINT, LIGHT_DIRECTIONAL = 1, 2, 3
LIGHTMAP_TEXELS, LIGHTMAP_SUNSKY = 0, 1
CAMERA_PERSPECTIVE, CAMERA_ENVIRONMENT = 0, 1
FILTER_BOX, FILTER_GAUSS, FILTER_MITCHELL, FILTER_SINC, FILTER_TRIANGLE = range(5)
DEVICE_PTRS, TIMED, COUNT, COHERENT = 1, 2, 4, 8
JOB_VARIANCE = 1
INTEGRATOR_PATH, INTEGRATOR_DIRECT = 0, 1
JOB_QUEUED, JOB_RUNNING, JOB_DONE, JOB_CANCELLED, JOB_FAILED = range(5)

F16 = C.c_float * 16
FP = C.POINTER(C.c_float)
IP = C.POINTER(C.c_int32)


class Shape(C.Structure):
    _fields_ = [("type", C.c_int32), ("material", C.c_int32), ("emissive", C.c_int32),
                ("light_id", C.c_int32), ("params", C.c_float * 6), ("o2w", F16),
                ("w2o", F16), ("emission", F16)]


TEX_NONE, TEX_CHECKER, TEX_GRAPHPAPER, TEX_NODES = 0, 1, 2, 3
MAP_UV, MAP_PLANAR, MAP_IDENTITY = 0, 1, 2
# texture node operations (bgpu_tex_node.op)
TN_SCONST, TN_SIMAGE, TN_CELL, TN_FBM, TN_PERLIN, TN_CRYSTAL, TN_SCALE = 1, 2, 3, 4, 5, 6, 7
TN_CONST, TN_IMAGE, TN_BLEND, TN_CHECKER, TN_GRADIENT, TN_GRAPHPAPER, TN_GSTEP = 16, 17, 18, 19, 20, 21, 22


class Texture(C.Structure):
    # `params` shares its storage with the int32 `roots` of BGPU_TEX_NODES
    # (a C union); texture_roots() views it as integers
    _fields_ = [("kind", C.c_int32), ("mapping", C.c_int32), ("slots", C.c_int32), ("pad", C.c_int32),
                ("params", C.c_float * 12), ("alt", (C.c_float * 16) * 2)]


def texture_roots(tex):
    """The roots[12] view of a Texture's params (BGPU_TEX_NODES)."""
    return (C.c_int32 * 12).from_buffer(tex, Texture.params.offset)


class TexNode(C.Structure):
    _fields_ = [("op", C.c_int32), ("a", C.c_int32), ("b", C.c_int32), ("c", C.c_int32), ("map", C.c_int32),
                ("pad", C.c_int32 * 3), ("q", C.c_float * 4), ("m", C.c_float * 16), ("spec", C.c_float * 16)]


class Image(C.Structure):
    _fields_ = [("width", C.c_int32), ("height", C.c_int32), ("channels", C.c_int32), ("pad", C.c_int32),
                ("texels", FP)]


class Material(C.Structure):
    _fields_ = [("type", C.c_int32), ("pad", C.c_int32), ("scalar", C.c_float * 4),
                ("spec", (C.c_float * 16) * 3), ("tex", Texture)]


class Mesh(C.Structure):
    _fields_ = [("material", C.c_int32), ("flags", C.c_int32)]


class Light(C.Structure):
    _fields_ = [("type", C.c_int32), ("width", C.c_int32), ("height", C.c_int32),
                ("map", C.c_int32), ("radiance", FP), ("w2l", F16), ("l2w", F16), ("vec", C.c_float * 4),
                ("sky", C.c_float * 8)]


class Camera(C.Structure):
    _fields_ = [("type", C.c_int32), ("pad", C.c_int32), ("c2w", F16), ("r2c", F16),
                ("lens_radius", C.c_float), ("focal_distance", C.c_float),
                ("xres", C.c_float), ("yres", C.c_float)]


class Film(C.Structure):
    _fields_ = [("width", C.c_int32), ("height", C.c_int32), ("filter", C.c_int32),
                ("pad", C.c_int32), ("fparams", C.c_float * 4)]


class SceneDesc(C.Structure):
    _fields_ = [("num_vertices", C.c_uint32), ("num_triangles", C.c_uint32),
                ("vertices", FP), ("normals", FP), ("uvs", FP),
                ("indices", IP), ("tri_mesh", IP),
                ("num_meshes", C.c_uint32), ("meshes", C.POINTER(Mesh)),
                ("num_shapes", C.c_uint32), ("shapes", C.POINTER(Shape)),
                ("num_materials", C.c_uint32), ("materials", C.POINTER(Material)),
                ("num_lights", C.c_uint32), ("lights", C.POINTER(Light)),
                ("camera", Camera), ("film", Film),
                ("num_tex_nodes", C.c_uint32), ("tex_nodes", C.POINTER(TexNode)),
                ("num_images", C.c_uint32), ("images", C.POINTER(Image))]


class Rays(C.Structure):
    _fields_ = [(n, C.c_void_p) for n in ("ox", "oy", "oz", "dx", "dy", "dz", "tmin", "tmax")]


class Hits(C.Structure):
    _fields_ = [("t", C.c_void_p), ("b1", C.c_void_p), ("b2", C.c_void_p), ("prim", C.c_void_p)]


class JobDesc(C.Structure):
    _fields_ = [("spp_nu", C.c_int32), ("spp_nv", C.c_int32), ("max_depth", C.c_int32),
                ("sample_depth", C.c_int32), ("rr_start_depth", C.c_int32),
                ("passes", C.c_int32), ("seed", C.c_uint64), ("flags", C.c_int32),
                ("integrator", C.c_int32), ("camera", C.POINTER(Camera)), ("film", C.POINTER(Film))]


PHASES = ("generate", "extend", "classify", "shade", "shadow", "mis", "finish")


class Profile(C.Structure):
    _fields_ = [("ms", C.c_double * 7), ("items", C.c_uint64 * 7), ("iterations", C.c_uint64)]


class BvhInfo(C.Structure):
    _fields_ = [("width", C.c_int32), ("nodes", C.c_int32), ("leaves", C.c_int32),
                ("max_depth", C.c_int32), ("stack_bound", C.c_int32), ("device", C.c_int32),
                ("build_ms", C.c_double)]


class Stats(C.Structure):
    _fields_ = [("rays_camera", C.c_uint64), ("rays_extend", C.c_uint64),
                ("rays_shadow", C.c_uint64), ("rays_mis", C.c_uint64),
                ("samples", C.c_uint64), ("jobs_done", C.c_uint64), ("gpu_ms", C.c_double),
                ("nodes_visited", C.c_uint64), ("prims_tested", C.c_uint64),
                ("branches_dropped", C.c_uint64), ("shape_nodes_visited", C.c_uint64),
                ("shapes_tested", C.c_uint64)]


# every symbol declared in include/bgpu.h
EXPORTS = [
    "bgpu_last_error", "bgpu_version", "bgpu_init", "bgpu_shutdown", "bgpu_scene_create",
    "bgpu_scene_destroy", "bgpu_scene_bytes", "bgpu_intersect_closest", "bgpu_intersect_any",
    "bgpu_last_kernel_ms", "bgpu_render", "bgpu_job_submit", "bgpu_job_poll", "bgpu_job_wait",
    "bgpu_job_cancel", "bgpu_job_release", "bgpu_job_latency_ms", "bgpu_stats_get",
    "bgpu_stats_reset", "bgpu_config_set", "bgpu_config_get", "bgpu_scene_bvh_info",
    "bgpu_scene_write", "bgpu_scene_read", "bgpu_scene_desc_free", "bgpu_scene_load",
    "bgpu_mem_info", "bgpu_profile_get", "bgpu_intersect_closest_noalloc", "bgpu_scene_texture_eval",
]

_lib = None


class BgpuError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("bgpu error %d: %s" % (code, msg))
        self.code = code


def load(path=LIB_PATH):
    """Loads libbgpu.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError("libbgpu.so not built: run `python -m bling_amd.build` (%s)" % path)
    lib = C.CDLL(path)
    vp = C.c_void_p
    lib.bgpu_last_error.restype = C.c_char_p
    lib.bgpu_version.restype = C.c_int
    lib.bgpu_init.argtypes = [C.c_int, C.c_size_t, C.POINTER(vp)]
    lib.bgpu_shutdown.argtypes = [vp]
    lib.bgpu_scene_create.argtypes = [vp, C.POINTER(SceneDesc), C.POINTER(vp)]
    lib.bgpu_scene_destroy.argtypes = [vp]
    lib.bgpu_scene_bytes.argtypes = [vp, C.POINTER(C.c_size_t)]
    lib.bgpu_intersect_closest.argtypes = [vp, C.POINTER(Rays), C.c_size_t, C.POINTER(Hits), C.c_int]
    lib.bgpu_intersect_any.argtypes = [vp, C.POINTER(Rays), C.c_size_t, vp, C.c_int]
    lib.bgpu_intersect_closest_noalloc.argtypes = [vp, C.POINTER(Rays), C.c_size_t, C.POINTER(Hits), C.c_int]
    lib.bgpu_scene_texture_eval.argtypes = [vp, C.c_int, C.c_size_t, vp, vp, vp]
    lib.bgpu_last_kernel_ms.argtypes = [vp, C.POINTER(C.c_double)]
    lib.bgpu_render.argtypes = [vp, C.POINTER(JobDesc), vp, vp]
    lib.bgpu_job_submit.argtypes = [vp, C.POINTER(JobDesc), C.POINTER(vp)]
    lib.bgpu_job_poll.argtypes = [vp, vp, vp, C.POINTER(C.c_int32), C.POINTER(C.c_int32)]
    lib.bgpu_job_wait.argtypes = [vp, C.c_double]
    lib.bgpu_job_cancel.argtypes = [vp]
    lib.bgpu_job_release.argtypes = [vp]
    lib.bgpu_job_latency_ms.argtypes = [vp, C.POINTER(C.c_double)]
    lib.bgpu_stats_get.argtypes = [vp, C.POINTER(Stats)]
    lib.bgpu_stats_reset.argtypes = [vp]
    lib.bgpu_mem_info.argtypes = [vp, C.POINTER(C.c_size_t), C.POINTER(C.c_size_t)]
    lib.bgpu_profile_get.argtypes = [vp, C.POINTER(Profile)]
    lib.bgpu_config_set.argtypes = [vp, C.c_char_p, C.c_int64]
    lib.bgpu_config_get.argtypes = [vp, C.c_char_p, C.POINTER(C.c_int64)]
    lib.bgpu_scene_bvh_info.argtypes = [vp, C.POINTER(BvhInfo)]
    lib.bgpu_scene_write.argtypes = [C.POINTER(SceneDesc), C.c_char_p]
    lib.bgpu_scene_read.argtypes = [C.c_char_p, C.POINTER(C.POINTER(SceneDesc))]
    lib.bgpu_scene_desc_free.argtypes = [C.POINTER(SceneDesc)]
    lib.bgpu_scene_load.argtypes = [vp, C.c_char_p, C.POINTER(vp)]
    _lib = lib
    return lib


def check(code):
    if code != OK:
        raise BgpuError(code, load().bgpu_last_error().decode(errors="replace"))
    return code
