"""Python host mirror of bling's accelerator and renderer interfaces over the
C ABI (include/bgpu.h).

  Accelerator  -- `GpuScene.intersect` / `GpuScene.intersects` batch the
                  `Primitive{intersect, intersects}` closures of the kd-tree
                  primitive (src/lib/Graphics/Bling/Primitive/KdTree.hs:236-246);
                  a miss is prim id -1, as `Nothing`.
  Renderer     -- `GpuRenderer.render(job, report)` follows
                  `class Renderer a where render :: a -> RenderJob ->
                  ProgressReporter -> IO ()` (Rendering.hs:77-78): it calls
                  report(PassDone(pass, image, 1)) after every pass and stops
                  when report returns False (Rendering.hs:111-140).

Everything runs on the GPU through libbgpu.so; there is no CPU fallback.
"""
import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import abi


def _ck(code):
    return abi.check(code)


class Context:
    def __init__(self, device=0, hbm_budget=0):
        self.lib = abi.load()
        h = C.c_void_p()
        _ck(self.lib.bgpu_init(device, hbm_budget, C.byref(h)))
        self.h = h

    def close(self):
        if self.h:
            _ck(self.lib.bgpu_shutdown(self.h))
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def last_kernel_ms(self):
        v = C.c_double()
        _ck(self.lib.bgpu_last_kernel_ms(self.h, C.byref(v)))
        return v.value

    def stats(self):
        s = abi.Stats()
        _ck(self.lib.bgpu_stats_get(self.h, C.byref(s)))
        return {k: getattr(s, k) for k, _ in abi.Stats._fields_}

    def mem_info(self):
        """(device bytes held, HBM budget) of the context."""
        used, budget = C.c_size_t(), C.c_size_t()
        _ck(self.lib.bgpu_mem_info(self.h, C.byref(used), C.byref(budget)))
        return used.value, budget.value

    def profile(self):
        """Per-phase device time of the bounce loop (pt.profile 1):
        {phase: (ms, items)} and the iteration count."""
        p = abi.Profile()
        _ck(self.lib.bgpu_profile_get(self.h, C.byref(p)))
        return {ph: (p.ms[k], p.items[k]) for k, ph in enumerate(abi.PHASES)}, p.iterations

    def reset_stats(self):
        _ck(self.lib.bgpu_stats_reset(self.h))

    def scene(self, flat):
        return GpuScene(self, flat)

    def load_scene(self, path):
        """A scene created from a flat binary scene file (bgpu_scene_load).
        Its `flat` is the file's header data (film, primitive counts)."""
        from .scene import SceneFile
        meta = SceneFile(path)
        meta.close()
        return GpuScene(self, meta, path=path)

    def config_set(self, key, value):
        _ck(self.lib.bgpu_config_set(self.h, key.encode(), int(value)))

    def config_get(self, key):
        v = C.c_int64()
        _ck(self.lib.bgpu_config_get(self.h, key.encode(), C.byref(v)))
        return v.value


def _rays_struct(arrs):
    r = abi.Rays()
    for name, a in zip(("ox", "oy", "oz", "dx", "dy", "dz", "tmin", "tmax"), arrs):
        setattr(r, name, a)
    return r


def _host_f32(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.float32))


class GpuScene:
    def __init__(self, ctx, flat, path=None):
        self.ctx = ctx
        self.flat = flat
        h = C.c_void_p()
        if path is not None:
            _ck(ctx.lib.bgpu_scene_load(ctx.h, str(path).encode(), C.byref(h)))
        else:
            _ck(ctx.lib.bgpu_scene_create(ctx.h, C.byref(flat.desc), C.byref(h)))
        self.h = h

    def close(self):
        if self.h:
            _ck(self.ctx.lib.bgpu_scene_destroy(self.h))
            self.h = None

    def bvh_info(self):
        b = abi.BvhInfo()
        _ck(self.ctx.lib.bgpu_scene_bvh_info(self.h, C.byref(b)))
        return {k: getattr(b, k) for k, _ in abi.BvhInfo._fields_ if k != "pad"}

    def device_bytes(self):
        v = C.c_size_t()
        _ck(self.ctx.lib.bgpu_scene_bytes(self.h, C.byref(v)))
        return v.value

    def texture_eval(self, node, xyz, uv):
        """Texture `node` of the scene's node list evaluated on the device at
        points xyz (n, 3) with surface parameters uv (n, 2): (n, 16) for a
        spectrum texture, (n,) for a scalar one (bgpu_scene_texture_eval)."""
        xyz = np.ascontiguousarray(xyz, np.float32)
        uv = np.ascontiguousarray(uv, np.float32)
        n = len(xyz)
        op = self.flat.tex_nodes[node].op if 0 <= node < self.flat.desc.num_tex_nodes else 0
        out = np.zeros((n, 16) if op >= abi.TN_CONST else n, np.float32)
        _ck(self.ctx.lib.bgpu_scene_texture_eval(self.h, node, n, xyz.ctypes.data, uv.ctypes.data,
                                                  out.ctypes.data))
        return out

    # --- host arrays ---------------------------------------------------
    def intersect(self, rays, coherent=False):
        """rays: (n, 8) float32 [ox oy oz dx dy dz tmin tmax]; returns
        (t, b1, b2, prim) numpy arrays.  `coherent`: the batch's rays are
        coherent (e.g. camera rays), the BGPU_COHERENT hint."""
        rays = _host_f32(rays)
        n = len(rays)
        cols = [np.ascontiguousarray(rays[:, k]) for k in range(8)]
        t = np.empty(n, np.float32)
        b1 = np.empty(n, np.float32)
        b2 = np.empty(n, np.float32)
        prim = np.empty(n, np.int32)
        r = _rays_struct([c.ctypes.data for c in cols])
        h = abi.Hits(t.ctypes.data, b1.ctypes.data, b2.ctypes.data, prim.ctypes.data)
        _ck(self.ctx.lib.bgpu_intersect_closest(self.h, C.byref(r), n, C.byref(h),
                                                abi.COHERENT if coherent else 0))
        return t, b1, b2, prim

    def intersects(self, rays, coherent=False):
        rays = _host_f32(rays)
        n = len(rays)
        cols = [np.ascontiguousarray(rays[:, k]) for k in range(8)]
        occ = np.empty(n, np.uint8)
        r = _rays_struct([c.ctypes.data for c in cols])
        _ck(self.ctx.lib.bgpu_intersect_any(self.h, C.byref(r), n, occ.ctypes.data,
                                            abi.COHERENT if coherent else 0))
        return occ.astype(bool)

    # --- device-resident arrays (integer device addresses) -------------
    def intersect_dev(self, ray_ptrs, n, hit_ptrs, timed=False, coherent=False):
        r = _rays_struct(ray_ptrs)
        h = abi.Hits(*hit_ptrs)
        flags = abi.DEVICE_PTRS | (abi.TIMED if timed else 0) | (abi.COHERENT if coherent else 0)
        _ck(self.ctx.lib.bgpu_intersect_closest(self.h, C.byref(r), n, C.byref(h), flags))
        return self.ctx.last_kernel_ms() if timed else None

    def intersects_dev(self, ray_ptrs, n, occ_ptr, timed=False, coherent=False):
        r = _rays_struct(ray_ptrs)
        flags = abi.DEVICE_PTRS | (abi.TIMED if timed else 0) | (abi.COHERENT if coherent else 0)
        _ck(self.ctx.lib.bgpu_intersect_any(self.h, C.byref(r), n, occ_ptr, flags))
        return self.ctx.last_kernel_ms() if timed else None


@dataclass
class JobSpec:
    """A render job: `sampler { stratified nu nv }` and
    `integrator { path maxDepth md sampleDepth sd }` (integrator "path") or
    `integrator { directLighting maxDepth md }` (integrator "direct"),
    `passes` passes."""
    spp_nu: int = 2
    spp_nv: int = 2
    max_depth: int = 7
    sample_depth: int = 3
    rr_start_depth: int = 7
    passes: int = 1
    seed: int = 1
    variance: bool = False
    view: object = None  # scene.View; None renders the scene's own camera and film
    integrator: str = "path"

    def desc(self):
        d = abi.JobDesc()
        d.spp_nu, d.spp_nv = self.spp_nu, self.spp_nv
        d.max_depth, d.sample_depth = self.max_depth, self.sample_depth
        d.rr_start_depth, d.passes = self.rr_start_depth, self.passes
        d.seed = self.seed & 0xFFFFFFFFFFFFFFFF
        d.flags = abi.JOB_VARIANCE if self.variance else 0
        if self.integrator not in ("path", "direct"):
            raise ValueError("integrator must be 'path' or 'direct'")
        d.integrator = abi.INTEGRATOR_DIRECT if self.integrator == "direct" else abi.INTEGRATOR_PATH
        if self.view is not None:
            d.camera = C.pointer(self.view.camera)
            d.film = C.pointer(self.view.film)
        return d

    def image_shape(self, flat):
        """(height, width) of the job's film on the FlatScene `flat`."""
        if self.view is not None:
            return self.view.film.height, self.view.film.width
        return flat.film[1], flat.film[0]


@dataclass
class PassDone:
    """Progress event after a pass (Rendering.hs:60-75)."""
    pass_num: int
    final_img: np.ndarray  # (h, w, 4) float32 = (weight, X, Y, Z)
    splat_weight: float = 1.0


class Job:
    def __init__(self, scene, spec):
        self.scene = scene
        self.spec = spec
        lib = scene.ctx.lib
        h = C.c_void_p()
        _ck(lib.bgpu_job_submit(scene.h, C.byref(spec.desc()), C.byref(h)))
        self.h = h
        self.shape = spec.image_shape(scene.flat)

    def poll(self, want_image=True):
        lib = self.scene.ctx.lib
        h, w = self.shape
        img = np.empty((h, w, 4), np.float32) if want_image else None
        sq = np.empty((h, w, 3), np.float32) if (want_image and self.spec.variance) else None
        done = C.c_int32()
        state = C.c_int32()
        _ck(lib.bgpu_job_poll(self.h, img.ctypes.data if img is not None else None,
                              sq.ctypes.data if sq is not None else None,
                              C.byref(done), C.byref(state)))
        return img, sq, done.value, state.value

    def wait(self, timeout=600.0):
        _ck(self.scene.ctx.lib.bgpu_job_wait(self.h, float(timeout)))

    def cancel(self):
        _ck(self.scene.ctx.lib.bgpu_job_cancel(self.h))

    def latency_ms(self):
        v = C.c_double()
        _ck(self.scene.ctx.lib.bgpu_job_latency_ms(self.h, C.byref(v)))
        return v.value

    def release(self):
        if self.h:
            _ck(self.scene.ctx.lib.bgpu_job_release(self.h))
            self.h = None


class GpuRenderer:
    """The `gpu` renderer: one job, progressive passes (Rendering.hs:111-140)."""

    def __init__(self, spec):
        self.spec = spec

    def render_plain(self, scene):
        """Blocking one-job path: returns ((h, w, 4) wxyz, (h, w, 3) sum of squares or None)."""
        lib = scene.ctx.lib
        h, w = self.spec.image_shape(scene.flat)
        img = np.zeros((h, w, 4), np.float32)
        sq = np.zeros((h, w, 3), np.float32) if self.spec.variance else None
        _ck(lib.bgpu_render(scene.h, C.byref(self.spec.desc()), img.ctypes.data,
                            sq.ctypes.data if sq is not None else None))
        return img, sq

    def render(self, scene, report, poll_interval=0.005):
        """Progressive render; report(PassDone) after each completed pass,
        stopping early when it returns False."""
        import time
        job = Job(scene, self.spec)
        try:
            seen = 0
            while True:
                _, _, done, state = job.poll(want_image=False)
                if done > seen:
                    seen = done
                    img, _, _, _ = job.poll()
                    if not report(PassDone(done, img, 1.0)):
                        job.cancel()
                        break
                if state in (abi.JOB_DONE, abi.JOB_CANCELLED, abi.JOB_FAILED):
                    break
                time.sleep(poll_interval)
        finally:
            job.release()
