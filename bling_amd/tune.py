"""Autotuner for the library's knobs (bgpu_config_set).

tune_trace()    picks the BVH width, the traversal mode (lane refill or one
                ray per thread), the block size and the LDS stack depth of the
                ray-batch kernels for a scene and a sample of its rays, timing
                each candidate with the library's own kernel timer
                (BGPU_TIMED);
tune_launches() the same over knobs that need no new scene, for launches the
                caller already holds (e.g. resident device batches);
tune_render()   picks the block size of the bounce kernels, the shading
                kernels' occupancy variants, the LDS stack depth, the
                path-pool size and (when asked) the shading-queue grouping
                for a render job, timing whole renders.

All three leave the winning configuration set on the context and return it with
the table of timings.  Results never depend on the knobs (the tests check
bit-exact intersections and equal images for every setting), so the choice
is purely about speed.
"""
import itertools
import time

import numpy as np

from . import abi


def _best(table):
    return min(table, key=lambda row: row["ms"])


def tune_launches(ctx, launches, space, repeats=3):
    """Times `launches` (callables taking timed=True and returning the kernel
    time in ms) for every combination of the knob values in `space`
    ({key: values}); a candidate's score is the sum over launches of the
    median of `repeats` timings.  Leaves the best combination set and returns
    (best, table)."""
    keys = list(space)
    keep = {k: ctx.config_get(k) for k in keys}
    table = []
    try:
        for combo in itertools.product(*(space[k] for k in keys)):
            for k, v in zip(keys, combo):
                ctx.config_set(k, v)
            for fn in launches:  # warm-up (and the grid-size query of a new depth)
                fn(timed=True)
            ms = sum(float(np.median([fn(timed=True) for _ in range(repeats)])) for fn in launches)
            row = dict(zip(keys, combo))
            row["ms"] = round(ms, 4)
            table.append(row)
    finally:
        for k, v in keep.items():
            ctx.config_set(k, v)
    best = _best(table)
    for k in keys:
        ctx.config_set(k, best[k])
    return best, table


def tune_trace(ctx, flat, rays, widths=(2, 4, 8), blocks=(64, 128, 256), modes=(1, 0), stacks=(8, 12, 16),
               walks=(0, 16), repeats=3, coherent=False):
    """Times closest + any hit over `rays` ((n, 8) float32, host) for every
    (bvh.width, trace.persistent, trace.block, trace.stack, trace.walk);
    returns (best, table).  A scene is built per width (the width is fixed at
    scene creation)."""
    table = []
    keys = ("bvh.width", "trace.persistent", "trace.block", "trace.stack", "trace.walk")
    keep = {k: ctx.config_get(k) for k in keys}
    flags = abi.TIMED | (abi.COHERENT if coherent else 0)
    try:
        for w in widths:
            ctx.config_set("bvh.width", w)
            gs = ctx.scene(flat)
            try:
                built = gs.bvh_info()["width"]
                dev = _HostRays(ctx, rays)
                for mode, block, stack, walk in itertools.product(modes, blocks, stacks, walks):
                    ctx.config_set("trace.persistent", mode)
                    ctx.config_set("trace.block", block)
                    ctx.config_set("trace.stack", stack)
                    ctx.config_set("trace.walk", walk)
                    dev.closest(gs, flags)  # warm-up (and the grid-size query)
                    dev.any(gs, flags)
                    ms = []
                    for _ in range(repeats):
                        ms.append(dev.closest(gs, flags) + dev.any(gs, flags))
                    table.append({"bvh.width": built, "trace.persistent": mode, "trace.block": block,
                                  "trace.stack": stack, "trace.walk": walk, "ms": float(np.median(ms))})
                dev.close()
            finally:
                gs.close()
    finally:
        for k, v in keep.items():
            ctx.config_set(k, v)
    best = _best(table)
    for k in keys:
        ctx.config_set(k, best[k])
    return best, table


def tune_render(ctx, gscene, spec, blocks=(64, 128, 256), shade_waves=(2, 0), diffuse_waves=(2,),
                stacks=(12,), walks=(8,), pools=(1 << 19, 1 << 20, 1 << 21), sorts=(2,), repeats=2):
    """Times the plain render of `spec` on `gscene` for every (pt.block,
    pt.shade_waves, pt.diffuse_waves, pt.stack, pt.walk, pt.pool, pt.sort);
    returns (best, table)."""
    from .runtime import GpuRenderer
    table = []
    keys = ("pt.block", "pt.shade_waves", "pt.diffuse_waves", "pt.stack", "pt.walk", "pt.pool", "pt.sort")
    keep = {k: ctx.config_get(k) for k in keys}
    try:
        for combo in itertools.product(blocks, shade_waves, diffuse_waves, stacks, walks, pools, sorts):
            for k, v in zip(keys, combo):
                ctx.config_set(k, v)
            GpuRenderer(spec).render_plain(gscene)  # engine build + warm-up
            ms = []
            for _ in range(repeats):
                t0 = time.perf_counter()
                GpuRenderer(spec).render_plain(gscene)
                ms.append((time.perf_counter() - t0) * 1e3)
            row = dict(zip(keys, combo))
            row["ms"] = float(np.median(ms))
            table.append(row)
    finally:
        for k, v in keep.items():
            ctx.config_set(k, v)
    best = _best(table)
    for k in keys:
        ctx.config_set(k, best[k])
    return best, table


class _HostRays:
    """A ray batch and its result arrays in host memory.  BGPU_TIMED times the
    kernels only (the batch is staged to the device before the start event),
    so the timings are those of resident inputs."""

    def __init__(self, ctx, rays):
        import ctypes as C
        self.C = C
        self.ctx = ctx
        self.n = len(rays)
        self.cols = [np.ascontiguousarray(rays[:, k], dtype=np.float32) for k in range(8)]
        self.outs = [np.empty(self.n, np.float32) for _ in range(3)] + [np.empty(self.n, np.int32)]
        self.occ = np.empty(self.n, np.uint8)
        self.r = abi.Rays(*[c.ctypes.data for c in self.cols])
        self.h = abi.Hits(*[o.ctypes.data for o in self.outs])

    def closest(self, gs, flags):
        abi.check(self.ctx.lib.bgpu_intersect_closest(gs.h, self.C.byref(self.r), self.n,
                                                      self.C.byref(self.h), flags))
        return self.ctx.last_kernel_ms()

    def any(self, gs, flags):
        abi.check(self.ctx.lib.bgpu_intersect_any(gs.h, self.C.byref(self.r), self.n,
                                                  self.occ.ctypes.data, flags))
        return self.ctx.last_kernel_ms()

    def close(self):
        self.cols = self.outs = self.occ = None
