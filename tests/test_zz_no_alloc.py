"""bgpu_intersect_closest_noalloc against bgpu_intersect_closest.

The allocation-free entry point runs the same kernels with their counters
in the library's static per-GPU scratch array; its results must equal the
allocating entry point's bit for bit, on two calls in a row (the second
with a smaller batch, so that what the first left in the array must not
leak into it), through the persistent lane-refill traversal (which uses
the counters) and the one-ray-per-thread one, and with BGPU_COUNT (whose
counters live in the same array) giving the same visit counts.
"""
import ctypes as C

import numpy as np
import pytest

from bling_amd import abi, scenes
from bling_amd.runtime import Context

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2  # hipMemcpyHostToDevice, hipMemcpyDeviceToHost


class _Hip:
    """Device buffers for the test through the HIP runtime that libbgpu.so
    already loaded (the same soname resolves to the loaded copy), so the test
    needs no other GPU library."""

    def __init__(self):
        self.rt = C.CDLL("libamdhip64.so.7")
        self.rt.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.rt.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.rt.hipFree.argtypes = [C.c_void_p]
        self.live = []

    def put(self, a):
        a = np.ascontiguousarray(a)
        p = C.c_void_p()
        assert self.rt.hipMalloc(C.byref(p), max(4, a.nbytes)) == 0
        self.live.append(p)
        assert self.rt.hipMemcpy(p, a.ctypes.data, a.nbytes, H2D) == 0
        return p.value

    def get(self, ptr, like):
        out = np.empty_like(like)
        assert self.rt.hipMemcpy(out.ctypes.data, ptr, out.nbytes, D2H) == 0
        return out

    def free(self):
        for p in self.live:
            self.rt.hipFree(p)
        self.live = []


def _run(hip, gs, rays, fn, flags):
    n = len(rays)
    rp = [hip.put(rays[:, k].astype(np.float32)) for k in range(8)]
    init = [np.full(n, 7.0, np.float32)] * 3 + [np.full(n, -7, np.int32)]
    hp = [hip.put(a) for a in init]
    r = abi.Rays(*rp)
    h = abi.Hits(*hp)
    abi.check(fn(gs.h, C.byref(r), n, C.byref(h), abi.DEVICE_PTRS | flags))
    out = [hip.get(p, a) for p, a in zip(hp, init)]
    hip.free()
    return out


def test_noalloc_closest_equals_closest():
    flat = scenes.small_scene(num_tris=3000, num_spheres=300, seed=12)
    with Context(0) as ctx:
        lib = ctx.lib
        hip = _Hip()
        gs = ctx.scene(flat)
        try:
            lo, hi = scenes.scene_bounds(flat)
            big = scenes.incoherent_rays(lo - 0.1, hi + 0.1, 300_000, seed=31)
            small = scenes.incoherent_rays(lo - 0.1, hi + 0.1, 1_777, seed=32)
            for persistent in (1, 0):
                ctx.config_set("trace.persistent", persistent)
                for rays in (big, small):  # two calls in a row, the second smaller
                    want = _run(hip, gs, rays, lib.bgpu_intersect_closest, 0)
                    got = _run(hip, gs, rays, lib.bgpu_intersect_closest_noalloc, 0)
                    assert (want[3] >= 0).sum() > len(rays) // 10
                    for a, b in zip(want, got):
                        assert np.array_equal(a.view(np.int32), b.view(np.int32))
            ctx.config_set("trace.persistent", 1)
            # BGPU_COUNT: the same node visits and primitive tests
            ctx.reset_stats()
            _run(hip, gs, small, lib.bgpu_intersect_closest, abi.COUNT)
            a = ctx.stats()
            ctx.reset_stats()
            _run(hip, gs, small, lib.bgpu_intersect_closest_noalloc, abi.COUNT)
            b = ctx.stats()
            keys = ("nodes_visited", "prims_tested", "shape_nodes_visited", "shapes_tested")
            assert a["nodes_visited"] > 0 and a["shapes_tested"] > 0
            assert [a[k] for k in keys] == [b[k] for k in keys]
            # host pointers would need staging buffers: refused
            host = np.zeros((4, 8), np.float32)
            cols = [np.ascontiguousarray(host[:, k]) for k in range(8)]
            r = abi.Rays(*[c.ctypes.data for c in cols])
            o = [np.zeros(4, np.float32) for _ in range(3)] + [np.zeros(4, np.int32)]
            h = abi.Hits(*[x.ctypes.data for x in o])
            assert lib.bgpu_intersect_closest_noalloc(gs.h, C.byref(r), 4, C.byref(h), 0) == abi.E_INVALID
        finally:
            gs.close()
