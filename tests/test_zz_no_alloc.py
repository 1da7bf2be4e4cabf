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


def _dev(torch, a, dtype):
    return torch.from_numpy(np.ascontiguousarray(a)).to("cuda:0", dtype)


def _run(lib, gs, torch, rays, fn, flags):
    n = len(rays)
    cols = [_dev(torch, rays[:, k], torch.float32) for k in range(8)]
    outs = [torch.full((n,), 7.0, dtype=torch.float32, device="cuda:0") for _ in range(3)]
    outs.append(torch.full((n,), -7, dtype=torch.int32, device="cuda:0"))
    r = abi.Rays(*[c.data_ptr() for c in cols])
    h = abi.Hits(*[o.data_ptr() for o in outs])
    abi.check(fn(gs.h, C.byref(r), n, C.byref(h), abi.DEVICE_PTRS | flags))
    torch.cuda.synchronize()
    return [o.cpu().numpy() for o in outs]


def test_noalloc_closest_equals_closest():
    import torch
    flat = scenes.small_scene(num_tris=3000, num_spheres=300, seed=12)
    with Context(0) as ctx:
        lib = ctx.lib
        gs = ctx.scene(flat)
        try:
            lo, hi = scenes.scene_bounds(flat)
            big = scenes.incoherent_rays(lo - 0.1, hi + 0.1, 300_000, seed=31)
            small = scenes.incoherent_rays(lo - 0.1, hi + 0.1, 1_777, seed=32)
            for persistent in (1, 0):
                ctx.config_set("trace.persistent", persistent)
                for rays in (big, small):  # two calls in a row, the second smaller
                    want = _run(lib, gs, torch, rays, lib.bgpu_intersect_closest, 0)
                    got = _run(lib, gs, torch, rays, lib.bgpu_intersect_closest_noalloc, 0)
                    assert (want[3] >= 0).sum() > len(rays) // 10
                    for a, b in zip(want, got):
                        assert np.array_equal(a.view(np.int32), b.view(np.int32))
            ctx.config_set("trace.persistent", 1)
            # BGPU_COUNT: the same node visits and primitive tests
            ctx.reset_stats()
            _run(lib, gs, torch, small, lib.bgpu_intersect_closest, abi.COUNT)
            a = ctx.stats()
            ctx.reset_stats()
            _run(lib, gs, torch, small, lib.bgpu_intersect_closest_noalloc, abi.COUNT)
            b = ctx.stats()
            assert a["nodes_visited"] > 0 and (a["nodes_visited"], a["prims_tested"]) == (
                b["nodes_visited"], b["prims_tested"])
            # host pointers would need staging buffers: refused
            host = np.zeros((4, 8), np.float32)
            cols = [np.ascontiguousarray(host[:, k]) for k in range(8)]
            r = abi.Rays(*[c.ctypes.data for c in cols])
            o = [np.zeros(4, np.float32) for _ in range(3)] + [np.zeros(4, np.int32)]
            h = abi.Hits(*[x.ctypes.data for x in o])
            assert lib.bgpu_intersect_closest_noalloc(gs.h, C.byref(r), 4, C.byref(h), 0) == abi.E_INVALID
        finally:
            gs.close()
