"""The device BVH build (bling_amd/csrc/bvh_device.hip) against the host
builder (bvh_build.cpp), which it restates on the GPU.

The two trees must be the same tree: equal node, leaf and depth counts and
stack bound at every width, bit-identical closest hits, identical any-hit
answers, and identical node visits and primitive tests for every ray
(traversal visits depend only on the boxes and the child order, which the
numbering of nodes does not change).  Inputs the host splits at the median
(many primitives with one centroid) must fall back to the host build.
"""
import ctypes as C

import numpy as np
import pytest

from bling_amd import abi, scenes
from bling_amd.runtime import GpuRenderer, JobSpec
from bling_amd.scene import Material, SceneBuilder, const

pytestmark = pytest.mark.gpu


def counted(gs, rays, any_hit, coherent):
    ctx = gs.ctx
    n = len(rays)
    cols = [np.ascontiguousarray(rays[:, k].astype(np.float32)) for k in range(8)]
    r = abi.Rays(*[c.ctypes.data for c in cols])
    flags = abi.COUNT | (abi.COHERENT if coherent else 0)
    ctx.reset_stats()
    if any_hit:
        occ = np.empty(n, np.uint8)
        abi.check(ctx.lib.bgpu_intersect_any(gs.h, C.byref(r), n, occ.ctypes.data, flags))
        out = (occ,)
    else:
        t, b1, b2 = (np.empty(n, np.float32) for _ in range(3))
        prim = np.empty(n, np.int32)
        h = abi.Hits(t.ctypes.data, b1.ctypes.data, b2.ctypes.data, prim.ctypes.data)
        abi.check(ctx.lib.bgpu_intersect_closest(gs.h, C.byref(r), n, C.byref(h), flags))
        out = (t.view(np.int32), b1.view(np.int32), b2.view(np.int32), prim)
    st = ctx.stats()
    return out, st["nodes_visited"], st["prims_tested"]


def both(ctx, flat, width):
    ctx.config_set("bvh.width", width)
    ctx.config_set("bvh.device", 0)
    hs = ctx.scene(flat)
    ctx.config_set("bvh.device", 1)
    ds = ctx.scene(flat)
    return hs, ds


def dedup_scene(copies=64, seed=3):
    """A mesh where `copies` triangles share one centroid (the host splits
    such a range at the median), among others."""
    rng = np.random.default_rng(seed)
    n = 3000
    v = rng.uniform(-1, 1, (3 * n, 3)).astype(np.float32)
    v[3:3 * copies] = np.tile(v[:3], (copies - 1, 1))
    sb = SceneBuilder()
    sb.material(Material.matte(const(0.5)))
    sb.mesh(v, np.arange(3 * n).reshape(n, 3))
    return sb.build()


CASES = {
    "icosphere": lambda: scenes.icosphere_scene(levels=6, num_spheres=300, seed=1),
    "stacks": lambda: scenes.stack_scene(120000, res=64, seed=2),
    "grid": lambda: scenes.grid_scene(200000, res=64, seed=1),
    "divergent": lambda: scenes.divergent_scene(res=64, hm=160, n_tori=4, torus_tris=4000, n_spheres=20,
                                                sky=(16, 8)),
}


@pytest.mark.parametrize("name", sorted(CASES))
def test_device_tree_is_the_host_tree(gpu_ctx, name):
    flat = CASES[name]()
    lo, hi = scenes.scene_bounds(flat)
    inc = scenes.incoherent_rays(lo, hi, 100000, seed=7)
    prim_rays = scenes.primary_rays(flat)
    try:
        for width in (2, 4, 8):
            hs, ds = both(gpu_ctx, flat, width)
            try:
                ih, idv = hs.bvh_info(), ds.bvh_info()
                assert ih["device"] == 0 and idv["device"] == 1, (ih, idv)
                for k in ("width", "nodes", "leaves", "max_depth", "stack_bound"):
                    assert ih[k] == idv[k], (name, width, k, ih[k], idv[k])
                for rays, coh in ((inc, False), (prim_rays, True)):
                    (a, an, ap), (b, bn, bp) = counted(hs, rays, False, coh), counted(ds, rays, False, coh)
                    for x, y in zip(a, b):
                        assert np.array_equal(x, y), (name, width, coh)
                    assert (an, ap) == (bn, bp), (name, width, coh, an, bn, ap, bp)
                    (a, an, _), (b, bn, _) = counted(hs, rays, True, coh), counted(ds, rays, True, coh)
                    assert np.array_equal(a[0], b[0]) and an == bn, (name, width, coh)
            finally:
                hs.close()
                ds.close()
    finally:
        gpu_ctx.config_set("bvh.width", 4)
        gpu_ctx.config_set("bvh.device", 1)


def test_median_ranges_fall_back_to_the_host(gpu_ctx):
    flat = dedup_scene()
    gpu_ctx.config_set("bvh.device", 1)
    gs = gpu_ctx.scene(flat)
    try:
        assert gs.bvh_info()["device"] == 0
        lo, hi = scenes.scene_bounds(flat)
        rays = scenes.incoherent_rays(lo, hi, 20000, seed=3)
        t, b1, b2, prim = gs.intersect(rays)
        assert (prim >= 0).sum() > 1000
    finally:
        gs.close()


def test_device_built_scene_renders_like_the_host_built(gpu_ctx):
    flat = scenes.divergent_scene(res=48, hm=128, n_tori=3, torus_tris=3000, n_spheres=12, sky=(16, 8))
    spec = JobSpec(spp_nu=2, spp_nv=2, max_depth=6, sample_depth=3, rr_start_depth=3, passes=1, seed=5)
    hs, ds = both(gpu_ctx, flat, 4)
    try:
        assert ds.bvh_info()["device"] == 1
        a, _ = GpuRenderer(spec).render_plain(hs)
        b, _ = GpuRenderer(spec).render_plain(ds)
    finally:
        hs.close()
        ds.close()
    assert np.array_equal(a[..., 0], b[..., 0])
    np.testing.assert_allclose(b, a, rtol=1e-5, atol=1e-6)
