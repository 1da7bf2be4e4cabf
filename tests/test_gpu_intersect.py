"""GPU parity for the accelerator queries (`intersect` / `intersects` of the
kd-tree primitive, src/lib/Graphics/Bling/Primitive/KdTree.hs:236-246) against
the oracle, through the C ABI.

Bar: hit/miss and primitive id identical; t, b1, b2 bit-identical (binary32).
"""
import numpy as np
import pytest

from bling_amd import abi, scenes
from bling_amd import transform as T
from bling_amd.scene import Material, SceneBuilder, const
from oracle.pyoracle import OracleScene

pytestmark = pytest.mark.gpu


def assert_same(gpu, orc, label):
    t, b1, b2, prim = gpu
    ot, ob1, ob2, oprim, _ = orc
    bad = np.nonzero(prim != oprim)[0]
    assert len(bad) == 0, "%s: %d prim mismatches, first %s gpu=%s oracle=%s" % (
        label, len(bad), bad[:5], prim[bad[:5]], oprim[bad[:5]])
    hit = prim >= 0
    for name, a, b in (("t", t, ot), ("b1", b1, ob1), ("b2", b2, ob2)):
        diff = np.nonzero(a[hit].view(np.int32) != b[hit].view(np.int32))[0]
        assert len(diff) == 0, "%s: %d %s bit mismatches, e.g. %r vs %r" % (
            label, len(diff), name, a[hit][diff[:3]], b[hit][diff[:3]])


def run_both(ctx, flat, rays, brute=False):
    gs = ctx.scene(flat)
    try:
        g = gs.intersect(rays)
        gocc = gs.intersects(rays)
    finally:
        gs.close()
    o = OracleScene(flat)
    oc = o.intersect(rays, brute=brute)
    oocc = o.intersects(rays, brute=brute)
    o.close()
    return g, oc, gocc, oocc


def test_small_scene_all_shapes(gpu_ctx):
    flat = scenes.small_scene(num_tris=3000, num_spheres=400, seed=11)
    lo, hi = scenes.scene_bounds(flat)
    rays = scenes.incoherent_rays(lo - 0.3, hi + 0.3, 60000, seed=4)
    g, o, gocc, oocc = run_both(gpu_ctx, flat, rays, brute=True)
    assert_same(g, o, "small/incoherent")
    assert (gocc == oocc).all(), "any-hit mismatches: %d" % (gocc != oocc).sum()
    assert (g[3] >= flat.num_triangles).sum() > 100  # shapes are actually hit


def test_cornell_primary_and_axis_rays(gpu_ctx):
    flat = scenes.cornell(128)
    rays = scenes.primary_rays(flat)
    g, o, gocc, oocc = run_both(gpu_ctx, flat, rays, brute=True)
    assert_same(g, o, "cornell/primary")
    assert (gocc == oocc).all()
    # axis-aligned rays in and along the wall planes (zero direction components,
    # origins exactly on faces): exercise the slab NaN/inf paths
    rng = np.random.default_rng(3)
    n = 20000
    o_ = rng.choice([0.0, 0.25, 0.5, 1.0], (n, 3)).astype(np.float32)
    d = np.zeros((n, 3), np.float32)
    ax = rng.integers(0, 3, n)
    d[np.arange(n), ax] = rng.choice([-1.0, 1.0], n)
    r = np.concatenate([o_, d, np.zeros((n, 1)), np.full((n, 1), np.inf)], 1).astype(np.float32)
    r[:, 6] = rng.choice([0.0, 1e-3], n)
    g, o, gocc, oocc = run_both(gpu_ctx, flat, r, brute=True)
    assert_same(g, o, "cornell/axis")
    assert (gocc == oocc).all()


def test_ray_intervals_and_edge_batches(gpu_ctx):
    flat = scenes.small_scene(num_tris=1500, num_spheres=150, seed=12)
    lo, hi = scenes.scene_bounds(flat)
    rays = scenes.incoherent_rays(lo, hi, 20000, seed=5)
    rng = np.random.default_rng(9)
    rays[:, 6] = rng.uniform(0, 0.5, len(rays)) * (rng.random(len(rays)) < 0.5)
    rays[:, 7] = np.where(rng.random(len(rays)) < 0.5, rng.uniform(0.1, 2.0, len(rays)), np.inf)
    # non-normalised directions, as transformed rays carry (Transform.hs:275-280)
    rays[:, 3:6] *= rng.uniform(0.2, 3.0, (len(rays), 1)).astype(np.float32)
    g, o, gocc, oocc = run_both(gpu_ctx, flat, rays, brute=True)
    assert_same(g, o, "intervals")
    assert (gocc == oocc).all()
    gs = gpu_ctx.scene(flat)
    try:
        t, b1, b2, prim = gs.intersect(np.zeros((0, 8), np.float32))
        assert len(prim) == 0
        one = gs.intersect(rays[:1])
        assert one[3][0] == g[3][0]
    finally:
        gs.close()


def test_empty_and_tiny_scenes(gpu_ctx):
    sb = SceneBuilder()
    sb.material(Material.matte(const(0.5)))
    sb.mesh(np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], np.float32), np.array([[0, 1, 2]]))
    flat = sb.build()
    r = np.array([[0.25, 0.25, -1, 0, 0, 1, 0, np.inf], [2, 2, -1, 0, 0, 1, 0, np.inf]], np.float32)
    g, o, gocc, oocc = run_both(gpu_ctx, flat, r, brute=True)
    assert_same(g, o, "one-triangle")
    # closed form: axis ray through (0.25, 0.25) hits at t = 1 with b1 = b2 = 0.25
    assert g[3][0] == 0 and g[0][0] == 1.0 and g[1][0] == 0.25 and g[2][0] == 0.25
    assert g[3][1] == -1 and gocc.tolist() == [True, False]
    sb = SceneBuilder()
    flat0 = sb.build()
    gs = gpu_ctx.scene(flat0)
    try:
        t, b1, b2, prim = gs.intersect(r)
        assert (prim == -1).all()
        assert not gs.intersects(r).any()
    finally:
        gs.close()


def test_invalid_scene_rejected(gpu_ctx):
    sb = SceneBuilder()
    sb.mesh(np.zeros((3, 3), np.float32), np.array([[0, 1, 2]]))
    flat = sb.build()
    flat.indices[0, 1] = 7  # out of bounds after validation in Python
    with pytest.raises(abi.BgpuError) as e:
        gpu_ctx.scene(flat)
    assert e.value.code == abi.E_INVALID
    flat.indices[0, 1] = -1
    with pytest.raises(abi.BgpuError):
        gpu_ctx.scene(flat)


def test_displaced_icosphere_medium(gpu_ctx):
    flat = scenes.icosphere_scene(levels=6, num_spheres=2000, res=256)
    o = OracleScene(flat)
    lo, hi = o.bounds()
    inc = scenes.incoherent_rays(lo, hi, 200000, seed=2)
    coh = scenes.primary_rays(flat)
    gs = gpu_ctx.scene(flat)
    try:
        for label, rays in (("incoherent", inc), ("coherent", coh)):
            g = gs.intersect(rays)
            oc = o.intersect(rays)
            assert_same(g, oc, "icosphere6/" + label)
            assert (gs.intersects(rays) == o.intersects(rays)).all()
    finally:
        gs.close()
        o.close()


@pytest.mark.parametrize("case", __import__("tests.cases", fromlist=["x"]).intersection_cases(),
                         ids=lambda c: c["name"])
def test_closed_form_cases_on_gpu(gpu_ctx, case):
    from tests import cases
    gs = gpu_ctx.scene(cases.case_scene(case))
    try:
        t, b1, b2, prim = gs.intersect(cases.case_ray(case))
        occ = gs.intersects(cases.case_ray(case))
    finally:
        gs.close()
    exp = case["expect"]
    if exp is None:
        assert prim[0] == -1
    else:
        assert prim[0] == 0 and t[0] == np.float32(exp["t"])
        if "b1" in exp:
            assert (b1[0], b2[0]) == (np.float32(exp["b1"]), np.float32(exp["b2"]))
    if "any" in case:
        assert occ[0] == case["any"]


@pytest.mark.parametrize("persistent", [0, 1])
@pytest.mark.parametrize("width", [2, 4, 8])
def test_bvh_widths_bit_exact(gpu_ctx, width, persistent):
    """Every BVH width and both traversal modes (one ray per thread; lanes
    refilled with new rays) give the oracle's results bit for bit (the
    structure changes the work, never the answer): a mixed triangle + shape
    scene under incoherent rays with random intervals, and a displaced
    icosphere with coherent and incoherent rays."""
    prev = gpu_ctx.config_get("bvh.width"), gpu_ctx.config_get("trace.persistent")
    gpu_ctx.config_set("bvh.width", width)
    gpu_ctx.config_set("trace.persistent", persistent)
    try:
        flat = scenes.small_scene(num_tris=4000, num_spheres=500, seed=21)
        lo, hi = scenes.scene_bounds(flat)
        rays = scenes.incoherent_rays(lo - 0.3, hi + 0.3, 50000, seed=6)
        rng = np.random.default_rng(width)
        rays[:, 6] = rng.uniform(0.0, 0.2, len(rays)).astype(np.float32)
        rays[:, 7] = np.where(rng.random(len(rays)) < 0.5, np.inf, rng.uniform(0.5, 3.0, len(rays))).astype(np.float32)
        gs = gpu_ctx.scene(flat)
        info = gs.bvh_info()
        g, gocc = gs.intersect(rays), gs.intersects(rays)
        gs.close()
        assert info["width"] == width and info["stack_bound"] <= 64, info
        o = OracleScene(flat)
        assert_same(g, o.intersect(rays, brute=True), "small/w%d" % width)
        assert (gocc == o.intersects(rays, brute=True)).all()
        o.close()
        ico = scenes.icosphere_scene(levels=5, num_spheres=300, seed=3, res=128)
        lo, hi = scenes.scene_bounds(ico)
        batches = [scenes.primary_rays(ico), scenes.incoherent_rays(lo, hi, 30000, seed=8)]
        gs = gpu_ctx.scene(ico)
        res = [(gs.intersect(b), gs.intersects(b)) for b in batches]
        gs.close()
        o = OracleScene(ico)
        for (g, gocc), b, name in zip(res, batches, ("coherent", "incoherent")):
            assert_same(g, o.intersect(b), "ico/%s/w%d" % (name, width))
            assert (gocc == o.intersects(b)).all()
        o.close()
    finally:
        gpu_ctx.config_set("bvh.width", prev[0])
        gpu_ctx.config_set("trace.persistent", prev[1])


def test_stack_depth_same_results(gpu_ctx):
    """trace.stack only moves the boundary between the LDS and the private
    part of the traversal stack: every depth gives the default's results bit
    for bit, on a binary tree deep enough that depth 8 spills, through both
    kernels (lane refill and one ray per thread).  trace.refill and
    trace.walk only schedule the traversal: same results too."""
    import ctypes as C
    keys = ("bvh.width", "trace.persistent", "trace.stack", "trace.refill", "trace.walk")
    keep = {k: gpu_ctx.config_get(k) for k in keys}
    try:
        gpu_ctx.config_set("bvh.width", 2)
        flat = scenes.icosphere_scene(levels=5, num_spheres=2000, seed=5, res=64)
        gs = gpu_ctx.scene(flat)
        try:
            assert gs.bvh_info()["stack_bound"] > 8
            rays = scenes.incoherent_rays(*scenes.scene_bounds(flat), 30000, seed=12)
            for persistent in (1, 0):
                gpu_ctx.config_set("trace.persistent", persistent)
                gpu_ctx.config_set("trace.stack", 12)
                base, occ = gs.intersect(rays), gs.intersects(rays)
                assert (base[3] >= 0).sum() > 3000
                for depth in (8, 16, 24):
                    gpu_ctx.config_set("trace.stack", depth)
                    assert gpu_ctx.config_get("trace.stack") == depth
                    got = gs.intersect(rays)
                    for a, b in zip(base, got):
                        assert np.array_equal(a.view(np.int32), b.view(np.int32)), (persistent, depth)
                    assert np.array_equal(occ, gs.intersects(rays)), (persistent, depth)
            # the scheduling knobs: same results (refill: lane-refill kernel
            # only; walk: both kernels)
            gpu_ctx.config_set("trace.stack", 12)
            for persistent, refill, walk in ((1, 1, 0), (1, 64, 0), (1, 32, 8), (1, 16, 32), (1, 32, 64), (1, 1, 1),
                                             (0, 32, 8), (0, 32, 16), (0, 32, 64), (0, 32, 1)):
                gpu_ctx.config_set("trace.persistent", persistent)
                gpu_ctx.config_set("trace.refill", refill)
                gpu_ctx.config_set("trace.walk", walk)
                got = gs.intersect(rays)
                for a, b in zip(base, got):
                    assert np.array_equal(a.view(np.int32), b.view(np.int32)), (persistent, refill, walk)
                assert np.array_equal(occ, gs.intersects(rays)), (persistent, refill, walk)
        finally:
            gs.close()
        for key, bad in (("trace.stack", 7), ("trace.stack", 25), ("trace.refill", 0), ("trace.refill", 65),
                         ("trace.walk", -1), ("trace.walk", 65)):
            with pytest.raises(abi.BgpuError) as e:
                gpu_ctx.config_set(key, bad)
            assert e.value.code == abi.E_INVALID
    finally:
        for k, v in keep.items():
            gpu_ctx.config_set(k, v)


def test_coherent_hint_same_results(gpu_ctx):
    """BGPU_COHERENT only chooses how lanes take rays: the results equal the
    default launch bit for bit."""
    import ctypes as C
    flat = scenes.icosphere_scene(levels=4, num_spheres=300, seed=9, res=96)
    gs = gpu_ctx.scene(flat)
    try:
        for rays in (scenes.primary_rays(flat),
                     scenes.incoherent_rays(*scenes.scene_bounds(flat), 20000, seed=3)):
            base = gs.intersect(rays)
            occ = gs.intersects(rays)
            n = len(rays)
            cols = [np.ascontiguousarray(rays[:, k]) for k in range(8)]
            out = [np.empty(n, np.float32) for _ in range(3)] + [np.empty(n, np.int32)]
            r = abi.Rays(*[c.ctypes.data for c in cols])
            h = abi.Hits(*[o.ctypes.data for o in out])
            abi.check(gpu_ctx.lib.bgpu_intersect_closest(gs.h, C.byref(r), n, C.byref(h), abi.COHERENT))
            occ2 = np.empty(n, np.uint8)
            abi.check(gpu_ctx.lib.bgpu_intersect_any(gs.h, C.byref(r), n, occ2.ctypes.data, abi.COHERENT))
            for a, b in zip(base, out):
                assert np.array_equal(a.view(np.int32), b.view(np.int32))
            assert np.array_equal(occ, occ2.astype(bool))
    finally:
        gs.close()


def test_autotuner_picks_a_valid_config(gpu_ctx):
    """The tuner times every candidate, leaves the fastest set on the
    context, and the tuned context still answers bit-exactly."""
    from bling_amd import tune
    keys = ("bvh.width", "trace.persistent", "trace.block", "trace.stack", "trace.walk")
    keep = {k: gpu_ctx.config_get(k) for k in keys}
    try:
        flat = scenes.small_scene(num_tris=3000, num_spheres=300, seed=17)
        rays = scenes.incoherent_rays(*scenes.scene_bounds(flat), 20000, seed=7)
        best, table = tune.tune_trace(gpu_ctx, flat, rays, widths=(2, 4), blocks=(64, 128), modes=(1, 0),
                                      stacks=(8, 16), walks=(0, 16), repeats=1)
        assert len(table) == 32 and all(row["ms"] > 0 for row in table)
        assert best == min(table, key=lambda r: r["ms"])
        for k in keys:
            assert gpu_ctx.config_get(k) == best[k]
        gs = gpu_ctx.scene(flat)
        g = gs.intersect(rays)
        gs.close()
        o = OracleScene(flat)
        assert_same(g, o.intersect(rays, brute=True), "tuned")
        o.close()
    finally:
        for k, v in keep.items():
            gpu_ctx.config_set(k, v)


@pytest.mark.parametrize("scaled", [False, True])
@pytest.mark.parametrize("width", [2, 4, 8])
def test_grazing_sphere_hits_bit_exact(gpu_ctx, width, scaled):
    """Rays that graze small spheres from far away, where the reference's
    binary32 sphere test reports hits outside the surface: the shape tree's
    grown boxes keep every such hit (plain and rotated, non-uniformly scaled
    spheres; every BVH width)."""
    from tests import cases
    flat, rays, _, _ = cases.grazing_spheres(n=40000, seed=23 + width, scaled=scaled)
    prev = gpu_ctx.config_get("bvh.width")
    gpu_ctx.config_set("bvh.width", width)
    try:
        g, o, gocc, oocc = run_both(gpu_ctx, flat, rays, brute=True)
    finally:
        gpu_ctx.config_set("bvh.width", prev)
    assert_same(g, o, "grazing/w%d/%s" % (width, "scaled" if scaled else "plain"))
    assert (gocc == oocc).all(), "any-hit mismatches: %d" % (gocc != oocc).sum()
    assert (g[3] >= 0).sum() > 1000


@pytest.mark.parametrize("coherent", [False, True])
def test_degenerate_rays_leave_valid_ones_exact(gpu_ctx, coherent):
    """Rays outside the bit-exact domain (zero, NaN or 1e+-30-scaled
    directions, NaN or infinite origins, empty intervals) mixed into a
    ragged batch: the reference's own answer for them depends on its
    kd-tree's visiting order (its binary32 quadric returns NaN or infinite
    roots), so only sanity is required of them; every valid ray in the same
    batch, with lanes refilled around the degenerate ones, stays bit-exact."""
    flat = scenes.small_scene(num_tris=1500, num_spheres=150, seed=12)
    lo, hi = scenes.scene_bounds(flat)
    n = 24007  # not a multiple of any block or slice size
    rays = scenes.incoherent_rays(lo, hi, n, seed=13)
    rng = np.random.default_rng(14)
    bad = rng.random(n) < 0.3
    kind = rng.integers(0, 7, n)
    for k, (sl, val) in enumerate(((slice(3, 6), 0.0), (slice(3, 4), np.nan), (slice(0, 1), np.nan),
                                   (slice(0, 3), np.inf), (None, None), (slice(3, 6), 1e-30),
                                   (slice(3, 6), 1e30))):
        m = bad & (kind == k)
        if k == 4:
            rays[m, 6], rays[m, 7] = 2.0, 1.0
        elif k in (5, 6):
            rays[m, sl] *= np.float32(val)
        else:
            rays[m, sl] = val
    gs = gpu_ctx.scene(flat)
    try:
        g = gs.intersect(rays, coherent=coherent)
        gocc = gs.intersects(rays, coherent=coherent)
    finally:
        gs.close()
    o = OracleScene(flat)
    oc = o.intersect(rays, brute=True)
    oocc = o.intersects(rays, brute=True)
    o.close()
    ok = ~bad
    assert_same(tuple(x[ok] for x in g), tuple(x[ok] for x in oc[:4]) + (0,), "valid rays among degenerate")
    assert (gocc[ok] == oocc[ok]).all()
    prim = g[3][bad]
    assert ((prim >= -1) & (prim < flat.num_prims)).all()
    empty = bad & (kind == 4)
    assert (g[3][empty] == -1).all() and not gocc[empty].any()
