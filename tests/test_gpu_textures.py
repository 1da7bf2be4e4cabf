"""Node textures on the device (bling_amd/csrc/pt_texture.h, k_texture in
pt.hip) against the oracle (oracle/orc_texture.h), the restatement of
Texture.hs:59-414.

- bgpu_scene_texture_eval: every node of a scene's node list at seeded
  points and surface parameters equals the oracle's value bit for bit,
  except where a quasi-crystal contributes (cosf/sinf may differ by an ulp
  between the device's and the host's libm: within 2e-5 absolute);
- a render of a scene whose materials take node textures (every op, 2d and
  3d mappings, images, textured scalars) tracks the oracle with the same
  seed (equal sample weights, median relative pixel difference < 1e-4) and
  is within 3 sigma of it with another seed; so does a scene of bump-mapped
  materials (bumpMapped, Reflection.hs:344-378);
- the served path (continuous batching, textured and untextured jobs in
  flight together) gives the plain path's image;
- evaluating a node that is not a texture (out of range, a gradient step)
  is refused with an error.
"""
import numpy as np
import pytest

from bling_amd import abi
from bling_amd.runtime import GpuRenderer, Job, JobSpec
from bling_amd.scene import Material, SceneBuilder, const, rgb_refl
from bling_amd.textures import Map2d, Map3d, Tex
from bling_amd import transform as T
from oracle import pyoracle
from tests import cases
from tests.imagestats import compare, mean_se

pytestmark = pytest.mark.gpu
F = np.float32


def _all_ops_scene():
    """One material per texture kind (and per cell distance) over a triangle."""
    g = (np.arange(8 * 8).reshape(8, 8) * 4).astype(np.uint8)
    m3 = Map3d.identity(T.scale((1.7, 1.7, 1.7)))
    texs = [Tex.perlin(m3), Tex.fbm(6, 0.5, m3), Tex.crystal(7, Map2d.uv(3.0, 3.0, 0.1, 0.2)),
            Tex.image_grey(g, Map2d.planar((1.0, 0.0, 0.0), (0.0, 0.0, 1.0), 0.5, 0.0)),
            Tex.scale(0.5, 2.0, Tex.perlin(m3))]
    texs += [Tex.cell(d, m3) for d in ("euclidian", "euclidian2", "manhattan", "chebyshev")]
    red, blue = Tex.spectrum(rgb_refl(0.8, 0.1, 0.1)), Tex.spectrum(rgb_refl(0.1, 0.2, 0.9))
    rgb = np.random.default_rng(5).integers(0, 256, (6, 10, 3)).astype(np.uint8)
    from bling_amd.textures import image_spectra
    texs += [Tex.image(image_spectra(rgb), Map2d.uv(1.0, 1.0, 0.0, 0.0)), Tex.blend(red, blue, Tex.fbm(3, 0.6, m3)),
             Tex.checker((2.0, 3.0, 4.0), red, blue),
             Tex.gradient(Tex.cell("manhattan", m3), [(0.1, rgb_refl(0.1, 0.1, 0.1)), (0.5, rgb_refl(0.9, 0.1, 0.3)),
                                                      (0.9, rgb_refl(0.2, 0.9, 0.9))]),
             Tex.graph_paper(0.1, Map2d.uv(5.0, 5.0, 0.0, 0.0), red, blue)]
    sb = SceneBuilder()
    for t in texs:
        m = Material.matte(const(0.5))
        sb.material(m.textured_inputs(scalar={0: t}) if t.scalar else m.textured_inputs(spec={0: t}))
    sb.mesh(np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0]], F), np.array([[0, 1, 2]], np.int32))
    return sb.build()


def _crystal_fed(flat):
    """Per node: whether a quasi-crystal node is among its descendants."""
    out = []
    for i in range(flat.desc.num_tex_nodes):
        n = flat.tex_nodes[i]
        kids = []
        if n.op == abi.TN_SCALE or n.op == abi.TN_GRADIENT:
            kids = [n.a]
        elif n.op in (abi.TN_CHECKER, abi.TN_GRAPHPAPER):
            kids = [n.a, n.b]
        elif n.op == abi.TN_BLEND:
            kids = [n.a, n.b, n.c]
        out.append(n.op == abi.TN_CRYSTAL or any(out[k] for k in kids))
    return out


@pytest.mark.parametrize("which", ["all_ops", "scene"])
def test_texture_eval_matches_oracle(gpu_ctx, which):
    flat = _all_ops_scene() if which == "all_ops" else cases.node_texture_scene(res=16)
    rng = np.random.default_rng(21)
    n = 4096
    xyz = rng.uniform(-6, 6, (n, 3)).astype(F)
    xyz[:64] = np.round(xyz[:64])          # lattice points and cell borders
    xyz[64:128] = 0.0
    uv = rng.uniform(-2, 2, (n, 2)).astype(F)
    uv[:32] = 0.0
    fed = _crystal_fed(flat)
    gs = gpu_ctx.scene(flat)
    o = pyoracle.OracleScene(flat)
    try:
        checked = 0
        for node in range(flat.desc.num_tex_nodes):
            if flat.tex_nodes[node].op == abi.TN_GSTEP:
                continue
            g = gs.texture_eval(node, xyz, uv)
            want = o.tex_eval(node, xyz, uv)
            assert g.shape == want.shape, node
            if fed[node]:
                np.testing.assert_allclose(g, want, rtol=0, atol=2e-5, err_msg="node %d" % node)
            else:
                assert np.array_equal(g.view(np.int32), want.view(np.int32)), \
                    (node, flat.tex_nodes[node].op, np.flatnonzero(g != want)[:5])
            checked += 1
        assert checked >= 10
    finally:
        o.close()
        gs.close()


def test_texture_eval_rejects_bad_nodes(gpu_ctx):
    flat = cases.node_texture_scene(res=16)
    gs = gpu_ctx.scene(flat)
    try:
        xyz = np.zeros((4, 3), F)
        uv = np.zeros((4, 2), F)
        for bad in (-1, flat.desc.num_tex_nodes):
            with pytest.raises(Exception):
                gs.texture_eval(bad, xyz, uv)
        steps = [i for i in range(flat.desc.num_tex_nodes) if flat.tex_nodes[i].op == abi.TN_GSTEP]
        with pytest.raises(Exception):
            gs.texture_eval(steps[0], xyz, uv)
        assert gs.texture_eval(0, xyz[:0], uv[:0]).shape[0] == 0
    finally:
        gs.close()


def _render(ctx, flat, spec):
    gs = ctx.scene(flat)
    try:
        return GpuRenderer(spec).render_plain(gs)
    finally:
        gs.close()


def test_node_texture_scene_matches_oracle(gpu_ctx):
    flat = cases.node_texture_scene(res=48)
    spec = JobSpec(spp_nu=4, spp_nv=4, max_depth=6, sample_depth=3, rr_start_depth=3, passes=4,
                   seed=41, variance=True)
    g, gsq = _render(gpu_ctx, flat, spec)
    o, osq, _ = pyoracle.render(flat, spec)
    assert np.isfinite(g).all()
    assert np.array_equal(g[..., 0], o[..., 0]), "sample weights per pixel differ"
    frac, zmean, zmax = compare(g, gsq, o, osq)
    assert frac < 0.01, (frac, zmean, zmax)
    mg, _ = mean_se(g, gsq)
    mo, _ = mean_se(o, osq)
    rel = np.abs(mg - mo) / np.maximum(np.abs(mo), 1e-3)
    assert np.median(rel) < 1e-4, np.median(rel)
    spec2 = JobSpec(spp_nu=4, spp_nv=4, max_depth=6, sample_depth=3, rr_start_depth=3, passes=4,
                    seed=42, variance=True)
    o2, osq2, _ = pyoracle.render(flat, spec2)
    frac, zmean, zmax = compare(g, gsq, o2, osq2)
    assert frac < 0.015, (frac, zmean, zmax)
    assert abs(zmean) < 0.08, zmean


def test_bump_mapped_scene_matches_oracle(gpu_ctx):
    """Bump maps over cell noise, fBm, a quasi-crystal (uv mapping) and
    Perlin noise on spheres, a quad and a uv mesh, one material also
    node-textured: the same seed tracks the oracle, and the bumps change
    the image."""
    flat = cases.bump_scene(res=48)
    spec = JobSpec(spp_nu=4, spp_nv=4, max_depth=5, sample_depth=3, rr_start_depth=3, passes=2,
                   seed=44, variance=True)
    g, gsq = _render(gpu_ctx, flat, spec)
    o, osq, _ = pyoracle.render(flat, spec)
    assert np.isfinite(g).all()
    assert np.array_equal(g[..., 0], o[..., 0])
    frac, zmean, zmax = compare(g, gsq, o, osq)
    assert frac < 0.01, (frac, zmean, zmax)
    mg, _ = mean_se(g, gsq)
    mo, _ = mean_se(o, osq)
    rel = np.abs(mg - mo) / np.maximum(np.abs(mo), 1e-3)
    assert np.median(rel) < 1e-4, np.median(rel)


def test_node_texture_direct_lighting_matches_oracle(gpu_ctx):
    flat = cases.node_texture_scene(res=32)
    spec = JobSpec(spp_nu=2, spp_nv=2, max_depth=4, sample_depth=3, rr_start_depth=3, passes=2,
                   seed=43, variance=True, integrator="direct")
    g, gsq = _render(gpu_ctx, flat, spec)
    o, osq, _ = pyoracle.render(flat, spec)
    assert np.array_equal(g[..., 0], o[..., 0])
    mg, _ = mean_se(g, gsq)
    mo, _ = mean_se(o, osq)
    rel = np.abs(mg - mo) / np.maximum(np.abs(mo), 1e-3)
    assert np.median(rel) < 1e-4, np.median(rel)


def test_textured_and_plain_jobs_served_together(gpu_ctx):
    """Jobs on a node-textured scene and on an untextured one, in flight
    together on the serving stream, each give the plain path's image."""
    from bling_amd import scenes
    flats = [cases.node_texture_scene(res=32), scenes.cornell(32)]
    specs = [JobSpec(spp_nu=2, spp_nv=2, max_depth=4, sample_depth=3, passes=2, seed=s, variance=True)
             for s in (5, 6)]
    gss = [gpu_ctx.scene(f) for f in flats]
    try:
        plain = {(si, k): GpuRenderer(sp).render_plain(gs) for si, gs in enumerate(gss) for k, sp in enumerate(specs)}
        jobs = {(si, k): Job(gs, sp) for si, gs in enumerate(gss) for k, sp in enumerate(specs)}
        for key, j in jobs.items():
            j.wait(timeout=120)
            img, sq, done, state = j.poll()
            assert state == abi.JOB_DONE
            p_img, p_sq = plain[key]
            assert np.array_equal(img[..., 0], p_img[..., 0])
            np.testing.assert_allclose(img, p_img, rtol=1e-4, atol=1e-5)
            j.release()
    finally:
        for gs in gss:
            gs.close()
