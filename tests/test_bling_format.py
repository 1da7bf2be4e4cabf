"""The `.bling` reader (bling_amd/bling_format.py) against the grammar of
src/lib/Graphics/Bling/IO/*.hs and against scenes built directly.

- A Cornell box written in .bling (tests/golden/cornell_c1.bling) flattens
  to the arrays of scenes.cornell(256): vertices, indices and shape records
  bit for bit, materials per primitive, the camera value for value.
- Grammar pieces: numbers rounded like GHC's `read`, transform composition,
  polygon fans, spectra (rgbR/rgbI/spd/temp/hex), statement and light
  order, defaults of the start state, errors.
- The reference's own example scenes (read in place when the reference
  checkout is present): the eight the GPU path renders build, with the
  primitive counts their text describes; the others fail either as
  Unsupported or, where the reference's current grammar rejects them too,
  as syntax errors.
"""
import glob
import os

import numpy as np
import pytest

from bling_amd import abi, scenes
from bling_amd import bling_format as B
from bling_amd import transform as T
from bling_amd.scene import rgb_illum, rgb_refl, spd
from bling_amd.textures import Map2d, Map3d, Tex

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
REF_EXAMPLES = "/root/reference/examples"
F = np.float32


def _mat_of_prims(flat):
    """Material record (type, scalars, spectra) per primitive: triangles
    through their mesh, then shapes."""
    mats = [bytes(flat.materials[i]) for i in range(flat.desc.num_materials)]
    tri = [mats[flat.meshes[m].material] for m in flat.tri_mesh]
    return tri + [mats[flat.shapes[s].material] for s in range(flat.num_shapes)]


def test_cornell_fixture_flattens_like_the_procedural_scene():
    b = B.load(os.path.join(GOLDEN, "cornell_c1.bling"))
    a = scenes.cornell(256)
    f = b.flat
    assert np.array_equal(a.vertices.view(np.int32), f.vertices.view(np.int32))
    assert np.array_equal(a.indices, f.indices)
    assert a.num_shapes == f.num_shapes == 1
    sa, sf = a.shapes[0], f.shapes[0]
    for name, _ in abi.Shape._fields_:
        if name != "material":
            assert bytes(getattr(sa, name)) == bytes(getattr(sf, name)) if hasattr(getattr(sa, name), "_length_") \
                else getattr(sa, name) == getattr(sf, name), name
    assert _mat_of_prims(a) == _mat_of_prims(f)
    # the reference composes every transform block with the current one, so
    # the camera matrix may carry -0.0 where the direct build has 0.0
    for name in ("c2w", "r2c"):
        assert list(getattr(a.desc.camera, name)) == list(getattr(f.desc.camera, name))
    assert bytes(a.desc.film) == bytes(f.desc.film)
    s = b.spec()
    assert (s.spp_nu, s.spp_nv, s.max_depth, s.sample_depth, s.rr_start_depth) == (4, 4, 5, 3, 7)
    v = f.view(128, 96)
    assert v.size == (128, 96)


def test_numbers_round_like_ghc_read():
    # decimal -> binary32 in one rounding; a double-rounding case
    # (1 + 2^-24 + 2^-60: nearest double is the midpoint, which rounds to even)
    assert B.f32("1.0000000596046447753906250000000000000000000000000000000000000867") == F(1.0000001)
    assert B.f32("0.1") == F(0.1)
    p = B._Parser("imageSize 320 200 filter mitchell 2 2.5 0.333333 0.333333")
    p.run()
    assert p.res == (320, 200)
    assert p.filter == (abi.FILTER_MITCHELL, (F(2), F(2.5), F(0.333333), F(0.333333)))


def test_transform_blocks_compose_like_the_reference():
    p = B._Parser("transform { rotateX 30 translate 1 2 3 } transform { scale 2 2 2 }")
    p.run()
    want = T.scale((2, 2, 2)) @ (T.rotate_x(30) @ T.translate((1, 2, 3)) @ T.Transform())
    assert p.transform.m == want.m and p.transform.inv == want.inv
    p = B._Parser("transform { rotateX 30 } newTransform { translate 1 0 0 }")
    p.run()
    assert p.transform.m == (T.translate((1, 0, 0)) @ T.Transform()).m
    p = B._Parser("transform { matrix { m 1 0 0 4 m 0 1 0 5 m 0 0 1 6 m 0 0 0 1 } }")
    p.run()
    assert p.transform.m == (T.from_matrix([[1, 0, 0, 4], [0, 1, 0, 5], [0, 0, 1, 6], [0, 0, 0, 1]])
                             @ T.Transform()).m


def test_polygon_faces_become_fans():
    b = B.parse("prim { mesh vertexCount 5 faceCount 2 v 0 0 0 v 1 0 0 v 1 1 0 v 0 1 0 v 2 2 0 "
                "f 0 1 2 3 4 f 4 3 2 }")
    assert b.flat.indices.tolist() == [[0, 1, 2], [0, 2, 3], [0, 3, 4], [4, 3, 2]]


def test_spectra():
    p = B._Parser("")
    p.toks = B._tokens("rgbR 0.5 0.25 1 rgbI %ff8000 spd { 500 1, 400 0.5 } temp 3000")
    assert np.array_equal(p.spectrum(), rgb_refl(0.5, 0.25, 1.0))
    assert np.array_equal(p.spectrum(), rgb_illum(F(1), F(F(128) / F(255)), F(0)))
    assert np.array_equal(p.spectrum(), spd([(400, 0.5), (500, 1)]))
    bb = p.spectrum()
    # Planck at 3000 K rises across 400-700 nm; band 0 is the mean of the
    # values at 400 and 418.75 nm
    assert np.all(np.diff(bb) > 0)
    pl = lambda w: 0.4e-9 * 3.74183e-16 * (w * 1e-9) ** -5 / (np.exp(1.4388e-2 / (w * 1e-9 * 3000)) - 1)
    assert abs(bb[0] / (0.5 * (pl(400) + pl(418.75))) - 1) < 1e-5
    assert abs(bb[15] / (0.5 * (pl(681.25) + pl(700))) - 1) < 1e-5


def test_statement_order_emission_and_light_ids():
    b = B.parse("""
        material { matte kd { constant rgbR 0.5 0.5 0.5 } sigma { constant 0.3 } }
        prim { shape { sphere radius 1 } }
        emission { rgbI 2 2 2 }
        prim { shape { cylinder radius 1 zmin 2 zmax -1 phiMax 400 } }
        emission { none }
        prim { shape { disk height 0.5 radius 0.2 innerRadius 1 phiMax 90 } }
        prim { mesh vertexCount 3 faceCount 1 v 0 0 0 v 1 0 0 v 0 1 0 f 0 1 2 }
        light { infinite { rotateX 90 } l { constant rgbI 1 1 1 } }
        light { infinite { } l { constant rgbI 3 3 3 } }
    """)
    f = b.flat
    kinds = [f.shapes[i].type for i in range(f.num_shapes)]
    assert kinds == [abi.SHAPE_DISK, abi.SHAPE_CYLINDER, abi.SHAPE_SPHERE]  # last statement first
    assert [f.shapes[i].light_id for i in range(3)] == [2, 1, 0]          # the parse-order counter
    assert [f.shapes[i].emissive for i in range(3)] == [0, 1, 0]
    cyl, disk = f.shapes[1], f.shapes[0]
    assert list(cyl.params)[:4] == [F(1), F(-1), F(2), T.radians(F(360))]  # mkCylinder orders z, clamps phi
    assert list(disk.params)[:4] == [F(0.5), F(1), F(0.2), T.radians(F(90))]
    m = f.materials[f.shapes[2].material]
    assert m.type == abi.MAT_MATTE and m.scalar[0] == F(0.3)
    assert f.num_triangles == 1
    assert [lm[0, 0] for lm in f.light_maps] == [F(rgb_illum(3, 3, 3)[0]), F(rgb_illum(1, 1, 1)[0])]


def test_start_state_defaults():
    b = B.parse("prim { shape { sphere radius 1 } }")
    assert (b.width, b.height) == (640, 480)
    s = b.spec()
    assert (s.spp_nu, s.spp_nv, s.max_depth, s.sample_depth) == (2, 2, 7, 3)
    c2w, r2c = T.perspective_camera(T.translate((0, 0, -5)), 0, 1, 90, 640, 480)
    assert list(b.flat.desc.camera.c2w) == c2w.m and list(b.flat.desc.camera.r2c) == r2c.m
    m = b.flat.materials[0]
    assert m.type == abi.MAT_MATTE and list(m.spec[0]) == list(rgb_refl(0.9, 0.9, 0.9))


def test_errors():
    with pytest.raises(B.BlingSyntaxError):
        B.parse("renderer { sampler sampled { sampler { stratified xSamples 2 ySamples 2 } } }")
    with pytest.raises(B.BlingSyntaxError):
        B.parse("prim { shape { torus radius 1 } }")
    with pytest.raises(B.BlingSyntaxError):
        B.parse("imageSize 10")
    with pytest.raises(B.Unsupported):
        B.parse("prim { mandelbulb order 8 epsilon 0.001 iterations 20 }")
    with pytest.raises(B.BlingSyntaxError):
        B.parse("material { matte kd { constant rgbR 1 1 1 } sigma { cellNoise taxicab map { identity { } } } }")
    with pytest.raises(B.BlingSyntaxError):
        B.parse("material { matte kd { gradient f { constant 0 } steps { } } sigma { constant 0 } }")
    with pytest.raises(B.BlingSyntaxError):  # graphPaper without its mapping (an older grammar)
        B.parse("material { matte kd { graphPaper 0.1 tex1 { constant rgbR 1 1 1 } "
                "tex2 { constant rgbR 0 0 0 } } sigma { constant 0 } }")
    b = B.parse("imageSize 32 16 transform { translate 0 1 0 } camera { environment } imageSize 64 32 "
                "prim { shape { sphere radius 1 } }")  # the raster size in effect when the camera is parsed
    cam = b.flat.desc.camera
    assert (cam.type, cam.xres, cam.yres) == (abi.CAMERA_ENVIRONMENT, 32.0, 16.0)
    assert (b.flat.desc.film.width, b.flat.desc.film.height) == (64, 32)
    with pytest.raises(ValueError, match="out of bounds"):
        B.parse("prim { mesh vertexCount 3 faceCount 1 v 0 0 0 v 1 0 0 v 0 1 0 f 0 1 5 }")
    b = B.parse("renderer { sppm photonCount 100 maxDepth 5 radius 1 } prim { shape { sphere radius 1 } }")
    with pytest.raises(B.Unsupported):
        b.spec()


def test_selector_textures():
    """checker / graphPaper over constants (MaterialParser.hs:207-241,
    159-178) become the material's selector texture; blend of constants by
    a constant folds to a constant like spectrumBlend (Texture.hs)."""
    hdr = "prim { shape { sphere radius 1 } } "
    t = B.parse("material { transMatte kr { constant rgbR 0.2 0.2 0.2 } kt { constant rgbR 0.5 0.5 0.5 } "
                "ks { constant 0.3 } } " + hdr).flat.materials[0]
    assert (t.type, t.scalar[0]) == (abi.MAT_TRANSMATTE, F(0.3))
    assert list(t.spec[1]) == list(rgb_refl(0.5, 0.5, 0.5))
    b = B.parse("material { matte kd { checker 1 2 3 tex1 { constant rgbR 0.5 0.5 0.5 } "
                "tex2 { constant rgbR 0.1 0.2 0.3 } } sigma { constant 0 } } " + hdr)
    m = b.flat.materials[0]
    assert (m.tex.kind, m.tex.mapping, m.tex.slots) == (abi.TEX_CHECKER, abi.MAP_UV, 1)
    assert list(m.tex.params)[:3] == [1.0, 2.0, 3.0]
    assert list(m.spec[0]) == list(rgb_refl(0.5, 0.5, 0.5))
    assert list(m.tex.alt[0]) == list(rgb_refl(0.1, 0.2, 0.3))
    b = B.parse("material { matte kd { graphPaper 0.05 map { planar 1 0 0 0 0 1 0.5 0.25 } "
                "tex1 { constant rgbR 0.6 0.6 0.6 } tex2 { constant rgbR 0.02 0.02 0.02 } } "
                "sigma { constant 0 } } " + hdr)
    m = b.flat.materials[0]
    assert (m.tex.kind, m.tex.mapping) == (abi.TEX_GRAPHPAPER, abi.MAP_PLANAR)
    assert list(m.tex.params)[:9] == [F(0.05), 1, 0, 0, 0, 0, 1, 0.5, 0.25]
    # both plastic spectra under one selector
    b = B.parse("material { plastic kd { graphPaper 0.1 map { uv 4 4 0 0 } tex1 { constant rgbR 1 1 1 } "
                "tex2 { constant rgbR 0 0 0 } } ks { graphPaper 0.1 map { uv 4 4 0 0 } "
                "tex1 { constant rgbR 0.5 0.5 0.5 } tex2 { constant rgbR 0.2 0.2 0.2 } } "
                "rough { constant 0.1 } } " + hdr)
    m = b.flat.materials[0]
    assert m.tex.slots == 3 and list(m.tex.alt[1]) == list(rgb_refl(0.2, 0.2, 0.2))
    # blend: (1 - f) tex1 + f tex2 per band; f clamps the ends
    for f, want in ((0.25, [F(F(F(0.8) * F(0.75)) + F(F(0.4) * F(0.25)))]), (1.5, [F(0.4)]), (-1, [F(0.8)])):
        b = B.parse("material { matte kd { blend tex1 { constant spd { 400 0.8, 700 0.8 } } "
                    "tex2 { constant spd { 400 0.4, 700 0.4 } } f { constant %s } } "
                    "sigma { constant 0 } } " % f + hdr)
        assert b.flat.materials[0].tex.kind == abi.TEX_NONE
        assert list(b.flat.materials[0].spec[0])[:3] == want * 3, f


EXPECT = {  # the examples the GPU path renders: (triangles, shapes, explicit lights, size)
    "cornell-box.bling": (30, 1, 0, (480, 480)),
    "cornell-box-specular.bling": (10, 3, 0, (480, 480)),
    "matte-test.bling": (2, 7, 0, (1280, 720)),
    "plastic-test.bling": (2, 7, 0, (640, 360)),
    "blackbody-emission.bling": (0, 16, 0, (640, 400)),
    "glass-torus.bling": (0, 4, 0, (640, 360)),
    "cornell-box-underwater.bling": (79210, 3, 0, (720, 720)),
    "trans-matte.bling": (28, 2, 0, (480, 480)),
    "substrate.bling": (0, 5, 0, (1280, 720)),
    "heightmap.bling": (1682, 1, 0, (640, 360)),
    "bumpmap.bling": (2, 1, 1, (640, 360)),      # sun/sky light maps
    "cellnoise.bling": (0, 9, 1, (640, 360)),
    "race.bling": (0, 14, 1, (480, 320)),
    "ducky.bling": (13456, 1, 1, (640, 360)),    # a Wavefront mesh (see test_wavefront_meshes)
}


@pytest.mark.skipif(not os.path.isdir(REF_EXAMPLES), reason="reference checkout not present")
def test_reference_examples():
    seen = {}
    for path in sorted(glob.glob(os.path.join(REF_EXAMPLES, "*.bling"))):
        name = os.path.basename(path)
        try:
            b = B.load(path)
            seen[name] = (b.flat.num_triangles, b.flat.num_shapes, len(b.flat.light_maps), (b.width, b.height))
        except B.Unsupported:
            seen[name] = "unsupported"
        except B.BlingSyntaxError:
            seen[name] = "syntax"
    for name, want in EXPECT.items():
        assert seen[name] == want, (name, seen[name])
    assert sum(v == "unsupported" for v in seen.values()) >= 2  # julia, mandelbulb
    # race.bling: light { infinite { rotateX -90 } l { sunSky east 1 0 0 sunDir 1 0.2 0 turbidity 5 } }
    lt = B.load(os.path.join(REF_EXAMPLES, "race.bling")).flat.desc.lights[0]
    assert (lt.type, lt.map) == (abi.LIGHT_INFINITE, abi.LIGHTMAP_SUNSKY)
    assert list(lt.sky)[:7] == [1, 0, 0, 1, F(0.2), 0, 5]
    assert list(lt.w2l) == T.rotate_x(-90).m
    # cornell-box.bling: camera from its lookAt, film from its first filter
    b = B.load(os.path.join(REF_EXAMPLES, "cornell-box.bling"))
    c2w, _ = T.perspective_camera(T.look_at((278, 273, -800), (278, 273, 0), (0, 1, 0)) @ T.Transform(),
                                  0, 10, 37.5, 480, 480)
    assert list(b.flat.desc.camera.c2w) == c2w.m
    assert b.flat.desc.film.filter == abi.FILTER_MITCHELL
    with pytest.raises(B.Unsupported):  # its last renderer block is sppm
        b.spec()
    # blackbody-emission.bling renders with `directLighting maxDepth 5`
    s = B.load(os.path.join(REF_EXAMPLES, "blackbody-emission.bling")).spec()
    assert (s.integrator, s.max_depth) == ("direct", 5)
    # glass-torus.bling: the floor's graph paper (0.05, uv 15 15 0 0)
    b = B.load(os.path.join(REF_EXAMPLES, "glass-torus.bling"))
    tex = [b.flat.materials[i].tex for i in range(b.flat.desc.num_materials)]
    paper = [t for t in tex if t.kind == abi.TEX_GRAPHPAPER]
    assert len(paper) == 1 and paper[0].mapping == abi.MAP_UV
    assert list(paper[0].params)[:5] == [F(0.05), 15, 15, 0, 0]
    # substrate.bling: one sphere's depth is `scale 0.03 0.03 (fbm 3 0.5)` over identity { scale 1 1 1 }
    b = B.load(os.path.join(REF_EXAMPLES, "substrate.bling"))
    ops = [b.flat.tex_nodes[i].op for i in range(b.flat.desc.num_tex_nodes)]
    assert ops == [abi.TN_FBM, abi.TN_SCALE]
    m = [b.flat.materials[i] for i in range(b.flat.desc.num_materials) if b.flat.materials[i].tex.kind == abi.TEX_NODES]
    assert len(m) == 1 and m[0].type == abi.MAT_SUBSTRATE
    assert abi.texture_roots(m[0].tex)[4 + 2] == 1


def test_noise_lattice_and_gradients():
    """perlin3d (Texture.hs:358-397) is 0 on the integer lattice, and its
    gradient at a lattice point is the gradient vector its hash selects:
    with h = perm[perm[perm[i] + j] + k] & 15, the 12 + 4 set
    u = dx if h < 8 or h in (12, 13) else dy, v = dy if h < 4 or h in (12, 13)
    else dz, signs from bits 0 and 1 (the fade weights have zero slope there,
    so the corner's own term is the whole first-order change)."""
    from bling_amd import noise
    from bling_amd.noise_table import PERM
    rng = np.random.default_rng(1)
    ijk = rng.integers(-40, 300, (64, 3)).astype(np.float32)
    assert (noise.perlin3d(ijk[:, 0], ijk[:, 1], ijk[:, 2]) == 0).all()
    P = PERM + PERM
    eps = np.float32(1e-3)
    for i, j, k in ijk.astype(np.int64)[:24]:
        h = P[P[P[i & 255] + (j & 255)] + (k & 255)] & 15
        g = np.zeros(3)
        ua = 0 if (h < 8 or h in (12, 13)) else 1
        va = 1 if (h < 4 or h in (12, 13)) else 2
        g[ua] += -1.0 if h & 1 else 1.0
        g[va] += -1.0 if h & 2 else 1.0
        fd = []
        for a in range(3):
            d = np.zeros(3, np.float32)
            d[a] = eps
            p = np.float32([i, j, k]) + d
            fd.append(float(noise.perlin3d(p[0:1], p[1:2], p[2:3])[0]) / float(eps))
        np.testing.assert_allclose(fd, g, atol=2e-2)
    x = rng.uniform(-50, 50, (4096, 3)).astype(np.float32)
    v = noise.perlin3d(x[:, 0], x[:, 1], x[:, 2])
    assert np.abs(v).max() < 1.5 and v.std() > 0.1
    # fbm's first octave is the noise itself; later octaves add omega^i-weighted terms
    np.testing.assert_array_equal(noise.fbm(1, 0.5, x[:, 0], x[:, 1], x[:, 2]), v)
    assert not np.array_equal(noise.fbm(2, 0.5, x[:, 0], x[:, 1], x[:, 2]), v)


def test_heightmap_mesh_closed_forms():
    """heightMap (Primitive/Heightmap.hs:14-47): ns x nt lattice on [0, 1]^2
    with z outer, two triangles per cell, uv divided by ns - 1 on both axes,
    and normals -normalize(dx, ex + ez, dz) by central differences; for a
    plane elev = a x + c the normal is -(-2 a ex, ex + ez, 0) / len."""
    ns, nt, a = 5, 3, 0.5
    v, f, n, uv = B.heightmap_mesh(lambda x, z: F(a) * x + F(0.25), ns, nt)
    assert v.shape == (15, 3) and f.shape == (2 * 4 * 2, 3)
    np.testing.assert_array_equal(f[:2], [[0, 1, 6], [0, 6, 5]])
    np.testing.assert_array_equal(v[7], F([F(2) / F(4), F(a) * (F(2) / F(4)) + F(0.25), F(1) / F(2)]))
    np.testing.assert_array_equal(uv[-1], F([F(1) / F(4), F(1) / F(4)]))  # (1, 1) / (ns - 1)
    ex, ez = 1.0 / ns, 1.0 / nt
    want = -np.array([-2 * a * ex, ex + ez, 0.0])
    want /= np.linalg.norm(want)
    np.testing.assert_allclose(n, np.broadcast_to(want, n.shape), rtol=1e-6, atol=1e-7)
    b = B.parse("material { matte kd { constant rgbR 0.5 0.5 0.5 } sigma { constant 0 } } "
                "prim { heightMap 6 4 { scale 0 { fbm 0 octaves 2 omega 0.5 } } { scale 2 1 2 } }")
    f2 = b.flat
    assert f2.num_triangles == 5 * 3 * 2 and f2.normals is not None and f2.uvs is not None
    assert (f2.vertices[:, 1] == 0).all() and np.allclose(f2.normals, [0, -1, 0])


TEX_TEXT = {  # .bling texture text -> the same texture built directly
    "cell": ("cellNoise manhattan map { identity { scale 2 2 2 } }",
             lambda: Tex.cell("manhattan", Map3d.identity(T.scale((2.0, 2.0, 2.0))))),
    "fbm": ("fbm octaves 4 omega 0.6 map { identity { translate 1 0 0 rotateY 30 } }",
            lambda: Tex.fbm(4, 0.6, Map3d.identity(T.translate((1.0, 0.0, 0.0)) @ T.rotate_y(30.0)))),
    "perlin": ("perlin map { identity { scale 3 3 3 } }",
               lambda: Tex.perlin(Map3d.identity(T.scale((3.0, 3.0, 3.0))))),
    "crystal": ("crystal octaves 5 map { planar 1 0 0 0 0 1 0.5 0.25 }",
                lambda: Tex.crystal(5, Map2d.planar((1.0, 0, 0), (0, 0, 1.0), 0.5, 0.25))),
    "scale": ("scale 0.1 0.3 tex { perlin map { identity { } } }",
              lambda: Tex.scale(0.1, 0.3, Tex.perlin(Map3d.identity()))),
}


@pytest.mark.parametrize("name", sorted(TEX_TEXT))
def test_scalar_textures_parse_to_their_trees(name):
    """Each scalar texture (MaterialParser.hs:115-157), as a matte's sigma,
    evaluates like the same tree built directly, at seeded points."""
    text, make = TEX_TEXT[name]
    b = B.parse("material { matte kd { constant rgbR 0.5 0.5 0.5 } sigma { %s } } "
                "prim { shape { sphere radius 1 } }" % text)
    m = b.flat.materials[0]
    assert m.tex.kind == abi.TEX_NODES
    root = abi.texture_roots(m.tex)[4]
    want = _direct(make())
    xyz = np.random.default_rng(1).uniform(-3, 3, (256, 3)).astype(F)
    uv = np.random.default_rng(2).uniform(-1, 1, (256, 2)).astype(F)
    assert np.array_equal(_eval(b.flat, root, xyz, uv), _eval(want[0], want[1], xyz, uv))


def test_spectrum_textures_parse_to_their_trees():
    """blend / checker / gradient / graphPaper over non-constant textures
    (MaterialParser.hs:189-228) become node textures: gradient steps are
    sorted by position, a constant blend or scale folds."""
    text = ("gradient f { cellNoise euclidian map { identity { scale 2 2 2 } } } "
            "steps { 0.9 rgbR 0.9 0.1 0.1, 0.1 rgbR 0.1 0.1 0.9,0.5 rgbR 0.2 0.8 0.2 }")
    b = B.parse("material { plastic kd { %s } ks { checker 1 2 3 tex1 { constant rgbR 1 1 1 } "
                "tex2 { blend tex1 { constant rgbR 0 0 0 } tex2 { constant rgbR 1 0 0 } "
                "f { perlin map { identity { } } } } } rough { scale 0.1 0.5 tex { constant 0.2 } } } "
                "prim { shape { sphere radius 1 } }" % text)
    m = b.flat.materials[0]
    assert m.tex.kind == abi.TEX_NODES and m.scalar[0] == F(F(0.1) + F(F(0.5) * F(0.2)))
    roots = abi.texture_roots(m.tex)
    grad = b.flat.tex_nodes[roots[0]]
    assert grad.op == abi.TN_GRADIENT and grad.c == 3
    assert [b.flat.tex_nodes[grad.b + k].q[0] for k in range(3)] == [F(0.1), F(0.5), F(0.9)]
    assert b.flat.tex_nodes[roots[1]].op == abi.TN_CHECKER and roots[4] < 0
    steps = [(0.9, rgb_refl(0.9, 0.1, 0.1)), (0.1, rgb_refl(0.1, 0.1, 0.9)), (0.5, rgb_refl(0.2, 0.8, 0.2))]
    want = _direct(Tex.gradient(Tex.cell("euclidian", Map3d.identity(T.scale((2.0, 2.0, 2.0)))), steps), spec=True)
    xyz = np.random.default_rng(3).uniform(-3, 3, (256, 3)).astype(F)
    uv = np.zeros((256, 2), F)
    assert np.array_equal(_eval(b.flat, roots[0], xyz, uv), _eval(want[0], want[1], xyz, uv))
    # two spectrum slots under different selectors: node textures too
    b = B.parse("material { plastic kd { checker 1 1 1 tex1 { constant rgbR 1 1 1 } "
                "tex2 { constant rgbR 0 0 0 } } ks { checker 2 1 1 tex1 { constant rgbR 1 1 1 } "
                "tex2 { constant rgbR 0 0 0 } } rough { constant 0.1 } } prim { shape { sphere radius 1 } }")
    assert b.flat.materials[0].tex.kind == abi.TEX_NODES


def _direct(t, spec=None):
    from bling_amd.scene import Material, SceneBuilder, const
    sb = SceneBuilder()
    m = Material.matte(const(0.5))
    sb.material(m.textured_inputs(spec={0: t}) if not t.scalar else m.textured_inputs(scalar={0: t}))
    sb.shape(abi.SHAPE_SPHERE, (1.0,))
    flat = sb.build()
    return flat, abi.texture_roots(flat.materials[0].tex)[0 if not t.scalar else 4]


def _eval(flat, root, xyz, uv):
    from oracle import pyoracle
    o = pyoracle.OracleScene(flat)
    try:
        return o.tex_eval(root, xyz, uv)
    finally:
        o.close()


def test_image_textures(tmp_path):
    """`image { file "x" map {...} }` (MaterialParser.hs:106-113, 189-196):
    the file is read next to the scene file; RGB and RGBA (alpha dropped)
    8-bit images give pixelSpectrum texels, 8-bit grey images x / 255; an
    image of another kind, or a missing file, is an error."""
    Image = pytest.importorskip("PIL.Image")
    rng = np.random.default_rng(4)
    rgb = rng.integers(0, 256, (5, 7, 3)).astype(np.uint8)
    rgba = np.concatenate([rgb, rng.integers(0, 256, (5, 7, 1)).astype(np.uint8)], 2)
    grey = rng.integers(0, 256, (3, 4)).astype(np.uint8)
    Image.fromarray(rgb, "RGB").save(tmp_path / "c.png")
    Image.fromarray(rgba, "RGBA").save(tmp_path / "ca.png")
    Image.fromarray(grey, "L").save(tmp_path / "g.png")
    (tmp_path / "sub").mkdir()
    scene = ('material { matte kd { image { file "%s" map { uv 1 1 0 0 } } } '
             'sigma { image { file "g.png" map { planar 1 0 0 0 1 0 0 0 } } } } prim { shape { sphere radius 1 } }')
    from bling_amd.textures import image_spectra
    for name in ("c.png", "ca.png"):
        path = tmp_path / "sub" / "s.bling"
        path.write_text((scene % ("../" + name)).replace('"g.png"', '"../g.png"'))
        b = B.load(str(path))
        d = b.flat.desc
        assert d.num_images == 2
        im = d.images[0]
        assert (im.width, im.height, im.channels) == (7, 5, 16)
        tex = np.ctypeslib.as_array(im.texels, (5 * 7 * 16,)).reshape(5, 7, 16)
        assert np.array_equal(tex, image_spectra(rgb))
        g = d.images[1]
        assert (g.width, g.height, g.channels) == (4, 3, 1)
        assert np.array_equal(np.ctypeslib.as_array(g.texels, (12,)).reshape(3, 4), grey.astype(F) / F(255))
    with pytest.raises(B.BlingSyntaxError, match="unsupported image type"):
        B.parse(scene % "g.png", base=str(tmp_path))       # a grey image as a spectrum texture
    with pytest.raises(B.BlingSyntaxError, match="unsupported image type"):
        B.parse((scene % "c.png").replace('"g.png"', '"c.png"'), base=str(tmp_path))
    with pytest.raises(B.BlingSyntaxError, match="cannot read"):
        B.parse(scene % "missing.png", base=str(tmp_path))


def test_bump_map_materials():
    """`bumpMap bump {d} <material>` (MaterialParser.hs:44-48) bump-maps the
    material it wraps: the displacement's root in slot 8, the wrapped
    material's textures (a checker over constants included) as node
    textures; a bump map over a bump-mapped material is refused."""
    b = B.parse("material { bumpMap bump { scale 0.2 0.2 tex { cellNoise euclidian map { identity { scale 2 2 2 } } } } "
                "plastic kd { checker 1 1 1 tex1 { constant rgbR 1 1 1 } tex2 { constant rgbR 0 0 0 } } "
                "ks { constant rgbR 0.5 0.5 0.5 } rough { constant 0.1 } } prim { shape { sphere radius 1 } }")
    m = b.flat.materials[0]
    roots = abi.texture_roots(m.tex)
    assert m.tex.kind == abi.TEX_NODES and m.tex.slots == (1 | 256) and m.type == abi.MAT_PLASTIC
    assert b.flat.tex_nodes[roots[0]].op == abi.TN_CHECKER and b.flat.tex_nodes[roots[8]].op == abi.TN_SCALE
    assert m.scalar[0] == F(0.1)
    b = B.parse("material { bumpMap bump { constant 0.5 } matte kd { constant rgbR 0.5 0.5 0.5 } "
                "sigma { constant 0 } } prim { shape { sphere radius 1 } }")
    assert b.flat.materials[0].tex.slots == 256
    with pytest.raises(B.Unsupported):
        B.parse("material { bumpMap bump { constant 0.5 } bumpMap bump { constant 0.1 } matte "
                "kd { constant rgbR 0.5 0.5 0.5 } sigma { constant 0 } } prim { shape { sphere radius 1 } }")
    assert B.parse("material { bumpMap bump { constant 0.5 } blackbody } "
                   "prim { shape { sphere radius 1 } }").flat.materials[0].type == abi.MAT_BLACK


def test_shiny_metal_material():
    """`shinyMetal kr {..} ks {..} rough {..}` (MaterialParser.hs:74-79)."""
    b = B.parse("material { shinyMetal kr { constant rgbR 0.9 0.6 0.3 } ks { constant rgbR 0.5 0.5 0.5 } "
                "rough { constant 0.05 } } prim { shape { sphere radius 1 } }")
    m = b.flat.materials[0]
    assert (m.type, m.scalar[0]) == (abi.MAT_SHINYMETAL, F(0.05))
    assert list(m.spec[0]) == list(rgb_refl(0.9, 0.6, 0.3)) and list(m.spec[1]) == list(rgb_refl(0.5, 0.5, 0.5))


def test_point_and_directional_lights():
    """`light { point intensity .. position .. }` and `light { directional
    intensity .. normal .. }` (LightParser.hs:35-45) become explicit delta
    lights, in the reference's light order (prepended, Scene.hs:37-43)."""
    b = B.parse("light { point intensity rgbI 1 2 3 position 0 4 0 } "
                "light { directional intensity rgbI 2 2 2 normal 0 2 1 } prim { shape { sphere radius 1 } }")
    d = b.flat.desc
    assert d.num_lights == 2
    first, second = d.lights[0], d.lights[1]
    assert (first.type, second.type) == (abi.LIGHT_DIRECTIONAL, abi.LIGHT_POINT)
    assert list(first.vec)[:3] == [0.0, 2.0, 1.0] and list(second.vec)[:3] == [0.0, 4.0, 0.0]
    got = np.ctypeslib.as_array(second.radiance, (16,))
    assert np.array_equal(got, rgb_illum(1, 2, 3))


def _rgbe_file(path, px, rle):
    """Writes (h, w, 4) RGBE bytes as a Radiance file, flat or in the
    per-channel run-length form."""
    h, w = px.shape[:2]
    out = bytearray(b"#?RADIANCE\nFORMAT=32-bit_rle_rgbe\n\n-Y %d +X %d\n" % (h, w))
    for y in range(h):
        if not rle:
            out += px[y].tobytes()
            continue
        out += bytes([2, 2, w >> 8, w & 255])
        for c in range(4):
            ch, x = px[y, :, c], 0
            while x < w:  # a run where values repeat, literals otherwise
                n = 1
                while x + n < w and n < 127 and ch[x + n] == ch[x]:
                    n += 1
                if n >= 3:
                    out += bytes([128 + n, ch[x]])
                else:
                    n = min(w - x, 128)
                    k = 1
                    while k < n and not (x + k + 2 < w and ch[x + k] == ch[x + k + 1] == ch[x + k + 2]):
                        k += 1
                    out += bytes([k]) + ch[x:x + k].tobytes()
                    n = k
                x += n
    with open(path, "wb") as f:
        f.write(bytes(out))


def test_hdr_light_maps(tmp_path):
    """`l { file "x.hdr" }` (IO/MaterialParser.hs:257-264, IO/Bitmap.hs:13-28):
    RGBE pixels decode to (r, g, b) 2^(e - 136) (0 when e = 0) and each
    becomes rgbToSpectrumIllum of its colour; flat and run-length files give
    the same map; non-RGBE files are refused like the reference's
    \"can't convert image format to texture\"."""
    rng = np.random.default_rng(5)
    h, w = 5, 12
    px = rng.integers(0, 256, size=(h, w, 4), dtype=np.uint8)
    px[:, :, 3] = rng.integers(120, 140, size=(h, w))
    px[1, 2:9] = px[1, 2]  # runs for the run-length form
    px[3, 4, 3] = 0        # e = 0: black whatever the mantissas
    want = px[..., :3].astype(np.float64) * np.exp2(px[..., 3:].astype(np.float64) - 136)
    want[px[..., 3] == 0] = 0
    for rle in (False, True):
        _rgbe_file(str(tmp_path / "sky.hdr"), px, rle)
        rgb = B.read_hdr(str(tmp_path / "sky.hdr"))
        assert rgb.dtype == np.float32 and np.array_equal(rgb, want.astype(np.float32))
        b = B.parse("light { infinite { rotateX -90 } l { file \"sky.hdr\" } }", base=str(tmp_path))
        lt = b.flat.desc.lights[0]
        assert (lt.type, lt.map, lt.width, lt.height) == (abi.LIGHT_INFINITE, abi.LIGHTMAP_TEXELS, w, h)
        m = b.flat.light_maps[0].reshape(h, w, 16)
        for (y, x) in ((0, 0), (1, 5), (3, 4), (4, 11)):
            assert np.array_equal(m[y, x], rgb_illum(*rgb[y, x]))
    (tmp_path / "flat.png").write_bytes(b"\x89PNG\r\n\x1a\n")
    with pytest.raises(B.BlingSyntaxError):
        B.parse("light { infinite { } l { file \"flat.png\" } }", base=str(tmp_path))
    with pytest.raises(B.BlingSyntaxError):
        B.parse("light { infinite { } l { file \"missing.hdr\" } }", base=str(tmp_path))


def test_sun_sky_light_map():
    """`l { sunSky east .. sunDir .. turbidity .. }` (IO/MaterialParser.hs:266-270)."""
    b = B.parse("light { infinite { rotateX -90 } l { sunSky east 0 0 1 sunDir 0.5 1 1 turbidity 5 } }")
    lt = b.flat.desc.lights[0]
    assert (lt.type, lt.map) == (abi.LIGHT_INFINITE, abi.LIGHTMAP_SUNSKY)
    assert list(lt.sky)[:7] == [0, 0, 1, 0.5, 1, 1, 5]
    assert not lt.radiance
    with pytest.raises(B.BlingSyntaxError):
        B.parse("light { infinite { } l { sunSky east 0 0 1 turbidity 5 } }")


OBJ = b"""# two quads, a triangle; materials by usemtl
v 0 0 0
v 1 0 0
v 1 1 0
v 0 1 0
v  0 0 1 1
vt 0.25 0.5
vt 0.75
vn 0 0 2
vn 1 0 0
f 1/1/1 2/2/1 3/1/2 4/2/2
usemtl red
f 1//2 2//1 5//2
f 2 3 5
usemtl blue
f 3/2 4/1 5/2
f 1 2 3
"""


def test_wavefront_meshes(tmp_path):
    """`waveFront "x.obj" materials {..}` (PrimitiveParser.hs:69-72,
    IO/WaveFront.hs): fans of the faces; the material intervals in corner
    units except the last, which ends at the triangle count; the third
    corner's uv and normal taken from the second; points through the
    transform, normals normalized and not transformed."""
    pts, nrm, uvs, corners, mtls = B.read_wavefront(OBJ)
    assert pts.shape == (5, 3) and list(uvs[1]) == [0.75, 1.0]
    assert np.array_equal(nrm, np.array([[0, 0, 1], [1, 0, 0]], np.float32))
    assert len(corners) == 3 * 6 and mtls == [("red", 6), ("blue", 12)]
    assert list(corners[3]) == [0, 0, 0] and list(corners[5]) == [3, 1, 1]  # fan (0, 2, 3) of the quad
    # intervals: default [0, 6) -> corners 0, 3; red [6, 12) -> 6, 9; blue [12, 6) (cnt = 6) -> none
    assert B.wavefront_triangles(corners, mtls) == [("default", 0), ("default", 3), ("red", 6), ("red", 9)]
    (tmp_path / "m.obj").write_bytes(OBJ)
    text = ("transform { translate 0 0 5 }\n"
            "material { matte kd { constant rgbR 0.5 0.5 0.5 } sigma { constant 0 } }\n"
            "prim { waveFront \"m.obj\" materials { \"red\" { matte kd { constant rgbR 0.9 0.1 0.1 } "
            "sigma { constant 0 } } } }")
    b = B.parse(text, base=str(tmp_path))
    fl = b.flat
    assert fl.num_triangles == 4 and fl.desc.num_meshes == 3
    v = fl.vertices.reshape(-1, 3, 3)
    assert np.array_equal(v[0], np.array([[0, 0, 5], [1, 0, 5], [1, 1, 5]], np.float32))
    # mesh 0: both default triangles with uvs and normals; the third corner repeats the second's
    assert list(fl.tri_mesh) == [0, 0, 1, 2]
    n = fl.normals.reshape(-1, 3, 3)
    assert np.array_equal(n[0], np.array([[0, 0, 1], [0, 0, 1], [0, 0, 1]], np.float32))
    assert np.array_equal(n[1], np.array([[0, 0, 1], [1, 0, 0], [1, 0, 0]], np.float32))
    assert np.array_equal(n[2], np.array([[1, 0, 0], [0, 0, 1], [0, 0, 1]], np.float32))
    u = fl.uvs.reshape(-1, 3, 2)
    assert np.array_equal(u[0], np.array([[0.25, 0.5], [0.75, 1.0], [0.75, 1.0]], np.float32))
    m = [fl.meshes[k] for k in range(fl.desc.num_meshes)]
    assert (m[0].flags, m[1].flags, m[2].flags) == (abi.MESH_NORMALS | abi.MESH_UVS, abi.MESH_NORMALS, 0)
    red = fl.materials[m[1].material]
    assert red.spec[0][0] == rgb_refl(0.9, 0.1, 0.1)[0] and m[2].material == m[1].material != m[0].material
    for bad in (b"v 1 2\n", b"v 1 2 3", b"f -1 2 3\n", b"vn 1e-3 0 1\n"):
        (tmp_path / "bad.obj").write_bytes(bad)
        with pytest.raises(B.BlingSyntaxError):
            B.parse("prim { waveFront \"bad.obj\" materials { } }", base=str(tmp_path))
