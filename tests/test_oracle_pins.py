"""Pins of the oracle (CPU restatement of the reference) against what the
mathematics fixes: closed-form intersections, structure independence of the
nearest-hit definition, analytic radiance, sampler stratification, spectral
and transform identities.  The reference has no tests or golden vectors for
this path, so these are the checks that make the oracle trustworthy."""
import numpy as np
import pytest

from bling_amd import abi, scenes
from bling_amd import transform as T
from bling_amd.runtime import JobSpec
from bling_amd.scene import (Material, SceneBuilder, const, rgb_illum, rgb_refl, spd,
                             spectrum_to_xyz)
from bling_amd.spectral_tables import CIE_Y_SUM, TABLES
from oracle import pyoracle
from oracle.pyoracle import OracleScene
from tests import cases
from tests.imagestats import mean_se


@pytest.mark.parametrize("case", cases.intersection_cases(), ids=lambda c: c["name"])
def test_closed_form_intersections(case):
    o = OracleScene(cases.case_scene(case))
    t, b1, b2, prim, _ = o.intersect(cases.case_ray(case), brute=True)
    exp = case["expect"]
    if exp is None:
        assert prim[0] == -1
    else:
        assert prim[0] == 0
        assert t[0] == np.float32(exp["t"])
        if "b1" in exp:
            assert (b1[0], b2[0]) == (np.float32(exp["b1"]), np.float32(exp["b2"]))
    if "any" in case:
        assert o.intersects(cases.case_ray(case), brute=True)[0] == case["any"]
    elif exp is not None:
        assert o.intersects(cases.case_ray(case), brute=True)[0]


def test_nearest_hit_is_structure_independent():
    """The oracle's pruning BVH returns exactly the brute-force result."""
    flat = scenes.small_scene(num_tris=1500, num_spheres=300, seed=21)
    o = OracleScene(flat)
    lo, hi = o.bounds()
    rays = scenes.incoherent_rays(lo - 0.2, hi + 0.2, 8000, seed=6)
    a = o.intersect(rays, brute=True)
    b = o.intersect(rays, brute=False)
    for x, y in zip(a[:4], b[:4]):
        assert np.array_equal(x.view(np.int32), y.view(np.int32))
    assert np.array_equal(o.intersects(rays, brute=True), o.intersects(rays, brute=False))


def test_ties_go_to_the_lowest_id():
    sb = SceneBuilder()
    sb.material(Material.matte(const(0.5)))
    tri = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], np.float32)
    sb.mesh(np.concatenate([tri, tri]), np.array([[0, 1, 2], [3, 4, 5]]))
    o = OracleScene(sb.build())
    t, _, _, prim, ties = o.intersect(np.array([[0.25, 0.25, -1, 0, 0, 1, 0, np.inf]], np.float32), brute=True)
    assert prim[0] == 0 and t[0] == 1.0 and ties == 1


def test_sky_pixels_equal_the_sky_radiance():
    """Camera rays that escape add le of the infinite light (Path.hs:44-47)."""
    flat = cases.sky_only(sky=2.0)
    img, sq, _ = pyoracle.render(flat, JobSpec(spp_nu=2, spp_nv=2, max_depth=3, passes=1, seed=3))
    mean = img[..., 1:] / img[..., :1]
    xyz = np.array(spectrum_to_xyz(const(2.0)), np.float32)
    np.testing.assert_allclose(mean, np.broadcast_to(xyz, mean.shape), rtol=1e-6)


def _analytic(flat, expect_y, spp=8, passes=8, tol_sigma=4.0):
    spec = JobSpec(spp_nu=spp, spp_nv=spp, max_depth=1, sample_depth=1, passes=passes, seed=5, variance=True)
    img, sq, _ = pyoracle.render(flat, spec)
    m, se = mean_se(img, sq)
    y = m[..., 1].mean()
    s = np.sqrt((se[..., 1] ** 2).sum()) / m[..., 1].size
    assert abs(y - expect_y) < tol_sigma * s + 1e-4 * expect_y, (y, expect_y, s)
    return y, s


def test_diffuse_plane_under_uniform_sky():
    """NEE + MIS against the infinite light: Lo = rho * L for a Lambertian
    plane with the whole upper hemisphere open (max depth 1: direct only)."""
    _analytic(cases.plane_under_sky(rho=0.5, sky=1.0), 0.5)


def test_translucent_plane_under_uniform_sky():
    """translucentMatte (Material.hs:43-53): a Lambertian reflection of
    r = kr and the same lobe turned to transmission (brdfToBtdf) with
    t = kt (1 - kr); under a sky that surrounds the plane the direct radiance
    is (r + t) L, the transmitted part coming from below."""
    kr, kt = 0.3, 0.7
    m = Material.trans_matte(const(kr), const(kt), 0.0)
    _analytic(cases.plane_under_sky(sky=1.0, material=m), kr + kt * (1 - kr))


def test_sphere_light_over_plane():
    """Cone-sampled sphere light (Shape.hs:364-375) with MIS: at the point
    below the light, Lo = rho * Le * (R / D)^2."""
    R, D, Le, rho = 0.25, 2.0, 4.0, 0.5
    _analytic(cases.sphere_light_over_plane(rho=rho, Le=Le, R=R, D=D), rho * Le * (R / D) ** 2)


def test_material_scene_is_finite_and_energy_bounded():
    flat = cases.materials_scene(res=24)
    img, sq, rays = pyoracle.render(flat, JobSpec(spp_nu=2, spp_nv=2, max_depth=8, passes=1, seed=2))
    assert np.isfinite(img).all() and rays > 0
    assert (img[..., 0] == 4).all()


def test_sampler_stratification():
    """Per pixel the camera offsets fill the nu x nv strata, and every 1D/2D
    table dimension puts exactly one sample in each of the N strata
    (stratified1D / stratified2D + shuffle, Sampling.hs:134-171)."""
    import ctypes as C
    lib = pyoracle.load()
    if not hasattr(lib, "orc_sampler_table"):
        pytest.skip("sampler export missing")
    nu = nv = 4
    N = nu * nv
    out = np.zeros((N, 8), np.float32)
    lib.orc_sampler_table.argtypes = [C.c_uint64, C.c_int, C.c_int, C.c_int, C.c_int, C.c_void_p]
    lib.orc_sampler_table(7, 3, 5, nu, nv, out.ctypes.data)
    cam_x, cam_y, d1, u2, v2, raw1 = out[:, 0], out[:, 1], out[:, 2], out[:, 3], out[:, 4], out[:, 5]
    assert ((out[:, :6] >= 0) & (out[:, :6] < 1)).all()
    cells = set(zip((cam_x * nu).astype(int), (cam_y * nv).astype(int)))
    assert len(cells) == N
    assert sorted((d1 * N).astype(int).tolist()) == list(range(N))
    assert len(set(zip((u2 * nu).astype(int), (v2 * nv).astype(int)))) == N
    assert len(set(np.round(raw1, 6))) > N // 2  # beyond the tables: plain uniform


def test_sampler_primitives_known_answers():
    """mix64 is the SplitMix64 finaliser (S. Vigna): fed the generator's
    states 1, 2, 3 x 0x9e3779b97f4a7c15 (seed 0) it must return the
    generator's published first outputs.  permute(., n, p) must be a
    bijection of [0, n) for every key (it stands in for `shuffle`)."""
    import ctypes as C
    lib = pyoracle.load()
    lib.orc_mix64.restype = C.c_uint64
    lib.orc_mix64.argtypes = [C.c_uint64]
    lib.orc_permute.restype = C.c_uint32
    lib.orc_permute.argtypes = [C.c_uint32, C.c_uint32, C.c_uint32]
    gamma = 0x9E3779B97F4A7C15
    got = [lib.orc_mix64((k * gamma) & (2 ** 64 - 1)) for k in (1, 2, 3)]
    assert got == [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F]
    rng = np.random.default_rng(3)
    for n in (1, 2, 3, 7, 16, 100, 1000, 4097):
        for p in rng.integers(0, 2 ** 32, 4):
            img = [lib.orc_permute(i, n, int(p)) for i in range(n)]
            assert sorted(img) == list(range(n)), (n, p)


def test_spectral_tables():
    y = np.array(TABLES["cie_y"], np.float32)
    acc = np.float32(0)
    for v in y:
        acc = np.float32(acc + v)
    assert acc == np.float32(CIE_Y_SUM)
    # sY (sConst 1) == 1 and XYZ is linear in the spectrum
    X, Y, Z = spectrum_to_xyz(const(1.0))
    assert abs(Y - 1.0) < 1e-6
    X2, Y2, Z2 = spectrum_to_xyz(const(2.0))
    assert (X2, Y2, Z2) == (2 * X, 2 * Y, 2 * Z)
    # grey RGB maps to the scaled white basis (rgbToSpectrum, Spectrum.hs:146-159)
    np.testing.assert_array_equal(rgb_refl(0.5, 0.5, 0.5), np.float32(0.5) * np.array(TABLES["rgb_refl_white"], np.float32))
    np.testing.assert_array_equal(rgb_illum(2, 2, 2), np.float32(2) * np.array(TABLES["rgb_illum_white"], np.float32))
    # a single-point SPD is constant (avgSpd of an IrregularSpd)
    np.testing.assert_array_equal(spd([(1, 0.78)]), const(0.78))
    s = rgb_refl(0.63, 0.065, 0.05)
    assert s[-1] > 0.5 and s[0] < 0.2  # red reflects long wavelengths


def test_transforms():
    rng = np.random.default_rng(0)
    m = rng.normal(size=(4, 4)).astype(np.float32)
    inv = np.array(T.invert(m.reshape(-1).tolist()), np.float32).reshape(4, 4)
    np.testing.assert_allclose(inv @ m, np.eye(4), atol=2e-5)
    # a @ b applies a first (concatTrans, Transform.hs:241-245)
    ab = T.translate((1, 0, 0)) @ T.scale((2, 2, 2))
    assert ab.point((0, 0, 0)) == (2.0, 0.0, 0.0)
    rt = T.rotate_z(90) @ T.translate((5, 0, 0))
    p = rt.point((1, 0, 0))
    assert abs(p[0] - 5) < 1e-6 and abs(p[1] - 1) < 1e-6
    l = T.look_at((1, 2, 3), (1, 2, 10), (0, 1, 0))
    c = np.array(l.m, np.float32).reshape(4, 4)[:3, :3]
    np.testing.assert_allclose(c.T @ c, np.eye(3), atol=1e-6)
    np.testing.assert_allclose(c[:, 2], [0, 0, 1], atol=1e-7)
    # the ray through the image centre points at the look-at point
    flat = scenes.cornell(64)
    rays = scenes.primary_rays(flat)
    centre = rays[32 * 64 + 32, 3:6]
    want = np.array([0, 0, 1.0])
    assert np.degrees(np.arccos(np.clip(np.dot(centre / np.linalg.norm(centre), want), -1, 1))) < 0.5


def test_mesh_validation_matches_reference_errors():
    sb = SceneBuilder()
    with pytest.raises(ValueError, match="negative"):
        sb.mesh(np.zeros((3, 3), np.float32), np.array([[0, -1, 2]]))
    with pytest.raises(ValueError, match="out of bounds"):
        sb.mesh(np.zeros((3, 3), np.float32), np.array([[0, 1, 3]]))
    flat = SceneBuilder().build()
    flat.indices = np.array([[0, 1, 5]], np.int32)
    flat.desc.num_triangles = 1
    flat.desc.num_vertices = 3
    flat.vertices = np.zeros((3, 3), np.float32)
    import ctypes as C
    flat.desc.indices = flat.indices.ctypes.data_as(C.POINTER(C.c_int32))
    flat.desc.vertices = flat.vertices.ctypes.data_as(C.POINTER(C.c_float))
    with pytest.raises(ValueError, match="out of bounds"):
        OracleScene(flat)


def test_job_view_equals_scene_built_at_that_size():
    """A job's own camera and film (the RenderJob's imageSize) give exactly
    the image of a scene described at that size (one thread: same order)."""
    from bling_amd import scenes
    from bling_amd.runtime import JobSpec
    direct = scenes.cornell(24, height=16)
    other = scenes.cornell(48)
    spec = JobSpec(spp_nu=2, spp_nv=1, max_depth=3, sample_depth=2, rr_start_depth=2, seed=3)
    a, asq, _ = pyoracle.render(direct, spec, threads=1)
    spec_v = JobSpec(spp_nu=2, spp_nv=1, max_depth=3, sample_depth=2, rr_start_depth=2, seed=3,
                     view=other.view(24, 16))
    b, bsq, _ = pyoracle.render(other, spec_v, threads=1)
    assert a.shape == (16, 24, 4)
    assert np.array_equal(a, b) and np.array_equal(asq, bsq)
    assert a[..., 2].sum() > 0


def test_sphere_false_hits_stay_within_the_traversal_margin():
    """The reference's binary32 sphere test (Shape.hs:173-187) reports hits
    slightly outside the surface when the ray passes close to it from far
    away (cancellation in c = |o|^2 - r^2).  The GPU's shape-tree boxes are
    grown by min(K0 |o - c|, K0^2 |o - c|^2 / (2 r)) with K0 = 2e-3; every
    hit the oracle reports, false ones included, must lie within that of the
    sphere.  Rays graze spheres of radius 1e-4..0.05 from 1..200 units away,
    at line distances up to twice the bound."""
    k0 = 2e-3
    flat, rays, cen, rad = cases.grazing_spheres(k0=k0)
    o = OracleScene(flat)
    t, _, _, prim, _ = o.intersect(rays, brute=True)
    hit = prim >= 0
    assert hit.sum() > 1000
    ro = rays[:, :3].astype(np.float64)
    rd = rays[:, 3:6].astype(np.float64)
    p = ro[hit] + t[hit].astype(np.float64)[:, None] * rd[hit]
    c = cen[prim[hit]]
    rr = rad[prim[hit]]
    dist_o = np.linalg.norm(ro[hit] - c, axis=1)
    excess = np.linalg.norm(p - c, axis=1) - rr
    allowed = np.minimum(k0 * dist_o, k0 * k0 * dist_o ** 2 / (2 * rr))
    # false hits exist (the margin is needed) ...
    line = np.linalg.norm(np.cross(c - ro[hit], rd[hit]), axis=1) / np.linalg.norm(rd[hit], axis=1)
    assert (line > rr * (1 + 1e-6) + 1e-7).sum() > 50
    # ... and every reported hit point lies within it
    assert np.all(excess <= allowed), float(np.max(excess / allowed))


def _tex(kind, params, mapping=None):
    t = abi.Texture()
    t.kind = kind
    t.mapping = abi.MAP_UV if mapping is None else mapping
    t.slots = 1
    for k, x in enumerate(params):
        t.params[k] = x
    return t


def test_texture_selectors_at_closed_form_points():
    """checkerBoard (Texture.hs:209-219) picks tex2 where
    floor(x sx) + floor(y sy) + floor(z sz) is odd (negative sums included);
    graphPaper (Texture.hs:191-207) picks tex2 (the line) where the
    fractional part of either mapped coordinate lies within lw/2 of an
    integer, measured by |x - truncate x| on both sides of zero."""
    import ctypes as C
    lib = pyoracle.load()
    lib.orc_tex_selects.restype = C.c_int
    lib.orc_tex_selects.argtypes = [C.POINTER(abi.Texture)] + [C.c_float] * 5

    def sel(t, p, u=0.0, v=0.0):
        return lib.orc_tex_selects(C.byref(t), *p, u, v)

    chk = _tex(abi.TEX_CHECKER, (1.0, 2.0, 1.0))
    assert sel(chk, (0.5, 0.2, 0.5)) == 0
    assert sel(chk, (1.5, 0.2, 0.5)) == 1
    assert sel(chk, (0.5, 0.7, 0.5)) == 1          # y scaled by 2: floor(1.4) = 1
    assert sel(chk, (-0.5, 0.2, 0.5)) == 1         # floor, not truncation: -1
    assert sel(chk, (-0.5, -0.2, 0.5)) == 0        # -1 + -1 = -2: even
    assert sel(chk, (-2.5, 0.2, 0.5)) == 1         # -3: odd under Haskell's mod
    # uv graph paper, su = 2, sv = 1, offsets 0, line width 0.1
    gp = _tex(abi.TEX_GRAPHPAPER, (0.1, 2.0, 1.0, 0.0, 0.0))
    p0 = (0.0, 0.0, 0.0)
    assert sel(gp, p0, 0.25, 0.5) == 0             # x = 0.5, z = 0.5
    assert sel(gp, p0, 0.51, 0.5) == 1             # x = 1.02: near a line
    assert sel(gp, p0, 0.25, 0.98) == 1            # z = 0.98 > 0.95
    assert sel(gp, p0, 0.5, 0.25) == 1             # u on a line (x = 1.0), v not
    assert sel(gp, p0, 0.25 / 2, 0.25) == 0        # x = 0.25
    assert sel(gp, p0, -0.01, 0.5) == 1            # x = -0.02: |frac| = 0.02
    assert sel(gp, p0, -0.24, 0.5) == 0            # x = -0.48
    # offsets: x = 2u + 0.5
    gpo = _tex(abi.TEX_GRAPHPAPER, (0.1, 2.0, 1.0, 0.5, 0.0))
    assert sel(gpo, p0, 0.25, 0.5) == 1 and sel(gpo, p0, 0.0, 0.5) == 0
    # planar: x = p . (1, 0, 0) + 0.25, z = p . (0, 0, 2)
    pl = _tex(abi.TEX_GRAPHPAPER, (0.1, 1.0, 0.0, 0.0, 0.0, 0.0, 2.0, 0.25, 0.0), abi.MAP_PLANAR)
    assert sel(pl, (0.25, 9.0, 0.25), 0.5, 0.5) == 0   # x = 0.5, z = 0.5
    assert sel(pl, (0.74, 9.0, 0.25), 0.5, 0.5) == 1   # x = 0.99
    assert sel(pl, (0.25, 9.0, 0.49), 0.5, 0.5) == 1   # z = 0.98
    assert sel(pl, (0.25, -3.0, 0.25), 0.0, 0.0) == 0  # uv and y are ignored


def test_textured_scene_differs_from_untextured():
    """The selectors reach the rendered image: the textured scene's mean
    differs from the same scene with every texture switched off."""
    from tests import cases
    flat = cases.textured_scene(res=16)
    spec = JobSpec(spp_nu=2, spp_nv=2, max_depth=4, sample_depth=3, rr_start_depth=3, passes=1,
                   seed=5, variance=True)
    a, _, _ = pyoracle.render(flat, spec)
    for i in range(flat.desc.num_materials):
        flat.materials[i].tex.kind = abi.TEX_NONE
    b, _, _ = pyoracle.render(flat, spec)
    assert np.isfinite(a).all() and np.isfinite(b).all()
    assert abs(a[..., 2].mean() - b[..., 2].mean()) > 0.02 * abs(b[..., 2].mean())


def _direct(flat, md, spp=4, passes=2, seed=5):
    spec = JobSpec(spp_nu=spp, spp_nv=spp, max_depth=md, sample_depth=0, passes=passes, seed=seed,
                   variance=True, integrator="direct")
    img, sq, _ = pyoracle.render(flat, spec)
    return mean_se(img, sq)


def test_direct_lighting_analytic_and_black_escapes():
    """DirectLighting (Integrator/DirectLighting.hs): a Lambertian plane
    under a uniform sky reflects rho * L and a sphere light gives
    rho * Le * (R / D)^2 below it, as for the path integrator's direct term;
    a ray that escapes is black (the integrator's `maybe black`), so a
    camera that sees only sky renders zero, where the path integrator
    renders the sky."""
    m, se = _direct(cases.plane_under_sky(rho=0.5, sky=1.0), md=3, spp=8, passes=4)
    y, s = m[..., 1].mean(), np.sqrt((se[..., 1] ** 2).sum()) / m[..., 1].size
    assert abs(y - 0.5) < 4 * s + 1e-4, (y, s)
    R, D, Le, rho = 0.25, 2.0, 4.0, 0.5
    m, se = _direct(cases.sphere_light_over_plane(rho=rho, Le=Le, R=R, D=D), md=2, spp=8, passes=4)
    y, s = m[..., 1].mean(), np.sqrt((se[..., 1] ** 2).sum()) / m[..., 1].size
    want = rho * Le * (R / D) ** 2
    assert abs(y - want) < 4 * s + 1e-4 * want, (y, want, s)
    sky = cases.sky_only(sky=2.0)
    m, _ = _direct(sky, md=4)
    assert (m[..., 1:] == 0).all()
    spec = JobSpec(spp_nu=2, spp_nv=2, max_depth=4, sample_depth=1, passes=1, seed=5)
    img, _, _ = pyoracle.render(sky, spec)
    assert img[..., 2].sum() > 0


def mirror_under_light(r=0.9, Le=3.0, theta=60.0, res=4):
    """A mirror (r, no Fresnel) at z = 0 seen at `theta` degrees from its
    normal by a narrow camera, under a large emissive quad at z = 3 facing
    down, behind the camera's rays, with the blackbody material."""
    sb = SceneBuilder()
    mir = sb.material(Material.mirror(const(r)))
    blk = sb.material(Material.black())
    sb.shape(abi.SHAPE_QUAD, (50.0, 50.0), material=mir)
    sb.shape(abi.SHAPE_QUAD, (100.0, 100.0), o2w=T.rotate_x(180) @ T.translate((0.0, 0.0, 3.0)), material=blk,
             emission=const(Le))
    t = np.radians(theta)
    c2w = T.look_at((0.0, -4.0 * np.sin(t), 4.0 * np.cos(t)), (0.0, 0.0, 0.0), (0, 0, 1))
    sb.perspective_camera(*T.perspective_camera(c2w, 0.0, 1.0, 0.02, res, res))
    sb.set_film(res, res)
    return sb.build()


def test_direct_lighting_specular_continuation():
    """At a mirror the light sample is zero (a specular BSDF evaluates to
    0) and the BSDF sample reaches the light with the power-heuristic weight
    w = 1 / (1 + pdf_light^2), pdf_light = dist^2 / (cos theta_l area); with
    maxDepth >= 2 the specular continuation (`cont`, DirectLighting.hs:45-59)
    adds f * Le, where f = r * fr(cos theta_o) = r is specRefl's sampled
    value with the no-op Fresnel of the mirror (Reflection/Specular.hs:11-20;
    no cosine and no pdf division).  The pixel is r Le (w + 1), and r Le w at
    maxDepth 1; the blackbody light adds no further branch."""
    r, Le, theta = 0.9, 3.0, 60.0
    c = np.cos(np.radians(theta))
    dist = 3.0 / c  # mirror point to the light along the reflected ray
    pdf_l = dist ** 2 / (c * (4 * 100.0 * 100.0))  # solid-angle pdf of the quad light
    w = 1.0 / (1.0 + pdf_l ** 2)
    flat = mirror_under_light(r=r, Le=Le, theta=theta)
    m1, _ = _direct(flat, md=1)
    m2, _ = _direct(flat, md=2)
    m5, _ = _direct(flat, md=5)
    np.testing.assert_allclose(m1[..., 1], r * Le * w, rtol=1e-5)
    np.testing.assert_allclose(m2[..., 1], r * Le * (w + 1.0), rtol=1e-5)
    np.testing.assert_allclose(m5[..., 1], m2[..., 1], rtol=1e-6)  # the light is black: no further branch
