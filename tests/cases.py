"""Shared scene builders for the closed-form fixtures and analytic renders."""
import json
import os

import numpy as np

from bling_amd import abi, scenes
from bling_amd import transform as T
from bling_amd.scene import Material, SceneBuilder, const, rgb_illum, rgb_refl
from bling_amd.textures import Map2d, Map3d, Tex, image_spectra

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
SHAPES = {"sphere": abi.SHAPE_SPHERE, "quad": abi.SHAPE_QUAD, "disk": abi.SHAPE_DISK,
          "box": abi.SHAPE_BOX, "cylinder": abi.SHAPE_CYLINDER}


def intersection_cases():
    with open(os.path.join(GOLDEN, "intersection_cases.json")) as f:
        return json.load(f)


def case_scene(c):
    sb = SceneBuilder()
    sb.material(Material.matte(const(0.5)))
    if c["kind"] == "triangle":
        sb.mesh(np.array(c["verts"], np.float32), np.array([[0, 1, 2]]))
    else:
        sb.shape(SHAPES[c["shape"]], c["params"], o2w=T.translate(c["translate"]))
    return sb.build()


def case_ray(c):
    o, d, tmin, tmax = c["ray"]
    return np.array([list(o) + list(d) + [tmin, np.inf if tmax == "inf" else tmax]], np.float32)


def sky_map(spec, w=1, h=1):
    return np.tile(np.asarray(spec, np.float32), (h, w, 1))


def plane_under_sky(rho=0.5, sky=1.0, res=8, fov=2.0, material=None):
    """A large diffuse quad at z = 0 under a uniform infinite light; the camera
    looks down at the origin.  Direct reflected radiance is rho * L (for a
    matte plane; `material` replaces it)."""
    sb = SceneBuilder()
    m = sb.material(material if material is not None else Material.matte(const(rho)))
    sb.shape(abi.SHAPE_QUAD, (1000.0, 1000.0), material=m)
    sb.infinite_light(sky_map(const(sky)))
    c2w = T.look_at((0.0, -3.0, 4.0), (0.0, 0.0, 0.0), (0, 0, 1))
    sb.perspective_camera(*T.perspective_camera(c2w, 0.0, 1.0, fov, res, res))
    sb.set_film(res, res)
    return sb.build()


def sphere_light_over_plane(rho=0.5, Le=4.0, R=0.25, D=2.0, res=8, fov=0.1):
    """A spherical area light of radius R at height D over a diffuse plane:
    at the point below it, E = pi Le (R/D)^2, so Lo = rho Le (R/D)^2.  The
    narrow field of view keeps the pixels' footprint within +-0.009 of that
    point, where E falls off by under 2e-5."""
    sb = SceneBuilder()
    m = sb.material(Material.matte(const(rho)))
    sb.shape(abi.SHAPE_QUAD, (1000.0, 1000.0), material=m)
    sb.shape(abi.SHAPE_SPHERE, (R,), o2w=T.translate((0.0, 0.0, D)), material=m,
             emission=const(Le))
    c2w = T.look_at((3.0, 0.0, 1.0), (0.0, 0.0, 0.0), (0, 0, 1))
    sb.perspective_camera(*T.perspective_camera(c2w, 0.0, 1.0, fov, res, res))
    sb.set_film(res, res)
    return sb.build()


def sky_only(sky=2.0, res=8):
    sb = SceneBuilder()
    sb.infinite_light(sky_map(const(sky), 4, 2))
    c2w = T.look_at((0.0, 0.0, 0.0), (0.0, 0.0, 1.0), (0, 1, 0))
    sb.perspective_camera(*T.perspective_camera(c2w, 0.0, 1.0, 60.0, res, res))
    sb.set_film(res, res)
    return sb.build()


def materials_scene(res=48, env=True, seed=3):
    """Spheres of every material over a matte floor, lit by a quad light, a
    sphere light and (optionally) a gradient sky."""
    from bling_amd.scene import rgb_refl, rgb_illum
    rng = np.random.default_rng(seed)
    sb = SceneBuilder()
    floor = sb.material(Material.matte(rgb_refl(0.6, 0.6, 0.55)))
    rough = sb.material(Material.matte(rgb_refl(0.7, 0.3, 0.2), sigma=0.4))
    plastic = sb.material(Material.plastic(rgb_refl(0.2, 0.3, 0.7), const(0.4), 0.02))
    metal = sb.material(Material.metal(const(0.2), const(3.9), 0.05))
    glass = sb.material(Material.glass(1.5, const(1.0), const(1.0)))
    mirror = sb.material(Material.mirror(const(0.9)))
    sb.mesh(np.array([[-4, -4, 0], [4, -4, 0], [4, 4, 0], [-4, 4, 0]], np.float32),
            np.array([[0, 1, 2], [0, 2, 3]]), material=floor)
    for k, m in enumerate((rough, plastic, metal, glass, mirror)):
        x = -2.0 + k * 1.0
        sb.shape(abi.SHAPE_SPHERE, (0.4,), o2w=T.translate((x, 0.0, 0.4)), material=m)
    lt = T.rotate_x(180) @ T.translate((0.0, 0.0, 3.0))
    sb.shape(abi.SHAPE_QUAD, (0.5, 0.5), o2w=lt, material=floor, emission=rgb_illum(6, 6, 6))
    sb.shape(abi.SHAPE_SPHERE, (0.15,), o2w=T.translate((2.0, -1.0, 1.5)), material=floor,
             emission=const(20.0))
    if env:
        h, w = 16, 32
        v = (np.arange(h, dtype=np.float32) + 0.5) / h
        sky = np.stack([np.outer(0.2 + 0.8 * v, np.ones(w)) * c for c in np.linspace(0.6, 1.0, 16)], -1)
        sb.infinite_light(sky.astype(np.float32))
    c2w = T.look_at((0.0, -5.0, 2.0), (0.0, 0.0, 0.4), (0, 0, 1))
    sb.perspective_camera(*T.perspective_camera(c2w, 0.0, 1.0, 45.0, res, res))
    sb.set_film(res, res)
    return sb.build()


def textured_scene(res=48, seed=3):
    """Selector textures on every mapping and primitive kind: a triangle
    floor with a planar-mapped checker, a uv graph-paper sphere, a plastic
    sphere whose two spectra are both textured, and a quad with planar
    graph paper, under a quad light."""
    from bling_amd.scene import rgb_refl, rgb_illum
    sb = SceneBuilder()
    floor = Material.matte(rgb_refl(0.7, 0.7, 0.7)).textured(
        0, abi.TEX_CHECKER, (1.0, 1.0, 1.0), rgb_refl(0.1, 0.2, 0.6))
    fm = sb.material(floor)
    paper = sb.material(Material.matte(rgb_refl(0.6, 0.6, 0.6)).textured(
        0, abi.TEX_GRAPHPAPER, (0.1, 8.0, 6.0, 0.0, 0.0), rgb_refl(0.02, 0.02, 0.02)))
    pl = Material.plastic(rgb_refl(0.2, 0.3, 0.7), const(0.4), 0.05).textured(
        0, abi.TEX_CHECKER, (4.0, 4.0, 4.0), rgb_refl(0.8, 0.1, 0.1))
    pl.tex_slots = 3
    pl.alt[1] = np.asarray(const(0.05), np.float32)
    plastic = sb.material(pl)
    wall = sb.material(Material.matte(rgb_refl(0.8, 0.8, 0.2)).textured(
        0, abi.TEX_GRAPHPAPER, (0.2, 2.0, 0.0, 0.0, 0.0, 0.0, 2.0, 0.25, 0.25),
        rgb_refl(0.1, 0.1, 0.1), mapping=abi.MAP_PLANAR))
    sb.mesh(np.array([[-4, -4, 0], [4, -4, 0], [4, 4, 0], [-4, 4, 0]], np.float32),
            np.array([[0, 1, 2], [0, 2, 3]]), material=fm,
            uvs=np.array([[0, 0], [1, 0], [1, 1], [0, 1]], np.float32))
    sb.shape(abi.SHAPE_SPHERE, (0.6,), o2w=T.translate((-0.8, 0.0, 0.6)), material=paper)
    sb.shape(abi.SHAPE_SPHERE, (0.5,), o2w=T.translate((0.8, 0.3, 0.5)), material=plastic)
    sb.shape(abi.SHAPE_QUAD, (2.0, 1.0), o2w=T.rotate_x(90) @ T.translate((0.0, 1.5, 1.0)),
             material=wall)
    lt = T.rotate_x(180) @ T.translate((0.0, 0.0, 3.0))
    sb.shape(abi.SHAPE_QUAD, (0.7, 0.7), o2w=lt, material=fm, emission=rgb_illum(8, 8, 8))
    c2w = T.look_at((0.0, -4.5, 2.0), (0.0, 0.0, 0.6), (0, 0, 1))
    sb.perspective_camera(*T.perspective_camera(c2w, 0.0, 1.0, 50.0, res, res))
    sb.set_film(res, res)
    return sb.build()


def grazing_spheres(n_sph=400, n=60000, seed=17, k0=2e-3, scaled=False):
    """Spheres of radius 1e-4..0.05 on a grid and rays that graze them from
    1..200 units away, at line distances from r - K/2 to r + 2K with
    K = min(k0 |o - c|, k0^2 |o - c|^2 / (2 r)): the regime where the
    reference's binary32 sphere test reports hits outside the surface.
    `scaled`: the spheres carry a rotation and a non-uniform scale (the
    stated radius is then the object-space one, and the targets are aimed at
    the unscaled centre, so the grazing distances are approximate).
    Returns (flat scene, rays (n, 8), centres (n_sph, 3), radii (n_sph,))."""
    rng = np.random.default_rng(seed)
    g = np.stack(np.meshgrid(*[np.arange(8)] * 3, indexing="ij"), -1).reshape(-1, 3)[:n_sph]
    cen = (g * 3.0).astype(np.float64) + rng.uniform(-0.5, 0.5, (n_sph, 3))
    cen = cen.astype(np.float32).astype(np.float64)
    rad = np.exp(rng.uniform(np.log(1e-4), np.log(0.05), n_sph)).astype(np.float32).astype(np.float64)
    sb = SceneBuilder()
    m = sb.material(Material.matte(const(0.5)))
    for k, (c, r) in enumerate(zip(cen, rad)):
        xf = T.translate(tuple(float(x) for x in c))
        if scaled:
            xf = (T.scale((1.0, 1.0 + 0.5 * (k % 3), 1.0)) @ T.rotate_x(float(rng.uniform(0, 360)))
                  @ T.rotate_y(float(rng.uniform(0, 360))) @ xf)
        sb.shape(abi.SHAPE_SPHERE, (float(r),), o2w=xf, material=m)
    tgt = rng.integers(n_sph, size=n)
    u = rng.normal(size=(n, 3))
    u /= np.linalg.norm(u, axis=1, keepdims=True)
    w = np.cross(u, rng.normal(size=(n, 3)))
    w /= np.linalg.norm(w, axis=1, keepdims=True)
    L = np.exp(rng.uniform(np.log(1.0), np.log(200.0), n))
    r = rad[tgt]
    bound = np.minimum(k0 * L, k0 * k0 * L * L / (2 * r))
    off = r + bound * rng.uniform(-0.5, 2.0, n)
    org = cen[tgt] + u * L[:, None]
    d = (cen[tgt] + w * off[:, None]) - org
    rays = np.zeros((n, 8), np.float32)
    rays[:, :3] = org
    rays[:, 3:6] = d
    rays[:, 7] = np.inf
    return sb.build(), rays, cen, rad


def material_abi(m):
    """The bgpu_material record of a scene.Material."""
    a = abi.Material()
    a.type = m.kind
    a.scalar[0], a.scalar[1], a.scalar[2] = m.s0, m.s1, m.s2
    for k in range(16):
        a.spec[0][k], a.spec[1][k], a.spec[2][k] = m.spec0[k], m.spec1[k], m.spec2[k]
    return a


def frame(n, p):
    """Object-to-world transform whose z axis is the unit normal n, placed at p."""
    n = np.asarray(n, np.float64)
    n = n / np.linalg.norm(n)
    a = np.array([0.0, 1.0, 0.0]) if abs(n[1]) < 0.9 else np.array([1.0, 0.0, 0.0])
    t1 = np.cross(a, n)
    t1 /= np.linalg.norm(t1)
    t2 = np.cross(n, t1)
    rows = [[t1[0], t2[0], n[0], p[0]], [t1[1], t2[1], n[1], p[1]], [t1[2], t2[2], n[2], p[2]], [0, 0, 0, 1]]
    return T.from_matrix(rows)


def two_mirrors_under_sky(r=0.9, sky=2.0, res=4, fov=1.0):
    """A camera at the origin looks along +z at mirror 1 (normal (-1, 0, -1)),
    which sends its rays along -x to mirror 2 at x = -10 (normal (1, 1, 0)),
    which sends them up (+y) into a uniform sky.  Both mirrors reflect r with
    no Fresnel term; every camera ray takes this two-bounce path."""
    sb = SceneBuilder()
    mir = sb.material(Material.mirror(const(r)))
    sb.shape(abi.SHAPE_QUAD, (2.0, 2.0), o2w=frame((-1.0, 0.0, -1.0), (0.0, 0.0, 5.0)), material=mir)
    sb.shape(abi.SHAPE_QUAD, (3.0, 3.0), o2w=frame((1.0, 1.0, 0.0), (-10.0, 0.0, 5.0)), material=mir)
    sb.infinite_light(sky_map(const(sky), 4, 2))
    c2w = T.look_at((0.0, 0.0, 0.0), (0.0, 0.0, 5.0), (0, 1, 0))
    sb.perspective_camera(*T.perspective_camera(c2w, 0.0, 1.0, fov, res, res))
    sb.set_film(res, res)
    return sb.build()


def checker_image(h=16, w=24, seed=3):
    """An 8-bit RGB test image: coloured blocks (for image textures)."""
    rng = np.random.default_rng(seed)
    blocks = rng.integers(0, 256, (4, 6, 3))
    return np.repeat(np.repeat(blocks, h // 4, axis=0), w // 6, axis=1).astype(np.uint8)


def node_texture_scene(res=48):
    """Every node texture on a lit floor (Texture.hs:59-414): a quad floor
    whose matte kd blends two colours by a quasi-crystal on a planar mapping
    and whose sigma is a scaled fbm; a plastic sphere coloured by a gradient
    of Euclidean cell noise; a matte box with an RGB image over the box's
    own uv; a mesh quad with graph paper over a Perlin blend (uv mapping); a
    substrate sphere with a checker kd, a grey-image ka and a textured
    urough; under a uniform sky and a quad light."""
    sb = SceneBuilder()
    floor = Material.matte(rgb_refl(0.5, 0.5, 0.5)).textured_inputs(
        spec={0: Tex.blend(Tex.spectrum(rgb_refl(0.8, 0.8, 0.75)), Tex.spectrum(rgb_refl(0.2, 0.25, 0.3)),
                           Tex.crystal(5, Map2d.planar((1.0, 0, 0), (0, 0, 1.0), -3.0, 0.0)))},
        scalar={0: Tex.scale(0.25, 0.25, Tex.fbm(4, 0.5, Map3d.identity(T.scale((2.0, 2.0, 2.0)))))})
    fm = sb.material(floor)
    sb.shape(abi.SHAPE_QUAD, (4.0, 4.0), o2w=T.rotate_x(90), material=fm)
    cellm = sb.material(Material.plastic(const(0), rgb_refl(0.6, 0.6, 0.6), 0.05).textured_inputs(
        spec={0: Tex.gradient(Tex.cell("euclidian", Map3d.identity(T.scale((3.0, 3.0, 3.0)))),
                              [(0.0, rgb_refl(0.1, 0.1, 0.1)), (0.6, rgb_refl(0.2, 0.7, 0.7)),
                               (1.0, rgb_refl(0.9, 0.6, 0.2))])}))
    sb.shape(abi.SHAPE_SPHERE, (0.8,), o2w=T.translate((-1.2, 0.8, 0.3)), material=cellm)
    img = sb.material(Material.matte(const(0.5)).textured_inputs(
        spec={0: Tex.image(image_spectra(checker_image()), Map2d.uv(1.0, 1.0, 0.0, 0.0))}))
    sb.shape(abi.SHAPE_BOX, (-0.5, 0.0, -0.5, 0.5, 1.0, 0.5), o2w=T.rotate_y(30) @ T.translate((1.3, 0.0, 0.4)),
             material=img)
    paper = sb.material(Material.matte(const(0.5)).textured_inputs(
        spec={0: Tex.graph_paper(0.08, Map2d.uv(8.0, 8.0, 0.0, 0.0),
                                 Tex.blend(Tex.spectrum(rgb_refl(0.7, 0.7, 0.2)), Tex.spectrum(rgb_refl(0.2, 0.3, 0.7)),
                                           Tex.scale(0.5, 0.5, Tex.perlin(Map3d.identity(T.scale((4.0, 4.0, 4.0)))))),
                                 Tex.spectrum(rgb_refl(0.05, 0.05, 0.05)))}))
    sb.mesh(np.array([[-2.5, 0.01, 1.2], [-0.5, 0.01, 1.2], [-0.5, 0.01, 2.4], [-2.5, 0.01, 2.4]], np.float32),
            np.array([[0, 1, 2], [0, 2, 3]], np.int32), material=paper,
            uvs=np.array([[0, 0], [1, 0], [1, 1], [0, 1]], np.float32))
    grey = np.clip(np.arange(12 * 12).reshape(12, 12) * 2, 0, 255).astype(np.uint8)
    sub = sb.material(Material.substrate(const(0.5), rgb_refl(0.1, 0.1, 0.1), const(0.0), 0.05, 0.05, 0.3)
                      .textured_inputs(spec={0: Tex.checker((4.0, 4.0, 4.0), Tex.spectrum(rgb_refl(0.8, 0.2, 0.2)),
                                                            Tex.spectrum(rgb_refl(0.9, 0.9, 0.9))),
                                             2: Tex.spectrum(rgb_refl(0.3, 0.1, 0.0))},
                                       scalar={0: Tex.scale(0.02, 0.2, Tex.image_grey(grey, Map2d.uv(1.0, 1.0, 0.0, 0.0)))}))
    sb.shape(abi.SHAPE_SPHERE, (0.6,), o2w=T.translate((0.2, 0.6, -1.0)), material=sub)
    lt = T.rotate_x(-90) @ T.translate((0.0, 4.0, 0.0))
    sb.shape(abi.SHAPE_QUAD, (1.0, 1.0), o2w=lt, material=fm, emission=rgb_illum(8.0, 8.0, 8.0))
    sb.infinite_light(np.tile(const(0.4), (2, 4, 1)))
    c2w = T.look_at((0.0, 3.5, -5.0), (0.0, 0.5, 0.3), (0, 1, 0))
    sb.look(c2w, 50.0, res, res)
    return sb.build()


def bump_scene(res=48):
    """Bump-mapped materials (bumpMapped, Reflection.hs:344-378): a plastic
    sphere under Euclidean cell noise, a matte floor quad under fBm, a mesh
    with uvs under a uv-mapped quasi-crystal and a checker-textured matte
    under Perlin noise, lit by a quad light and a uniform sky."""
    sb = SceneBuilder()
    floor = sb.material(Material.matte(rgb_refl(0.6, 0.6, 0.6)).textured_inputs(
        bump=Tex.scale(0.0, 0.05, Tex.fbm(4, 0.5, Map3d.identity(T.scale((3.0, 3.0, 3.0)))))))
    sb.shape(abi.SHAPE_QUAD, (4.0, 4.0), o2w=T.rotate_x(90), material=floor)
    cell = sb.material(Material.plastic(rgb_refl(0.2, 0.5, 0.7), rgb_refl(0.5, 0.5, 0.5), 0.05).textured_inputs(
        bump=Tex.scale(0.0, 0.2, Tex.cell("euclidian", Map3d.identity(T.scale((2.0, 2.0, 2.0)))))))
    sb.shape(abi.SHAPE_SPHERE, (0.8,), o2w=T.translate((-1.0, 0.8, 0.2)), material=cell)
    cry = sb.material(Material.matte(rgb_refl(0.8, 0.5, 0.3)).textured_inputs(
        bump=Tex.scale(0.0, 0.03, Tex.crystal(5, Map2d.uv(6.0, 6.0, 0.0, 0.0)))))
    sb.mesh(np.array([[0.2, 0.0, -0.5], [1.8, 0.0, -0.5], [1.8, 1.5, 0.2], [0.2, 1.5, 0.2]], np.float32),
            np.array([[0, 1, 2], [0, 2, 3]], np.int32), material=cry,
            uvs=np.array([[0, 0], [1, 0], [1, 1], [0, 1]], np.float32))
    chk = sb.material(Material.matte(const(0.5)).textured_inputs(
        spec={0: Tex.checker((4.0, 4.0, 4.0), Tex.spectrum(rgb_refl(0.9, 0.9, 0.2)),
                             Tex.spectrum(rgb_refl(0.1, 0.1, 0.6)))},
        bump=Tex.scale(0.0, 0.1, Tex.perlin(Map3d.identity(T.scale((5.0, 5.0, 5.0)))))))
    sb.shape(abi.SHAPE_SPHERE, (0.5,), o2w=T.translate((0.6, 0.5, 1.2)), material=chk)
    lt = T.rotate_x(-90) @ T.translate((0.0, 4.0, 0.0))
    sb.shape(abi.SHAPE_QUAD, (1.0, 1.0), o2w=lt, material=floor, emission=rgb_illum(8.0, 8.0, 8.0))
    sb.infinite_light(np.tile(const(0.3), (2, 4, 1)))
    sb.look(T.look_at((0.0, 3.0, -5.0), (0.0, 0.5, 0.3), (0, 1, 0)), 50.0, res, res)
    return sb.build()


def delta_light_over_plane(kind, rho=0.5, I=3.0, D=2.0, theta=60.0, res=4, fov=0.1):
    """A Lambertian plane z = 0 lit by one delta light, seen near the origin
    by a narrow camera: a point light of intensity I at height D (Lo = rho I
    / (pi D^2) below it), or a directional light of intensity I from
    `theta` degrees off the normal (Lo = rho I cos^2(theta) / pi: the light's
    |n . d| (Light.hs:142-144) and the BSDF's cosine)."""
    sb = SceneBuilder()
    m = sb.material(Material.matte(const(rho)))
    sb.shape(abi.SHAPE_QUAD, (1000.0, 1000.0), material=m)
    if kind == "point":
        sb.point_light(const(I), (0.0, 0.0, D))
    else:
        t = np.radians(theta)
        sb.directional_light(const(I), (2.0 * np.sin(t), 0.0, 2.0 * np.cos(t)))  # normalized by the library
    c2w = T.look_at((3.0, 0.0, 1.0), (0.0, 0.0, 0.0), (0, 0, 1))
    sb.perspective_camera(*T.perspective_camera(c2w, 0.0, 1.0, fov, res, res))
    sb.set_film(res, res)
    return sb.build()


def sun_sky_scene(env_camera=False, res=48):
    """Matte, plastic and glass spheres on a floor under a sun/sky light map
    (the light of race.bling, infinite { rotateX -90 } with sunSky east 1 0 0,
    the sun raised and turned toward the camera), seen by a perspective
    camera or, from above the floor, by the environment camera (2:1 raster)."""
    sb = SceneBuilder()
    floor = sb.material(Material.matte(rgb_refl(0.6, 0.6, 0.6)))
    sb.shape(abi.SHAPE_QUAD, (6.0, 6.0), o2w=T.rotate_x(90), material=floor)
    mats = [Material.matte(rgb_refl(0.7, 0.3, 0.2), 0.3), Material.plastic(rgb_refl(0.2, 0.4, 0.7),
            rgb_refl(0.5, 0.5, 0.5), 0.05), Material.glass(1.5, const(1.0), const(1.0))]
    for x, m in zip((-1.2, 0.0, 1.2), mats):
        sb.shape(abi.SHAPE_SPHERE, (0.5,), o2w=T.translate((x, 0.5, 0.0)), material=sb.material(m))
    sb.sun_sky_light((1.0, 0.0, 0.0), (0.4, 0.5, 1.0), 5.0, T.rotate_x(-90))
    if env_camera:
        sb.environment_camera(T.translate((0.0, 1.0, -2.0)), 2 * res, res)
        sb.set_film(2 * res, res)
    else:
        sb.look(T.look_at((0.0, 1.2, -4.0), (0.0, 0.6, 0.0), (0, 1, 0)), 60.0, res, res)
    return sb.build()
