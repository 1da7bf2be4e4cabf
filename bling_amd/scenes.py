"""Procedural workloads (synthetic, seeded):

  cornell()          Cornell-box-style scene: 5 unit-cube walls as 10 triangles,
                     two boxes of 12 triangles rotated +-17 degrees, a 0.25x0.25
                     emissive quad (rgbI 15 15 15) under the ceiling, Lambertian
                     red / green / 0.75 grey, pinhole camera with 40 degree FOV.
  icosphere_scene()  intersection workload: an icosphere subdivided 8 times
                     (1,310,720 triangles) displaced by 0.05 * fBm (6 octaves,
                     seed 1) plus 10,000 spheres with radii log-uniform in
                     [0.001, 0.05]; coherent (pinhole primary) and incoherent
                     (uniform origin in the bounds, uniform direction) rays.
  stack_scene()      random axis-aligned box stacks (12 triangles per box) on a
                     floor, lit by a quad and a small sphere light.
  grid_scene()       a displaced heightmap grid (Heightmap.hs) with shading
                     normals, lit the same way.

The meshes go through SceneBuilder.mesh, i.e. the same transPoint as the
reference's mkTriangleMesh.
"""
import numpy as np

from . import abi
from . import transform as T
from .scene import Material, SceneBuilder, const, rgb_illum, rgb_refl

F = np.float32


# ---------------------------------------------------------------- noise
class Perlin:
    """Improved-noise style gradient noise (Perlin 2002): 256-entry permutation
    (here drawn from a seeded generator), quintic fade, 12+4 gradient set as
    in perlin3d (Texture.hs:357-397)."""

    def __init__(self, seed):
        p = np.random.default_rng(seed).permutation(256).astype(np.int64)
        self.perm = np.concatenate([p, p])

    def _grad(self, h, x, y, z):
        h = h & 15
        u = np.where((h < 8) | (h == 12) | (h == 13), x, y)
        v = np.where((h < 4) | (h == 12) | (h == 13), y, z)
        return np.where(h & 1, -u, u) + np.where(h & 2, -v, v)

    def __call__(self, x, y, z):
        xi, yi, zi = np.floor(x), np.floor(y), np.floor(z)
        dx, dy, dz = x - xi, y - yi, z - zi
        ix, iy, iz = (xi.astype(np.int64) & 255, yi.astype(np.int64) & 255, zi.astype(np.int64) & 255)
        P = self.perm

        def h(a, b, c):
            return P[P[P[a] + b] + c]

        fade = lambda t: t * t * t * (t * (t * 6 - 15) + 10)
        wx, wy, wz = fade(dx), fade(dy), fade(dz)
        g = self._grad
        x00 = (1 - wx) * g(h(ix, iy, iz), dx, dy, dz) + wx * g(h(ix + 1, iy, iz), dx - 1, dy, dz)
        x10 = (1 - wx) * g(h(ix, iy + 1, iz), dx, dy - 1, dz) + wx * g(h(ix + 1, iy + 1, iz), dx - 1, dy - 1, dz)
        x01 = (1 - wx) * g(h(ix, iy, iz + 1), dx, dy, dz - 1) + wx * g(h(ix + 1, iy, iz + 1), dx - 1, dy, dz - 1)
        x11 = (1 - wx) * g(h(ix, iy + 1, iz + 1), dx, dy - 1, dz - 1) + wx * g(h(ix + 1, iy + 1, iz + 1), dx - 1, dy - 1, dz - 1)
        y0 = (1 - wy) * x00 + wy * x10
        y1 = (1 - wy) * x01 + wy * x11
        return (1 - wz) * y0 + wz * y1


def fbm(p, octaves=6, omega=0.5, seed=1):
    """fbm octaves omega (Texture.hs:342-350): sum of omega^i * noise(p * 1.99^i)."""
    noise = Perlin(seed)
    off = np.random.default_rng(seed + 7919).uniform(0, 64, 3)
    out = np.zeros(len(p))
    l, o = 1.0, 1.0
    for _ in range(octaves):
        q = (p + off) * l
        out += o * noise(q[:, 0], q[:, 1], q[:, 2])
        l *= 1.99
        o *= omega
    return out


# ------------------------------------------------------------- geometry
def icosphere(levels):
    t = (1.0 + 5 ** 0.5) / 2.0
    v = np.array([[-1, t, 0], [1, t, 0], [-1, -t, 0], [1, -t, 0], [0, -1, t], [0, 1, t],
                  [0, -1, -t], [0, 1, -t], [t, 0, -1], [t, 0, 1], [-t, 0, -1], [-t, 0, 1]], float)
    f = np.array([[0, 11, 5], [0, 5, 1], [0, 1, 7], [0, 7, 10], [0, 10, 11], [1, 5, 9], [5, 11, 4],
                  [11, 10, 2], [10, 7, 6], [7, 1, 8], [3, 9, 4], [3, 4, 2], [3, 2, 6], [3, 6, 8],
                  [3, 8, 9], [4, 9, 5], [2, 4, 11], [6, 2, 10], [8, 6, 7], [9, 8, 1]], np.int64)
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    for _ in range(levels):
        e = np.concatenate([f[:, [0, 1]], f[:, [1, 2]], f[:, [2, 0]]])
        e.sort(axis=1)
        key = e[:, 0] * (len(v) + 1) + e[:, 1]
        uk, inv = np.unique(key, return_inverse=True)
        a, b = uk // (len(v) + 1), uk % (len(v) + 1)
        mid = v[a] + v[b]
        mid /= np.linalg.norm(mid, axis=1, keepdims=True)
        base = len(v)
        v = np.concatenate([v, mid])
        nf = len(f)
        m01, m12, m20 = inv[:nf] + base, inv[nf:2 * nf] + base, inv[2 * nf:] + base
        f = np.concatenate([np.stack([f[:, 0], m01, m20], 1), np.stack([f[:, 1], m12, m01], 1),
                            np.stack([f[:, 2], m20, m12], 1), np.stack([m01, m12, m20], 1)])
    return v, f


def unit_box():
    """Unit cube [-0.5, 0.5]^3 as 12 triangles."""
    v = np.array([[x, y, z] for x in (-0.5, 0.5) for y in (-0.5, 0.5) for z in (-0.5, 0.5)], F)
    quads = [(0, 1, 3, 2), (4, 6, 7, 5), (0, 4, 5, 1), (2, 3, 7, 6), (0, 2, 6, 4), (1, 5, 7, 3)]
    f = []
    for a, b, c, d in quads:
        f += [(a, b, c), (a, c, d)]
    return v, np.array(f, np.int32)


# --------------------------------------------------------------- scenes
def cornell(res=256, light_radiance=15.0, height=None):
    """Cornell box (unit cube, quad light under the ceiling); a `res` x
    `height` (default `res`) image."""
    sb = SceneBuilder()
    red = sb.material(Material.matte(rgb_refl(0.63, 0.065, 0.05)))
    green = sb.material(Material.matte(rgb_refl(0.14, 0.45, 0.091)))
    grey = sb.material(Material.matte(rgb_refl(0.75, 0.75, 0.75)))
    lightm = sb.material(Material.matte(rgb_refl(0.78, 0.78, 0.78)))
    quad = lambda a, b, c, d: (np.array([a, b, c, d], F), np.array([[0, 1, 2], [0, 2, 3]], np.int32))
    v, f = quad((0, 0, 0), (1, 0, 0), (1, 0, 1), (0, 0, 1))        # floor y=0
    sb.mesh(v, f, material=grey)
    v, f = quad((0, 1, 0), (0, 1, 1), (1, 1, 1), (1, 1, 0))        # ceiling y=1
    sb.mesh(v, f, material=grey)
    v, f = quad((0, 0, 1), (1, 0, 1), (1, 1, 1), (0, 1, 1))        # back z=1
    sb.mesh(v, f, material=grey)
    v, f = quad((0, 0, 0), (0, 0, 1), (0, 1, 1), (0, 1, 0))        # left x=0
    sb.mesh(v, f, material=red)
    v, f = quad((1, 0, 0), (1, 1, 0), (1, 1, 1), (1, 0, 1))        # right x=1
    sb.mesh(v, f, material=green)
    bv, bf = unit_box()
    short = T.scale((0.3, 0.3, 0.3)) @ T.rotate_y(-17) @ T.translate((0.35, 0.15, 0.33))
    tall = T.scale((0.3, 0.6, 0.3)) @ T.rotate_y(17) @ T.translate((0.66, 0.3, 0.64))
    sb.mesh(bv, bf, o2w=short, material=grey)
    sb.mesh(bv, bf, o2w=tall, material=grey)
    lt = T.rotate_x(-90) @ T.translate((0.5, 0.999, 0.5))
    sb.shape(abi.SHAPE_QUAD, (0.125, 0.125), o2w=lt, material=lightm,
             emission=rgb_illum(light_radiance, light_radiance, light_radiance))
    c2w = T.look_at((0.5, 0.5, -1.4), (0.5, 0.5, 0.5), (0, 1, 0))
    sb.look(c2w, 40.0, res, height or res)
    return sb.build()


def icosphere_scene(levels=8, num_spheres=10000, seed=1, amplitude=0.05, res=2048, lit=False):
    """A displaced icosphere (amplitude * fBm) among spheres of log-uniform
    radius; `lit` adds a uniform sky (for rendering; the ray queries do not
    need light)."""
    v, f = icosphere(levels)
    r = 1.0 + amplitude * fbm(v, octaves=6, omega=0.5, seed=seed)
    v = (v * r[:, None]).astype(F)
    sb = SceneBuilder()
    m = sb.material(Material.matte(const(0.6)))
    sb.mesh(v, f.astype(np.int32), material=m)
    rng = np.random.default_rng(seed)
    cen = rng.uniform(-1.5, 1.5, (num_spheres, 3))
    rad = np.exp(rng.uniform(np.log(0.001), np.log(0.05), num_spheres))
    for c, rr in zip(cen, rad):
        sb.shape(abi.SHAPE_SPHERE, (float(rr),), o2w=T.translate(tuple(c)), material=m)
    if lit:
        sb.infinite_light(np.tile(const(1.0), (2, 4, 1)))
    c2w = T.look_at((0.6, 0.9, -3.4), (0.0, 0.0, 0.0), (0, 1, 0))
    sb.look(c2w, 50.0, res, res)
    return sb.build()


def small_scene(num_tris=2000, num_spheres=200, seed=5, shapes_all=True):
    """Random soup of triangles plus every shape type; for parity tests."""
    rng = np.random.default_rng(seed)
    sb = SceneBuilder()
    m = sb.material(Material.matte(const(0.5)))
    c = rng.uniform(-1, 1, (num_tris, 1, 3))
    v = (c + rng.normal(0, 0.08, (num_tris, 3, 3))).reshape(-1, 3).astype(F)
    sb.mesh(v, np.arange(3 * num_tris, dtype=np.int32).reshape(-1, 3), material=m)
    kinds = [abi.SHAPE_SPHERE, abi.SHAPE_BOX, abi.SHAPE_QUAD, abi.SHAPE_DISK, abi.SHAPE_CYLINDER]
    for i in range(num_spheres):
        k = kinds[i % len(kinds)] if shapes_all else abi.SHAPE_SPHERE
        pos = tuple(rng.uniform(-1.2, 1.2, 3))
        xf = T.rotate_x(float(rng.uniform(0, 360))) @ T.rotate_y(float(rng.uniform(0, 360))) @ T.translate(pos)
        s = float(rng.uniform(0.02, 0.15))
        if k == abi.SHAPE_SPHERE:
            p = (s,)
        elif k == abi.SHAPE_BOX:
            p = (-s, -s * 0.5, -s * 0.7, s, s * 0.5, s * 0.7)
        elif k == abi.SHAPE_QUAD:
            p = (s, s * 0.6)
        elif k == abi.SHAPE_DISK:
            p = (s * 0.2, s, s * 0.3 if i % 2 else 0.0, 360.0 if i % 3 else 270.0)
            p = (p[0], p[1], p[2], float(T.radians(min(max(p[3], 0), 360))))
        else:
            p = (s * 0.5, -s, s, float(T.radians(360.0 if i % 3 else 300.0)))
        sb.shape(k, p, o2w=xf, material=m)
    return sb.build()


def heightmap_mesh(elev, ns, nt):
    """heightMap elev (ns, nt) (Primitive/Heightmap.hs:15-50): an ns x nt
    vertex grid over [0,1]^2 with y = elev(x, z), two triangles per cell,
    shading normals from central differences and uv = (x, z) / (ns - 1)."""
    fns, fnt = float(ns), float(nt)
    xs = np.arange(ns) / (fns - 1)
    zs = np.arange(nt) / (fnt - 1)
    X, Z = np.meshgrid(xs, zs)  # row z, column x: vertex index x + z * ns
    X, Z = X.reshape(-1), Z.reshape(-1)
    Y = elev(X, Z)
    verts = np.stack([X, Y, Z], 1).astype(F)
    ex, ez = 1.0 / fns, 1.0 / fnt
    dx = elev(X - ex, Z) - elev(X + ex, Z)
    dz = elev(X, Z - ez) - elev(X, Z + ez)
    n = np.stack([dx, np.full_like(X, ex + ez), dz], 1)
    n = -(n / np.linalg.norm(n, axis=1, keepdims=True))
    uvs = np.stack([X / (fns - 1), Z / (fns - 1)], 1)
    x, y = np.meshgrid(np.arange(ns - 1), np.arange(nt - 1))
    x, y = x.reshape(-1), y.reshape(-1)
    v = lambda a, b: a + b * ns
    first = np.stack([v(x, y), v(x + 1, y), v(x + 1, y + 1)], 1)
    second = np.stack([v(x, y), v(x + 1, y + 1), v(x, y + 1)], 1)
    faces = np.stack([first, second], 1).reshape(-1, 3).astype(np.int32)
    return verts, faces, n.astype(F), uvs.astype(F)


def _lit_room(sb, size, light_radiance=12.0, height=None):
    """Floor, a quad light above, a small sphere light; Lambertian."""
    floor = sb.material(Material.matte(rgb_refl(0.7, 0.7, 0.7)))
    h = height if height is not None else size
    sb.mesh(np.array([[-size, 0, -size], [size, 0, -size], [size, 0, size], [-size, 0, size]], F),
            np.array([[0, 1, 2], [0, 2, 3]], np.int32), material=floor)
    # light sampling emits towards the quad's local -z (Shape.hs:405, Light.hs:153)
    lt = T.rotate_x(-90) @ T.translate((0.0, h * 1.5, 0.0))
    sb.shape(abi.SHAPE_QUAD, (size * 0.3, size * 0.3), o2w=lt, material=floor,
             emission=rgb_illum(light_radiance, light_radiance, light_radiance))
    sb.shape(abi.SHAPE_SPHERE, (size * 0.05,), o2w=T.translate((size * 0.6, h * 0.8, -size * 0.4)),
             material=floor, emission=const(light_radiance * 4))
    return floor


def stack_scene(num_tris, res=256, seed=0):
    """Random axis-aligned box stacks (12 triangles per box) on a lit floor."""
    rng = np.random.default_rng(seed)
    sb = SceneBuilder()
    _lit_room(sb, 2.0)
    mats = [sb.material(Material.matte(rgb_refl(*rng.uniform(0.2, 0.9, 3)))) for _ in range(4)]
    nbox = max(1, num_tris // 12)
    bv, bf = unit_box()
    cols = int(np.ceil(np.sqrt(max(1, nbox // 8))))
    centres, sizes = [], []
    for k in range(nbox):
        c = k % max(1, cols * cols)
        i, j = divmod(c, cols)
        level = k // max(1, cols * cols)
        s = rng.uniform(0.3, 0.9) * 3.2 / cols
        centres.append((-1.6 + (i + 0.5) * 3.2 / cols, (level + 0.5) * s, -1.6 + (j + 0.5) * 3.2 / cols))
        sizes.append(s)
    centres = np.array(centres, F)
    sizes = np.array(sizes, F)
    verts = (bv[None, :, :] * sizes[:, None, None] + centres[:, None, :]).reshape(-1, 3)
    faces = (bf[None, :, :] + 8 * np.arange(nbox)[:, None, None]).reshape(-1, 3)
    m = np.array(mats)[np.arange(nbox) % len(mats)]
    # one mesh per material keeps mkTriangleMesh's one-material-per-mesh rule
    for mi in mats:
        sel = np.nonzero(m == mi)[0]
        if len(sel):
            sb.mesh(verts.reshape(nbox, 8, 3)[sel].reshape(-1, 3),
                    (bf[None] + 8 * np.arange(len(sel))[:, None, None]).reshape(-1, 3), material=mi)
    c2w = T.look_at((3.5, 3.0, -4.5), (0.0, 0.6, 0.0), (0, 1, 0))
    sb.look(c2w, 45.0, res, res)
    return sb.build()


def fbm_elevation(n, seed, scale=0.25, freq=4.0):
    """elev(x, z) on [0,1]^2: scale * fBm(freq * (x, z, 0)) (6 octaves), sampled
    once on a lattice of at most 513^2 points and interpolated bilinearly (the
    heightmap evaluates elev five times per vertex)."""
    L = min(n, 512) + 1
    lx, lz = np.meshgrid(np.arange(L) / (L - 1.0), np.arange(L) / (L - 1.0))
    lat = scale * fbm(np.stack([lx.reshape(-1) * freq, lz.reshape(-1) * freq, np.zeros(L * L)], 1),
                      octaves=6, omega=0.5, seed=seed).reshape(L, L)

    def elev(x, z):
        u = np.clip(x, 0.0, 1.0) * (L - 1)
        w = np.clip(z, 0.0, 1.0) * (L - 1)
        i = np.minimum(u.astype(np.int64), L - 2)
        k = np.minimum(w.astype(np.int64), L - 2)
        fu, fw = u - i, w - k
        return ((1 - fu) * (1 - fw) * lat[k, i] + fu * (1 - fw) * lat[k, i + 1]
                + (1 - fu) * fw * lat[k + 1, i] + fu * fw * lat[k + 1, i + 1])
    return elev


def grid_scene(num_tris, res=256, seed=0):
    """A displaced heightmap grid (fBm heights) with shading normals, lit."""
    n = max(2, int(np.sqrt(num_tris / 2.0)) + 1)
    sb = SceneBuilder()
    _lit_room(sb, 2.0, height=1.0)
    m = sb.material(Material.matte(rgb_refl(0.5, 0.6, 0.4)))
    v, f, nrm, uv = heightmap_mesh(fbm_elevation(n, seed + 1), n, n)
    o2w = T.scale((3.0, 1.0, 3.0)) @ T.translate((-1.5, 0.1, -1.5))
    sb.mesh(v, f, o2w=o2w, material=m, normals=nrm, uvs=uv)
    c2w = T.look_at((0.0, 2.5, -3.5), (0.0, 0.2, 0.0), (0, 1, 0))
    sb.look(c2w, 50.0, res, res)
    return sb.build()


def torus(major, minor, nu, nv):
    """A torus around the y axis tessellated into 2 * nu * nv triangles."""
    u = 2 * np.pi * np.arange(nu) / nu
    v = 2 * np.pi * np.arange(nv) / nv
    U, V = np.meshgrid(u, v, indexing="ij")
    ring = major + minor * np.cos(V)
    verts = np.stack([ring * np.cos(U), minor * np.sin(V), ring * np.sin(U)], -1).reshape(-1, 3)
    i, j = np.meshgrid(np.arange(nu), np.arange(nv), indexing="ij")
    a = i * nv + j
    b = ((i + 1) % nu) * nv + j
    c = ((i + 1) % nu) * nv + (j + 1) % nv
    d = i * nv + (j + 1) % nv
    faces = np.stack([np.stack([a, b, c], -1), np.stack([a, c, d], -1)], 2).reshape(-1, 3)
    return verts.astype(F), faces.astype(np.int32)


def sky_gradient(w=512, h=256):
    """A w x h spectral sky map: a bluish zenith over a pale horizon and a
    dark ground.  A light map's row y covers v in ((h-1-y)/h, (h-y)/h]
    (rgbfToTexMap, IO/Bitmap.hs:21-28), so the rows run from the -z pole
    (v = 1, ground) at row 0 to the +z pole (v = 0, zenith) at row h - 1."""
    zen, hor, gnd = rgb_illum(0.3, 0.5, 1.0), rgb_illum(1.0, 0.95, 0.85), rgb_illum(0.15, 0.12, 0.1)
    zen, hor, gnd = np.asarray(zen, F), np.asarray(hor, F), np.asarray(gnd, F)
    v = (np.arange(h, dtype=np.float64) + 0.5) / h
    rows = np.where((v < 0.5)[:, None], zen[None] + (hor - zen)[None] * (v / 0.5)[:, None] ** 0.7,
                    gnd[None] + (hor - gnd)[None] * np.exp(-(v - 0.5) * 12.0)[:, None])
    rows = rows[::-1]  # v = (h - 1 - y + 0.5) / h at row y
    return np.ascontiguousarray(np.broadcast_to(rows[:, None, :], (h, w, 16)).astype(F))


def divergent_scene(res=1024, hm=1024, n_tori=20, torus_tris=20000, n_spheres=200, sky=(512, 256),
                    seed=4):
    """Divergent materials: a fBm heightmap of hm x hm vertices, n_tori tori of
    torus_tris triangles and n_spheres spheres, each object with a material
    drawn uniformly from matte, plastic (Blinn exponent 10..1000), metal
    (roughness 0..0.3) and glass (IOR 1.5), lit only by a sky-gradient
    infinite light."""
    rng = np.random.default_rng(seed)
    sb = SceneBuilder()

    def pick():
        k = rng.integers(4)
        if k == 0:
            return sb.material(Material.matte(rgb_refl(*rng.uniform(0.2, 0.9, 3))))
        if k == 1:
            return sb.material(Material.plastic(rgb_refl(*rng.uniform(0.1, 0.8, 3)), const(0.3),
                                                1.0 / rng.uniform(10.0, 1000.0)))
        if k == 2:
            return sb.material(Material.metal(const(rng.uniform(0.2, 1.5)), const(rng.uniform(2.0, 4.0)),
                                              rng.uniform(0.0, 0.3)))
        return sb.material(Material.glass(1.5, const(1.0), const(1.0)))

    terrain = sb.material(Material.matte(rgb_refl(0.45, 0.5, 0.35)))
    v, f, nrm, uv = heightmap_mesh(fbm_elevation(hm, seed + 11, scale=0.06), hm, hm)
    o2w = T.scale((8.0, 8.0, 8.0)) @ T.translate((-4.0, 0.0, -4.0))
    sb.mesh(v, f, o2w=o2w, material=terrain, normals=nrm, uvs=uv)
    nside = max(3, int(np.sqrt(torus_tris / 2.0)))
    tv, tf = torus(0.35, 0.1, nside, max(3, torus_tris // (2 * nside)))
    for _ in range(n_tori):
        o = (T.rotate_x(rng.uniform(0, 360)) @ T.rotate_y(rng.uniform(0, 360))
             @ T.translate((rng.uniform(-3, 3), rng.uniform(0.6, 1.6), rng.uniform(-3, 3))))
        sb.mesh(tv, tf, o2w=o, material=pick())
    for _ in range(n_spheres):
        r = float(np.exp(rng.uniform(np.log(0.05), np.log(0.3))))
        c = (rng.uniform(-3.5, 3.5), rng.uniform(0.3, 2.0), rng.uniform(-3.5, 3.5))
        sb.shape(abi.SHAPE_SPHERE, (r,), o2w=T.translate(c), material=pick())
    sb.infinite_light(sky_gradient(*sky), w2l=T.rotate_x(90))
    c2w = T.look_at((0.0, 2.6, -5.5), (0.0, 0.6, 0.0), (0, 1, 0))
    sb.look(c2w, 50.0, res, res)
    return sb.build()


# ----------------------------------------------------------------- rays
def primary_rays(flat, res=None):
    """Pinhole camera rays through pixel centres, fireRay (Camera.hs:49-68):
    pCamera = transPoint r2c (x, y, 0); Ray 0 (normalize pCamera) 0 inf; transRay c2w."""
    w, h = (flat.film[0], flat.film[1]) if res is None else (res, res)
    ys, xs = np.mgrid[0:h, 0:w]
    n = w * h
    pr = np.stack([xs.reshape(-1) + 0.5, ys.reshape(-1) + 0.5, np.zeros(n)], 1).astype(F)
    pc = flat.r2c.points(pr)
    sl = (pc[:, 0] * pc[:, 0] + pc[:, 1] * pc[:, 1]).astype(F) + (pc[:, 2] * pc[:, 2]).astype(F)
    sl = sl.astype(F)
    il = (F(1) / np.sqrt(sl)).astype(F)
    d = (pc * il[:, None]).astype(F)
    o = flat.c2w.points(np.zeros((1, 3), F))[0]
    m = flat.c2w.np_m()
    dw = np.stack([((m[4 * r] * d[:, 0]).astype(F) + (m[4 * r + 1] * d[:, 1]).astype(F)).astype(F)
                   + (m[4 * r + 2] * d[:, 2]).astype(F) for r in range(3)], 1).astype(F)
    rays = np.zeros((n, 8), F)
    rays[:, 0:3] = o
    rays[:, 3:6] = dw
    rays[:, 6] = 0.0
    rays[:, 7] = np.inf
    return rays


def incoherent_rays(lo, hi, n, seed=2):
    """Origins uniform in the box [lo, hi], directions uniform on the sphere."""
    rng = np.random.default_rng(seed)
    o = rng.uniform(lo, hi, (n, 3))
    u = rng.uniform(-1.0, 1.0, n)
    phi = rng.uniform(0.0, 2 * np.pi, n)
    s = np.sqrt(1 - u * u)
    d = np.stack([s * np.cos(phi), s * np.sin(phi), u], 1)
    rays = np.zeros((n, 8), F)
    rays[:, 0:3] = o
    rays[:, 3:6] = d
    rays[:, 7] = np.inf
    return rays


def scene_bounds(flat):
    v = flat.vertices
    lo = v.min(axis=0) if len(v) else np.full(3, -1.0)
    hi = v.max(axis=0) if len(v) else np.full(3, 1.0)
    return lo.astype(F), hi.astype(F)
