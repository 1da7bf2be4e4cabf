"""Scene description on the host side: the flattened arrays handed to
`bgpu_scene_create`, built the way bling's `.bling` front end builds its
primitives (src/lib/Graphics/Bling/IO/PrimitiveParser.hs:28-139):

  * meshes are pre-transformed to world space with transPoint/transNormal
    (Primitive/TriangleMesh.hs:39-59);
  * shapes keep their object-to-world transform and its inverse and get an
    AreaLight with a fresh id when an emission is active (Primitive/Geometry.hs:14-37,
    IO/PrimitiveParser.hs:63-67);
  * spectra come from rgbR / rgbI / spd / constant exactly as Spectrum.hs
    converts them (rgbToSpectrum Spectrum.hs:146-159, fromSpd Spectrum.hs:291-335).
"""
import ctypes as C

import numpy as np

from . import abi, textures
from .spectral_tables import TABLES
from .transform import Transform

F = np.float32
BANDS = 16

_T = {k: np.array(v, dtype=np.float32) for k, v in TABLES.items()}


def const(c):
    """sConst (Spectrum.hs:375-377)."""
    return np.full(BANDS, F(c), dtype=np.float32)


def _rgb_to_spectrum(kind, r, g, b):
    """rgbToSpectrum (Spectrum.hs:146-159), elementwise binary32."""
    r, g, b = F(r), F(g), F(b)
    t = lambda n: _T["rgb_%s_%s" % (kind, n)]
    wb, cb, mb, yb = t("white"), t("cyan"), t("magenta"), t("yellow")
    rb, gb, bb = t("red"), t("green"), t("blue")
    if r <= g and r <= b:
        if g <= b:
            rest = (cb * F(g - r)).astype(F) + (bb * F(b - g)).astype(F)
        else:
            rest = (cb * F(b - r)).astype(F) + (gb * F(g - b)).astype(F)
        return ((wb * r).astype(F) + rest.astype(F)).astype(F)
    if g <= r and g <= b:
        if r <= b:
            rest = (mb * F(r - g)).astype(F) + (bb * F(b - r)).astype(F)
        else:
            rest = (mb * F(b - g)).astype(F) + (rb * F(r - b)).astype(F)
        return ((wb * g).astype(F) + rest.astype(F)).astype(F)
    if r <= b:
        rest = (yb * F(r - b)).astype(F) + (gb * F(g - r)).astype(F)
    else:
        rest = (yb * F(g - b)).astype(F) + (rb * F(r - g)).astype(F)
    return ((wb * b).astype(F) + rest.astype(F)).astype(F)


def rgb_illum_array(rgb):
    """rgbToSpectrumIllum of every (r, g, b) row of an (..., 3) float32
    array: (..., 16), the same binary32 operations as _rgb_to_spectrum."""
    a = np.asarray(rgb, F)
    r, g, b = a[..., 0:1], a[..., 1:2], a[..., 2:3]
    t = lambda n: _T["rgb_illum_%s" % n]
    c1 = (r <= g) & (r <= b)
    c2 = ~c1 & (g <= r) & (g <= b)
    c3 = ~c1 & ~c2
    out = np.zeros(a.shape[:-1] + (BANDS,), F)

    def put(mask, w, cw, b1, k1, b2, k2):
        m = mask[..., 0]
        if m.any():
            rest = (t(b1) * k1[m]).astype(F) + (t(b2) * k2[m]).astype(F)
            out[m] = (t("white") * cw[m]).astype(F) + rest.astype(F)

    put(c1 & (g <= b), None, r, "cyan", g - r, "blue", b - g)
    put(c1 & ~(g <= b), None, r, "cyan", b - r, "green", g - b)
    put(c2 & (r <= b), None, g, "magenta", r - g, "blue", b - r)
    put(c2 & ~(r <= b), None, g, "magenta", b - g, "red", r - b)
    put(c3 & (r <= b), None, b, "yellow", r - b, "green", g - r)
    put(c3 & ~(r <= b), None, b, "yellow", g - b, "red", r - g)
    return out


def rgb_refl(r, g, b):
    return _rgb_to_spectrum("refl", r, g, b)


def rgb_illum(r, g, b):
    return _rgb_to_spectrum("illum", r, g, b)


def _lerp(t, a, b):
    return F(F(F(F(1) - t) * a) + F(t * b))


def spd(pairs):
    """fromSpd (mkSpd pairs) for an irregular SPD (Spectrum.hs:199-207,307-314)."""
    pairs = sorted(((F(l), F(v)) for l, v in pairs), key=lambda p: p[0])
    ls = [p[0] for p in pairs]
    vs = [p[1] for p in pairs]
    out = []
    for i in range(BANDS):
        l0 = _lerp(F(F(i) / F(BANDS)), F(400), F(700))
        l1 = _lerp(F(F(i + 1) / F(BANDS)), F(400), F(700))
        if l1 <= ls[0]:
            out.append(vs[0])
            continue
        if l0 >= ls[-1]:
            out.append(vs[-1])
            continue
        i0 = next((k for k, l in enumerate(ls) if l >= l0), 0)
        i1 = next((k for k, l in enumerate(ls) if l >= l1), len(vs) - 1)
        sl = vs[i0:i1 + 1]
        acc = F(0)
        for v in sl:
            acc = F(acc + v)
        out.append(F(acc / F(len(sl))))
    return np.array(out, dtype=np.float32)


def spectrum_to_xyz(s):
    """spectrumToXYZ (Spectrum.hs:349-355), binary32, left fold."""
    x = y = z = F(0)
    for v, a, b, c in zip(np.asarray(s, F), _T["cie_x"], _T["cie_y"], _T["cie_z"]):
        x = F(x + F(a * v))
        y = F(y + F(b * v))
        z = F(z + F(c * v))
    from .spectral_tables import CIE_Y_SUM
    ys = F(CIE_Y_SUM)
    return F(x / ys), F(y / ys), F(z / ys)


def xyz_to_rgb(xyz):
    """xyzToRgb (Spectrum.hs:162-168), vectorised over (..., 3)."""
    x, y, z = xyz[..., 0], xyz[..., 1], xyz[..., 2]
    r = 3.240479 * x - 1.537150 * y - 0.498535 * z
    g = -0.969256 * x + 1.875991 * y + 0.041556 * z
    b = 0.055648 * x - 0.204043 * y + 1.057311 * z
    return np.stack([r, g, b], axis=-1)


class Material:
    """Material parameters in the descriptor's layout (bgpu_material)."""

    def __init__(self, kind, spec0=None, spec1=None, s0=0.0, s1=0.0, spec2=None, s2=0.0):
        self.kind = kind
        self.spec0 = const(0) if spec0 is None else np.asarray(spec0, F)
        self.spec1 = const(0) if spec1 is None else np.asarray(spec1, F)
        self.spec2 = const(0) if spec2 is None else np.asarray(spec2, F)
        self.s0, self.s1, self.s2 = F(s0), F(s1), F(s2)
        # selector texture (bgpu_texture): none by default
        self.tex_kind, self.tex_map, self.tex_slots = abi.TEX_NONE, abi.MAP_UV, 0
        self.tex_params = [0.0] * 12
        self.alt = [const(0), const(0)]
        # inputs taken from node textures (TEX_NODES): slot -> textures.Tex
        self.node_spec, self.node_scalar, self.node_bump = {}, {}, None

    def textured(self, slot, kind, params, alt, mapping=abi.MAP_UV):
        """This material with spectrum slot `slot` (0 or 1) textured: `kind`
        (TEX_CHECKER: params sx sy sz; TEX_GRAPHPAPER: lw and the mapping's
        parameters) selects per hit between the slot's spectrum (tex1) and
        `alt` (tex2)."""
        m = Material(self.kind, self.spec0, self.spec1, self.s0, self.s1, self.spec2, self.s2)
        m.tex_kind, m.tex_map, m.tex_slots = kind, mapping, 1 << slot
        m.tex_params = [float(F(x)) for x in params] + [0.0] * (12 - len(params))
        m.alt = [self.spec0.copy(), self.spec1.copy()]
        m.alt[slot] = np.asarray(alt, F)
        return m

    def textured_inputs(self, spec=None, scalar=None, bump=None):
        """This material with inputs taken from textures (textures.Tex,
        evaluated per hit): `spec` maps spectrum slots (0..2) and `scalar`
        scalar slots (0..3) to textures; the other inputs stay constant.
        `bump`: a scalar texture the material is bump-mapped by
        (bumpMapped, Reflection.hs:344-378)."""
        spec, scalar = dict(spec or {}), dict(scalar or {})
        if not spec and not scalar and bump is None:
            return self
        if any(k not in (0, 1, 2) or t.scalar for k, t in spec.items()) or \
                any(k not in (0, 1, 2, 3) or not t.scalar for k, t in scalar.items()):
            raise ValueError("spectrum slots 0..2 take spectrum textures, scalar slots 0..3 scalar ones")
        if bump is not None and not bump.scalar:
            raise ValueError("a bump map takes a scalar texture")
        m = Material(self.kind, self.spec0, self.spec1, self.s0, self.s1, self.spec2, self.s2)
        m.tex_kind = abi.TEX_NODES
        m.tex_slots = sum(1 << k for k in spec) + sum(1 << (4 + k) for k in scalar) + (256 if bump is not None else 0)
        m.node_spec, m.node_scalar, m.node_bump = spec, scalar, bump
        return m

    @staticmethod
    def matte(kd, sigma=0.0):
        return Material(abi.MAT_MATTE, kd, None, sigma)

    @staticmethod
    def plastic(kd, ks, rough):
        return Material(abi.MAT_PLASTIC, kd, ks, rough)

    @staticmethod
    def metal(eta, k, rough):
        return Material(abi.MAT_METAL, eta, k, rough)

    @staticmethod
    def glass(ior, kr, kt):
        return Material(abi.MAT_GLASS, kr, kt, ior)

    @staticmethod
    def mirror(kr):
        return Material(abi.MAT_MIRROR, kr)

    @staticmethod
    def trans_matte(kr, kt, sigma=0.0):
        """translucentMatte kr kt sigma (Material.hs:43-53)."""
        return Material(abi.MAT_TRANSMATTE, kr, kt, sigma)

    @staticmethod
    def substrate(kd, ks, ka, urough, vrough, depth=0.0):
        """mkSubstrate kd ks ka urough vrough depth (Material.hs:111-127): a
        FresnelBlend coating over a diffuse base with an anisotropic
        microfacet distribution of exponents 1/urough and 1/vrough."""
        return Material(abi.MAT_SUBSTRATE, kd, ks, urough, vrough, ka, depth)

    @staticmethod
    def shiny_metal(kr, ks, rough):
        """mkShinyMetal kr ks rough (Material.hs:98-109): a Blinn microfacet
        and a specular reflection with conductor Fresnel terms approximated
        from ks and kr (frApproxEta, frApproxK, Fresnel.hs:72-78)."""
        return Material(abi.MAT_SHINYMETAL, kr, ks, rough)

    @staticmethod
    def black():
        return Material(abi.MAT_BLACK)


class SceneBuilder:
    """Accumulates primitives like bling's parser state (PState,
    IO/ParserCore.hs:45-58) and flattens them into a bgpu_scene_desc."""

    def __init__(self):
        self.verts = []      # world-space (n, 3) blocks
        self.norms = []
        self.uvs = []
        self.idx = []
        self.tri_mesh = []
        self.meshes = []     # (material, flags)
        self.nvert = 0
        self.ntri = 0
        self.shapes = []     # abi.Shape
        self.materials = []
        self.lights = []     # (abi.Light, radiance array)
        self.next_id = 0
        self.camera = None
        self.film = (256, 256, abi.FILTER_BOX, (0.0, 0.0, 0.0, 0.0))

    def material(self, m):
        self.materials.append(m)
        return len(self.materials) - 1

    def mesh(self, verts, faces, o2w=None, material=0, normals=None, uvs=None, world=False):
        """mkTriangleMesh o2w mat p i n uv: faces are index triples (0-based).
        world=True: points and normals are already what the triangles use
        (a Wavefront file's transformed points and untransformed normals),
        and no transform is applied."""
        if world and o2w is not None:
            raise ValueError("a world-space mesh takes no transform")
        o2w = o2w or Transform()
        v = np.asarray(verts, F).reshape(-1, 3)
        i = np.asarray(faces, np.int64).reshape(-1, 3)
        if (i < 0).any():
            raise ValueError("mkTriangleMesh: contains negative indices")
        if (i >= len(v)).any():
            raise ValueError("mkTriangleMesh: contains out of bounds indices")
        flags = 0
        wv = v.copy() if world else o2w.points(v)
        self.verts.append(wv)
        if normals is not None:
            flags |= abi.MESH_NORMALS
            n = np.asarray(normals, F).reshape(-1, 3)
            self.norms.append(n.copy() if world else o2w.normals(n))
        else:
            self.norms.append(None)
        if uvs is not None:
            flags |= abi.MESH_UVS
            self.uvs.append(np.asarray(uvs, F).reshape(-1, 2))
        else:
            self.uvs.append(None)
        self.idx.append((i + self.nvert).astype(np.int32))
        mid = len(self.meshes)
        self.meshes.append((material, flags))
        self.tri_mesh.append(np.full(len(i), mid, np.int32))
        self.nvert += len(v)
        self.ntri += len(i)
        return mid

    def shape(self, kind, params, o2w=None, material=0, emission=None):
        """mkGeom o2w False mat emission shape id (Primitive/Geometry.hs:14-37)."""
        o2w = o2w or Transform()
        s = abi.Shape()
        s.type = kind
        s.material = material
        p = list(params) + [0.0] * (6 - len(params))
        if kind == abi.SHAPE_BOX:  # mkBox orders the corners (Shape.hs:41-44)
            lo = [min(F(p[k]), F(p[3 + k])) for k in range(3)]
            hi = [max(F(p[k]), F(p[3 + k])) for k in range(3)]
            p = lo + hi
        for k in range(6):
            s.params[k] = F(p[k])
        for k in range(16):
            s.o2w[k] = o2w.m[k]
            s.w2o[k] = o2w.inv[k]
        s.light_id = self.next_id
        self.next_id += 1
        if emission is not None:
            s.emissive = 1
            em = np.asarray(emission, F)
            for k in range(16):
                s.emission[k] = em[k]
        self.shapes.append(s)
        return len(self.shapes) - 1

    def infinite_light(self, radiance_map, w2l=None):
        """mkInfiniteAreaLight rmap t (Light.hs:72-82); radiance_map is
        (height, width, 16) spectral pixels, looked up as rgbfToTexMap does
        (IO/Bitmap.hs:21-28): pixel (x, y) covers u in ((w-1-x)/w, (w-x)/w],
        v likewise, so row 0 holds the largest v (theta near pi) and column 0
        the largest u (phi near 2 pi)."""
        w2l = w2l or Transform()
        m = np.ascontiguousarray(np.asarray(radiance_map, F).reshape(-1, 16))
        h, w = np.asarray(radiance_map).shape[:2]
        l = abi.Light()
        l.type = abi.LIGHT_INFINITE
        l.width, l.height = int(w), int(h)
        for k in range(16):
            l.w2l[k] = w2l.m[k]
            l.l2w[k] = w2l.inv[k]
        self.lights.append((l, m))

    def sun_sky_light(self, east, sun_dir, turbidity, w2l=None):
        """mkInfiniteAreaLight (mkSunSkyLight east sunDir turbidity) t
        (SunSky.hs:12-24): Perez' sky and the attenuated sun, evaluated per
        direction by the library."""
        w2l = w2l or Transform()
        l = abi.Light()
        l.type = abi.LIGHT_INFINITE
        l.map = abi.LIGHTMAP_SUNSKY
        l.width, l.height = 640, 480
        for k in range(16):
            l.w2l[k] = w2l.m[k]
            l.l2w[k] = w2l.inv[k]
        for k, v in enumerate(list(east) + list(sun_dir) + [turbidity]):
            l.sky[k] = float(v)
        self.lights.append((l, None))

    def point_light(self, intensity, position):
        """mkPointLight intensity position (Light.hs:57-61): a delta light
        whose radiance falls off with the squared distance."""
        self._delta_light(abi.LIGHT_POINT, intensity, position)

    def directional_light(self, intensity, normal):
        """mkDirectional intensity normal (Light.hs:53-54): a delta light
        from direction `normal` (normalized by the library)."""
        self._delta_light(abi.LIGHT_DIRECTIONAL, intensity, normal)

    def _delta_light(self, kind, intensity, vec):
        l = abi.Light()
        l.type = kind
        l.width = l.height = 1
        ident = Transform()
        for k in range(16):
            l.w2l[k] = ident.m[k]
            l.l2w[k] = ident.inv[k]
        for k in range(3):
            l.vec[k] = float(F(vec[k]))
        self.lights.append((l, np.ascontiguousarray(np.asarray(intensity, F).reshape(1, 16))))

    def perspective_camera(self, c2w, r2c, lens_radius=0.0, focal_distance=1.0):
        cam = abi.Camera()
        cam.type = abi.CAMERA_PERSPECTIVE
        for k in range(16):
            cam.c2w[k] = c2w.m[k]
            cam.r2c[k] = r2c.m[k]
        cam.lens_radius = F(lens_radius)
        cam.focal_distance = F(focal_distance)
        self.camera = cam
        self.c2w, self.r2c = c2w, r2c

    def environment_camera(self, c2w, width, height):
        """mkEnvironmentCamera c2w sx sy (Camera.hs:70-76,149-155): every
        direction over the raster, at the image size in effect."""
        self.camera = View.environment(c2w, width, height).camera
        self.c2w, self.r2c = c2w, None
        self.lens = ("environment", c2w)

    def set_film(self, width, height, filt=abi.FILTER_BOX, fparams=(0.0, 0.0, 0.0, 0.0)):
        self.film = (int(width), int(height), int(filt), tuple(fparams))

    def look(self, c2w, fov, width, height, lens_radius=0.0, focal_distance=1.0,
             filt=abi.FILTER_BOX, fparams=(0.0, 0.0, 0.0, 0.0)):
        """`camera { perspective ... }` plus `film`: sets both and remembers
        the lens so that views at other image sizes can be made later."""
        from . import transform as T
        self.perspective_camera(*T.perspective_camera(c2w, lens_radius, focal_distance, fov, width, height),
                                lens_radius=lens_radius, focal_distance=focal_distance)
        self.set_film(width, height, filt, fparams)
        self.lens = (c2w, float(fov), float(lens_radius), float(focal_distance), int(filt), tuple(fparams))

    def build(self):
        return FlatScene(self)


class View:
    """A job's camera and film (bgpu_camera, bgpu_film)."""

    def __init__(self, camera, film):
        self.camera, self.film = camera, film

    @property
    def size(self):
        return self.film.width, self.film.height

    @staticmethod
    def perspective(c2w, fov, width, height, lens_radius=0.0, focal_distance=1.0,
                    filt=abi.FILTER_BOX, fparams=(0.0, 0.0, 0.0, 0.0)):
        from . import transform as T
        c2w, r2c = T.perspective_camera(c2w, lens_radius, focal_distance, fov, width, height)
        cam = abi.Camera()
        cam.type = abi.CAMERA_PERSPECTIVE
        for k in range(16):
            cam.c2w[k] = c2w.m[k]
            cam.r2c[k] = r2c.m[k]
        cam.lens_radius = F(lens_radius)
        cam.focal_distance = F(focal_distance)
        film = abi.Film()
        film.width, film.height, film.filter = int(width), int(height), int(filt)
        for k in range(4):
            film.fparams[k] = fparams[k]
        return View(cam, film)

    @staticmethod
    def environment(c2w, width, height, filt=abi.FILTER_BOX, fparams=(0.0, 0.0, 0.0, 0.0)):
        cam = abi.Camera()
        cam.type = abi.CAMERA_ENVIRONMENT
        for k in range(16):
            cam.c2w[k] = c2w.m[k]
        cam.xres, cam.yres = F(width), F(height)
        film = abi.Film()
        film.width, film.height, film.filter = int(width), int(height), int(filt)
        for k in range(4):
            film.fparams[k] = fparams[k]
        return View(cam, film)


def _ptr(a, t):
    return a.ctypes.data_as(C.POINTER(t)) if a is not None else None


class FlatScene:
    """Owns the flattened arrays and the ctypes descriptor pointing at them."""

    def __init__(self, b):
        self.vertices = (np.concatenate(b.verts) if b.verts else np.zeros((0, 3), F)).astype(F)
        self.indices = (np.concatenate(b.idx) if b.idx else np.zeros((0, 3), np.int32)).astype(np.int32)
        self.tri_mesh = (np.concatenate(b.tri_mesh) if b.tri_mesh else np.zeros(0, np.int32)).astype(np.int32)
        nv = len(self.vertices)
        self.normals = None
        if any(n is not None for n in b.norms):
            parts = [n if n is not None else np.zeros((len(v), 3), F) for n, v in zip(b.norms, b.verts)]
            self.normals = np.ascontiguousarray(np.concatenate(parts).astype(F))
        self.uvs = None
        if any(u is not None for u in b.uvs):
            parts = [u if u is not None else np.zeros((len(v), 2), F) for u, v in zip(b.uvs, b.verts)]
            self.uvs = np.ascontiguousarray(np.concatenate(parts).astype(F))
        self.vertices = np.ascontiguousarray(self.vertices)
        self.indices = np.ascontiguousarray(self.indices)
        mats = b.materials or [Material.matte(const(0.5))]
        self.materials = (abi.Material * len(mats))()
        tex_nodes, tex_images = [], []
        for i, m in enumerate(mats):
            self.materials[i].type = m.kind
            self.materials[i].scalar[0] = m.s0
            self.materials[i].scalar[1] = m.s1
            self.materials[i].scalar[2] = m.s2
            for k in range(16):
                self.materials[i].spec[0][k] = m.spec0[k]
                self.materials[i].spec[1][k] = m.spec1[k]
                self.materials[i].spec[2][k] = m.spec2[k]
            t = self.materials[i].tex
            t.kind, t.mapping, t.slots = m.tex_kind, m.tex_map, m.tex_slots
            for k in range(12):
                t.params[k] = m.tex_params[k]
            for k in range(16):
                t.alt[0][k] = m.alt[0][k]
                t.alt[1][k] = m.alt[1][k]
            if m.tex_kind == abi.TEX_NODES:  # roots into the scene's node list
                roots = abi.texture_roots(t)
                for k in range(12):
                    roots[k] = -1
                for k, tex in m.node_spec.items():
                    roots[k] = textures.emit(tex, tex_nodes, tex_images)
                for k, tex in m.node_scalar.items():
                    roots[4 + k] = textures.emit(tex, tex_nodes, tex_images)
                if m.node_bump is not None:
                    roots[8] = textures.emit(m.node_bump, tex_nodes, tex_images)
        meshes = b.meshes or [(0, 0)]
        self.meshes = (abi.Mesh * len(meshes))()
        for i, (mat, fl) in enumerate(meshes):
            self.meshes[i].material = mat
            self.meshes[i].flags = fl
        self.shapes = (abi.Shape * max(1, len(b.shapes)))()
        for i, s in enumerate(b.shapes):
            self.shapes[i] = s
        self.light_maps = [m for _, m in b.lights]
        self.lights = (abi.Light * max(1, len(b.lights)))()
        for i, (l, m) in enumerate(b.lights):
            l.radiance = m.ctypes.data_as(abi.FP) if m is not None else None
            self.lights[i] = l
        d = abi.SceneDesc()
        d.num_vertices = nv
        d.num_triangles = len(self.indices)
        d.vertices = _ptr(self.vertices, C.c_float)
        d.normals = _ptr(self.normals, C.c_float)
        d.uvs = _ptr(self.uvs, C.c_float)
        d.indices = _ptr(self.indices, C.c_int32)
        d.tri_mesh = _ptr(self.tri_mesh, C.c_int32)
        d.num_meshes = len(meshes)
        d.meshes = C.cast(self.meshes, C.POINTER(abi.Mesh))
        d.num_shapes = len(b.shapes)
        d.shapes = C.cast(self.shapes, C.POINTER(abi.Shape))
        d.num_materials = len(mats)
        d.materials = C.cast(self.materials, C.POINTER(abi.Material))
        d.num_lights = len(b.lights)
        d.lights = C.cast(self.lights, C.POINTER(abi.Light))
        # the texture node list and its images (texels kept alive here)
        self.tex_nodes = (abi.TexNode * max(1, len(tex_nodes)))(*tex_nodes)
        self.image_texels = [np.ascontiguousarray(t.reshape(-1), F) for _, t in tex_images]
        self.images = (abi.Image * max(1, len(tex_images)))()
        for i, ((ch, t), flat) in enumerate(zip(tex_images, self.image_texels)):
            self.images[i].width, self.images[i].height, self.images[i].channels = t.shape[1], t.shape[0], ch
            self.images[i].texels = flat.ctypes.data_as(abi.FP)
        d.num_tex_nodes = len(tex_nodes)
        d.tex_nodes = C.cast(self.tex_nodes, C.POINTER(abi.TexNode))
        d.num_images = len(tex_images)
        d.images = C.cast(self.images, C.POINTER(abi.Image))
        if b.camera is not None:
            d.camera = b.camera
        fw, fh, ft, fp = b.film
        d.film.width, d.film.height, d.film.filter = fw, fh, ft
        for k in range(4):
            d.film.fparams[k] = fp[k]
        self.desc = d
        self.num_prims = len(self.indices) + len(b.shapes)
        self.num_triangles = len(self.indices)
        self.num_shapes = len(b.shapes)
        self.c2w = getattr(b, "c2w", None)
        self.r2c = getattr(b, "r2c", None)
        self.film = b.film
        self.lens = getattr(b, "lens", None)

    def save(self, path):
        """Writes the flattened scene as a flat binary scene file
        (bgpu_scene_write); `SceneFile(path)` reads it back."""
        abi.check(abi.load().bgpu_scene_write(C.byref(self.desc), str(path).encode()))

    def view(self, width, height):
        """The scene's camera re-made for a `width` x `height` image (a
        RenderJob's camera follows its imageSize, IO/CameraParser.hs:30-38)."""
        if self.lens is None:
            raise ValueError("scene was not built with SceneBuilder.look")
        if self.lens[0] == "environment":
            ft, fp = self.film[2], self.film[3]
            return View.environment(self.lens[1], width, height, ft, fp)
        c2w, fov, lr, fd, ft, fp = self.lens
        return View.perspective(c2w, fov, width, height, lr, fd, ft, fp)


class SceneFile:
    """A flat binary scene file read by bgpu_scene_read: `desc` is the
    descriptor (its arrays live in the library's block until `close`), and
    `film`, `num_triangles`, `num_shapes` and `num_prims` mirror FlatScene's."""

    def __init__(self, path):
        self._lib = abi.load()
        p = C.POINTER(abi.SceneDesc)()
        abi.check(self._lib.bgpu_scene_read(str(path).encode(), C.byref(p)))
        self._p = p
        self.desc = p.contents
        d = self.desc
        self.film = (d.film.width, d.film.height, d.film.filter, list(d.film.fparams))
        self.num_triangles = d.num_triangles
        self.num_shapes = d.num_shapes
        self.num_prims = d.num_triangles + d.num_shapes

    def array(self, name, dtype, cols):
        """The descriptor's array `name` (vertices, normals, uvs, indices,
        tri_mesh) as a numpy copy, or None."""
        ptr = getattr(self.desc, name)
        if not ptr:
            return None
        n = self.desc.num_vertices if name in ("vertices", "normals", "uvs") else self.desc.num_triangles
        a = np.ctypeslib.as_array(C.cast(ptr, C.POINTER(np.ctypeslib.as_ctypes_type(dtype))), (n * cols,))
        return a.reshape(n, cols).copy() if cols > 1 else a.copy()

    def close(self):
        if self._p:
            abi.check(self._lib.bgpu_scene_desc_free(self._p))
            self._p = None
            self.desc = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()
