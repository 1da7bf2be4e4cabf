"""Textures as node lists (bgpu_tex_node): the textures of Texture.hs:59-414
that the material parser builds (IO/MaterialParser.hs:106-226), as values
a material's inputs can take (Material.textured_inputs).

    kd = Tex.gradient(Tex.cell("euclidian", Map3d.identity(T.scale(2, 2, 2))),
                      [(0.0, rgb_refl(0.1, 0.1, 0.1)), (1.0, rgb_refl(0.7, 0.2, 0.7))])
    m = Material.plastic(const(0), const(0.8), 0.002).textured_inputs(spec={0: kd})

A texture is a tree of Tex nodes; `emit` appends a tree to a scene's node
list, children first, and returns the root's index.  Scalar textures
(constant, image, cellNoise, fbm, perlin, crystal, scale) feed scalar inputs
and the weights of spectrum textures (constant, image, blend, checker,
gradient, graphPaper).
"""
import numpy as np

from . import abi

F = np.float32

CELL_DISTANCES = {"euclidian": 0, "euclidian2": 1, "manhattan": 2, "chebyshev": 3}


class Map2d:
    """A 2d texture mapping (Texture.hs:162-179)."""

    def __init__(self, kind, params):
        self.kind, self.params = kind, [float(F(x)) for x in params]

    @staticmethod
    def uv(su, sv, ou, ov):
        """uvMapping (su, sv) (ou, ov): (su u + ou, sv v + ov)."""
        return Map2d(abi.MAP_UV, [su, sv, ou, ov])

    @staticmethod
    def planar(vu, vv, ou, ov):
        """planarMapping (vu, vv) (ou, ov): (p . vu + ou, p . vv + ov)."""
        return Map2d(abi.MAP_PLANAR, list(vu) + list(vv) + [ou, ov])


class Map3d:
    """The 3d identity mapping (Texture.hs:150-156): the hit point by the
    world-to-texture transform (a 4x4 row-major matrix)."""

    def __init__(self, w2t):
        self.kind = abi.MAP_IDENTITY
        m = w2t.m if hasattr(w2t, "m") else w2t  # a transform.Transform, or a matrix
        self.params = [float(F(x)) for x in np.asarray(m, F).reshape(16)]

    @staticmethod
    def identity(w2t=None):
        """identityMapping3d w2t: transPoint w2t of the hit point (the
        transform's own matrix, as the parser hands it over)."""
        return Map3d(np.eye(4, dtype=F) if w2t is None else w2t)


class Tex:
    """One texture node and its children."""

    def __init__(self, op, kids=(), a=0, mapping=None, q=(), spec=None, image=None, steps=None):
        self.op, self.kids, self.a, self.mapping = op, list(kids), a, mapping
        self.q = [float(F(x)) for x in q]
        self.spec = None if spec is None else np.asarray(spec, F).reshape(16)
        self.image = image    # (channels, texels (h, w, channels) float32)
        self.steps = steps    # gradient: [(position, spectrum)] ascending

    @property
    def scalar(self):
        return self.op < abi.TN_CONST

    # -- scalar textures
    @staticmethod
    def constant(v):
        return Tex(abi.TN_SCONST, q=[v])

    @staticmethod
    def image_grey(pixels, mapping):
        """readImageScalarMap: an 8-bit grey image (h, w) over a 2d mapping;
        a texel is x / 255 (Texture.hs:96-104)."""
        g = np.asarray(pixels)
        if g.ndim != 2:
            raise ValueError("a grey image is (height, width)")
        tex = g.astype(F) / F(255)
        return Tex(abi.TN_SIMAGE, mapping=mapping, image=(1, np.ascontiguousarray(tex[..., None])))

    @staticmethod
    def cell(distance, mapping):
        """cellNoise with a distance ("euclidian", "euclidian2", "manhattan",
        "chebyshev") over a 3d mapping (Texture.hs:253-307)."""
        return Tex(abi.TN_CELL, a=CELL_DISTANCES[distance], mapping=mapping)

    @staticmethod
    def fbm(octaves, omega, mapping):
        """texMap3dToTexture (fbm octaves omega) over a 3d mapping."""
        return Tex(abi.TN_FBM, a=int(octaves), q=[omega], mapping=mapping)

    @staticmethod
    def perlin(mapping):
        """noiseTexture over a 3d mapping."""
        return Tex(abi.TN_PERLIN, mapping=mapping)

    @staticmethod
    def crystal(octaves, mapping):
        """quasiCrystal octaves over a 2d mapping (Texture.hs:309-340)."""
        return Tex(abi.TN_CRYSTAL, a=int(octaves), mapping=mapping)

    @staticmethod
    def scale(a, s, t):
        """scaleTexture a s t: a + s * t (Texture.hs:184-185)."""
        if not t.scalar:
            raise ValueError("scale takes a scalar texture")
        return Tex(abi.TN_SCALE, kids=[t], q=[a, s])

    # -- spectrum textures
    @staticmethod
    def spectrum(s):
        return Tex(abi.TN_CONST, spec=s)

    @staticmethod
    def image(spectra, mapping):
        """readImageTextureMap: an image of 16-band texels (h, w, 16), the
        pixelSpectrum of each RGB pixel (image_spectra), over a 2d mapping."""
        t = np.ascontiguousarray(np.asarray(spectra, F))
        if t.ndim != 3 or t.shape[2] != 16:
            raise ValueError("an image texture is (height, width, 16)")
        return Tex(abi.TN_IMAGE, mapping=mapping, image=(16, t))

    @staticmethod
    def blend(t1, t2, f):
        """spectrumBlend t1 t2 f (Texture.hs:133-145)."""
        _spectra(t1, t2)
        _scalars(f)
        return Tex(abi.TN_BLEND, kids=[t1, t2, f])

    @staticmethod
    def checker(scale, t1, t2):
        """checkerBoard (sx, sy, sz) t1 t2 (Texture.hs:209-219)."""
        _spectra(t1, t2)
        return Tex(abi.TN_CHECKER, kids=[t1, t2], q=list(scale))

    @staticmethod
    def gradient(f, steps):
        """gradient (mkGradient steps) f (Texture.hs:221-247): steps are
        (position, spectrum), sorted by position as mkGradient sorts them
        (stable)."""
        _scalars(f)
        if not steps:
            raise ValueError("empty list given to mkGradient")
        st = sorted([(float(F(p)), np.asarray(c, F).reshape(16)) for p, c in steps], key=lambda s: s[0])
        return Tex(abi.TN_GRADIENT, kids=[f], steps=st)

    @staticmethod
    def graph_paper(lw, mapping, t1, t2):
        """graphPaper lw mapping t1 t2 (Texture.hs:191-207): t2 on the lines."""
        _spectra(t1, t2)
        return Tex(abi.TN_GRAPHPAPER, kids=[t1, t2], q=[lw], mapping=mapping)


def _spectra(*ts):
    for t in ts:
        if t.scalar:
            raise ValueError("expected a spectrum texture")


def _scalars(*ts):
    for t in ts:
        if not t.scalar:
            raise ValueError("expected a scalar texture")


def image_spectra(rgb8):
    """pixelSpectrum of every pixel of an 8-bit RGB image (h, w, 3)
    (Texture.hs:84-86): rgbToSpectrumRefl (unGamma (r/255, g/255, b/255)),
    unGamma raising each channel to 2.2 in binary32."""
    from .scene import rgb_refl
    a = np.asarray(rgb8)
    if a.ndim != 3 or a.shape[2] < 3:
        raise ValueError("an RGB image is (height, width, 3)")
    c = np.power(a[..., :3].astype(F) / F(255), F(2.2)).astype(F)
    out = np.empty(a.shape[:2] + (16,), F)
    for y in range(a.shape[0]):
        for x in range(a.shape[1]):
            out[y, x] = rgb_refl(c[y, x, 0], c[y, x, 1], c[y, x, 2])
    return out


def emit(t, nodes, images):
    """Appends texture `t` (children first) to `nodes` (a list of
    abi.TexNode) and its images to `images` (a list of (channels, texels));
    returns the root's index."""
    kids = [emit(k, nodes, images) for k in t.kids]
    n = abi.TexNode()
    n.op = t.op
    if t.image is not None:
        images.append(t.image)
        n.a = len(images) - 1
    else:
        n.a = t.a
    if t.op in (abi.TN_SCALE, abi.TN_BLEND, abi.TN_CHECKER, abi.TN_GRAPHPAPER, abi.TN_GRADIENT):
        n.a = kids[0]
    if t.op in (abi.TN_BLEND, abi.TN_CHECKER, abi.TN_GRAPHPAPER):
        n.b = kids[1]
    if t.op == abi.TN_BLEND:
        n.c = kids[2]
    for k, v in enumerate(t.q):
        n.q[k] = v
    if t.mapping is not None:
        n.map = t.mapping.kind
        for k, v in enumerate(t.mapping.params):
            n.m[k] = v
    if t.spec is not None:
        for k in range(16):
            n.spec[k] = t.spec[k]
    if t.op == abi.TN_GRADIENT:
        first = len(nodes)
        for pos, col in t.steps:
            s = abi.TexNode()
            s.op = abi.TN_GSTEP
            s.q[0] = pos
            for k in range(16):
                s.spec[k] = col[k]
            nodes.append(s)
        n.b, n.c = first, len(t.steps)
    nodes.append(n)
    return len(nodes) - 1
