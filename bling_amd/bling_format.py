"""Reader for bling's `.bling` scene language, for the part of it the GPU
path renders.

The grammar follows the reference's Parsec parser:

  IO/RenderJob.hs:31-103      top-level objects, imageSize, filter; the start
                              state (640x480, box filter, default camera,
                              identity transform, default matte, no emission)
  IO/ParserCore.hs:104-205    whitespace and `#` comments, numbers, spectra
                              (rgbR, rgbI, spd, temp, `%rrggbb` colours), blocks
  IO/TransformParser.hs       transform / newTransform blocks: rotateX/Y/Z,
                              scale, translate, matrix, lookAt
  IO/CameraParser.hs          perspective and environment cameras (built at parse time with
                              the image size then in effect)
  IO/MaterialParser.hs        matte, plastic, metal, glass, mirror,
                              transMatte, substrate, shinyMetal, blackbody,
                              bumpMap; every texture
                              (constant, image, cellNoise, fbm, perlin,
                              crystal, scale, blend, checker, gradient,
                              graphPaper; uv, planar and identity mappings):
                              checker / graphPaper over constants become the
                              material's selector texture, anything else a
                              node texture (textures.Tex) evaluated per hit
  IO/PrimitiveParser.hs       mesh (polygon faces triangulated as fans,
                              TriangleMesh.hs:23-29), heightMap, waveFront
                              (IO/WaveFront.hs) and shape primitives
  IO/LightParser.hs           infinite lights with a constant, file (RGBE)
                              or sunSky map, point and directional lights;
                              emission
  IO/RendererParser.hs,       renderer blocks (every kind is read; only the
  IO/IntegratorParser.hs      sampler renderer with a stratified sampler and
                              the path or directLighting integrator maps
                              to a GPU job)

The reader builds a `SceneBuilder` in the order `mkScene` sees the
primitives: the parser prepends each statement's primitives
(RenderJob.hs:49-52), so the statements are replayed last to first, each
keeping its own primitive order; explicit lights are prepended the same way
(LightParser.hs:24-25).

Two kinds of error:
  BlingSyntaxError  the reference's parser rejects the text as well;
  Unsupported       valid `.bling` that the GPU path does not render
                    (fractals, Bezier patches), and light map files other
                    than Radiance RGBE.

Numbers are read like GHC's `read` (the exact decimal rounded once to
binary32), so the scene carries the values the Haskell front end computes.
"""
import ctypes
import ctypes.util
import os
import re
from fractions import Fraction

import numpy as np

from . import abi
from . import noise
from . import transform as T
from .scene import Material, SceneBuilder, rgb_illum, rgb_illum_array, rgb_refl, spd
from .textures import CELL_DISTANCES, Map2d, Map3d, Tex, image_spectra

F = np.float32
BANDS = 16

_libm = ctypes.CDLL(ctypes.util.find_library("m") or "libm.so.6")
_libm.expf.restype = ctypes.c_float
_libm.expf.argtypes = [ctypes.c_float]


class BlingSyntaxError(ValueError):
    """Text the reference parser rejects too."""


class Unsupported(ValueError):
    """Valid .bling that the GPU path does not render."""


# ------------------------------------------------------------------ lexing
_TOKEN = re.compile(r"""
    (?P<ws>\s+|\#[^\n]*(?:\n|$))
  | (?P<num>[+-]?\d+(?:\.\d*)?)(?![A-Za-z])
  | (?P<word>[A-Za-z0-9]+)
  | (?P<str>"(?:[^"\\\r\n]|\\[^\r\n])*")
  | (?P<punct>[{},%])
""", re.X)


def _tokens(text):
    out = []
    pos = 0
    line = 1
    while pos < len(text):
        m = _TOKEN.match(text, pos)
        if not m:
            raise BlingSyntaxError("line %d: unexpected %r" % (line, text[pos:pos + 20]))
        kind = m.lastgroup
        if kind != "ws":
            out.append((kind, m.group(kind), line))
        line += m.group(0).count("\n")
        pos = m.end()
    return out


def f32(s):
    """GHC's `read` for Float on a decimal literal: the exact value rounded
    to the nearest binary32 (ties to even)."""
    q = Fraction(s)
    c = F(float(q))
    best = None
    for cand in (np.nextafter(c, F(-np.inf)), c, np.nextafter(c, F(np.inf))):
        if not np.isfinite(cand):
            continue
        d = abs(Fraction(float(cand)) - q)
        key = (d, int(np.array(cand, F).view(np.uint32)) & 1)
        if best is None or key < best[0]:
            best = (key, cand)
    return F(best[1])


def read_hdr(path):
    """A Radiance RGBE (.hdr) image as float RGB rows, (h, w, 3) binary32,
    for `file` light maps: readTexture (IO/Bitmap.hs:13-19) takes what
    JuicyPixels' readImage decodes to ImageRGBF, which for these files is
    the RGBE decode.  JuicyPixels is not part of the reference checkout
    (a Stackage dependency, stack.yaml:18); this follows the published
    format: a text header ending in an empty line, the resolution line
    `-Y h +X w` (rows top to bottom), then scanlines flat, in the old
    run-length form or the newer per-channel run-length form, and a pixel
    (r, g, b, e) decodes to (r, g, b) * 2^(e - 136), or 0 when e = 0."""
    with open(path, "rb") as f:
        data = f.read()
    if not (data.startswith(b"#?RADIANCE") or data.startswith(b"#?RGBE")):
        raise ValueError("not a Radiance RGBE file")
    pos = data.find(b"\n\n")
    if pos < 0:
        raise ValueError("RGBE header without an end")
    head = data[:pos].split(b"\n")
    for line in head:
        if line.startswith(b"FORMAT=") and line.strip() != b"FORMAT=32-bit_rle_rgbe":
            raise Unsupported("light map format %r" % line.decode("latin-1"))
    end = data.find(b"\n", pos + 2)
    res = data[pos + 2:end].split()
    if len(res) != 4 or res[0] != b"-Y" or res[2] != b"+X":
        raise Unsupported("RGBE orientation %r (only -Y h +X w)" % data[pos + 2:end].decode("latin-1"))
    h, w = int(res[1]), int(res[3])
    if h <= 0 or w <= 0:
        raise ValueError("RGBE image of size %d x %d" % (w, h))
    buf = np.frombuffer(data, np.uint8)
    at = end + 1
    out = np.zeros((h, w, 4), np.uint8)
    for y in range(h):
        row = out[y]
        if (8 <= w < 0x8000 and at + 4 <= len(buf) and buf[at] == 2 and buf[at + 1] == 2
                and (int(buf[at + 2]) << 8 | int(buf[at + 3])) == w and buf[at + 2] < 128):
            at += 4
            for c in range(4):  # per-channel runs
                x = 0
                while x < w:
                    n = int(buf[at])
                    if n > 128:
                        n -= 128
                        if x + n > w:
                            raise ValueError("RGBE run past the scanline")
                        row[x:x + n, c] = buf[at + 1]
                        at += 2
                    else:
                        if n == 0 or x + n > w:
                            raise ValueError("bad RGBE literal run")
                        row[x:x + n, c] = buf[at + 1:at + 1 + n]
                        at += 1 + n
                    x += n
            continue
        x, shift = 0, 0
        while x < w:  # flat pixels, (1, 1, 1, n) repeating the previous one n << shift times
            px = buf[at:at + 4]
            if len(px) < 4:
                raise ValueError("RGBE data ends early")
            at += 4
            if px[0] == 1 and px[1] == 1 and px[2] == 1:
                if x == 0:
                    raise ValueError("RGBE repeat at a scanline's start")
                n = int(px[3]) << shift
                if x + n > w:
                    raise ValueError("RGBE repeat past the scanline")
                row[x:x + n] = row[x - 1]
                x += n
                shift += 8
            else:
                row[x] = px
                x += 1
                shift = 0
    e = out[..., 3].astype(np.int32)
    f = np.where(e == 0, F(0), np.ldexp(F(1), e - 136).astype(F)).astype(F)
    return (out[..., :3].astype(F) * f[..., None]).astype(F)


_OBJ_NUM = re.compile(rb"[+-]?[0-9]+(?:\.[0-9]+)?")
_OBJ_INT = re.compile(rb"[0-9]+")


def read_wavefront(data):
    """waveFrontParser (IO/WaveFront.hs:118-211) over the bytes of an .obj
    file: (points (n, 3), normals (m, 3) normalized, uvs (k, 2), corners
    (c, 3) of 0-based (point, uv, normal) indices, -1 where absent, faces
    triangulated as fans, and [(material name, first corner)] in file
    order).  Lines are `v x y z [w]`, `vn x y z`, `vt u [v [w]]` (v
    defaults to 1), `f a[/[t][/n]] ...`, `usemtl name`; anything else is
    skipped.  Numbers are decimals without exponents (bytestring-lexing's
    readDecimal), indices positive; spaces only between fields, lines end
    in a newline.  A malformed v/vn/vt/f line is an error, as it is for the
    reference's parser."""
    pts, nrm, uvs, corners, mtls = [], [], [], [], []

    def nums(fields, lo, hi, line):
        if not (lo <= len(fields) <= hi) or not all(_OBJ_NUM.fullmatch(x) for x in fields):
            raise ValueError("bad line %r" % line[:60])
        return [f32(x.decode()) for x in fields]

    lines = data.split(b"\n")
    if lines and lines[-1] == b"":
        lines.pop()
    elif lines:  # the last line lacks its newline: v, vn, vt and f need one
        last = lines[-1]
        if last[:1] in (b"v", b"f") and not last.startswith(b"vp"):
            raise ValueError("line without a newline: %r" % last[:60])
    for line in lines:
        if line.startswith(b"vn"):
            x, y, z = nums([t for t in line[2:].split(b" ") if t] if line[2:3] == b" " else [b"?"], 3, 3, line)
            sq = F(F(F(x * x) + F(y * y)) + F(z * z))  # normalize (Math.hs:349-353)
            if sq != 0:
                il = F(F(1) / F(np.sqrt(sq)))
                nrm.append((F(x * il), F(y * il), F(z * il)))
            else:
                nrm.append((F(0), F(1), F(0)))
        elif line.startswith(b"vt"):
            f = nums([t for t in line[2:].split(b" ") if t] if line[2:3] == b" " else [b"?"], 1, 3, line)
            uvs.append((f[0], f[1] if len(f) > 1 else F(1)))
        elif line.startswith(b"usemtl"):
            if line[6:7] != b" ":
                raise ValueError("bad line %r" % line[:60])
            mtls.append((line[6:].lstrip(b" ").decode("latin-1"), len(corners)))
        elif line.startswith(b"v"):
            if line[1:2] != b" ":
                continue  # `vp` and the like: skipped
            f = nums([t for t in line[1:].split(b" ") if t], 3, 4, line)
            pts.append(tuple(f[:3]))
        elif line.startswith(b"f"):
            if line[1:2] != b" ":
                raise ValueError("bad line %r" % line[:60])
            face = []
            for tok in [t for t in line[1:].split(b" ") if t]:
                parts = tok.split(b"/")
                if len(parts) > 3 or not _OBJ_INT.fullmatch(parts[0]) or \
                        any(p and not _OBJ_INT.fullmatch(p) for p in parts[1:]) or (len(parts) == 3 and not parts[2]):
                    raise ValueError("bad face %r" % line[:60])
                t_ = int(parts[1]) if len(parts) > 1 and parts[1] else 0
                n_ = int(parts[2]) if len(parts) > 2 else 0
                face.append((int(parts[0]) - 1, t_ - 1, n_ - 1))  # pred
            if not face:
                raise ValueError("bad face %r" % line[:60])
            for k in range(1, len(face) - 1):  # triangulate (Primitive/TriangleMesh.hs:23-29)
                corners += [face[0], face[k], face[k + 1]]
    return (np.array(pts, F).reshape(-1, 3), np.array(nrm, F).reshape(-1, 3), np.array(uvs, F).reshape(-1, 2),
            np.array(corners, np.int64).reshape(-1, 3), mtls)


def wavefront_triangles(corners, mtls):
    """The triangles parseWaveFront makes (IO/WaveFront.hs:101-116): the
    material intervals start at ("default", 0) and at each usemtl's corner
    count, and end at the next start, the last one at the number of
    TRIANGLES (cnt = corners / 3, mixing units as the reference does); each
    interval makes a triangle at every third corner index from its start
    while below start + length.  Returns [(material name, first corner)]."""
    cnt = len(corners) // 3
    starts = [("default", 0)] + list(mtls)
    ends = [c for _, c in mtls] + [cnt]
    out = []
    for (name, st), en in zip(starts, ends):
        ln = en - st
        if ln > 0:
            out += [(name, i) for i in range(st, st + ln, 3)]
    return out


def _powi5(x):
    """x ^ 5 as GHC's (^) computes it: ((x*x)*(x*x))*x."""
    x2 = F(x * x)
    return F(F(x2 * x2) * x)


def blackbody(temp):
    """sBlackBody t (Spectrum.hs:480-493): fromSpd of the Planck function;
    avgSpd of a function averages its values at the band edges."""
    t = F(temp)

    def planck(w):
        w1 = F(F(w) * F(1e-9))
        num = F(F(0.4e-9) * F(F(3.74183e-16) * F(F(1) / _powi5(w1))))
        den = F(F(_libm.expf(float(F(F(1.4388e-2) / F(w1 * t))))) - F(1))
        return F(num / den)

    out = []
    for i in range(BANDS):
        l0 = _lerp(F(F(i) / F(BANDS)), F(400), F(700))
        l1 = _lerp(F(F(i + 1) / F(BANDS)), F(400), F(700))
        out.append(F(F(planck(l0) + planck(l1)) * F(0.5)))
    return np.array(out, F)


def _lerp(t, a, b):  # Math.hs lerp: (1 - t) * a + t * b
    return F(F(F(F(1) - t) * a) + F(t * b))


def _hmin(x, y):  # Haskell `min` on Float
    return x if x <= y else y


def _hmax(x, y):
    return y if x <= y else x


def _clamp(v, lo, hi):  # Math.hs:84-87
    return lo if v < lo else (hi if v > hi else v)


# ----------------------------------------------------------------- parsing
def heightmap_mesh(elev, ns, nt):
    """heightMap elev (ns, nt) (Primitive/Heightmap.hs:14-47) in binary32:
    vertices (x, elev(x, z), z) on the (ns x nt) lattice of [0, 1]^2 (z outer),
    two triangles per cell, shading normals -normalize(dx, ex + ez, dz) from
    central differences with ex = 1 / ns, ez = 1 / nt, and uv
    (x / (ns - 1), z / (ns - 1)) -- both divided by ns - 1, as the reference
    writes it.  Returns (vertices, faces, normals, uvs) in object space."""
    fns, fnt = F(ns), F(nt)
    xs = np.arange(ns, dtype=F) / (fns - F(1))
    zs = np.arange(nt, dtype=F) / (fnt - F(1))
    cx = np.tile(xs, nt)
    cz = np.repeat(zs, ns)
    verts = np.stack([cx, elev(cx, cz), cz], 1).astype(F)
    ex, ez = F(1) / fns, F(1) / fnt
    dx = elev(cx - ex, cz) - elev(cx + ex, cz)
    dz = elev(cx, cz - ez) - elev(cx, cz + ez)
    dy = np.full_like(dx, ex + ez)
    sq = (dx * dx + dy * dy) + dz * dz
    inv = F(1) / np.sqrt(sq)
    nrm = np.where((sq != 0)[:, None], np.stack([dx * inv, dy * inv, dz * inv], 1),
                   np.array([0, 1, 0], F)[None, :])
    nrm = (-nrm).astype(F)
    uvs = np.stack([cx / (fns - F(1)), cz / (fns - F(1))], 1).astype(F)
    x, y = np.meshgrid(np.arange(ns - 1), np.arange(nt - 1))
    x, y = x.reshape(-1), y.reshape(-1)
    v = lambda a, b: a + b * ns
    faces = np.stack([v(x, y), v(x + 1, y), v(x + 1, y + 1), v(x, y), v(x + 1, y + 1), v(x, y + 1)], 1)
    return verts, faces.reshape(-1, 3).astype(np.int64), nrm, uvs


class _Parser:
    def __init__(self, text, base="."):
        self.toks = _tokens(text)
        self.i = 0
        self.base = base  # basePath: the scene file's folder (RenderJob.hs:31-34)
        # start state (RenderJob.hs:19-29)
        self.res = (640, 480)
        self.renderer = ("sampler", ("stratified", 2, 2), ("path", 7, 3))
        self.filter = (abi.FILTER_BOX, (0.0, 0.0, 0.0, 0.0))
        self.camera = self._perspective(T.translate((0, 0, -5)), F(0), F(1), F(90), 640, 480)
        self.transform = T.Transform()
        self.material = Material.matte(rgb_refl(0.9, 0.9, 0.9), 0.0)
        self.emit = None
        self.lights = []  # in reference order (prepended)
        self.prims = []   # statements in parse order
        self.curr_id = 0

    # -- token helpers
    def peek(self, k=0):
        j = self.i + k
        return self.toks[j] if j < len(self.toks) else ("eof", "", -1)

    def fail(self, msg):
        line = self.peek()[2]
        raise BlingSyntaxError("line %s: %s" % (line if line >= 0 else "end", msg))

    def take(self, kind=None, value=None):
        t = self.peek()
        if t[0] == "eof":
            self.fail("unexpected end of input")
        if (kind and t[0] != kind) or (value is not None and t[1] != value):
            self.fail("expected %s, found %r" % (value or kind, t[1]))
        self.i += 1
        return t[1]

    def word(self):
        t = self.peek()
        if t[0] == "word" or (t[0] == "num" and t[1].isdigit()):
            self.i += 1
            return t[1]
        self.fail("expected a name, found %r" % (t[1],))

    def flt(self):
        return f32(self.take("num"))

    def integ(self):
        s = self.take("num")
        if not s.isdigit():
            self.fail("expected a non-negative integer, found %r" % s)
        return int(s)

    def vec(self):
        return (self.flt(), self.flt(), self.flt())

    def open(self):
        self.take("punct", "{")

    def close(self):
        self.take("punct", "}")

    def named(self, name):
        self.take("word", name)

    def named_flt(self, name):
        self.named(name)
        return self.flt()

    def named_int(self, name):
        self.named(name)
        return self.integ()

    def named_vec(self, name):
        self.named(name)
        return self.vec()

    # -- spectra (ParserCore.hs:133-173)
    def spectrum(self):
        t = self.word()
        if t in ("rgbR", "rgbI"):
            r, g, b = self.rgb_value()
            return rgb_refl(r, g, b) if t == "rgbR" else rgb_illum(r, g, b)
        if t == "spd":
            self.open()
            pairs = [(self.flt(), self.flt())]
            while self.peek()[1] == ",":
                self.take("punct", ",")
                pairs.append((self.flt(), self.flt()))
            self.close()
            return spd(pairs)
        if t == "temp":
            return blackbody(self.flt())
        self.fail("unknown spectrum type " + t)

    def rgb_value(self):
        if self.peek()[1] == "%":
            self.take("punct", "%")
            digits = ""
            while len(digits) < 6:
                t = self.peek()
                if t[0] not in ("word", "num") or not re.fullmatch(r"[0-9A-Fa-f]+", t[1]):
                    self.fail("expected hex digits")
                digits += self.take()
            if len(digits) != 6:
                self.fail("expected three hex pairs")
            return tuple(F(F(int(digits[k:k + 2], 16)) / F(255)) for k in (0, 2, 4))
        return self.vec()

    # -- transforms (TransformParser.hs)
    def transform_block(self):
        self.open()
        acc = T.Transform()
        while True:
            t = self.peek()
            if t[1] in ("rotateX", "rotateY", "rotateZ"):
                self.take()
                deg = self.flt()
                acc = acc @ {"rotateX": T.rotate_x, "rotateY": T.rotate_y, "rotateZ": T.rotate_z}[t[1]](deg)
            elif t[1] == "scale":
                self.take()
                acc = acc @ T.scale(self.vec())
            elif t[1] == "translate":
                self.take()
                acc = acc @ T.translate(self.vec())
            elif t[1] == "matrix":
                self.take()
                self.open()
                rows = []
                for _ in range(4):
                    self.take("word", "m")
                    rows.append([self.flt() for _ in range(4)])
                self.close()
                acc = acc @ T.from_matrix(rows)
            elif t[1] == "lookAt":
                self.take()
                self.open()
                pos = self.named_vec("pos")
                look = self.named_vec("look")
                up = self.named_vec("up")
                self.close()
                acc = acc @ T.look_at(pos, look, up)
            else:
                break
        self.close()
        return acc

    # -- camera (CameraParser.hs)
    @staticmethod
    def _perspective(c2w, lr, fd, fov, sx, sy):
        cw, r2c = T.perspective_camera(c2w, lr, fd, fov, sx, sy)
        return ("perspective", cw, r2c, lr, fd, fov)

    def camera_block(self):
        self.open()
        t = self.word()
        if t == "perspective":
            fov = self.named_flt("fov")
            lr = self.named_flt("lensRadius")
            fd = self.named_flt("focalDistance")
            self.camera = self._perspective(self.transform, lr, fd, fov, self.res[0], self.res[1])
        elif t == "environment":  # pEnvironmentCam (CameraParser.hs:40-44)
            self.camera = ("environment", self.transform, self.res)
        else:
            self.fail("unknown camera type " + t)
        self.close()

    # -- textures and materials (MaterialParser.hs)
    def qstring(self):
        """pQString (ParserCore.hs:87-101): the text between the quotes, a
        backslash pair kept as written."""
        return self.take("str")[1:-1]

    def spectrum_texture(self, name):
        """pSpectrumTexture (MaterialParser.hs:189-228) as a textures.Tex
        tree; a blend of constants by a constant folds to a constant."""
        self.named(name)
        self.open()
        t = self.word()
        if t == "constant":
            v = Tex.spectrum(self.spectrum())
        elif t == "blend":
            t1 = self.spectrum_texture("tex1")
            t2 = self.spectrum_texture("tex2")
            f = self.scalar_texture("f")
            if t1.op == abi.TN_CONST and t2.op == abi.TN_CONST and f.op == abi.TN_SCONST:
                a, b, w = t1.spec, t2.spec, F(f.q[0])
                if w <= F(0):  # spectrumBlend (Texture.hs:135-145)
                    v = t1
                elif w >= F(1):
                    v = t2
                else:
                    v = Tex.spectrum(np.array([F(F(x * F(F(1) - w)) + F(y * w)) for x, y in zip(a, b)], F))
            else:
                v = Tex.blend(t1, t2, f)
        elif t == "image":
            self.open()
            v = self.image_texture(grey=False)
            self.close()
        elif t == "checker":
            scale = self.vec()
            v = Tex.checker(scale, self.spectrum_texture("tex1"), self.spectrum_texture("tex2"))
        elif t == "gradient":
            f = self.scalar_texture("f")
            self.named("steps")
            self.open()
            steps = []
            if self.peek()[0] == "num":  # sepBy: zero or more (pos, spectrum)
                steps.append((self.flt(), self.spectrum()))
                while self.peek()[1] == ",":
                    self.take("punct", ",")
                    steps.append((self.flt(), self.spectrum()))
            self.close()
            if not steps:
                self.fail("empty list given to mkGradient")
            v = Tex.gradient(f, steps)
        elif t == "graphPaper":
            lw = self.flt()
            mapping = self.mapping_2d("map")
            v = Tex.graph_paper(lw, mapping, self.spectrum_texture("tex1"), self.spectrum_texture("tex2"))
        else:
            self.fail("unknown texture type " + t)
        self.close()
        return v

    def mapping_2d(self, name):
        """pTextureMapping2d (MaterialParser.hs:159-178)."""
        self.named(name)
        self.open()
        t = self.word()
        if t == "uv":
            m = Map2d.uv(self.flt(), self.flt(), self.flt(), self.flt())
        elif t == "planar":
            vu, vv = self.vec(), self.vec()
            m = Map2d.planar(vu, vv, self.flt(), self.flt())
        else:
            self.fail("unknown 2d mapping " + t)
        self.close()
        return m

    def mapping_3d(self, name):
        """pTextureMapping3d (MaterialParser.hs:180-187): `identity` and the
        world-to-texture transform block."""
        self.named(name)
        self.open()
        t = self.word()
        if t != "identity":
            self.fail("unknown 3d mapping " + t)
        m = Map3d.identity(self.transform_block())
        self.close()
        return m

    def image_texture(self, grey):
        """pImageTexture / pImageScalar (MaterialParser.hs:106-113, 189-196):
        `file "name" map {...}`, the file relative to the scene file's
        folder, decoded as readImageTextureMap / readImageScalarMap
        (Texture.hs:106-124) accept it: RGB, RGBA (alpha dropped) or YCbCr
        (converted to RGB) 8-bit images for spectra, 8-bit grey images for
        scalars."""
        self.named("file")
        name = self.qstring()
        path = os.path.join(self.base, name)
        try:
            from PIL import Image
        except ImportError:
            raise Unsupported("image textures need PIL to decode %r" % name)
        try:
            with Image.open(path) as im:
                im.load()
                mode, bits = im.mode, getattr(im, "bits", 8)
                if grey:
                    if mode != "L":
                        self.fail("unsupported image type %s for a scalar image texture (%r)" % (mode, name))
                    px = np.asarray(im, np.uint8)
                else:
                    if mode == "P":
                        im = im.convert("RGBA" if "transparency" in im.info else "RGB")
                        mode = im.mode
                    if mode not in ("RGB", "RGBA", "YCbCr") or bits != 8:
                        self.fail("unsupported image type %s for an image texture (%r)" % (mode, name))
                    px = np.asarray(im.convert("RGB") if mode == "YCbCr" else im, np.uint8)[..., :3]
        except OSError as e:
            self.fail("cannot read image %r: %s" % (name, e))
        mapping = self.mapping_2d("map")
        return Tex.image_grey(px, mapping) if grey else Tex.image(image_spectra(px), mapping)

    def scalar_texture(self, name):
        """pScalarTexture (MaterialParser.hs:115-157) as a textures.Tex tree;
        scale of a constant folds to a constant (a + s * v in binary32)."""
        self.named(name)
        self.open()
        t = self.word()
        if t == "constant":
            v = Tex.constant(self.flt())
        elif t == "image":
            self.open()
            v = self.image_texture(grey=True)
            self.close()
        elif t == "cellNoise":
            d = self.word()
            if d not in CELL_DISTANCES:
                self.fail("unknown distance function " + d)
            v = Tex.cell(d, self.mapping_3d("map"))
        elif t == "fbm":
            octaves = self.named_int("octaves")
            omega = self.named_flt("omega")
            v = Tex.fbm(octaves, omega, self.mapping_3d("map"))
        elif t == "perlin":
            v = Tex.perlin(self.mapping_3d("map"))
        elif t == "crystal":
            octaves = self.named_int("octaves")
            v = Tex.crystal(octaves, self.mapping_2d("map"))
        elif t == "scale":
            a = self.flt()
            sc = self.flt()
            inner = self.scalar_texture("tex")
            if inner.op == abi.TN_SCONST:  # scaleTexture (Texture.hs:184-185)
                v = Tex.constant(F(F(a) + F(F(sc) * F(inner.q[0]))))
            else:
                v = Tex.scale(a, sc, inner)
        else:
            self.fail("unknown texture type " + t)
        self.close()
        return v

    @staticmethod
    def _selector(t):
        """A checker or graph paper over two constants, as the material's
        selector texture (kind, mapping, params, tex1, tex2); None otherwise."""
        if t.op not in (abi.TN_CHECKER, abi.TN_GRAPHPAPER) or \
                any(k.op != abi.TN_CONST for k in t.kids):
            return None
        if t.op == abi.TN_CHECKER:
            return ("checker", abi.MAP_UV, list(t.q[:3]), t.kids[0].spec, t.kids[1].spec)
        return ("graphpaper", t.mapping.kind, [t.q[0]] + list(t.mapping.params), t.kids[0].spec, t.kids[1].spec)

    @classmethod
    def make_material(cls, make, spectra, scalars, bump=None):
        """The material `make(spectra, scalars)` with its inputs from
        textures.  Constants stay constants; when every textured input is a
        checker or graph paper over constants in spectrum slot 0 or 1 and all
        of them use the same selector, the material takes the selector
        texture (bgpu_texture); any other texture, or a bump map, makes the
        material's textured inputs node textures, evaluated per hit."""
        zero = np.zeros(BANDS, F)
        sel = [cls._selector(t) for t in spectra]
        simple = bump is None and all(t.op == abi.TN_SCONST for t in scalars) and \
            all(t.op == abi.TN_CONST or (sel[k] is not None and k < 2) for k, t in enumerate(spectra))
        if simple:
            tex = [v for v in sel if v is not None]
            if tex and any(v[:3] != tex[0][:3] for v in tex[1:]):
                simple = False
        if not simple:
            m = make([t.spec if t.op == abi.TN_CONST else zero for t in spectra],
                     [t.q[0] if t.op == abi.TN_SCONST else 0.0 for t in scalars])
            return m.textured_inputs(spec={k: t for k, t in enumerate(spectra) if t.op != abi.TN_CONST},
                                     scalar={k: t for k, t in enumerate(scalars) if t.op != abi.TN_SCONST},
                                     bump=bump)
        m = make([v[3] if v is not None else t.spec for v, t in zip(sel, spectra)], [t.q[0] for t in scalars])
        if not tex:
            return m
        kind, mapping, params = tex[0][:3]
        bits = sum(1 << k for k, v in enumerate(sel) if v is not None)
        alt = [v[4] if v is not None else t.spec for v, t in zip(sel, spectra)]
        code = abi.TEX_CHECKER if kind == "checker" else abi.TEX_GRAPHPAPER
        out = m.textured(0, code, params, alt[0], mapping) if bits & 1 else m.textured(1, code, params, alt[1], mapping)
        if bits == 3:
            out.tex_slots = 3
            out.alt = [np.asarray(alt[0], F), np.asarray(alt[1], F)]
        return out

    def material_body(self, bump=None):
        """pMaterial' (MaterialParser.hs:30-104): each material's spectrum
        and scalar inputs in the slots of bgpu_material; `bumpMap bump {d}
        <material>` bump-maps the material it wraps (pBumpMap, :44-48)."""
        t = self.word()
        if t == "blackbody":
            return Material.black()
        if t == "bumpMap":
            d = self.scalar_texture("bump")
            if bump is not None:
                raise Unsupported("a bump map over a bump-mapped material")
            return self.material_body(bump=d)
        st, sc = self.spectrum_texture, self.scalar_texture

        def mk(make, spectra, scalars):
            return self.make_material(make, spectra, scalars, bump)
        if t == "matte":
            kd = st("kd")
            return mk(lambda s, x: Material.matte(s[0], x[0]), [kd], [sc("sigma")])
        if t == "plastic":
            kd, ks = st("kd"), st("ks")
            return mk(lambda s, x: Material.plastic(s[0], s[1], x[0]), [kd, ks], [sc("rough")])
        if t == "metal":
            eta, k = st("eta"), st("k")
            return mk(lambda s, x: Material.metal(s[0], s[1], x[0]), [eta, k], [sc("rough")])
        if t == "glass":
            ior = sc("ior")
            kr, kt = st("kr"), st("kt")
            return mk(lambda s, x: Material.glass(x[0], s[0], s[1]), [kr, kt], [ior])
        if t == "mirror":
            return mk(lambda s, x: Material.mirror(s[0]), [st("kr")], [])
        if t == "transMatte":  # pMatteTranslucent (MaterialParser.hs:92-97)
            kr, kt = st("kr"), st("kt")
            return mk(lambda s, x: Material.trans_matte(s[0], s[1], x[0]), [kr, kt], [sc("ks")])
        if t == "substrate":  # pSubstrateMaterial (MaterialParser.hs:50-58)
            kd, ks, ka = st("kd"), st("ks"), st("ka")
            ur, vr, depth = sc("urough"), sc("vrough"), sc("depth")
            return mk(lambda s, x: Material.substrate(s[0], s[1], s[2], x[0], x[1], x[2]),
                      [kd, ks, ka], [ur, vr, depth])
        if t == "shinyMetal":  # pShinyMetal (MaterialParser.hs:74-79)
            kr, ks = st("kr"), st("ks")
            return mk(lambda s, x: Material.shiny_metal(s[0], s[1], x[0]), [kr, ks], [sc("rough")])
        self.fail("unknown material type " + t)

    def scalar_map_2d(self):
        """pScalarMap2d (MaterialParser.hs:232-245): `fbm z octaves n omega w`
        (texMap3dTo2d (fbm n w) z) or `scale f map`; returns elev(x, y) over
        binary32 arrays."""
        self.open()
        t = self.word()
        if t == "fbm":
            z = self.flt()
            octaves = self.named_int("octaves")
            omega = self.named_flt("omega")

            def elev(x, y):
                return noise.fbm(octaves, omega, x, y, np.full(np.shape(x), z, F))
        elif t == "scale":
            k = F(self.flt())
            inner = self.scalar_map_2d()

            def elev(x, y):
                return k * inner(x, y)
        else:
            self.fail("unknown scalar map type " + t)
        self.close()
        return elev

    # -- primitives (PrimitiveParser.hs)
    def prim_block(self):
        self.open()
        t = self.word()
        if t == "mesh":
            vc = self.named_int("vertexCount")
            fc = self.named_int("faceCount")
            verts = []
            for _ in range(vc):
                self.take("word", "v")
                verts.append(self.vec())
            flat = []
            for _ in range(fc):
                self.take("word", "f")
                face = [self.integ()]
                while self.peek()[0] == "num" and self.peek()[1].isdigit():
                    face.append(self.integ())
                f0, rest = face[0], face[1:]  # triangulate (TriangleMesh.hs:23-29)
                for k in range(len(rest) - 1):
                    flat += [f0, rest[k], rest[k + 1]]
            stmt = [("mesh", np.array(verts, F).reshape(-1, 3), np.array(flat, np.int64).reshape(-1, 3),
                     self.transform, self.material, None, None)]
        elif t == "heightMap":  # PrimitiveParser.hs:39-45: its own transform block
            ns = self.integ()
            nt = self.integ()
            if ns < 2 or nt < 2:  # the lattice divides by ns - 1 and nt - 1
                raise Unsupported("heightMap with fewer than 2 x 2 samples")
            elev = self.scalar_map_2d()
            xf = self.transform_block()
            verts, faces, norms, uvs = heightmap_mesh(elev, ns, nt)
            stmt = [("mesh", verts, faces, xf, self.material, norms, uvs)]
        elif t == "shape":
            kind, params = self.shape_block()
            sid = self.curr_id
            self.curr_id += 1
            stmt = [("shape", kind, params, self.transform, self.material, self.emit, sid)]
        elif t == "waveFront":  # PrimitiveParser.hs:69-72, IO/WaveFront.hs
            stmt = self.wavefront()
        elif t in ("julia", "mandelbulb", "bezier"):
            raise Unsupported("primitive %r" % t)
        else:
            self.fail("unknown primitive type " + t)
        self.close()
        self.prims.append(stmt)

    def wavefront(self):
        """`waveFront "file.obj" materials { "name" { <material> } ... }`:
        the file relative to the scene's folder (readFileBS), the named
        materials with the current one as the default (pNamedMaterialMap,
        PrimitiveParser.hs:76-84).  Points go through the current transform;
        normals and uvs are used as the file gives them (the triangles shade
        with an identity transform, Primitive.hs:65), and a triangle takes
        its third corner's uv and normal index from its second corner
        (wfTriUVs / wfTriNormals read corner i + 1 twice, WaveFront.hs:52-79)."""
        name = self.qstring()
        self.named("materials")
        self.open()
        mats = {}
        while self.peek()[0] == "str":
            mn = self.qstring()
            self.open()
            mats[mn] = self.material_body()
            self.close()
        self.close()
        path = os.path.join(self.base, name)
        try:
            with open(path, "rb") as f:
                data = f.read()
        except OSError as e:
            self.fail("cannot read %r: %s" % (name, e))
        try:
            pts, nrm, uvs, corners, mtls = read_wavefront(data)
        except ValueError as e:
            self.fail("%s: %s" % (name, e))
        tris = wavefront_triangles(corners, mtls)
        if not tris:
            return []
        xf = self.transform
        wp = xf.points(pts) if len(pts) else pts
        runs = []  # (material, has uv, has normals) runs of consecutive triangles
        for mname, c in tris:
            if c + 2 >= len(corners):
                self.fail("%s: a material interval runs past the faces" % name)
            (p0, t0, n0), (p1, t1, n1), (p2, _, _) = corners[c], corners[c + 1], corners[c + 2]
            ti = (t0, t1, t1)
            ni = (n0, n1, n1)
            if max(p0, p1, p2) >= len(pts):
                self.fail("%s: point index out of range" % name)
            has_uv = min(ti) >= 0
            has_n = not (max(ni) < 0)
            if has_n and (min(ni) < 0 or max(ni) >= len(nrm)):
                self.fail("%s: normal index out of range" % name)
            if has_uv and max(ti) >= len(uvs):
                self.fail("%s: uv index out of range" % name)
            m = mats.get(mname, self.material)
            key = (id(m), has_uv, has_n)
            if not runs or runs[-1][0] != key:
                runs.append((key, m, has_uv, has_n, []))
            runs[-1][4].append(((p0, p1, p2), ti, ni))
        out = []
        for _, m, has_uv, has_n, ts in runs:
            v = np.concatenate([wp[list(p)] for p, _, _ in ts])
            faces = np.arange(3 * len(ts), dtype=np.int64).reshape(-1, 3)
            uv = np.concatenate([uvs[list(t)] for _, t, _ in ts]) if has_uv else None
            nn = np.concatenate([nrm[list(n)] for _, _, n in ts]) if has_n else None
            out.append(("objmesh", v, faces, m, nn, uv))
        return out

    def shape_block(self):
        self.open()
        t = self.word()
        if t == "box":
            lo = self.named_vec("pmin")
            hi = self.named_vec("pmax")
            kind, params = abi.SHAPE_BOX, list(lo) + list(hi)
        elif t == "cylinder":  # mkCylinder (Shape.hs:46-56)
            r = self.named_flt("radius")
            z0 = self.named_flt("zmin")
            z1 = self.named_flt("zmax")
            pm = self.named_flt("phiMax")
            kind = abi.SHAPE_CYLINDER
            params = [r, _hmin(z0, z1), _hmax(z0, z1), T.radians(_clamp(pm, F(0), F(360)))]
        elif t == "disk":  # mkDisk (Shape.hs:58-68)
            h = self.named_flt("height")
            r0 = self.named_flt("radius")
            r1 = self.named_flt("innerRadius")
            pm = self.named_flt("phiMax")
            kind = abi.SHAPE_DISK
            params = [h, _hmax(r0, r1), _hmin(r0, r1), T.radians(_clamp(pm, F(0), F(360)))]
        elif t == "quad":
            kind, params = abi.SHAPE_QUAD, [self.flt(), self.flt()]
        elif t == "sphere":
            kind, params = abi.SHAPE_SPHERE, [self.named_flt("radius")]
        else:
            self.fail("unknown shape type " + t)
        self.close()
        return kind, params

    # -- lights (LightParser.hs)
    def light_block(self):
        self.open()
        t = self.word()
        if t == "infinite":
            w2l = self.transform_block()
            self.named("l")
            self.open()
            m = self.word()
            if m == "constant":  # constSpectrumMap2d: a 1 x 1 map (Texture.hs:131-132)
                s = self.spectrum()
                self.lights.insert(0, ("infinite", np.asarray(s, F).reshape(1, 1, BANDS), w2l))
            elif m == "file":  # readTexture: rgbToSpectrumIllum per pixel (IO/Bitmap.hs:13-28)
                name = self.qstring()
                path = os.path.join(self.base, name)
                try:
                    rgb = read_hdr(path)
                except OSError as e:
                    self.fail("cannot read light map %r: %s" % (name, e))
                except ValueError as e:
                    if isinstance(e, Unsupported):
                        raise
                    self.fail("can't convert image format to texture (%r: %s)" % (name, e))
                self.lights.insert(0, ("infinite", rgb_illum_array(rgb), w2l))
            elif m == "sunSky":  # mkSunSkyLight east sunDir turbidity (IO/MaterialParser.hs:266-270)
                east = self.named_vec("east")
                sun = self.named_vec("sunDir")
                turb = self.named_flt("turbidity")
                self.lights.insert(0, ("sunsky", (east, sun, turb), w2l))
            else:
                self.fail("unknown map type " + m)
            self.close()
        elif t == "point":  # pPointLight (LightParser.hs:35-39)
            self.named("intensity")
            r = self.spectrum()
            self.lights.insert(0, ("point", r, self.named_vec("position")))
        elif t == "directional":  # pDirectionalLight (LightParser.hs:41-45)
            self.named("intensity")
            r = self.spectrum()
            self.lights.insert(0, ("directional", r, self.named_vec("normal")))
        else:
            self.fail("unknown light type " + t)
        self.close()

    def emission_block(self):
        self.open()
        if self.peek()[1] == "none":
            self.take()
            self.emit = None
        else:
            self.emit = self.spectrum()
        self.close()

    # -- renderer (RendererParser.hs, IntegratorParser.hs)
    def renderer_block(self):
        self.open()
        t = self.word()
        if t == "light":
            self.renderer = ("light", self.named_int("passPhotons"))
        elif t == "metropolis":
            self.renderer = ("metropolis", self.named_int("maxDepth"), self.named_flt("mpp"),
                             self.named_int("bootstrap"), self.named_flt("plarge"),
                             self.named_int("directSamples"))
        elif t == "sppm":
            n = self.named_int("photonCount")
            md = self.named_int("maxDepth")
            r = self.named_flt("radius")
            a = self.named_flt("alpha") if self.peek()[1] == "alpha" else F(0.8)
            self.renderer = ("sppm", n, md, r, a)
        elif t == "sampler":
            self.named("sampled")
            self.open()
            self.named("sampler")
            self.open()
            st = self.word()
            if st == "stratified":
                sampler = ("stratified", self.integ(), self.integ())
            elif st == "random":
                sampler = ("random", self.integ())
            else:
                self.fail("unknown sampler type " + st)
            self.close()
            self.named("integrator")
            self.open()
            it = self.word()
            if it in ("bidir", "bidirnod", "path"):
                integ = (it, self.named_int("maxDepth"), self.named_int("sampleDepth"))
            elif it == "directLighting":
                integ = (it, self.named_int("maxDepth"))
            elif it == "debug":
                d = self.word()
                if d not in ("kdtree", "normals", "reference"):
                    self.fail("unknown debug integrator " + d)
                integ = (it, d)
            else:
                self.fail("unknown integrator type " + it)
            self.close()
            self.close()
            self.renderer = ("sampler", sampler, integ)
        else:
            self.fail("unknown renderer type " + t)
        self.close()

    def filter_stmt(self):
        t = self.word()
        if t == "box":
            self.filter = (abi.FILTER_BOX, (0.0, 0.0, 0.0, 0.0))
        elif t == "gauss":
            self.filter = (abi.FILTER_GAUSS, (self.flt(), self.flt(), self.flt(), 0.0))
        elif t == "sinc":
            self.filter = (abi.FILTER_SINC, (self.flt(), self.flt(), self.flt(), 0.0))
        elif t == "triangle":
            self.filter = (abi.FILTER_TRIANGLE, (self.flt(), self.flt(), 0.0, 0.0))
        elif t == "mitchell":
            self.filter = (abi.FILTER_MITCHELL, (self.flt(), self.flt(), self.flt(), self.flt()))
        else:
            self.fail('unknown pixel filter function "%s"' % t)

    def run(self):
        while self.peek()[0] != "eof":
            name = self.word()
            if name == "filter":
                self.filter_stmt()
            elif name == "prim":
                self.prim_block()
            elif name == "imageSize":
                self.res = (self.integ(), self.integ())
            elif name == "renderer":
                self.renderer_block()
            elif name == "transform":
                self.transform = self.transform_block() @ self.transform
            elif name == "newTransform":
                self.transform = T.Transform()
                self.transform = self.transform_block() @ self.transform
            elif name == "camera":
                self.camera_block()
            elif name == "light":
                self.light_block()
            elif name == "material":
                self.open()
                self.material = self.material_body()
                self.close()
            elif name == "emission":
                self.emission_block()
            else:
                self.fail("unknwon object type " + name)


class BlingScene:
    """A parsed `.bling` file: the flattened scene and the job its renderer
    block describes."""

    def __init__(self, p):
        sb = SceneBuilder()
        mat_index = {}

        def mat(m):
            if id(m) not in mat_index:
                mat_index[id(m)] = sb.material(m)
            return mat_index[id(m)]

        for stmt in reversed(p.prims):  # mkScene order (RenderJob.hs:49-52)
            for rec in stmt:
                if rec[0] == "mesh":
                    _, verts, faces, xf, m, norms, uvs = rec
                    if len(faces):
                        sb.mesh(verts, faces, o2w=xf, material=mat(m), normals=norms, uvs=uvs)
                elif rec[0] == "objmesh":  # points already in world space, normals as the file gives them
                    _, verts, faces, m, norms, uvs = rec
                    sb.mesh(verts, faces, material=mat(m), normals=norms, uvs=uvs, world=True)
                else:
                    _, kind, params, xf, m, emit, sid = rec
                    sb.shape(kind, params, o2w=xf, material=mat(m), emission=emit)
                    sb.shapes[-1].light_id = sid
        for kind, a, b in p.lights:
            if kind == "infinite":
                sb.infinite_light(a, b)
            elif kind == "sunsky":
                sb.sun_sky_light(a[0], a[1], a[2], b)
            elif kind == "point":
                sb.point_light(a, b)
            else:
                sb.directional_light(a, b)
        filt, fparams = p.filter
        if p.camera[0] == "environment":
            _, c2w, (sx, sy) = p.camera
            sb.environment_camera(c2w, sx, sy)
        else:
            _, c2w, r2c, lr, fd, fov = p.camera
            sb.perspective_camera(c2w, r2c, lens_radius=lr, focal_distance=fd)
            sb.lens = (c2w, float(fov), float(lr), float(fd), int(filt), tuple(float(x) for x in fparams))
        sb.set_film(p.res[0], p.res[1], filt, fparams)
        self.flat = sb.build()
        self.width, self.height = p.res
        self.renderer = p.renderer
        self.num_statements = len(p.prims)

    def spec(self, passes=1, seed=1):
        """The GPU job of the file's renderer block: `sampled { sampler {
        stratified nu nv } integrator { path maxDepth md sampleDepth sd } }`,
        with the reference's hard-coded Russian-roulette depth 7
        (Integrator/Path.hs:68), or `integrator { directLighting maxDepth md }`."""
        from .runtime import JobSpec
        r = self.renderer
        if r[0] != "sampler":
            raise Unsupported("renderer %r (the GPU path renders the sampler renderer)" % r[0])
        if r[1][0] != "stratified":
            raise Unsupported("sampler %r (the GPU path uses the stratified sampler)" % r[1][0])
        if r[2][0] == "directLighting":  # IntegratorParser.hs:37-39
            return JobSpec(spp_nu=r[1][1], spp_nv=r[1][2], max_depth=r[2][1], sample_depth=0,
                           rr_start_depth=7, passes=passes, seed=seed, integrator="direct")
        if r[2][0] != "path":
            raise Unsupported("integrator %r (the GPU path runs the path and direct-lighting integrators)"
                              % r[2][0])
        return JobSpec(spp_nu=r[1][1], spp_nv=r[1][2], max_depth=r[2][1], sample_depth=r[2][2],
                       rr_start_depth=7, passes=passes, seed=seed)


def parse(text, base="."):
    """Parses `.bling` text; returns a BlingScene.  Files the text names
    (image textures) are read relative to `base`."""
    p = _Parser(text, base)
    p.run()
    return BlingScene(p)


def load(path):
    with open(path) as f:
        return parse(f.read(), os.path.dirname(os.path.abspath(path)))
