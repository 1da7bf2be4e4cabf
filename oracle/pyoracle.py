"""ctypes access to oracle/liboracle.so (built by oracle/Makefile).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
CPU-baseline leg of bench.py, never by the product package.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")
_lib = None


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        lib = C.CDLL(LIB)
        vp = C.c_void_p
        lib.orc_scene_create.restype = vp
        lib.orc_scene_create.argtypes = [vp, C.c_char_p, C.c_int]
        lib.orc_scene_destroy.argtypes = [vp]
        lib.orc_scene_bounds.argtypes = [vp, vp]
        lib.orc_intersect_closest.argtypes = [vp, vp, C.c_size_t, vp, C.c_int, C.c_int,
                                              C.POINTER(C.c_int64)]
        lib.orc_intersect_any.argtypes = [vp, vp, C.c_size_t, vp, C.c_int, C.c_int]
        for name in ("orc_render", "orc_render_samples"):
            if hasattr(lib, name):
                getattr(lib, name).restype = C.c_int
        _lib = lib
    return _lib


class OracleScene:
    def __init__(self, flat):
        from bling_amd import abi  # descriptor layout only
        self.flat = flat
        lib = load()
        err = C.create_string_buffer(256)
        h = lib.orc_scene_create(C.addressof(flat.desc), err, 256)
        if not h:
            raise ValueError(err.value.decode())
        self.h = h
        self._abi = abi

    def close(self):
        if self.h:
            load().orc_scene_destroy(self.h)
            self.h = None

    def tex_eval(self, node, xyz, uv):
        """Texture `node` of the scene's node list at points xyz (n, 3) with
        surface parameters uv (n, 2): (n, 16) for a spectrum texture, (n,)
        for a scalar one."""
        lib = load()
        lib.orc_tex_eval.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p]
        xyz = np.ascontiguousarray(xyz, np.float32)
        uv = np.ascontiguousarray(uv, np.float32)
        n = len(xyz)
        op = self.flat.tex_nodes[node].op if node < self.flat.desc.num_tex_nodes else 0
        out = np.zeros((n, 16) if op >= 16 else n, np.float32)
        if lib.orc_tex_eval(self.h, node, n, xyz.ctypes.data, uv.ctypes.data, out.ctypes.data) != 0:
            raise ValueError("not a texture node: %d" % node)
        return out

    def bump(self, material, g, ng):
        """bump d dgg dgs for bump-mapped `material` at the shading DG g = (p,
        n, dpdu, dpdv, u, v) (14 floats) with geometric normal ng: the bumped
        (n, dpdu, dpdv) as a (3, 3) array."""
        lib = load()
        lib.orc_bump.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p]
        g = np.ascontiguousarray(g, np.float32).reshape(14)
        ng = np.ascontiguousarray(ng, np.float32).reshape(3)
        out = np.zeros(9, np.float32)
        if lib.orc_bump(self.h, material, g.ctypes.data, ng.ctypes.data, out.ctypes.data) != 0:
            raise ValueError("material %d is not bump-mapped" % material)
        return out.reshape(3, 3)

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def bounds(self):
        out = np.zeros(6, np.float32)
        load().orc_scene_bounds(self.h, out.ctypes.data)
        return out[:3], out[3:]

    def _rays(self, rays):
        rays = np.ascontiguousarray(np.asarray(rays, np.float32))
        cols = [np.ascontiguousarray(rays[:, k]) for k in range(8)]
        r = self._abi.Rays(*[c.ctypes.data for c in cols])
        return rays, cols, r

    def intersect(self, rays, brute=False, threads=None):
        rays, cols, r = self._rays(rays)
        n = len(rays)
        t = np.empty(n, np.float32)
        b1 = np.empty(n, np.float32)
        b2 = np.empty(n, np.float32)
        prim = np.empty(n, np.int32)
        h = self._abi.Hits(t.ctypes.data, b1.ctypes.data, b2.ctypes.data, prim.ctypes.data)
        ties = C.c_int64()
        load().orc_intersect_closest(self.h, C.addressof(r), n, C.addressof(h), int(brute),
                                     threads or os.cpu_count(), C.byref(ties))
        return t, b1, b2, prim, ties.value

    def intersects(self, rays, brute=False, threads=None):
        rays, cols, r = self._rays(rays)
        n = len(rays)
        occ = np.empty(n, np.uint8)
        load().orc_intersect_any(self.h, C.addressof(r), n, occ.ctypes.data, int(brute),
                                 threads or os.cpu_count())
        return occ.astype(bool)


def render(flat, spec, threads=None, passes=None, scene=None):
    """Oracle render of a whole job: returns ((h, w, 4) wxyz, (h, w, 3) sum of
    squared XYZ, ray queries traced).  `scene`: an OracleScene of `flat` to
    reuse (its build is then not repeated)."""
    lib = load()
    o = scene if scene is not None else OracleScene(flat)
    h, w = spec.image_shape(flat)
    img = np.zeros((h, w, 4), np.float32)
    sq = np.zeros((h, w, 3), np.float32)
    rays = C.c_uint64()
    d = spec.desc()
    p0, p1 = (0, spec.passes) if passes is None else passes
    lib.orc_render.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int,
                               C.POINTER(C.c_uint64), C.c_int, C.c_int]
    lib.orc_render(o.h, C.addressof(d), img.ctypes.data, sq.ctypes.data,
                   threads or os.cpu_count(), C.byref(rays), p0, p1)
    if scene is None:
        o.close()
    return img, sq, rays.value


def aniso(ex, ey, wh=None, u=None):
    """The anisotropic microfacet distribution of exponents (ex, ey)
    (Microfacet.hs:136-196).  With half vectors wh (n, 3): (D, pdf) there.
    With sample points u (n, 2): the sampler's (wh, d, pdf)."""
    lib = load()
    vp = C.c_void_p
    lib.orc_aniso.argtypes = [C.c_float, C.c_float, C.c_int, vp, vp, vp, vp]
    if u is not None:
        u = np.ascontiguousarray(u, np.float32)
        n = len(u)
        wh = np.zeros((n, 3), np.float32)
    else:
        wh = np.ascontiguousarray(wh, np.float32)
        n = len(wh)
    d = np.zeros(n, np.float32)
    pdf = np.zeros(n, np.float32)
    lib.orc_aniso(ex, ey, n, None if u is None else u.ctypes.data, wh.ctypes.data, d.ctypes.data, pdf.ctypes.data)
    return (wh, d, pdf) if u is not None else (d, pdf)


def bxdf_eval(material, wo, wi, comp=0):
    """Component `comp` of `material` (an abi.Material) in its local frame:
    bxdfEval wo wi as (n, 16) spectra and bxdfPdf wo wi."""
    return _bxdf(material, 0, wo, wi, None, comp)[1:]


def bxdf_sample(material, wo, u, comp=0):
    """bxdfSample (adj False) of component `comp` of `material` at sample
    points u (n, 2): returns (wi, f, pdf)."""
    return _bxdf(material, 1, wo, None, u, comp)


def _bxdf(material, op, wo, wi, u, comp=0):
    lib = load()
    vp = C.c_void_p
    lib.orc_bxdf_comp.argtypes = [vp, C.c_int, C.c_int, C.c_int, vp, vp, vp, vp, vp]
    lib.orc_bxdf_comp.restype = C.c_int
    wo = np.ascontiguousarray(wo, np.float32)
    n = len(wo)
    wi = np.zeros((n, 3), np.float32) if wi is None else np.ascontiguousarray(wi, np.float32).copy()
    u = None if u is None else np.ascontiguousarray(u, np.float32)
    f = np.zeros((n, 16), np.float32)
    pdf = np.zeros(n, np.float32)
    rc = lib.orc_bxdf_comp(C.addressof(material), comp, op, n, wo.ctypes.data,
                           None if u is None else u.ctypes.data, wi.ctypes.data, f.ctypes.data, pdf.ctypes.data)
    if rc != 0:
        raise ValueError("material has no BxDF component %d" % comp)
    return wi, f, pdf


def fresnel_dielectric(etai, etat, cosi):
    """frDielectric etai etat cosi (Fresnel.hs:21-55), band 0."""
    lib = load()
    lib.orc_fresnel.argtypes = [C.c_int, C.c_float, C.c_float, C.c_void_p, C.c_void_p, C.c_float, C.c_void_p]
    out = np.zeros(16, np.float32)
    lib.orc_fresnel(0, etai, etat, None, None, cosi, out.ctypes.data)
    return out[0]


def fresnel_conductor(eta, k, cosi):
    """frConductor eta k cosi (Fresnel.hs:58-70) for 16-band eta and k."""
    lib = load()
    lib.orc_fresnel.argtypes = [C.c_int, C.c_float, C.c_float, C.c_void_p, C.c_void_p, C.c_float, C.c_void_p]
    eta = np.ascontiguousarray(np.broadcast_to(np.asarray(eta, np.float32), (16,)))
    k = np.ascontiguousarray(np.broadcast_to(np.asarray(k, np.float32), (16,)))
    out = np.zeros(16, np.float32)
    lib.orc_fresnel(1, 0.0, 0.0, eta.ctypes.data, k.ctypes.data, cosi, out.ctypes.data)
    return out


def blinn(e, wh):
    """The Blinn distribution's D and half-vector pdf at half vectors wh (n, 3)."""
    lib = load()
    vp = C.c_void_p
    lib.orc_blinn.argtypes = [C.c_float, C.c_int, vp, vp, vp]
    wh = np.ascontiguousarray(wh, np.float32)
    d = np.zeros(len(wh), np.float32)
    pdf = np.zeros(len(wh), np.float32)
    lib.orc_blinn(e, len(wh), wh.ctypes.data, d.ctypes.data, pdf.ctypes.data)
    return d, pdf


def envlight(radiance_map, u=None, wi=None, w2l=None):
    """An infinite light with the (h, w, 16) map: with u (n, 2) its samples
    (wi, pdf, li); with wi (n, 3) its (pdf, le) there."""
    lib = load()
    vp = C.c_void_p
    lib.orc_envlight.argtypes = [C.c_int, C.c_int, vp, vp, vp, C.c_int, C.c_int, vp, vp, vp, vp]
    m = np.ascontiguousarray(radiance_map, np.float32)
    h, w = m.shape[:2]
    eye = np.eye(4, dtype=np.float32)
    w2l = eye if w2l is None else np.ascontiguousarray(w2l, np.float32)
    l2w = np.ascontiguousarray(np.linalg.inv(w2l.astype(np.float64)).astype(np.float32))
    if u is not None:
        u = np.ascontiguousarray(u, np.float32)
        n = len(u)
        wi = np.zeros((n, 3), np.float32)
    else:
        wi = np.ascontiguousarray(wi, np.float32)
        n = len(wi)
    pdf = np.zeros(n, np.float32)
    li = np.zeros((n, 16), np.float32)
    lib.orc_envlight(w, h, m.ctypes.data, w2l.ctypes.data, l2w.ctypes.data, 0 if u is not None else 1, n,
                     None if u is None else u.ctypes.data, wi.ctypes.data, pdf.ctypes.data, li.ctypes.data)
    return (wi, pdf, li) if u is not None else (pdf, li)


def skylight(sky, u=None, wi=None, w2l=None):
    """An infinite light with the sun/sky map of sky = (east xyz, sunDir xyz,
    turbidity): with u (n, 2) its samples (wi, pdf, li); with wi (n, 3) its
    (pdf, le) there."""
    lib = load()
    vp = C.c_void_p
    lib.orc_skylight.argtypes = [vp, vp, vp, C.c_int, C.c_int, vp, vp, vp, vp]
    k = np.ascontiguousarray(sky, np.float32)
    eye = np.eye(4, dtype=np.float32)
    w2l = eye if w2l is None else np.ascontiguousarray(w2l, np.float32)
    l2w = np.ascontiguousarray(np.linalg.inv(w2l.astype(np.float64)).astype(np.float32))
    if u is not None:
        u = np.ascontiguousarray(u, np.float32)
        n = len(u)
        wi = np.zeros((n, 3), np.float32)
    else:
        wi = np.ascontiguousarray(wi, np.float32)
        n = len(wi)
    pdf = np.zeros(n, np.float32)
    li = np.zeros((n, 16), np.float32)
    lib.orc_skylight(k.ctypes.data, w2l.ctypes.data, l2w.ctypes.data, 0 if u is not None else 1, n,
                     None if u is None else u.ctypes.data, wi.ctypes.data, pdf.ctypes.data, li.ctypes.data)
    return (wi, pdf, li) if u is not None else (pdf, li)


def sky_eval(sky, uv):
    """The sun/sky map of sky = (east xyz, sunDir xyz, turbidity) at the map
    coordinates uv (n, 2): (spectra (n, 16), data) with data a dict of the
    model's sunR, sunDir, sunTheta and zenith x, y, Y."""
    lib = load()
    vp = C.c_void_p
    lib.orc_sky_eval.argtypes = [vp, C.c_int, vp, vp, vp]
    k = np.ascontiguousarray(sky, np.float32)
    uv = np.ascontiguousarray(uv, np.float32).reshape(-1, 2)
    out = np.zeros((len(uv), 16), np.float32)
    data = np.zeros(23, np.float32)
    lib.orc_sky_eval(k.ctypes.data, len(uv), uv.ctypes.data, out.ctypes.data, data.ctypes.data)
    return out, dict(sunR=data[:16].copy(), sunDir=data[16:19].copy(), sunTheta=data[19], zx=data[20],
                     zy=data[21], zY=data[22])


class mutant:
    """Context manager planting mistake `m` (orc_render.cpp's MUT_* list) in
    the oracle for the duration of a block."""
    FR_DIEL_ETA, FR_COND_K, BLINN_NORM, MF_G, TRANS_ETA2, TRANS_DIR, OREN_SWAP = 1, 2, 3, 4, 5, 6, 7
    ENV_SIN, ENV_ROW, PATH_PC, PATH_T, POWER, FILTER_CLIP = 8, 9, 10, 11, 12, 13
    PERLIN_WEIGHT, CELL_OWN, GRAD_SWAP, CRYSTAL_WRAP, BLEND_SWAP, IMAGE_FLIP = 14, 15, 16, 17, 18, 19
    FBM_FREQ, CHECKER_TRUNC, BUMP_FACE, BUMP_SHIFT, SHINY_SQRT = 20, 21, 22, 23, 24
    POINT_FALLOFF, DIR_COS = 25, 26
    SKY_UP, SKY_DISK, SKY_SUN_BELOW, LIGHT_TEXEL = 27, 28, 29, 30

    def __init__(self, m):
        self.m = m

    def __enter__(self):
        lib = load()
        lib.orc_set_mutant.argtypes = [C.c_int]
        lib.orc_set_mutant(self.m)
        return self

    def __exit__(self, *exc):
        load().orc_set_mutant(0)
