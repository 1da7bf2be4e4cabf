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
