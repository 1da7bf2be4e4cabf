"""Binary32 gradient noise and fBm, evaluated the way the reference's
Float code evaluates them, for maps the scene reader bakes into geometry
(heightmaps).

  perlin3d   Texture.hs:358-397 (noiseWeight, grad, trilinear lerps)
  fbm        Texture.hs:342-350: sum (foldl, from 0) of o_i * perlin3d(p * l_i)
             with l_{i+1} = 1.99 * l_i and o_{i+1} = omega * o_i, from 1
  lerp       Math.hs:108-110: (1 - t) * v1 + t * v2

Every operation is one numpy binary32 operation in the reference's order, so
the results are the values a GHC `Float` evaluation gives.
"""
import numpy as np

from .noise_table import PERM

F = np.float32
_P = np.array(PERM + PERM, np.int64)


def _lerp(t, a, b):
    return (F(1) - t) * a + t * b


def _weight(t):  # noiseWeight: 6 t^4 t - 15 t^4 + 10 t^3
    t3 = (t * t) * t
    t4 = t3 * t
    return ((F(6) * t4) * t - F(15) * t4) + F(10) * t3


def _grad(x, y, z, dx, dy, dz):
    h = _P[_P[_P[x] + y] + z] & 15
    sel = (h == 12) | (h == 13)
    u = np.where((h < 8) | sel, dx, dy)
    v = np.where((h < 4) | sel, dy, dz)
    u = np.where((h & 1) != 0, -u, u)
    v = np.where((h & 2) != 0, -v, v)
    return u + v


def perlin3d(x, y, z):
    """perlin3d (x, y, z) for binary32 arrays."""
    x, y, z = (np.asarray(a, F) for a in (x, y, z))
    fx, fy, fz = np.floor(x), np.floor(y), np.floor(z)
    dx, dy, dz = x - fx, y - fy, z - fz  # x - fromIntegral (floor x)
    ix, iy, iz = (f.astype(np.int64) & 255 for f in (fx, fy, fz))
    one = F(1)
    w000 = _grad(ix, iy, iz, dx, dy, dz)
    w100 = _grad(ix + 1, iy, iz, dx - one, dy, dz)
    w010 = _grad(ix, iy + 1, iz, dx, dy - one, dz)
    w110 = _grad(ix + 1, iy + 1, iz, dx - one, dy - one, dz)
    w001 = _grad(ix, iy, iz + 1, dx, dy, dz - one)
    w101 = _grad(ix + 1, iy, iz + 1, dx - one, dy, dz - one)
    w011 = _grad(ix, iy + 1, iz + 1, dx, dy - one, dz - one)
    w111 = _grad(ix + 1, iy + 1, iz + 1, dx - one, dy - one, dz - one)
    wx, wy, wz = _weight(dx), _weight(dy), _weight(dz)
    x00 = _lerp(wx, w000, w100)
    x10 = _lerp(wx, w010, w110)
    x01 = _lerp(wx, w001, w101)
    x11 = _lerp(wx, w011, w111)
    y0 = _lerp(wy, x00, x10)
    y1 = _lerp(wy, x01, x11)
    return _lerp(wz, y0, y1)


def fbm(octaves, omega, x, y, z):
    """fbm octaves omega (x, y, z) for binary32 arrays."""
    x, y, z = (np.asarray(a, F) for a in (x, y, z))
    acc = np.zeros(np.broadcast(x, y, z).shape, F)
    l, o = F(1), F(1)
    for _ in range(max(0, octaves)):
        acc = acc + o * perlin3d(x * l, y * l, z * l)
        l = F(1.99) * l
        o = F(omega) * o
    return acc
