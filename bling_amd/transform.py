"""Binary32 transforms and cameras, built the way bling builds them, so a
scene flattened here carries the same matrices the Haskell front end computes.

  Transform.hs:123-306   matrices, `mul`, Gauss-Jordan `invert`, translate,
                         scale, rotateX/Y/Z, perspective, lookAt, concatTrans
  Camera.hs:108-147      mkProjective / mkPerspectiveCamera

A `Transform` holds (m, inv) as 16 float32 values each, row-major.  `a @ b`
is the reference's `a <> b` (concatTrans): apply `a` first, then `b`.
Trigonometry goes through the C library's sinf/cosf/tanf, which GHC's Float
primops call.
"""
import ctypes
import ctypes.util

import numpy as np

F = np.float32
_libm = ctypes.CDLL(ctypes.util.find_library("m") or "libm.so.6")
for _n in ("sinf", "cosf", "tanf"):
    getattr(_libm, _n).restype = ctypes.c_float
    getattr(_libm, _n).argtypes = [ctypes.c_float]

PI = F(np.pi)  # `pi :: Float`


def sinf(x):
    return F(_libm.sinf(float(F(x))))


def cosf(x):
    return F(_libm.cosf(float(F(x))))


def tanf(x):
    return F(_libm.tanf(float(F(x))))


def radians(x):  # Math.hs:62-66: x / 180 * pi
    return F(F(F(x) / F(180)) * PI)


def _mat(*vals):
    assert len(vals) == 16
    return [F(v) for v in vals]


IDENTITY = _mat(1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1)


def mat_mul(m1, m2):
    """`mul m1 m2` (Transform.hs:196-199): element n=(i,j) is the left fold
    sum over k of m1[k][j] * m2[i][k]."""
    out = []
    for n in range(16):
        i, j = divmod(n, 4)
        acc = F(0)
        for k in range(4):
            acc = F(acc + F(m1[4 * k + j] * m2[4 * i + k]))
        out.append(acc)
    return out


def transpose(m):
    return [m[4 * (n % 4) + n // 4] for n in range(16)]


def invert(m):
    """Gauss-Jordan with full pivoting in the reference's order
    (Transform.hs:138-181); storage is indexed column-major there."""
    minv = [F(v) for v in m]

    def idx(r, c):
        return c * 4 + r

    ipiv = [0, 1, 2, 3]
    indx = []
    for _ in range(4):
        best, bv = None, None
        for j in ipiv:
            for k in ipiv:
                v = abs(minv[idx(j, k)])
                if best is None or not (bv > v):  # maximumBy keeps the last maximum
                    best, bv = (j, k), v
        irow, icol = best
        ipiv = [x for x in ipiv if x != icol]
        if irow != icol:
            for k in range(4):
                a, b = idx(irow, k), idx(icol, k)
                minv[a], minv[b] = minv[b], minv[a]
        pivinv = F(F(1) / minv[idx(icol, icol)])
        minv[idx(icol, icol)] = F(1)
        for j in range(4):
            minv[idx(icol, j)] = F(minv[idx(icol, j)] * pivinv)
        for j in range(4):
            if j == icol:
                continue
            save = minv[idx(j, icol)]
            minv[idx(j, icol)] = F(0)
            for k in range(4):
                x = minv[idx(icol, k)]
                minv[idx(j, k)] = F(minv[idx(j, k)] - F(x * save))
        indx.append((irow, icol))
    for ir, ic in reversed(indx):
        if ir != ic:
            for k in range(4):
                a, b = idx(k, ir), idx(k, ic)
                minv[a], minv[b] = minv[b], minv[a]
    return minv


class Transform:
    __slots__ = ("m", "inv")

    def __init__(self, m=None, inv=None):
        self.m = list(IDENTITY) if m is None else [F(v) for v in m]
        self.inv = list(IDENTITY) if inv is None else [F(v) for v in inv]

    def __matmul__(self, other):  # concatTrans (Transform.hs:241-245)
        return Transform(mat_mul(self.m, other.m), mat_mul(other.inv, self.inv))

    def inverse(self):
        return Transform(self.inv, self.m)

    def np_m(self):
        return np.array(self.m, dtype=np.float32)

    def np_inv(self):
        return np.array(self.inv, dtype=np.float32)

    def point(self, p):  # transPoint (Transform.hs:247-258)
        m = self.m
        x, y, z = (F(v) for v in p)
        xp = F(F(F(F(m[0] * x) + F(m[1] * y)) + F(m[2] * z)) + m[3])
        yp = F(F(F(F(m[4] * x) + F(m[5] * y)) + F(m[6] * z)) + m[7])
        zp = F(F(F(F(m[8] * x) + F(m[9] * y)) + F(m[10] * z)) + m[11])
        wp = F(F(F(F(m[12] * x) + F(m[13] * y)) + F(m[14] * z)) + m[15])
        if wp == F(1):
            return (xp, yp, zp)
        return (F(xp / wp), F(yp / wp), F(zp / wp))

    def points(self, pts):
        """Vectorised transPoint over an (n, 3) float32 array, same op order."""
        p = np.asarray(pts, dtype=np.float32)
        m = self.np_m()
        x, y, z = p[:, 0], p[:, 1], p[:, 2]
        rows = []
        for r in range(4):
            v = m[4 * r] * x
            v = v + m[4 * r + 1] * y
            v = v + m[4 * r + 2] * z
            v = v + m[4 * r + 3]
            rows.append(v.astype(np.float32))
        xp, yp, zp, wp = rows
        one = wp == np.float32(1)
        with np.errstate(divide="ignore", invalid="ignore"):
            out = np.stack([np.where(one, xp, xp / wp), np.where(one, yp, yp / wp),
                            np.where(one, zp, zp / wp)], axis=1)
        return out.astype(np.float32)

    def normals(self, ns):
        """transNormal (Transform.hs:267-274) over (n, 3): uses the inverse, transposed."""
        n = np.asarray(ns, dtype=np.float32)
        i = self.np_inv()
        x, y, z = n[:, 0], n[:, 1], n[:, 2]
        out = []
        for c in range(3):
            v = i[c] * x
            v = v + i[4 + c] * y
            v = v + i[8 + c] * z
            out.append(v.astype(np.float32))
        return np.stack(out, axis=1)


def translate(d):
    dx, dy, dz = (F(v) for v in d)
    return Transform(_mat(1, 0, 0, dx, 0, 1, 0, dy, 0, 0, 1, dz, 0, 0, 0, 1),
                     _mat(1, 0, 0, -dx, 0, 1, 0, -dy, 0, 0, 1, -dz, 0, 0, 0, 1))


def scale(s):
    sx, sy, sz = (F(v) for v in s)
    return Transform(_mat(sx, 0, 0, 0, 0, sy, 0, 0, 0, 0, sz, 0, 0, 0, 0, 1),
                     _mat(F(1) / sx, 0, 0, 0, 0, F(1) / sy, 0, 0, 0, 0, F(1) / sz, 0, 0, 0, 0, 1))


def rotate_x(deg):
    s, c = sinf(radians(deg)), cosf(radians(deg))
    m = _mat(1, 0, 0, 0, 0, c, -s, 0, 0, s, c, 0, 0, 0, 0, 1)
    return Transform(m, transpose(m))


def rotate_y(deg):
    c, s = cosf(radians(deg)), sinf(radians(deg))
    m = _mat(c, 0, s, 0, 0, 1, 0, 0, -s, 0, c, 0, 0, 0, 0, 1)
    return Transform(m, transpose(m))


def rotate_z(deg):
    s, c = sinf(radians(deg)), cosf(radians(deg))
    m = _mat(c, -s, 0, 0, s, c, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1)
    return Transform(m, transpose(m))


def from_matrix(rows):
    """fromMatrix' (Transform.hs:140-141): matrix plus its computed inverse."""
    m = [F(v) for r in rows for v in r]
    return Transform(m, invert(m))


def _v_norm(v):  # normalize (Math.hs:349-353)
    x, y, z = v
    sl = F(F(F(x * x) + F(y * y)) + F(z * z))
    if sl != F(0):
        il = F(F(1) / F(np.sqrt(sl)))
        return (F(x * il), F(y * il), F(z * il))
    return (F(0), F(1), F(0))


def _v_cross(u, v):  # cross (Math.hs:336-339)
    return (F(F(u[1] * v[2]) - F(u[2] * v[1])), F(-F(F(u[0] * v[2]) - F(u[2] * v[0]))),
            F(F(u[0] * v[1]) - F(u[1] * v[0])))


def look_at(pos, look, up):  # Transform.hs:222-237
    p = tuple(F(v) for v in pos)
    l = tuple(F(v) for v in look)
    d = _v_norm((F(l[0] - p[0]), F(l[1] - p[1]), F(l[2] - p[2])))
    left = _v_norm(_v_cross(_v_norm(tuple(F(v) for v in up)), d))
    u = _v_cross(d, left)
    m = _mat(left[0], u[0], d[0], p[0], left[1], u[1], d[1], p[1],
             left[2], u[2], d[2], p[2], 0, 0, 0, 1)
    return Transform(m, invert(m))


def perspective(fov, n, f):  # Transform.hs:207-221
    n, f = F(n), F(f)
    i_tan = F(F(1) / tanf(F(radians(fov) / F(2))))
    fn = F(f - n)
    m = _mat(1, 0, 0, 0, 0, 1, 0, 0, 0, 0, F(f / fn), F(-F(F(f * n) / fn)), 0, 0, 1, 0)
    return scale((i_tan, i_tan, F(1))) @ Transform(m, invert(m))


def perspective_camera(c2w, lens_radius, focal_distance, fov, sx, sy):
    """mkPerspectiveCamera (Camera.hs:136-147) through mkProjective
    (Camera.hs:108-133); returns (c2w, r2c) transforms."""
    p = perspective(fov, 1e-2, 1000)
    sx, sy = F(sx), F(sy)
    aspect = F(sx / sy)
    if aspect > F(1):
        s0, s1, s2, s3 = -aspect, aspect, F(-1), F(1)
    else:
        s0, s1, s2, s3 = F(-1), F(1), F(F(-1) / aspect), F(F(1) / aspect)
    st1 = scale((sx, sy, F(1)))
    st2 = scale((F(F(1) / F(s1 - s0)), F(F(1) / F(s2 - s3)), F(1)))
    t = translate((-s0, -s3, F(0)))
    s2r = t @ (st2 @ st1)  # t <> st2 <> st1, infixr
    r2s = s2r.inverse()
    r2c = r2s @ p.inverse()
    return c2w, r2c
