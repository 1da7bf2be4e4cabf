"""The oracle's node textures (oracle/orc_texture.h) against what the
mathematics and Texture.hs's definitions fix.

- perlin3d: zero on the lattice, and its gradient there is Ken Perlin's
  published gradient vector of the cell's hash (improved noise, 2002; the
  reference's grad differs from it for two hashes, see PERLIN_GRADS);
- fbm: one octave is perlin3d; octave k adds omega^k perlin3d at 1.99^k p;
- cellNoise: a minimum of distances, so Chebyshev <= Euclidean <= Manhattan,
  the squared Euclidean is the square, and the Euclidean one is 1-Lipschitz;
- quasiCrystal: at the origin every wave is 1, so n waves wrap to n mod 2;
  one wave is (cos x + 1) / 2; even in (x, y);
- gradient, blend, checker, graphPaper, scale: their definitions at points
  that fix them; image: texel (px, py) at its centre, wrapped, rows from the
  top (v negated);
- bump mapping: a constant displacement leaves the shading frame alone; a
  displacement (cos x + 1) / 2 tilts dpdu by its finite difference along the
  normal (over u, and over p for a planar mapping); the normal faces the
  geometric one.
Each pin fails when a plausible mistake is planted in the oracle.
The cell hash and its Poisson table are the reference's constants; beyond
the properties above they are parity unpinned.
"""
import numpy as np
import pytest

from bling_amd import abi
from bling_amd.scene import Material, SceneBuilder, const, rgb_refl
from bling_amd.textures import Map2d, Map3d, Tex
from bling_amd.noise_table import PERM
from oracle import pyoracle
from oracle.pyoracle import mutant

F = np.float32
# Ken Perlin's improved-noise gradients (2002), indexed by the hash's low 4
# bits, except for hashes 13 and 14: there the reference's grad
# (Texture.hs:393-400) takes (dx, dy) and (dy, dz) where Perlin's takes
# (dy, dz) and (dy, dx), giving (-1, 1, 0) and (0, 1, -1)
PERLIN_GRADS = [(1, 1, 0), (-1, 1, 0), (1, -1, 0), (-1, -1, 0), (1, 0, 1), (-1, 0, 1), (1, 0, -1), (-1, 0, -1),
                (0, 1, 1), (0, -1, 1), (0, 1, -1), (0, -1, -1), (1, 1, 0), (-1, 1, 0), (0, 1, -1), (0, -1, -1)]


def texture_scene(texs):
    """A one-triangle scene whose materials take the textures (spectrum
    textures as a matte's kd, scalar ones as its sigma); returns (oracle
    scene, flat scene, root of each texture)."""
    sb = SceneBuilder()
    for t in texs:
        m = Material.matte(const(0.5))
        sb.material(m.textured_inputs(scalar={0: t}) if t.scalar else m.textured_inputs(spec={0: t}))
    sb.mesh(np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0]], F), np.array([[0, 1, 2]], np.int32))
    flat = sb.build()
    roots = []
    for i, t in enumerate(texs):
        r = abi.texture_roots(flat.materials[i].tex)
        roots.append(r[4] if t.scalar else r[0])
    return pyoracle.OracleScene(flat), flat, roots


def evaluate(tex, xyz, uv=None):
    o, _, roots = texture_scene([tex])
    try:
        uv = np.zeros((len(xyz), 2), F) if uv is None else uv
        return o.tex_eval(roots[0], np.asarray(xyz, F), np.asarray(uv, F))
    finally:
        o.close()


def perlin_at(p, mapping=None):
    return evaluate(Tex.perlin(mapping or Map3d.identity()), p)


# ------------------------------------------------------------------- pins
def pin_perlin():
    rng = np.random.default_rng(3)
    lat = rng.integers(-300, 300, (40, 3)).astype(F)
    assert np.all(perlin_at(lat) == 0.0)
    # the gradient at a lattice point: (f(p + h e) - f(p - h e)) / 2h
    h = F(1e-3)
    for p in lat[:12]:
        ix, iy, iz = (int(v) & 255 for v in p)
        g = PERLIN_GRADS[PERM[(PERM[(PERM[ix] + iy) % 256] + iz) % 256] & 15]
        pts = []
        for a in range(3):
            e = np.zeros(3, F)
            e[a] = h
            pts += [p + e, p - e]
        v = perlin_at(np.array(pts, F))
        fd = [(v[2 * a] - v[2 * a + 1]) / (2 * h) for a in range(3)]
        assert np.allclose(fd, g, atol=2e-2), (p, fd, g)


def pin_fbm():
    rng = np.random.default_rng(4)
    p = rng.uniform(-20, 20, (200, 3)).astype(F)
    idm = Map3d.identity()
    one = evaluate(Tex.fbm(1, 0.5, idm), p)
    assert np.array_equal(one, perlin_at(p))
    three = evaluate(Tex.fbm(3, 0.5, idm), p)
    l1, l2 = F(1.99), F(F(1.99) * F(1.99))
    want = (perlin_at(p) + F(0.5) * perlin_at(p * l1)).astype(F) + F(0.25) * perlin_at(p * l2)
    assert np.allclose(three, want, rtol=0, atol=1e-6)
    assert np.array_equal(evaluate(Tex.fbm(5, 0.0, idm), p), perlin_at(p))


def pin_cell():
    rng = np.random.default_rng(5)
    m = Map3d.identity()
    p = rng.uniform(-10, 10, (400, 3)).astype(F)
    e, e2, mh, ch = (evaluate(Tex.cell(d, m), p) for d in ("euclidian", "euclidian2", "manhattan", "chebyshev"))
    assert np.all(e >= 0) and np.all(ch <= e * F(1 + 1e-6)) and np.all(e <= mh * F(1 + 1e-6))
    assert np.allclose(e2, e * e, rtol=1e-5)
    # 1-Lipschitz, also across cell faces: pairs straddling x = integer
    a = p.copy()
    a[:, 0] = np.floor(a[:, 0]) + F(0.9999)
    b = a.copy()
    b[:, 0] += F(0.0002)
    d = np.abs(evaluate(Tex.cell("euclidian", m), a) - evaluate(Tex.cell("euclidian", m), b))
    assert np.all(d <= 0.0002 + 1e-6), d.max()


def pin_crystal():
    m = Map2d.uv(1, 1, 0, 0)
    zero = np.zeros((1, 3), F)
    for n in range(1, 9):
        assert evaluate(Tex.crystal(n, m), zero, np.zeros((1, 2), F))[0] == (1.0 if n % 2 else 0.0), n
    rng = np.random.default_rng(6)
    uv = rng.uniform(-30, 30, (300, 2)).astype(F)
    pts = np.zeros((300, 3), F)
    one = evaluate(Tex.crystal(1, m), pts, uv)
    assert np.allclose(one, (np.cos(uv[:, 0].astype(np.float64)) + 1) / 2, atol=2e-6)
    c7 = evaluate(Tex.crystal(7, m), pts, uv)
    assert np.allclose(c7, evaluate(Tex.crystal(7, m), pts, -uv), atol=2e-5)
    assert np.all((c7 >= 0) & (c7 <= 1))


def pin_gradient_blend():
    c0, c1, c2 = rgb_refl(0.1, 0.2, 0.3), rgb_refl(0.9, 0.5, 0.1), rgb_refl(0.3, 0.8, 0.6)
    steps = [(1.0, c2), (0.0, c0), (0.5, c1)]  # mkGradient sorts them
    pts = np.zeros((6, 3), F)
    fs = [-1.0, 0.0, 0.25, 0.5, 0.75, 2.0]
    got = np.concatenate([evaluate(Tex.gradient(Tex.constant(f), steps), pts[:1]) for f in fs])
    want = [c0, c0, F(0.5) * c0 + F(0.5) * c1, c1, F(0.5) * c1 + F(0.5) * c2, c2]
    for g, w in zip(got, want):
        assert np.allclose(g, w, rtol=0, atol=1e-7)
    t1, t2 = Tex.spectrum(c0), Tex.spectrum(c2)
    for x, w in ((-0.5, c0), (0.0, c0), (0.25, F(0.75) * c0 + F(0.25) * c2), (1.0, c2), (3.0, c2)):
        g = evaluate(Tex.blend(t1, t2, Tex.constant(x)), pts[:1])[0]
        assert np.allclose(g, w, rtol=0, atol=1e-7), x


def pin_selectors_scale():
    a, b = rgb_refl(0.8, 0.1, 0.1), rgb_refl(0.1, 0.1, 0.8)
    pts = np.array([[0.2, 0.2, 0.2], [1.2, 0.2, 0.2], [-0.2, 0.2, 0.2], [1.5, 1.5, 0.5]], F)
    got = evaluate(Tex.checker((1, 1, 1), Tex.spectrum(a), Tex.spectrum(b)), pts)
    for g, even in zip(got, [True, False, False, True]):
        assert np.array_equal(g, a if even else b)
    uv = np.array([[0.5, 0.5], [0.01, 0.5], [0.5, 0.995], [3.5, -2.5]], F)
    got = evaluate(Tex.graph_paper(0.1, Map2d.uv(1, 1, 0, 0), Tex.spectrum(a), Tex.spectrum(b)), np.zeros((4, 3), F), uv)
    for g, line in zip(got, [False, True, True, False]):
        assert np.array_equal(g, b if line else a)
    s = evaluate(Tex.scale(0.5, 2.0, Tex.constant(0.25)), pts[:1])[0]
    assert s == F(1.0)


def pin_image():
    h, w = 3, 4
    spectra = np.arange(h * w * 16, dtype=F).reshape(h, w, 16)
    tex = Tex.image(spectra, Map2d.uv(1, 1, 0, 0))
    py, px = np.mgrid[0:h, 0:w]
    u = ((px + 0.5) / w).reshape(-1).astype(F)
    v = (-(py + 0.5) / h).reshape(-1).astype(F)
    uv = np.stack([u, v], 1)
    pts = np.zeros((h * w, 3), F)
    assert np.array_equal(evaluate(tex, pts, uv), spectra.reshape(-1, 16))
    wrapped = np.stack([u + 2, v - 1], 1).astype(F)
    assert np.array_equal(evaluate(tex, pts, wrapped), spectra.reshape(-1, 16))
    grey = np.arange(h * w, dtype=np.uint8).reshape(h, w) * 20
    g = evaluate(Tex.image_grey(grey, Map2d.uv(1, 1, 0, 0)), pts, uv)
    assert np.array_equal(g, grey.reshape(-1).astype(F) / F(255))


def pin_mappings():
    rng = np.random.default_rng(7)
    p = rng.uniform(-5, 5, (100, 3)).astype(F)
    t = np.eye(4, dtype=F)
    t[:3, 3] = (1, 0, 0)  # translate by x + 1
    assert np.array_equal(perlin_at(p, Map3d.identity(t)), perlin_at(p + np.array([1, 0, 0], F)))
    spectra = np.arange(4 * 16, dtype=F).reshape(1, 4, 16)
    tex = Tex.image(spectra, Map2d.planar((0.25, 0, 0), (0, 0, 1), 0.0, -0.5))
    xs = np.array([[0.5, 0, 0], [1.5, 0, 0], [2.5, 0, 0], [3.5, 0, 0]], F)
    assert np.array_equal(evaluate(tex, xs, np.zeros((4, 2), F)), spectra[0])


def bump_frame(disp, g, ng):
    """The oracle's bumped (n, dpdu, dpdv) for a matte bump-mapped by `disp`."""
    sb = SceneBuilder()
    sb.material(Material.matte(const(0.5)).textured_inputs(bump=disp))
    sb.mesh(np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0]], F), np.array([[0, 1, 2]], np.int32))
    o = pyoracle.OracleScene(sb.build())
    try:
        return o.bump(0, np.asarray(g, F), np.asarray(ng, F))
    finally:
        o.close()


def pin_bump():
    """bump d dgg dgs (Reflection.hs:347-378) on a plane z = 0 with dpdu =
    (2, 0, 0), dpdv = (0, 3, 0): a constant displacement leaves the frame
    alone; (cos x + 1) / 2 over x = u (uv mapping) or x = p.x (planar
    mapping, where the u-shifted point moves by 0.01 dpdu = 0.02 in x) tilts
    dpdu by the finite difference along n; the normal faces ng."""
    n, dpdu, dpdv = (0, 0, 1), (2, 0, 0), (0, 3, 0)
    for u in (0.3, 1.1, 2.5):
        g = [0.4, 0.7, 0.0, *n, *dpdu, *dpdv, u, 0.2]
        r = bump_frame(Tex.constant(0.3), g, n)
        assert np.array_equal(r, np.array([n, dpdu, dpdv], F))
        for mapping, x, h in ((Map2d.uv(1, 1, 0, 0), u, 0.01), (Map2d.planar((1, 0, 0), (0, 1, 0), 0, 0), 0.4, 0.02)):
            r = bump_frame(Tex.crystal(1, mapping), g, n)
            us = ((np.cos(x + h) + 1) / 2 - (np.cos(x) + 1) / 2) / 0.01
            assert np.allclose(r[1], [2, 0, us], atol=2e-4), (r[1], us)
            assert np.array_equal(r[2], np.array(dpdv, F))  # nothing varies along v
            want = np.array([-3 * us, 0, 6]) / np.hypot(3 * us, 6)
            assert np.allclose(r[0], want, atol=2e-4), (r[0], want)
            flipped = bump_frame(Tex.crystal(1, mapping), g, (0, 0, -1))
            assert np.allclose(flipped[0], -want, atol=2e-4)


PINS = {"perlin": pin_perlin, "fbm": pin_fbm, "cell": pin_cell, "crystal": pin_crystal,
        "gradient_blend": pin_gradient_blend, "selectors_scale": pin_selectors_scale, "image": pin_image,
        "mappings": pin_mappings, "bump": pin_bump}


@pytest.mark.parametrize("name", sorted(PINS))
def test_texture_pin(name):
    PINS[name]()


MUTATIONS = [
    (mutant.PERLIN_WEIGHT, "perlin"),
    (mutant.CELL_OWN, "cell"),
    (mutant.CRYSTAL_WRAP, "crystal"),
    (mutant.GRAD_SWAP, "gradient_blend"),
    (mutant.BLEND_SWAP, "gradient_blend"),
    (mutant.IMAGE_FLIP, "image"),
    (mutant.FBM_FREQ, "fbm"),
    (mutant.CHECKER_TRUNC, "selectors_scale"),
    (mutant.BUMP_FACE, "bump"),
    (mutant.BUMP_SHIFT, "bump"),
]


@pytest.mark.parametrize("m,name", MUTATIONS)
def test_each_texture_pin_fails_on_its_planted_mistake(m, name):
    with mutant(m):
        with pytest.raises(AssertionError):
            PINS[name]()
    PINS[name]()


def test_invalid_node_lists_are_refused():
    """A child after its node, a scalar where a spectrum is taken, a
    gradient without steps: the oracle refuses the scene."""
    o, flat, roots = texture_scene([Tex.blend(Tex.spectrum(const(0.2)), Tex.spectrum(const(0.4)),
                                              Tex.constant(0.5))])
    o.close()
    n = flat.desc.num_tex_nodes
    for mutate in (lambda t: setattr(t[n - 1], "a", n - 1),       # own child
                   lambda t: setattr(t[n - 1], "c", 0),           # a spectrum as the weight
                   lambda t: setattr(t[n - 1], "op", 99)):        # unknown operation
        saved = [(t.op, t.a, t.b, t.c) for t in flat.tex_nodes[:n]]
        mutate(flat.tex_nodes)
        with pytest.raises(ValueError):
            pyoracle.OracleScene(flat)
        for t, (op, a, b, c) in zip(flat.tex_nodes[:n], saved):
            t.op, t.a, t.b, t.c = op, a, b, c
    pyoracle.OracleScene(flat).close()
