"""Pins of the oracle's anisotropic microfacet distribution and FresnelBlend
BRDF (Microfacet.hs:56-104,136-196), the substrate material's only
component, against what the mathematics fixes:

- both normalisations, D (projected) and the sampling pdf, by quadrature;
- the isotropic case, which is the Blinn distribution and its textbook
  sampler in closed form;
- the sampler's half vectors, distributed as its pdf says (moments and
  quadrant occupancy against quadrature);
- FresnelBlend's sample, consistent with its own eval and pdf;
- the coating's absorption, Beer-Lambert over the two paths through it.

The diffuse and specular factors of FresnelBlend's value are the
reference's own expressions (not a textbook form); no check fixes them
beyond the consistency above."""
import numpy as np
import pytest

from bling_amd.scene import Material, rgb_refl
from oracle import pyoracle


def _hemisphere_grid(nt=3000, nphi=720):
    """Midpoint grid over the upper hemisphere: directions and solid angles."""
    th = (np.arange(nt) + 0.5) * (np.pi / 2 / nt)
    ph = (np.arange(nphi) + 0.5) * (2 * np.pi / nphi)
    T, P = np.meshgrid(th, ph, indexing="ij")
    w = np.stack([np.sin(T) * np.cos(P), np.sin(T) * np.sin(P), np.cos(T)], -1).reshape(-1, 3)
    dw = (np.sin(T) * (np.pi / 2 / nt) * (2 * np.pi / nphi)).reshape(-1)
    return w, dw


EXPONENTS = [(2.0, 30.0), (10.0, 100.0), (200.0, 50.0), (20.0, 20.0)]


@pytest.mark.parametrize("ex,ey", EXPONENTS)
def test_anisotropic_distribution_is_normalised(ex, ey):
    """The integrals of D cos(theta_h) and of the sampling pdf over the
    hemisphere are both 1.  Integrating over theta leaves the azimuthal
    integral of 1 / (ex cos^2 phi + ey sin^2 phi + k), which is
    2 pi / sqrt((ex + k)(ey + k)), with k = 2 for D cos and k = 1 for the pdf:
    exactly what the constants sqrt((ex+2)(ey+2)) / 2pi and
    sqrt((ex+1)(ey+1)) / 2pi cancel.  Swapping the two constants, or a wrong
    exponent, moves one integral by more than 1e-2 at these exponents."""
    w, dw = _hemisphere_grid()
    d, pdf = pyoracle.aniso(ex, ey, wh=w)
    assert abs(np.sum(d.astype(np.float64) * w[:, 2] * dw) - 1.0) < 2e-3
    assert abs(np.sum(pdf.astype(np.float64) * dw) - 1.0) < 2e-3


@pytest.mark.parametrize("e", [1.0, 20.0, 500.0])
def test_isotropic_case_is_blinn(e):
    """With ex = ey = e the exponent is e everywhere: D = (e+2)/2pi cos^e,
    pdf = (e+1)/2pi cos^e, and the sampler is the textbook Blinn sampler
    cos(theta) = u2^(1/(e+1)), phi = 2 pi u1 in every quadrant."""
    rng = np.random.default_rng(7)
    v = rng.normal(size=(4000, 3))
    v[:, 2] = np.abs(v[:, 2]) + 0.3
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    d, pdf = pyoracle.aniso(e, e, wh=v.astype(np.float32))
    c = v[:, 2].astype(np.float32).astype(np.float64) ** e
    keep = c > 1e-20
    assert np.allclose(d[keep], ((e + 2) / (2 * np.pi) * c)[keep], rtol=2e-4)
    assert np.allclose(pdf[keep], ((e + 1) / (2 * np.pi) * c)[keep], rtol=2e-4)
    u = rng.random((4000, 2)).astype(np.float32)
    wh, _, _ = pyoracle.aniso(e, e, u=u)
    u64 = u.astype(np.float64)
    cost = u64[:, 1] ** (1.0 / (e + 1))
    sint = np.sqrt(np.maximum(0, 1 - cost * cost))
    want = np.stack([sint * np.cos(2 * np.pi * u64[:, 0]), sint * np.sin(2 * np.pi * u64[:, 0]), cost], -1)
    assert np.abs(wh - want).max() < 2e-4


@pytest.mark.parametrize("ex,ey", [(2.0, 30.0), (200.0, 50.0)])
def test_anisotropic_sampler_follows_its_pdf(ex, ey):
    """Half vectors drawn by the sampler have the moments that the pdf gives
    by quadrature (E[x^2], E[y^2], E[z]), E[x y] = 0, one quarter in each
    quadrant; the returned pdf is the pdf at the returned half vector.  A
    swapped exponent or a quadrant mapped wrong moves a moment by many
    standard errors."""
    rng = np.random.default_rng(11)
    n = 400_000
    u = rng.random((n, 2)).astype(np.float32)
    wh, _, spdf = pyoracle.aniso(ex, ey, u=u)
    assert np.all(np.isfinite(wh)) and np.all(wh[:, 2] >= 0)
    _, pdf_at = pyoracle.aniso(ex, ey, wh=wh)
    assert np.allclose(spdf, pdf_at, rtol=1e-5)
    w, dw = _hemisphere_grid()
    _, pdf = pyoracle.aniso(ex, ey, wh=w)
    p = pdf.astype(np.float64) * dw
    for g in (lambda v: v[:, 0] ** 2, lambda v: v[:, 1] ** 2, lambda v: v[:, 2]):
        s = g(wh.astype(np.float64))
        want = np.sum(g(w) * p)
        assert abs(s.mean() - want) < 6 * s.std() / np.sqrt(n) + 1e-4, (ex, ey, s.mean(), want)
    xy = wh[:, 0].astype(np.float64) * wh[:, 1]
    assert abs(xy.mean()) < 6 * xy.std() / np.sqrt(n)
    quad = (wh[:, 0] < 0).astype(int) * 2 + (wh[:, 1] < 0)
    frac = np.bincount(quad, minlength=4) / n
    assert np.all(np.abs(frac - 0.25) < 6 * np.sqrt(0.25 * 0.75 / n))


def _substrate(depth=0.3, urough=0.05, vrough=0.2):
    return Material.substrate(rgb_refl(0.6, 0.3, 0.1), rgb_refl(0.05, 0.2, 0.4), rgb_refl(0.9, 0.5, 0.2),
                              urough, vrough, depth)


def _abi(m):
    from bling_amd import abi
    a = abi.Material()
    a.type = m.kind
    a.scalar[0], a.scalar[1], a.scalar[2] = m.s0, m.s1, m.s2
    for k in range(16):
        a.spec[0][k], a.spec[1][k], a.spec[2][k] = m.spec0[k], m.spec1[k], m.spec2[k]
    return a


def test_fresnel_blend_sample_matches_its_eval_and_pdf():
    """FresnelBlend's sample (adj False) returns e wo wi / pdf with pdf = its
    p wo wi, from either strategy (cosine-weighted for u1 < 0.5, the
    distribution's half vector otherwise); the cosine branch stays in wo's
    hemisphere."""
    rng = np.random.default_rng(5)
    n = 20000
    wo = rng.normal(size=(n, 3))
    wo /= np.linalg.norm(wo, axis=1, keepdims=True)
    u = rng.random((n, 2)).astype(np.float32)
    m = _abi(_substrate())
    wi, f, pdf = pyoracle.bxdf_sample(m, wo, u)
    ok = pdf > 0
    assert ok.mean() > 0.9
    fe, pe = pyoracle.bxdf_eval(m, wo[ok], wi[ok])
    same = wo[ok, 2] * wi[ok, 2] > 0
    assert np.all(same[u[ok, 0] < 0.5])
    assert np.allclose(pe[same], pdf[ok][same], rtol=1e-4)  # the half vector is recomputed from wi
    assert np.all(pe[~same] == 0)
    fin = np.all(np.isfinite(fe), axis=1) & np.all(np.isfinite(f[ok]), axis=1)
    assert fin.mean() > 0.99
    got = f[ok][fin].astype(np.float64) * pdf[ok][fin, None]
    assert np.allclose(got, fe[fin], rtol=1e-4, atol=1e-30)


def test_coating_absorption_is_beer_lambert():
    """The coating of thickness d scales the diffuse layer by
    exp(-ra d (1/cos_i + 1/cos_o)): the attenuation of a path into the layer
    and back out.  The specular part does not depend on d, so with A = f(d),
    B = f(0) and S = f(large d): (A - S) / (B - S) is that factor."""
    rng = np.random.default_rng(3)
    wo = rng.normal(size=(64, 3))
    wi = rng.normal(size=(64, 3))
    wo[:, 2], wi[:, 2] = np.abs(wo[:, 2]) + 0.4, np.abs(wi[:, 2]) + 0.4
    wo /= np.linalg.norm(wo, axis=1, keepdims=True)
    wi /= np.linalg.norm(wi, axis=1, keepdims=True)
    ra = np.clip(np.asarray(rgb_refl(0.9, 0.5, 0.2), np.float64), 0, 1)  # sClamp 0 1
    d = 0.25
    fa, _ = pyoracle.bxdf_eval(_abi(_substrate(depth=d)), wo, wi)
    fb, _ = pyoracle.bxdf_eval(_abi(_substrate(depth=0.0)), wo, wi)
    fs, _ = pyoracle.bxdf_eval(_abi(_substrate(depth=1e4)), wo, wi)
    ci = np.abs(wi[:, 2].astype(np.float32)).astype(np.float64)
    co = np.abs(wo[:, 2].astype(np.float32)).astype(np.float64)
    want = np.exp(-ra[None, :] * d * (1 / ci + 1 / co)[:, None])
    got = (fa.astype(np.float64) - fs) / (fb.astype(np.float64) - fs)
    big = (fb - fs) > 1e-3 * np.abs(fb)
    assert big.mean() > 0.9
    assert np.allclose(got[big], want[big], rtol=2e-3)
