"""Closed-form pins of the oracle's sun/sky light map (SunSky.hs), each shown
to fail on a mistake planted in the oracle.

mkSunSkyLight east sunDir turbidity is a map of (u, v): the direction
d = sphToDir (cartToSph (u, v)) gets Perez' sky where -d.z >= 1e-4 (the
sky's zenith is the map's -z), converted through the CIE daylight
chromaticity and sRGB to rgbToSpectrumIllum, plus the attenuated solar
spectrum inside a disk around the sun direction with its z flipped
(SunSky.hs:80-113).  What the pins fix:

  * zenith: with the sun on the horizon (sunTheta = pi/2, chi = 0) the
    zenith direction has gamma = sunTheta and Perez' ratio is 1, so the sky
    there has luminance Y_z = (2.4192 - 0.2155 T) * 1000 * 1e-4 and the
    zenith chromaticity (SunSky.hs:60-67); the CIE daylight construction
    keeps the chromaticity, checked through the spectrum's XYZ (to the
    accuracy of the sRGB -> spectrum bases);
  * the lower half of the map (-d.z < 1e-4) is black away from the sun;
  * the sun disk: cos bound sqrt(1 - 6.955e5 / 1.496e8) around
    (sd.x, sd.y, -sd.z); inside it the map is the sky plus sunR;
  * sunR: black with the sun below the horizon, positive above it, falling
    with the turbidity and with the sun's height (the air mass grows);
  * the light: the pdf integrates to 1 over the sphere and a sample's
    radiance is the map at the sample's own direction.
Parity unpinned: the per-band values of sunR beyond those orderings (the
solar and absorption curves and the transmittance forms are the
reference's own, SunSky.hs:115-158).
"""
import numpy as np
import pytest

from oracle import pyoracle
from oracle.pyoracle import mutant
from bling_amd.scene import spectrum_to_xyz

F = np.float32


def _uv_of(d):
    """sphToCart (dirToSph d) for local directions d (n, 3)."""
    d = np.asarray(d, np.float64)
    d = d / np.linalg.norm(d, axis=1, keepdims=True)
    phi = np.mod(np.arctan2(d[:, 1], d[:, 0]), 2 * np.pi)
    return np.stack([phi / (2 * np.pi), np.arccos(np.clip(d[:, 2], -1, 1)) / np.pi], 1).astype(F)


def _dir_of(uv):
    uv = np.asarray(uv, np.float64)
    ph, th = uv[:, 0] * 2 * np.pi, uv[:, 1] * np.pi
    return np.stack([np.sin(th) * np.cos(ph), np.sin(th) * np.sin(ph), np.cos(th)], 1)


def _local_sun(east, sun):
    """The sky's frame: coordinateSystem' up east, up = +y (SunSky.hs:21-22)."""
    up = np.array([0.0, 1.0, 0.0])
    e = np.asarray(east, np.float64) / np.linalg.norm(east)
    u = np.cross(e, up)
    u /= np.linalg.norm(u)
    v = np.cross(up, u)
    s = np.asarray(sun, np.float64) / np.linalg.norm(sun)
    return np.array([s @ u, s @ v, s @ up])


def pin_zenith_with_the_sun_on_the_horizon():
    for turb in (2.0, 5.0, 9.0):
        sky = [1, 0, 0, 0.3, 0.0, 1.0, turb]  # sunDir.y = 0: on the horizon
        s, d = pyoracle.sky_eval(sky, [[0.1, 1.0]])  # v = 1: straight down the map's -z
        assert d["sunTheta"] == F(np.pi / 2)
        assert abs(d["zY"] - (2.4192 - 0.2155 * turb) * 1000) < 1e-3 * abs(d["zY"])
        X, Y, Z = (float(c) for c in spectrum_to_xyz(s[0]))
        yz = d["zY"] * 1e-4
        # XYZ -> sRGB -> rgbToSpectrumIllum -> XYZ keeps the luminance within
        # 8% (the illuminant bases add 6-7%) and the chromaticity within 0.005
        assert abs(Y / yz - 1) < 0.08, (turb, Y, yz)
        assert abs(X / (X + Y + Z) - d["zx"]) < 0.005 and abs(Y / (X + Y + Z) - d["zy"]) < 0.005


def pin_lower_half_is_black():
    rng = np.random.default_rng(3)
    for sky in ([1, 0, 0, 0.2, 0.6, 1.0, 4.0], [0, 0, 1, 0.5, 1.0, 1.0, 6.0]):
        sd = _local_sun(sky[:3], sky[3:6])
        uv = np.stack([rng.random(4000), rng.random(4000) * 0.49], 1).astype(F)  # -d.z < 0
        d = _dir_of(uv)
        far = d @ np.array([sd[0], sd[1], -sd[2]]) < 0.99
        s, _ = pyoracle.sky_eval(sky, uv)
        assert np.all(s[far] == 0)
        uv[:, 1] = 0.52 + rng.random(4000) * 0.47  # -d.z > 0: the sky
        s, _ = pyoracle.sky_eval(sky, uv)
        assert np.all(s.sum(axis=1) > 0)


def pin_sun_disk():
    stm = np.sqrt(1 - 6.955e5 / 1.496e8)
    sky = [1, 0, 0, 0.4, 0.7, 1.0, 4.0]
    sd = _local_sun(sky[:3], sky[3:6])
    c = np.array([sd[0], sd[1], -sd[2]])  # the disk's centre in the map's frame
    a = np.arccos(stm)
    t = np.cross(c, [0.0, 0.0, 1.0])
    t /= np.linalg.norm(t)
    dirs = [np.cos(x) * c + np.sin(x) * t for x in (0.0, a - 2e-3, a + 2e-3)]
    s, data = pyoracle.sky_eval(sky, _uv_of(np.array(dirs)))
    sun = data["sunR"]
    assert np.all(sun > 0)
    # inside: sky + sunR; just outside: the sky alone (continuous over 4e-3 rad)
    np.testing.assert_allclose(s[1] - s[2], sun, rtol=0.02, atol=1e-3)
    np.testing.assert_allclose(s[0] - s[2], sun, rtol=0.1)


def pin_sun_spectrum_orderings():
    east = [1, 0, 0]
    _, below = pyoracle.sky_eval(east + [0.2, -0.1, 1.0, 4.0], [[0.5, 0.5]])
    assert np.all(below["sunR"] == 0)
    prev = None
    for turb in (2.0, 4.0, 8.0):
        _, d = pyoracle.sky_eval(east + [0.2, 0.8, 1.0, turb], [[0.5, 0.5]])
        assert np.all(d["sunR"] > 0)
        if prev is not None:
            assert np.all(d["sunR"] < prev)
        prev = d["sunR"]
    prev = None
    for h in (2.0, 0.8, 0.3):  # the sun sinking
        _, d = pyoracle.sky_eval(east + [0.0, h, 1.0, 4.0], [[0.5, 0.5]])
        if prev is not None:
            assert np.all(d["sunR"] < prev)
        prev = d["sunR"]


def pin_sky_light_sampling():
    sky = [1, 0, 0, 0.3, 0.5, 1.0, 4.0]
    nt, nphi = 360, 720
    th = (np.arange(nt) + 0.5) * (np.pi / nt)
    ph = (np.arange(nphi) + 0.5) * (2 * np.pi / nphi)
    T_, P_ = np.meshgrid(th, ph, indexing="ij")
    w = np.stack([np.sin(T_) * np.cos(P_), np.sin(T_) * np.sin(P_), np.cos(T_)], -1).reshape(-1, 3)
    dw = (np.sin(T_) * (np.pi / nt) * (2 * np.pi / nphi)).reshape(-1)
    pdf, le = pyoracle.skylight(sky, wi=w.astype(F))
    assert abs(np.sum(pdf.astype(np.float64) * dw) - 1.0) < 5e-3
    rng = np.random.default_rng(7)
    wi, pdf, li = pyoracle.skylight(sky, u=rng.random((20000, 2)).astype(F))
    ok = pdf > 0
    assert ok.mean() > 0.999
    s, _ = pyoracle.sky_eval(sky, _uv_of(wi[ok]))
    assert np.mean(np.all(np.isclose(s, li[ok], rtol=1e-4, atol=1e-6), axis=1)) > 0.99
    # the map's lower half is black (the sun's disk lies in the upper one):
    # Dist2D rows whose corner is at v <= 1/2 hold no mass, so no sample has d.z >= 0
    assert np.all(wi[ok][:, 2] < 0)


PINS = [pin_zenith_with_the_sun_on_the_horizon, pin_lower_half_is_black, pin_sun_disk,
        pin_sun_spectrum_orderings, pin_sky_light_sampling]


@pytest.mark.parametrize("pin", PINS, ids=lambda f: f.__name__)
def test_sky_pins(pin):
    pin()


@pytest.mark.parametrize("m,pin", [
    (mutant.SKY_UP, pin_lower_half_is_black),
    (mutant.SKY_UP, pin_zenith_with_the_sun_on_the_horizon),
    (mutant.SKY_DISK, pin_sun_disk),
    (mutant.SKY_SUN_BELOW, pin_sun_spectrum_orderings),
], ids=lambda x: str(x) if isinstance(x, int) else "")
def test_each_sky_pin_fails_on_its_planted_mistake(m, pin):
    with mutant(m):
        with pytest.raises(AssertionError):
            pin()
    pin()
