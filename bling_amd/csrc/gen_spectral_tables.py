#!/usr/bin/env python3
"""Generate the 16-band spectral constant tables used by the film and by scene
flattening.

The reference bins the CIE 1931 colour matching functions (360..830 nm, 1 nm
steps) and Smits-style RGB->spectrum basis functions (380..720 nm, 32 samples)
into 16 bands over 400..700 nm with `fromSpd`/`avgSpd`
(src/lib/Graphics/Bling/Spectrum.hs:291-335), then derives `spectrumCieYSum`
(Spectrum.hs:346-347).  This script restates that binning in IEEE binary32
arithmetic in the same operation order (Haskell `Float`), reading the source
lists from the reference checkout, and writes the binned results, the CIE
daylight basis functions' XYZ (`s0XYZ`..`s2XYZ`, Spectrum.hs:234-241, summed
over 360..830 nm at 1 nm as `spdToXYZ` does, :319-326) and, unchanged, the
sun/sky model's solar spectrum and absorption curves (SunSky.hs:122-158),
which the library evaluates per scene:

  bling_amd/csrc/spectral_tables.h   (C header, hex-float literals)
  bling_amd/spectral_tables.py       (the same numbers for the Python host side)

Apart from those four curves only derived values are written; the raw lists
are read at generation time.  Run:  python3 bling_amd/csrc/gen_spectral_tables.py
"""
import os
import re
import sys
from fractions import Fraction

import numpy as np

REF = "/root/reference/src/lib/Graphics/Bling/Spectrum.hs"
REF_SKY = "/root/reference/src/lib/Graphics/Bling/SunSky.hs"
HERE = os.path.dirname(os.path.abspath(__file__))
OUT_H = os.path.join(HERE, "spectral_tables.h")
OUT_PY = os.path.join(os.path.dirname(HERE), "spectral_tables.py")

F = np.float32
BANDS = 16


def f32_exact(text):
    """Round a decimal literal to the nearest binary32 (ties to even), as GHC's
    `fromRational :: Rational -> Float` does for a Float literal."""
    q = Fraction(text)
    d = float(q)                       # nearest double
    a = np.float32(d)
    # double rounding is only wrong when d sits exactly between two floats;
    # resolve every candidate against the exact rational instead.
    cands = {a, np.nextafter(a, F(np.inf)), np.nextafter(a, F(-np.inf))}
    best = None
    for c in cands:
        err = abs(Fraction(float(c)) - q)
        if best is None or err < best[0] or (
                err == best[0] and (int(np.float32(c).view(np.uint32)) & 1) == 0):
            best = (err, c)
    return F(best[1])


def parse_list(src, name):
    m = re.search(r"^%s\s*=\s*\[(.*?)\]" % re.escape(name), src, re.S | re.M)
    if not m:
        raise SystemExit("list %s not found" % name)
    body = re.sub(r"--[^\n]*", "", m.group(1))
    return [f32_exact(t) for t in re.findall(r"-?[0-9][0-9.eE+\-]*", body)]


def parse_rgbfunc(src, name):
    m = re.search(r"^%s\s*=\s*rgbFunc\s*\[(.*?)\]" % re.escape(name), src, re.S | re.M)
    if not m:
        raise SystemExit("rgbFunc %s not found" % name)
    return [f32_exact(t) for t in re.findall(r"-?[0-9][0-9.eE+\-]*", m.group(1))]


def lerp(t, a, b):                     # Math.hs:108-110
    return F(F(F(1) - t) * a) + F(t * b)


def floor_i(x):
    return int(np.floor(np.float64(x)))


def avg_regular(s0, s1, amps, l0, l1):  # Spectrum.hs:297-305
    n = len(amps)
    if l1 <= s0:
        return amps[0]
    if l0 >= s1:
        return amps[-1]
    fn = F(n)
    i0 = max(0, min(n, floor_i(F(fn * F(F(l0 - s0) / F(s1 - s0))))))
    i1 = max(0, min(n, floor_i(F(fn * F(F(l1 - s0) / F(s1 - s0))))))
    sl = amps[i0:i1 + 1]
    acc = F(0)
    for v in sl:                       # Data.Vector sum = strict left fold
        acc = F(acc + v)
    return F(acc / F(len(sl)))


def eval_regular(s0, s1, amps, l):      # Spectrum.hs:271-280
    if l <= s0:
        return amps[0]
    if l >= s1:
        return amps[-1]
    n = len(amps)
    d1 = F(F(1) / F(F(s1 - s0) / F(n - 1)))
    x = F(F(l - s0) * d1)
    b0 = floor_i(x)
    b1 = min(b0 + 1, n - 1)
    dx = F(x - F(b0))
    return F(F(F(F(1) - dx) * amps[b0]) + F(dx * amps[b1]))


def spd_to_xyz(cie, s0, s1, amps):     # Spectrum.hs:319-326
    xs = ys = zs = yint = F(0)
    for l in range(360, 831):
        v = eval_regular(F(s0), F(s1), amps, F(l))
        cx, cy, cz = (cie[k][l - 360] for k in "XYZ")
        yint = F(yint + cy)
        xs = F(xs + F(cx * v))
        ys = F(ys + F(cy * v))
        zs = F(zs + F(cz * v))
    return [F(xs / yint), F(ys / yint), F(zs / yint)]


def parse_regular(src, name):
    """`name = mkSpd' lo hi [..]` -> (lo, hi, amplitudes)."""
    m = re.search(r"^%s\s*=\s*mkSpd'\s+(\S+)\s+(\S+)\s*\[(.*?)\]" % re.escape(name), src, re.S | re.M)
    if not m:
        raise SystemExit("regular SPD %s not found" % name)
    return (f32_exact(m.group(1)), f32_exact(m.group(2)),
            [f32_exact(t) for t in re.findall(r"-?[0-9][0-9.eE+\-]*", m.group(3))])


def parse_pairs(src, name):
    """`name = mkSpd $ zip ls as` with the lists inline or in a where clause."""
    m = re.search(r"^%s\s*=\s*mkSpd \$ zip (.*?)(?=^\S)" % re.escape(name), src + "\nEND", re.S | re.M)
    if not m:
        raise SystemExit("irregular SPD %s not found" % name)
    body = m.group(1)
    lists = re.findall(r"\[(.*?)\]", body, re.S)
    if len(lists) != 2:
        raise SystemExit("irregular SPD %s: expected two lists" % name)
    ls, vs = ([f32_exact(t) for t in re.findall(r"-?[0-9][0-9.eE+\-]*", x)] for x in lists)
    assert len(ls) == len(vs), name
    return ls, vs


def from_spd(s0, s1, amps):            # Spectrum.hs:329-335
    out = []
    for i in range(BANDS):
        l0 = lerp(F(F(i) / F(BANDS)), F(400), F(700))
        l1 = lerp(F(F(i + 1) / F(BANDS)), F(400), F(700))
        out.append(avg_regular(F(s0), F(s1), amps, l0, l1))
    return out


def fhex(x):
    return float(x).hex()


def main():
    if not os.path.exists(REF):
        sys.exit("reference checkout not found at %s" % REF)
    src = open(REF).read()
    cie = {k: parse_list(src, "cie%sValues" % k) for k in "XYZ"}
    for k, v in cie.items():
        assert len(v) == 471, (k, len(v))
    tables = {}
    for k in "XYZ":
        tables["cie_" + k.lower()] = from_spd(360, 830, cie[k])
    ysum = F(0)
    for v in tables["cie_y"]:
        ysum = F(ysum + v)
    names = ["White", "Cyan", "Magenta", "Yellow", "Red", "Green", "Blue"]
    for kind in ("Refl", "Illum"):
        for nm in names:
            lst = parse_rgbfunc(src, "rgb%s%s" % (kind, nm))
            assert len(lst) == 32, (kind, nm, len(lst))
            tables["rgb_%s_%s" % (kind.lower(), nm.lower())] = from_spd(380, 720, lst)
    daylight = {}
    for k in range(3):
        lo, hi, amps = parse_regular(src, "cieS%d" % k)
        assert (lo, hi, len(amps)) == (300, 830, 54), k
        daylight["s%d_xyz" % k] = spd_to_xyz(cie, lo, hi, amps)
    if not os.path.exists(REF_SKY):
        sys.exit("reference checkout not found at %s" % REF_SKY)
    sky = open(REF_SKY).read()
    lo, hi, sol = parse_regular(sky, "solCurve")
    assert (lo, hi, len(sol)) == (380, 750, 38)
    curves = {"sol": sol}
    for nm, n in (("ko", 64), ("kg", 4), ("kwa", 13)):
        ls, vs = parse_pairs(sky, nm + "Curve")
        assert len(ls) == n, (nm, len(ls))
        assert all(a < b for a, b in zip(ls, ls[1:])), nm
        curves[nm + "_l"] = ls
        curves[nm + "_a"] = vs

    with open(OUT_H, "w") as fh:
        fh.write("/* Generated by gen_spectral_tables.py: 16-band binnings (400..700 nm)\n"
                 " * of the CIE 1931 matching functions and the RGB->spectrum bases, in\n"
                 " * binary32, following Spectrum.hs:291-347.  Do not edit. */\n")
        fh.write("#pragma once\n#define BGPU_BANDS 16\n")
        for k, v in tables.items():
            fh.write("static const float BGPU_%s[16] = {%s};\n"
                     % (k.upper(), ", ".join(fhex(x) for x in v)))
        fh.write("static const float BGPU_CIE_Y_SUM = %s;\n" % fhex(ysum))
        for k, v in daylight.items():
            fh.write("static const float BGPU_%s[3] = {%s};\n" % (k.upper(), ", ".join(fhex(x) for x in v)))
        fh.write("/* SunSky.hs:122-158: solCurve (380..750 nm, regular), and the ozone,\n"
                 " * mixed-gas and water-vapour absorption curves (wavelength, value). */\n")
        for k, v in curves.items():
            fh.write("static const float BGPU_SKY_%s[%d] = {%s};\n"
                     % (k.upper(), len(v), ", ".join(fhex(x) for x in v)))
    with open(OUT_PY, "w") as fp:
        fp.write('"""Generated by csrc/gen_spectral_tables.py: 16-band spectral constants\n'
                 '(CIE 1931 matching functions and RGB->spectrum bases binned to 400..700 nm,\n'
                 'the daylight basis functions\' XYZ), binary32 values as hex floats.  Do not edit."""\n')
        fp.write("BANDS = 16\nTABLES = {\n")
        for k, v in tables.items():
            fp.write("    %r: [%s],\n" % (k, ", ".join("float.fromhex(%r)" % fhex(x) for x in v)))
        fp.write("}\nCIE_Y_SUM = float.fromhex(%r)\n" % fhex(ysum))
        for k, v in daylight.items():
            fp.write("%s = [%s]\n" % (k.upper(), ", ".join("float.fromhex(%r)" % fhex(x) for x in v)))
    print("wrote", OUT_H, OUT_PY)


if __name__ == "__main__":
    main()
