#!/usr/bin/env python3
"""Writes tests/golden/intersection_cases.json: ray/primitive cases whose
results follow in closed form from the reference's formulas
(Primitive/TriangleMesh.hs:160-207 Moeller-Trumbore, Shape.hs:81-284 shapes,
Math.hs:124-139 solveQuadric), with every quantity dyadic so binary32
arithmetic is exact.  Expected values are computed here with exact rationals;
nothing is computed by the oracle or the GPU.

    python3 tests/golden/make_intersection_cases.py
"""
import json
import os
from fractions import Fraction as Fr

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "intersection_cases.json")
INF = "inf"


def det3(m):
    return (m[0][0] * (m[1][1] * m[2][2] - m[1][2] * m[2][1])
            - m[0][1] * (m[1][0] * m[2][2] - m[1][2] * m[2][0])
            + m[0][2] * (m[1][0] * m[2][1] - m[1][1] * m[2][0]))


def tri_case(name, p1, p2, p3, o, d, tmin=0, tmax=INF):
    """Solves o + t d = p1 + b1 (p2 - p1) + b2 (p3 - p1) exactly (Cramer's
    rule); a hit needs b1, b2 >= 0, b1 + b2 <= 1 and t in [tmin, tmax]."""
    P1, P2, P3, O, D = ([Fr(v) for v in p] for p in (p1, p2, p3, o, d))
    e1 = [P2[i] - P1[i] for i in range(3)]
    e2 = [P3[i] - P1[i] for i in range(3)]
    rhs = [O[i] - P1[i] for i in range(3)]
    cols = [e1, e2, [-v for v in D]]
    A = [[cols[c][r] for c in range(3)] for r in range(3)]
    den = det3(A)
    exp = None
    if den != 0:
        sol = []
        for c in range(3):
            Ac = [[rhs[r] if cc == c else A[r][cc] for cc in range(3)] for r in range(3)]
            sol.append(det3(Ac) / den)
        b1, b2, t = sol
        ok = b1 >= 0 and b2 >= 0 and b1 + b2 <= 1 and t >= Fr(tmin) and (tmax == INF or t <= Fr(tmax))
        if ok:
            exp = {"t": float(t), "b1": float(b1), "b2": float(b2)}
    return {"name": name, "kind": "triangle", "verts": [p1, p2, p3], "ray": [o, d, tmin, tmax],
            "expect": exp}


def shape_case(name, shape, params, translate, o, d, expect_t, any_hit, tmin=0, tmax=INF):
    return {"name": name, "kind": "shape", "shape": shape, "params": params, "translate": translate,
            "ray": [o, d, tmin, tmax], "expect": None if expect_t is None else {"t": expect_t},
            "any": any_hit}


def main():
    T = [(0, 0, 0), (1, 0, 0), (0, 1, 0)]
    cases = []
    for x, y in ((0.25, 0.5), (0.125, 0.375), (0.0, 0.5), (0.5, 0.5), (0.75, 0.0)):
        cases.append(tri_case("tri_bary_%g_%g" % (x, y), *T, (x, y, -2), (0, 0, 1)))
    cases.append(tri_case("tri_swapped_winding", T[0], T[2], T[1], (0.25, 0.5, -2), (0, 0, 1)))
    cases.append(tri_case("tri_unnormalised_dir", *T, (0.25, 0.5, -2), (0, 0, 2)))
    cases.append(tri_case("tri_backface", *T, (0.25, 0.5, 2), (0, 0, -1)))
    cases.append(tri_case("tri_outside_edge", *T, (-0.0009765625, 0.5, -2), (0, 0, 1)))
    cases.append(tri_case("tri_beyond_tmax", *T, (0.25, 0.5, -2), (0, 0, 1), 0, 1.5))
    cases.append(tri_case("tri_before_tmin", *T, (0.25, 0.5, -2), (0, 0, 1), 2.5, INF))
    cases.append(tri_case("tri_parallel", *T, (0.25, 0.5, -2), (1, 0, 0)))
    # sphere: a=|d|^2, b=2 o.d, c=|o|^2-r^2 (Shape.hs:173-187)
    cases.append(shape_case("sphere_outside", "sphere", [1.0], [0, 0, 0], (0, 0, -4), (0, 0, 1), 3.0, True))
    cases.append(shape_case("sphere_inside", "sphere", [1.0], [0, 0, 0], (0, 0, 0.5), (0, 0, 1), 0.5, True))
    cases.append(shape_case("sphere_translated", "sphere", [0.5], [2, 0, 0], (2, 0, -3), (0, 0, 1), 2.5, True))
    cases.append(shape_case("sphere_miss", "sphere", [1.0], [0, 0, 0], (2, 0, -4), (0, 0, 1), None, False))
    # quad: z = 0 plane, |x| <= sx, |y| <= sy (Shape.hs:157-171)
    cases.append(shape_case("quad_hit", "quad", [1.0, 0.5], [0, 0, 0], (0.25, 0.25, 2), (0, 0, -1), 2.0, True))
    cases.append(shape_case("quad_outside", "quad", [1.0, 0.5], [0, 0, 0], (0.25, 0.75, 2), (0, 0, -1), None, False))
    cases.append(shape_case("quad_parallel", "quad", [1.0, 0.5], [0, 0, 0], (0.25, 0.25, 2), (1, 0, 0), None, False))
    # disk (Shape.hs:142-155): height 0, radius 1, inner 0.5, full circle
    cases.append(shape_case("disk_ring", "disk", [0.0, 1.0, 0.5, 6.2831854820251465], [0, 0, 0],
                            (0.75, 0, 3), (0, 0, -1), 3.0, True))
    cases.append(shape_case("disk_hole", "disk", [0.0, 1.0, 0.5, 6.2831854820251465], [0, 0, 0],
                            (0.25, 0, 3), (0, 0, -1), None, False))
    # box (Shape.hs:86-111): entry face; a ray starting inside misses intersect
    # (t0 < tmin) while intersects (intersectAABB) reports the box
    cases.append(shape_case("box_entry", "box", [-1, -1, -1, 1, 1, 1], [0, 0, 0], (0, 0, -5), (0, 0, 1), 4.0, True))
    cases.append(shape_case("box_from_inside", "box", [-1, -1, -1, 1, 1, 1], [0, 0, 0], (0, 0, 0), (0, 0, 1), None, True))
    # cylinder (Shape.hs:113-140): radius 1, z in [-1, 1]
    cases.append(shape_case("cylinder_side", "cylinder", [1.0, -1.0, 1.0, 6.2831854820251465], [0, 0, 0],
                            (-5, 0, 0), (1, 0, 0), 4.0, True))
    cases.append(shape_case("cylinder_above", "cylinder", [1.0, -1.0, 1.0, 6.2831854820251465], [0, 0, 0],
                            (-5, 0, 2), (1, 0, 0), None, False))
    with open(OUT, "w") as f:
        json.dump(cases, f, indent=1)
    print("wrote %d cases to %s" % (len(cases), OUT))


if __name__ == "__main__":
    main()
