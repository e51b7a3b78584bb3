"""Pin the oracle's refinement (intersects, intersection area, DE-9IM) and point-in-polygon against exact rational
geometry and textbook answers, on small random and hand-built configurations placed at lon/lat coordinates."""
import math
import random
from types import SimpleNamespace

import numpy as np
import pytest

from oracle import oracle as O
import exact as X

X0, Y0 = -95.5, 38.5


def hostset(polys):
    """polys: list of polygons, each a list of closed rings [(x, y), ...]"""
    gro, rvo, xy = [0], [0], []
    for rings in polys:
        for r in rings:
            xy.extend(r)
            rvo.append(len(xy))
        gro.append(len(rvo) - 1)
    return SimpleNamespace(gro=np.array(gro, np.uint64), rvo=np.array(rvo, np.uint64), xy=np.array(xy, float).reshape(-1, 2),
                           parent=None, syn=None)


def ring(pts):
    pts = [(float(x), float(y)) for x, y in pts]
    return pts + [pts[0]]


def convex(rng, cx, cy, r, k):
    angs = sorted(rng.uniform(0, 2 * math.pi) for _ in range(k))
    return ring([(cx + r * math.cos(a), cy + r * math.sin(a)) for a in angs])


def star(rng, cx, cy, r, k):
    """simple star-shaped polygon: one radius per vertex at evenly spaced angles (CCW)"""
    pts = []
    for i in range(k):
        ri = r * rng.uniform(0.3, 1.0)
        pts.append((cx + ri * math.cos(2 * math.pi * i / k), cy + ri * math.sin(2 * math.pi * i / k)))
    return ring(pts)


def refine(polysA, polysB, pairs, **kw):
    A, B = hostset(polysA), hostset(polysB)
    return O.refine(O.Layer.polys(A), O.Layer.polys(B), np.array(pairs, np.uint32).reshape(-1, 2), **kw)


def exact_intersects(ra, rb):
    for i in range(len(ra) - 1):
        for j in range(len(rb) - 1):
            if X.segments_intersect(ra[i], ra[i + 1], rb[j], rb[j + 1]):
                return True
    return X.point_in_rings(ra[0], [rb]) > 0 or X.point_in_rings(rb[0], [ra]) > 0


def test_area_and_intersects_random_convex():
    rng = random.Random(11)
    A, B = [], []
    for _ in range(400):
        cx, cy = X0 + rng.uniform(0, 0.01), Y0 + rng.uniform(0, 0.01)
        A.append([convex(rng, cx, cy, rng.uniform(1e-4, 5e-4), rng.randint(3, 12))])
        B.append([convex(rng, cx + rng.uniform(-6e-4, 6e-4), cy + rng.uniform(-6e-4, 6e-4), rng.uniform(1e-4, 5e-4), rng.randint(3, 12))])
    pairs = [(i, i) for i in range(len(A))]
    isect, area, _, _ = refine(A, B, pairs)
    for i in range(len(A)):
        ex = float(X.clip_convex(A[i][0], B[i][0]))
        scale = float(X.shoelace(A[i][0]))
        assert abs(area[i] - ex) <= 1e-12 * scale, (i, area[i], ex)
        assert (isect[i] == 1) == exact_intersects(A[i][0], B[i][0]), i


def test_area_and_intersects_star_vs_convex():
    rng = random.Random(12)
    A, B = [], []
    for _ in range(300):
        cx, cy = X0 + rng.uniform(0, 0.01), Y0 + rng.uniform(0, 0.01)
        A.append([star(rng, cx, cy, 4e-4, rng.randint(5, 20))])
        B.append([convex(rng, cx + rng.uniform(-5e-4, 5e-4), cy + rng.uniform(-5e-4, 5e-4), rng.uniform(1e-4, 4e-4), rng.randint(3, 9))])
    isect, area, _, _ = refine(A, B, [(i, i) for i in range(len(A))])
    for i in range(len(A)):
        ex = float(X.clip_convex(A[i][0], B[i][0]))   # star subject clipped by the convex clipper
        assert abs(area[i] - ex) <= 1e-12 * float(X.shoelace(A[i][0])), i
        assert (isect[i] == 1) == exact_intersects(A[i][0], B[i][0]), i


def test_area_closed_forms():
    sq = ring([(X0, Y0), (X0 + 1e-3, Y0), (X0 + 1e-3, Y0 + 1e-3), (X0, Y0 + 1e-3)])
    inner = ring([(X0 + 2e-4, Y0 + 3e-4), (X0 + 6e-4, Y0 + 3e-4), (X0 + 5e-4, Y0 + 7e-4)])
    right = ring([(X0 + 1e-3, Y0), (X0 + 2e-3, Y0), (X0 + 2e-3, Y0 + 1e-3), (X0 + 1e-3, Y0 + 1e-3)])
    A = [[sq], [inner], [sq], [sq]]
    B = [[sq], [sq], [right], [inner]]
    isect, area, _, _ = refine(A, B, [(0, 0), (1, 1), (2, 2), (3, 3)])
    a_sq, a_in = float(X.shoelace(sq)), float(X.shoelace(inner))
    assert isect.tolist() == [1, 1, 1, 1]
    assert abs(area[0] - a_sq) <= 1e-15 * a_sq          # identical polygons: own area
    assert abs(area[1] - a_in) <= 1e-15 * a_sq          # contained: the inner area
    assert abs(area[2]) <= 1e-15 * a_sq                 # edge-adjacent: zero area, but intersecting
    assert abs(area[3] - a_in) <= 1e-15 * a_sq


def test_area_with_hole_and_tiling():
    # square with a square hole, against a polygon covering the hole and part of the ring
    outer = ring([(X0, Y0), (X0 + 4e-3, Y0), (X0 + 4e-3, Y0 + 4e-3), (X0, Y0 + 4e-3)])
    hole = ring([(X0 + 1e-3, Y0 + 1e-3), (X0 + 1e-3, Y0 + 3e-3), (X0 + 3e-3, Y0 + 3e-3), (X0 + 3e-3, Y0 + 1e-3)])   # CW
    cover = ring([(X0 + 5e-4, Y0 + 5e-4), (X0 + 3.5e-3, Y0 + 5e-4), (X0 + 3.5e-3, Y0 + 2e-3), (X0 + 5e-4, Y0 + 2e-3)])
    isect, area, _, _ = refine([[outer, hole]], [[cover]], [(0, 0)])
    # exact: cover within the outer square, minus cover within the hole (both convex clips of the stored floats)
    ex = float(X.clip_convex(cover, outer) - X.clip_convex(cover, hole[::-1]))
    assert isect[0] == 1 and abs(area[0] - ex) <= 1e-12 * ex
    # tiling: a parcel split by a vertical line into two regions; the two areas sum to the parcel's
    rng = random.Random(13)
    for _ in range(50):
        P = star(rng, X0 + 1e-3, Y0, 4e-4, rng.randint(5, 12))
        m = X0 + 1e-3 + rng.uniform(-2e-4, 2e-4)
        Ra = ring([(X0, Y0 - 1e-3), (m, Y0 - 1e-3), (m, Y0 + 1e-3), (X0, Y0 + 1e-3)])
        Rb = ring([(m, Y0 - 1e-3), (X0 + 2e-3, Y0 - 1e-3), (X0 + 2e-3, Y0 + 1e-3), (m, Y0 + 1e-3)])
        _, ar, _, _ = refine([[P]], [[Ra], [Rb]], [(0, 0), (0, 1)])
        whole = float(X.shoelace(P))
        assert abs(ar[0] + ar[1] - whole) <= 1e-12 * whole


DE9 = {"identical": "2FFF1FFF2", "edge": "FF2F11212", "corner": "FF2F01212", "overlap": "212101212", "within": "2FF1FF212",
       "contains": "212FF1FF2", "within_touch": "2FF11F212"}


def de9(code):
    return "".join("F012"[(int(code) >> (2 * k)) & 3] for k in range(9))


def test_de9im_textbook_matrices():
    s = 1e-3
    sq = ring([(X0, Y0), (X0 + s, Y0), (X0 + s, Y0 + s), (X0, Y0 + s)])
    right = ring([(X0 + s, Y0), (X0 + 2 * s, Y0), (X0 + 2 * s, Y0 + s), (X0 + s, Y0 + s)])
    diag = ring([(X0 + s, Y0 + s), (X0 + 2 * s, Y0 + s), (X0 + 2 * s, Y0 + 2 * s), (X0 + s, Y0 + 2 * s)])
    shifted = ring([(X0 + s / 2, Y0 + s / 2), (X0 + 1.5 * s, Y0 + s / 2), (X0 + 1.5 * s, Y0 + 1.5 * s), (X0 + s / 2, Y0 + 1.5 * s)])
    small = ring([(X0 + s / 4, Y0 + s / 4), (X0 + s / 2, Y0 + s / 4), (X0 + s / 2, Y0 + s / 2), (X0 + s / 4, Y0 + s / 2)])
    touching = ring([(X0, Y0 + s / 4), (X0 + s / 2, Y0 + s / 4), (X0 + s / 2, Y0 + s / 2), (X0, Y0 + s / 2)])   # inside, on A's left edge
    cases = [("identical", sq, sq), ("edge", sq, right), ("corner", sq, diag), ("overlap", sq, shifted), ("within", small, sq),
             ("contains", sq, small), ("within_touch", touching, sq)]
    A = [[c[1]] for c in cases]
    B = [[c[2]] for c in cases]
    A_, B_ = hostset(A), hostset(B)
    LA, LB = O.Layer.polys(A_), O.Layer.polys(B_)
    isect, area, ev, evb = O.refine(LA, LB, np.array([(i, i) for i in range(len(cases))], np.uint32), want_area=True, want_relate=True)
    blA, blB = LA.export()["blen"], LB.export()["blen"]
    for i, (name, _, _) in enumerate(cases):
        assert isect[i] == 1, name
        assert de9(O.de9im_code(ev[i], int(evb[i]), blA[i], blB[i])) == DE9[name], name


def test_de9im_transpose_and_match_tables():
    import gsj
    rng = random.Random(14)
    for _ in range(2000):
        c = sum(rng.randrange(4) << (2 * k) for k in range(9))
        s, t = de9(c), de9(gsj.de9im_transpose(c))
        assert t == "".join(s[3 * (k % 3) + k // 3] for k in range(9))          # matrix transpose of the 3x3 string
        assert gsj.de9im_transpose(gsj.de9im_transpose(c)) == c
        pat = "".join(rng.choice("TF*012") for _ in range(9))
        want = all(p == "*" or (p == "T" and ch != "F") or (p == "F" and ch == "F") or p == ch for p, ch in zip(pat, s))
        assert gsj.de9im_match(c, pat) == want
        assert gsj.de9im_match(c, pat.lower()) == want


def test_point_in_polygon_exact():
    rng = random.Random(15)
    polys = [[star(rng, X0 + 1e-3 * (i % 10), Y0 + 1e-3 * (i // 10), 4e-4, rng.randint(4, 15))] for i in range(100)]
    hs = hostset(polys)
    R = O.Layer.polys(hs)
    pts = []
    for i, p in enumerate(polys):
        r = p[0]
        for _ in range(20):
            pts.append((X0 + 1e-3 * (i % 10) + rng.uniform(-5e-4, 5e-4), Y0 + 1e-3 * (i // 10) + rng.uniform(-5e-4, 5e-4)))
        pts.extend(r[:3])                                            # vertices: on the boundary
        pts.append(((r[0][0] + r[1][0]) / 2, (r[0][1] + r[1][1]) / 2))  # near an edge
    P = O.Layer.points_layer(np.array(pts))
    for k, pt in enumerate(pts):
        g = min(len(polys) - 1, k // 24)
        assert O.pip_pair(P, R, k, g) == X.point_in_rings(pt, polys[g]), (k, g)
