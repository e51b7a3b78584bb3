"""CPU tests pinning the oracle's eroded intersects (ST_Intersects(ST_Buffer(A::geography, -ra), ST_Buffer(B, -rb)))
by closed forms: rectangles erode to rectangles (no fillets), so two overlapping rectangles' eroded polygons meet iff
the overlap exceeds ra + rb (in the UTM frame, r / k true metres); a rectangle eroded past half its width vanishes;
points near a reflex vertex are excluded inside the GEOS fillet chords and kept outside the arc; the eroded polygon of
a parent does not depend on how it is subdivided."""
import math

import numpy as np
import pytest

import gsj
from oracle import oracle as O

LAT = 34.0
KX, KY = O.m_per_deg(LAT)
K0 = 0.9996   # UTM scale on the central meridian (-117: zone 11)


def _set(rings):
    xy = np.array([p for r in rings for p in r], float)
    n = len(rings)
    return gsj.HostSet(np.arange(n + 1, dtype=np.uint64), np.cumsum([0] + [len(r) for r in rings]).astype(np.uint64), xy)


def _rect(x0, y0, w_m, h_m):
    w, h = w_m / KX, h_m / KY
    return [(x0, y0), (x0 + w, y0), (x0 + w, y0 + h), (x0, y0 + h), (x0, y0)]


def _two_rects(o_m, w_m=200.0, h_m=100.0):
    x0 = -117.0 - w_m / KX / 2
    return _set([_rect(x0, LAT, w_m, h_m), _rect(x0 + (w_m - o_m) / KX, LAT, w_m, h_m)])


@pytest.mark.parametrize("ra,rb", [(2.0, 1.0), (1.0, 2.0), (0.0, 2.9), (1.5, 0.0), (0.6, 2.4)])
def test_overlapping_rectangles_meet_iff_overlap_exceeds_the_radii(ra, rb):
    t = (ra + rb) / K0   # true metres in the frame at the zone's central meridian
    for f, expect in ((0.99, False), (0.998, False), (1.002, True), (1.02, True)):
        L = O.Layer.polys(_two_rects(t * f))
        assert O.er_pair(L, L, 0, 1, ra, rb) == expect, (ra, rb, f)
        rows = {tuple(p) for p in O.erode(L, L, ra, rb)["pairs"].tolist()}
        assert ((0, 1) in rows) == expect and ((1, 0) in rows) == expect


def test_a_rectangle_eroded_past_half_its_height_vanishes():
    L = O.Layer.polys(_two_rects(50.0))   # 200 m x 100 m
    assert O.er_pair(L, L, 0, 0, 49.0, 49.0) and not O.er_pair(L, L, 0, 0, 51.0, 51.0)
    # need_self: a geometry's pair with itself only where requested
    r = O.erode(L, L, 5.0, 5.0, same=True, upper=True, need_self=np.array([0, 1], np.uint8))
    assert {tuple(p) for p in r["pairs"].tolist()} == {(0, 1), (1, 1)}


def test_reflex_vertex_fillet_chords():
    """L-shaped polygon with a 270-degree interior angle at (x0 + 100 m, y0 + 100 m): around that reflex vertex the
    eroded boundary follows round(theta / (pi / 16)) chords inscribed in the radius-r arc (theta = pi / 2 -> 8 chords);
    points on the bisector closer than the chords' distance r cos(pi / 32) are excluded, points beyond r are kept."""
    x0, y0 = -117.0, LAT
    a, b = 200.0, 100.0

    def P(xm, ym):
        return (x0 + xm / KX, y0 + ym / KY)

    ring = [P(0, 0), P(a, 0), P(a, b), P(b, b), P(b, a), P(0, a), P(0, 0)]
    R = O.Layer.polys(_set([ring]))
    r = 10.0
    rt = r / K0   # the point scale factor is ~0.9996 here (within 1e-6 over this patch)
    chord = rt * math.cos(math.pi / 32)
    # the reflex vertex (b, b); the interior bisector points away from the notch, towards (b - t, b - t)
    pts = []
    for dist in (0.97 * chord, 0.995 * chord, 1.003 * rt, 1.05 * rt):
        t = dist / math.sqrt(2)
        pts.append(P(b - t, b - t))
    pts.append(P(50, 50))          # deep inside: kept
    pts.append(P(b - 5, 150))      # within 10 m of the edge x = b: excluded
    res = O.points_erode(O.Layer.points_layer(np.array(pts)), R, r)
    kept = {int(p) for p, _ in res["pairs"].tolist()}
    assert kept == {2, 3, 4}


def test_erosion_does_not_depend_on_subdivision():
    M = gsj.generate("parcels_mesh", nx=24, ny=24, x0=-95.2, y0=38.6, w=0.6 / 1095, h=0.6 / 1095, seed=7)
    S = gsj.subdivide(M, 6)
    assert S.n > M.n
    for ra, rb in ((1.0, 1.0), (0.5, 2.5), (0.0, 1.5)):
        a = O.erode(O.Layer.polys(M), O.Layer.polys(M), ra, rb, same=True, upper=True)
        b = O.erode(O.Layer.polys(S), O.Layer.polys(S), ra, rb, same=True, upper=True)
        assert {tuple(p) for p in a["pairs"].tolist()} == {tuple(p) for p in b["pairs"].tolist()}
        assert 0 < len(a["pairs"]) < a["candidates"]


POLAR = 75.0


def _rects_at(lat, lon0, o_m, w_m=200.0, h_m=20.0):
    """two w x h metre rectangles at (lon0, lat) overlapping by o_m metres east-west (metres of the local degree
    lengths at lat)"""
    kx, ky = O.m_per_deg(lat)

    def rect(x0):
        w, h = w_m / kx, h_m / ky
        return [(x0, lat), (x0 + w, lat), (x0 + w, lat + h), (x0, lat + h), (x0, lat)]

    x0 = lon0 - w_m / kx / 2
    return _set([rect(x0), rect(x0 + (w_m - o_m) / kx)])


@pytest.mark.parametrize("ra,rb", [(2.0, 1.0), (0.0, 2.5), (1.5, 1.5)])
def test_polar_erosion_takes_scale_one(ra, rb):
    """Above 70 degrees PostGIS's _ST_BestSRID picks a Lambert azimuthal equal-area projection, whose scale the
    reference takes as 1 (docs/technical-details.md:468-473, engine/buffer.cuh:55-57), and so does this build: at
    75 degrees eroded rectangles meet iff their overlap exceeds ra + rb true metres, while on a UTM zone's central
    meridian the threshold is (ra + rb) / 0.9996. An overlap of 1.0002 (ra + rb) therefore meets at 75 degrees and
    does not at 34 degrees on the meridian (-117)."""
    t = ra + rb
    for f, expect in ((0.999, False), (1.0002, True), (1.001, True)):
        L = O.Layer.polys(_rects_at(POLAR, 20.0, t * f))
        assert O.er_pair(L, L, 0, 1, ra, rb) == expect, (ra, rb, f)
    L = O.Layer.polys(_rects_at(LAT, -117.0, t * 1.0002))
    assert not O.er_pair(L, L, 0, 1, ra, rb)
    L = O.Layer.polys(_rects_at(LAT, -117.0, t / K0 * 1.0002))
    assert O.er_pair(L, L, 0, 1, ra, rb)
