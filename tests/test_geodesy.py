"""CPU tests pinning the portable elementary functions and the WGS84 geodesy that the distance joins share between
the engine and the oracle: against mpmath at high precision, closed forms, and published test lines."""
import math
import random

import mpmath as mp
import numpy as np
import pytest

from oracle import oracle as O

A = 6378137.0
F = 1 / 298.257223563


def ulps(got, exact):
    exact = float(exact)
    if exact == 0:
        return abs(got) / math.ulp(0.0)
    return abs(got - exact) / math.ulp(exact)


# ------------------------------------------------------------------------------------------- elementary functions
def test_cody_waite_split_of_half_pi():
    mp.mp.prec = 300
    p1, p2, p3 = 1.5707963267341256, 6.077100506303966e-11, 2.0222662487959506e-21
    assert abs(mp.mpf(p1) + mp.mpf(p2) + mp.mpf(p3) - mp.pi / 2) < mp.mpf(2) ** -120
    for p in (p1, p2):   # 33 significant bits: k * part is exact for |k| < 2^20
        m, e = math.frexp(p)
        assert (m * 2 ** 33).is_integer()


@pytest.mark.parametrize("name,lo,hi", [("sin", -7.0, 7.0), ("cos", -7.0, 7.0), ("sin", -1e-3, 1e-3), ("cos", 1.4, 1.75),
                                        ("tan", -1.5, 1.5), ("sin", -400.0, 400.0), ("cos", -400.0, 400.0)])
def test_sin_cos_tan_within_two_ulp(name, lo, hi):
    mp.mp.dps = 40
    rng = random.Random(hash((name, lo)) & 0xFFFF)
    fn = {"sin": mp.sin, "cos": mp.cos, "tan": mp.tan}[name]
    worst = 0.0
    for _ in range(3000):
        x = rng.uniform(lo, hi)
        worst = max(worst, ulps(O.gm(name, x), fn(mp.mpf(x))))
    assert worst <= (3.0 if name == "tan" else 2.0), worst   # tan is the quotient of two 2-ulp values
    # quadrant boundaries and exact points
    for x in (0.0, -0.0, math.pi / 2, math.pi, -math.pi, 3 * math.pi / 2, 2 * math.pi, 1e-300):
        assert ulps(O.gm(name, x), fn(mp.mpf(x))) <= 2.0 or abs(O.gm(name, x) - float(fn(mp.mpf(x)))) < 1e-15
    assert math.isnan(O.gm(name, float("nan"))) and math.isnan(O.gm(name, 2.0 ** 21))


def test_atan_and_atan2_within_two_ulp():
    mp.mp.dps = 40
    rng = random.Random(7)
    worst = 0.0
    for _ in range(3000):
        x = math.copysign(10 ** rng.uniform(-12, 12), rng.uniform(-1, 1))
        worst = max(worst, ulps(O.gm("atan", x), mp.atan(mp.mpf(x))))
    for k in range(-40, 41):   # the reduction breakpoints k/8 and their neighbours
        for x in (k / 16, math.nextafter(k / 16, 1), math.nextafter(k / 16, -1)):
            worst = max(worst, ulps(O.gm("atan", x), mp.atan(mp.mpf(x))))
    assert worst <= 2.0, worst
    worst = 0.0
    for _ in range(3000):
        y, x = rng.uniform(-1, 1) * 10 ** rng.uniform(-5, 5), rng.uniform(-1, 1) * 10 ** rng.uniform(-5, 5)
        worst = max(worst, ulps(O.gm("atan2", y, x), mp.atan2(mp.mpf(y), mp.mpf(x))))
    assert worst <= 2.0, worst
    # signed zeros, axes and infinities as IEEE atan2
    assert O.gm("atan2", 0.0, 0.0) == 0.0 and math.copysign(1, O.gm("atan2", -0.0, 0.0)) < 0
    assert O.gm("atan2", 0.0, -0.0) == math.pi and O.gm("atan2", -0.0, -0.0) == -math.pi
    assert O.gm("atan2", 1.0, 0.0) == math.pi / 2 and O.gm("atan2", -1.0, 0.0) == -math.pi / 2
    assert O.gm("atan2", 0.0, -1.0) == math.pi and O.gm("atan", math.inf) == math.pi / 2
    assert O.gm("atan2", 1.0, math.inf) == 0.0 and O.gm("atan2", 1.0, -math.inf) == math.pi


# ------------------------------------------------------------------------------------------- geodesy
def test_metres_per_degree_from_the_radii_of_curvature():
    mp.mp.dps = 40
    e2 = mp.mpf(F) * (2 - mp.mpf(F))
    for lat in np.linspace(-89.0, 89.0, 61):
        phi = mp.radians(mp.mpf(lat))
        w2 = 1 - e2 * mp.sin(phi) ** 2
        kx = A / mp.sqrt(w2) * mp.cos(phi) * mp.pi / 180   # prime vertical N cos(phi)
        ky = A * (1 - e2) / w2 ** mp.mpf(1.5) * mp.pi / 180   # meridional M
        gx, gy = O.m_per_deg(float(lat))
        assert abs(gx - kx) <= 1e-14 * kx and abs(gy - ky) <= 1e-14 * ky
    kx, ky = O.m_per_deg(0.0)
    assert abs(kx - 111319.49079327357) < 1e-8 and abs(ky - 110574.27582159435) < 1e-8


def test_degree_cover_covers_and_is_covered():
    """The widened cover holds d metres at every latitude up to d metres beyond the range (d over the smallest metres
    per degree of latitude, a (1 - e^2) pi / 180 at the equator); the shrunk cover is held by d at every latitude of
    the range; and a geodesic of length d from any point of the range ends inside the widened box."""
    rng = random.Random(3)
    ky_min = A * (1 - F * (2 - F)) * math.pi / 180
    for _ in range(200):
        y0 = rng.uniform(-80, 80)
        y1 = y0 + rng.uniform(0, 0.5)
        d = 10 ** rng.uniform(-1, 4)
        dx, dy = O.deg_cover(y0, y1, d, True)
        pad = d / ky_min
        for lat in np.linspace(y0 - pad, y1 + pad, 41):
            kx, ky = O.m_per_deg(float(lat))
            assert dx * kx >= d and dy * ky >= d
        sx, sy = O.deg_cover(y0, y1, d, False)
        for lat in np.linspace(y0, y1, 41):
            kx, ky = O.m_per_deg(float(lat))
            assert sx * kx <= d and sy * ky <= d
    assert O.deg_cover(10.0, 11.0, 0.0, True) == (0.0, 0.0)
    assert O.deg_cover(89.995, 89.999, 5.0, True)[0] == 360.0   # a range at a pole: every longitude
    # end points of geodesics of length just under d leave the start's latitude range by less than the cover
    for _ in range(300):
        lat, lon = rng.uniform(-75, 75), rng.uniform(-170, 170)
        d = 10 ** rng.uniform(0, 4)
        dx, dy = O.deg_cover(lat, lat, d, True)
        ang = rng.uniform(0, 2 * math.pi)
        # walk outwards along a bearing until Vincenty says d: the end lies within (dx, dy) of the start
        lo_, hi_ = 0.0, 3 * d / ky_min
        for _ in range(60):
            t = 0.5 * (lo_ + hi_)
            q = (lon + t * math.cos(ang) / max(math.cos(math.radians(lat)), 1e-3), lat + t * math.sin(ang))
            if O.vincenty(lon, lat, q[0], q[1]) < d:
                lo_ = t
            else:
                hi_ = t
        q = (lon + lo_ * math.cos(ang) / max(math.cos(math.radians(lat)), 1e-3), lat + lo_ * math.sin(ang))
        assert abs(q[0] - lon) <= dx and abs(q[1] - lat) <= dy


def test_frame_band_holds_the_geography_distance():
    """Near the threshold the frame distance of two straight lon/lat edges (metres per degree at the window's middle
    latitude) is within the band of the geography distance of the same edges (great-circle arcs, Vincenty between the
    closest points): random edge pairs of 1-300 m at latitudes up to 70 degrees, windows as the distance join builds
    them. The band is also not wasteful: within 1e-3 of d plus the edges' curvature term."""
    rng = random.Random(17)
    worst = 0.0
    for _ in range(3000):
        lat, lon = rng.uniform(-70, 70), rng.uniform(-170, 170)
        kx, ky = O.m_per_deg(lat)
        d = 10 ** rng.uniform(0, 2.7)

        def edge(cx, cy):
            L, ang = 10 ** rng.uniform(0, 2.5), rng.uniform(0, math.pi)
            return ((cx - 0.5 * L * math.cos(ang) / kx, cy - 0.5 * L * math.sin(ang) / ky),
                    (cx + 0.5 * L * math.cos(ang) / kx, cy + 0.5 * L * math.sin(ang) / ky))

        a0, a1 = edge(lon, lat)
        g = d * rng.uniform(0.8, 1.2) + 0.6 * 10 ** rng.uniform(0, 2.5)
        ang = rng.uniform(0, 2 * math.pi)
        b0, b1 = edge(lon + g * math.cos(ang) / kx, lat + g * math.sin(ang) / ky)
        if O.seg_isect(*a0, *a1, *b0, *b1):
            continue
        ys = [a0[1], a1[1], b0[1], b1[1]]
        lo_, hi_ = min(ys), max(ys)
        fkx, fky, rel, curv = O.geo_frame(lo_, hi_)
        ox, oy = a0   # the join measures in the frame of the first geometry's origin
        fr = math.sqrt(O.seg_seg_d2(((a0[0] - ox) * fkx, (a0[1] - oy) * fky), ((a1[0] - ox) * fkx, (a1[1] - oy) * fky),
                                    ((b0[0] - ox) * fkx, (b0[1] - oy) * fky), ((b1[0] - ox) * fkx, (b1[1] - oy) * fky)))
        La2 = ((a1[0] - a0[0]) * fkx) ** 2 + ((a1[1] - a0[1]) * fky) ** 2
        Lb2 = ((b1[0] - b0[0]) * fkx) ** 2 + ((b1[1] - b0[1]) * fky) ** 2
        band = O.geog_band(lo_, hi_, fr, La2, Lb2)
        geo = O.geog_dist(a0, a1, b0, b1)
        worst = max(worst, abs(fr - geo) / band)
        assert abs(fr - geo) <= band, (lat, d, fr, geo, band)
        assert band <= 1e-3 * fr + 2 * (La2 + Lb2) * curv + 1e-5
    assert worst > 0.01   # the pins are not vacuous: the two distances do differ at the band's scale


def _vincenty_mp(lon1, lat1, lon2, lat2):
    """Vincenty's inverse method in mpmath at 50 digits (converged to 1e-40)."""
    mp.mp.dps = 50
    a, f = mp.mpf(A), mp.mpf(1) / mp.mpf("298.257223563")
    b = a * (1 - f)
    L = mp.radians(mp.mpf(lon2) - mp.mpf(lon1))
    U1, U2 = mp.atan((1 - f) * mp.tan(mp.radians(mp.mpf(lat1)))), mp.atan((1 - f) * mp.tan(mp.radians(mp.mpf(lat2))))
    lam = L
    for _ in range(200):
        sl, cl = mp.sin(lam), mp.cos(lam)
        sinS = mp.sqrt((mp.cos(U2) * sl) ** 2 + (mp.cos(U1) * mp.sin(U2) - mp.sin(U1) * mp.cos(U2) * cl) ** 2)
        if sinS == 0:
            return mp.mpf(0)
        cosS = mp.sin(U1) * mp.sin(U2) + mp.cos(U1) * mp.cos(U2) * cl
        sig = mp.atan2(sinS, cosS)
        sinA = mp.cos(U1) * mp.cos(U2) * sl / sinS
        cos2A = 1 - sinA ** 2
        cos2Sm = cosS - 2 * mp.sin(U1) * mp.sin(U2) / cos2A if cos2A != 0 else mp.mpf(0)
        C = f / 16 * cos2A * (4 + f * (4 - 3 * cos2A))
        prev = lam
        lam = L + (1 - C) * f * sinA * (sig + C * sinS * (cos2Sm + C * cosS * (-1 + 2 * cos2Sm ** 2)))
        if abs(lam - prev) < mp.mpf(10) ** -40:
            break
    u2 = cos2A * (a * a - b * b) / (b * b)
    Ak = 1 + u2 / 16384 * (4096 + u2 * (-768 + u2 * (320 - 175 * u2)))
    Bk = u2 / 1024 * (256 + u2 * (-128 + u2 * (74 - 47 * u2)))
    dS = Bk * sinS * (cos2Sm + Bk / 4 * (cosS * (-1 + 2 * cos2Sm ** 2) - Bk / 6 * cos2Sm * (-3 + 4 * sinS ** 2) * (-3 + 4 * cos2Sm ** 2)))
    return b * Ak * (sig - dS)


def test_vincenty_against_high_precision_and_closed_forms():
    rng = random.Random(11)
    for _ in range(60):   # parcel-scale and regional distances
        lon1, lat1 = rng.uniform(-179, 179), rng.uniform(-75, 75)
        s = 10 ** rng.uniform(-4, 1.5)
        lon2, lat2 = lon1 + rng.uniform(-s, s), max(-80.0, min(80.0, lat1 + rng.uniform(-s, s)))
        exact = float(_vincenty_mp(lon1, lat1, lon2, lat2))
        # double rounding plus the 1e-14 rad convergence test on lambda: ~1e-7 m
        assert abs(O.vincenty(lon1, lat1, lon2, lat2) - exact) <= 2e-7 + 1e-12 * exact
    # equator: a * dlambda; meridian: the meridian arc integral of M(phi)
    mp.mp.dps = 30
    for dl in (1e-6, 1e-3, 0.5, 20.0):
        exact = A * math.radians((10.0 + dl) - 10.0)
        assert abs(O.vincenty(10.0, 0.0, 10.0 + dl, 0.0) - exact) <= 2e-7 + 1e-13 * exact
    e2 = mp.mpf(F) * (2 - mp.mpf(F))
    for p1, p2 in ((0.0, 1e-3), (34.0, 34.01), (-60.0, -59.0), (10.0, 40.0)):
        arc = mp.quad(lambda t: A * (1 - e2) / (1 - e2 * mp.sin(t) ** 2) ** mp.mpf(1.5), [mp.radians(p1), mp.radians(p2)])
        assert abs(O.vincenty(5.0, p1, 5.0, p2) - float(arc)) <= 2e-7 + 1e-12 * float(arc)
    # Vincenty (1975) test line, Flinders Peak -> Buninyong: 54972.271 m (published on GRS80; WGS84 differs < 1 mm)
    lat1 = -(37 + 57 / 60 + 3.72030 / 3600)
    lon1 = 144 + 25 / 60 + 29.52440 / 3600
    lat2 = -(37 + 39 / 60 + 10.15610 / 3600)
    lon2 = 143 + 55 / 60 + 35.38390 / 3600
    assert abs(O.vincenty(lon1, lat1, lon2, lat2) - 54972.271) < 1e-3
    assert O.vincenty(3.0, 4.0, 3.0, 4.0) == 0.0


def test_utm_scale_factor():
    # central meridian of the zone: k0 = 0.9996 exactly; zone chosen from the geometry's centre (6-degree zones)
    assert O.utm_k(-118.2, 34.0, -117.0, 34.0) == 0.9996          # zone 11: central meridian -117
    assert O.utm_k(-95.5, 38.5, -93.0, 40.0) == 0.9996            # zone 15: central meridian -93
    assert O.utm_k(179.9, 10.0, 177.0, 10.0) == 0.9996            # zone 60 (index 59 clamp)
    assert O.utm_k(10.0, 75.0, 13.0, 75.0) == 1.0                 # polar: equal-area projection, k = 1
    # off the central meridian: Snyder's series (Map Projections, eq. 8-11) evaluated in mpmath
    mp.mp.dps = 40
    e2 = mp.mpf(F) * (2 - mp.mpf(F))
    ep2 = e2 / (1 - e2)
    for lon, lat in ((-114.5, 34.0), (-119.9, 20.0), (-112.0, 60.0)):
        phi = mp.radians(lat)
        T, Cc = mp.tan(phi) ** 2, ep2 * mp.cos(phi) ** 2
        Aa = mp.radians(lon - (-117.0)) * mp.cos(phi)
        k = mp.mpf("0.9996") * (1 + (1 + Cc) * Aa ** 2 / 2 + (5 - 4 * T + 42 * Cc + 13 * Cc ** 2 - 28 * ep2) * Aa ** 4 / 24 +
                                (61 - 148 * T + 16 * T ** 2) * Aa ** 6 / 720)
        assert abs(O.utm_k(-117.5, lat, lon, lat) - float(k)) < 1e-15
    assert O.utm_k(-114.5, 34.0, -114.5, 34.0) > 1.0 > O.utm_k(-117.1, 34.0, -117.1, 34.0)


def test_great_circle_closest_point_and_geography_distance():
    # a degenerate edge reduces to Vincenty between the points
    for p, q in (((10.0, 20.0), (10.001, 20.0005)), ((-95.0, 38.9), (-95.0004, 38.9003))):
        assert O.geog_dist(p, p, q) == O.vincenty(p[0], p[1], q[0], q[1])
    # point above an equatorial edge: closest point is the foot on the equator; distance = meridian arc
    mp.mp.dps = 30
    e2 = mp.mpf(F) * (2 - mp.mpf(F))
    arc = mp.quad(lambda t: A * (1 - e2) / (1 - e2 * mp.sin(t) ** 2) ** mp.mpf(1.5), [0, mp.radians(0.001)])
    assert abs(O.geog_dist((0.0, 0.0), (0.01, 0.0), (0.005, 0.001)) - float(arc)) < 1e-7
    # the foot on a general great-circle arc minimises the angle to p over the arc (golden-section search in mpmath)
    rng = random.Random(5)
    mp.mp.dps = 30
    for _ in range(20):
        s = (rng.uniform(-100, 100), rng.uniform(-60, 60))
        e = (s[0] + rng.uniform(-0.01, 0.01), s[1] + rng.uniform(-0.01, 0.01))
        p = ((s[0] + e[0]) / 2 + rng.uniform(-0.003, 0.003), (s[1] + e[1]) / 2 + rng.uniform(-0.003, 0.003))
        got = O.geog_dist(s, e, p)

        def v3(lon, lat):
            lo, la = mp.radians(lon), mp.radians(lat)
            return mp.matrix([mp.cos(la) * mp.cos(lo), mp.cos(la) * mp.sin(lo), mp.sin(la)])

        vs, ve, vp = v3(*s), v3(*e), v3(*p)
        om = mp.acos(sum(vs[i] * ve[i] for i in range(3)))

        def at(t):
            return (mp.sin((1 - t) * om) * vs + mp.sin(t * om) * ve) / mp.sin(om)

        def ang(t):
            w = at(t)
            c = mp.matrix([w[1] * vp[2] - w[2] * vp[1], w[2] * vp[0] - w[0] * vp[2], w[0] * vp[1] - w[1] * vp[0]])
            return mp.atan2(mp.sqrt(sum(c[i] ** 2 for i in range(3))), sum(w[i] * vp[i] for i in range(3)))

        lo, hi = mp.mpf(0), mp.mpf(1)
        g = (mp.sqrt(5) - 1) / 2
        for _ in range(120):
            m1, m2 = hi - g * (hi - lo), lo + g * (hi - lo)
            if ang(m1) < ang(m2):
                hi = m2
            else:
                lo = m1
        w = at((lo + hi) / 2)
        lon, lat = float(mp.degrees(mp.atan2(w[1], w[0]))), float(mp.degrees(mp.asin(w[2])))
        assert abs(got - O.vincenty(lon, lat, p[0], p[1])) < 1e-6


def test_dwithin_closed_forms_on_the_equator():
    """Two unit squares 0.001 deg wide, separated along the equator by a gap g: the distance is a * g (in radians)
    between the facing meridian edges at the equator; a square strictly inside another is at distance 0."""
    import gsj

    def squares(gap):
        w = 0.001
        rings = [[(0, 0), (w, 0), (w, w), (0, w), (0, 0)], [(w + gap, 0), (2 * w + gap, 0), (2 * w + gap, w), (w + gap, w), (w + gap, 0)]]
        xy = np.array([p for r in rings for p in r], np.float64)
        return gsj.HostSet(np.array([0, 1, 2], np.uint64), np.array([0, 5, 10], np.uint64), xy)

    for gap in (1e-5, 3e-5, 4e-5):
        L = O.Layer.polys(squares(gap))
        exact = A * math.radians(gap)
        for d in (exact * 0.999, exact * 1.001):
            hit, dist = O.dw_pair(L, L, 0, 1, d)
            assert hit == (d >= exact)
            if hit:
                assert abs(dist - exact) < 1e-6
        r = O.dwithin(L, L, exact * 1.01, same=True, upper=True, report=True)
        assert r["pairs"].tolist() == [[0, 0], [0, 1], [1, 1]] and abs(r["dist"][1] - exact) < 1e-6 and r["dist"][0] == 0.0
    # containment: inner square strictly inside, 10 m from no boundary of interest -> distance 0
    xy = np.array([(0, 0), (0.01, 0), (0.01, 0.01), (0, 0.01), (0, 0), (0.004, 0.004), (0.006, 0.004), (0.006, 0.006), (0.004, 0.006), (0.004, 0.004)], np.float64)
    L = O.Layer.polys(gsj.HostSet(np.array([0, 1, 2], np.uint64), np.array([0, 5, 10], np.uint64), xy))
    assert O.dw_pair(L, L, 0, 1, 1.0) == (True, 0.0)
    # points: inside -> 0; above the equatorial edge at height h -> the meridian arc
    pts = O.Layer.points_layer(np.array([[0.005, 0.005], [0.005, -0.0001], [0.02, 0.02]], np.float64))
    r = O.points_dwithin(pts, L, 15.0, report=True)
    got = {tuple(p): dd for p, dd in zip(r["pairs"].tolist(), r["dist"].tolist())}
    assert got[(0, 0)] == 0.0 and got[(0, 1)] == 0.0 and (2, 0) not in got
    mp.mp.dps = 30
    e2 = mp.mpf(F) * (2 - mp.mpf(F))
    arc = float(mp.quad(lambda t: A * (1 - e2) / (1 - e2 * mp.sin(t) ** 2) ** mp.mpf(1.5), [0, mp.radians(0.0001)]))
    assert abs(got[(1, 0)] - arc) < 1e-6
