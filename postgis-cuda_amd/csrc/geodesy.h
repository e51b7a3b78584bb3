// WGS84 geodesy of the distance joins (the role of engine/buffer.cuh:29-181). Host + device (one source over the
// portable elementary functions of gmath.h), so the GPU and the CPU restatement decide every threshold identically.
//
// Each helper is derived here from the ellipsoid's geometry; the constants below are the ones that derivation needs.
//
//  * m_per_deg: metres per degree along a parallel and along a meridian, from the radii of curvature N (prime
//    vertical) and M (meridian): kx = N cos(phi), ky = M, both times pi/180.
//  * deg_cover: longitude / latitude spans (degrees) that cover, or are covered by, d metres at every latitude of a
//    range. A geodesic of length s moves |dlon| <= s / min(kx) and |dlat| <= s / min(ky) along its path, whose
//    latitudes stay within s / min(ky) of the range; kx falls and ky grows with |latitude|.
//  * utm_k: point scale factor of the transverse Mercator zone PostGIS's _ST_BestSRID picks for a geometry (Snyder,
//    Map Projections - A Working Manual, 1987, eq. 8-11; zone of the geometry's centre, 6 degrees wide from -180);
//    polar geometries (|lat| > 70) use a Lambert azimuthal equal-area projection there, whose scale is taken as 1.
//  * vincenty_m: Vincenty's inverse method on the ellipsoid (Survey Review 23(176), 1975).
//  * pt_seg_d2 / seg_seg_d2: planar distances in a metric frame (the distance of two non-crossing segments is the
//    smallest distance of an endpoint of one to the other segment).
//  * gc_closest / geog_dist: closest point of a great-circle arc to a point, by projecting the point's unit vector on
//    the arc's plane; the arc normal comes from the endpoint chord, whose coordinates are formed from sum-to-product
//    identities so that nearby endpoints keep full relative accuracy; the ellipsoidal distance between the closest
//    points then decides (geography edges are great-circle arcs).
//  * Frame / frame_of / geog_band: the local metric frame of a pair and the half-width of the band around d in which
//    the frame distance may disagree with the geography distance (derivation at geog_band).
#pragma once
#include "gmath.h"

namespace gsj {
namespace geo {

constexpr double WGS84_A = 6378137.0;
constexpr double WGS84_F = 1.0 / 298.257223563;
constexpr double WGS84_E2 = WGS84_F * (2.0 - WGS84_F);
constexpr double DEG = gm::PI_HI / 180.0;
// the smallest radius of curvature of the ellipsoid: the meridian radius at the equator, M(0) = a (1 - e^2)
constexpr double M_EQ = WGS84_A * (1.0 - WGS84_E2);
// metres per degree of latitude at the equator, the smallest ky
constexpr double KY_MIN = M_EQ * DEG;
// bounds on how fast the frame scales change with latitude (per radian, relative): d ln kx / dphi = -tan(phi) +
// e^2 sin cos / W^2, so |.| <= |tan(phi)| + e^2 / 2; d ln ky / dphi = 3 e^2 sin cos / W^2 <= 1.5 e^2 / (1 - e^2)
constexpr double SCALE_SLOPE = 1.5 * WGS84_E2 / (1.0 - WGS84_E2);

// metres per degree of longitude (kx) and of latitude (ky) at latitude lat (degrees)
GSJ_PORT void m_per_deg(double lat, double& kx, double& ky) {
  double s, c;
  gm::sincos(lat * DEG, s, c);
  const double w2 = 1.0 - WGS84_E2 * s * s;
  const double n = WGS84_A / sqrt(w2);   // prime-vertical radius N
  kx = n * c * DEG;
  ky = n * (1.0 - WGS84_E2) / w2 * DEG;   // meridian radius M = N (1 - e^2) / W^2
}

// Degrees covering d metres at every latitude in [y0, y1] (larger: for widening boxes, every point within d of the
// range lies within (dx, dy) degrees), or covered by d metres there (smaller: for shrinking boxes). The pad of the
// larger cover is the most latitude d metres can add (d / KY_MIN); 1e-12 relative covers the rounding of the scales.
// Ranges reaching within 0.01 degrees of a pole get the whole circle of longitude.
GSJ_PORT void deg_cover(double y0, double y1, double d, bool larger, double& dx, double& dy) {
  if (d <= 0) { dx = dy = 0; return; }
  const double pad = larger ? d / KY_MIN * (1.0 + 1e-12) : 0.0;
  const double a = y0 - pad, b = y1 + pad;
  const double far = fmax(fabs(a), fabs(b));                     // |lat| where kx is smallest and ky largest
  const double near = (a <= 0.0 && b >= 0.0) ? 0.0 : fmin(fabs(a), fabs(b));   // where kx is largest, ky smallest
  double kx_far, ky_far, kx_near, ky_near;
  m_per_deg(fmin(far, 90.0), kx_far, ky_far);
  m_per_deg(near, kx_near, ky_near);
  if (larger) {
    dx = far >= 89.99 ? 360.0 : d / kx_far * (1.0 + 1e-12);
    dy = d / ky_near * (1.0 + 1e-12);
  } else {
    dx = d / kx_near * (1.0 - 1e-12);
    dy = d / ky_far * (1.0 - 1e-12);
  }
}

// scale factor at (lon, lat) of the transverse Mercator zone chosen for a geometry centred at (glon, glat)
GSJ_PORT double utm_k(double glon, double glat, double lon, double lat) {
  if (fabs(glat) > 70.0) return 1.0;   // polar equal-area projection: scale 1 taken there
  const int zone = (int)fmin(59.0, floor((glon + 180.0) / 6.0));
  const double lon0 = -177.0 + 6.0 * zone;   // the zone's central meridian
  double s, c;
  gm::sincos(lat * DEG, s, c);
  const double ep2 = WGS84_E2 / (1.0 - WGS84_E2);   // second eccentricity squared
  const double T = (s / c) * (s / c), C = ep2 * c * c;
  const double A2 = ((lon - lon0) * DEG * c) * ((lon - lon0) * DEG * c);
  const double series = 1.0 + A2 * ((1.0 + C) / 2.0 + A2 * ((5.0 - 4.0 * T + 42.0 * C + 13.0 * C * C - 28.0 * ep2) / 24.0 +
                                                             A2 * (61.0 - 148.0 * T + 16.0 * T * T) / 720.0));
  return 0.9996 * series;
}

// ellipsoidal distance in metres (Vincenty's inverse method: iterate the longitude on the auxiliary sphere until it
// changes by less than 1e-14 rad, at most 64 times; then the series in u^2 for the distance)
GSJ_PORT double vincenty_m(double lon1, double lat1, double lon2, double lat2) {
  const double a = WGS84_A, f = WGS84_F, b = a * (1.0 - f);
  const double L = (lon2 - lon1) * DEG;
  double sU1, cU1, sU2, cU2;   // reduced latitudes
  gm::sincos(gm::atan((1.0 - f) * gm::tan(lat1 * DEG)), sU1, cU1);
  gm::sincos(gm::atan((1.0 - f) * gm::tan(lat2 * DEG)), sU2, cU2);
  double lambda = L, sin_sig = 0, cos_sig = 1, sig = 0, cos2_alpha = 1, cos_2sm = 0;
  for (int it = 0; it < 64; it++) {
    double sl, cl;
    gm::sincos(lambda, sl, cl);
    const double p = cU2 * sl, q = cU1 * sU2 - sU1 * cU2 * cl;
    sin_sig = sqrt(p * p + q * q);
    if (sin_sig == 0) return 0.0;   // coincident points
    cos_sig = sU1 * sU2 + cU1 * cU2 * cl;
    sig = gm::atan2(sin_sig, cos_sig);
    const double sin_alpha = cU1 * cU2 * sl / sin_sig;
    cos2_alpha = 1.0 - sin_alpha * sin_alpha;
    cos_2sm = cos2_alpha != 0 ? cos_sig - 2.0 * sU1 * sU2 / cos2_alpha : 0.0;   // equatorial line: 0
    const double C = f / 16.0 * cos2_alpha * (4.0 + f * (4.0 - 3.0 * cos2_alpha));
    const double next = L + (1.0 - C) * f * sin_alpha * (sig + C * sin_sig * (cos_2sm + C * cos_sig * (2.0 * cos_2sm * cos_2sm - 1.0)));
    const bool done = fabs(next - lambda) < 1e-14;
    lambda = next;
    if (done) break;
  }
  const double u2 = cos2_alpha * (a * a - b * b) / (b * b);
  const double A = 1.0 + u2 / 16384.0 * (4096.0 + u2 * (-768.0 + u2 * (320.0 - 175.0 * u2)));
  const double B = u2 / 1024.0 * (256.0 + u2 * (-128.0 + u2 * (74.0 - 47.0 * u2)));
  const double c2 = cos_2sm * cos_2sm;
  const double dsig = B * sin_sig * (cos_2sm + B / 4.0 * (cos_sig * (2.0 * c2 - 1.0) - B / 6.0 * cos_2sm * (4.0 * sin_sig * sin_sig - 3.0) * (4.0 * c2 - 3.0)));
  return b * A * (sig - dsig);
}

// squared distance from p to the segment a-b (metric frame coordinates)
GSJ_PORT double pt_seg_d2(double px, double py, double ax, double ay, double bx, double by) {
  const double ux = bx - ax, uy = by - ay, wx = px - ax, wy = py - ay;
  const double uu = ux * ux + uy * uy;
  const double t = uu > 0 ? fmin(1.0, fmax(0.0, (wx * ux + wy * uy) / uu)) : 0.0;
  const double rx = wx - t * ux, ry = wy - t * uy;
  return rx * rx + ry * ry;
}

// squared distance of two segments that do not cross
GSJ_PORT double seg_seg_d2(double a0x, double a0y, double a1x, double a1y, double b0x, double b0y, double b1x, double b1y) {
  return fmin(fmin(pt_seg_d2(a0x, a0y, b0x, b0y, b1x, b1y), pt_seg_d2(a1x, a1y, b0x, b0y, b1x, b1y)),
              fmin(pt_seg_d2(b0x, b0y, a0x, a0y, a1x, a1y), pt_seg_d2(b1x, b1y, a0x, a0y, a1x, a1y)));
}

// ---- great-circle arcs on the unit sphere ---------------------------------------------------------------------------
struct V3 { double x, y, z; };
GSJ_PORT V3 unit_of(double lon, double lat) {   // radians
  double sl, cl, sp, cp;
  gm::sincos(lon, sl, cl);
  gm::sincos(lat, sp, cp);
  return V3{cp * cl, cp * sl, sp};
}
GSJ_PORT double dot3(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
GSJ_PORT V3 cross3(V3 a, V3 b) { return V3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
GSJ_PORT V3 unit3(V3 a) {
  const double l = sqrt(dot3(a, a));
  return l > 0 ? V3{a.x / l, a.y / l, a.z / l} : V3{0, 0, 0};
}
GSJ_PORT double angle3(V3 a, V3 b) { const V3 c = cross3(a, b); return gm::atan2(sqrt(dot3(c, c)), dot3(a, b)); }

// e - s for the unit vectors of (slon, slat) and (elon, elat) (radians), each coordinate difference written as
// products of sines of half differences (cos A - cos B = -2 sin((A+B)/2) sin((A-B)/2), sin A - sin B =
// 2 cos((A+B)/2) sin((A-B)/2)), so that the chord of nearby points carries full relative precision
GSJ_PORT V3 chord(double slon, double slat, double elon, double elat) {
  double s_hl, c_hl, s_ml, c_ml, s_hp, c_hp, s_mp, c_mp, sls, cls, sps, cps, spe, cpe;
  gm::sincos(0.5 * (elon - slon), s_hl, c_hl);   // half the longitude difference
  gm::sincos(0.5 * (elon + slon), s_ml, c_ml);   // mean longitude
  gm::sincos(0.5 * (elat - slat), s_hp, c_hp);
  gm::sincos(0.5 * (elat + slat), s_mp, c_mp);
  gm::sincos(slon, sls, cls);
  gm::sincos(slat, sps, cps);
  gm::sincos(elat, spe, cpe);
  const double dcos_lon = -2.0 * s_ml * s_hl, dsin_lon = 2.0 * c_ml * s_hl;   // cos(elon) - cos(slon), sin(elon) - sin(slon)
  const double dcos_lat = -2.0 * s_mp * s_hp, dsin_lat = 2.0 * c_mp * s_hp;   // cos(elat) - cos(slat), sin(elat) - sin(slat)
  return V3{cpe * dcos_lon + cls * dcos_lat, cpe * dsin_lon + sls * dcos_lat, dsin_lat};
}

// closest point (clon, clat, radians) of the great-circle arc s-e to p, and its angle from p. The foot of p on the
// arc's great circle is p minus its component along the arc normal n = s x (e - s); it lies on the arc when it is on
// the inner side of both endpoints ((s x k).n >= 0 and (k x e).n >= 0); otherwise the nearer endpoint is closest.
GSJ_PORT double gc_closest(double slon, double slat, double elon, double elat, double plon, double plat, double& clon, double& clat) {
  const V3 s = unit_of(slon, slat), e = unit_of(elon, elat), p = unit_of(plon, plat);
  const double as = angle3(p, s), ae = angle3(p, e);
  double best = as;
  clon = slon; clat = slat;
  if (ae < best) { best = ae; clon = elon; clat = elat; }
  if (slon == elon && slat == elat) return best;
  const V3 n = unit3(cross3(s, chord(slon, slat, elon, elat)));
  const double h = dot3(p, n);
  const V3 k = unit3(V3{p.x - h * n.x, p.y - h * n.y, p.z - h * n.z});
  if (dot3(k, k) == 0 || dot3(cross3(s, k), n) < 0 || dot3(cross3(k, e), n) < 0) return best;
  const double ak = angle3(p, k);
  if (ak < best) {
    best = ak;
    clat = gm::atan2(k.z, sqrt(k.x * k.x + k.y * k.y));
    clon = gm::atan2(k.y, k.x);
  }
  return best;
}

// geography distance (metres) between the arcs a0-a1 and b0-b1 that do not cross, or of the arc a from the point b0
// (b_is_point); degrees in. The closest pair of points has an endpoint of one arc: of the four (endpoint, arc)
// candidates the one at the smallest angle is measured on the ellipsoid. Not inlined: called only for edge pairs
// inside the band, it would otherwise dominate the callers' register use.
__attribute__((noinline)) GSJ_PORT double geog_dist(double a0x, double a0y, double a1x, double a1y, double b0x, double b0y, double b1x, double b1y,
                                                     bool b_is_point) {
  const double r = DEG;
  const double A[4] = {a0x * r, a0y * r, a1x * r, a1y * r}, B[4] = {b0x * r, b0y * r, b1x * r, b1y * r};
  double best = INFINITY, p1x = B[0], p1y = B[1], p2x = B[0], p2y = B[1];
#pragma unroll
  for (int k = 0; k < 4; k++) {
    if (b_is_point && k > 0) break;
    const double* arc = k < 2 ? A : B;   // endpoint k % 2 of the other side against this arc
    const double* other = k < 2 ? B : A;
    const double qx = other[2 * (k & 1)], qy = other[2 * (k & 1) + 1];
    double cx, cy;
    const double ang = gc_closest(arc[0], arc[1], arc[2], arc[3], qx, qy, cx, cy);
    if (ang < best) { best = ang; p1x = cx; p1y = cy; p2x = qx; p2y = qy; }
  }
  return vincenty_m(p1x / r, p1y / r, p2x / r, p2y / r);
}

// ---- frame distance versus geography distance ------------------------------------------------------------------------
// A pair is measured in a local frame: metres per degree (kx, ky) at the middle latitude of its window. Two effects
// separate the frame distance of two straight lon/lat edges from the geography distance of the same edges:
//  1. the frame's scales are those of one latitude: across a window of half-width h radians they are off by at most
//     h (|tan(phi)| + SCALE_SLOPE) relatively at the window's most poleward latitude (to first order; the factor 1.25
//     covers the second order for windows of a few km), so a distance near d is off by at most d * rel;
//  2. geography edges are great-circle arcs: a straight lon/lat edge of length L has geodesic curvature of at most
//     (|tan(phi)| + SCALE_SLOPE) / M_EQ, so it departs from its arc by at most L^2 curv with curv = that / 8;
//  3. the geography distance takes its closest points on the sphere, whose local metric differs from the ellipsoid's
//     by a stretch along the parallel of N / M = 1 + e^2 cos^2(phi) / (1 - e^2): the spherical foot of the
//     perpendicular is not the ellipsoidal one, and the distance between them is longer by at most (N/M - 1)^2 / 2
//     relatively (ANISO; 2.25e-5 is the largest excess measured near the equator, where the stretch is largest).
// The band doubles the sum, and adds 1e-7 of d and 1e-6 m for Vincenty's convergence and the rounding of the closest
// points (both well below 1e-6 m at these distances).
constexpr double ANISO = 0.5 * (WGS84_E2 / (1.0 - WGS84_E2)) * (WGS84_E2 / (1.0 - WGS84_E2));
struct Frame {
  double kx, ky;   // metres per degree at the window's middle latitude
  double rel;      // relative scale error over the window
  double curv;     // departure of a straight edge from its arc, per squared metre of edge length
};
GSJ_PORT Frame frame_of(double lat_lo, double lat_hi) {   // the window's latitude range, degrees
  Frame F;
  m_per_deg(0.5 * (lat_lo + lat_hi), F.kx, F.ky);
  const double far = fmin(fmax(fabs(lat_lo), fabs(lat_hi)), 89.0);
  const double slope = fabs(gm::tan(far * DEG)) + SCALE_SLOPE;
  F.rel = 1.25 * 0.5 * (lat_hi - lat_lo) * DEG * slope;
  F.curv = slope / (8.0 * M_EQ);
  return F;
}
// half-width (metres) of the band around d: La2 / Lb2 are the squared frame lengths of the two edges (0 for a point)
GSJ_PORT double geog_band(double d, const Frame& F, double La2, double Lb2) {
  return 2.0 * (d * (F.rel + ANISO + 1e-7) + (La2 + Lb2) * F.curv) + 1e-6;
}

// float box rounded outward (largest float <= lo, smallest float >= hi)
GSJ_PORT float f_down(double x) {
  float f = (float)x;
  if ((double)f > x) f = nextafterf(f, -INFINITY);
  return f;
}
GSJ_PORT float f_up(double x) {
  float f = (float)x;
  if ((double)f < x) f = nextafterf(f, INFINITY);
  return f;
}

}  // namespace geo
}  // namespace gsj
