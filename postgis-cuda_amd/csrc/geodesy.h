// WGS84 geodesy for the distance joins (reference: engine/buffer.cuh:29-180). Host + device (one source, the
// portable elementary functions of gmath.h), so the GPU and the CPU restatement decide every threshold identically.
//
//  * m_per_deg: metres per degree of longitude / latitude from the prime-vertical and meridional radii of curvature.
//  * deg_cover: degrees that cover d metres everywhere in a latitude range (outer cover for expanding boxes, inner
//    for shrinking).
//  * utm_k: point scale factor of the UTM zone PostGIS's _ST_BestSRID picks for a geometry (Snyder's series).
//  * vincenty_m: Vincenty's inverse formula on the ellipsoid (at most 64 iterations, |dlambda| < 1e-14).
//  * pt_seg_d2 / seg_seg_d2: planar closest points in a metric frame.
//  * arc_pt / geog_dist: closest points of great-circle edges on the unit sphere (the construction of PostGIS's
//    edge_distance_to_point, with its robust_cross_product for the edge normal), then the ellipsoidal distance
//    between those points.
//  * geog_band: half-width of the band around d in which the straight-edge frame distance may disagree with the
//    geography distance.
#pragma once
#include "gmath.h"

namespace gsj {
namespace geo {

constexpr double WGS84_A = 6378137.0;
constexpr double WGS84_F = 1.0 / 298.257223563;
constexpr double WGS84_E2 = WGS84_F * (2.0 - WGS84_F);
constexpr double DEG = gm::PI_HI / 180.0;

// metres per degree of longitude (kx) and of latitude (ky) at latitude lat (degrees) (buffer.cuh:33-39)
GSJ_PORT void m_per_deg(double lat, double& kx, double& ky) {
  double s, c;
  gm::sincos(lat * DEG, s, c);
  const double w2 = 1.0 - WGS84_E2 * s * s, w = sqrt(w2);
  kx = WGS84_A / w * c * DEG;
  ky = WGS84_A * (1.0 - WGS84_E2) / (w2 * w) * DEG;
}

// degrees covering d metres at every latitude in [y0, y1]: larger (expansion) or smaller (shrink) (buffer.cuh:42-54)
GSJ_PORT void deg_cover(double y0, double y1, double d, bool larger, double& dx, double& dy) {
  if (d <= 0) { dx = dy = 0; return; }
  const double pad = larger ? d / 110000.0 + 1e-9 : 0.0;   // points covered may lie up to about d beyond the range
  double lo = (y0 - pad <= 0 && y1 + pad >= 0) ? 0.0 : fmin(fabs(y0), fabs(y1)) - pad;
  const double hi = fmin(fmax(fabs(y0), fabs(y1)) + pad, 89.9);
  lo = fmax(lo, 0.0);
  double kx_hi, ky_hi, kx_lo, ky_lo;
  m_per_deg(hi, kx_hi, ky_hi);   // at the highest |lat|: kx smallest, ky largest
  m_per_deg(lo, kx_lo, ky_lo);
  if (larger) { dx = d / kx_hi * (1 + 1e-6); dy = d / ky_lo * (1 + 1e-6); }
  else { dx = d / kx_lo * (1 - 1e-6); dy = d / ky_hi * (1 - 1e-6); }
}

// scale factor at (lon, lat) of the UTM zone chosen for a geometry centred at (glon, glat); polar caps use an
// equal-area projection there, k = 1 (buffer.cuh:58-67)
GSJ_PORT double utm_k(double glon, double glat, double lon, double lat) {
  if (glat > 70.0 || glat < -70.0) return 1.0;
  int zone = (int)floor((glon + 180.0) / 6.0);
  if (zone > 59) zone = 59;
  const double lon0 = zone * 6.0 - 177.0;
  double s, c;
  gm::sincos(lat * DEG, s, c);
  const double t = s / c, T = t * t;
  const double ep2 = WGS84_E2 / (1.0 - WGS84_E2), C = ep2 * c * c, A = (lon - lon0) * DEG * c, A2 = A * A;
  return 0.9996 * (1.0 + (1.0 + C) * A2 / 2.0 + (5.0 - 4.0 * T + 42.0 * C + 13.0 * C * C - 28.0 * ep2) * A2 * A2 / 24.0 +
                   (61.0 - 148.0 * T + 16.0 * T * T) * A2 * A2 * A2 / 720.0);
}

// ellipsoidal distance (metres) by Vincenty's inverse method (buffer.cuh:70-99)
GSJ_PORT double vincenty_m(double lon1, double lat1, double lon2, double lat2) {
  const double a = WGS84_A, f = WGS84_F, b = a * (1.0 - f);
  const double L = (lon2 - lon1) * DEG;
  const double U1 = gm::atan((1.0 - f) * gm::tan(lat1 * DEG)), U2 = gm::atan((1.0 - f) * gm::tan(lat2 * DEG));
  double sU1, cU1, sU2, cU2;
  gm::sincos(U1, sU1, cU1);
  gm::sincos(U2, sU2, cU2);
  double lam = L, sinS = 0, cosS = 1, sig = 0, cos2A = 1, cos2Sm = 0;
  for (int it = 0; it < 64; it++) {
    double sl, cl;
    gm::sincos(lam, sl, cl);
    const double t1 = cU2 * sl, t2 = cU1 * sU2 - sU1 * cU2 * cl;
    sinS = sqrt(t1 * t1 + t2 * t2);
    if (sinS == 0) return 0.0;
    cosS = sU1 * sU2 + cU1 * cU2 * cl;
    sig = gm::atan2(sinS, cosS);
    const double sinA = cU1 * cU2 * sl / sinS;
    cos2A = 1.0 - sinA * sinA;
    cos2Sm = cos2A != 0 ? cosS - 2.0 * sU1 * sU2 / cos2A : 0.0;
    const double C = f / 16.0 * cos2A * (4.0 + f * (4.0 - 3.0 * cos2A));
    const double prev = lam;
    lam = L + (1.0 - C) * f * sinA * (sig + C * sinS * (cos2Sm + C * cosS * (-1.0 + 2.0 * cos2Sm * cos2Sm)));
    if (fabs(lam - prev) < 1e-14) break;
  }
  const double u2 = cos2A * (a * a - b * b) / (b * b);
  const double A = 1.0 + u2 / 16384.0 * (4096.0 + u2 * (-768.0 + u2 * (320.0 - 175.0 * u2)));
  const double B = u2 / 1024.0 * (256.0 + u2 * (-128.0 + u2 * (74.0 - 47.0 * u2)));
  const double dS = B * sinS * (cos2Sm + B / 4.0 * (cosS * (-1.0 + 2.0 * cos2Sm * cos2Sm) -
                                                    B / 6.0 * cos2Sm * (-3.0 + 4.0 * sinS * sinS) * (-3.0 + 4.0 * cos2Sm * cos2Sm)));
  return b * A * (sig - dS);
}

// squared distance from p to segment a-b and the closest point q (buffer.cuh:101-107)
GSJ_PORT double pt_seg_d2(double px, double py, double ax, double ay, double bx, double by, double& qx, double& qy) {
  const double dx = bx - ax, dy = by - ay, l2 = dx * dx + dy * dy;
  double t = l2 > 0 ? ((px - ax) * dx + (py - ay) * dy) / l2 : 0.0;
  t = fmin(1.0, fmax(0.0, t));
  qx = ax + t * dx;
  qy = ay + t * dy;
  return (px - qx) * (px - qx) + (py - qy) * (py - qy);
}

// squared distance of two non-crossing segments a and b, closest points p (on a) and q (on b) (buffer.cuh:110-122):
// the minimum of the four endpoint-to-segment distances, first minimum kept
GSJ_PORT double seg_seg_d2(double a0x, double a0y, double a1x, double a1y, double b0x, double b0y, double b1x, double b1y, double& px, double& py,
                           double& qx, double& qy) {
  double tx, ty;
  double best = pt_seg_d2(a0x, a0y, b0x, b0y, b1x, b1y, qx, qy);
  px = a0x; py = a0y;
  double d = pt_seg_d2(a1x, a1y, b0x, b0y, b1x, b1y, tx, ty);
  if (d < best) { best = d; px = a1x; py = a1y; qx = tx; qy = ty; }
  d = pt_seg_d2(b0x, b0y, a0x, a0y, a1x, a1y, tx, ty);
  if (d < best) { best = d; px = tx; py = ty; qx = b0x; qy = b0y; }
  d = pt_seg_d2(b1x, b1y, a0x, a0y, a1x, a1y, tx, ty);
  if (d < best) { best = d; px = tx; py = ty; qx = b1x; qy = b1y; }
  return best;
}

// ---- great-circle edges on the unit sphere (buffer.cuh:128-176) ------------------------------------------------
struct V3 { double x, y, z; };
GSJ_PORT V3 v3_ll(double lon, double lat) {   // radians
  double sl, cl, sp, cp;
  gm::sincos(lon, sl, cl);
  gm::sincos(lat, sp, cp);
  return V3{cp * cl, cp * sl, sp};
}
GSJ_PORT double v3_dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
GSJ_PORT V3 v3_norm(V3 a) {
  const double l = sqrt(v3_dot(a, a));
  return l > 0 ? V3{a.x / l, a.y / l, a.z / l} : a;
}
GSJ_PORT double v3_ang(V3 a, V3 b) {
  const V3 c{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
  return gm::atan2(sqrt(v3_dot(c, c)), v3_dot(a, b));
}
// p x q for nearby points from lon/lat (radians), written with half-sum / half-difference angles so the result keeps
// its relative accuracy as q -> p (PostGIS's robust_cross_product)
GSJ_PORT V3 robust_cross(double plon, double plat, double qlon, double qlat) {
  const double lon_qpp = (qlon + plon) / -2.0, lon_qmp = (qlon - plon) / 2.0;
  const double smq = gm::sin(plat - qlat), spq = gm::sin(plat + qlat);
  double s1, c1, s2, c2;
  gm::sincos(lon_qpp, s1, c1);
  gm::sincos(lon_qmp, s2, c2);
  return V3{smq * s1 * c2 - spq * c1 * s2, smq * c1 * c2 + spq * s1 * s2, gm::cos(plat) * gm::cos(qlat) * gm::sin(qlon - plon)};
}
// angle from p to the great-circle edge s-e (radians in and out) and the closest point (clon, clat)
GSJ_PORT double arc_pt(double slon, double slat, double elon, double elat, double plon, double plat, double& clon, double& clat) {
  const V3 vs = v3_ll(slon, slat), ve = v3_ll(elon, elat), vp = v3_ll(plon, plat);
  double best = 1e9;
  if (!(slon == elon && slat == elat)) {
    const V3 n = v3_norm(robust_cross(slon, slat, elon, elat));
    const double t = v3_dot(vp, n);
    const V3 k = v3_norm(V3{vp.x - t * n.x, vp.y - t * n.y, vp.z - t * n.z});
    const V3 vc = v3_norm(V3{vs.x + ve.x, vs.y + ve.y, vs.z + ve.z});   // the projection lies inside the edge's cone
    const double ds = v3_dot(vs, vc), dk = v3_dot(k, vc);
    if (dk > ds || fabs(dk - ds) < 2e-16) {
      best = v3_ang(vp, k);
      clat = gm::atan2(k.z, sqrt(k.x * k.x + k.y * k.y));
      clon = gm::atan2(k.y, k.x);
    }
  }
  const double d2 = v3_ang(vp, vs), d3 = v3_ang(vp, ve);
  if (d2 < best) { best = d2; clon = slon; clat = slat; }
  if (d3 < best) { best = d3; clon = elon; clat = elat; }
  return best;
}
// geography distance (metres) of edge a from edge b, or of a from point b0 (b_is_point); degrees in
// (buffer.cuh:162-176); not inlined: called only for edge pairs inside the band, it would otherwise dominate the callers' register use
__attribute__((noinline)) GSJ_PORT double geog_dist(double a0x, double a0y, double a1x, double a1y, double b0x, double b0y, double b1x, double b1y, bool b_is_point) {
  const double R = DEG;
  double cx = 0, cy = 0, p1x, p1y, p2x, p2y;
  double best = arc_pt(a0x * R, a0y * R, a1x * R, a1y * R, b0x * R, b0y * R, cx, cy);
  p1x = cx; p1y = cy; p2x = b0x * R; p2y = b0y * R;
  if (!b_is_point) {
    double d = arc_pt(a0x * R, a0y * R, a1x * R, a1y * R, b1x * R, b1y * R, cx, cy);
    if (d < best) { best = d; p1x = cx; p1y = cy; p2x = b1x * R; p2y = b1y * R; }
    d = arc_pt(b0x * R, b0y * R, b1x * R, b1y * R, a0x * R, a0y * R, cx, cy);
    if (d < best) { best = d; p1x = a0x * R; p1y = a0y * R; p2x = cx; p2y = cy; }
    d = arc_pt(b0x * R, b0y * R, b1x * R, b1y * R, a1x * R, a1y * R, cx, cy);
    if (d < best) { best = d; p1x = a1x * R; p1y = a1y * R; p2x = cx; p2y = cy; }
  }
  return vincenty_m(p1x / R, p1y / R, p2x / R, p2y / R);
}
// half-width (metres) of the band around d where the frame distance may disagree with geography; La2 / Lb2 are the
// squared edge lengths in the frame, tl = |tan(lat)| + 0.01 (buffer.cuh:179-181)
GSJ_PORT double geog_band(double d, double rel_frame, double La2, double Lb2, double tl) {
  return d * (rel_frame + 5e-5) + (La2 + Lb2) * tl / (8.0 * 6.3e6) + 1e-9;
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
