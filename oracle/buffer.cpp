// ORACLE — test infrastructure only. CPU restatement of the reference's distance joins (engine/buffer.cuh), the
// checker for the HIP engine's dwithin paths. Nothing in the product links or calls it.
//
// The WGS84 primitives (m_per_deg, deg_cover, vincenty, great-circle closest points, frame band) come from the
// engine's host+device header geodesy.h, built here with g++ -ffp-contract=off, so that CPU and GPU decide every
// threshold with identical bits; those primitives are pinned independently of both by the tests (mpmath Vincenty at
// 50 digits, closed forms on the equator and meridian, radii of curvature, UTM scale on the central meridian,
// elementary functions against mpmath). The pair logic below is restated here from the reference.
//
// Reported distances are new (the reference reports NaN, daemon/gpujoind.cu:468). Definition used by oracle and
// engine alike: for a piece pair within d, 0 if real boundaries cross or a real probe vertex of one lies inside the
// other; else the minimum geography distance over the real edge pairs whose frame distance could hold the minimum
// (frame - band <= min(frame + band)); per parent pair the minimum over its piece pairs. Parity of this definition
// with PostGIS ST_Distance(geography) is unpinned here (no PostGIS).
#include "core.h"
#include "../postgis-cuda_amd/csrc/geodesy.h"

namespace orc {

namespace geo = gsj::geo;

struct DwOut { uint8_t hit = 0; double dist = NAN; };

// real vertices of ring r of S as probes (engine/engine.cuh:322-327)
static inline bool probe_real(const View& S, uint64_t r, double& x, double& y) {
  uint64_t v = S.L->rvo[r], e = S.L->rvo[r + 1];
  if (S.L->has_syn()) { while (v + 1 < e && S.syn(v)) v++; if (v + 1 >= e) return false; }
  if (v >= e) return false;   // empty ring
  Real rx, ry; S.v(v, rx, ry);
  x = val(rx); y = val(ry);
  return true;
}

// engine/buffer.cuh:185-237 (dw_pair), plus the reported distance
static DwOut dw_pair(const Layer& P, const Layer& R, uint32_t a, uint32_t b, double d, bool report) {
  DwOut out;
  const double oax = P.ox[a], oay = P.oy[a], obx = R.ox[b], oby = R.oy[b], eax = P.ex[a], eay = P.ey[a], ebx = R.ex[b], eby = R.ey[b];
  const double sx = obx - oax, sy = oby - oay;
  double gx, gy;
  geo::deg_cover(std::fmin(oay, oby), std::fmax(oay + eay, oby + eby), d, true, gx, gy);
  Box ov{std::fmax(-gx, sx - gx), std::fmax(-gy, sy - gy), std::fmin(eax + gx, sx + ebx + gx), std::fmin(eay + gy, sy + eby + gy)};
  if (ov.x0 > ov.x1 || ov.y0 > ov.y1) return out;
  View A{&P, a, oax - oax, oay - oay}, B{&R, b, sx, sy};
  std::vector<Edge> LA, LB;
  collect(A, ov, LA);
  collect(B, ov, LB);
  const double phi0 = oay + 0.5 * (val(ov.y0) + val(ov.y1));
  double kx, ky;
  geo::m_per_deg(phi0, kx, ky);
  const double tl = std::fabs(gsj::gm::tan(phi0 * geo::DEG)) + 0.01;
  const double delta = 2.0 * 0.5 * (val(ov.y1) - val(ov.y0)) * geo::DEG * (tl + 0.04) + 1e-9;
  bool hit = false, cross = false;
  double m_up = INFINITY;
  // per real edge pair near each other: crossing, frame distance and band (the reference's hit loop, run to the end)
  struct Near { double lo; size_t i, j; };
  std::vector<Near> near;
  for (size_t i = 0; i < LA.size(); i++) {
    const Edge& e = LA[i];
    if (e.syn) continue;
    const double ex0 = val(e.x0), ey0 = val(e.y0), ex1 = val(e.x1), ey1 = val(e.y1);
    const double bx0 = std::fmin(ex0, ex1) - gx, bx1 = std::fmax(ex0, ex1) + gx, by0 = std::fmin(ey0, ey1) - gy, by1 = std::fmax(ey0, ey1) + gy;
    const double ux = (ex1 - ex0) * kx, uy = (ey1 - ey0) * ky, La2 = ux * ux + uy * uy;
    for (size_t j = 0; j < LB.size(); j++) {
      const Edge& f = LB[j];
      const double fx0 = val(f.x0), fy0 = val(f.y0), fx1 = val(f.x1), fy1 = val(f.y1);
      if (f.syn || std::fmax(fx0, fx1) < bx0 || std::fmin(fx0, fx1) > bx1 || std::fmax(fy0, fy1) < by0 || std::fmin(fy0, fy1) > by1) continue;
      if (seg_isect(ex0, ey0, ex1, ey1, fx0, fy0, fx1, fy1)) { hit = cross = true; continue; }
      double px, py, qx, qy;
      const double d2 = geo::seg_seg_d2(ex0 * kx, ey0 * ky, ex1 * kx, ey1 * ky, fx0 * kx, fy0 * ky, fx1 * kx, fy1 * ky, px, py, qx, qy);
      const double vx = (fx1 - fx0) * kx, vy = (fy1 - fy0) * ky;
      const double band = geo::geog_band(d, delta, La2, vx * vx + vy * vy, tl);
      const double dlo = d - band, dhi = d + band;
      if (d2 <= dlo * dlo) hit = true;
      else if (d2 <= dhi * dhi &&
               geo::geog_dist(oax + ex0, oay + ey0, oax + ex1, oay + ey1, oax + fx0, oay + fy0, oax + fx1, oay + fy1, false) <= d)
        hit = true;
      if (report) {
        const double fr = std::sqrt(d2);
        m_up = std::fmin(m_up, fr + band);
        near.push_back({fr - band, i, j});
      }
    }
  }
  bool contained = false;
  if (!hit || report) {
    // no boundary within d: within d iff one contains the other (a real vertex of one inside the other piece)
    const Real cx = ov.x0 + 0.381966 * (ov.x1 - ov.x0), cy = ov.y0 + 0.414213 * (ov.y1 - ov.y0);
    const bool inB_c = pip_full(B, cx, cy), inA_c = pip_full(A, cx, cy);
    for (uint64_t r = A.r0(); r < A.r1() && !contained; r++) {
      double x, y;
      if (probe_real(A, r, x, y) && x >= ov.x0 && x <= ov.x1 && y >= ov.y0 && y <= ov.y1 && inside_from_c(LB, cx, cy, inB_c, x, y)) contained = true;
    }
    for (uint64_t r = B.r0(); r < B.r1() && !contained; r++) {
      double x, y;
      if (probe_real(B, r, x, y) && x >= ov.x0 && x <= ov.x1 && y >= ov.y0 && y <= ov.y1 && inside_from_c(LA, cx, cy, inA_c, x, y)) contained = true;
    }
  }
  out.hit = (hit || contained) ? 1 : 0;
  if (report && out.hit) {
    if (cross || contained) out.dist = 0.0;
    else {
      double best = INFINITY;
      for (const Near& n : near) {
        if (n.lo > m_up) continue;
        const Edge &e = LA[n.i], &f = LB[n.j];
        best = std::fmin(best, geo::geog_dist(oax + val(e.x0), oay + val(e.y0), oax + val(e.x1), oay + val(e.y1), oax + val(f.x0), oay + val(f.y0),
                                              oax + val(f.x1), oay + val(f.y1), false));
      }
      out.dist = best;
    }
  }
  return out;
}

// engine/buffer.cuh:703-715 (k_expand_bb): query boxes of P widened by d metres, rounded outward to float
static void expand_boxes(const Layer& P, double d, std::vector<float>& qbb) {
  qbb.resize(4 * P.n);
  for (uint64_t i = 0; i < P.n; i++) {
    double dx, dy;
    geo::deg_cover(P.oy[i], P.oy[i] + P.ey[i], d, true, dx, dy);
    qbb[4 * i] = geo::f_down(P.ox[i] - dx); qbb[4 * i + 1] = geo::f_down(P.oy[i] - dy);
    qbb[4 * i + 2] = geo::f_up(P.ox[i] + P.ex[i] + dx); qbb[4 * i + 3] = geo::f_up(P.oy[i] + P.ey[i] + dy);
  }
}

struct DwResult {
  std::vector<uint32_t> pairs;
  std::vector<double> dist;
  uint64_t candidates = 0;
};

// unique (key) rows with the minimum distance per key
static void fold_min(std::vector<std::pair<uint64_t, double>>& kv, DwResult& res) {
  std::sort(kv.begin(), kv.end(), [](const auto& x, const auto& y) { return x.first < y.first; });
  res.pairs.clear(); res.dist.clear();
  for (size_t s = 0; s < kv.size();) {
    size_t e = s;
    double m = kv[s].second;
    while (e < kv.size() && kv[e].first == kv[s].first) { m = std::fmin(m, kv[e].second); e++; }
    res.pairs.push_back((uint32_t)(kv[s].first >> 32)); res.pairs.push_back((uint32_t)kv[s].first); res.dist.push_back(m);
    s = e;
  }
}

// engine/buffer.cuh:892-991 (dwithin_join_stream): candidates of the widened query boxes, dw_pair per piece pair,
// pieces of one parent within any distance of each other on a self-join, unique parent pairs
static void dwithin(const Layer& P, const Layer& R, double d, bool same, bool upper, bool report, DwResult& res) {
  std::vector<float> qbb;
  const std::vector<float>* q = &P.bb;
  if (d > 0) { expand_boxes(P, d, qbb); q = &qbb; }
  std::vector<uint32_t> cand;
  candidates(*q, P.n, R.bb, R.n, upper, cand);
  const uint64_t np = cand.size() / 2;
  res.candidates = np;
  std::vector<DwOut> o(np);
#pragma omp parallel for schedule(dynamic, 256)
  for (int64_t i = 0; i < (int64_t)np; i++) {
    const uint32_t a = cand[2 * i], b = cand[2 * i + 1];
    const uint32_t pa = P.parent.empty() ? a : P.parent[a], pb = R.parent.empty() ? b : R.parent[b];
    if (same && pa == pb) { o[i].hit = 1; o[i].dist = 0.0; continue; }
    o[i] = dw_pair(P, R, a, b, d, report);
  }
  std::vector<std::pair<uint64_t, double>> kv;
  for (uint64_t i = 0; i < np; i++) {
    if (!o[i].hit) continue;
    const uint64_t pa = P.parent.empty() ? cand[2 * i] : P.parent[cand[2 * i]], pb = R.parent.empty() ? cand[2 * i + 1] : R.parent[cand[2 * i + 1]];
    kv.push_back({(pa << 32) | pb, o[i].dist});
  }
  fold_min(kv, res);
}

// engine/buffer.cuh:1018-1049 (k_pt_dwithin): inside the piece, or within d of its real boundary; with report, the
// distance (0 inside, else the minimum geography distance over the edges that could hold it)
static DwOut pt_dwithin(const Layer& pts, const Layer& R, uint32_t pi, uint32_t g, double d, bool report) {
  DwOut out;
  const double ptx = pts.xw[2 * pi], pty = pts.xw[2 * pi + 1], ox_ = R.ox[g], oy_ = R.oy[g];
  View B{&R, g, ox_ - ox_, oy_ - oy_};
  const double px = ptx - ox_, py = pty - oy_;
  double kx, ky;
  geo::m_per_deg(pty, kx, ky);
  const double gy = d / ky;
  const double tl = std::fabs(gsj::gm::tan(pty * geo::DEG)) + 0.01;
  const double delta = 2.0 * gy * geo::DEG * (tl + 0.04) + 1e-9;
  const double ox = val(B.shx) - 1.0;
  bool par = false, h = false;
  double m_up = INFINITY;
  struct Near { double lo; uint64_t j; };
  std::vector<Near> near;
  for (uint64_t r = B.r0(); r < B.r1() && (report || !h); r++) {
    const uint64_t s0 = R.rvo[r], e = R.rvo[r + 1];
    Real rx0, ry0; B.v(s0, rx0, ry0);
    double x0 = val(rx0), y0 = val(ry0);
    for (uint64_t j = s0 + 1; j < e && (report || !h); j++) {
      Real rx1, ry1; B.v(j, rx1, ry1);
      const double x1 = val(rx1), y1 = val(ry1);
      if (!B.syn(j - 1)) {
        double qx, qy;
        const double d2 = geo::pt_seg_d2(px * kx, py * ky, x0 * kx, y0 * ky, x1 * kx, y1 * ky, qx, qy);
        const double ex = (x1 - x0) * kx, ey = (y1 - y0) * ky;
        const double band = geo::geog_band(d, delta, ex * ex + ey * ey, 0.0, tl), dlo = d - band, dhi = d + band;
        if (d2 <= dlo * dlo || (d2 <= dhi * dhi && geo::geog_dist(ox_ + x0, oy_ + y0, ox_ + x1, oy_ + y1, ptx, pty, ptx, pty, true) <= d)) h = true;
        if (report) {
          const double fr = std::sqrt(d2);
          m_up = std::fmin(m_up, fr + band);
          near.push_back({fr - band, j});
        }
      }
      if (!((y0 > py) == (y1 > py) && y0 != py && y1 != py)) {
        const bool so = left_of(x0, y0, x1, y1, ox, py), sp = left_of(x0, y0, x1, y1, px, py);
        if (so != sp) {
          const bool su = orient(ox, py, px, py, x0, y0) > 0, sv = orient(ox, py, px, py, x1, y1) > 0;
          if (su != sv) par = !par;
        }
      }
      x0 = x1; y0 = y1;
    }
  }
  out.hit = (h || par) ? 1 : 0;
  if (report && out.hit) {
    if (par) out.dist = 0.0;
    else {
      double best = INFINITY;
      for (const Near& n : near) {
        if (n.lo > m_up) continue;
        Real ax, ay, bx, by;
        B.v(n.j - 1, ax, ay); B.v(n.j, bx, by);
        best = std::fmin(best, geo::geog_dist(ox_ + val(ax), oy_ + val(ay), ox_ + val(bx), oy_ + val(by), ptx, pty, ptx, pty, true));
      }
      out.dist = best;
    }
  }
  return out;
}

// engine/buffer.cuh:1004-1016 (k_expand_pts) and 1137-1229 (points_buffer_join_stream, dwithin mode): unique
// (point, parent) rows
static void points_dwithin(const Layer& pts, const Layer& R, double d, bool report, DwResult& res) {
  std::vector<float> qbb(4 * pts.n);
  for (uint64_t i = 0; i < pts.n; i++) {
    const double x = pts.xw[2 * i], y = pts.xw[2 * i + 1];
    double dx = 0, dy = 0;
    geo::deg_cover(y, y, d, true, dx, dy);
    qbb[4 * i] = geo::f_down(x - dx); qbb[4 * i + 1] = geo::f_down(y - dy); qbb[4 * i + 2] = geo::f_up(x + dx); qbb[4 * i + 3] = geo::f_up(y + dy);
  }
  std::vector<uint32_t> cand;
  candidates(d > 0 ? qbb : pts.bb, pts.n, R.bb, R.n, false, cand);
  const uint64_t np = cand.size() / 2;
  res.candidates = np;
  std::vector<DwOut> o(np);
#pragma omp parallel for schedule(dynamic, 1024)
  for (int64_t i = 0; i < (int64_t)np; i++) o[i] = pt_dwithin(pts, R, cand[2 * i], cand[2 * i + 1], d, report);
  std::vector<std::pair<uint64_t, double>> kv;
  for (uint64_t i = 0; i < np; i++) {
    if (!o[i].hit) continue;
    const uint64_t pb = R.parent.empty() ? cand[2 * i + 1] : R.parent[cand[2 * i + 1]];
    kv.push_back({((uint64_t)cand[2 * i] << 32) | pb, o[i].dist});
  }
  fold_min(kv, res);
}

}  // namespace orc

using namespace orc;

extern "C" {

// portable elementary functions and geodesy (for the tests that pin them)
double orc_gm_sin(double x) { return gsj::gm::sin(x); }
double orc_gm_cos(double x) { return gsj::gm::cos(x); }
double orc_gm_tan(double x) { return gsj::gm::tan(x); }
double orc_gm_atan(double x) { return gsj::gm::atan(x); }
double orc_gm_atan2(double y, double x) { return gsj::gm::atan2(y, x); }
void orc_geo_m_per_deg(double lat, double* kx, double* ky) { geo::m_per_deg(lat, *kx, *ky); }
void orc_geo_deg_cover(double y0, double y1, double d, int larger, double* dx, double* dy) { geo::deg_cover(y0, y1, d, larger != 0, *dx, *dy); }
double orc_geo_utm_k(double glon, double glat, double lon, double lat) { return geo::utm_k(glon, glat, lon, lat); }
double orc_geo_vincenty(double lon1, double lat1, double lon2, double lat2) { return geo::vincenty_m(lon1, lat1, lon2, lat2); }
double orc_geo_geog_dist(double a0x, double a0y, double a1x, double a1y, double b0x, double b0y, double b1x, double b1y, int b_is_point) {
  return geo::geog_dist(a0x, a0y, a1x, a1y, b0x, b0y, b1x, b1y, b_is_point != 0);
}

int orc_dw_pair(const void* P, const void* R, uint32_t a, uint32_t b, double d, int report, double* dist) {
  DwOut o = dw_pair(*(const Layer*)P, *(const Layer*)R, a, b, d, report != 0);
  if (dist) *dist = o.dist;
  return o.hit;
}

void* orc_dw_new() { return new DwResult(); }
void orc_dw_free(void* r) { delete (DwResult*)r; }
uint64_t orc_dw_rows(const void* r) { return ((const DwResult*)r)->pairs.size() / 2; }
uint64_t orc_dw_candidates(const void* r) { return ((const DwResult*)r)->candidates; }
void orc_dw_copy(const void* rp, uint32_t* pairs, double* dist) {
  const DwResult& r = *(const DwResult*)rp;
  if (pairs && !r.pairs.empty()) std::memcpy(pairs, r.pairs.data(), r.pairs.size() * 4);
  if (dist && !r.dist.empty()) std::memcpy(dist, r.dist.data(), r.dist.size() * 8);
}
void orc_dwithin(const void* P, const void* R, double d, int same, int upper, int report, void* res) {
  dwithin(*(const Layer*)P, *(const Layer*)R, d, same != 0, upper != 0, report != 0, *(DwResult*)res);
}
void orc_points_dwithin(const void* pts, const void* R, double d, int report, void* res) {
  points_dwithin(*(const Layer*)pts, *(const Layer*)R, d, report != 0, *(DwResult*)res);
}

}  // extern "C"
