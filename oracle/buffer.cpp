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
  const geo::Frame F = geo::frame_of(oay + val(ov.y0), oay + val(ov.y1));   // metric frame of the pair's window
  const double kx = F.kx, ky = F.ky;
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
      const double d2 = geo::seg_seg_d2(ex0 * kx, ey0 * ky, ex1 * kx, ey1 * ky, fx0 * kx, fy0 * ky, fx1 * kx, fy1 * ky);
      const double vx = (fx1 - fx0) * kx, vy = (fy1 - fy0) * ky;
      const double band = geo::geog_band(d, F, La2, vx * vx + vy * vy);
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
  const double reach = d / geo::KY_MIN;   // latitude the closest boundary point can be from the point
  const geo::Frame F = geo::frame_of(pty - reach, pty + reach);
  const double kx = F.kx, ky = F.ky;
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
        const double d2 = geo::pt_seg_d2(px * kx, py * ky, x0 * kx, y0 * ky, x1 * kx, y1 * ky);
        const double ex = (x1 - x0) * kx, ey = (y1 - y0) * ky;
        const double band = geo::geog_band(d, F, ex * ex + ey * ey, 0.0), dlo = d - band, dhi = d + band;
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

// =========================================================================================== eroded intersects
// engine/buffer.cuh:275-588 (er_pair and its parts), 717-889 (filters and driver), 1053-1131 (point side).
// Lists are unbounded here; the engine's capacity handling (LDS lists, then global-memory lists of the size a pair
// needs) evaluates the same window, so both give the same answer.
constexpr double FILLET_QUANTUM = gsj::gm::PI_HI / 16.0;   // GEOS: 8 quadrant segments

// real boundary edge in the pair's metric frame; vt: 0 no region at its end vertex, 1 convex, m + 1 reflex (m chords)
struct BEd { double x0, y0, x1, y1, ux, uy, L, ci, si; int succ, vt; };
struct BEl { double x0, y0, x1, y1; };   // element: piece of the raw offset curve

// buffer.cuh:285-320: strictly inside (by more than eps) the region within r of the boundary
static bool in_excl(const std::vector<BEd>& E, double r, double x, double y, double eps) {
  if (r <= 0) return false;
  const double ri = r - eps;
  for (const BEd& e : E) {
    if (x < std::fmin(e.x0, e.x1) - r || x > std::fmax(e.x0, e.x1) + r || y < std::fmin(e.y0, e.y1) - r || y > std::fmax(e.y0, e.y1) + r) continue;
    const double dx = x - e.x0, dy = y - e.y0;
    const double t = dx * e.ux + dy * e.uy, s = dy * e.ux - dx * e.uy;
    if (std::fabs(s) < ri && t > -eps && t < e.L + eps) return true;
    if (e.vt == 0) continue;
    const double vx = x - e.x1, vy = y - e.y1, d2 = vx * vx + vy * vy;
    if (d2 >= r * r) continue;
    const BEd& f = E[e.succ];
    if (e.vt == 1) {
      const double s2 = vy * f.ux - vx * f.uy;
      if (d2 < ri * ri && !(s > eps && s2 > eps)) return true;
      continue;
    }
    const double nkx = -e.uy, nky = e.ux, nsx = -f.uy, nsy = f.ux;
    if (nkx * vy - nky * vx > eps) continue;
    if (nsx * vy - nsy * vx < -eps) continue;
    const double chh = std::sqrt(0.5 * (1.0 + e.ci)), shh = std::sqrt(std::fmax(0.0, 0.5 * (1.0 - e.ci)));
    const double ch = r * chh - eps;
    if (d2 < ch * ch) return true;
    double mx = nkx * chh + nky * shh, my = -nkx * shh + nky * chh;
    bool in = true;
    for (int i = 0, m = e.vt - 1; i < m && in; i++) {
      if (mx * vx + my * vy >= ch) in = false;
      const double t2 = mx * e.ci + my * e.si;
      my = -mx * e.si + my * e.ci;
      mx = t2;
    }
    if (in) return true;
  }
  return false;
}

// buffer.cuh:323-335: inside(x) = inside(c) xor parity of the real edges crossing c -> x
static bool inside_m(const std::vector<BEd>& E, double cx, double cy, bool in_c, double x, double y) {
  bool in = in_c;
  const double sx0 = std::fmin(cx, x), sx1 = std::fmax(cx, x), sy0 = std::fmin(cy, y), sy1 = std::fmax(cy, y);
  for (const BEd& e : E) {
    if (std::fmax(e.x0, e.x1) < sx0 || std::fmin(e.x0, e.x1) > sx1 || std::fmax(e.y0, e.y1) < sy0 || std::fmin(e.y0, e.y1) > sy1) continue;
    const bool sc = left_of(e.x0, e.y0, e.x1, e.y1, cx, cy), sx = left_of(e.x0, e.y0, e.x1, e.y1, x, y);
    if (sc == sx) continue;
    const bool su = orient(cx, cy, x, y, e.x0, e.y0) > 0, sv = orient(cx, cy, x, y, e.x1, e.y1) > 0;
    if (su != sv) in = !in;
  }
  return in;
}

struct ESide { uint32_t g0, g1; double r, bx0, by0, bx1, by1; };
struct EFrame { double ox, oy, kx, ky; };

// frame metres of vertex v of a piece whose origin is (frame origin + (shx, shy)) (buffer.cuh:343-347)
static inline void fv(const Layer& S, double shx, double shy, uint64_t v, const EFrame& F, double& x, double& y) {
  x = (shx + S.q[2 * v]) * F.kx;
  y = (shy + S.q[2 * v + 1]) * F.ky;
}

// buffer.cuh:351-374: c inside the parent (c lies in one piece): parity over the pieces whose bbox holds c
static bool pip_parent(const Layer& S, const ESide& s, const EFrame& F, double cx, double cy) {
  bool in = false;
  for (uint32_t g = s.g0; g < s.g1 && !in; g++) {
    const double bx0 = (S.ox[g] - F.ox) * F.kx, by0 = (S.oy[g] - F.oy) * F.ky, bx1 = (S.ox[g] + S.ex[g] - F.ox) * F.kx,
                 by1 = (S.oy[g] + S.ey[g] - F.oy) * F.ky;
    if (cx < bx0 - 1e-6 || cx > bx1 + 1e-6 || cy < by0 - 1e-6 || cy > by1 + 1e-6) continue;
    const double shx = S.ox[g] - F.ox, shy = S.oy[g] - F.oy, ox = bx0 - 1.0;
    bool par = false;
    for (uint64_t r = S.gro[g]; r < S.gro[g + 1]; r++)
      for (uint64_t j = S.rvo[r] + 1; j < S.rvo[r + 1]; j++) {
        double x0, y0, x1, y1;
        fv(S, shx, shy, j - 1, F, x0, y0);
        fv(S, shx, shy, j, F, x1, y1);
        if ((y0 > cy) == (y1 > cy) && y0 != cy && y1 != cy) continue;
        const bool so = left_of(x0, y0, x1, y1, ox, cy), sc = left_of(x0, y0, x1, y1, cx, cy);
        if (so == sc) continue;
        const bool su = orient(ox, cy, cx, cy, x0, y0) > 0, sv = orient(ox, cy, cx, cy, x1, y1) > 0;
        if (su != sv) par = !par;
      }
    in = par;
  }
  return in;
}

// buffer.cuh:377-405: the parent's real, non-degenerate edges meeting box C (frame metres), in piece/ring order
static void collect_side(const Layer& S, const ESide& s, const EFrame& F, double cx0, double cy0, double cx1, double cy1, std::vector<BEd>& E) {
  E.clear();
  for (uint32_t g = s.g0; g < s.g1; g++) {
    if ((S.ox[g] + S.ex[g] - F.ox) * F.kx < cx0 || (S.ox[g] - F.ox) * F.kx > cx1 || (S.oy[g] + S.ey[g] - F.oy) * F.ky < cy0 ||
        (S.oy[g] - F.oy) * F.ky > cy1)
      continue;
    const double shx = S.ox[g] - F.ox, shy = S.oy[g] - F.oy;
    for (uint64_t r = S.gro[g]; r < S.gro[g + 1]; r++)
      for (uint64_t i = S.rvo[r]; i + 1 < S.rvo[r + 1]; i++) {
        if (S.is_syn(i)) continue;
        BEd ed{};
        fv(S, shx, shy, i, F, ed.x0, ed.y0);
        fv(S, shx, shy, i + 1, F, ed.x1, ed.y1);
        if (ed.x0 == ed.x1 && ed.y0 == ed.y1) continue;
        if (std::fmax(ed.x0, ed.x1) < cx0 || std::fmin(ed.x0, ed.x1) > cx1 || std::fmax(ed.y0, ed.y1) < cy0 || std::fmin(ed.y0, ed.y1) > cy1) continue;
        E.push_back(ed);
      }
  }
}

// buffer.cuh:408-438: directions, successors and end-vertex regions
static void prepare_side(std::vector<BEd>& E, double r) {
  const int n = (int)E.size();
  for (BEd& e : E) {
    const double dx = e.x1 - e.x0, dy = e.y1 - e.y0, L = std::sqrt(dx * dx + dy * dy);
    e.ux = dx / L; e.uy = dy / L; e.L = L;
  }
  for (int k = 0; k < n; k++) {
    const BEd e = E[k];
    int s = -1;
    if (k + 1 < n && E[k + 1].x0 == e.x1 && E[k + 1].y0 == e.y1) s = k + 1;
    else
      for (int j = 0; j < n; j++)
        if (j != k && E[j].x0 == e.x1 && E[j].y0 == e.y1) { s = j; break; }
    int vt = 0;
    double ci = 1, si = 0;
    if (s >= 0 && r > 0) {
      const BEd& f = E[s];
      const double cr = e.ux * f.uy - e.uy * f.ux, dt = e.ux * f.ux + e.uy * f.uy;
      if (cr > 0) vt = 1;
      else if (cr < 0 || dt < 0) {   // right turn, or a reversal: fillet (GEOS joins a reversal with a half circle)
        double th = gsj::gm::atan2(-cr, dt);
        if (cr == 0) th = gsj::gm::PI_HI;
        int m = (int)(th / FILLET_QUANTUM + 0.5);
        if (m < 1) m = 1;
        const double inc = th / m;
        gsj::gm::sincos(inc, si, ci);
        vt = m + 1;
      }
    }
    E[k].succ = s; E[k].vt = vt; E[k].ci = ci; E[k].si = si;
  }
}

// buffer.cuh:441-475: offset edges and reflex fillet chords meeting the window (+- eps)
static void build_elements(const std::vector<BEd>& E, double r, double wx0, double wy0, double wx1, double wy1, double eps, std::vector<BEl>& L) {
  L.clear();
  for (const BEd& e : E) {
    const int ne = 1 + (e.vt >= 2 ? e.vt - 1 : 0);
    const double nkx = -e.uy, nky = e.ux;
    double wxr = nkx, wyr = nky, px = e.x1 + r * nkx, py = e.y1 + r * nky, nsx = 0, nsy = 0;
    if (e.vt >= 2) { nsx = -E[e.succ].uy; nsy = E[e.succ].ux; }
    for (int it = 0; it < ne; it++) {
      BEl el;
      if (it == 0) { el.x0 = e.x0 + r * nkx; el.y0 = e.y0 + r * nky; el.x1 = e.x1 + r * nkx; el.y1 = e.y1 + r * nky; }
      else {
        const double t2 = wxr * e.ci + wyr * e.si;
        wyr = -wxr * e.si + wyr * e.ci;
        wxr = t2;
        double qx, qy;
        if (it == ne - 1) { qx = e.x1 + r * nsx; qy = e.y1 + r * nsy; }
        else { qx = e.x1 + r * wxr; qy = e.y1 + r * wyr; }
        el.x0 = px; el.y0 = py; el.x1 = qx; el.y1 = qy;
        px = qx; py = qy;
      }
      if (!(std::fmax(el.x0, el.x1) < wx0 - eps || std::fmin(el.x0, el.x1) > wx1 + eps || std::fmax(el.y0, el.y1) < wy0 - eps ||
            std::fmin(el.y0, el.y1) > wy1 + eps))
        L.push_back(el);
    }
  }
}

// buffer.cuh:478-498: intersection points of two intersecting segments (1, or 3 for a collinear overlap)
static int seg_points(const BEl& p, const BEl& q, double* xs, double* ys) {
  const double rx = p.x1 - p.x0, ry = p.y1 - p.y0, sx = q.x1 - q.x0, sy = q.y1 - q.y0;
  const double den = rx * sy - ry * sx;
  if (den != 0) {
    double t = ((q.x0 - p.x0) * sy - (q.y0 - p.y0) * sx) / den;
    t = std::fmin(1.0, std::fmax(0.0, t));
    xs[0] = p.x0 + t * rx; ys[0] = p.y0 + t * ry;
    return 1;
  }
  const double l2 = rx * rx + ry * ry;
  if (l2 == 0) { xs[0] = p.x0; ys[0] = p.y0; return 1; }
  const double t0 = ((q.x0 - p.x0) * rx + (q.y0 - p.y0) * ry) / l2, t1 = ((q.x1 - p.x0) * rx + (q.y1 - p.y0) * ry) / l2;
  double lo = std::fmax(0.0, std::fmin(t0, t1)), hi = std::fmin(1.0, std::fmax(t0, t1));
  if (lo > hi) { const double m = 0.5 * (lo + hi); lo = hi = std::fmin(1.0, std::fmax(0.0, m)); }
  xs[0] = p.x0 + lo * rx; ys[0] = p.y0 + lo * ry;
  xs[1] = p.x0 + hi * rx; ys[1] = p.y0 + hi * ry;
  xs[2] = p.x0 + 0.5 * (lo + hi) * rx; ys[2] = p.y0 + 0.5 * (lo + hi) * ry;
  return 3;
}

// pieces of parent p: [pf[p], pf[p + 1]) of a subdivided layer, or p itself
static inline void piece_range(const Layer& S, const std::vector<uint32_t>& pf, uint32_t p, uint32_t& g0, uint32_t& g1) {
  if (pf.empty()) { g0 = p; g1 = p + 1; }
  else { g0 = pf[p]; g1 = pf[p + 1]; }
}
static std::vector<uint32_t> piece_first(const Layer& S) {
  std::vector<uint32_t> pf;
  if (S.parent.empty()) return pf;
  pf.assign(S.nparent + 1, (uint32_t)S.n);
  for (uint64_t i = S.n; i-- > 0;) pf[S.parent[i]] = (uint32_t)i;
  for (uint32_t q = S.nparent; q-- > 0;) if (pf[q] > pf[q + 1]) pf[q] = pf[q + 1];   // parents without pieces: empty range
  return pf;
}

// buffer.cuh:502-588 (er_pair): X_A and X_B intersect within the window
static bool er_pair(const Layer& P, const Layer& R, uint32_t a0, uint32_t a1, uint32_t b0, uint32_t b1, double ra_nom, double rb_nom) {
  ESide A{a0, a1, 0, 1e300, 1e300, -1e300, -1e300}, B{b0, b1, 0, 1e300, 1e300, -1e300, -1e300};
  if (A.g0 >= A.g1 || B.g0 >= B.g1) return false;
  auto bbox = [](const Layer& S, ESide& s) {
    for (uint32_t g = s.g0; g < s.g1; g++) {
      s.bx0 = std::fmin(s.bx0, S.ox[g]); s.by0 = std::fmin(s.by0, S.oy[g]);
      s.bx1 = std::fmax(s.bx1, S.ox[g] + S.ex[g]); s.by1 = std::fmax(s.by1, S.oy[g] + S.ey[g]);
    }
  };
  bbox(P, A);
  bbox(R, B);
  // window (world degrees): both bboxes shrunk by a lower bound of the true radius (k <= 1.001 in a UTM zone)
  double dax, day, dbx, dby;
  geo::deg_cover(A.by0, A.by1, ra_nom * 0.998, false, dax, day);
  geo::deg_cover(B.by0, B.by1, rb_nom * 0.998, false, dbx, dby);
  const double wx0 = std::fmax(A.bx0 + dax, B.bx0 + dbx), wx1 = std::fmin(A.bx1 - dax, B.bx1 - dbx);
  const double wy0 = std::fmax(A.by0 + day, B.by0 + dby), wy1 = std::fmin(A.by1 - day, B.by1 - dby);
  if (wx0 > wx1 || wy0 > wy1) return false;
  const double clon = 0.5 * (wx0 + wx1), clat = 0.5 * (wy0 + wy1);
  A.r = ra_nom > 0 ? ra_nom / geo::utm_k(0.5 * (A.bx0 + A.bx1), 0.5 * (A.by0 + A.by1), clon, clat) : 0.0;
  B.r = rb_nom > 0 ? rb_nom / geo::utm_k(0.5 * (B.bx0 + B.bx1), 0.5 * (B.by0 + B.by1), clon, clat) : 0.0;
  EFrame F{P.ox[A.g0], P.oy[A.g0], 0, 0};
  geo::m_per_deg(clat, F.kx, F.ky);
  const double Wx0 = (wx0 - F.ox) * F.kx, Wx1 = (wx1 - F.ox) * F.kx, Wy0 = (wy0 - F.oy) * F.ky, Wy1 = (wy1 - F.oy) * F.ky;
  const double ext = std::fmax(std::fmax(std::fabs(Wx0), std::fabs(Wx1)), std::fmax(std::fabs(Wy0), std::fabs(Wy1))) + std::fmax(A.r, B.r);
  const double eps = 1e-9 + 1e-12 * ext;
  // the frame metric is exact at the window's centre latitude; widen the collect margin for its error
  const double fe = 1.0 + 2.0 * 0.5 * (wy1 - wy0) * geo::DEG * (std::fabs(gsj::gm::tan(clat * geo::DEG)) + 0.05);
  const double ma = A.r * fe + 1e-6, mb = B.r * fe + 1e-6;
  std::vector<BEd> EA, EB;
  std::vector<BEl> LA, LB;
  collect_side(P, A, F, Wx0 - ma, Wy0 - ma, Wx1 + ma, Wy1 + ma, EA);
  collect_side(R, B, F, Wx0 - mb, Wy0 - mb, Wx1 + mb, Wy1 + mb, EB);
  prepare_side(EA, A.r);
  prepare_side(EB, B.r);
  build_elements(EA, A.r, Wx0, Wy0, Wx1, Wy1, eps, LA);
  build_elements(EB, B.r, Wx0, Wy0, Wx1, Wy1, eps, LB);
  auto inW = [&](double x, double y) { return x >= Wx0 - eps && x <= Wx1 + eps && y >= Wy0 - eps && y <= Wy1 + eps; };
  // (a) the eroded boundaries cross outside both exclusion zones
  for (const BEl& p : LA) {
    const double px0 = std::fmin(p.x0, p.x1), px1 = std::fmax(p.x0, p.x1), py0 = std::fmin(p.y0, p.y1), py1 = std::fmax(p.y0, p.y1);
    for (const BEl& q : LB) {
      if (std::fmax(q.x0, q.x1) < px0 || std::fmin(q.x0, q.x1) > px1 || std::fmax(q.y0, q.y1) < py0 || std::fmin(q.y0, q.y1) > py1) continue;
      if (!seg_isect(p.x0, p.y0, p.x1, p.y1, q.x0, q.y0, q.x1, q.y1)) continue;
      double xs[3], ys[3];
      const int k = seg_points(p, q, xs, ys);
      for (int t = 0; t < k; t++)
        if (inW(xs[t], ys[t]) && !in_excl(EA, A.r, xs[t], ys[t], eps) && !in_excl(EB, B.r, xs[t], ys[t], eps)) return true;
    }
  }
  // (b), (c): a vertex of one eroded boundary inside the other eroded polygon
  const double cx = Wx0 + 0.381966 * (Wx1 - Wx0), cy = Wy0 + 0.414213 * (Wy1 - Wy0);
  const bool inA_c = pip_parent(P, A, F, cx, cy), inB_c = pip_parent(R, B, F, cx, cy);
  for (int side = 0; side < 2; side++) {
    const std::vector<BEd>& ES = side ? EB : EA;
    const std::vector<BEd>& EO = side ? EA : EB;
    const std::vector<BEl>& LS = side ? LB : LA;
    const double rs = side ? B.r : A.r, ro = side ? A.r : B.r;
    const bool inO_c = side ? inA_c : inB_c;
    auto vertex_in_other = [&](double x, double y) {
      return inW(x, y) && !in_excl(ES, rs, x, y, eps) && inside_m(EO, cx, cy, inO_c, x, y) && !in_excl(EO, ro, x, y, eps);
    };
    for (size_t i = 0; i < LS.size(); i++) {
      const BEl& p = LS[i];
      if (vertex_in_other(p.x0, p.y0) || vertex_in_other(p.x1, p.y1)) return true;
      const double px0 = std::fmin(p.x0, p.x1), px1 = std::fmax(p.x0, p.x1), py0 = std::fmin(p.y0, p.y1), py1 = std::fmax(p.y0, p.y1);
      for (size_t j = i + 1; j < LS.size(); j++) {
        const BEl& q = LS[j];
        if (std::fmax(q.x0, q.x1) < px0 || std::fmin(q.x0, q.x1) > px1 || std::fmax(q.y0, q.y1) < py0 || std::fmin(q.y0, q.y1) > py1) continue;
        if (!seg_isect(p.x0, p.y0, p.x1, p.y1, q.x0, q.y0, q.x1, q.y1)) continue;
        double xs[3], ys[3];
        const int k = seg_points(p, q, xs, ys);
        for (int t = 0; t < k; t++)
          if (vertex_in_other(xs[t], ys[t])) return true;
      }
    }
  }
  return false;
}

// buffer.cuh:751-776 (k_er_prefilter): a candidate piece pair is kept unless both pieces are whole geometries and their
// bboxes shrunk by (a lower bound of) the radii do not overlap
static bool er_prefilter(const Layer& P, const Layer& R, uint32_t a, uint32_t b, double ra, double rb) {
  auto whole = [](const Layer& S, uint32_t g) {
    if (S.parent.empty()) return true;
    const uint32_t q = S.parent[g];
    return (g == 0 || S.parent[g - 1] != q) && (g + 1 >= S.n || S.parent[g + 1] != q);
  };
  if (!(whole(P, a) && whole(R, b))) return true;
  double dax, day, dbx, dby;
  geo::deg_cover(P.oy[a], P.oy[a] + P.ey[a], ra * 0.998, false, dax, day);
  geo::deg_cover(R.oy[b], R.oy[b] + R.ey[b], rb * 0.998, false, dbx, dby);
  const double wx0 = std::fmax(P.ox[a] + dax, R.ox[b] + dbx), wx1 = std::fmin(P.ox[a] + P.ex[a] - dax, R.ox[b] + R.ex[b] - dbx);
  const double wy0 = std::fmax(P.oy[a] + day, R.oy[b] + dby), wy1 = std::fmin(P.oy[a] + P.ey[a] - day, R.oy[b] + R.ey[b] - dby);
  return wx0 <= wx1 && wy0 <= wy1;
}

// buffer.cuh:722-748 (k_er_filter): intersecting parent pairs whose overlap area can hold a disk of radius rmin
// (rmin 0: any overlap; < 0: every intersecting pair), self pairs only where needed
static bool er_filter(const Layer& P, const std::vector<uint32_t>& pfP, uint32_t a, uint32_t b, double area, double rmin, bool same,
                      const uint8_t* need_self) {
  bool k = true;
  if (same && a == b && need_self) k = need_self[a] != 0;
  if (k && rmin >= 0) {
    const uint32_t g = pfP.empty() ? a : pfP[a];
    const double y0 = P.oy[g], y1 = y0 + P.ey[g];
    const double lat = (y0 <= 0 && y1 >= 0) ? 0.0 : (std::fabs(y0) < std::fabs(y1) ? y0 : y1);
    double kx, ky;
    geo::m_per_deg(lat, kx, ky);
    const double a_m2 = area * kx * ky * 1.001;   // upper bound over the pair's latitudes
    // GEOS's eroded polygon reaches up to ~r/100 (input simplification) + 0.5% (chord fillets) closer to the boundary
    const double r = 0.98 * rmin;
    k = rmin > 0 ? a_m2 >= gsj::gm::PI_HI * r * r : area > 0;
  }
  return k;
}

// buffer.cuh:779-889 (erode_join_stream): pass 1 = the exact intersects + area join over the prefiltered candidates,
// pass 2 = er_pair on the filtered intersecting parent pairs
static void erode(const Layer& P, const Layer& R, double ra, double rb, bool same, bool upper, const uint8_t* need_self, DwResult& res) {
  Params rp;
  rp.want_area = 1;
  JoinResult j;
  std::function<bool(uint32_t, uint32_t)> pre;
  if (ra > 0 || rb > 0) pre = [&](uint32_t a, uint32_t b) { return er_prefilter(P, R, a, b, ra, rb); };
  join(P, R, rp, upper, false, j, nullptr, nullptr, pre);
  res.candidates = j.candidates;
  const std::vector<uint32_t> pfP = piece_first(P), pfR = piece_first(R);
  const double rmin = (ra > 0 && rb > 0) ? std::min(ra, rb) : (ra > 0 || rb > 0 ? 0.0 : -1.0);
  const size_t nrows = j.pairs.size() / 2;
  std::vector<uint8_t> hit(nrows, 0);
#pragma omp parallel for schedule(dynamic, 16)
  for (int64_t i = 0; i < (int64_t)nrows; i++) {
    const uint32_t a = j.pairs[2 * i], b = j.pairs[2 * i + 1];
    if (!er_filter(P, pfP, a, b, j.area[i], rmin, same, need_self)) continue;
    uint32_t a0, a1, b0, b1;
    piece_range(P, pfP, a, a0, a1);
    piece_range(R, pfR, b, b0, b1);
    hit[i] = er_pair(P, R, a0, a1, b0, b1, ra, rb) ? 1 : 0;
  }
  res.pairs.clear(); res.dist.clear();
  for (size_t i = 0; i < nrows; i++)
    if (hit[i]) { res.pairs.push_back(j.pairs[2 * i]); res.pairs.push_back(j.pairs[2 * i + 1]); }
}

// buffer.cuh:1056-1078 (vertex_excl): w (metres from vertex v) strictly inside v's region between unit directions
// e -> f: convex, the disk outside the interior wedge; reflex, the GEOS fillet polygon
static bool vertex_excl(double ex, double ey, double fx, double fy, double r, double wx, double wy, double eps) {
  const double d2 = wx * wx + wy * wy;
  if (d2 >= r * r) return false;
  const double cr = ex * fy - ey * fx, dt = ex * fx + ey * fy;
  if (cr > 0) {
    const double ri = r - eps, s1 = wy * ex - wx * ey, s2 = wy * fx - wx * fy;
    return d2 < ri * ri && !(s1 > eps && s2 > eps);
  }
  if (cr == 0 && dt >= 0) return false;   // straight
  const double th = cr == 0 ? gsj::gm::PI_HI : gsj::gm::atan2(-cr, dt);
  int m = (int)(th / FILLET_QUANTUM + 0.5);
  if (m < 1) m = 1;
  const double inc = th / m;
  double si, ci;
  gsj::gm::sincos(inc, si, ci);
  const double nkx = -ey, nky = ex, nsx = -fy, nsy = fx;
  if (nkx * wy - nky * wx > eps || nsx * wy - nsy * wx < -eps) return false;
  const double chh = std::sqrt(0.5 * (1.0 + ci)), shh = std::sqrt(std::fmax(0.0, 0.5 * (1.0 - ci)));
  const double ch = r * chh - eps;
  if (d2 < ch * ch) return true;
  double mx = nkx * chh + nky * shh, my = -nkx * shh + nky * chh;
  for (int i = 0; i < m; i++) {
    if (mx * wx + my * wy >= ch) return false;
    const double t2 = mx * ci + my * si;
    my = -mx * si + my * ci;
    mx = t2;
  }
  return true;
}

// buffer.cuh:1080-1128 (k_pt_erode): 1 = inside the piece, 2 = within r of its real boundary (excluded), 0 = outside
static uint8_t pt_erode(const Layer& pts, const Layer& R, uint32_t pi, uint32_t g, double r_nom) {
  const double ptx = pts.xw[2 * pi], pty = pts.xw[2 * pi + 1];
  const double r = r_nom / geo::utm_k(R.ox[g] + 0.5 * R.ex[g], R.oy[g] + 0.5 * R.ey[g], ptx, pty);
  double kx, ky;
  geo::m_per_deg(pty, kx, ky);
  const double eps = 1e-9;
  const double shx = R.ox[g] - ptx, shy = R.oy[g] - pty, ox = shx * kx - 1.0;
  auto vx = [&](uint64_t j, double& x, double& y) { x = (shx + R.q[2 * j]) * kx; y = (shy + R.q[2 * j + 1]) * ky; };
  bool par = false, ex = false;
  for (uint64_t rr = R.gro[g]; rr < R.gro[g + 1] && !ex; rr++) {
    const uint64_t s0 = R.rvo[rr], e = R.rvo[rr + 1];
    if (e - s0 < 2) continue;
    // previous real edge direction (region of the vertex that starts the current edge); the ring closes
    bool pv = false;
    double pux = 0, puy = 0;
    {
      const uint64_t j = e - 2;
      if (!R.is_syn(j)) {
        double a, b, c, dd;
        vx(j, a, b); vx(j + 1, c, dd);
        const double L = std::sqrt((c - a) * (c - a) + (dd - b) * (dd - b));
        if (L > 0) { pv = true; pux = (c - a) / L; puy = (dd - b) / L; }
      }
    }
    double x0, y0;
    vx(s0, x0, y0);
    for (uint64_t j = s0; j + 1 < e && !ex; j++) {
      double x1, y1;
      vx(j + 1, x1, y1);
      if (!((y0 > 0) == (y1 > 0) && y0 != 0 && y1 != 0)) {
        const bool so = left_of(x0, y0, x1, y1, ox, 0.0), sp = left_of(x0, y0, x1, y1, 0.0, 0.0);
        if (so != sp) {
          const bool su = orient(ox, 0.0, 0.0, 0.0, x0, y0) > 0, sv = orient(ox, 0.0, 0.0, 0.0, x1, y1) > 0;
          if (su != sv) par = !par;
        }
      }
      const double L = std::sqrt((x1 - x0) * (x1 - x0) + (y1 - y0) * (y1 - y0));
      if (R.is_syn(j) || L == 0) {
        if (R.is_syn(j)) pv = false;
        x0 = x1; y0 = y1;
        continue;
      }
      const double ux = (x1 - x0) / L, uy = (y1 - y0) / L;
      if (r > 0) {
        const double t = -x0 * ux - y0 * uy, s = -y0 * ux + x0 * uy;
        if (std::fabs(s) < r - eps && t > -eps && t < L + eps) ex = true;
        else if (pv && vertex_excl(pux, puy, ux, uy, r, -x0, -y0, eps)) ex = true;
      }
      pv = true; pux = ux; puy = uy;
      x0 = x1; y0 = y1;
    }
  }
  return ex ? 2 : (par ? 1 : 0);
}

// buffer.cuh:1137-1229 (points_buffer_join_stream, erode mode): point boxes widened by 1.002 r, bits OR-ed per
// (point, parent); a row where the point is inside some piece and near no real boundary (bits == 1)
static void points_erode(const Layer& pts, const Layer& R, double r, DwResult& res) {
  const double grow = r * 1.002;
  std::vector<float> qbb(4 * pts.n);
  for (uint64_t i = 0; i < pts.n; i++) {
    const double x = pts.xw[2 * i], y = pts.xw[2 * i + 1];
    double dx = 0, dy = 0;
    geo::deg_cover(y, y, grow, true, dx, dy);
    qbb[4 * i] = geo::f_down(x - dx); qbb[4 * i + 1] = geo::f_down(y - dy); qbb[4 * i + 2] = geo::f_up(x + dx); qbb[4 * i + 3] = geo::f_up(y + dy);
  }
  std::vector<uint32_t> cand;
  candidates(grow > 0 ? qbb : pts.bb, pts.n, R.bb, R.n, false, cand);
  const uint64_t np = cand.size() / 2;
  res.candidates = np;
  std::vector<uint8_t> bits(np);
#pragma omp parallel for schedule(dynamic, 1024)
  for (int64_t i = 0; i < (int64_t)np; i++) bits[i] = pt_erode(pts, R, cand[2 * i], cand[2 * i + 1], r);
  std::vector<std::pair<uint64_t, uint8_t>> kv;
  for (uint64_t i = 0; i < np; i++) {
    const uint64_t pb = R.parent.empty() ? cand[2 * i + 1] : R.parent[cand[2 * i + 1]];
    kv.push_back({((uint64_t)cand[2 * i] << 32) | pb, bits[i]});
  }
  std::sort(kv.begin(), kv.end());
  res.pairs.clear(); res.dist.clear();
  for (size_t s = 0; s < kv.size();) {
    size_t e = s;
    uint8_t b = 0;
    while (e < kv.size() && kv[e].first == kv[s].first) { b |= kv[e].second; e++; }
    if (b == 1) { res.pairs.push_back((uint32_t)(kv[s].first >> 32)); res.pairs.push_back((uint32_t)kv[s].first); }
    s = e;
  }
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
// frame of a window's latitude range: out = {kx, ky, rel, curv}; band of a distance d for two edges in that frame
void orc_geo_frame(double lat_lo, double lat_hi, double* out) {
  const geo::Frame F = geo::frame_of(lat_lo, lat_hi);
  out[0] = F.kx; out[1] = F.ky; out[2] = F.rel; out[3] = F.curv;
}
double orc_geo_band(double lat_lo, double lat_hi, double d, double La2, double Lb2) { return geo::geog_band(d, geo::frame_of(lat_lo, lat_hi), La2, Lb2); }
double orc_geo_seg_seg_d2(double a0x, double a0y, double a1x, double a1y, double b0x, double b0y, double b1x, double b1y) {
  return geo::seg_seg_d2(a0x, a0y, a1x, a1y, b0x, b0y, b1x, b1y);
}
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
int orc_er_pair(const void* P, const void* R, uint32_t a, uint32_t b, double ra, double rb) {
  const Layer& p = *(const Layer*)P; const Layer& r = *(const Layer*)R;
  const std::vector<uint32_t> pfP = piece_first(p), pfR = piece_first(r);
  uint32_t a0, a1, b0, b1;
  piece_range(p, pfP, a, a0, a1);
  piece_range(r, pfR, b, b0, b1);
  return er_pair(p, r, a0, a1, b0, b1, ra, rb) ? 1 : 0;
}
void orc_erode(const void* P, const void* R, double ra, double rb, int same, int upper, const uint8_t* need_self, void* res) {
  erode(*(const Layer*)P, *(const Layer*)R, ra, rb, same != 0, upper != 0, need_self, *(DwResult*)res);
}
void orc_points_erode(const void* pts, const void* R, double r, void* res) { points_erode(*(const Layer*)pts, *(const Layer*)R, r, *(DwResult*)res); }

}  // extern "C"
