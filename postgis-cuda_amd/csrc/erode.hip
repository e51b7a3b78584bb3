// Eroded intersects of polygon layers: ST_Intersects(ST_Buffer(A::geography, -ra)::geometry,
// ST_Buffer(B::geography, -rb)::geometry) without building buffers (reference: engine/buffer.cuh:275-588 er_pair and
// its parts, 590-684 main pass, 717-889 filters and erode_join_stream).
//
// X_A = A minus E_A, E_A = the points within r of A's real boundary: a two-sided strip per edge, a disk per convex
// vertex, GEOS's chord fillet (round(angle / (pi/16)) chords) per reflex vertex. X_A and X_B meet iff their boundaries
// (inward offset edges and fillet chords, the "elements") cross outside E_A and E_B, or an element vertex or element
// self-crossing of one lies inside the other eroded polygon, all inside the window W = shrink(bbox A, ra) and
// shrink(bbox B, rb), in a metric frame at W's centre latitude with r divided by the UTM point scale factor.
//
// Pipeline: pass 1 is the exact intersects + area join (refine.hip) over the candidates that survive a box prefilter;
// its parent rows that could hold a disk of radius min(ra, rb) in their overlap go to pass 2, one wave per parent pair
// (all pieces of both parents at once: a piece's erosion can come from a neighbouring piece's boundary). Lists live
// in LDS (64 edges / 128 elements per side); pairs that need more are re-run with global-memory lists of the size they
// reported, on the same window. The reference splits such windows into tiles; evaluating the whole window instead
// gives exactly the CPU restatement's answer.
#include "pairs.h"
#include "geodesy.h"

namespace gsj {

namespace {

constexpr double FILLET_QUANTUM = gm::PI_HI / 16.0;   // GEOS: 8 quadrant segments
constexpr int ER_CAPE = 64, ER_CAPL = 128;            // LDS list capacities per side (wave pass)
constexpr int ER_CAPE16 = 16, ER_CAPL16 = 48;         // the same for the 16-lane first pass
constexpr int ER_WAVES = 2;                           // waves per block

// real boundary edge in the frame (metres); vt: 0 no region at the end vertex, 1 convex, m + 1 reflex (m chords)
struct BEd { double x0, y0, x1, y1, ux, uy, L, ci, si; int succ, vt; };
struct BEl { double x0, y0, x1, y1; };   // element: piece of the raw offset curve

struct ErScratch { BEd* EA; BEd* EB; BEl* LA; BEl* LB; int capE, capL; };
__host__ __device__ constexpr size_t er_bytes(int capE, int capL) { return 2 * (size_t)capE * sizeof(BEd) + 2 * (size_t)capL * sizeof(BEl); }
__device__ __forceinline__ ErScratch er_carve(unsigned char* base, int capE, int capL) {
  ErScratch s;
  s.EA = reinterpret_cast<BEd*>(base);
  s.EB = s.EA + capE;
  s.LA = reinterpret_cast<BEl*>(s.EB + capE);
  s.LB = s.LA + capL;
  s.capE = capE;
  s.capL = capL;
  return s;
}

// the point is strictly inside (by more than eps) the part of the exclusion region that edge E[k] contributes: its
// two-sided strip, or the disk (convex) or fillet polygon (reflex) at its end vertex; the box test against the edge's
// box widened by r goes first (buffer.cuh:285-320)
__device__ __forceinline__ bool in_zone(const BEd* E, int k, double r, double ri, double x, double y, double eps) {
  const BEd& e = E[k];
  if (x < fmin(e.x0, e.x1) - r || x > fmax(e.x0, e.x1) + r || y < fmin(e.y0, e.y1) - r || y > fmax(e.y0, e.y1) + r) return false;
  const double dx = x - e.x0, dy = y - e.y0;
  const double t = dx * e.ux + dy * e.uy, s = dy * e.ux - dx * e.uy;   // s > 0: interior (left) side
  if (fabs(s) < ri && t > -eps && t < e.L + eps) return true;         // two-sided strip along the edge
  if (e.vt == 0) return false;
  const double vx = x - e.x1, vy = y - e.y1, d2 = vx * vx + vy * vy;
  if (d2 >= r * r) return false;
  const BEd& f = E[e.succ];
  if (e.vt == 1) {   // convex vertex: disk outside the interior wedge
    const double s2 = vy * f.ux - vx * f.uy;
    return d2 < ri * ri && !(s > eps && s2 > eps);
  }
  // reflex vertex: fillet polygon (v, P0 .. Pm) from n_k clockwise to n_s
  const double nkx = -e.uy, nky = e.ux, nsx = -f.uy, nsy = f.ux;
  if (nkx * vy - nky * vx > eps) return false;
  if (nsx * vy - nsy * vx < -eps) return false;
  const double chh = sqrt(0.5 * (1.0 + e.ci)), shh = sqrt(fmax(0.0, 0.5 * (1.0 - e.ci)));   // cos, sin of half the increment
  const double ch = r * chh - eps;                                                         // chord distance from v
  if (d2 < ch * ch) return true;
  double mx = nkx * chh + nky * shh, my = -nkx * shh + nky * chh;   // chord normals
  bool in = true;
  for (int i = 0, m = e.vt - 1; i < m && in; i++) {
    if (mx * vx + my * vy >= ch) in = false;
    const double t2 = mx * e.ci + my * e.si;
    my = -mx * e.si + my * e.ci;
    mx = t2;
  }
  return in;
}

// strictly inside (by more than eps) the region within r of the boundary: an OR of in_zone over the edges
__device__ bool in_excl(const BEd* E, int n, double r, double x, double y, double eps) {
  if (r <= 0) return false;
  const double ri = r - eps;
  // the box test marks 64 edges at a time, then the region tests run over the marked ones (a wave's lanes take those
  // iterations together)
  for (int base = 0; base < n; base += 64) {
    uint64_t cand = 0;
    for (int j = 0, m = min(64, n - base); j < m; j++) {
      const BEd& e = E[base + j];
      if (!(x < fmin(e.x0, e.x1) - r || x > fmax(e.x0, e.x1) + r || y < fmin(e.y0, e.y1) - r || y > fmax(e.y0, e.y1) + r)) cand |= 1ull << j;
    }
    while (cand != 0) {
      const int k = base + __builtin_ctzll(cand);
      cand &= cand - 1;
      if (in_zone(E, k, r, ri, x, y, eps)) return true;
    }
  }
  return false;
}

// whether edge E[k] crosses the segment c -> x (parity step of inside_m)
__device__ __forceinline__ bool crosses_cx(const BEd* E, int k, double cx, double cy, double x, double y, double sx0, double sx1, double sy0,
                                           double sy1) {
  const BEd& e = E[k];
  if (fmax(e.x0, e.x1) < sx0 || fmin(e.x0, e.x1) > sx1 || fmax(e.y0, e.y1) < sy0 || fmin(e.y0, e.y1) > sy1) return false;
  const bool sc = left_of(e.x0, e.y0, e.x1, e.y1, cx, cy), sx = left_of(e.x0, e.y0, e.x1, e.y1, x, y);
  if (sc == sx) return false;
  const bool su = orient(cx, cy, x, y, e.x0, e.y0) > 0, sv = orient(cx, cy, x, y, e.x1, e.y1) > 0;
  return su != sv;
}

// inside(x) = inside(c) xor parity of the real edges crossing c -> x (buffer.cuh:323-335)
__device__ bool inside_m(const BEd* E, int n, double cx, double cy, bool in_c, double x, double y) {
  bool in = in_c;
  const double sx0 = fmin(cx, x), sx1 = fmax(cx, x), sy0 = fmin(cy, y), sy1 = fmax(cy, y);
  for (int k = 0; k < n; k++)
    if (crosses_cx(E, k, cx, cy, x, y, sx0, sx1, sy0, sy1)) in = !in;
  return in;
}

// in_excl and inside_m of one point on a G-lane group, each lane a share of the edges (k = lane, lane + G, ...),
// combined by ballot: the same OR and the same parity as the one-lane loops, in 1/G of their latency
template <int G>
__device__ __forceinline__ bool in_excl_g(const BEd* E, int n, double r, double x, double y, double eps, int lane) {
  if (r <= 0) return false;
  const double ri = r - eps;
  bool hit = false;
  for (int k = lane; k < n && !hit; k += G) hit = in_zone(E, k, r, ri, x, y, eps);
  return gany<G>(hit);
}
template <int G>
__device__ __forceinline__ bool inside_m_g(const BEd* E, int n, double cx, double cy, bool in_c, double x, double y, int lane) {
  const double sx0 = fmin(cx, x), sx1 = fmax(cx, x), sy0 = fmin(cy, y), sy1 = fmax(cy, y);
  bool par = false;
  for (int k = lane; k < n; k += G)
    if (crosses_cx(E, k, cx, cy, x, y, sx0, sx1, sy0, sy1)) par = !par;
  return in_c != ((__popcll(gballot<G>(par)) & 1) != 0);
}

struct ESide { uint32_t g0, g1; double r, bx0, by0, bx1, by1; };
struct EFrame { double ox, oy, kx, ky; };

__device__ __forceinline__ void fv(const DPolys& S, double shx, double shy, uint64_t v, const EFrame& F, double& x, double& y) {
  const double2 q = S.xy[v];
  x = (shx + q.x) * F.kx;
  y = (shy + q.y) * F.ky;
}

// c inside the parent: parity over the pieces whose bbox holds c (buffer.cuh:351-374); a G-lane group
template <int G>
__device__ bool pip_parent(const DPolys& S, const ESide& s, const EFrame& F, double cx, double cy, int lane) {
  bool in = false;
  for (uint32_t g = s.g0; g < s.g1 && !in; g++) {
    const double2 o = S.origin[g], e = S.ext[g];
    const double bx0 = (o.x - F.ox) * F.kx, by0 = (o.y - F.oy) * F.ky, bx1 = (o.x + e.x - F.ox) * F.kx, by1 = (o.y + e.y - F.oy) * F.ky;
    if (cx < bx0 - 1e-6 || cx > bx1 + 1e-6 || cy < by0 - 1e-6 || cy > by1 + 1e-6) continue;
    const double shx = o.x - F.ox, shy = o.y - F.oy, ox = bx0 - 1.0;
    bool par = false;
    // runs whose chunk lies strictly above or below the row y = cy in the frame cross nothing
    auto skip = [&](const float4 c) { return (shy + (double)c.y) * F.ky > cy || (shy + (double)c.w) * F.ky < cy; };
    walk_edge_runs<G>(S, g, lane, skip, [&](uint64_t, uint64_t, uint64_t i0, uint64_t i1) {
      for (uint64_t i = i0 + lane; i < i1; i += G) {
        double x0, y0, x1, y1;
        fv(S, shx, shy, i, F, x0, y0);
        fv(S, shx, shy, i + 1, F, x1, y1);
        if ((y0 > cy) == (y1 > cy) && y0 != cy && y1 != cy) continue;
        const bool so = left_of(x0, y0, x1, y1, ox, cy), sc = left_of(x0, y0, x1, y1, cx, cy);
        if (so == sc) continue;
        const bool su = orient(ox, cy, cx, cy, x0, y0) > 0, sv = orient(ox, cy, cx, cy, x1, y1) > 0;
        if (su != sv) par = !par;
      }
    });
    in = (__popcll(gballot<G>(par)) & 1) != 0;
  }
  return in;
}

// the parent's real, non-degenerate edges meeting box C, ballot-compacted; returns the full count (buffer.cuh:377-405)
template <int G>
__device__ int collect_side(const DPolys& S, const ESide& s, const EFrame& F, double cx0, double cy0, double cx1, double cy1, BEd* E, int cap,
                            int lane) {
  int cnt = 0;
  for (uint32_t g = s.g0; g < s.g1; g++) {
    const double2 o = S.origin[g], ex = S.ext[g];
    if ((o.x + ex.x - F.ox) * F.kx < cx0 || (o.x - F.ox) * F.kx > cx1 || (o.y + ex.y - F.oy) * F.ky < cy0 || (o.y - F.oy) * F.ky > cy1) continue;
    const double shx = o.x - F.ox, shy = o.y - F.oy;
    // edge-chunk runs whose box, in the frame, misses C hold no edge meeting C (the frame map is monotonic)
    auto skip = [&](const float4 c) {
      return (shx + (double)c.z) * F.kx < cx0 || (shx + (double)c.x) * F.kx > cx1 || (shy + (double)c.w) * F.ky < cy0 || (shy + (double)c.y) * F.ky > cy1;
    };
    walk_edge_runs<G>(S, g, lane, skip, [&](uint64_t, uint64_t, uint64_t i0, uint64_t i1) {
      for (uint64_t base = i0; base < i1; base += G) {
        const uint64_t i = base + lane;
        bool v = false;
        BEd ed;
        if (i < i1 && !syn_bit(S.synb, i)) {
          fv(S, shx, shy, i, F, ed.x0, ed.y0);
          fv(S, shx, shy, i + 1, F, ed.x1, ed.y1);
          v = !(ed.x0 == ed.x1 && ed.y0 == ed.y1) &&
              !(fmax(ed.x0, ed.x1) < cx0 || fmin(ed.x0, ed.x1) > cx1 || fmax(ed.y0, ed.y1) < cy0 || fmin(ed.y0, ed.y1) > cy1);
        }
        const uint64_t m = gballot<G>(v);
        const int pos = cnt + __popcll(m & ((1ull << lane) - 1));
        if (v && pos < cap) E[pos] = ed;
        cnt += __popcll(m);
      }
    });
  }
  return cnt;
}

// directions, successors and end-vertex regions (buffer.cuh:408-438)
template <int G>
__device__ void prepare_side(BEd* E, int n, double r, int lane) {
  for (int k = lane; k < n; k += G) {
    const double dx = E[k].x1 - E[k].x0, dy = E[k].y1 - E[k].y0, L = sqrt(dx * dx + dy * dy);
    E[k].ux = dx / L;
    E[k].uy = dy / L;
    E[k].L = L;
  }
  wave_lds_sync();
  // the successor search reads only positions and directions; the fields written below are disjoint from them
  for (int k = lane; k < n; k += G) {
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
        double th = gm::atan2(-cr, dt);
        if (cr == 0) th = gm::PI_HI;
        int m = (int)(th / FILLET_QUANTUM + 0.5);
        if (m < 1) m = 1;
        const double inc = th / m;
        gm::sincos(inc, si, ci);
        vt = m + 1;
      }
    }
    E[k].succ = s;
    E[k].vt = vt;
    E[k].ci = ci;
    E[k].si = si;
  }
  wave_lds_sync();
}

template <int G>
__device__ __forceinline__ int group_max(int v) {
#pragma unroll
  for (int o = G / 2; o > 0; o >>= 1) v = max(v, __shfl_xor(v, o, G));
  return v;
}

// offset edges and reflex fillet chords meeting the window (+- eps), ballot-compacted; returns the full count
// (buffer.cuh:441-475)
template <int G>
__device__ int build_elements(const BEd* E, int n, double r, double wx0, double wy0, double wx1, double wy1, double eps, BEl* L, int cap, int lane) {
  int cnt = 0;
  for (int base = 0; base < n; base += G) {
    const int k = base + lane;
    const bool valid = k < n;
    BEd e{};
    if (valid) e = E[k];
    const int ne = valid ? 1 + (e.vt >= 2 ? e.vt - 1 : 0) : 0;
    const int mx = group_max<G>(ne);
    double nkx = 0, nky = 0, wxr = 0, wyr = 0, px = 0, py = 0, nsx = 0, nsy = 0;
    if (valid) {
      nkx = -e.uy; nky = e.ux; wxr = nkx; wyr = nky;
      px = e.x1 + r * nkx; py = e.y1 + r * nky;
      if (e.vt >= 2) { nsx = -E[e.succ].uy; nsy = E[e.succ].ux; }
    }
    for (int it = 0; it < mx; it++) {
      bool has = it < ne;
      BEl el{};
      if (has) {
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
        has = !(fmax(el.x0, el.x1) < wx0 - eps || fmin(el.x0, el.x1) > wx1 + eps || fmax(el.y0, el.y1) < wy0 - eps || fmin(el.y0, el.y1) > wy1 + eps);
      }
      const uint64_t m = gballot<G>(has);
      const int pos = cnt + __popcll(m & ((1ull << lane) - 1));
      if (has && pos < cap) L[pos] = el;
      cnt += __popcll(m);
    }
  }
  return cnt;
}

// intersection points of two intersecting segments: 1 (proper / touching) or 3 (collinear overlap) (buffer.cuh:478-498)
__device__ __forceinline__ int seg_points(const BEl& p, const BEl& q, double* xs, double* ys) {
  const double rx = p.x1 - p.x0, ry = p.y1 - p.y0, sx = q.x1 - q.x0, sy = q.y1 - q.y0;
  const double den = rx * sy - ry * sx;
  if (den != 0) {
    double t = ((q.x0 - p.x0) * sy - (q.y0 - p.y0) * sx) / den;
    t = fmin(1.0, fmax(0.0, t));
    xs[0] = p.x0 + t * rx; ys[0] = p.y0 + t * ry;
    return 1;
  }
  const double l2 = rx * rx + ry * ry;
  if (l2 == 0) { xs[0] = p.x0; ys[0] = p.y0; return 1; }
  const double t0 = ((q.x0 - p.x0) * rx + (q.y0 - p.y0) * ry) / l2, t1 = ((q.x1 - p.x0) * rx + (q.y1 - p.y0) * ry) / l2;
  double lo = fmax(0.0, fmin(t0, t1)), hi = fmin(1.0, fmax(t0, t1));
  if (lo > hi) { const double m = 0.5 * (lo + hi); lo = hi = fmin(1.0, fmax(0.0, m)); }
  xs[0] = p.x0 + lo * rx; ys[0] = p.y0 + lo * ry;
  xs[1] = p.x0 + hi * rx; ys[1] = p.y0 + hi * ry;
  xs[2] = p.x0 + 0.5 * (lo + hi) * rx; ys[2] = p.y0 + 0.5 * (lo + hi) * ry;
  return 3;
}

__device__ __forceinline__ void piece_range(const uint32_t* pf, uint32_t p, uint32_t& g0, uint32_t& g1) {
  if (pf) { g0 = pf[p]; g1 = pf[p + 1]; }
  else { g0 = p; g1 = p + 1; }
}

// A parent pair made ready for the tests: the window and frame, both sides' edge and element lists in the scratch.
struct ErPair {
  ESide A, B;
  EFrame F;
  int na, nb, la, lb;
  double Wx0, Wy0, Wx1, Wy1, eps, cx, cy;
};

// Window, frame, edge lists and element lists of a parent pair (buffer.cuh:502-541); a G-lane group. Returns 1 when ready,
// 0 when the window is empty (no intersection), -1 when a list overflows (needE / needL: sizes required; needL is 0
// when the edge lists already overflowed).
template <int G>
__device__ int er_setup(const DPolys& P, const DPolys& R, const uint32_t* pfP, const uint32_t* pfR, uint32_t pa, uint32_t pb, double ra_nom,
                        double rb_nom, const ErScratch& sc, int lane, ErPair& E, int& needE, int& needL) {
  ESide& A = E.A;
  ESide& B = E.B;
  A = ESide{0, 0, 0, 1e300, 1e300, -1e300, -1e300};
  B = ESide{0, 0, 0, 1e300, 1e300, -1e300, -1e300};
  piece_range(pfP, pa, A.g0, A.g1);
  piece_range(pfR, pb, B.g0, B.g1);
  if (A.g0 >= A.g1 || B.g0 >= B.g1) return 0;
  for (uint32_t g = A.g0; g < A.g1; g++) {
    const double2 o = P.origin[g], e = P.ext[g];
    A.bx0 = fmin(A.bx0, o.x); A.by0 = fmin(A.by0, o.y); A.bx1 = fmax(A.bx1, o.x + e.x); A.by1 = fmax(A.by1, o.y + e.y);
  }
  for (uint32_t g = B.g0; g < B.g1; g++) {
    const double2 o = R.origin[g], e = R.ext[g];
    B.bx0 = fmin(B.bx0, o.x); B.by0 = fmin(B.by0, o.y); B.bx1 = fmax(B.bx1, o.x + e.x); B.by1 = fmax(B.by1, o.y + e.y);
  }
  // window (world degrees): both bboxes shrunk by a lower bound of the true radius (k <= 1.001 in a UTM zone)
  double dax, day, dbx, dby;
  geo::deg_cover(A.by0, A.by1, ra_nom * 0.998, false, dax, day);
  geo::deg_cover(B.by0, B.by1, rb_nom * 0.998, false, dbx, dby);
  const double wx0 = fmax(A.bx0 + dax, B.bx0 + dbx), wx1 = fmin(A.bx1 - dax, B.bx1 - dbx);
  const double wy0 = fmax(A.by0 + day, B.by0 + dby), wy1 = fmin(A.by1 - day, B.by1 - dby);
  if (wx0 > wx1 || wy0 > wy1) return 0;
  const double clon = 0.5 * (wx0 + wx1), clat = 0.5 * (wy0 + wy1);
  A.r = ra_nom > 0 ? ra_nom / geo::utm_k(0.5 * (A.bx0 + A.bx1), 0.5 * (A.by0 + A.by1), clon, clat) : 0.0;
  B.r = rb_nom > 0 ? rb_nom / geo::utm_k(0.5 * (B.bx0 + B.bx1), 0.5 * (B.by0 + B.by1), clon, clat) : 0.0;
  const double2 fo = P.origin[A.g0];
  EFrame& F = E.F;
  F = EFrame{fo.x, fo.y, 0, 0};
  geo::m_per_deg(clat, F.kx, F.ky);
  E.Wx0 = (wx0 - F.ox) * F.kx; E.Wx1 = (wx1 - F.ox) * F.kx; E.Wy0 = (wy0 - F.oy) * F.ky; E.Wy1 = (wy1 - F.oy) * F.ky;
  const double ext = fmax(fmax(fabs(E.Wx0), fabs(E.Wx1)), fmax(fabs(E.Wy0), fabs(E.Wy1))) + fmax(A.r, B.r);
  E.eps = 1e-9 + 1e-12 * ext;
  // the frame metric is exact at the window's centre latitude; widen the collect margin for its error
  const double fe = 1.0 + 2.0 * 0.5 * (wy1 - wy0) * geo::DEG * (fabs(gm::tan(clat * geo::DEG)) + 0.05);
  const double ma = A.r * fe + 1e-6, mb = B.r * fe + 1e-6;
  E.na = collect_side<G>(P, A, F, E.Wx0 - ma, E.Wy0 - ma, E.Wx1 + ma, E.Wy1 + ma, sc.EA, sc.capE, lane);
  E.nb = collect_side<G>(R, B, F, E.Wx0 - mb, E.Wy0 - mb, E.Wx1 + mb, E.Wy1 + mb, sc.EB, sc.capE, lane);
  if (E.na > sc.capE || E.nb > sc.capE) { needE = max(E.na, E.nb); needL = 0; return -1; }
  wave_lds_sync();
  prepare_side<G>(sc.EA, E.na, A.r, lane);
  prepare_side<G>(sc.EB, E.nb, B.r, lane);
  E.la = build_elements<G>(sc.EA, E.na, A.r, E.Wx0, E.Wy0, E.Wx1, E.Wy1, E.eps, sc.LA, sc.capL, lane);
  E.lb = build_elements<G>(sc.EB, E.nb, B.r, E.Wx0, E.Wy0, E.Wx1, E.Wy1, E.eps, sc.LB, sc.capL, lane);
  if (E.la > sc.capL || E.lb > sc.capL) { needE = max(E.na, E.nb); needL = max(E.la, E.lb); return -1; }
  E.cx = E.Wx0 + 0.381966 * (E.Wx1 - E.Wx0);
  E.cy = E.Wy0 + 0.414213 * (E.Wy1 - E.Wy0);
  wave_lds_sync();
  return 1;
}

// elements L[base .. min(base + 64, n)) whose boxes meet [x0, x1] x [y0, y1], one bit each: the crossing tests below
// are ORs over element pairs, so they run over these bits, the lanes of a wave together
__device__ __forceinline__ uint64_t box_mask(const BEl* L, int base, int n, double x0, double x1, double y0, double y1) {
  uint64_t cand = 0;
  for (int j = 0, m = min(64, n - base); j < m; j++) {
    const BEl q = L[base + j];
    if (!(fmax(q.x0, q.x1) < x0 || fmin(q.x0, q.x1) > x1 || fmax(q.y0, q.y1) < y0 || fmin(q.y0, q.y1) > y1)) cand |= 1ull << j;
  }
  return cand;
}

// (a) the eroded boundaries cross outside both exclusion zones: this thread's share (element i of A, i = t, t + NT, ...)
__device__ bool er_cross(const ErScratch& sc, const ErPair& E, int t, int NT, const volatile int* stop) {
  auto inW = [&](double x, double y) { return x >= E.Wx0 - E.eps && x <= E.Wx1 + E.eps && y >= E.Wy0 - E.eps && y <= E.Wy1 + E.eps; };
  for (int i = t; i < E.la; i += NT) {
    if (stop && *stop) return false;
    const BEl p = sc.LA[i];
    const double px0 = fmin(p.x0, p.x1), px1 = fmax(p.x0, p.x1), py0 = fmin(p.y0, p.y1), py1 = fmax(p.y0, p.y1);
    for (int base = 0; base < E.lb; base += 64) {
      uint64_t cand = box_mask(sc.LB, base, E.lb, px0, px1, py0, py1);
      while (cand != 0) {
        const BEl q = sc.LB[base + __builtin_ctzll(cand)];
        cand &= cand - 1;
        if (!seg_isect(p.x0, p.y0, p.x1, p.y1, q.x0, q.y0, q.x1, q.y1)) continue;
        double xs[3], ys[3];
        const int k = seg_points(p, q, xs, ys);
        for (int u = 0; u < k; u++)
          if (inW(xs[u], ys[u]) && !in_excl(sc.EA, E.na, E.A.r, xs[u], ys[u], E.eps) && !in_excl(sc.EB, E.nb, E.B.r, xs[u], ys[u], E.eps)) return true;
      }
    }
  }
  return false;
}

// (b), (c) a vertex or self-crossing of one side's eroded boundary inside the other eroded polygon: this thread's share
__device__ bool er_inside(const ErScratch& sc, const ErPair& E, bool inA_c, bool inB_c, int side, int t, int NT, const volatile int* stop) {
  auto inW = [&](double x, double y) { return x >= E.Wx0 - E.eps && x <= E.Wx1 + E.eps && y >= E.Wy0 - E.eps && y <= E.Wy1 + E.eps; };
  const BEd* ES = side ? sc.EB : sc.EA;
  const BEd* EO = side ? sc.EA : sc.EB;
  const BEl* LS = side ? sc.LB : sc.LA;
  const int ns = side ? E.nb : E.na, no = side ? E.na : E.nb, ls = side ? E.lb : E.la;
  const double rs = side ? E.B.r : E.A.r, ro = side ? E.A.r : E.B.r;
  const bool inO_c = side ? inA_c : inB_c;
  auto vertex_in_other = [&](double x, double y) {
    return inW(x, y) && !in_excl(ES, ns, rs, x, y, E.eps) && inside_m(EO, no, E.cx, E.cy, inO_c, x, y) && !in_excl(EO, no, ro, x, y, E.eps);
  };
  for (int i = t; i < ls; i += NT) {
    if (stop && *stop) return false;
    const BEl p = LS[i];
    if (vertex_in_other(p.x0, p.y0) || vertex_in_other(p.x1, p.y1)) return true;
    const double px0 = fmin(p.x0, p.x1), px1 = fmax(p.x0, p.x1), py0 = fmin(p.y0, p.y1), py1 = fmax(p.y0, p.y1);
    for (int base = i + 1; base < ls; base += 64) {
      uint64_t cand = box_mask(LS, base, ls, px0, px1, py0, py1);
      while (cand != 0) {
        const BEl q = LS[base + __builtin_ctzll(cand)];
        cand &= cand - 1;
        if (!seg_isect(p.x0, p.y0, p.x1, p.y1, q.x0, q.y0, q.x1, q.y1)) continue;
        double xs[3], ys[3];
        const int k = seg_points(p, q, xs, ys);
        for (int u = 0; u < k; u++)
          if (vertex_in_other(xs[u], ys[u])) return true;
      }
    }
  }
  return false;
}

// The vertex part of (b) / (c) for one side: an element end of that side's eroded boundary inside the other eroded
// polygon (a subset of er_inside's tests, so a hit here is a hit of the pair). Each element end is checked by a whole
// G-lane group, the lanes sharing its edges (in_excl_g, inside_m_g): group gi of ng takes elements gi, gi + ng, ... in
// order. A parent's self pair is settled by its first qualifying end, so the group's latency per end, not one lane's
// walk over all edges per end, sets its time (windowed erosion on the national layer: 648 -> 362 ms for 209
// requests). stop (another group's success) is polled between elements, group-uniformly.
template <int G>
__device__ bool er_vertices_g(const ErScratch& sc, const ErPair& E, bool inA_c, bool inB_c, int side, int gi, int ng, int lane,
                              const volatile int* stop) {
  const BEd* ES = side ? sc.EB : sc.EA;
  const BEd* EO = side ? sc.EA : sc.EB;
  const BEl* LS = side ? sc.LB : sc.LA;
  const int ns = side ? E.nb : E.na, no = side ? E.na : E.nb, ls = side ? E.lb : E.la;
  const double rs = side ? E.B.r : E.A.r, ro = side ? E.A.r : E.B.r;
  const bool inO_c = side ? inA_c : inB_c;
  for (int i = gi; i < ls; i += ng) {
    if (stop && gany<G>(*stop != 0)) return false;
    const BEl p = LS[i];
    for (int end = 0; end < 2; end++) {
      const double x = end ? p.x1 : p.x0, y = end ? p.y1 : p.y0;
      if (!(x >= E.Wx0 - E.eps && x <= E.Wx1 + E.eps && y >= E.Wy0 - E.eps && y <= E.Wy1 + E.eps)) continue;
      if (in_excl_g<G>(ES, ns, rs, x, y, E.eps, lane)) continue;
      if (!inside_m_g<G>(EO, no, E.cx, E.cy, inO_c, x, y, lane)) continue;
      if (!in_excl_g<G>(EO, no, ro, x, y, E.eps, lane)) return true;
    }
  }
  return false;
}

// A parent paired with itself: X_A(ra) and X_A(rb) meet wherever the more eroded one is non-empty, so an element end of
// the larger radius's boundary inside the other usually settles it at once; only when none is found do the full tests
// run (their answer is an OR over all tests, so taking this subset first changes no result).
__device__ __forceinline__ int self_side(const ErPair& E) { return E.B.r > E.A.r ? 1 : 0; }

// X_A and X_B intersect within the window (buffer.cuh:502-588). A G-lane group; returns 1 / 0, or -1 when a list
// overflows (needE / needL: sizes required; needL is 0 when the edge lists already overflowed).
template <int G>
__device__ int er_pair(const DPolys& P, const DPolys& R, const uint32_t* pfP, const uint32_t* pfR, uint32_t pa, uint32_t pb, double ra_nom,
                       double rb_nom, const ErScratch& sc, int lane, int& needE, int& needL) {
  ErPair E;
  const int st = er_setup<G>(P, R, pfP, pfR, pa, pb, ra_nom, rb_nom, sc, lane, E, needE, needL);
  if (st <= 0) return st;
  const bool self = pa == pb && P.xy == R.xy;
  bool inA_c = false, inB_c = false, have_c = false;
  if (self) {   // one parent, one point c: both containments are the same walk
    inA_c = pip_parent<G>(P, E.A, E.F, E.cx, E.cy, lane);
    inB_c = inA_c;
    have_c = true;
    if (er_vertices_g<G>(sc, E, inA_c, inB_c, self_side(E), 0, 1, lane, nullptr)) return 1;
  }
  if (gany<G>(er_cross(sc, E, lane, G, nullptr))) return 1;
  if (!have_c) {
    inA_c = pip_parent<G>(P, E.A, E.F, E.cx, E.cy, lane);
    inB_c = pip_parent<G>(R, E.B, E.F, E.cx, E.cy, lane);
  }
  for (int side = 0; side < 2; side++)
    if (gany<G>(er_inside(sc, E, inA_c, inB_c, side, lane, G, nullptr))) return 1;
  return 0;
}

struct ErCtl { uint32_t n; int needE, needL; uint32_t pad; };

// Radii of the pair (or candidate, or row) at index p: one request's (seg null), or, for batched requests, those of
// request r with seg[r] <= p < seg[r + 1]
struct ErRadii {
  const uint64_t* seg;
  uint32_t k;
  double ra[MAX_WINDOWS], rb[MAX_WINDOWS];
};
__device__ __forceinline__ uint32_t radii_of(const ErRadii& R, uint64_t p, double& ra, double& rb) {
  uint32_t r = 0;
  if (R.seg)
    while (r + 1 < R.k && R.seg[r + 1] <= p) r++;
  ra = R.ra[r];
  rb = R.rb[r];
  return r;
}

// LDS passes: a G-lane group per parent pair with CAPE / CAPL lists per side; overflowing pairs are queued with the
// sizes they need. The pairs are 0 .. np (items null) or items[0 .. items_ctl->n) (the previous pass's queue).
template <int G, int CAPE, int CAPL>
__global__ void __launch_bounds__(ER_WAVES * 64) k_erode_main(const DPolys P, const DPolys R, const uint32_t* __restrict__ pfP,
                                                              const uint32_t* __restrict__ pfR, const uint32_t* __restrict__ pairs, uint32_t np,
                                                              const uint32_t* __restrict__ items, const ErCtl* __restrict__ items_ctl, const ErRadii rr,
                                                              uint8_t* __restrict__ flag, uint32_t* __restrict__ heavy, ErCtl* ctl) {
  constexpr int NG = ER_WAVES * 64 / G;
  __shared__ __align__(16) unsigned char lds[NG][er_bytes(CAPE, CAPL)];
  const int lane = threadIdx.x % G, gib = threadIdx.x / G;
  const ErScratch sc = er_carve(lds[gib], CAPE, CAPL);
  const uint32_t n = items ? items_ctl->n : np;
  for (uint32_t k = blockIdx.x * NG + gib; k < n; k += gridDim.x * NG) {
    const uint32_t p = items ? items[k] : k;
    int ne = 0, nl = 0;
    double ra, rb;
    radii_of(rr, p, ra, rb);
    const uint32_t a = pairs[2 * p], b = pairs[2 * p + 1];
    // a parent of an unsubdivided layer paired with itself: its window holds nearly all its edges, so a piece of more
    // edges than CAPE overflows this pass's lists; it goes on to the next pass untried (every pass decides a pair the
    // same way; needE = its edge count bounds what it collects)
    int r;
    const uint32_t gi = !pfP && a == b && P.xy == R.xy ? P.ginfo[a] : 0u;
    const int edges = (int)(gi & 0xFFFFu) - (int)(gi >> 16);
    if (edges > CAPE) { r = -1; ne = edges; nl = 0; }
    else r = er_pair<G>(P, R, pfP, pfR, a, b, ra, rb, sc, lane, ne, nl);
    if (lane == 0) {
      flag[p] = r > 0 ? 1 : 0;
      if (r < 0) {
        heavy[atomicAdd(&ctl->n, 1u)] = p;
        atomicMax(&ctl->needE, ne);
        atomicMax(&ctl->needL, nl);
      }
    }
    wave_lds_sync();
  }
}

// pairs that overflowed: a block of T threads per pair with lists in global memory of the size the queue reported;
// its first wave builds the lists, then the whole block runs the crossing and containment tests; pairs that overflow
// the element lists again are re-queued with their exact element count. T is chosen per launch from the number of
// pairs (slab_threads): a few pairs take 512 threads each, many pairs one wave each, so that the other waves of a
// block do not wait through the list building while the queue still holds pairs for them.
template <int T>
__global__ void __launch_bounds__(T) k_erode_slab(const DPolys P, const DPolys R, const uint32_t* __restrict__ pfP,
                                                          const uint32_t* __restrict__ pfR, const uint32_t* __restrict__ pairs,
                                                          const uint32_t* __restrict__ items, uint32_t ni, const ErRadii rr,
                                                          uint8_t* __restrict__ flag, unsigned char* __restrict__ slab, int capE, int capL,
                                                          uint32_t* __restrict__ again, ErCtl* ctl) {
  __shared__ ErPair SE;
  __shared__ int s_state, s_hit, s_inA, s_inB;
  const int t = threadIdx.x, lane = t & 63;
  const ErScratch sc = er_carve(slab + (uint64_t)blockIdx.x * er_bytes(capE, capL), capE, capL);
  for (uint32_t i = blockIdx.x; i < ni; i += gridDim.x) {
    const uint32_t p = items[i];
    if (t < 64) {   // first wave: window, frame and lists
      ErPair E;
      int ne = 0, nl = 0;
      double ra, rb;
      radii_of(rr, p, ra, rb);
      const int st = er_setup<64>(P, R, pfP, pfR, pairs[2 * p], pairs[2 * p + 1], ra, rb, sc, lane, E, ne, nl);
      if (lane == 0) {
        SE = E;
        s_state = st;
        s_hit = 0;
        if (st < 0) {
          flag[p] = 0;
          again[atomicAdd(&ctl->n, 1u)] = p;
          atomicMax(&ctl->needE, ne);
          atomicMax(&ctl->needL, nl);
        } else if (st == 0) {
          flag[p] = 0;
        }
      }
    }
    __syncthreads();
    if (s_state > 0) {
      const ErPair E = SE;
      const bool self = pairs[2 * p] == pairs[2 * p + 1] && P.xy == R.xy;
      if (t < 64) {
        const bool a = pip_parent<64>(P, E.A, E.F, E.cx, E.cy, lane);
        const bool b = self ? a : pip_parent<64>(R, E.B, E.F, E.cx, E.cy, lane);   // one parent, one point c
        if (lane == 0) { s_inA = a; s_inB = b; }
      }
      __syncthreads();
      if (self && er_vertices_g<64>(sc, E, s_inA != 0, s_inB != 0, self_side(E), t / 64, T / 64, lane, &s_hit)) s_hit = 1;
      __syncthreads();
      if (!s_hit && er_cross(sc, E, t, T, &s_hit)) s_hit = 1;
      __syncthreads();
      if (!s_hit) {
        for (int side = 0; side < 2; side++) {
          if (!s_hit && er_inside(sc, E, s_inA != 0, s_inB != 0, side, t, T, &s_hit)) s_hit = 1;
          __syncthreads();
        }
      }
      if (t == 0) flag[p] = s_hit ? 1 : 0;
    }
    __syncthreads();   // the lists and SE are rewritten for the next pair
  }
}

// threads per overflowing pair: one wave each once the pairs fill every CU's waves (8 per CU at these kernels'
// register counts), two below that, a block of 512 for a few pairs
int slab_threads(uint64_t nh, int cu) {
  if (nh >= 8ull * cu) return 64;
  if (nh >= 2ull * cu) return 128;
  return 512;
}

__global__ void k_ctl_clear(ErCtl* ctl) { ctl->n = 0; ctl->needE = 0; ctl->needL = 0; }

// candidate piece pairs whose eroded polygons can meet: not both whole geometries, or shrunk bboxes that overlap
// (buffer.cuh:751-776)
__global__ void k_er_prefilter(const DPolys P, const DPolys R, const uint32_t* __restrict__ pa, const uint32_t* __restrict__ pb, uint64_t n,
                               const ErRadii rr, uint32_t* __restrict__ keep) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double ra, rb;
  radii_of(rr, i, ra, rb);
  if (ra <= 0 && rb <= 0) { keep[i] = 1u; return; }   // no erosion: every candidate is refined
  const uint32_t a = pa[i], b = pb[i];
  auto whole = [](const DPolys& S, uint32_t g) {
    if (!S.parent) return true;
    const uint32_t q = S.parent[g];
    return (g == 0 || S.parent[g - 1] != q) && (g + 1 >= S.n || S.parent[g + 1] != q);
  };
  bool k = true;
  if (whole(P, a) && whole(R, b)) {
    const double2 oa = P.origin[a], ea = P.ext[a], ob = R.origin[b], eb = R.ext[b];
    double dax, day, dbx, dby;
    geo::deg_cover(oa.y, oa.y + ea.y, ra * 0.998, false, dax, day);
    geo::deg_cover(ob.y, ob.y + eb.y, rb * 0.998, false, dbx, dby);
    const double wx0 = fmax(oa.x + dax, ob.x + dbx), wx1 = fmin(oa.x + ea.x - dax, ob.x + eb.x - dbx);
    const double wy0 = fmax(oa.y + day, ob.y + dby), wy1 = fmin(oa.y + ea.y - day, ob.y + eb.y - dby);
    k = wx0 <= wx1 && wy0 <= wy1;
  }
  keep[i] = k ? 1u : 0u;
}
__global__ void k_compact_pairs(const uint32_t* __restrict__ pa, const uint32_t* __restrict__ pb, const uint32_t* __restrict__ keep,
                                const uint64_t* __restrict__ off, uint64_t n, uint32_t* __restrict__ oa, uint32_t* __restrict__ ob) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n || !keep[i]) return;
  oa[off[i]] = pa[i];
  ob[off[i]] = pb[i];
}
__global__ void k_copy_u32(const uint32_t* __restrict__ a, uint64_t n, uint32_t* __restrict__ b) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) b[i] = a[i];
}

// whether parent a is in the ascending list
__device__ __forceinline__ bool in_sorted(const uint32_t* __restrict__ list, uint32_t n, uint32_t a) {
  uint32_t lo = 0, hi = n;
  while (lo < hi) {
    const uint32_t mid = (lo + hi) / 2;
    if (list[mid] < a) lo = mid + 1;
    else hi = mid;
  }
  return lo < n && list[lo] == a;
}

// parent rows of pass 1 that go to pass 2: overlap area able to hold a disk of radius rmin (rmin 0: any overlap,
// < 0: every intersecting pair), self pairs only where needed (buffer.cuh:722-748): by the flags, or by the list
__global__ void k_er_filter(const uint32_t* __restrict__ rows, const double* __restrict__ area, uint64_t n, const DPolys P,
                            const uint32_t* __restrict__ pfP, const ErRadii rr, int same, const SelfNeed need, uint32_t* __restrict__ keep) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double ra, rb;
  radii_of(rr, i, ra, rb);
  const double rmin = (ra > 0 && rb > 0) ? fmin(ra, rb) : (ra > 0 || rb > 0 ? 0.0 : -1.0);
  const uint32_t a = rows[2 * i], b = rows[2 * i + 1];
  bool k = true;
  if (same && a == b) {
    if (need.listed) k = in_sorted(need.list, need.n, a);
    else if (need.flags) k = need.flags[a] != 0;
  }
  if (k && rmin >= 0) {
    const uint32_t g = pfP ? pfP[a] : a;
    const double y0 = P.origin[g].y, y1 = y0 + P.ext[g].y;
    const double lat = (y0 <= 0 && y1 >= 0) ? 0.0 : (fabs(y0) < fabs(y1) ? y0 : y1);
    double kx, ky;
    geo::m_per_deg(lat, kx, ky);
    const double a_m2 = area[i] * kx * ky * 1.001;   // upper bound over the pair's latitudes
    const double r = 0.98 * rmin;                    // GEOS's eroded polygon reaches up to ~2% closer to the boundary
    k = rmin > 0 ? a_m2 >= gm::PI_HI * r * r : area[i] > 0;
  }
  keep[i] = k ? 1u : 0u;
}
__global__ void k_compact_rows(const uint32_t* __restrict__ rows, const uint32_t* __restrict__ keep, const uint64_t* __restrict__ off, uint64_t n,
                               uint32_t* __restrict__ out) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n || !keep[i]) return;
  out[2 * off[i]] = rows[2 * i];
  out[2 * off[i] + 1] = rows[2 * i + 1];
}
__global__ void k_flag_u32(const uint8_t* __restrict__ f, uint64_t n, uint32_t* __restrict__ o) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) o[i] = f[i] ? 1u : 0u;
}

}  // namespace

namespace {

// Pass 2 over n pass-1 parent rows (rows, overlap areas) on the device: the rows whose eroded polygons meet,
// compacted into *out in pass-1 order; returns their count. Batched requests: lo_in[0 .. k] are the rows' request
// boundaries (rr.seg == lo_in), and lo_mid, lo_out receive those of the filtered pairs and of the output.
uint64_t erode_pass2(Context& c, const DPolys& P, const DPolys& R, const uint32_t* pfP, const uint32_t* pfR, const uint32_t* rows, const double* area,
                     uint64_t n, ErRadii rr, bool same, const SelfNeed& d_need, ErCtl* ctl, gsj_join_stats& s, uint32_t*& out, uint64_t* lo_mid,
                     uint64_t* lo_out) {
  hipStream_t stream = c.stream;
  uint32_t* keep = c.s_er_keep.as<uint32_t>(n + 16);
  uint64_t* off = c.s_er_off.as<uint64_t>(n + 16);
  hipLaunchKernelGGL(k_er_filter, dim3(div_up(n, 256)), dim3(256), 0, stream, rows, area, n, P, pfP, rr, same ? 1 : 0, d_need, keep);
  GSJ_LAUNCHED();
  const uint64_t ne = scan_u32_to_u64(c, keep, off, n);
  if (rr.seg) {
    seg_bounds_dev(c, rr.seg, rr.k, off, n, ne, lo_mid);
    rr.seg = lo_mid;
  }
  if (ne == 0) {
    if (lo_out) seg_bounds_dev(c, lo_mid, rr.k, off, 0, 0, lo_out);
    return 0;
  }
  uint32_t* ep = c.s_er_pairs.as<uint32_t>(2 * ne + 16);
  hipLaunchKernelGGL(k_compact_rows, dim3(div_up(n, 256)), dim3(256), 0, stream, rows, keep, off, n, ep);
  GSJ_LAUNCHED();
  uint8_t* flag = c.s_er_flag.as<uint8_t>(ne + 16);
  uint32_t* lists = c.s_er_list.as<uint32_t>(3 * ne + 48);
  uint32_t* heavy = lists;
  uint32_t* again = lists + ne + 16;
  uint32_t* first = again + ne + 16;   // the 16-lane pass's overflow, for the wave pass
  ErCtl* ctl16 = ctl + 1;
  hipLaunchKernelGGL(k_ctl_clear, dim3(1), dim3(1), 0, stream, ctl);
  hipLaunchKernelGGL(k_ctl_clear, dim3(1), dim3(1), 0, stream, ctl16);
  GSJ_HIP(hipEventRecord(c.ev[0], stream));
  // 16-lane groups with small lists first (most parent pairs), then a wave per pair for those that overflowed them
  const uint32_t g16 = (uint32_t)std::min<uint64_t>(div_up(ne, ER_WAVES * 4), (uint64_t)c.cu_count * 32);
  hipLaunchKernelGGL((k_erode_main<16, ER_CAPE16, ER_CAPL16>), dim3(g16), dim3(ER_WAVES * 64), 0, stream, P, R, pfP, pfR, ep, (uint32_t)ne, nullptr,
                     nullptr, rr, flag, first, ctl16);
  GSJ_LAUNCHED();
  const uint32_t g = (uint32_t)std::min<uint64_t>(div_up(ne, ER_WAVES), (uint64_t)c.cu_count * 32);
  hipLaunchKernelGGL((k_erode_main<64, ER_CAPE, ER_CAPL>), dim3(g), dim3(ER_WAVES * 64), 0, stream, P, R, pfP, pfR, ep, (uint32_t)ne, first, ctl16, rr,
                     flag, heavy, ctl);
  GSJ_LAUNCHED();
  GSJ_HIP(hipEventRecord(c.ev[1], stream));
  ErCtl* h = c.h_scal.as<ErCtl>(2);
  GSJ_HIP(hipMemcpyAsync(h, ctl, sizeof(ErCtl), hipMemcpyDeviceToHost, stream));
  GSJ_HIP(hipStreamSynchronize(stream));
  float ms = 0;
  GSJ_HIP(hipEventElapsedTime(&ms, c.ev[0], c.ev[1]));
  s.k_large_ms += ms;
  s.n_large += ne;
  s.launches++;
  uint32_t nh = h->n;
  int capE = std::max(h->needE, ER_CAPE), capL = std::max(h->needL, 4 * capE);
  s.heavy += nh;
  s.n_over += nh;
  for (int round = 0; nh; round++) {
    GSJ_REQUIRE(round < 4, GSJ_E_INTERNAL, "erosion: overflowing pairs did not converge");
    const size_t per = er_bytes(capE, capL);
    const uint64_t blocks = std::max<uint64_t>(1, std::min<uint64_t>(nh, (1ull << 30) / per));   // one list set per block
    unsigned char* slab = c.s_er_slab.as<unsigned char>(per * blocks + 64);
    hipLaunchKernelGGL(k_ctl_clear, dim3(1), dim3(1), 0, stream, ctl);
    GSJ_HIP(hipEventRecord(c.ev[2], stream));
    const int T = slab_threads(nh, c.cu_count);
    if (T == 64)
      hipLaunchKernelGGL(k_erode_slab<64>, dim3((uint32_t)blocks), dim3(64), 0, stream, P, R, pfP, pfR, ep, heavy, nh, rr, flag, slab, capE, capL, again, ctl);
    else if (T == 128)
      hipLaunchKernelGGL(k_erode_slab<128>, dim3((uint32_t)blocks), dim3(128), 0, stream, P, R, pfP, pfR, ep, heavy, nh, rr, flag, slab, capE, capL, again,
                         ctl);
    else
      hipLaunchKernelGGL(k_erode_slab<512>, dim3((uint32_t)blocks), dim3(512), 0, stream, P, R, pfP, pfR, ep, heavy, nh, rr, flag, slab, capE, capL, again,
                         ctl);
    GSJ_LAUNCHED();
    GSJ_HIP(hipEventRecord(c.ev[3], stream));
    GSJ_HIP(hipMemcpyAsync(h, ctl, sizeof(ErCtl), hipMemcpyDeviceToHost, stream));
    GSJ_HIP(hipStreamSynchronize(stream));
    float mo = 0;
    GSJ_HIP(hipEventElapsedTime(&mo, c.ev[2], c.ev[3]));
    s.k_over_ms += mo;
    nh = h->n;
    capE = std::max(capE, h->needE);
    capL = std::max(capL, h->needL);
    std::swap(heavy, again);
  }
  // rows: flagged parent pairs in pass-1 order
  uint32_t* f = c.s_er_keep.as<uint32_t>(ne + 16);
  hipLaunchKernelGGL(k_flag_u32, dim3(div_up(ne, 256)), dim3(256), 0, stream, flag, ne, f);
  GSJ_LAUNCHED();
  const uint64_t nhit = scan_u32_to_u64(c, f, off, ne);
  if (lo_out) seg_bounds_dev(c, lo_mid, rr.k, off, ne, nhit, lo_out);
  if (nhit == 0) return 0;
  out = c.s_er_out.as<uint32_t>(2 * nhit + 16);
  hipLaunchKernelGGL(k_compact_rows, dim3(div_up(ne, 256)), dim3(256), 0, stream, ep, f, off, ne, out);
  GSJ_LAUNCHED();
  return nhit;
}

// the self-pair rule on the device (the flags, one per parent, or the list) and the overflow queue's control blocks
// (one allocation, reused per call)
ErCtl* er_scratch(Context& c, const DPolys& P, bool same, const SelfNeedHost& need, SelfNeed& d_need) {
  const bool flags = same && !need.listed && need.flags, list = same && need.listed && need.n;
  const size_t need_bytes = flags ? (((size_t)P.nparent + 16 + 63) & ~size_t(63)) : list ? ((need.n * 4 + 63) & ~uint64_t(63)) : 0;
  uint8_t* er_misc = c.s_er_need.as<uint8_t>(need_bytes + 64);
  d_need = SelfNeed{};
  d_need.listed = same && need.listed ? 1 : 0;
  if (flags) {
    GSJ_HIP(hipMemcpyAsync(er_misc, need.flags, P.nparent, hipMemcpyHostToDevice, c.stream));
    d_need.flags = er_misc;
  } else if (list) {
    GSJ_HIP(hipMemcpyAsync(er_misc, need.list, need.n * 4, hipMemcpyHostToDevice, c.stream));
    d_need.list = reinterpret_cast<const uint32_t*>(er_misc);
    d_need.n = (uint32_t)need.n;
  }
  return reinterpret_cast<ErCtl*>(er_misc + need_bytes);
}

}  // namespace

void erode_run(Context& c, const Layer& PL, const Layer& RL, double ra, double rb, bool same, const SelfNeedHost& need, const gsj_join_opts& o,
               gsj_dist_cb cb, void* user, gsj_join_stats* st) {
  GSJ_REQUIRE(!PL.points && !RL.points, GSJ_E_ARG, "intersects_buffered: both layers must be polygon layers (points: the points entry point)");
  GSJ_REQUIRE(RL.has_bvh, GSJ_E_ARG, "intersects_buffered: build layer has no BVH");
  GSJ_REQUIRE(ra >= 0 && rb >= 0 && ra < 1e5 && rb < 1e5, GSJ_E_ARG, "intersects_buffered: radii must be in [0, 1e5) metres");
  gsj_join_stats s{};
  if (PL.n == 0 || RL.n == 0) { if (st) *st = s; return; }
  hipStream_t stream = c.stream;
  const DPolys& P = PL.poly;
  const DPolys& R = RL.poly;
  const uint32_t* pfP = P.pfirst;   // first piece per parent (subdivided layers), as k_piece_first builds it
  const uint32_t* pfR = R.pfirst;
  GSJ_REQUIRE(!need.listed || need.n < (1ull << 32), GSJ_E_ARG, "intersects_buffered: self-pair list too long");
  SelfNeed d_need;
  ErCtl* ctl = er_scratch(c, P, same, need, d_need);
  ErRadii rr{};
  rr.k = 1;
  rr.ra[0] = ra;
  rr.rb[0] = rb;
  gsj_join_opts o1 = o;
  o1.want_area = 1;
  o1.want_relate = 0;
  o1.no_agg = 0;
  JoinHooks hooks;
  hooks.device_only = true;
  if (ra > 0 || rb > 0)
    hooks.prefilter = [&](uint32_t* pa, uint32_t* pb, uint64_t n) -> uint64_t {
      uint32_t* keep = c.s_er_keep.as<uint32_t>(n + 16);
      uint64_t* off = c.s_er_off.as<uint64_t>(n + 16);
      uint32_t* tmp = c.s_er_pairs.as<uint32_t>(2 * n + 16);
      hipLaunchKernelGGL(k_er_prefilter, dim3(div_up(n, 256)), dim3(256), 0, stream, P, R, pa, pb, n, rr, keep);
      GSJ_LAUNCHED();
      const uint64_t k = scan_u32_to_u64(c, keep, off, n);
      hipLaunchKernelGGL(k_compact_pairs, dim3(div_up(n, 256)), dim3(256), 0, stream, pa, pb, keep, off, n, tmp, tmp + n);
      GSJ_LAUNCHED();
      if (k) {
        hipLaunchKernelGGL(k_copy_u32, dim3(div_up(k, 256)), dim3(256), 0, stream, tmp, k, pa);
        hipLaunchKernelGGL(k_copy_u32, dim3(div_up(k, 256)), dim3(256), 0, stream, tmp + n, k, pb);
        GSJ_LAUNCHED();
      }
      GSJ_HIP(hipStreamSynchronize(stream));
      return k;
    };
  uint64_t rows_out = 0;
  double er_ms = 0;
  hooks.on_device = [&](const uint32_t* rows, const double* area, uint64_t n) {
    WallTimer t0;
    uint32_t* out = nullptr;
    const uint64_t nhit = erode_pass2(c, P, R, pfP, pfR, rows, area, n, rr, same, d_need, ctl, s, out, nullptr, nullptr);
    er_ms += t0.ms();
    if (nhit == 0) return;
    rows_out += nhit;
    WallTimer t4;
    uint32_t* hrow = c.h_pairs.as<uint32_t>(2 * nhit);
    GSJ_HIP(hipMemcpyAsync(hrow, out, nhit * 8, hipMemcpyDeviceToHost, stream));
    GSJ_HIP(hipStreamSynchronize(stream));
    s.copy_ms += t4.ms();
    if (cb && cb(user, hrow, nullptr, (size_t)nhit)) throw Error(GSJ_E_CANCEL, "intersects_buffered cancelled by the row callback");
  };
  gsj_join_stats js{};
  join_run(c, PL, RL, o1, nullptr, nullptr, &js, &hooks);
  s.traverse_ms = js.traverse_ms;
  s.candidates = js.candidates;
  s.batches = js.batches;
  s.refine_ms = js.refine_ms + er_ms;
  s.compact_ms += js.compact_ms;
  s.k_small_ms = js.k_small_ms;
  s.n_small = js.n_small + js.n_large;
  s.hits = rows_out;
  if (st) *st = s;
}

bool erode_batchable(const Layer& P, const Layer& R, const gsj_join_opts& o, const SelfNeedHost& need) {
  // one self-pair rule for the whole batch: every self pair, or (an empty list) none
  return o.has_window && !o.cancel && !need.flags && !(need.listed && need.n) && !P.points && !R.points && R.has_bvh && P.n && R.n && !P.poly.parent && !R.poly.parent &&
         probes_batchable(probes_of(P));
}

void erode_batch_run(Context& c, const Layer& PL, const Layer& RL, bool same, bool no_self, std::vector<BatchErode>& reqs) {
  const uint32_t k = (uint32_t)reqs.size();
  GSJ_REQUIRE(k >= 1 && k <= MAX_WINDOWS, GSJ_E_ARG, "batch size out of range");
  const gsj_join_opts& o = *reqs[0].o;
  WindowSet ws;
  ErRadii rr{};
  ws.k = rr.k = k;
  for (uint32_t r = 0; r < k; r++) {
    const gsj_join_opts& q = *reqs[r].o;
    GSJ_REQUIRE(erode_batchable(PL, RL, q, SelfNeedHost{}) && q.upper == o.upper && q.batch_q == o.batch_q && q.max_pairs == o.max_pairs &&
                    q.snap_half_ulp == o.snap_half_ulp,
                GSJ_E_ARG, "request cannot join this batch");
    GSJ_REQUIRE(reqs[r].ra >= 0 && reqs[r].rb >= 0 && reqs[r].ra < 1e5 && reqs[r].rb < 1e5, GSJ_E_ARG,
                "intersects_buffered: radii must be in [0, 1e5) metres");
    for (int j = 0; j < 4; j++) ws.w[r][j] = q.window[j];
    rr.ra[r] = reqs[r].ra;
    rr.rb[r] = reqs[r].rb;
  }
  hipStream_t stream = c.stream;
  const DPolys& P = PL.poly;
  const DPolys& R = RL.poly;
  SelfNeedHost need;
  need.listed = no_self;   // an empty list: no parent's pair with itself
  SelfNeed d_need;
  ErCtl* ctl = er_scratch(c, P, same, need, d_need);
  RefineParams rp;   // pass 1: intersection areas
  rp.want_area = 1;
  rp.want_relate = 0;
  rp.snap_half_ulp = o.snap_half_ulp ? 1 : 0;
  gsj_join_stats s{};
  std::vector<uint64_t> cand(k, 0), hits(k, 0), rlo(k + 1);
  uint64_t* lo1 = wseg(c, 4);   // kept candidates
  uint64_t* lo2 = wseg(c, 5);   // pass-1 rows
  uint64_t* lo3 = wseg(c, 6);   // pass-2 pairs
  uint64_t* lo4 = wseg(c, 7);   // output rows
  for_each_candidate_multi(c, probes_of(PL), RL.bvh, o.upper != 0, o, ws, s, [&](uint32_t* pa, uint32_t* pb, uint64_t np, const SegBounds& sb) {
    for (uint32_t r = 0; r < k; r++) cand[r] += sb.h[r + 1] - sb.h[r];
    // pass 1: candidates that the erosion can leave intersecting, refined for their overlap areas
    WallTimer tp;
    uint32_t* keep = c.s_er_keep.as<uint32_t>(np + 16);
    uint64_t* off = c.s_er_off.as<uint64_t>(np + 16);
    uint32_t* tmp = c.s_er_pairs.as<uint32_t>(2 * np + 16);
    ErRadii r1 = rr;
    r1.seg = sb.d;
    hipLaunchKernelGGL(k_er_prefilter, dim3(div_up(np, 256)), dim3(256), 0, stream, P, R, pa, pb, np, r1, keep);
    GSJ_LAUNCHED();
    const uint64_t nk = scan_u32_to_u64(c, keep, off, np);
    seg_bounds_dev(c, sb.d, k, off, np, nk, lo1);
    hipLaunchKernelGGL(k_compact_pairs, dim3(div_up(np, 256)), dim3(256), 0, stream, pa, pb, keep, off, np, tmp, tmp + np);
    GSJ_LAUNCHED();
    s.compact_ms += tp.ms();
    if (nk == 0) return;
    hipLaunchKernelGGL(k_copy_u32, dim3(div_up(nk, 256)), dim3(256), 0, stream, tmp, nk, pa);
    hipLaunchKernelGGL(k_copy_u32, dim3(div_up(nk, 256)), dim3(256), 0, stream, tmp + np, nk, pb);
    GSJ_LAUNCHED();
    WallTimer t2;
    RefineOut ro{};
    ro.isect = c.s_flag.as<uint8_t>(nk + 16);
    ro.area = c.s_area.as<double>(nk + 16);
    s.heavy += refine_pairs(c, P, R, pa, pb, nk, rp, ro, &s);
    uint32_t* orow = c.s_out_a.as<uint32_t>(2 * nk + 16);
    double* oarea = c.s_out_area.as<double>(nk + 16);
    const uint64_t* offs = nullptr;
    const uint64_t nrows = compact_single_rows(c, PL, RL, pa, pb, nk, rp, ro, orow, oarea, &offs);
    seg_bounds_dev(c, lo1, k, offs, nk, nrows, lo2);
    // pass 2
    ErRadii r2 = rr;
    r2.seg = lo2;
    uint32_t* out = nullptr;
    const uint64_t nhit = nrows ? erode_pass2(c, P, R, nullptr, nullptr, orow, oarea, nrows, r2, same, d_need, ctl, s, out, lo3, lo4) : 0;
    if (nrows == 0) seg_bounds_dev(c, lo2, k, offs, 0, 0, lo4);
    GSJ_HIP(hipMemcpyAsync(rlo.data(), lo4, (size_t)(k + 1) * 8, hipMemcpyDeviceToHost, stream));
    GSJ_HIP(hipStreamSynchronize(stream));
    s.refine_ms += t2.ms();
    if (nhit == 0) return;
    WallTimer t4;
    uint32_t* hrow = c.h_pairs.as<uint32_t>(2 * nhit);
    GSJ_HIP(hipMemcpyAsync(hrow, out, nhit * 8, hipMemcpyDeviceToHost, stream));
    GSJ_HIP(hipStreamSynchronize(stream));
    s.copy_ms += t4.ms();
    for (uint32_t r = 0; r < k; r++) {
      const uint64_t a = rlo[r], n = rlo[r + 1] - rlo[r];
      hits[r] += n;
      BatchErode& q = reqs[r];
      if (n == 0 || !q.cb || q.rc != GSJ_OK) continue;
      if (q.cb(q.user, hrow + 2 * a, nullptr, (size_t)n)) q.rc = GSJ_E_CANCEL;
    }
  });
  for (uint32_t r = 0; r < k; r++) {
    if (!reqs[r].st) continue;
    gsj_join_stats& st = *reqs[r].st;
    st = s;
    st.candidates = cand[r];
    st.hits = hits[r];
  }
}

}  // namespace gsj
