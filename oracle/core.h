// ORACLE (test infrastructure): shared core of the CPU restatement — arithmetic type, exact predicates, layers,
// candidate enumeration, edge lists and containment. Included by oracle.cpp and buffer.cpp.
#pragma once
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <numeric>
#include <vector>
#include <string>
#include <unordered_map>
#include <functional>
#ifdef _OPENMP
#include <omp.h>
#endif

// Planted mistakes (tests only): a build with -DORC_MUTANT=k changes one rule of the restatement, so that a test can
// show that its checks fail on that mistake. 0 (every shipped build) changes nothing.
//   1  relate evidence: a real edge running along the other piece's cut edge counts its full length inside (not half)
//   2  area: a sub-segment on a coincident boundary with interiors on one side weighs 1 (not 1/2)
//   3  cut flags: the first edge of a ring takes its start-vertex cut flag from the closing vertex (off by one)
//   4  piece -> parent fold: area of area-only piece pairs (flag 2) left out of the parent's sum
//   5  coincidence: a cut endpoint gets no tolerance (rounded cut points must lie exactly on the other line)
#ifndef ORC_MUTANT
#define ORC_MUTANT 0
#endif

namespace orc {

// ------------------------------------------------------------------------------------------- arithmetic type
// Real is double; the FLOP-counting build (-DORC_COUNT_FLOPS, liboracle_flops.so) swaps in a double wrapper that
// counts every add/sub, mul, fma, div and sqrt the refinement executes: the algorithmic FLOPs of a pair.
#ifdef ORC_COUNT_FLOPS
struct FlopCount { uint64_t add = 0, mul = 0, fma = 0, div = 0, sqrt = 0; };
static thread_local FlopCount g_fc;
[[maybe_unused]] static thread_local FlopCount g_fc_edge;   // the part spent in edge contributions (area / relate terms)
[[maybe_unused]] static thread_local uint64_t g_edge_pairs;  // pairs that reached the edge contributions
struct Real {
  double v;
  Real() = default;
  Real(double x) : v(x) {}
};
static inline Real operator+(Real a, Real b) { g_fc.add++; return Real(a.v + b.v); }
static inline Real operator-(Real a, Real b) { g_fc.add++; return Real(a.v - b.v); }
static inline Real operator-(Real a) { return Real(-a.v); }
static inline Real operator*(Real a, Real b) { g_fc.mul++; return Real(a.v * b.v); }
static inline Real operator/(Real a, Real b) { g_fc.div++; return Real(a.v / b.v); }
static inline Real& operator+=(Real& a, Real b) { a = a + b; return a; }
static inline bool operator<(Real a, Real b) { return a.v < b.v; }
static inline bool operator>(Real a, Real b) { return a.v > b.v; }
static inline bool operator<=(Real a, Real b) { return a.v <= b.v; }
static inline bool operator>=(Real a, Real b) { return a.v >= b.v; }
static inline bool operator==(Real a, Real b) { return a.v == b.v; }
static inline bool operator!=(Real a, Real b) { return a.v != b.v; }
static inline Real r_fma(Real a, Real b, Real c) { g_fc.fma++; return Real(std::fma(a.v, b.v, c.v)); }
static inline Real r_min(Real a, Real b) { return Real(std::fmin(a.v, b.v)); }
static inline Real r_max(Real a, Real b) { return Real(std::fmax(a.v, b.v)); }
static inline Real r_abs(Real a) { return Real(std::fabs(a.v)); }
static inline Real r_sqrt(Real a) { g_fc.sqrt++; return Real(std::sqrt(a.v)); }
static inline double val(Real a) { return a.v; }
#else
using Real = double;
static inline Real r_fma(Real a, Real b, Real c) { return std::fma(a, b, c); }
static inline Real r_min(Real a, Real b) { return std::fmin(a, b); }
static inline Real r_max(Real a, Real b) { return std::fmax(a, b); }
static inline Real r_abs(Real a) { return std::fabs(a); }
static inline Real r_sqrt(Real a) { return std::sqrt(a); }
static inline double val(Real a) { return a; }
#endif

// ------------------------------------------------------------------------------------------- predicates
// engine/engine.cuh:283-289: 2x2 determinant (b-a) x (c-a) with the product rounding error recovered by fma
static inline Real orient(Real ax, Real ay, Real bx, Real by, Real cx, Real cy) {
  Real a = bx - ax, b = cy - ay, c = by - ay, d = cx - ax;
  Real w = c * d;
  Real err = r_fma(-c, d, w);
  Real f = r_fma(a, b, -w);
  return f + err;
}
// engine/engine.cuh:290-292: closed bbox test for a point known to be collinear
static inline bool onseg(Real ax, Real ay, Real bx, Real by, Real px, Real py) {
  return px >= r_min(ax, bx) && px <= r_max(ax, bx) && py >= r_min(ay, by) && py <= r_max(ay, by);
}
// engine/engine.cuh:298-304: p left of a->b, with p displaced by (eps, eps^2) when exactly on the line
static inline bool left_of(Real ax, Real ay, Real bx, Real by, Real px, Real py) {
  Real o = orient(ax, ay, bx, by, px, py);
  if (o != 0) return o > 0;
  Real dy = by - ay;
  if (dy != 0) return dy < 0;
  return bx > ax;
}
// engine/engine.cuh:307-316: closed-segment intersection (touching counts)
static inline bool seg_isect(Real p0x, Real p0y, Real p1x, Real p1y, Real q0x, Real q0y, Real q1x, Real q1y) {
  Real d1 = orient(q0x, q0y, q1x, q1y, p0x, p0y), d2 = orient(q0x, q0y, q1x, q1y, p1x, p1y);
  Real d3 = orient(p0x, p0y, p1x, p1y, q0x, q0y), d4 = orient(p0x, p0y, p1x, p1y, q1x, q1y);
  bool straddle_q = (d1 > 0 && d2 < 0) || (d1 < 0 && d2 > 0);
  bool straddle_p = (d3 > 0 && d4 < 0) || (d3 < 0 && d4 > 0);
  if (straddle_q && straddle_p) return true;
  if (d1 == 0 && onseg(q0x, q0y, q1x, q1y, p0x, p0y)) return true;
  if (d2 == 0 && onseg(q0x, q0y, q1x, q1y, p1x, p1y)) return true;
  if (d3 == 0 && onseg(p0x, p0y, p1x, p1y, q0x, q0y)) return true;
  if (d4 == 0 && onseg(p0x, p0y, p1x, p1y, q1x, q1y)) return true;
  return false;
}

// engine/engine.cuh:34-35: outward rounding of a double to float
static inline float f_down(double v) { float f = (float)v; if ((double)f > v) f = std::nextafter(f, -INFINITY); return f; }
static inline float f_up(double v) { float f = (float)v; if ((double)f < v) f = std::nextafter(f, INFINITY); return f; }

// ------------------------------------------------------------------------------------------- layers
// The prepared form of engine/engine.cuh:19-32 / 42-89 (DSet, upload_set), kept on the host.
struct Layer {
  bool points = false;
  uint64_t n = 0, nr = 0, nv = 0;
  std::vector<uint64_t> gro, rvo;          // [n+1], [nr+1]
  std::vector<double> xw;                  // world xy [2nv] (points: [2n])
  std::vector<double> q;                   // xy - origin of its geometry [2nv]
  std::vector<double> ox, oy, ex, ey;      // origin (bbox min) and extent per geometry
  std::vector<float> bb;                   // [4n] outward float bbox
  std::vector<uint32_t> parent;            // empty: not subdivided
  std::vector<uint8_t> syn;                // [nv] edge v->v+1 was made by subdivision
  std::vector<double> blen;                // per parent: length of the real boundary
  uint32_t nparent = 0;
  bool has_syn() const { return !syn.empty(); }
  bool is_syn(uint64_t v) const { return !syn.empty() && syn[v]; }
};

static void prep_polys(Layer& L) {
  const uint64_t n = L.n;
  L.ox.resize(n); L.oy.resize(n); L.ex.resize(n); L.ey.resize(n); L.bb.resize(4 * n); L.q.resize(2 * L.nv);
#pragma omp parallel for schedule(dynamic, 4096)
  for (int64_t g = 0; g < (int64_t)n; g++) {
    uint64_t v0 = L.rvo[L.gro[g]], v1 = L.rvo[L.gro[g + 1]];
    double x0 = 1e300, y0 = 1e300, x1 = -1e300, y1 = -1e300;
    for (uint64_t v = v0; v < v1; v++) {
      double x = L.xw[2 * v], y = L.xw[2 * v + 1];
      x0 = std::min(x0, x); x1 = std::max(x1, x); y0 = std::min(y0, y); y1 = std::max(y1, y);
    }
    L.ox[g] = x0; L.oy[g] = y0; L.ex[g] = x1 - x0; L.ey[g] = y1 - y0;
    L.bb[4 * g] = f_down(x0); L.bb[4 * g + 1] = f_down(y0); L.bb[4 * g + 2] = f_up(x1); L.bb[4 * g + 3] = f_up(y1);
    for (uint64_t v = v0; v < v1; v++) { L.q[2 * v] = L.xw[2 * v] - x0; L.q[2 * v + 1] = L.xw[2 * v + 1] - y0; }
  }
  // engine/engine.cuh:69-82: real boundary length per geometry, summed per parent in geometry order
  std::vector<double> glen(n, 0.0);
#pragma omp parallel for schedule(dynamic, 4096)
  for (int64_t g = 0; g < (int64_t)n; g++) {
    double s = 0;
    for (uint64_t r = L.gro[g]; r < L.gro[g + 1]; r++)
      for (uint64_t v = L.rvo[r]; v + 1 < L.rvo[r + 1]; v++)
        if (!L.is_syn(v)) s += std::hypot(L.xw[2 * v + 2] - L.xw[2 * v], L.xw[2 * v + 3] - L.xw[2 * v + 1]);
    glen[g] = s;
  }
  L.nparent = L.parent.empty() ? (uint32_t)n : (n ? L.parent[n - 1] + 1 : 0);
  L.blen.assign(L.nparent, 0.0);
  for (uint64_t g = 0; g < n; g++) L.blen[L.parent.empty() ? g : L.parent[g]] += glen[g];
}

static void prep_points(Layer& L) {
  L.bb.resize(4 * L.n);
  for (uint64_t i = 0; i < L.n; i++) {
    double x = L.xw[2 * i], y = L.xw[2 * i + 1];
    L.bb[4 * i] = f_down(x); L.bb[4 * i + 1] = f_down(y); L.bb[4 * i + 2] = f_up(x); L.bb[4 * i + 3] = f_up(y);
  }
}

// ------------------------------------------------------------------------------------------- candidates
// The LBVH traversal (engine/engine.cuh:233-259) returns exactly the leaves whose float box overlaps the query
// box (closed test, line 241), restricted to leaf >= query for self-joins (line 248). That set does not depend
// on the tree, so the oracle enumerates it with a uniform grid instead.
struct Grid {
  float x0, y0, cw, ch; int nx, ny;
  std::vector<uint32_t> start, items;
};
static Grid build_grid(const std::vector<float>& bb, uint64_t n) {
  Grid G{};
  float x0 = INFINITY, y0 = INFINITY, x1 = -INFINITY, y1 = -INFINITY;
  std::vector<float> w;
  w.reserve(n);
  for (uint64_t i = 0; i < n; i++) {
    x0 = std::min(x0, bb[4 * i]); y0 = std::min(y0, bb[4 * i + 1]); x1 = std::max(x1, bb[4 * i + 2]); y1 = std::max(y1, bb[4 * i + 3]);
    w.push_back(std::max(bb[4 * i + 2] - bb[4 * i], bb[4 * i + 3] - bb[4 * i + 1]));
  }
  if (n == 0) { G.nx = G.ny = 1; G.x0 = G.y0 = 0; G.cw = G.ch = 1; G.start.assign(2, 0); return G; }
  std::nth_element(w.begin(), w.begin() + w.size() / 2, w.end());
  double med = std::max((double)w[w.size() / 2], 1e-9);
  double spanx = std::max((double)x1 - x0, 1e-9), spany = std::max((double)y1 - y0, 1e-9);
  double cell = std::max(med, std::sqrt(spanx * spany / std::max<double>(1.0, (double)n)));
  G.nx = (int)std::min(4096.0, std::max(1.0, spanx / cell)); G.ny = (int)std::min(4096.0, std::max(1.0, spany / cell));
  G.x0 = x0; G.y0 = y0; G.cw = (float)(spanx / G.nx); G.ch = (float)(spany / G.ny);
  auto cx = [&](float x) { int c = (int)std::floor((x - G.x0) / G.cw); return std::min(std::max(c, 0), G.nx - 1); };
  auto cy = [&](float y) { int c = (int)std::floor((y - G.y0) / G.ch); return std::min(std::max(c, 0), G.ny - 1); };
  std::vector<uint32_t> cnt((size_t)G.nx * G.ny + 1, 0);
  for (uint64_t i = 0; i < n; i++)
    for (int j = cy(bb[4 * i + 1]); j <= cy(bb[4 * i + 3]); j++)
      for (int k = cx(bb[4 * i]); k <= cx(bb[4 * i + 2]); k++) cnt[(size_t)j * G.nx + k + 1]++;
  for (size_t c = 1; c < cnt.size(); c++) cnt[c] += cnt[c - 1];
  G.start = cnt;
  G.items.resize(cnt.back());
  std::vector<uint32_t> fill(cnt.begin(), cnt.end() - 1);
  for (uint64_t i = 0; i < n; i++)
    for (int j = cy(bb[4 * i + 1]); j <= cy(bb[4 * i + 3]); j++)
      for (int k = cx(bb[4 * i]); k <= cx(bb[4 * i + 2]); k++) G.items[fill[(size_t)j * G.nx + k]++] = (uint32_t)i;
  return G;
}

// candidate pairs (q, r) with q in [0, nq): query boxes qbb[4nq] against R's boxes; sorted by (q, r)
static void candidates(const std::vector<float>& qbb, uint64_t nq, const std::vector<float>& rbb, uint64_t nr, bool upper,
                       std::vector<uint32_t>& out) {
  Grid G = build_grid(rbb, nr);
  int T = 1;
#ifdef _OPENMP
  T = omp_get_max_threads();
#endif
  std::vector<std::vector<uint32_t>> part(T);
#pragma omp parallel
  {
    int t = 0;
#ifdef _OPENMP
    t = omp_get_thread_num();
#endif
    std::vector<uint32_t> hits;
#pragma omp for schedule(dynamic, 1024)
    for (int64_t q = 0; q < (int64_t)nq; q++) {
      float bx0 = qbb[4 * q], by0 = qbb[4 * q + 1], bx1 = qbb[4 * q + 2], by1 = qbb[4 * q + 3];
      if (nr == 0) continue;
      int i0 = (int)std::floor((bx0 - G.x0) / G.cw), i1 = (int)std::floor((bx1 - G.x0) / G.cw);
      int j0 = (int)std::floor((by0 - G.y0) / G.ch), j1 = (int)std::floor((by1 - G.y0) / G.ch);
      // clamp both ends into the grid: a box touching the extent's far border registers in the last cell
      i0 = std::min(std::max(i0, 0), G.nx - 1); i1 = std::min(std::max(i1, 0), G.nx - 1);
      j0 = std::min(std::max(j0, 0), G.ny - 1); j1 = std::min(std::max(j1, 0), G.ny - 1);
      hits.clear();
      for (int j = j0; j <= j1; j++)
        for (int i = i0; i <= i1; i++)
          for (uint32_t s = G.start[(size_t)j * G.nx + i]; s < G.start[(size_t)j * G.nx + i + 1]; s++) {
            uint32_t r = G.items[s];
            const float* c = &rbb[4 * (size_t)r];
            if (c[0] > bx1 || c[2] < bx0 || c[1] > by1 || c[3] < by0) continue;
            if (upper && r < (uint64_t)q) continue;
            hits.push_back(r);
          }
      // boxes outside the grid's extent clamp into the border cells, so every overlap is found
      std::sort(hits.begin(), hits.end());
      hits.erase(std::unique(hits.begin(), hits.end()), hits.end());
      for (uint32_t r : hits) { part[t].push_back((uint32_t)q); part[t].push_back(r); }
    }
  }
  out.clear();
  for (auto& p : part) out.insert(out.end(), p.begin(), p.end());
  // order by (q, r)
  std::vector<uint64_t> k(out.size() / 2);
  for (size_t i = 0; i < k.size(); i++) k[i] = ((uint64_t)out[2 * i] << 32) | out[2 * i + 1];
  std::sort(k.begin(), k.end());
  for (size_t i = 0; i < k.size(); i++) { out[2 * i] = (uint32_t)(k[i] >> 32); out[2 * i + 1] = (uint32_t)k[i]; }
}

// ------------------------------------------------------------------------------------------- pair refinement
constexpr uint32_t REL_COIN_SAME = 1, REL_COIN_OPP = 2, REL_TOUCH = 4;   // engine/engine.cuh:334-336
constexpr int MAXS = 24, MAXO = 8;                                        // engine/engine.cuh:339-340

struct Params {   // engine/engine.cuh:517-518 (RefineParams)
  double tol_rel = 1e-12, eps_par = 1e-7, tol_abs = 1e-11;
  int want_area = 1, want_relate = 0, snap_half_ulp = 0;
};

struct Edge { Real x0, y0, x1, y1; int syn, cut; };   // engine/engine.cuh:356
struct Box { Real x0, y0, x1, y1; };

// a geometry seen from the frame of another geometry's origin (engine/engine.cuh:265-279)
struct View {
  const Layer* L; uint64_t g; Real shx, shy;
  void v(uint64_t i, Real& x, Real& y) const { x = shx + L->q[2 * i]; y = shy + L->q[2 * i + 1]; }
  int syn(uint64_t i) const { return L->is_syn(i) ? 1 : 0; }
  uint64_t r0() const { return L->gro[g]; }
  uint64_t r1() const { return L->gro[g + 1]; }
};

static inline bool edge_outside(Real px, Real py, Real qx, Real qy, const Box& o) {   // engine/engine.cuh:329-331
  return r_max(px, qx) < o.x0 || r_min(px, qx) > o.x1 || r_max(py, qy) < o.y0 || r_min(py, qy) > o.y1;
}

// engine/engine.cuh:358-377 without the capacity limit: edges meeting the box, in ring order
static void collect(const View& A, const Box& ov, std::vector<Edge>& out) {
  out.clear();
  for (uint64_t r = A.r0(); r < A.r1(); r++) {
    uint64_t s = A.L->rvo[r], e = A.L->rvo[r + 1];
    for (uint64_t i = s; i + 1 < e; i++) {
      Edge ed;
      A.v(i, ed.x0, ed.y0); A.v(i + 1, ed.x1, ed.y1);
      ed.syn = A.syn(i);
      ed.cut = A.L->has_syn() ? ((A.syn(i > s ? i - 1 : e - (ORC_MUTANT == 3 ? 1 : 2)) ? 1 : 0) | (A.syn(i + 2 < e ? i + 1 : s) ? 2 : 0)) : 0;
      if (!edge_outside(ed.x0, ed.y0, ed.x1, ed.y1, ov)) out.push_back(ed);
    }
  }
}

// engine/engine.cuh:382-397: parity of the segment from a point left of the geometry's bbox to m, over every edge
static bool pip_full(const View& B, Real mx, Real my) {
  Real ox = B.shx - 1.0, oy = my;
  bool par = false;
  for (uint64_t r = B.r0(); r < B.r1(); r++) {
    uint64_t s = B.L->rvo[r], e = B.L->rvo[r + 1];
    for (uint64_t j = s + 1; j < e; j++) {
      Real x0, y0, x1, y1; B.v(j - 1, x0, y0); B.v(j, x1, y1);
      if ((y0 > my) == (y1 > my) && y0 != my && y1 != my) continue;
      bool so = left_of(x0, y0, x1, y1, ox, oy), sc = left_of(x0, y0, x1, y1, mx, my);
      if (so == sc) continue;
      bool su = orient(ox, oy, mx, my, x0, y0) > 0, sv = orient(ox, oy, mx, my, x1, y1) > 0;
      if (su != sv) par = !par;
    }
  }
  return par;
}

// engine/engine.cuh:400-412: inside(x) = inside(c) xor parity of segment c->x across the local edges
static bool inside_from_c(const std::vector<Edge>& L, Real cx, Real cy, bool in_c, Real x, Real y) {
  bool in = in_c;
  Real sx0 = r_min(cx, x), sx1 = r_max(cx, x), sy0 = r_min(cy, y), sy1 = r_max(cy, y);
  for (const Edge& e : L) {
    if (r_max(e.x0, e.x1) < sx0 || r_min(e.x0, e.x1) > sx1 || r_max(e.y0, e.y1) < sy0 || r_min(e.y0, e.y1) > sy1) continue;
    bool sc = left_of(e.x0, e.y0, e.x1, e.y1, cx, cy), sx = left_of(e.x0, e.y0, e.x1, e.y1, x, y);
    if (sc == sx) continue;
    bool su = orient(cx, cy, x, y, e.x0, e.y0) > 0, sv = orient(cx, cy, x, y, e.x1, e.y1) > 0;
    if (su != sv) in = !in;
  }
  return in;
}
// ------------------------------------------------------------------------------------------- join (oracle.cpp)
struct JoinResult {
  std::vector<uint32_t> pairs;   // (a, b) parent ids
  std::vector<double> area;
  std::vector<uint32_t> code;
  std::vector<uint8_t> where;    // points: 1 interior, 2 boundary
  std::vector<uint32_t> root;    // cluster: root per parent
  uint64_t candidates = 0;
};
struct PairOut;
void join(const Layer& P, const Layer& R, const Params& rp, bool upper, bool no_agg, JoinResult& res, std::vector<uint32_t>* cand_out,
          std::vector<PairOut>* refined_out, const std::function<bool(uint32_t, uint32_t)>& prefilter = {});

}  // namespace orc
