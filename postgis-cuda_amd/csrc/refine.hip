// Exact FP64 refinement of candidate polygon pairs (reference: k_refine and the heavy path,
// engine/engine.cuh:342-792). Per pair: the edges of both polygons inside the pair's overlap box, boundary contact by
// exact segment tests, containment by crossing parity from a probe point c, and (for area / relate) Green's-theorem
// contributions of every such edge weighted by its position relative to the other polygon.
//
// MI355X design:
//  * three work classes, chosen per pair from the vertex and ring counts of its polygons:
//      small (both <= 32 stored vertices, <= 8 rings): a 16-lane group per pair (four pairs per wave64); both
//        polygons are loaded once into LDS as vertex lists with per-vertex flags (edge exists / synthetic / cut ends /
//        meets the overlap box), so the full-ring parity scans read LDS, not HBM; the contact test covers the whole
//        A x B edge grid across the group; A's and B's edge contributions run in one combined pass;
//      large: a wave64 per pair, compacted lists of the edges meeting the overlap box in LDS;
//      overflow (lists beyond the LDS capacity): a wave64 per pair (four waves per pair when a launch holds few
//        pairs) with lists in global memory sized by the largest overflow.
//  * split points of an edge live in eight named registers kept sorted by a branch-free min/max insertion, with the
//    overlaps in two register slots. A pair with an edge of more than six interior splits or two overlaps is handed to
//    the overflow kernel, which runs the array path with the reference's caps (24 / 8); the two paths give the same
//    bits wherever both apply, and the fast kernels carry no call and no scratch arrays.
//  * deterministic sums: each edge's partial sums start at zero and are stored per edge; group lanes 0..4 add them in
//    list order (A's edges, then B's). The result is the same in every class and bit-identical to the CPU
//    restatement, so a pair gives the same bits alone or batched with any other request.
//  * FP64 on the vector ALUs only; nothing here is a dense contraction, so MFMA is not used.
#include "pairs.h"

namespace gsj {

static_assert(MCAP <= (int)PACK_VCAP && RCAP <= (int)PACK_RCAP, "every piece of the vertex-list classes has a packed record");

namespace {

// ------------------------------------------------------------------------------------------- small-class source
// The small class reads a polygon's edges from its LDS vertex list (the list classes use ListSrc).
// per-vertex flags in LDS: the layer's VF_SYN / VF_CUT0 / VF_CUT1 bits plus these two
constexpr uint8_t F_VALID = 1, F_SYN = VF_SYN, F_INOV = 16;
static_assert(VF_CUT0 == 4 && VF_CUT1 == 8, "Ed::cut = (flags >> 2) & 3");

struct VertSrc {   // the edges of a small polygon meeting the overlap box: list entry j is edge (V[I[j]], V[I[j] + 1])
  const double2* V;
  const uint8_t* F;
  const uint8_t* I;
  int n;
  __device__ __forceinline__ int size() const { return n; }
  __device__ __forceinline__ bool get(int j, Ed& e) const {
    const int k = I[j];
    const uint8_t f = F[k];
    const double2 a = V[k], b = V[k + 1];
    e.x0 = a.x; e.y0 = a.y; e.x1 = b.x; e.y1 = b.y;
    e.syn = (f & F_SYN) ? 1 : 0;
    e.cut = (f >> 2) & 3;
    return true;
  }
};

struct Contrib { double acc, lin, lon; uint32_t bits; };
constexpr uint32_t SPILL = 0x80000000u;   // Contrib.bits: the edge exceeded the register path (pair goes to overflow)

struct EdgeCtx {   // what an edge's contribution needs besides the other polygon's edges
  double cx, cy, tol, eps_par, snap;
  Box ov;
  bool in_c, rel;
};

// One sub-segment [t0, t1] of edge p->q: relate lengths and the Green's-theorem term (engine/engine.cuh:464-513).
// `net`, `first_*` describe the collinear overlaps containing the midpoint.
template <class Src>
__device__ __forceinline__ void subsegment(const Src& L, const EdgeCtx& c, double px, double py, double dx, double dy, double len2, bool real,
                                           double t0, double t1, int net, int net_real, int n_real, int first_idx, bool first_same,
                                           Contrib& out) {
  const double tm = 0.5 * (t0 + t1);
  const double mx = px + dx * tm, my = py + dy * tm;
  const bool in_ov = mx >= c.ov.x0 && mx <= c.ov.x1 && my >= c.ov.y0 && my <= c.ov.y1;
  int vin = -1;   // midpoint inside the other polygon, computed on first use
  auto inside_m = [&]() {
    if (vin < 0) vin = (in_ov && inside_from_c(L, c.cx, c.cy, c.in_c, mx, my)) ? 1 : 0;
    return vin;
  };
  if (c.rel && real) {
    const double seg = sqrt(len2) * (t1 - t0);
    if (net_real != 0) { out.lon += seg; out.bits |= (net_real > 0 ? REL_COIN_SAME : REL_COIN_OPP) | REL_TOUCH; }
    else if (n_real == 0 && (net == 1 || net == -1)) out.lin += 0.5 * seg;
    else if (inside_m()) out.lin += seg;
  }
  double w;
  if (first_idx < 0) w = inside_m() ? 1.0 : 0.0;
  else if (net == 1 || net == -1) w = 0.5;
  else if (net == 0) w = inside_m() ? 1.0 : 0.0;
  else {   // |net| >= 2 only for inconsistent input: derive the other side from the classified one and clamp
    Ed eb;
    L.get(first_idx, eb);
    const bool sb = left_of(eb.x0, eb.y0, eb.x1, eb.y1, mx, my);
    const bool m_left = first_same ? sb : !sb;
    const int v = inside_m();
    int other = m_left ? v - net : v + net;
    other = other < 0 ? 0 : (other > 1 ? 1 : other);
    w = 0.5 * (v + other);
  }
  if (w != 0.0) {
    const double ax = px + dx * t0 - c.cx, ay = py + dy * t0 - c.cy, bx = px + dx * t1 - c.cx, by = py + dy * t1 - c.cy;
    out.acc += w * (ax * by - ay * bx);
  }
}

// Classification of one other-polygon edge against edge p->q (engine/engine.cuh:427-461): returns 0 (nothing),
// 1 (collinear overlap [lo, hi] in edge parameter, same direction in *same), 2 (proper split at *t).
__device__ __forceinline__ int classify_edge(const Ed& e, double px, double py, double qx, double qy, double dx, double dy, double len2, int cut,
                                             const EdgeCtx& c, double& lo, double& hi, bool& same, double& t) {
  const double tol = c.tol;
  const double bx0 = fmin(px, qx) - tol, bx1 = fmax(px, qx) + tol, by0 = fmin(py, qy) - tol, by1 = fmax(py, qy) + tol;
  if (fmax(e.x0, e.x1) < bx0 || fmin(e.x0, e.x1) > bx1 || fmax(e.y0, e.y1) < by0 || fmin(e.y0, e.y1) > by1) return 0;
  const double fx = e.x1 - e.x0, fy = e.y1 - e.y0;
  const double r = dx * fy - dy * fx;
  const double flen2 = fx * fx + fy * fy;
  const double snap2 = c.snap * c.snap, tol2len2 = tol * tol * len2, eps2 = c.eps_par * c.eps_par;
  // coincident if either edge's endpoints lie on the other's line (within snap, or within tol at cut points)
  const double c0 = orient(px, py, qx, qy, e.x0, e.y0), c1 = orient(px, py, qx, qy, e.x1, e.y1);
  bool col = (c0 * c0 <= snap2 * len2 || ((e.cut & 1) && c0 * c0 <= tol2len2)) && (c1 * c1 <= snap2 * len2 || ((e.cut & 2) && c1 * c1 <= tol2len2));
  if (!col && flen2 > 0) {
    const double g0 = orient(e.x0, e.y0, e.x1, e.y1, px, py), g1 = orient(e.x0, e.y0, e.x1, e.y1, qx, qy), tf = tol * tol * flen2;
    col = (g0 * g0 <= snap2 * flen2 || ((cut & 1) && g0 * g0 <= tf)) && (g1 * g1 <= snap2 * flen2 || ((cut & 2) && g1 * g1 <= tf));
  }
  if (col) {
    const double tc = ((e.x0 - px) * dx + (e.y0 - py) * dy) / len2, te = ((e.x1 - px) * dx + (e.y1 - py) * dy) / len2;
    lo = fmax(0.0, fmin(tc, te));
    hi = fmin(1.0, fmax(tc, te));
    same = (dx * fx + dy * fy) > 0;
    return hi > lo ? 1 : 0;
  }
  if (r * r > eps2 * len2 * flen2) {
    const double wx = e.x0 - px, wy = e.y0 - py;
    const double tn = wx * fy - wy * fx;
    if (r > 0 ? (tn <= 0 || tn >= r) : (tn >= 0 || tn <= r)) return 0;
    t = tn / r;
    const double u = (wx * dy - wy * dx) / r;
    const double du = tol / (flen2 > 0 ? sqrt(flen2) : 1.0);
    return (u >= -du && u <= 1 + du) ? 2 : 0;
  }
  return 0;
}

// Array path with the reference's caps (MAXS split points, MAXO overlaps per edge, in list order): taken only by
// edges whose splits or overlaps exceed the register path's capacity.
template <class Src>
__device__ Contrib edge_contrib_capped(const Ed& E, const Src& L, const EdgeCtx& c) {
  Contrib out{0.0, 0.0, 0.0, 0u};
  const double px = E.x0, py = E.y0, qx = E.x1, qy = E.y1;
  const double dx = qx - px, dy = qy - py, len2 = dx * dx + dy * dy;
  if (len2 == 0) return out;
  double ts[MAXS];
  int ns = 0;
  double ol0[MAXO], ol1[MAXO];
  bool osame[MAXO], oreal[MAXO];
  int oidx[MAXO];
  int no = 0;
  ts[ns++] = 0; ts[ns++] = 1;
  // box test first, then the classification of the marked entries in list order (as in the register path below)
  const int n = L.size();
  const double tx0 = fmin(px, qx) - c.tol, tx1 = fmax(px, qx) + c.tol, ty0 = fmin(py, qy) - c.tol, ty1 = fmax(py, qy) + c.tol;
  for (int base = 0; base < n; base += 64) {
    const int m = min(64, n - base);
    uint64_t cand = 0;
    for (int j = 0; j < m; j++) {
      Ed e;
      if (!L.get(base + j, e)) continue;
      if (!(fmax(e.x0, e.x1) < tx0 || fmin(e.x0, e.x1) > tx1 || fmax(e.y0, e.y1) < ty0 || fmin(e.y0, e.y1) > ty1)) cand |= 1ull << j;
    }
    while (cand != 0) {
      const int k = base + __builtin_ctzll(cand);
      cand &= cand - 1;
      Ed e;
      L.get(k, e);
      double lo, hi, t;
      bool same;
      const int kind = classify_edge(e, px, py, qx, qy, dx, dy, len2, E.cut, c, lo, hi, same, t);
      if (kind == 1) {
        if (lo > 0 && ns < MAXS) ts[ns++] = lo;
        if (hi < 1 && ns < MAXS) ts[ns++] = hi;
        if (no < MAXO) { ol0[no] = lo; ol1[no] = hi; osame[no] = same; oreal[no] = !e.syn; oidx[no] = k; no++; }
      } else if (kind == 2 && ns < MAXS) {
        ts[ns++] = t;
      }
    }
  }
  for (int a = 1; a < ns; a++) {
    const double v = ts[a];
    int b = a - 1;
    while (b >= 0 && ts[b] > v) { ts[b + 1] = ts[b]; b--; }
    ts[b + 1] = v;
  }
  for (int k = 0; k + 1 < ns; k++) {
    const double t0 = ts[k], t1 = ts[k + 1];
    if (!(t1 > t0)) continue;
    const double tm = 0.5 * (t0 + t1);
    int net = 0, first = -1, net_real = 0, n_real = 0;
    for (int o = 0; o < no; o++)
      if (tm > ol0[o] && tm < ol1[o]) {
        const int sgn = osame[o] ? 1 : -1;
        net += sgn;
        if (first < 0) first = o;
        if (oreal[o]) { net_real += sgn; n_real++; }
      }
    subsegment<Src>(L, c, px, py, dx, dy, len2, !E.syn, t0, t1, net, net_real, n_real, first < 0 ? -1 : oidx[first], first < 0 ? false : osame[first],
                    out);
  }
  return out;
}

// One edge's Green's-theorem term (relative to c) and, for a real edge with relate on, its length inside / on the
// other polygon (engine/engine.cuh:416-515). Partial sums start at zero for every edge. Register path: up to eight
// split values (0 and 1 included) kept sorted, up to two overlaps; beyond that the result carries SPILL.
template <class Src, bool CAPPED>
__device__ Contrib edge_contrib(const Ed& E, const Src& L, const EdgeCtx& c) {
  if constexpr (CAPPED) return edge_contrib_capped<Src>(E, L, c);
  Contrib out{0.0, 0.0, 0.0, 0u};
  const double px = E.x0, py = E.y0, qx = E.x1, qy = E.y1;
  const double dx = qx - px, dy = qy - py, len2 = dx * dx + dy * dy;
  if (len2 == 0) return out;
  // sorted split values; 2.0 marks an empty slot (every split lies in [0, 1])
  double s0 = 0.0, s1 = 1.0, s2 = 2.0, s3 = 2.0, s4 = 2.0, s5 = 2.0, s6 = 2.0, s7 = 2.0;
  int ns = 2, no = 0;
  double o0lo = 0, o0hi = 0, o1lo = 0, o1hi = 0;
  bool o0same = false, o1same = false, o0real = false, o1real = false;
  int o0idx = 0, o1idx = 0;
  bool spill = false;
  auto insert = [&](double v) {   // insertion into the sorted slots: b0 = min(a0, v), bj = max(a(j-1), min(aj, v))
    if (ns >= 8) { spill = true; return; }
    const double n7 = fmax(s6, fmin(s7, v)), n6 = fmax(s5, fmin(s6, v)), n5 = fmax(s4, fmin(s5, v)), n4 = fmax(s3, fmin(s4, v));
    const double n3 = fmax(s2, fmin(s3, v)), n2 = fmax(s1, fmin(s2, v)), n1 = fmax(s0, fmin(s1, v)), n0 = fmin(s0, v);
    s0 = n0; s1 = n1; s2 = n2; s3 = n3; s4 = n4; s5 = n5; s6 = n6; s7 = n7;
    ns++;
  };
  // Two passes per run of 64 list entries: the box test alone marks the entries whose box meets this edge's box grown
  // by tol (the first test of classify_edge, same expressions), then the classification runs for the marked entries
  // in list order. Every lane of a wave takes the second loop's iterations together, so the classification runs
  // about as many times as the most crowded lane has neighbours, not once per list entry whenever any lane has one.
  const int n = L.size();
  const double tx0 = fmin(px, qx) - c.tol, tx1 = fmax(px, qx) + c.tol, ty0 = fmin(py, qy) - c.tol, ty1 = fmax(py, qy) + c.tol;
  for (int base = 0; base < n && !spill; base += 64) {
    const int m = min(64, n - base);
    uint64_t cand = 0;
    for (int j = 0; j < m; j++) {
      Ed e;
      if (!L.get(base + j, e)) continue;
      if (!(fmax(e.x0, e.x1) < tx0 || fmin(e.x0, e.x1) > tx1 || fmax(e.y0, e.y1) < ty0 || fmin(e.y0, e.y1) > ty1)) cand |= 1ull << j;
    }
    while (cand != 0 && !spill) {
      const int k = base + __builtin_ctzll(cand);
      cand &= cand - 1;
      Ed e;
      L.get(k, e);
      double lo, hi, t;
      bool same;
      const int kind = classify_edge(e, px, py, qx, qy, dx, dy, len2, E.cut, c, lo, hi, same, t);
      if (kind == 1) {
        if (lo > 0) insert(lo);
        if (hi < 1) insert(hi);
        if (no == 0) { o0lo = lo; o0hi = hi; o0same = same; o0real = !e.syn; o0idx = k; }
        else if (no == 1) { o1lo = lo; o1hi = hi; o1same = same; o1real = !e.syn; o1idx = k; }
        else spill = true;
        no++;
      } else if (kind == 2) {
        insert(t);
      }
    }
  }
  if (spill) { out.bits = SPILL; return out; }
  const bool real = !E.syn;
  while (ns >= 2) {
    const double t0 = s0, t1 = s1;
    if (t1 > t0) {
      const double tm = 0.5 * (t0 + t1);
      int net = 0, net_real = 0, n_real = 0, first = -1;
      bool first_same = false;
      if (no > 0 && tm > o0lo && tm < o0hi) {
        const int sgn = o0same ? 1 : -1;
        net += sgn; first = o0idx; first_same = o0same;
        if (o0real) { net_real += sgn; n_real++; }
      }
      if (no > 1 && tm > o1lo && tm < o1hi) {
        const int sgn = o1same ? 1 : -1;
        net += sgn;
        if (first < 0) { first = o1idx; first_same = o1same; }
        if (o1real) { net_real += sgn; n_real++; }
      }
      subsegment<Src>(L, c, px, py, dx, dy, len2, real, t0, t1, net, net_real, n_real, first, first_same, out);
    }
    s0 = s1; s1 = s2; s2 = s3; s3 = s4; s4 = s5; s5 = s6; s6 = s7; s7 = 2.0;
    ns--;
  }
  return out;
}

__device__ __forceinline__ double pair_snap(double2 o, double2 e) {   // engine/engine.cuh:524-527
  const double m = fmax(fmax(fabs(o.x), fabs(o.x + e.x)), fmax(fabs(o.y), fabs(o.y + e.y)));
  return 0.5 * (nextafter(m, INFINITY) - m);
}

// a polygon paired with itself (a self-join's diagonal): both sides see the same edges in the same frame, so every
// per-edge result of side B equals side A's and is computed once
__device__ __forceinline__ bool is_self(const DPolys& P, const DPolys& R, uint32_t a, uint32_t b) { return a == b && P.xy == R.xy; }

struct PairSetup {
  Box ov;
  double tol, snap, cx, cy, bshx, bshy, ashx, ashy;
};

__device__ __forceinline__ PairSetup setup_pair_of(const double2 oa, const double2 ea, const double2 ob, const double2 eb, const RefineParams& rp) {
  PairSetup s;
  const double sx = ob.x - oa.x, sy = ob.y - oa.y;
  const double pad = 4.0 * fmax(rp.tol_abs, rp.tol_rel * fmax(fmax(ea.x, ea.y), fmax(eb.x, eb.y)));
  s.ov = Box{fmax(0.0, sx) - pad, fmax(0.0, sy) - pad, fmin(ea.x, sx + eb.x) + pad, fmin(ea.y, sy + eb.y) + pad};
  s.bshx = sx; s.bshy = sy;
  s.ashx = oa.x - oa.x; s.ashy = oa.y - oa.y;   // A in its own frame
  s.tol = fmax(rp.tol_abs, rp.tol_rel * fmax(s.ov.x1 - s.ov.x0, s.ov.y1 - s.ov.y0));
  s.snap = rp.snap_half_ulp ? pair_snap(oa, ea) : 0.0;
  s.cx = s.ov.x0 + 0.381966 * (s.ov.x1 - s.ov.x0);
  s.cy = s.ov.y0 + 0.414213 * (s.ov.y1 - s.ov.y0);
  return s;
}
__device__ __forceinline__ PairSetup setup_pair(const DPolys& P, const DPolys& R, uint32_t a, uint32_t b, const RefineParams& rp) {
  return setup_pair_of(P.origin[a], P.ext[a], R.origin[b], R.ext[b], rp);
}
// the vertex-list classes: origin and extent from the pieces' packed records
__device__ __forceinline__ PairSetup setup_pair_rec(const uint4* ra, const uint4* rb, const RefineParams& rp) {
  const double2* a = reinterpret_cast<const double2*>(ra);
  const double2* b = reinterpret_cast<const double2*>(rb);
  return setup_pair_of(a[0], a[1], b[0], b[1], rp);
}

// group epilogue: deterministic sums of the per-edge partials (stored at acc/lin/lon[0 .. n)), lane 0 writes
template <int G>
__device__ __forceinline__ void finish_pair(int gl, uint64_t p, bool hit, bool rtouch, bool computed, bool rel, int nsum, int na, const double* acc,
                                            const double* lin, const double* lon, uint32_t bits, const RefineParams& rp, const RefineOut& out) {
  double ar = 0.0, ai = 0.0, ao = 0.0, bi = 0.0, bo = 0.0;
  uint32_t eb4 = rtouch ? REL_TOUCH : 0u;
  if (computed) {
    double s = 0.0;
    if (gl == 0) { for (int k = 0; k < nsum; k++) s += acc[k]; }
    else if (rel && gl == 1) { for (int k = 0; k < na; k++) s += lin[k]; }
    else if (rel && gl == 2) { for (int k = 0; k < na; k++) s += lon[k]; }
    else if (rel && gl == 3) { for (int k = na; k < nsum; k++) s += lin[k]; }
    else if (rel && gl == 4) { for (int k = na; k < nsum; k++) s += lon[k]; }
    ar = 0.5 * __shfl(s, 0, G);
    if (rel) { ai = __shfl(s, 1, G); ao = __shfl(s, 2, G); bi = __shfl(s, 3, G); bo = __shfl(s, 4, G); }
    eb4 |= (gany<G>(bits & REL_COIN_SAME) ? REL_COIN_SAME : 0u) | (gany<G>(bits & REL_COIN_OPP) ? REL_COIN_OPP : 0u) |
           (gany<G>(bits & REL_TOUCH) ? REL_TOUCH : 0u);
  }
  if (gl == 0) {
    const bool ev_any = ai != 0 || ao != 0 || bi != 0 || bo != 0;
    const uint8_t v = hit ? 1 : ((ar != 0.0 || ev_any) ? 2 : 0);
    out.isect[p] = v;
    if (v) {   // compaction reads a pair's area and evidence only where its flag is set
      if (out.area) out.area[p] = ar;
      if (out.ev) { out.ev[p] = make_double4(ai, ao, bi, bo); out.evb[p] = eb4; }
    }
  }
}

// ------------------------------------------------------------------------------------------- small class
template <int CAP>
struct SmallPoly {        // a polygon of at most CAP stored vertices (<= 255) and RCAP rings
  double2 V[CAP];
  uint8_t F[CAP];
  uint8_t I[CAP];         // edges meeting the overlap box, ascending (the reference's compacted edge list)
  uint8_t rs[RCAP + 1];   // ring start slots, rs[nrings] = vertex count
  uint8_t nv, nr, ni;
};
template <int CAP>
constexpr size_t small_bytes(bool rel) { return ((2 * sizeof(SmallPoly<CAP>) + 15) & ~size_t(15)) + (size_t)2 * CAP * 8 * (rel ? 3 : 1); }

// parity of the segment from a point left of the geometry's bbox to m over every edge (engine/engine.cuh:382-397)
template <int G, int CAP>
__device__ bool pip_small(const SmallPoly<CAP>& sp, double shx, double mx, double my, int gl) {
  const double ox = shx - 1.0, oy = my;
  bool par = false;
  for (int k = gl; k < sp.nv; k += G) {
    if (!(sp.F[k] & F_VALID)) continue;
    const double2 a = sp.V[k], b = sp.V[k + 1];
    const double x0 = a.x, y0 = a.y, x1 = b.x, y1 = b.y;
    if ((y0 > my) == (y1 > my) && y0 != my && y1 != my) continue;
    const bool so = left_of(x0, y0, x1, y1, ox, oy), sc = left_of(x0, y0, x1, y1, mx, my);
    if (so == sc) continue;
    const bool su = orient(ox, oy, mx, my, x0, y0) > 0, sv = orient(ox, oy, mx, my, x1, y1) > 0;
    if (su != sv) par = !par;
  }
  return (__popcll(gballot<G>(par)) & 1) != 0;
}

// containment probe of ring r: first vertex that starts a real edge, else the first vertex (engine/engine.cuh:322-327);
// returns 0 for an empty ring (no probe), 1 for a probe on a synthetic-only ring, 2 for a real probe
template <int CAP>
__device__ __forceinline__ int probe_small(const SmallPoly<CAP>& sp, int r, bool has_syn, double& x, double& y) {
  const int s = sp.rs[r], e = sp.rs[r + 1];
  if (s == e) return 0;
  int v = s;
  bool real = true;
  if (has_syn) {
    while (v + 1 < e && (sp.F[v] & F_SYN)) v++;
    real = v + 1 < e;
  }
  const double2 p = sp.V[real ? v : s];
  x = p.x; y = p.y;
  return real ? 2 : 1;
}

// Vertex-list classes: a G-lane group per pair, both polygons in LDS (at most CAP stored vertices each); class
// WHICH of the partition (0 small, 1 medium). Two kernels: the screen decides contact and containment for every pair
// and finishes the pairs that need no edge contributions; the pairs that do are queued for the edge kernel, so that
// its waves hold only pairs with edge work (the screen's verdict travels in isect[p]).
constexpr uint8_t ST_HIT = 1, ST_RTOUCH = 2, ST_INA = 4, ST_INB = 8;

struct EdgeQueue {   // pairs with edge work: flag[k] per position k of the class list (set by the screen), then the
  uint32_t* flag;    // flagged pairs in class-list order: class c's at list[first[c] .. first[c] + count[c])
  uint64_t* off;
  uint32_t* list;
  uint32_t* count;   // [3]: tiny, small, medium
  uint32_t* first;   // [3]
};

// both polygons of a pair (B only when it is not A itself) into LDS with per-vertex flags (reference edge collection,
// engine/engine.cuh:358-377), read from their packed records (DPolys::pk): the headers' counts and ring starts, the
// vertices, their flag bytes. Every global load of both records is issued into registers before the first LDS write,
// so the pair pays one memory latency for its vertices instead of four (A's vertices, A's flags, B's vertices, B's
// flags, one polygon after the other), and three wave barriers instead of six.
template <int G, int CAP>
__device__ __forceinline__ void load_pair(const DPolys& P, const DPolys& R, uint32_t ra, uint32_t rb, const PairSetup& f, SmallPoly<CAP>& SA,
                                          SmallPoly<CAP>& SB, int gl, bool self) {
  constexpr int PER = (CAP + G - 1) / G;   // vertices per lane and polygon
  const uint4* reca = P.pk + ra;
  const uint4* recb = R.pk + rb;
  const uint8_t* ha = reinterpret_cast<const uint8_t*>(reca + 2);
  const uint8_t* hb = reinterpret_cast<const uint8_t*>(recb + 2);
  const int nva = ha[0], nra = ha[1];
  const int nvb = self ? 0 : hb[0], nrb = self ? 0 : hb[1];
  const double2* VA = reinterpret_cast<const double2*>(reca + 3);
  const double2* VB = reinterpret_cast<const double2*>(recb + 3);
  const uint8_t* FA = reinterpret_cast<const uint8_t*>(reca + 3 + nva);
  const uint8_t* FB = reinterpret_cast<const uint8_t*>(recb + 3 + nvb);
  double2 va[PER], vb[PER];
  uint8_t fa[PER], fb[PER];
#pragma unroll
  for (int j = 0; j < PER; j++) {
    const int k = gl + j * G;
    if (k < nva) { va[j] = VA[k]; fa[j] = FA[k]; }
    if (k < nvb) { vb[j] = VB[k]; fb[j] = FB[k]; }
  }
#pragma unroll
  for (int j = 0; j < PER; j++) {
    const int k = gl + j * G;
    if (k < nva) SA.V[k] = make_double2(f.ashx + va[j].x, f.ashy + va[j].y);
    if (k < nvb) SB.V[k] = make_double2(f.bshx + vb[j].x, f.bshy + vb[j].y);
  }
  for (int r = gl; r <= nra; r += G) SA.rs[r] = ha[2 + r];
  if (!self)
    for (int r = gl; r <= nrb; r += G) SB.rs[r] = hb[2 + r];
  if (gl == 0) {
    SA.nv = (uint8_t)nva; SA.nr = (uint8_t)nra;
    if (!self) { SB.nv = (uint8_t)nvb; SB.nr = (uint8_t)nrb; }
  }
  wave_lds_sync();
  // per-vertex flags: valid (not the ring's last vertex), the record's flag bits, meets the overlap box
  auto flag = [&](const SmallPoly<CAP>& sp, int k, int nrings, uint8_t rf) -> uint8_t {
    int r = 0;
    while (r + 1 < nrings && sp.rs[r + 1] <= k) r++;
    if (k + 1 >= sp.rs[r + 1]) return 0;
    uint8_t fl = F_VALID | rf;
    const double2 a = sp.V[k], b = sp.V[k + 1];
    if (!edge_outside(a.x, a.y, b.x, b.y, f.ov)) fl |= F_INOV;
    return fl;
  };
#pragma unroll
  for (int j = 0; j < PER; j++) {
    const int k = gl + j * G;
    if (k < nva) SA.F[k] = flag(SA, k, nra, fa[j]);
    if (k < nvb) SB.F[k] = flag(SB, k, nrb, fb[j]);
  }
  wave_lds_sync();
  // compacted lists of the edges meeting the overlap box, in vertex order (group ballot)
  int cnta = 0, cntb = 0;
  for (int base = 0; base < max(nva, nvb); base += G) {
    const int k = base + gl;
    const bool ina = k < nva && (SA.F[k] & (F_VALID | F_INOV)) == (F_VALID | F_INOV);
    const bool inb = k < nvb && (SB.F[k] & (F_VALID | F_INOV)) == (F_VALID | F_INOV);
    const uint64_t ma = gballot<G>(ina), mb = gballot<G>(inb);
    const uint64_t below = (1ull << gl) - 1;
    if (ina) SA.I[cnta + __popcll(ma & below)] = (uint8_t)k;
    if (inb) SB.I[cntb + __popcll(mb & below)] = (uint8_t)k;
    cnta += __popcll(ma);
    cntb += __popcll(mb);
  }
  if (gl == 0) {
    SA.ni = (uint8_t)cnta;
    if (!self) SB.ni = (uint8_t)cntb;
  }
  wave_lds_sync();
}

template <int G, int CAP, int WHICH, int OCC>
__global__ void __launch_bounds__(256, OCC) k_screen_small(const DPolys P, const DPolys R, const uint32_t* __restrict__ pa, const uint32_t* __restrict__ pb,
                                                      const uint32_t* __restrict__ list, const uint32_t* __restrict__ counts, RefineParams rp,
                                                      RefineOut out, EdgeQueue q) {
  extern __shared__ __align__(16) unsigned char smem[];
  using SP = SmallPoly<CAP>;
  const int gpb = blockDim.x / G;
  const int gib = threadIdx.x / G, gl = threadIdx.x % G;
  const bool rel = rp.want_relate != 0;
  SP& SA = *reinterpret_cast<SP*>(smem + (size_t)gib * 2 * sizeof(SP));
  SP& SB = *(&SA + 1);
  const uint32_t nlist = counts[WHICH];
  const uint32_t base = class_first(counts, WHICH);
  const uint32_t* cl = list + base;
  PairStream ps(cl, pa, pb, nlist, blockIdx.x * gpb + gib, gridDim.x * gpb);
  for (; ps.k < nlist; ps.advance()) {
    ps.prefetch();
    const uint32_t k = ps.k, p = ps.p, a = ps.a, b = ps.b;
    const uint32_t ra = P.pk_off[a], rb = R.pk_off[b];   // record offsets (32-bit: fewer registers than pointers)
    const PairSetup f = setup_pair_rec(P.pk + ra, R.pk + rb, rp);
    const bool self = is_self(P, R, a, b);
    load_pair<G, CAP>(P, R, ra, rb, f, SA, SB, gl, self);
    const SP& SBx = self ? SA : SB;
    const int na = SA.ni, nb = SBx.ni;   // edges meeting the overlap box
    const VertSrc LA{SA.V, SA.F, SA.I, na}, LB{SBx.V, SBx.F, SBx.I, nb};
    // boundary contact over the A x B edge grid
    bool hit = false, hit_any = false;
    const int grid = na * nb;
    for (int q0 = 0; q0 < grid; q0 += G) {
      const int qi = q0 + gl;
      if (qi < grid) {
        Ed ea, eb;
        LA.get(qi / nb, ea);
        LB.get(qi % nb, eb);
        if (!(fmax(eb.x0, eb.x1) < fmin(ea.x0, ea.x1) || fmin(eb.x0, eb.x1) > fmax(ea.x0, ea.x1) || fmax(eb.y0, eb.y1) < fmin(ea.y0, ea.y1) ||
              fmin(eb.y0, eb.y1) > fmax(ea.y0, ea.y1)) &&
            seg_isect(ea.x0, ea.y0, ea.x1, ea.y1, eb.x0, eb.y0, eb.x1, eb.y1)) {
          hit_any = true;
          if (!(ea.syn | eb.syn)) hit = true;
        }
      }
      if (gany<G>(hit)) break;
    }
    hit = gany<G>(hit);
    hit_any = gany<G>(hit_any || hit);
    const bool rtouch = hit;
    bool inB_c = false, inA_c = false;
    if (!hit || rp.want_area || rel) {
      inA_c = pip_small<G, CAP>(SA, f.ashx, f.cx, f.cy, gl);
      inB_c = self ? inA_c : pip_small<G, CAP>(SB, f.bshx, f.cx, f.cy, gl);
    }
    bool contain_any = false;
    if (!hit) {
      bool h = false, ha = false;
      const int nra = SA.nr, nrb = SBx.nr;
      for (int r = gl; r < nra + nrb; r += G) {
        const bool sideA = r < nra;
        double x, y;
        const int pr = probe_small(sideA ? SA : SBx, sideA ? r : r - nra, (sideA ? P.synb : R.synb) != nullptr, x, y);
        const bool real = pr == 2;
        if (pr && x >= f.ov.x0 && x <= f.ov.x1 && y >= f.ov.y0 && y <= f.ov.y1 &&
            inside_from_c(sideA ? LB : LA, f.cx, f.cy, sideA ? inB_c : inA_c, x, y)) {
          ha = true;
          if (real) h = true;
        }
      }
      hit = gany<G>(h);
      contain_any = gany<G>(ha);
    }
    const bool computed = (hit || hit_any || contain_any) && (rp.want_area || rel) && na + nb > 0;
    if (!computed) {
      finish_pair<G>(gl, p, hit, rtouch, false, rel, 0, 0, nullptr, nullptr, nullptr, 0u, rp, out);
    } else if (gl == 0) {
      out.isect[p] = (hit ? ST_HIT : 0) | (rtouch ? ST_RTOUCH : 0) | (inA_c ? ST_INA : 0) | (inB_c ? ST_INB : 0);
    }
    if (gl == 0) q.flag[base + k] = computed ? 1u : 0u;
    wave_lds_sync();
  }
}

template <int G, int CAP, int WHICH>
__global__ void __launch_bounds__(256, 3) k_edges_small(const DPolys P, const DPolys R, const uint32_t* __restrict__ pa, const uint32_t* __restrict__ pb,
                                                        const uint32_t* __restrict__ counts, RefineParams rp, RefineOut out, EdgeQueue q,
                                                        uint32_t* __restrict__ ovf_list, OverflowCtl* ctl) {
  extern __shared__ __align__(16) unsigned char smem[];
  using SP = SmallPoly<CAP>;
  const int gpb = blockDim.x / G;
  const int gib = threadIdx.x / G, gl = threadIdx.x % G;
  const bool rel = rp.want_relate != 0;
  unsigned char* gbase = smem + (size_t)gib * small_bytes<CAP>(rel);
  SP& SA = *reinterpret_cast<SP*>(gbase);
  SP& SB = *(reinterpret_cast<SP*>(gbase) + 1);
  double* acc = reinterpret_cast<double*>(gbase + ((2 * sizeof(SP) + 15) & ~size_t(15)));
  double* lin = acc + 2 * CAP;
  double* lon = lin + 2 * CAP;
  const uint32_t nlist = q.count[WHICH];
  const uint32_t* cl = q.list + q.first[WHICH];
  PairStream ps(cl, pa, pb, nlist, blockIdx.x * gpb + gib, gridDim.x * gpb);
  for (; ps.k < nlist; ps.advance()) {
    ps.prefetch();
    const uint32_t p = ps.p, a = ps.a, b = ps.b;
    const uint8_t st = out.isect[p];
    const uint32_t ra = P.pk_off[a], rb = R.pk_off[b];   // record offsets (32-bit: fewer registers than pointers)
    const PairSetup f = setup_pair_rec(P.pk + ra, R.pk + rb, rp);
    const bool self = is_self(P, R, a, b);
    load_pair<G, CAP>(P, R, ra, rb, f, SA, SB, gl, self);
    const SP& SBx = self ? SA : SB;
    const int na = SA.ni, nb = SBx.ni;
    const VertSrc LA{SA.V, SA.F, SA.I, na}, LB{SBx.V, SBx.F, SBx.I, nb};
    EdgeCtx ca{f.cx, f.cy, f.tol, rp.eps_par, f.snap, f.ov, (st & ST_INB) != 0, rel}, cb = ca;
    cb.in_c = (st & ST_INA) != 0;
    uint32_t bits = 0;
    for (int s = gl; s < (self ? na : na + nb); s += G) {
      // one call with the side chosen per lane: a wave holds A- and B-edge lanes in the same round, and two calls
      // would run one after the other. A polygon with itself: B's edge s computes exactly A's (same edge, same list,
      // same containment of c), so A's result fills both slots.
      const bool sideA = s < na;
      Ed e;
      (sideA ? LA : LB).get(sideA ? s : s - na, e);
      const Contrib c = edge_contrib<VertSrc, false>(e, sideA ? LB : LA, sideA ? ca : cb);
      acc[s] = c.acc;
      if (rel) { lin[s] = c.lin; lon[s] = c.lon; }
      if (self) {
        acc[na + s] = c.acc;
        if (rel) { lin[na + s] = c.lin; lon[na + s] = c.lon; }
      }
      bits |= c.bits;
    }
    wave_lds_sync();
    if (gany<G>(bits & SPILL)) {
      if (gl == 0) queue_overflow(p, max(na, nb), ovf_list, ctl);
    } else {
      finish_pair<G>(gl, p, (st & ST_HIT) != 0, (st & ST_RTOUCH) != 0, true, rel, na + nb, na, acc, lin, lon, bits, rp, out);
    }
    wave_lds_sync();
  }
}

// the flagged pairs of the vertex-list classes (tiny, small, medium) in class-list order (stable compaction of the
// screens' flags)
__global__ void k_queue_scatter(const uint32_t* __restrict__ list, const uint32_t* __restrict__ counts, EdgeQueue q) {
  const uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const uint64_t n = (uint64_t)counts[CL_TINY] + counts[CL_SMALL] + counts[CL_MEDIUM];
  if (k == 0) {
    uint64_t lo = 0;
    for (int cl = CL_TINY; cl <= CL_MEDIUM; cl++) {
      const uint64_t hi = lo + counts[cl];
      q.first[cl] = (uint32_t)q.off[lo];
      q.count[cl] = (uint32_t)(q.off[hi] - q.off[lo]);
      lo = hi;
    }
  }
  if (k < n && q.flag[k]) q.list[q.off[k]] = list[k];
}

// ------------------------------------------------------------------------------------------- large / overflow
// Per-group scratch for compacted lists and per-edge partial sums (LDS in the large kernel, global in overflow).
struct Scratch {
  Ed* LA; Ed* LB;
  double* acc;   // [2 cap] area terms: A edges then B edges
  double* lin;   // [2 cap] inside lengths (relate only)
  double* lon;   // [2 cap] on lengths (relate only)
};
__device__ __forceinline__ Scratch carve(unsigned char* base, int cap, bool rel) {
  Scratch s;
  s.LA = reinterpret_cast<Ed*>(base);
  s.LB = s.LA + cap;
  s.acc = reinterpret_cast<double*>(s.LB + cap);
  s.lin = rel ? s.acc + 2 * cap : nullptr;
  s.lon = rel ? s.lin + 2 * cap : nullptr;
  return s;
}
__host__ __device__ constexpr size_t list_bytes(int cap, bool rel) { return (size_t)cap * (2 * 40 + 2 * 8 + (rel ? 4 * 8 : 0)); }

// returns true (group-uniform) when an edge spilled the register path: nothing is written for the pair then
template <int G, bool CAPPED>
__device__ bool refine_lists(const DPolys& P, const DPolys& R, uint32_t a, uint32_t b, uint64_t p, const RefineParams& rp, const PairSetup& f,
                             const Scratch& S, int na, int nb, int gl, const RefineOut& out) {
  const ListSrc LA{S.LA, na}, LB{S.LB, nb};
  bool hit = false, hit_any = false;
  for (int i = gl; i < na && !hit; i += G) {
    const Ed ea = S.LA[i];
    const double ax0 = fmin(ea.x0, ea.x1), ax1 = fmax(ea.x0, ea.x1), ay0 = fmin(ea.y0, ea.y1), ay1 = fmax(ea.y0, ea.y1);
    for (int j = 0; j < nb; j++) {
      const Ed eb = S.LB[j];
      if (fmax(eb.x0, eb.x1) < ax0 || fmin(eb.x0, eb.x1) > ax1 || fmax(eb.y0, eb.y1) < ay0 || fmin(eb.y0, eb.y1) > ay1) continue;
      if (seg_isect(ea.x0, ea.y0, ea.x1, ea.y1, eb.x0, eb.y0, eb.x1, eb.y1)) {
        hit_any = true;
        if (!(ea.syn | eb.syn)) { hit = true; break; }
      }
    }
  }
  hit = gany<G>(hit);
  hit_any = gany<G>(hit_any || hit);
  const bool rtouch = hit;
  const bool rel = rp.want_relate != 0;
  const bool self = is_self(P, R, a, b);
  bool inB_c = false, inA_c = false;
  if (!hit || rp.want_area || rel) {
    inA_c = pip_full<G>(P, a, f.ashx, f.ashy, f.cx, f.cy, gl);
    inB_c = self ? inA_c : pip_full<G>(R, b, f.bshx, f.bshy, f.cx, f.cy, gl);
  }
  bool contain_any = false;
  if (!hit) {
    bool h = false, ha = false;
    for (uint32_t r = P.gro[a] + gl; r < P.gro[a + 1] && !h; r += G) {
      if (P.rvo[r] == P.rvo[r + 1]) continue;   // an empty ring has no probe vertex
      double x, y;
      const bool real = probe_vertex(P, r, f.ashx, f.ashy, x, y);
      if (!real) { const double2 v = P.xy[P.rvo[r]]; x = f.ashx + v.x; y = f.ashy + v.y; }
      if (x >= f.ov.x0 && x <= f.ov.x1 && y >= f.ov.y0 && y <= f.ov.y1 && inside_from_c(LB, f.cx, f.cy, inB_c, x, y)) { ha = true; if (real) h = true; }
    }
    for (uint32_t r = R.gro[b] + gl; r < R.gro[b + 1] && !h; r += G) {
      if (R.rvo[r] == R.rvo[r + 1]) continue;   // an empty ring has no probe vertex
      double x, y;
      const bool real = probe_vertex(R, r, f.bshx, f.bshy, x, y);
      if (!real) { const double2 v = R.xy[R.rvo[r]]; x = f.bshx + v.x; y = f.bshy + v.y; }
      if (x >= f.ov.x0 && x <= f.ov.x1 && y >= f.ov.y0 && y <= f.ov.y1 && inside_from_c(LA, f.cx, f.cy, inA_c, x, y)) { ha = true; if (real) h = true; }
    }
    hit = gany<G>(h);
    contain_any = gany<G>(ha);
  }
  const bool computed = (hit || hit_any || contain_any) && (rp.want_area || rel);
  uint32_t bits = 0;
  if (computed) {
    EdgeCtx ca{f.cx, f.cy, f.tol, rp.eps_par, f.snap, f.ov, inB_c, rel}, cb = ca;
    cb.in_c = inA_c;
    for (int s = gl; s < (self ? na : na + nb); s += G) {   // self: A's results fill B's slots too
      const bool sideA = s < na;
      const Ed e = (sideA ? S.LA : S.LB)[sideA ? s : s - na];
      const Contrib c = edge_contrib<ListSrc, CAPPED>(e, ListSrc{sideA ? S.LB : S.LA, sideA ? nb : na}, sideA ? ca : cb);
      S.acc[s] = c.acc;
      if (rel) { S.lin[s] = c.lin; S.lon[s] = c.lon; }
      if (self) {
        S.acc[na + s] = c.acc;
        if (rel) { S.lin[na + s] = c.lin; S.lon[na + s] = c.lon; }
      }
      bits |= c.bits;
    }
    wave_lds_sync();
    if (!CAPPED && gany<G>(bits & SPILL)) return true;
  }
  finish_pair<G>(gl, p, hit, rtouch, computed, rel, na + nb, na, S.acc, S.lin, S.lon, bits, rp, out);
  return false;
}

// large class: a wave per pair, compacted lists in LDS (cap edges per side); overflowing pairs are queued. Its pairs
// are the partition's last class followed by the mixed pairs promoted because their lists outgrew the mixed class.
__global__ void __launch_bounds__(256, 3) k_refine_large(const DPolys P, const DPolys R, const uint32_t* __restrict__ pa, const uint32_t* __restrict__ pb,
                                                      const uint32_t* __restrict__ list, const uint32_t* __restrict__ counts, const uint32_t* __restrict__ promo,
                                                      const OverflowCtl* __restrict__ promo_ctl, RefineParams rp, RefineOut out, int cap,
                                                      uint32_t* __restrict__ ovf_list, OverflowCtl* ctl) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int wpb = blockDim.x / 64;
  const int wib = threadIdx.x / 64, gl = threadIdx.x % 64;
  const bool rel = rp.want_relate != 0;
  const Scratch S = carve(smem + (size_t)wib * list_bytes(cap, rel), cap, rel);
  const uint32_t nclass = counts[CL_LARGE], nlist = nclass + promo_ctl->count;
  const uint32_t* large = list + class_first(counts, CL_LARGE);   // the large class comes last
  for (uint32_t k = blockIdx.x * wpb + wib; k < nlist; k += gridDim.x * wpb) {
    const uint32_t p = k < nclass ? large[k] : promo[k - nclass];
    const uint32_t a = pa[p], b = pb[p];
    const PairSetup f = setup_pair(P, R, a, b, rp);
    const int na = collect<64>(P, a, f.ashx, f.ashy, f.ov, S.LA, cap, gl);
    const int nb = collect<64>(R, b, f.bshx, f.bshy, f.ov, S.LB, cap, gl);
    if (na > cap || nb > cap) {
      if (gl == 0) queue_overflow(p, max(na, nb), ovf_list, ctl);
      continue;
    }
    wave_lds_sync();
    if (refine_lists<64, false>(P, R, a, b, p, rp, f, S, na, nb, gl, out) && gl == 0) queue_overflow(p, max(na, nb), ovf_list, ctl);
    wave_lds_sync();
  }
}

// mixed class (one polygon small, the other large): a 16-lane group per pair, compacted lists of CAP_MIXED edges per
// side in LDS (the large polygon's edges meeting the overlap box, found through its edge-chunk boxes); pairs whose
// lists overflow are queued
constexpr int CAP_MIXED = 32, MIXED_T = 128;
constexpr uint32_t OVF_LDS_CAP = 448;   // overflow lists up to this many edges per side go to LDS (448 x 128 B = 56 KiB)
constexpr uint32_t OVF_SIDE_CAP = 256;  // the list classes' own overflow launch (ovf_lds 2): LDS lists of up to 256 edges
__global__ void __launch_bounds__(MIXED_T) k_refine_mixed(const DPolys P, const DPolys R, const uint32_t* __restrict__ pa, const uint32_t* __restrict__ pb,
                                                          const uint32_t* __restrict__ list, const uint32_t* __restrict__ counts, RefineParams rp,
                                                          RefineOut out, uint32_t* __restrict__ promo, OverflowCtl* promo_ctl,
                                                          uint32_t* __restrict__ ovf_list, OverflowCtl* ctl) {
  extern __shared__ __align__(16) unsigned char smem[];
  constexpr int G = 16;
  const int gpb = blockDim.x / G;
  const int gib = threadIdx.x / G, gl = threadIdx.x % G;
  const bool rel = rp.want_relate != 0;
  const Scratch S = carve(smem + (size_t)gib * list_bytes(CAP_MIXED, rel), CAP_MIXED, rel);
  const uint32_t nlist = counts[CL_MIXED];
  const uint32_t* cl = list + class_first(counts, CL_MIXED);
  for (uint32_t k = blockIdx.x * gpb + gib; k < nlist; k += gridDim.x * gpb) {
    const uint32_t p = cl[k];
    const uint32_t a = pa[p], b = pb[p];
    const PairSetup f = setup_pair(P, R, a, b, rp);
    const int na = collect<G>(P, a, f.ashx, f.ashy, f.ov, S.LA, CAP_MIXED, gl);
    const int nb = collect<G>(R, b, f.bshx, f.bshy, f.ov, S.LB, CAP_MIXED, gl);
    if (na > CAP_MIXED || nb > CAP_MIXED) {   // to the large class (which sends it on if its lists overflow too)
      if (gl == 0) queue_overflow(p, max(na, nb), promo, promo_ctl);
      continue;
    }
    wave_lds_sync();
    if (refine_lists<G, false>(P, R, a, b, p, rp, f, S, na, nb, gl, out) && gl == 0) queue_overflow(p, max(na, nb), ovf_list, ctl);
    wave_lds_sync();
  }
}

// overflow: a block of T threads per pair (T = 64: one wave), lists and partial sums in LDS when the launch's largest
// lists fit (IN_LDS; the edge loops then read LDS instead of global memory), else in a per-block slab of global
// memory, array path for every edge. The first wave collects the lists and runs the steps linear in the list sizes
// (containment of c, the ring probes) and the ordered sums; the whole block runs the contact grid and the edge
// contributions, whose work grows with the product of the list sizes. T is chosen per launch from the number of pairs: four
// waves per pair when the pairs are too few to fill the GPU (a windowed request's few largest pairs), one wave per pair
// otherwise. Every barrier sits in block-uniform control flow (the loop bounds and branches around them depend only on
// the pair and on values the block has read back from shared memory after a barrier).
// With `dev_n` the queue length is read on the device (a launch enqueued behind the kernels that fill the queue, with
// no host read in between), and with `defer` a pair whose lists exceed `cap` is handed on to that queue (its count
// of such pairs kept in dev_n->pad0) instead of being refined here.
template <int T, bool IN_LDS>
__global__ void __launch_bounds__(T) k_refine_overflow(const DPolys P, const DPolys R, const uint32_t* __restrict__ pa, const uint32_t* __restrict__ pb,
                                                           const uint32_t* __restrict__ list, uint32_t nlist, RefineParams rp, RefineOut out, int cap,
                                                           unsigned char* __restrict__ slab, OverflowCtl* dev_n, uint32_t* __restrict__ defer_list,
                                                           OverflowCtl* defer) {
  extern __shared__ __align__(16) unsigned char smem[];
  __shared__ int s_na, s_nb, s_hit, s_hit_any, s_found, s_contain, s_inA, s_inB;   // s_hit: contact; s_found: contact or containment
  __shared__ int s_defer;
  __shared__ uint32_t s_bits;
  const int t = threadIdx.x, gl = t & 63, w = t >> 6;
  const bool rel = rp.want_relate != 0;
  const Scratch S = carve(IN_LDS ? smem : slab + (uint64_t)blockIdx.x * list_bytes(cap, rel), cap, rel);
  if (dev_n) nlist = dev_n->count;
  for (uint32_t k = blockIdx.x; k < nlist; k += gridDim.x) {
    const uint32_t p = list[k];
    const uint32_t a = pa[p], b = pb[p];
    const PairSetup f = setup_pair(P, R, a, b, rp);
    if (w == 0) {
      const int na = collect<64>(P, a, f.ashx, f.ashy, f.ov, S.LA, cap, gl);
      const int nb = collect<64>(R, b, f.bshx, f.bshy, f.ov, S.LB, cap, gl);
      if (gl == 0) {
        s_na = min(na, cap); s_nb = min(nb, cap); s_hit = 0; s_hit_any = 0; s_contain = 0; s_bits = 0;
        s_defer = defer && (na > cap || nb > cap);
        if (s_defer) { queue_overflow(p, max(na, nb), defer_list, defer); atomicAdd(&dev_n->pad0, 1u); }
      }
    }
    __syncthreads();
    const int na = s_na, nb = s_nb;
    if (s_defer) { __syncthreads(); continue; }   // every wave has read s_defer before the next pair rewrites it
    // boundary contact over the A x B edge grid (engine/engine.cuh:378-381): a real contact anywhere, else any contact.
    // The grid is cut into T strips of whole rows when A has T edges or more, else each row into T / na pieces, so
    // that every thread of the block gets a share however few edges A has.
    {
      bool hit = false, hit_any = false;
      const int parts = na >= T || na == 0 ? 1 : T / na;   // pieces per row
      const int chunk = (nb + parts - 1) / parts;
      for (int r = t; r < na * parts && !hit; r += T) {
        const int i = r / parts, j0 = (r % parts) * chunk, j1 = min(nb, j0 + chunk);
        const Ed ea = S.LA[i];
        const double ax0 = fmin(ea.x0, ea.x1), ax1 = fmax(ea.x0, ea.x1), ay0 = fmin(ea.y0, ea.y1), ay1 = fmax(ea.y0, ea.y1);
        for (int j = j0; j < j1; j++) {
          const Ed eb = S.LB[j];
          if (fmax(eb.x0, eb.x1) < ax0 || fmin(eb.x0, eb.x1) > ax1 || fmax(eb.y0, eb.y1) < ay0 || fmin(eb.y0, eb.y1) > ay1) continue;
          if (seg_isect(ea.x0, ea.y0, ea.x1, ea.y1, eb.x0, eb.y0, eb.x1, eb.y1)) {
            hit_any = true;
            if (!(ea.syn | eb.syn)) { hit = true; break; }
          }
        }
      }
      if (hit) atomicOr(&s_hit, 1);
      if (hit_any) atomicOr(&s_hit_any, 1);
    }
    __syncthreads();
    const bool rtouch = s_hit != 0;
    // containment of c and the ring probes: the first wave
    if (w == 0) {
      bool hit = rtouch;
      bool inB_c = false, inA_c = false;
      if (!hit || rp.want_area || rel) {
        inA_c = pip_full<64>(P, a, f.ashx, f.ashy, f.cx, f.cy, gl);
        inB_c = is_self(P, R, a, b) ? inA_c : pip_full<64>(R, b, f.bshx, f.bshy, f.cx, f.cy, gl);
      }
      bool contain_any = false;
      if (!hit) {
        const ListSrc LA{S.LA, na}, LB{S.LB, nb};
        bool h = false, ha = false;
        for (uint32_t r = P.gro[a] + gl; r < P.gro[a + 1] && !h; r += 64) {
          if (P.rvo[r] == P.rvo[r + 1]) continue;
          double x, y;
          const bool real = probe_vertex(P, r, f.ashx, f.ashy, x, y);
          if (!real) { const double2 v = P.xy[P.rvo[r]]; x = f.ashx + v.x; y = f.ashy + v.y; }
          if (x >= f.ov.x0 && x <= f.ov.x1 && y >= f.ov.y0 && y <= f.ov.y1 && inside_from_c(LB, f.cx, f.cy, inB_c, x, y)) { ha = true; if (real) h = true; }
        }
        for (uint32_t r = R.gro[b] + gl; r < R.gro[b + 1] && !h; r += 64) {
          if (R.rvo[r] == R.rvo[r + 1]) continue;
          double x, y;
          const bool real = probe_vertex(R, r, f.bshx, f.bshy, x, y);
          if (!real) { const double2 v = R.xy[R.rvo[r]]; x = f.bshx + v.x; y = f.bshy + v.y; }
          if (x >= f.ov.x0 && x <= f.ov.x1 && y >= f.ov.y0 && y <= f.ov.y1 && inside_from_c(LA, f.cx, f.cy, inA_c, x, y)) { ha = true; if (real) h = true; }
        }
        hit = gany<64>(h);
        contain_any = gany<64>(ha);
      }
      if (gl == 0) { s_found = hit ? 1 : 0; s_contain = contain_any ? 1 : 0; s_inA = inA_c ? 1 : 0; s_inB = inB_c ? 1 : 0; }
    }
    __syncthreads();
    const bool hit = s_found != 0;
    const bool computed = (hit || s_hit_any != 0 || s_contain != 0) && (rp.want_area || rel);
    if (computed) {
      EdgeCtx ca{f.cx, f.cy, f.tol, rp.eps_par, f.snap, f.ov, s_inB != 0, rel}, cb = ca;
      cb.in_c = s_inA != 0;
      uint32_t bits = 0;
      const bool self = is_self(P, R, a, b);
      for (int s = t; s < (self ? na : na + nb); s += T) {   // self: A's results fill B's slots too
        const bool sideA = s < na;
        const Ed e = (sideA ? S.LA : S.LB)[sideA ? s : s - na];
        const Contrib c = edge_contrib<ListSrc, true>(e, ListSrc{sideA ? S.LB : S.LA, sideA ? nb : na}, sideA ? ca : cb);
        S.acc[s] = c.acc;
        if (rel) { S.lin[s] = c.lin; S.lon[s] = c.lon; }
        if (self) {
          S.acc[na + s] = c.acc;
          if (rel) { S.lin[na + s] = c.lin; S.lon[na + s] = c.lon; }
        }
        bits |= c.bits;
      }
      if (bits) atomicOr(&s_bits, bits);
    }
    __syncthreads();
    if (w == 0) finish_pair<64>(gl, p, hit, rtouch, computed, rel, na + nb, na, S.acc, S.lin, S.lon, s_bits, rp, out);
    __syncthreads();   // the lists and the shared votes are rewritten for the next pair
  }
}

}  // namespace

namespace {
// the three screens at one register budget (OCC waves per SIMD)
template <int OCC>
void launch_screens(uint32_t gt, uint32_t gs, uint32_t gm, hipStream_t st, const DPolys& P, const DPolys& R, const uint32_t* pa, const uint32_t* pb,
                    const uint32_t* list, const uint32_t* counts, const RefineParams& rp, const RefineOut& out, const EdgeQueue& eq) {
  constexpr int T = 256;
  hipLaunchKernelGGL((k_screen_small<8, TCAP, CL_TINY, OCC>), dim3(gt), dim3(T), (T / 8) * 2 * sizeof(SmallPoly<TCAP>), st, P, R, pa, pb, list, counts,
                     rp, out, eq);
  GSJ_LAUNCHED();
  hipLaunchKernelGGL((k_screen_small<16, VCAP, CL_SMALL, OCC>), dim3(gs), dim3(T), (T / 16) * 2 * sizeof(SmallPoly<VCAP>), st, P, R, pa, pb, list,
                     counts, rp, out, eq);
  GSJ_LAUNCHED();
  hipLaunchKernelGGL((k_screen_small<32, MCAP, CL_MEDIUM, OCC>), dim3(gm), dim3(T), (T / 32) * 2 * sizeof(SmallPoly<MCAP>), st, P, R, pa, pb, list,
                     counts, rp, out, eq);
  GSJ_LAUNCHED();
}
}  // namespace

uint64_t refine_pairs(Context& c, const DPolys& P, const DPolys& R, const uint32_t* pa, const uint32_t* pb, uint64_t np, const RefineParams& rp,
                      const RefineOut& out, gsj_join_stats* stats) {
  if (np == 0) return 0;
  GSJ_REQUIRE(np < (1ull << 32), GSJ_E_LIMIT, "refine: more than 2^32 pairs in one launch");
  hipStream_t st = c.stream;
  const bool rel = rp.want_relate != 0;
  const LaunchShape& sh = c.shape(np);
  const PairClasses pc = partition_pairs(c, P, R, pa, pb, np, sh.mcap, sh.mixed != 0, sh.tcap, sh.size_sort);
  uint32_t* list = pc.list;
  uint32_t* counts = pc.counts;
  OverflowCtl* ctl = pc.ctl;
  uint32_t* ovf = c.s_heavy.as<uint32_t>(3 * np + 192);
  uint32_t* promo = ovf + np + 64;   // mixed pairs promoted to the large class
  OverflowCtl* promo_ctl = reinterpret_cast<OverflowCtl*>(reinterpret_cast<char*>(ctl) + 64);
  // ovf_lds 2: the list classes' overflow pairs are refined on their own chain, right behind the large kernel, by a
  // launch that reads the queue length on the device; the vertex-list classes' spills and the pairs whose lists
  // exceed OVF_SIDE_CAP go to a second queue (ovf_main) that the host reads once both chains are done
  const bool side_ovf = sh.ovf_lds == 2;
  uint32_t* ovf_main = side_ovf ? promo + np + 64 : ovf;
  OverflowCtl* ctl_main = side_ovf ? reinterpret_cast<OverflowCtl*>(reinterpret_cast<char*>(ctl) + 16) : ctl;
  const int T = 256;
  const uint32_t cu = (uint32_t)c.cu_count;
  // The list classes (mixed, then large: the large kernel takes the pairs mixed promotes) depend only on the
  // partition. With sh.fork they run on the side stream while the vertex-list classes run here, so that a small
  // launch pays the longer of the two chains instead of their sum; both chains append to the overflow queue with
  // atomics, and every pair's outputs are written by the one kernel of its class.
  EdgeQueue eq;   // allocated before the fork: a growing buffer would synchronise the device mid-chain
  eq.flag = c.s_edgeq.as<uint32_t>(2 * np + 32);
  eq.list = eq.flag + np + 16;
  eq.off = c.s_tmp2.as<uint64_t>(np + 17);
  eq.count = counts + 8;
  eq.first = eq.count + 3;
  hipStream_t sl = st;
  if (sh.fork) {
    sl = c.side;
    GSJ_HIP(hipEventRecord(c.ev[10], st));
    GSJ_HIP(hipStreamWaitEvent(sl, c.ev[10], 0));
  }
  GSJ_HIP(hipEventRecord(c.ev[5], sl));
  {   // mixed: 16-lane groups with edge lists, 8 pairs per block
    const uint32_t g = std::min<uint32_t>(div_up(np, MIXED_T / 16), cu * sh.mixed_bpc);
    hipLaunchKernelGGL(k_refine_mixed, dim3(g), dim3(MIXED_T), (MIXED_T / 16) * list_bytes(CAP_MIXED, rel), sl, P, R, pa, pb, list, counts, rp, out,
                       promo, promo_ctl, ovf, ctl);
    GSJ_LAUNCHED();
  }
  GSJ_HIP(hipEventRecord(c.ev[7], sl));
  // large: one wave per pair, 4 pairs per block, 96-edge lists (3 blocks per CU fit in LDS; 128 allowed 2); the pairs
  // that need more go to the overflow kernel (2.4M relate: large 6.1 -> 3.6 ms, overflow 2.7 -> 4.3 ms; a second LDS
  // tier of 128-edge lists for those was slower than the overflow kernel)
  // large_cap 0 (launches of a few pairs): every large pair goes to the overflow kernel, whose blocks of eight waves
  // per pair finish a few big pairs sooner than one wave each
  const int cap_large = (int)sh.large_cap;
  {
    const size_t sm = (size_t)(T / 64) * list_bytes(std::max(cap_large, 1), rel);
    const uint32_t g = std::min<uint32_t>(div_up(np, T / 64), cu * sh.large_bpc);
    hipLaunchKernelGGL(k_refine_large, dim3(g), dim3(T), sm, sl, P, R, pa, pb, list, counts, promo, promo_ctl, rp, out, cap_large, ovf, ctl);
    GSJ_LAUNCHED();
  }
  GSJ_HIP(hipEventRecord(c.ev[2], sl));
  if (side_ovf) {   // one wave per pair over a grid sized for the largest queue the launch can have, idle blocks leave at once
    const uint32_t g = std::min<uint32_t>((uint32_t)std::min<uint64_t>(np, 1u << 30), cu * 8);
    hipLaunchKernelGGL((k_refine_overflow<64, true>), dim3(g), dim3(64), list_bytes(OVF_SIDE_CAP, rel), sl, P, R, pa, pb, ovf, 0u, rp, out,
                       (int)OVF_SIDE_CAP, nullptr, ctl, ovf_main, ctl_main);
    GSJ_LAUNCHED();
    GSJ_HIP(hipEventRecord(c.ev[16], sl));
  }
  GSJ_HIP(hipEventRecord(c.ev[0], st));
  // grid-stride kernels: grids of one group per pair up to the shape's cap in blocks per CU
  auto grid = [&](uint32_t lanes, uint32_t bpc) { return std::min<uint32_t>(div_up(np, T / lanes), cu * bpc); };
  // screens (every tiny, small and medium pair), then the edge work of the flagged ones in class-list order
  if (sh.screen_occ == 4) {
    launch_screens<4>(grid(8, sh.screen_bpc), grid(16, sh.screen_bpc), grid(32, sh.screen_bpc), st, P, R, pa, pb, list, counts, rp, out, eq);
  } else {
    launch_screens<5>(grid(8, sh.screen_bpc), grid(16, sh.screen_bpc), grid(32, sh.screen_bpc), st, P, R, pa, pb, list, counts, rp, out, eq);
  }
  // over np + 1 positions so that off[n] exists for n = np; flags past the two classes are stale and only reach
  // offsets that are never read
  scan_u32_to_u64_async(c, eq.flag, eq.off, np + 1, nullptr);
  hipLaunchKernelGGL(k_queue_scatter, dim3(div_up(np, 256)), dim3(256), 0, st, list, counts, eq);
  GSJ_LAUNCHED();
  GSJ_HIP(hipEventRecord(c.ev[6], st));
  const uint32_t gt = grid(8, sh.edge_bpc), gs = grid(16, sh.edge_bpc), gm = grid(32, sh.edge_bpc);
  hipLaunchKernelGGL((k_edges_small<8, TCAP, CL_TINY>), dim3(gt), dim3(T), (T / 8) * small_bytes<TCAP>(rel), st, P, R, pa, pb, counts, rp, out, eq,
                     ovf_main, ctl_main);
  GSJ_LAUNCHED();
  GSJ_HIP(hipEventRecord(c.ev[8], st));
  hipLaunchKernelGGL((k_edges_small<16, VCAP, CL_SMALL>), dim3(gs), dim3(T), (T / 16) * small_bytes<VCAP>(rel), st, P, R, pa, pb, counts, rp, out, eq,
                     ovf_main, ctl_main);
  GSJ_LAUNCHED();
  GSJ_HIP(hipEventRecord(c.ev[1], st));
  hipLaunchKernelGGL((k_edges_small<32, MCAP, CL_MEDIUM>), dim3(gm), dim3(T), (T / 32) * small_bytes<MCAP>(rel), st, P, R, pa, pb, counts, rp, out, eq,
                     ovf_main, ctl_main);
  GSJ_LAUNCHED();
  GSJ_HIP(hipEventRecord(c.ev[9], st));
  if (sh.fork) {   // the overflow queue is complete once both chains are done
    GSJ_HIP(hipEventRecord(c.ev[11], sl));
    GSJ_HIP(hipStreamWaitEvent(st, c.ev[11], 0));
  }
  // class sizes [0, 5), edge queue sizes [8, 11); the overflow queues (the list classes', the host-read one)
  struct Host { uint32_t counts[16]; OverflowCtl ctl[2]; };
  Host* h = c.h_scal.as<Host>(2);
  static_assert(sizeof(Host) == 96, "counts and the two queues' controls are copied as one block");
  GSJ_HIP(hipMemcpyAsync(h, counts, sizeof(Host), hipMemcpyDeviceToHost, st));   // the counts, then both controls
  GSJ_HIP(hipStreamSynchronize(st));
  const OverflowCtl& qm = h->ctl[side_ovf ? 1 : 0];
  const uint32_t nover = qm.count, need = qm.need;
  // pairs refined by an overflow kernel, each once: the list classes' queue, plus the host-read queue less the pairs
  // the first handed on to it
  const uint32_t n_over = side_ovf ? h->ctl[0].count - h->ctl[0].pad0 + nover : nover;
  if (stats) {
    float ms_q = 0, ms_t = 0, ms_s = 0, ms_m = 0, ms_x = 0, ms_l = 0;
    GSJ_HIP(hipEventElapsedTime(&ms_q, c.ev[0], c.ev[6]));
    GSJ_HIP(hipEventElapsedTime(&ms_t, c.ev[6], c.ev[8]));
    GSJ_HIP(hipEventElapsedTime(&ms_s, c.ev[8], c.ev[1]));
    GSJ_HIP(hipEventElapsedTime(&ms_m, c.ev[1], c.ev[9]));
    GSJ_HIP(hipEventElapsedTime(&ms_x, c.ev[5], c.ev[7]));
    GSJ_HIP(hipEventElapsedTime(&ms_l, c.ev[7], c.ev[2]));
    stats->k_screen_ms += ms_q; stats->k_tiny_ms += ms_t; stats->k_small_ms += ms_s; stats->k_medium_ms += ms_m; stats->k_mixed_ms += ms_x;
    stats->k_large_ms += ms_l;
    stats->n_tiny += h->counts[CL_TINY]; stats->n_small += h->counts[CL_SMALL]; stats->n_medium += h->counts[CL_MEDIUM];
    stats->n_mixed += h->counts[CL_MIXED]; stats->n_large += h->counts[CL_LARGE];
    stats->n_over += n_over;
    if (side_ovf) {
      float ms_v = 0;
      GSJ_HIP(hipEventElapsedTime(&ms_v, c.ev[2], c.ev[16]));
      stats->k_over_ms += ms_v;
    }
    stats->n_tiny_edge += h->counts[8 + CL_TINY]; stats->n_small_edge += h->counts[8 + CL_SMALL]; stats->n_medium_edge += h->counts[8 + CL_MEDIUM];
    stats->launches++;
  }
  if (nover && sh.ovf_lds && need <= OVF_LDS_CAP) {   // lists in LDS: no slab
    const size_t lds = list_bytes((int)need, rel);
    GSJ_HIP(hipEventRecord(c.ev[3], st));
    // pairs too few to fill the GPU (a windowed request's few largest): eight waves each, so that the edge
    // contributions of both lists (up to 2 x 448 edges) take one or two rounds
    const bool few = nover <= (uint64_t)c.cu_count;
    const uint32_t blocks = (uint32_t)std::min<uint64_t>(nover, (uint64_t)c.cu_count * (few ? 1 : 8));
    if (few)
      hipLaunchKernelGGL((k_refine_overflow<512, true>), dim3(blocks), dim3(512), lds, st, P, R, pa, pb, ovf_main, nover, rp, out, (int)need, nullptr, nullptr, nullptr, nullptr);
    else
      hipLaunchKernelGGL((k_refine_overflow<64, true>), dim3(blocks), dim3(64), lds, st, P, R, pa, pb, ovf_main, nover, rp, out, (int)need, nullptr, nullptr, nullptr, nullptr);
    GSJ_LAUNCHED();
    GSJ_HIP(hipEventRecord(c.ev[4], st));
    if (stats) {
      GSJ_HIP(hipEventSynchronize(c.ev[4]));
      float ms_o = 0;
      GSJ_HIP(hipEventElapsedTime(&ms_o, c.ev[3], c.ev[4]));
      stats->k_over_ms += ms_o;
    }
  } else if (nover) {
    const size_t per = list_bytes((int)need, rel);
    // slabs of at most 1 GiB (half of what is left when the scratch is a fixed arena: compaction still needs room)
    const uint64_t budget = c.dev_arena ? std::min<uint64_t>(1ull << 30, c.dev_arena->left() / 2) : (1ull << 30);
    GSJ_REQUIRE(!c.dev_arena || per <= budget, GSJ_E_LIMIT, "overflow pair lists exceed the no-allocation scratch left for them");
    const uint64_t blocks = std::max<uint64_t>(1, std::min<uint64_t>(nover, budget / std::max<size_t>(per, 1)));   // one slab per block
    unsigned char* slab = c.s_heavy_edges.as<unsigned char>(per * blocks + 64);
    GSJ_HIP(hipEventRecord(c.ev[3], st));
    if (nover <= 2ull * c.cu_count)   // a few pairs: four waves each (2.4M relate: one wave each, 2.5 vs 4.5 ms)
      hipLaunchKernelGGL((k_refine_overflow<256, false>), dim3((uint32_t)blocks), dim3(256), 0, st, P, R, pa, pb, ovf_main, nover, rp, out, (int)need, slab, nullptr, nullptr, nullptr);
    else
      hipLaunchKernelGGL((k_refine_overflow<64, false>), dim3((uint32_t)blocks), dim3(64), 0, st, P, R, pa, pb, ovf_main, nover, rp, out, (int)need, slab, nullptr, nullptr, nullptr);
    GSJ_LAUNCHED();
    GSJ_HIP(hipEventRecord(c.ev[4], st));
    if (stats) {
      GSJ_HIP(hipEventSynchronize(c.ev[4]));
      float ms_o = 0;
      GSJ_HIP(hipEventElapsedTime(&ms_o, c.ev[3], c.ev[4]));
      stats->k_over_ms += ms_o;
    }
  }
  return n_over;
}

namespace {
constexpr int CLASS_BIT = 5;   // partition key: class above, size bucket below

// sort key per candidate, 8 bits (one radix pass): class (PairClass, bits 5-7) — tiny (both polygons within tcap stored vertices; none when
// tcap is 0), small (both within VCAP), medium (both within mcap; none when mcap is 0), all three with rings within
// RCAP; mixed (when enabled: none of those, one polygon within VCAP stored vertices); large — and the pair's stored vertex
// count in buckets (saturated at 31) below, so that the groups of a wave get pairs of similar work
__global__ void k_class_keys(const DPolys P, const DPolys R, const uint32_t* __restrict__ pa, const uint32_t* __restrict__ pb, uint64_t np,
                             uint32_t mcap, int mixed, uint32_t tcap, uint32_t size_sort, uint64_t* __restrict__ key, uint32_t* __restrict__ val) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= np) return;
  const uint32_t ia = P.ginfo[pa[i]], ib = R.ginfo[pb[i]];
  const uint32_t va = ia & 0xFFFFu, vb = ib & 0xFFFFu;
  const bool rings = (ia >> 16) <= RCAP && (ib >> 16) <= RCAP;
  const uint32_t cls = (rings && va <= tcap && vb <= tcap) ? CL_TINY
                       : (rings && va <= VCAP && vb <= VCAP) ? CL_SMALL
                       : (rings && va <= mcap && vb <= mcap) ? CL_MEDIUM
                       : (mixed && min(va, vb) <= (uint32_t)VCAP) ? CL_MIXED : CL_LARGE;
  // within a class, pairs of similar size side by side (balanced groups), at a base resolution per class: the large
  // class spans hundreds of vertices, so its buckets are 8 vertices wide (1 for tiny, 2 for small and mixed, 4 for
  // medium); size_sort > 1 widens them by powers of two (fewer, longer runs of candidate order)
  const uint32_t shift = (cls == CL_LARGE ? 3u : (cls == CL_MEDIUM ? 2u : (cls == CL_TINY ? 0u : 1u))) + size_sort - 1;
  key[i] = (uint64_t)((cls << CLASS_BIT) | (size_sort ? min(31u, (va + vb) >> shift) : 0u));
  val[i] = (uint32_t)i;
}

// class sizes from the sorted keys: first index of every class after the first by binary search
__global__ void k_class_counts(const uint64_t* __restrict__ key, uint64_t np, uint32_t* __restrict__ counts) {
  uint64_t first[CL_COUNT + 1];
  first[0] = 0;
  first[CL_COUNT] = np;
  for (int c = 1; c < CL_COUNT; c++) {
    uint64_t lo = 0, hi = np;
    while (lo < hi) {
      const uint64_t mid = (lo + hi) / 2;
      if ((key[mid] >> CLASS_BIT) < (uint64_t)c) lo = mid + 1;
      else hi = mid;
    }
    first[c] = lo;
  }
  for (int c = 0; c < CL_COUNT; c++) counts[c] = (uint32_t)(first[c + 1] - first[c]);
}

// the overflow queue's control at ctl, the host-read queue's (ovf_lds 2) right after it, the promotion queue's 64 bytes on
__global__ void k_ctl_reset(OverflowCtl* ctl, uint32_t* counts) {
  ctl[0] = OverflowCtl{0, 0, 0, 0};
  ctl[1] = OverflowCtl{0, 0, 0, 0};
  ctl[4] = OverflowCtl{0, 0, 0, 0};
  for (int k = 0; k < 16; k++) counts[k] = 0;   // class sizes [0, 5) and the edge queues' counts and firsts [8, 14)
}
}  // namespace

PairClasses partition_pairs(Context& c, const DPolys& P, const DPolys& R, const uint32_t* pa, const uint32_t* pb, uint64_t np, uint32_t mcap,
                            bool mixed, uint32_t tcap, uint32_t size_sort) {
  hipStream_t st = c.stream;
  char* misc = c.s_misc.as<char>(256);   // [0, 32) belongs to the scan / sort primitives
  PairClasses pc;
  pc.counts = reinterpret_cast<uint32_t*>(misc + 64);
  pc.ctl = reinterpret_cast<OverflowCtl*>(misc + 128);
  pc.list = nullptr;
  hipLaunchKernelGGL(k_ctl_reset, dim3(1), dim3(1), 0, st, pc.ctl, pc.counts);
  GSJ_LAUNCHED();
  if (np == 0) return pc;
  uint64_t* keys = c.s_keys.as<uint64_t>(np + 16);
  uint64_t* keys2 = c.s_keys2.as<uint64_t>(np + 16);
  uint32_t* vals = c.s_vals.as<uint32_t>(np + 16);
  uint32_t* vals2 = c.s_vals2.as<uint32_t>(np + 16);
  hipLaunchKernelGGL(k_class_keys, dim3(div_up(np, 256)), dim3(256), 0, st, P, R, pa, pb, np, mcap, mixed ? 1 : 0, tcap, size_sort, keys, vals);
  GSJ_LAUNCHED();
  // stable (candidate order within a key); one 8-bit digit, sorted without probing whether it varies: no host
  // synchronisation
  radix_sort_pairs(c, keys, keys2, vals, vals2, np, 0, 8, false);
  hipLaunchKernelGGL(k_class_counts, dim3(1), dim3(1), 0, st, keys, np, pc.counts);
  GSJ_LAUNCHED();
  pc.list = vals;
  return pc;
}

}  // namespace gsj
