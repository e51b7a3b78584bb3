// ORACLE — test infrastructure only. A plain CPU restatement of the reference engine's semantics, used by
// tests/ and bench.py (cpu_baseline) as the checker for the HIP engine. Nothing in the product links or calls it.
//
// Every function cites the reference file:line it restates. Build: oracle/Makefile (g++ -O2 -fopenmp
// -ffp-contract=off), so that every product and sum is rounded exactly as written: the HIP engine is compiled with
// the same contraction rule and must agree bit for bit on every predicate.
//
// Pinning status (what checks each part; see tests/test_oracle_*.py):
//   orient/left_of/seg_isect/onseg ... exact rational brute force (fractions) incl. adversarial near-collinear
//   layer prep (origin/ext/bb/blen) .. f_down/f_up checked against numpy float32 nextafter; blen against math.hypot
//   candidates ....................... brute-force all-pairs float-box overlap
//   refine: intersects ............... exact rational polygon intersection test on random small polygons
//   refine: area ..................... closed forms (containment, identical, adjacent, tilings) + exact rational
//                                       convex clipping
//   relate (DE-9IM) .................. textbook matrices for canonical configurations, transpose/match tables
//   point in polygon ................. exact rational crossing parity + on-boundary brute force
//   union-find ....................... BFS components over the same edge list
//   subdivide ........................ area conservation, vertex bound, syn flags only on cut edges
//   dedup ............................ partition of rows, exact equality of members
#include "core.h"

namespace orc {

// engine/engine.cuh:416-515: Green's-theorem contribution of edge p->q (weighted by its position relative to the
// other polygon), plus ST_Relate boundary lengths (inside / on the other polygon) when rel is set and the edge is
// real. Each edge's partial sums start at zero; the caller adds the per-edge partials in list order. (The reference
// accumulates per lane and reduces across a warp, an order with no canonical form; this per-edge order is the one
// the HIP engine uses, so both produce bit-identical sums.)
struct Contrib { Real acc = 0, lin = 0, lon = 0; uint32_t bits = 0; };
static Contrib edge_contrib(Real px, Real py, Real qx, Real qy, const std::vector<Edge>& L, Real cx, Real cy, bool in_c,
                            const Box& ov, Real tol, Real eps_par, bool real, bool rel, int cut, Real snap) {
  Contrib out;
  const Real dx = qx - px, dy = qy - py;
  const Real len2 = dx * dx + dy * dy;
  if (len2 == 0) return out;
  const Real tol2len2 = tol * tol * len2, eps2 = eps_par * eps_par;
  Real ts[MAXS]; int ns = 0;
  Real ol0[MAXO], ol1[MAXO]; bool osame[MAXO]; int oidx[MAXO]; int no = 0;
  ts[ns++] = 0; ts[ns++] = 1;
  const Real bx0 = r_min(px, qx) - tol, bx1 = r_max(px, qx) + tol, by0 = r_min(py, qy) - tol, by1 = r_max(py, qy) + tol;
  const Real snap2 = snap * snap;
  const int cnt = (int)L.size();
  for (int k = 0; k < cnt; k++) {
    const Edge& e = L[k];
    if (r_max(e.x0, e.x1) < bx0 || r_min(e.x0, e.x1) > bx1 || r_max(e.y0, e.y1) < by0 || r_min(e.y0, e.y1) > by1) continue;
    Real fx = e.x1 - e.x0, fy = e.y1 - e.y0;
    Real r = dx * fy - dy * fx;
    Real flen2 = fx * fx + fy * fy;
    Real c0 = orient(px, py, qx, qy, e.x0, e.y0), c1 = orient(px, py, qx, qy, e.x1, e.y1);
    const bool cut_tol = ORC_MUTANT != 5;
    bool col = (c0 * c0 <= snap2 * len2 || (cut_tol && (e.cut & 1) && c0 * c0 <= tol2len2)) &&
               (c1 * c1 <= snap2 * len2 || (cut_tol && (e.cut & 2) && c1 * c1 <= tol2len2));
    if (!col && flen2 > 0) {
      Real g0 = orient(e.x0, e.y0, e.x1, e.y1, px, py), g1 = orient(e.x0, e.y0, e.x1, e.y1, qx, qy), tf = tol * tol * flen2;
      col = (g0 * g0 <= snap2 * flen2 || (cut_tol && (cut & 1) && g0 * g0 <= tf)) &&
            (g1 * g1 <= snap2 * flen2 || (cut_tol && (cut & 2) && g1 * g1 <= tf));
    }
    if (col) {
      Real tc = ((e.x0 - px) * dx + (e.y0 - py) * dy) / len2, te = ((e.x1 - px) * dx + (e.y1 - py) * dy) / len2;
      Real lo = r_max(0.0, r_min(tc, te)), hi = r_min(1.0, r_max(tc, te));
      if (hi > lo) {
        if (lo > 0 && ns < MAXS) ts[ns++] = lo;
        if (hi < 1 && ns < MAXS) ts[ns++] = hi;
        if (no < MAXO) { ol0[no] = lo; ol1[no] = hi; osame[no] = (dx * fx + dy * fy) > 0; oidx[no] = k; no++; }
      }
    } else if (r * r > eps2 * len2 * flen2) {
      Real wx = e.x0 - px, wy = e.y0 - py;
      Real tn = wx * fy - wy * fx;
      if (r > 0 ? (tn <= 0 || tn >= r) : (tn >= 0 || tn <= r)) continue;
      Real t = tn / r, u = (wx * dy - wy * dx) / r;
      Real du = tol / (flen2 > 0 ? r_sqrt(flen2) : 1.0);
      if (u >= -du && u <= 1 + du && ns < MAXS) ts[ns++] = t;
    }
  }
  for (int a = 1; a < ns; a++) { Real v = ts[a]; int b = a - 1; while (b >= 0 && ts[b] > v) { ts[b + 1] = ts[b]; b--; } ts[b + 1] = v; }
  for (int k = 0; k + 1 < ns; k++) {
    Real t0 = ts[k], t1 = ts[k + 1];
    if (!(t1 > t0)) continue;
    Real tm = 0.5 * (t0 + t1);
    Real mx = px + dx * tm, my = py + dy * tm;
    bool in_ov = mx >= ov.x0 && mx <= ov.x1 && my >= ov.y0 && my <= ov.y1;
    int net = 0, first = -1, net_real = 0, n_real = 0;
    for (int o = 0; o < no; o++)
      if (tm > ol0[o] && tm < ol1[o]) {
        int sgn = osame[o] ? 1 : -1;
        net += sgn; if (first < 0) first = o;
        if (!L[oidx[o]].syn) { net_real += sgn; n_real++; }
      }
    int vin = -1;
    auto inside_m = [&]() { if (vin < 0) vin = (in_ov && inside_from_c(L, cx, cy, in_c, mx, my)) ? 1 : 0; return vin; };
    if (rel && real) {
      Real seg = r_sqrt(len2) * (t1 - t0);
      if (net_real != 0) { out.lon += seg; out.bits |= (net_real > 0 ? REL_COIN_SAME : REL_COIN_OPP) | REL_TOUCH; }
      else if (n_real == 0 && (net == 1 || net == -1)) out.lin += (ORC_MUTANT == 1 ? 1.0 : 0.5) * seg;
      else if (inside_m()) out.lin += seg;
    }
    Real w;
    if (first < 0) w = inside_m() ? 1.0 : 0.0;
    else if (net == 1 || net == -1) w = ORC_MUTANT == 2 ? 1.0 : 0.5;
    else if (net == 0) w = inside_m() ? 1.0 : 0.0;
    else {
      const Edge& eb = L[oidx[first]];
      bool sb = left_of(eb.x0, eb.y0, eb.x1, eb.y1, mx, my);
      bool m_left = osame[first] ? sb : !sb;
      int v = inside_m();
      int other = m_left ? v - net : v + net;
      other = other < 0 ? 0 : (other > 1 ? 1 : other);
      w = 0.5 * (v + other);
    }
    if (w != 0.0) {
      Real ax = px + dx * t0 - cx, ay = py + dy * t0 - cy, bx = px + dx * t1 - cx, by = py + dy * t1 - cy;
      out.acc += w * (ax * by - ay * bx);
    }
  }
  return out;
}

// engine/engine.cuh:524-527: half an ulp of the largest world coordinate of the pair
static inline Real pair_snap(Real ox, Real oy, Real ex, Real ey) {
  Real m = r_max(r_max(r_abs(ox), r_abs(ox + ex)), r_max(r_abs(oy), r_abs(oy + ey)));
  return 0.5 * (Real(std::nextafter(val(m), INFINITY)) - m);
}

// probe vertex of ring r (engine/engine.cuh:322-327): first vertex that starts a real edge
static inline bool probe_vertex(const View& A, uint64_t r, Real& x, Real& y) {
  uint64_t v = A.L->rvo[r], e = A.L->rvo[r + 1];
  if (A.L->has_syn()) { while (v + 1 < e && A.syn(v)) v++; if (v + 1 >= e) return false; }
  A.v(v, x, y);
  return true;
}

struct PairOut { uint8_t isect; double area; double ev[4]; uint32_t bits; };

// engine/engine.cuh:530-636 (k_refine; its heavy path 638-792 computes the same quantities with unbounded lists)
static PairOut refine_pair(const Layer& P, const Layer& R, uint32_t a, uint32_t b, const Params& rp) {
  std::vector<Edge> LA, LB;
  PairOut out{0, 0.0, {0, 0, 0, 0}, 0};
  const Real oax = P.ox[a], oay = P.oy[a], obx = R.ox[b], oby = R.oy[b];
  const Real eax = P.ex[a], eay = P.ey[a], ebx = R.ex[b], eby = R.ey[b];
  const Real sx = obx - oax, sy = oby - oay;
  const Real pad = 4.0 * r_max(rp.tol_abs, rp.tol_rel * r_max(r_max(eax, eay), r_max(ebx, eby)));
  Box ov{r_max(0.0, sx) - pad, r_max(0.0, sy) - pad, r_min(eax, sx + ebx) + pad, r_min(eay, sy + eby) + pad};
  View A{&P, a, oax - oax, oay - oay}, B{&R, b, obx - oax, oby - oay};
  const Real tol = r_max(rp.tol_abs, rp.tol_rel * r_max(ov.x1 - ov.x0, ov.y1 - ov.y0));
  const Real snap = rp.snap_half_ulp ? pair_snap(oax, oay, eax, eay) : 0.0;
  collect(A, ov, LA); collect(B, ov, LB);
  bool hit = false, hit_any = false;
  for (const Edge& ea : LA) {
    Real ax0 = r_min(ea.x0, ea.x1), ax1 = r_max(ea.x0, ea.x1), ay0 = r_min(ea.y0, ea.y1), ay1 = r_max(ea.y0, ea.y1);
    for (const Edge& eb : LB) {
      if (r_max(eb.x0, eb.x1) < ax0 || r_min(eb.x0, eb.x1) > ax1 || r_max(eb.y0, eb.y1) < ay0 || r_min(eb.y0, eb.y1) > ay1) continue;
      if (seg_isect(ea.x0, ea.y0, ea.x1, ea.y1, eb.x0, eb.y0, eb.x1, eb.y1)) {
        hit_any = true;
        if (!(ea.syn | eb.syn)) { hit = true; break; }
      }
    }
    if (hit) break;
  }
  const bool rtouch = hit;
  const Real cx = ov.x0 + 0.381966 * (ov.x1 - ov.x0), cy = ov.y0 + 0.414213 * (ov.y1 - ov.y0);
  bool inB_c = false, inA_c = false;
  if (!hit || rp.want_area || rp.want_relate) { inB_c = pip_full(B, cx, cy); inA_c = pip_full(A, cx, cy); }
  bool contain_any = false;
  if (!hit) {
    bool h = false, ha = false;
    auto probe = [&](const View& S, const std::vector<Edge>& Lother, bool in_other) {
      for (uint64_t r = S.r0(); r < S.r1() && !h; r++) {
        if (S.L->rvo[r] == S.L->rvo[r + 1]) continue;   // an empty ring has no probe vertex
        Real x, y; bool real = probe_vertex(S, r, x, y);
        if (!real) S.v(S.L->rvo[r], x, y);
        if (x >= ov.x0 && x <= ov.x1 && y >= ov.y0 && y <= ov.y1 && inside_from_c(Lother, cx, cy, in_other, x, y)) { ha = true; if (real) h = true; }
      }
    };
    probe(A, LB, inB_c);
    probe(B, LA, inA_c);
    hit = h; contain_any = ha;
  }
  Real ar = 0.0, e4[4] = {0, 0, 0, 0}; uint32_t eb4 = 0;
  if ((hit || hit_any || contain_any) && (rp.want_area || rp.want_relate)) {
#ifdef ORC_COUNT_FLOPS
    const FlopCount f0 = g_fc;
    struct EdgePhase {   // adds this block's counts to the edge-phase counters when it ends
      FlopCount f0;
      ~EdgePhase() {
        g_fc_edge.add += g_fc.add - f0.add; g_fc_edge.mul += g_fc.mul - f0.mul; g_fc_edge.fma += g_fc.fma - f0.fma;
        g_fc_edge.div += g_fc.div - f0.div; g_fc_edge.sqrt += g_fc.sqrt - f0.sqrt; g_edge_pairs++;
      }
    } edge_phase{f0};
#endif
    Real s2 = 0.0, ai = 0, ao = 0, bi = 0, bo = 0; uint32_t bits = rtouch ? REL_TOUCH : 0;
    const bool rel = rp.want_relate != 0;
    for (const Edge& e : LA) {
      Contrib c = edge_contrib(e.x0, e.y0, e.x1, e.y1, LB, cx, cy, inB_c, ov, tol, rp.eps_par, !e.syn, rel, e.cut, snap);
      s2 += c.acc; ai += c.lin; ao += c.lon; bits |= c.bits;
    }
#ifdef ORC_COUNT_FLOPS
    // a polygon with itself: side B repeats side A's computation exactly, and the engine copies A's results into B's
    // slots instead, so B's operations are not counted
    const FlopCount before_b = g_fc;
    const bool self_pair = &P == &R && a == b;
#endif
    for (const Edge& e : LB) {
      Contrib c = edge_contrib(e.x0, e.y0, e.x1, e.y1, LA, cx, cy, inA_c, ov, tol, rp.eps_par, !e.syn, rel, e.cut, snap);
      s2 += c.acc; bi += c.lin; bo += c.lon; bits |= c.bits;
    }
#ifdef ORC_COUNT_FLOPS
    if (self_pair) g_fc = before_b;
#endif
    ar = 0.5 * s2;
    e4[0] = ai; e4[1] = ao; e4[2] = bi; e4[3] = bo;
    eb4 = bits;
  } else if (rtouch) eb4 = REL_TOUCH;
  bool ev_any = e4[0] != 0 || e4[1] != 0 || e4[2] != 0 || e4[3] != 0;
  out.isect = hit ? 1 : ((ar != 0.0 || ev_any) ? 2 : 0);
  out.area = val(ar);
  for (int k = 0; k < 4; k++) out.ev[k] = val(e4[k]);
  out.bits = eb4;
  return out;
}

// engine/engine.cuh:848-856: DE-9IM of two areal geometries from the boundary evidence
static inline uint32_t de9im_code(const double* e, uint32_t bits, double LA, double LB) {
  const double ta = 1e-9 * LA, tb = 1e-9 * LB;
  bool aIn = e[0] > ta, aOut = (LA - e[0] - e[1]) > ta, bIn = e[2] > tb, bOut = (LB - e[2] - e[3]) > tb;
  bool same = bits & REL_COIN_SAME, opp = bits & REL_COIN_OPP, touch = bits & REL_TOUCH;
  uint32_t II = (aIn || bIn || same) ? 3 : 0, IB = bIn ? 2 : 0, IE = (aOut || bIn || opp) ? 3 : 0;
  uint32_t BI = aIn ? 2 : 0, BB = (same || opp) ? 2 : (touch ? 1 : 0), BE = aOut ? 2 : 0;
  uint32_t EI = (bOut || aIn || opp) ? 3 : 0, EB = bOut ? 2 : 0, EE = 3;
  return II | IB << 2 | IE << 4 | BI << 6 | BB << 8 | BE << 10 | EI << 12 | EB << 14 | EE << 16;
}

// ------------------------------------------------------------------------------------------- join
// engine/engine.cuh:895-1061 (join_stream) without batching: every piece pair of one parent pair is aggregated in
// one place, which the reference guarantees by keeping a probe parent's pieces in one sub-batch (lines 909-913, 925)
// (prefilter: candidate piece pairs it rejects are dropped before refinement, JoinOpts::prefilter, engine.cuh:935)
void join(const Layer& P, const Layer& R, const Params& rp, bool upper, bool no_agg, JoinResult& res, std::vector<uint32_t>* cand_out,
          std::vector<PairOut>* refined_out, const std::function<bool(uint32_t, uint32_t)>& prefilter) {
  std::vector<uint32_t> cand;
  candidates(P.bb, P.n, R.bb, R.n, upper, cand);
  res.candidates = cand.size() / 2;
  if (prefilter) {
    size_t k = 0;
    for (size_t i = 0; i < cand.size(); i += 2)
      if (prefilter(cand[i], cand[i + 1])) { cand[k++] = cand[i]; cand[k++] = cand[i + 1]; }
    cand.resize(k);
  }
  const uint64_t np = cand.size() / 2;
  std::vector<PairOut> ref(np);
#pragma omp parallel for schedule(dynamic, 256)
  for (int64_t i = 0; i < (int64_t)np; i++) ref[i] = refine_pair(P, R, cand[2 * i], cand[2 * i + 1], rp);
  const bool agg = (!P.parent.empty() || !R.parent.empty()) && !no_agg;
  const bool rel = rp.want_relate != 0;
  res.pairs.clear(); res.area.clear(); res.code.clear();
  if (!agg) {
    for (uint64_t i = 0; i < np; i++) {
      if (no_agg ? ref[i].isect == 0 : ref[i].isect != 1) continue;
      uint32_t a = cand[2 * i], b = cand[2 * i + 1];
      res.pairs.push_back(a); res.pairs.push_back(b);
      if (rp.want_area) res.area.push_back(ref[i].area);
      if (rel) res.code.push_back(de9im_code(ref[i].ev, ref[i].bits, P.blen[a], R.blen[b]));
    }
  } else {
    // engine/engine.cuh:963-1036: keep flags 1|2, fold by (parent a, parent b): area and evidence summed in
    // candidate order, bits or-ed, flag min; keep parents whose flag is 1
    std::vector<std::pair<uint64_t, uint64_t>> keyed;
    for (uint64_t i = 0; i < np; i++) {
      if (ref[i].isect == 0) continue;
      uint32_t a = cand[2 * i], b = cand[2 * i + 1];
      uint64_t pa = P.parent.empty() ? a : P.parent[a], pb = R.parent.empty() ? b : R.parent[b];
      keyed.push_back({(pa << 32) | pb, i});
    }
    std::stable_sort(keyed.begin(), keyed.end(), [](const auto& x, const auto& y) { return x.first < y.first; });
    for (size_t s = 0; s < keyed.size();) {
      size_t e = s;
      double ar = 0, ev[4] = {0, 0, 0, 0}; uint32_t bits = 0; uint8_t f = 255;
      while (e < keyed.size() && keyed[e].first == keyed[s].first) {
        const PairOut& o = ref[keyed[e].second];
        if (ORC_MUTANT != 4 || o.isect == 1) ar += o.area;
        for (int k = 0; k < 4; k++) ev[k] += o.ev[k];
        bits |= o.bits; f = std::min(f, o.isect);
        e++;
      }
      if (f == 1) {
        uint32_t pa = (uint32_t)(keyed[s].first >> 32), pb = (uint32_t)keyed[s].first;
        res.pairs.push_back(pa); res.pairs.push_back(pb);
        if (rp.want_area) res.area.push_back(ar);
        if (rel) res.code.push_back(de9im_code(ev, bits, P.blen[pa], R.blen[pb]));
      }
      s = e;
    }
  }
  if (cand_out) cand_out->swap(cand);
  if (refined_out) refined_out->swap(ref);
}

// ------------------------------------------------------------------------------------------- points in polygons
// engine/engine.cuh:1079-1106: 1 = interior, 2 = on a real edge, 0 = outside (the piece)
static uint8_t pip_pair(const Layer& pts, const Layer& R, uint32_t pi, uint32_t g) {
  const Real ox_ = R.ox[g], oy_ = R.oy[g];
  const Real px = pts.xw[2 * pi] - ox_, py = pts.xw[2 * pi + 1] - oy_;
  View B{&R, g, ox_ - ox_, oy_ - oy_};
  const Real ox = B.shx - 1.0;
  bool par = false, on = false;
  for (uint64_t r = B.r0(); r < B.r1() && !on; r++) {
    uint64_t s0 = R.rvo[r], e = R.rvo[r + 1];
    Real x0, y0; B.v(s0, x0, y0);
    for (uint64_t j = s0 + 1; j < e; j++) {
      Real x1, y1; B.v(j, x1, y1);
      if (!((y0 > py) == (y1 > py) && y0 != py && y1 != py)) {
        Real o_ = orient(x0, y0, x1, y1, px, py);
        if (o_ == 0 && !B.syn(j - 1) && onseg(x0, y0, x1, y1, px, py)) { on = true; break; }
        bool so = left_of(x0, y0, x1, y1, ox, py), sp = left_of(x0, y0, x1, y1, px, py);
        if (so != sp) {
          bool su = orient(ox, py, px, py, x0, y0) > 0, sv = orient(ox, py, px, py, x1, y1) > 0;
          if (su != sv) par = !par;
        }
      }
      x0 = x1; y0 = y1;
    }
  }
  return on ? 2 : (par ? 1 : 0);
}

// engine/engine.cuh:1112-1178: hits per (point, parent), boundary wins over interior
static void points_join(const Layer& pts, const Layer& R, JoinResult& res) {
  std::vector<uint32_t> cand;
  candidates(pts.bb, pts.n, R.bb, R.n, false, cand);
  const uint64_t np = cand.size() / 2;
  res.candidates = np;
  std::vector<uint8_t> hit(np);
#pragma omp parallel for schedule(dynamic, 1024)
  for (int64_t i = 0; i < (int64_t)np; i++) hit[i] = pip_pair(pts, R, cand[2 * i], cand[2 * i + 1]);
  std::vector<std::pair<uint64_t, uint8_t>> kv;
  for (uint64_t i = 0; i < np; i++)
    if (hit[i]) {
      uint64_t pb = R.parent.empty() ? cand[2 * i + 1] : R.parent[cand[2 * i + 1]];
      kv.push_back({((uint64_t)cand[2 * i] << 32) | pb, hit[i]});
    }
  std::sort(kv.begin(), kv.end());
  res.pairs.clear(); res.where.clear();
  for (size_t s = 0; s < kv.size();) {
    size_t e = s; uint8_t w = 0;
    while (e < kv.size() && kv[e].first == kv[s].first) { w = std::max(w, kv[e].second); e++; }
    res.pairs.push_back((uint32_t)(kv[s].first >> 32)); res.pairs.push_back((uint32_t)kv[s].first); res.where.push_back(w);
    s = e;
  }
}

// ------------------------------------------------------------------------------------------- cluster
// engine/engine.cuh:1216-1234: components of the intersects graph over parents; root = smallest member
static void cluster(const Layer& P, JoinResult& res) {
  Params rp; rp.want_area = 0; rp.want_relate = 0;
  JoinResult j;
  join(P, P, rp, true, false, j, nullptr, nullptr);
  const uint32_t n = P.nparent;
  std::vector<uint32_t> par(n);
  std::iota(par.begin(), par.end(), 0u);
  auto find = [&](uint32_t x) { while (par[x] != x) { par[x] = par[par[x]]; x = par[x]; } return x; };
  for (size_t i = 0; i < j.pairs.size(); i += 2) {
    uint32_t a = find(j.pairs[i]), b = find(j.pairs[i + 1]);
    if (a == b) continue;
    if (a < b) par[b] = a; else par[a] = b;
  }
  res.root.resize(n);
  for (uint32_t i = 0; i < n; i++) res.root[i] = find(i);
  res.candidates = j.candidates;
  res.pairs = j.pairs;
}

// ------------------------------------------------------------------------------------------- subdivide
// engine/subdivide.h:14-129 (ST_Subdivide analogue): recursive split at the bbox middle of the longer side,
// Sutherland-Hodgman clipping per ring, cut edges flagged synthetic
struct Ring { std::vector<double> xy; std::vector<uint8_t> syn; };

static void clip(const Ring& r, int axis, double m, bool keep_le, Ring& out) {   // engine/subdivide.h:14-48
  out.xy.clear(); out.syn.clear();
  const size_t n = r.xy.size() / 2;
  auto inside = [&](size_t i) { double v = r.xy[2 * i + axis]; return keep_le ? v <= m : v >= m; };
  auto push = [&](double x, double y, uint8_t f) {
    size_t k = out.xy.size();
    if (k >= 2 && out.xy[k - 2] == x && out.xy[k - 1] == y) { out.syn.back() = f; return; }
    out.xy.push_back(x); out.xy.push_back(y); out.syn.push_back(f);
  };
  for (size_t i = 0; i + 1 < n; i++) {
    double px = r.xy[2 * i], py = r.xy[2 * i + 1], qx = r.xy[2 * i + 2], qy = r.xy[2 * i + 3];
    bool pin = inside(i), qin = inside(i + 1);
    uint8_t f = r.syn[i];
    double ix, iy;
    if (axis == 0) { ix = m; iy = py + (m - px) * (qy - py) / (qx - px); }
    else { iy = m; ix = px + (m - py) * (qx - px) / (qy - py); }
    if (pin) push(px, py, f);
    if (pin && !qin) push(ix, iy, 1);
    else if (!pin && qin) push(ix, iy, f);
  }
  size_t k = out.xy.size() / 2;
  if (k < 3) { out.xy.clear(); out.syn.clear(); return; }
  if (out.xy[0] != out.xy[2 * k - 2] || out.xy[1] != out.xy[2 * k - 1]) { double x0 = out.xy[0], y0 = out.xy[1]; out.xy.push_back(x0); out.xy.push_back(y0); out.syn.push_back(0); }
  else out.syn.back() = 0;
  k = out.xy.size() / 2;
  double a = 0;
  for (size_t i = 0; i + 1 < k; i++) a += out.xy[2 * i] * out.xy[2 * i + 3] - out.xy[2 * i + 2] * out.xy[2 * i + 1];
  if (k < 4 || a == 0.0) { out.xy.clear(); out.syn.clear(); }
}

static void subdivide_rec(std::vector<Ring>& rings, int maxv, int depth, std::vector<std::vector<Ring>>& out) {   // :52-67
  size_t nv = 0; for (auto& r : rings) nv += r.xy.size() / 2;
  if (nv == 0) return;
  double x0 = 1e300, y0 = 1e300, x1 = -1e300, y1 = -1e300;
  for (auto& r : rings) for (size_t i = 0; i < r.xy.size(); i += 2) {
    x0 = std::min(x0, r.xy[i]); x1 = std::max(x1, r.xy[i]); y0 = std::min(y0, r.xy[i + 1]); y1 = std::max(y1, r.xy[i + 1]);
  }
  if ((int)nv <= maxv || depth >= 24 || (x1 - x0 <= 0 && y1 - y0 <= 0)) { out.push_back(std::move(rings)); return; }
  int axis = (x1 - x0) >= (y1 - y0) ? 0 : 1;
  double m = axis == 0 ? 0.5 * (x0 + x1) : 0.5 * (y0 + y1);
  for (int side = 0; side < 2; side++) {
    std::vector<Ring> part; Ring c;
    for (auto& r : rings) { clip(r, axis, m, side == 0, c); if (!c.xy.empty()) part.push_back(c); }
    if (!part.empty()) subdivide_rec(part, maxv, depth + 1, out);
  }
}

struct Flat { std::vector<uint64_t> gro{0}, rvo{0}; std::vector<double> xy; std::vector<uint8_t> syn; std::vector<uint32_t> parent; };

static void subdivide(const uint64_t* gro, const uint64_t* rvo, const double* xy, uint64_t n, int maxv, Flat& o) {   // :71-129
  o = Flat();
  for (uint64_t g = 0; g < n; g++) {
    uint64_t r0 = gro[g], r1 = gro[g + 1];
    uint64_t nv = rvo[r1] - rvo[r0];
    if ((int64_t)nv <= maxv) {
      for (uint64_t r = r0; r < r1; r++) {
        o.xy.insert(o.xy.end(), xy + 2 * rvo[r], xy + 2 * rvo[r + 1]);
        o.syn.insert(o.syn.end(), rvo[r + 1] - rvo[r], 0);
        o.rvo.push_back(o.xy.size() / 2);
      }
      o.gro.push_back(o.rvo.size() - 1); o.parent.push_back((uint32_t)g);
      continue;
    }
    std::vector<Ring> rings;
    for (uint64_t r = r0; r < r1; r++) {
      Ring sr; sr.xy.assign(xy + 2 * rvo[r], xy + 2 * rvo[r + 1]); sr.syn.assign(rvo[r + 1] - rvo[r], 0);
      rings.push_back(std::move(sr));
    }
    std::vector<std::vector<Ring>> pieces;
    subdivide_rec(rings, maxv, 0, pieces);
    for (auto& p : pieces) {
      for (auto& r : p) { o.xy.insert(o.xy.end(), r.xy.begin(), r.xy.end()); o.syn.insert(o.syn.end(), r.syn.begin(), r.syn.end()); o.rvo.push_back(o.xy.size() / 2); }
      o.gro.push_back(o.rvo.size() - 1); o.parent.push_back((uint32_t)g);
    }
  }
}

// ------------------------------------------------------------------------------------------- dedup
// engine/dedup.h:23-124 (first-occurrence order): distinct geometries and member lists
static uint64_t geom_hash(const uint64_t* gro, const uint64_t* rvo, const double* xy, uint64_t g) {   // :23-31
  uint64_t h = 1469598103934665603ull ^ (gro[g + 1] - gro[g]);
  for (uint64_t r = gro[g]; r < gro[g + 1]; r++) {
    h = (h ^ (rvo[r + 1] - rvo[r])) * 1099511628211ull;
    const uint64_t* w = reinterpret_cast<const uint64_t*>(xy + 2 * rvo[r]);
    for (uint64_t i = 0; i < 2 * (rvo[r + 1] - rvo[r]); i++) { h ^= w[i]; h *= 0x100000001b3ull; h ^= h >> 29; }
  }
  return h;
}
static bool geom_equal(const uint64_t* gro, const uint64_t* rvo, const double* xy, uint64_t a, uint64_t b) {   // :33-42
  if (gro[a + 1] - gro[a] != gro[b + 1] - gro[b]) return false;
  for (uint64_t k = 0; k < gro[a + 1] - gro[a]; k++)
    if (rvo[gro[a] + k + 1] - rvo[gro[a] + k] != rvo[gro[b] + k + 1] - rvo[gro[b] + k]) return false;
  uint64_t va = rvo[gro[a]], vb = rvo[gro[b]], nv = rvo[gro[a + 1]] - va;
  return std::memcmp(xy + 2 * va, xy + 2 * vb, nv * 16) == 0;
}
struct Dedup { std::vector<uint32_t> first; std::vector<uint64_t> moff; std::vector<uint32_t> midx; };
static void dedup(const uint64_t* gro, const uint64_t* rvo, const double* xy, uint64_t n, Dedup& d) {
  std::vector<uint64_t> h(n);
  for (uint64_t g = 0; g < n; g++) h[g] = geom_hash(gro, rvo, xy, g);
  std::vector<uint32_t> ord(n);
  std::iota(ord.begin(), ord.end(), 0u);
  std::sort(ord.begin(), ord.end(), [&](uint32_t a, uint32_t b) { return h[a] != h[b] ? h[a] < h[b] : a < b; });
  std::vector<uint32_t> rep(n);
  for (uint64_t i = 0; i < n;) {
    uint64_t j = i;
    while (j < n && h[ord[j]] == h[ord[i]]) j++;
    for (uint64_t k = i; k < j; k++) {
      uint32_t g = ord[k]; rep[g] = g;
      for (uint64_t m = i; m < k; m++) if (rep[ord[m]] == ord[m] && geom_equal(gro, rvo, xy, ord[m], g)) { rep[g] = ord[m]; break; }
    }
    i = j;
  }
  std::vector<uint32_t> uid(n, UINT32_MAX);
  d.first.clear();
  for (uint64_t g = 0; g < n; g++) if (rep[g] == g) { uid[g] = (uint32_t)d.first.size(); d.first.push_back((uint32_t)g); }
  const uint64_t un = d.first.size();
  d.moff.assign(un + 1, 0);
  for (uint64_t g = 0; g < n; g++) d.moff[uid[rep[g]] + 1]++;
  for (uint64_t i = 0; i < un; i++) d.moff[i + 1] += d.moff[i];
  d.midx.resize(n);
  std::vector<uint64_t> fill(d.moff.begin(), d.moff.end() - 1);
  for (uint64_t g = 0; g < n; g++) d.midx[fill[uid[rep[g]]]++] = (uint32_t)g;
}

}  // namespace orc

// =========================================================================================== C ABI (ctypes)
using namespace orc;

extern "C" {

double orc_orient(double ax, double ay, double bx, double by, double cx, double cy) { return val(orient(ax, ay, bx, by, cx, cy)); }
int orc_left_of(double ax, double ay, double bx, double by, double px, double py) { return left_of(ax, ay, bx, by, px, py); }
int orc_seg_isect(double a, double b, double c, double d, double e, double f, double g, double h) { return seg_isect(a, b, c, d, e, f, g, h); }
float orc_f_down(double v) { return f_down(v); }
float orc_f_up(double v) { return f_up(v); }
uint32_t orc_de9im_code(const double* ev, uint32_t bits, double la, double lb) { return de9im_code(ev, bits, la, lb); }

// FLOP counters of this thread (counting build only; zeros otherwise): add/sub, mul, fma, div, sqrt
// out[5 .. 10): the same for the edge-contribution phase alone, out[10]: pairs that reached it
void orc_flops(uint64_t* out, int reset) {
#ifdef ORC_COUNT_FLOPS
  out[0] = g_fc.add; out[1] = g_fc.mul; out[2] = g_fc.fma; out[3] = g_fc.div; out[4] = g_fc.sqrt;
  out[5] = g_fc_edge.add; out[6] = g_fc_edge.mul; out[7] = g_fc_edge.fma; out[8] = g_fc_edge.div; out[9] = g_fc_edge.sqrt;
  out[10] = g_edge_pairs;
  if (reset) { g_fc = FlopCount(); g_fc_edge = FlopCount(); g_edge_pairs = 0; }
#else
  (void)reset;
  for (int k = 0; k < 11; k++) out[k] = 0;
#endif
}

void orc_set_threads(int t) {
#ifdef _OPENMP
  if (t > 0) omp_set_num_threads(t);
#else
  (void)t;
#endif
}

void* orc_layer_polys(const uint64_t* gro, const uint64_t* rvo, const double* xy, uint64_t n, uint64_t nr, uint64_t nv,
                      const uint32_t* parent, const uint8_t* syn) {
  Layer* L = new Layer();
  L->n = n; L->nr = nr; L->nv = nv;
  L->gro.assign(gro, gro + n + 1); L->rvo.assign(rvo, rvo + nr + 1); L->xw.assign(xy, xy + 2 * nv);
  if (parent) L->parent.assign(parent, parent + n);
  if (syn) L->syn.assign(syn, syn + nv);
  prep_polys(*L);
  return L;
}
void* orc_layer_points(const double* xy, uint64_t n) {
  Layer* L = new Layer();
  L->points = true; L->n = n; L->xw.assign(xy, xy + 2 * n);
  prep_points(*L);
  return L;
}
void orc_layer_free(void* p) { delete (Layer*)p; }

// prepared arrays for bit-exact comparison with the engine's device layer
void orc_layer_export(const void* p, double* origin, double* ext, float* bb, double* q, double* blen) {
  const Layer& L = *(const Layer*)p;
  for (uint64_t g = 0; g < L.n && !L.points; g++) {
    if (origin) { origin[2 * g] = L.ox[g]; origin[2 * g + 1] = L.oy[g]; }
    if (ext) { ext[2 * g] = L.ex[g]; ext[2 * g + 1] = L.ey[g]; }
  }
  if (bb) std::memcpy(bb, L.bb.data(), L.bb.size() * 4);
  if (q && !L.points) std::memcpy(q, L.q.data(), L.q.size() * 8);
  if (blen && !L.points) std::memcpy(blen, L.blen.data(), L.blen.size() * 8);
}
uint32_t orc_layer_nparent(const void* p) { return ((const Layer*)p)->nparent; }

// results
void* orc_result_new() { return new JoinResult(); }
void orc_result_free(void* r) { delete (JoinResult*)r; }
uint64_t orc_result_rows(const void* r) { return ((const JoinResult*)r)->pairs.size() / 2; }
uint64_t orc_result_candidates(const void* r) { return ((const JoinResult*)r)->candidates; }
void orc_result_copy(const void* rp, uint32_t* pairs, double* area, uint32_t* code, uint8_t* where) {
  const JoinResult& r = *(const JoinResult*)rp;
  if (pairs && !r.pairs.empty()) std::memcpy(pairs, r.pairs.data(), r.pairs.size() * 4);
  if (area && !r.area.empty()) std::memcpy(area, r.area.data(), r.area.size() * 8);
  if (code && !r.code.empty()) std::memcpy(code, r.code.data(), r.code.size() * 4);
  if (where && !r.where.empty()) std::memcpy(where, r.where.data(), r.where.size());
}
uint64_t orc_result_nroot(const void* r) { return ((const JoinResult*)r)->root.size(); }
void orc_result_roots(const void* rp, uint32_t* root) { const JoinResult& r = *(const JoinResult*)rp; std::memcpy(root, r.root.data(), r.root.size() * 4); }

// candidate pairs (sorted by (q, r)); query boxes from P (or qbb if given); returns the count; copies up to cap
uint64_t orc_candidates(const void* P, const void* R, int upper, uint32_t* out, uint64_t cap) {
  const Layer& p = *(const Layer*)P; const Layer& r = *(const Layer*)R;
  std::vector<uint32_t> c;
  candidates(p.bb, p.n, r.bb, r.n, upper != 0, c);
  uint64_t n = c.size() / 2;
  if (out) std::memcpy(out, c.data(), std::min(n, cap) * 8);
  return n;
}

// per-pair refinement of the given candidate pairs
void orc_refine(const void* P, const void* R, const uint32_t* pairs, uint64_t np, int want_area, int want_relate, int snap_half_ulp,
                uint8_t* isect, double* area, double* ev, uint32_t* evb) {
  Params rp; rp.want_area = want_area; rp.want_relate = want_relate; rp.snap_half_ulp = snap_half_ulp;
  const Layer& p = *(const Layer*)P; const Layer& r = *(const Layer*)R;
#pragma omp parallel for schedule(dynamic, 256)
  for (int64_t i = 0; i < (int64_t)np; i++) {
    PairOut o = refine_pair(p, r, pairs[2 * i], pairs[2 * i + 1], rp);
    if (isect) isect[i] = o.isect;
    if (area) area[i] = o.area;
    if (ev) for (int k = 0; k < 4; k++) ev[4 * i + k] = o.ev[k];
    if (evb) evb[i] = o.bits;
  }
}

// join (intersects [+ area] [+ relate]) with piece -> parent aggregation; rows are parent pairs
void orc_join(const void* P, const void* R, int want_area, int want_relate, int upper, int no_agg, void* res) {
  Params rp; rp.want_area = want_area; rp.want_relate = want_relate;
  join(*(const Layer*)P, *(const Layer*)R, rp, upper != 0, no_agg != 0, *(JoinResult*)res, nullptr, nullptr);
}

uint8_t orc_pip_pair(const void* pts, const void* R, uint32_t pi, uint32_t g) { return pip_pair(*(const Layer*)pts, *(const Layer*)R, pi, g); }
void orc_points_join(const void* pts, const void* R, void* res) { points_join(*(const Layer*)pts, *(const Layer*)R, *(JoinResult*)res); }
void orc_cluster(const void* P, void* res) { cluster(*(const Layer*)P, *(JoinResult*)res); }

// subdivision: returns a handle to the flat piece set
void* orc_subdivide(const uint64_t* gro, const uint64_t* rvo, const double* xy, uint64_t n, int maxv) {
  Flat* f = new Flat();
  subdivide(gro, rvo, xy, n, maxv, *f);
  return f;
}
void orc_flat_sizes(const void* p, uint64_t* n, uint64_t* nr, uint64_t* nv) {
  const Flat& f = *(const Flat*)p; *n = f.parent.size(); *nr = f.rvo.size() - 1; *nv = f.xy.size() / 2;
}
void orc_flat_copy(const void* p, uint64_t* gro, uint64_t* rvo, double* xy, uint8_t* syn, uint32_t* parent) {
  const Flat& f = *(const Flat*)p;
  std::memcpy(gro, f.gro.data(), f.gro.size() * 8); std::memcpy(rvo, f.rvo.data(), f.rvo.size() * 8);
  std::memcpy(xy, f.xy.data(), f.xy.size() * 8); std::memcpy(syn, f.syn.data(), f.syn.size());
  std::memcpy(parent, f.parent.data(), f.parent.size() * 4);
}
void orc_flat_free(void* p) { delete (Flat*)p; }

// dedup: returns a handle; first[un] (representative per distinct geometry), moff[un+1], midx[n]
void* orc_dedup(const uint64_t* gro, const uint64_t* rvo, const double* xy, uint64_t n) { Dedup* d = new Dedup(); dedup(gro, rvo, xy, n, *d); return d; }
uint64_t orc_dedup_n(const void* p) { return ((const Dedup*)p)->first.size(); }
void orc_dedup_copy(const void* p, uint32_t* first, uint64_t* moff, uint32_t* midx) {
  const Dedup& d = *(const Dedup*)p;
  std::memcpy(first, d.first.data(), d.first.size() * 4); std::memcpy(moff, d.moff.data(), d.moff.size() * 8);
  std::memcpy(midx, d.midx.data(), d.midx.size() * 4);
}
void orc_dedup_free(void* p) { delete (Dedup*)p; }

}  // extern "C"
