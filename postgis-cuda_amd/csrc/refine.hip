// Exact FP64 refinement of candidate polygon pairs (reference: k_refine and the heavy path,
// engine/engine.cuh:342-792). Per pair: local edge lists of both polygons inside the pair's overlap box, boundary
// contact by exact segment tests, containment by crossing parity from a probe point c, and (for area / relate)
// Green's-theorem contributions of every local edge weighted by its position relative to the other polygon.
//
// MI355X design:
//  * a lane GROUP per pair, not a 32-lane warp: 16 lanes for pairs whose polygons have <= 32 edges each (four pairs
//    per wave64; the bulk of parcel data), a full wave64 for larger pairs. Group-local ballots are slices of the
//    64-bit wave ballot.
//  * edge lists live in LDS (40 B per edge); pairs whose lists exceed the LDS capacity go to a second pass with
//    per-wave lists in global memory (sized by the largest overflow), all on the device without host round trips
//    except one read of the overflow count.
//  * sums are deterministic and independent of the group size: each lane stores its edges' contributions in LDS and
//    group lanes 0..4 add them sequentially in list order (A's edges, then B's). The result is bit-identical between
//    the 16-lane, 64-lane and overflow paths, and with the CPU restatement, so areas compare exactly and a pair gives
//    the same bits alone or batched with other requests.
//  * FP64 on the vector ALUs only; nothing here is a dense contraction, so MFMA is not used.
#include "engine.h"

namespace gsj {

namespace {

struct Ed { double x0, y0, x1, y1; int syn, cut; };   // 40 B; cut bit 1/2: start/end vertex is a cut point
struct Box { double x0, y0, x1, y1; };

__device__ __forceinline__ int syn_bit(const uint32_t* synb, uint64_t i) {
  return synb ? (int)((synb[i >> 5] >> (i & 31)) & 1u) : 0;
}

template <int G>
__device__ __forceinline__ uint64_t gballot(bool p) {
  const uint64_t m = __ballot(p);
  if constexpr (G == 64) return m;
  const int sh = (threadIdx.x & 63) & ~(G - 1);
  return (m >> sh) & ((1ull << G) - 1);
}
template <int G> __device__ __forceinline__ bool gany(bool p) { return gballot<G>(p) != 0; }

// make this wave's scratch writes (LDS or global) visible to its other lanes before they read them; also a code
// motion barrier. Scratch is never shared between waves.
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
}

__device__ __forceinline__ bool edge_outside(double px, double py, double qx, double qy, const Box& o) {
  return fmax(px, qx) < o.x0 || fmin(px, qx) > o.x1 || fmax(py, qy) < o.y0 || fmin(py, qy) > o.y1;
}

// edges of geometry g meeting box ov, in ring order, compacted by group ballot (engine/engine.cuh:358-377);
// returns the full count (may exceed cap: then the list is incomplete and the pair overflows)
template <int G>
__device__ int collect(const DPolys& S, uint32_t g, double shx, double shy, const Box& ov, Ed* out, int cap, int gl) {
  int cnt = 0;
  const uint32_t r0 = S.gro[g], r1 = S.gro[g + 1];
  for (uint32_t r = r0; r < r1; r++) {
    const uint64_t s = S.rvo[r], e = S.rvo[r + 1];
    for (uint64_t base = s; base + 1 < e; base += G) {
      const uint64_t i = base + gl;
      bool v = false;
      Ed ed;
      if (i + 1 < e) {
        const double2 p = S.xy[i], q = S.xy[i + 1];
        ed.x0 = shx + p.x; ed.y0 = shy + p.y; ed.x1 = shx + q.x; ed.y1 = shy + q.y;
        ed.syn = syn_bit(S.synb, i);
        ed.cut = S.synb ? ((syn_bit(S.synb, i > s ? i - 1 : e - 2) ? 1 : 0) | (syn_bit(S.synb, i + 2 < e ? i + 1 : s) ? 2 : 0)) : 0;
        v = !edge_outside(ed.x0, ed.y0, ed.x1, ed.y1, ov);
      }
      const uint64_t m = gballot<G>(v);
      const int pos = cnt + __popcll(m & ((1ull << gl) - 1));
      if (v && pos < cap) out[pos] = ed;
      cnt += __popcll(m);
    }
  }
  return cnt;
}

// inside(m) for geometry g: parity of the segment from a point left of its bbox to m, over every edge
// (engine/engine.cuh:382-397)
template <int G>
__device__ bool pip_full(const DPolys& S, uint32_t g, double shx, double shy, double mx, double my, int gl) {
  const double ox = shx - 1.0, oy = my;
  bool par = false;
  const uint32_t r0 = S.gro[g], r1 = S.gro[g + 1];
  for (uint32_t r = r0; r < r1; r++) {
    const uint64_t s = S.rvo[r], e = S.rvo[r + 1];
    for (uint64_t j = s + 1 + gl; j < e; j += G) {
      const double2 p = S.xy[j - 1], q = S.xy[j];
      const double x0 = shx + p.x, y0 = shy + p.y, x1 = shx + q.x, y1 = shy + q.y;
      if ((y0 > my) == (y1 > my) && y0 != my && y1 != my) continue;
      const bool so = left_of(x0, y0, x1, y1, ox, oy), sc = left_of(x0, y0, x1, y1, mx, my);
      if (so == sc) continue;
      const bool su = orient(ox, oy, mx, my, x0, y0) > 0, sv = orient(ox, oy, mx, my, x1, y1) > 0;
      if (su != sv) par = !par;
    }
  }
  return (__popcll(gballot<G>(par)) & 1) != 0;
}

// inside(x) = inside(c) xor parity of segment c->x across the local edges (engine/engine.cuh:400-412)
__device__ bool inside_from_c(const Ed* L, int cnt, double cx, double cy, bool in_c, double x, double y) {
  bool in = in_c;
  const double sx0 = fmin(cx, x), sx1 = fmax(cx, x), sy0 = fmin(cy, y), sy1 = fmax(cy, y);
  for (int k = 0; k < cnt; k++) {
    const Ed e = L[k];
    if (fmax(e.x0, e.x1) < sx0 || fmin(e.x0, e.x1) > sx1 || fmax(e.y0, e.y1) < sy0 || fmin(e.y0, e.y1) > sy1) continue;
    const bool sc = left_of(e.x0, e.y0, e.x1, e.y1, cx, cy), sx = left_of(e.x0, e.y0, e.x1, e.y1, x, y);
    if (sc == sx) continue;
    const bool su = orient(cx, cy, x, y, e.x0, e.y0) > 0, sv = orient(cx, cy, x, y, e.x1, e.y1) > 0;
    if (su != sv) in = !in;
  }
  return in;
}

struct Contrib { double acc, lin, lon; uint32_t bits; };

// One edge's Green's-theorem term (relative to c) and, for a real edge with relate on, its length inside / on the
// other polygon (engine/engine.cuh:416-515). Partial sums start at zero for every edge.
__device__ Contrib edge_contrib(const Ed& E, const Ed* L, int cnt, double cx, double cy, bool in_c, const Box& ov, double tol, double eps_par,
                                bool rel, double snap) {
  Contrib out{0.0, 0.0, 0.0, 0u};
  const double px = E.x0, py = E.y0, qx = E.x1, qy = E.y1;
  const int cut = E.cut;
  const bool real = !E.syn;
  const double dx = qx - px, dy = qy - py;
  const double len2 = dx * dx + dy * dy;
  if (len2 == 0) return out;
  const double tol2len2 = tol * tol * len2, eps2 = eps_par * eps_par;
  const double snap2 = snap * snap;
  double ts[MAXS];
  int ns = 0;
  double ol0[MAXO], ol1[MAXO];
  bool osame[MAXO];
  int oidx[MAXO];
  int no = 0;
  ts[ns++] = 0; ts[ns++] = 1;
  const double bx0 = fmin(px, qx) - tol, bx1 = fmax(px, qx) + tol, by0 = fmin(py, qy) - tol, by1 = fmax(py, qy) + tol;
  for (int k = 0; k < cnt; k++) {
    const Ed e = L[k];
    if (fmax(e.x0, e.x1) < bx0 || fmin(e.x0, e.x1) > bx1 || fmax(e.y0, e.y1) < by0 || fmin(e.y0, e.y1) > by1) continue;
    const double fx = e.x1 - e.x0, fy = e.y1 - e.y0;
    const double r = dx * fy - dy * fx;
    const double flen2 = fx * fx + fy * fy;
    // coincident if either edge's endpoints lie on the other's line (within snap, or within tol at cut points)
    const double c0 = orient(px, py, qx, qy, e.x0, e.y0), c1 = orient(px, py, qx, qy, e.x1, e.y1);
    bool col = (c0 * c0 <= snap2 * len2 || ((e.cut & 1) && c0 * c0 <= tol2len2)) && (c1 * c1 <= snap2 * len2 || ((e.cut & 2) && c1 * c1 <= tol2len2));
    if (!col && flen2 > 0) {
      const double g0 = orient(e.x0, e.y0, e.x1, e.y1, px, py), g1 = orient(e.x0, e.y0, e.x1, e.y1, qx, qy), tf = tol * tol * flen2;
      col = (g0 * g0 <= snap2 * flen2 || ((cut & 1) && g0 * g0 <= tf)) && (g1 * g1 <= snap2 * flen2 || ((cut & 2) && g1 * g1 <= tf));
    }
    if (col) {
      const double tc = ((e.x0 - px) * dx + (e.y0 - py) * dy) / len2, te = ((e.x1 - px) * dx + (e.y1 - py) * dy) / len2;
      const double lo = fmax(0.0, fmin(tc, te)), hi = fmin(1.0, fmax(tc, te));
      if (hi > lo) {
        if (lo > 0 && ns < MAXS) ts[ns++] = lo;
        if (hi < 1 && ns < MAXS) ts[ns++] = hi;
        if (no < MAXO) { ol0[no] = lo; ol1[no] = hi; osame[no] = (dx * fx + dy * fy) > 0; oidx[no] = k; no++; }
      }
    } else if (r * r > eps2 * len2 * flen2) {
      const double wx = e.x0 - px, wy = e.y0 - py;
      const double tn = wx * fy - wy * fx;
      if (r > 0 ? (tn <= 0 || tn >= r) : (tn >= 0 || tn <= r)) continue;
      const double t = tn / r, u = (wx * dy - wy * dx) / r;
      const double du = tol / (flen2 > 0 ? sqrt(flen2) : 1.0);
      if (u >= -du && u <= 1 + du && ns < MAXS) ts[ns++] = t;
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
    const double mx = px + dx * tm, my = py + dy * tm;
    const bool in_ov = mx >= ov.x0 && mx <= ov.x1 && my >= ov.y0 && my <= ov.y1;
    int net = 0, first = -1, net_real = 0, n_real = 0;
    for (int o = 0; o < no; o++)
      if (tm > ol0[o] && tm < ol1[o]) {
        const int sgn = osame[o] ? 1 : -1;
        net += sgn;
        if (first < 0) first = o;
        if (!L[oidx[o]].syn) { net_real += sgn; n_real++; }
      }
    int vin = -1;   // midpoint inside the other polygon, computed on first use
    auto inside_m = [&]() {
      if (vin < 0) vin = (in_ov && inside_from_c(L, cnt, cx, cy, in_c, mx, my)) ? 1 : 0;
      return vin;
    };
    if (rel && real) {
      const double seg = sqrt(len2) * (t1 - t0);
      if (net_real != 0) { out.lon += seg; out.bits |= (net_real > 0 ? REL_COIN_SAME : REL_COIN_OPP) | REL_TOUCH; }
      else if (n_real == 0 && (net == 1 || net == -1)) out.lin += 0.5 * seg;
      else if (inside_m()) out.lin += seg;
    }
    double w;
    if (first < 0) w = inside_m() ? 1.0 : 0.0;
    else if (net == 1 || net == -1) w = 0.5;
    else if (net == 0) w = inside_m() ? 1.0 : 0.0;
    else {
      const Ed eb = L[oidx[first]];
      const bool sb = left_of(eb.x0, eb.y0, eb.x1, eb.y1, mx, my);
      const bool m_left = osame[first] ? sb : !sb;
      const int v = inside_m();
      int other = m_left ? v - net : v + net;
      other = other < 0 ? 0 : (other > 1 ? 1 : other);
      w = 0.5 * (v + other);
    }
    if (w != 0.0) {
      const double ax = px + dx * t0 - cx, ay = py + dy * t0 - cy, bx = px + dx * t1 - cx, by = py + dy * t1 - cy;
      out.acc += w * (ax * by - ay * bx);
    }
  }
  return out;
}

__device__ __forceinline__ double pair_snap(double2 o, double2 e) {   // engine/engine.cuh:524-527
  const double m = fmax(fmax(fabs(o.x), fabs(o.x + e.x)), fmax(fabs(o.y), fabs(o.y + e.y)));
  return 0.5 * (nextafter(m, INFINITY) - m);
}

// containment probe of ring r: first vertex that starts a real edge (engine/engine.cuh:322-327)
__device__ __forceinline__ bool probe_vertex(const DPolys& S, uint32_t r, double shx, double shy, double& x, double& y) {
  uint64_t v = S.rvo[r];
  const uint64_t e = S.rvo[r + 1];
  if (S.synb) {
    while (v + 1 < e && syn_bit(S.synb, v)) v++;
    if (v + 1 >= e) return false;
  }
  const double2 p = S.xy[v];
  x = shx + p.x; y = shy + p.y;
  return true;
}

// Per-group scratch: edge lists and per-edge partial sums. LDS for the fast kernels, global memory for overflow.
struct Scratch {
  Ed* LA; Ed* LB;
  double* cacc;   // [capA + capB] area terms: A edges then B edges
  double* clin;   // [capA + capB] inside lengths (relate only)
  double* clon;   // [capA + capB] on lengths (relate only)
  int cap;
};

__device__ __forceinline__ Scratch carve(unsigned char* base, int cap, bool rel) {
  Scratch s;
  s.LA = reinterpret_cast<Ed*>(base);
  s.LB = s.LA + cap;
  s.cacc = reinterpret_cast<double*>(s.LB + cap);
  s.clin = rel ? s.cacc + 2 * cap : nullptr;
  s.clon = rel ? s.clin + 2 * cap : nullptr;
  s.cap = cap;
  return s;
}
__host__ __device__ constexpr size_t scratch_bytes(int cap, bool rel) { return (size_t)cap * (2 * 40 + 2 * 8 + (rel ? 4 * 8 : 0)); }

struct PairSetup {
  Box ov;
  double tol, snap, cx, cy, bshx, bshy;
};

__device__ __forceinline__ PairSetup setup_pair(const DPolys& P, const DPolys& R, uint32_t a, uint32_t b, const RefineParams& rp) {
  PairSetup s;
  const double2 oa = P.origin[a], ob = R.origin[b], ea = P.ext[a], eb = R.ext[b];
  const double sx = ob.x - oa.x, sy = ob.y - oa.y;
  const double pad = 4.0 * fmax(rp.tol_abs, rp.tol_rel * fmax(fmax(ea.x, ea.y), fmax(eb.x, eb.y)));
  s.ov = Box{fmax(0.0, sx) - pad, fmax(0.0, sy) - pad, fmin(ea.x, sx + eb.x) + pad, fmin(ea.y, sy + eb.y) + pad};
  s.bshx = sx; s.bshy = sy;
  s.tol = fmax(rp.tol_abs, rp.tol_rel * fmax(s.ov.x1 - s.ov.x0, s.ov.y1 - s.ov.y0));
  s.snap = rp.snap_half_ulp ? pair_snap(oa, ea) : 0.0;
  s.cx = s.ov.x0 + 0.381966 * (s.ov.x1 - s.ov.x0);
  s.cy = s.ov.y0 + 0.414213 * (s.ov.y1 - s.ov.y0);
  return s;
}

// the refinement of one pair by one lane group after its lists are collected; lane 0 of the group writes
template <int G>
__device__ void refine_body(const DPolys& P, const DPolys& R, uint32_t a, uint32_t b, uint64_t p, const RefineParams& rp, const PairSetup& f,
                            const Scratch& S, int na, int nb, int gl, const RefineOut& out) {
  const double2 oa = P.origin[a];
  const double a_shx = oa.x - oa.x, a_shy = oa.y - oa.y;   // A in its own frame
  // boundary contact
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
  bool inB_c = false, inA_c = false;
  if (!hit || rp.want_area || rp.want_relate) {
    inB_c = pip_full<G>(R, b, f.bshx, f.bshy, f.cx, f.cy, gl);
    inA_c = pip_full<G>(P, a, a_shx, a_shy, f.cx, f.cy, gl);
  }
  bool contain_any = false;
  if (!hit) {
    // no real boundary contact: intersects iff one contains the other; one probe vertex per ring
    bool h = false, ha = false;
    for (uint32_t r = P.gro[a] + gl; r < P.gro[a + 1] && !h; r += G) {
      double x, y;
      const bool real = probe_vertex(P, r, a_shx, a_shy, x, y);
      if (!real) { const double2 v = P.xy[P.rvo[r]]; x = a_shx + v.x; y = a_shy + v.y; }
      if (x >= f.ov.x0 && x <= f.ov.x1 && y >= f.ov.y0 && y <= f.ov.y1 && inside_from_c(S.LB, nb, f.cx, f.cy, inB_c, x, y)) { ha = true; if (real) h = true; }
    }
    for (uint32_t r = R.gro[b] + gl; r < R.gro[b + 1] && !h; r += G) {
      double x, y;
      const bool real = probe_vertex(R, r, f.bshx, f.bshy, x, y);
      if (!real) { const double2 v = R.xy[R.rvo[r]]; x = f.bshx + v.x; y = f.bshy + v.y; }
      if (x >= f.ov.x0 && x <= f.ov.x1 && y >= f.ov.y0 && y <= f.ov.y1 && inside_from_c(S.LA, na, f.cx, f.cy, inA_c, x, y)) { ha = true; if (real) h = true; }
    }
    hit = gany<G>(h);
    contain_any = gany<G>(ha);
  }
  double ar = 0.0, ai = 0.0, ao = 0.0, bi = 0.0, bo = 0.0;
  uint32_t eb4 = 0;
  const bool rel = rp.want_relate != 0;
  if ((hit || hit_any || contain_any) && (rp.want_area || rel)) {
    uint32_t bits = 0;
    for (int i = gl; i < na; i += G) {
      const Contrib c = edge_contrib(S.LA[i], S.LB, nb, f.cx, f.cy, inB_c, f.ov, f.tol, rp.eps_par, rel, f.snap);
      S.cacc[i] = c.acc;
      if (rel) { S.clin[i] = c.lin; S.clon[i] = c.lon; }
      bits |= c.bits;
    }
    for (int i = gl; i < nb; i += G) {
      const Contrib c = edge_contrib(S.LB[i], S.LA, na, f.cx, f.cy, inA_c, f.ov, f.tol, rp.eps_par, rel, f.snap);
      S.cacc[na + i] = c.acc;
      if (rel) { S.clin[na + i] = c.lin; S.clon[na + i] = c.lon; }
      bits |= c.bits;
    }
    wave_lds_sync();
    // deterministic sums in list order: lane 0 area, lanes 1..4 the relate lengths
    double s = 0.0;
    if (gl == 0) { for (int k = 0; k < na + nb; k++) s += S.cacc[k]; }
    else if (rel && gl == 1) { for (int k = 0; k < na; k++) s += S.clin[k]; }
    else if (rel && gl == 2) { for (int k = 0; k < na; k++) s += S.clon[k]; }
    else if (rel && gl == 3) { for (int k = na; k < na + nb; k++) s += S.clin[k]; }
    else if (rel && gl == 4) { for (int k = na; k < na + nb; k++) s += S.clon[k]; }
    ar = 0.5 * __shfl(s, 0, G);
    ai = __shfl(s, 1, G); ao = __shfl(s, 2, G); bi = __shfl(s, 3, G); bo = __shfl(s, 4, G);
    if (!rel) { ai = ao = bi = bo = 0.0; }
    eb4 = (rtouch ? REL_TOUCH : 0u) | (gany<G>(bits & REL_COIN_SAME) ? REL_COIN_SAME : 0u) | (gany<G>(bits & REL_COIN_OPP) ? REL_COIN_OPP : 0u) |
          (gany<G>(bits & REL_TOUCH) ? REL_TOUCH : 0u);
  } else if (rtouch) {
    eb4 = REL_TOUCH;
  }
  if (gl == 0) {
    const bool ev_any = ai != 0 || ao != 0 || bi != 0 || bo != 0;
    out.isect[p] = hit ? 1 : ((ar != 0.0 || ev_any) ? 2 : 0);
    if (out.area) out.area[p] = ar;
    if (out.ev) { out.ev[p] = make_double4(ai, ao, bi, bo); out.evb[p] = eb4; }
  }
}

struct OverflowCtl { uint32_t count; uint32_t need; uint32_t n_small; uint32_t n_large; };

// fast path: a group of G lanes per pair, lists in LDS (cap edges per side); overflowing pairs are queued
template <int G>
__global__ void __launch_bounds__(256) k_refine(const DPolys P, const DPolys R, const uint32_t* __restrict__ pa, const uint32_t* __restrict__ pb,
                                                const uint32_t* __restrict__ list, int which, RefineParams rp,
                                                RefineOut out, int cap, uint32_t* __restrict__ ovf_list, OverflowCtl* ctl) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int gpb = blockDim.x / G;
  const int gib = threadIdx.x / G, gl = threadIdx.x % G;
  const bool rel = rp.want_relate != 0;
  const Scratch S = carve(smem + (size_t)gib * scratch_bytes(cap, rel), cap, rel);
  const uint32_t nlist = which == 0 ? ctl->n_small : ctl->n_large;
  for (uint32_t k = blockIdx.x * gpb + gib; k < nlist; k += gridDim.x * gpb) {
    const uint32_t p = list[k];
    const uint32_t a = pa[p], b = pb[p];
    const PairSetup f = setup_pair(P, R, a, b, rp);
    const double2 oa = P.origin[a];
    const int na = collect<G>(P, a, oa.x - oa.x, oa.y - oa.y, f.ov, S.LA, cap, gl);
    const int nb = collect<G>(R, b, f.bshx, f.bshy, f.ov, S.LB, cap, gl);
    if (na > cap || nb > cap) {
      if (gl == 0) {
        const uint32_t slot = atomicAdd(&ctl->count, 1u);
        ovf_list[slot] = p;
        atomicMax(&ctl->need, (uint32_t)max(na, nb));
      }
      continue;
    }
    wave_lds_sync();
    refine_body<G>(P, R, a, b, p, rp, f, S, na, nb, gl, out);
    wave_lds_sync();
  }
}

// overflow path: one wave per pair, lists and partial sums in a per-wave slab of global memory
__global__ void __launch_bounds__(256) k_refine_overflow(const DPolys P, const DPolys R, const uint32_t* __restrict__ pa, const uint32_t* __restrict__ pb,
                                                         const uint32_t* __restrict__ list, uint32_t nlist, RefineParams rp, RefineOut out, int cap,
                                                         unsigned char* __restrict__ slab) {
  const int wpb = blockDim.x / 64;
  const int wib = threadIdx.x / 64, gl = threadIdx.x % 64;
  const bool rel = rp.want_relate != 0;
  const uint64_t wave = (uint64_t)blockIdx.x * wpb + wib;
  const Scratch S = carve(slab + wave * scratch_bytes(cap, rel), cap, rel);
  for (uint32_t k = (uint32_t)wave; k < nlist; k += gridDim.x * wpb) {
    const uint32_t p = list[k];
    const uint32_t a = pa[p], b = pb[p];
    const PairSetup f = setup_pair(P, R, a, b, rp);
    const double2 oa = P.origin[a];
    const int na = collect<64>(P, a, oa.x - oa.x, oa.y - oa.y, f.ov, S.LA, cap, gl);
    const int nb = collect<64>(R, b, f.bshx, f.bshy, f.ov, S.LB, cap, gl);
    wave_lds_sync();
    refine_body<64>(P, R, a, b, p, rp, f, S, min(na, cap), min(nb, cap), gl, out);
    wave_lds_sync();
  }
}

constexpr uint32_t SMALL_EDGES = 32;   // both polygons at most this many edges: 16-lane groups, LDS cap 32
constexpr int CAP_SMALL = 32, CAP_LARGE = 128;

__global__ void k_ctl_reset(OverflowCtl* ctl) { ctl->count = 0; ctl->need = 0; ctl->n_small = 0; ctl->n_large = 0; }

// split candidates into the 16-lane class and the 64-lane class (wave-aggregated appends)
__global__ void k_classify(const DPolys P, const DPolys R, const uint32_t* __restrict__ pa, const uint32_t* __restrict__ pb, uint64_t np,
                           uint32_t* __restrict__ small, uint32_t* __restrict__ large, OverflowCtl* ctl) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const bool valid = i < np;
  bool is_small = false;
  if (valid) is_small = max(P.nedge[pa[i]], R.nedge[pb[i]]) <= SMALL_EDGES;
  const int lane = threadIdx.x & 63;
  const uint64_t lt = lane ? (~0ull >> (64 - lane)) : 0ull;
  const uint64_t ms = __ballot(valid && is_small), ml = __ballot(valid && !is_small);
  uint32_t bs = 0, bl = 0;
  const int leader = __ffsll((long long)(ms | ml)) - 1;
  if (lane == leader) {
    if (ms) bs = atomicAdd(&ctl->n_small, (uint32_t)__popcll(ms));
    if (ml) bl = atomicAdd(&ctl->n_large, (uint32_t)__popcll(ml));
  }
  if (leader < 0) return;
  bs = __shfl(bs, leader, 64);
  bl = __shfl(bl, leader, 64);
  if (!valid) return;
  if (is_small) small[bs + __popcll(ms & lt)] = (uint32_t)i;
  else large[bl + __popcll(ml & lt)] = (uint32_t)i;
}

}  // namespace

uint64_t refine_pairs(Context& c, const DPolys& P, const DPolys& R, const uint32_t* pa, const uint32_t* pb, uint64_t np, const RefineParams& rp,
                      const RefineOut& out, gsj_join_stats* stats) {
  if (np == 0) return 0;
  GSJ_REQUIRE(np < (1ull << 32), GSJ_E_LIMIT, "refine: more than 2^32 pairs in one launch");
  hipStream_t st = c.stream;
  const bool rel = rp.want_relate != 0;
  uint32_t* lists = c.s_cls.as<uint32_t>(2 * np + 64);
  uint32_t* small = lists;
  uint32_t* large = lists + np;
  uint32_t* ovf = c.s_heavy.as<uint32_t>(np + 64);
  OverflowCtl* ctl = c.s_misc.as<OverflowCtl>(4) + 1;
  hipLaunchKernelGGL(k_ctl_reset, dim3(1), dim3(1), 0, st, ctl);
  hipLaunchKernelGGL(k_classify, dim3(div_up(np, 256)), dim3(256), 0, st, P, R, pa, pb, np, small, large, ctl);
  GSJ_LAUNCHED();
  const int T = 256;
  const uint32_t max_blocks = (uint32_t)c.cu_count * 16;
  GSJ_HIP(hipEventRecord(c.ev[0], st));
  {   // 16-lane groups: 16 pairs per block
    const size_t sm = (size_t)(T / 16) * scratch_bytes(CAP_SMALL, rel);
    const uint32_t g = std::min<uint32_t>(div_up(np, T / 16), max_blocks);
    hipLaunchKernelGGL(k_refine<16>, dim3(g), dim3(T), sm, st, P, R, pa, pb, small, 0, rp, out, CAP_SMALL, ovf, ctl);
    GSJ_LAUNCHED();
  }
  GSJ_HIP(hipEventRecord(c.ev[1], st));
  {   // 64-lane groups: 4 pairs per block
    const size_t sm = (size_t)(T / 64) * scratch_bytes(CAP_LARGE, rel);
    const uint32_t g = std::min<uint32_t>(div_up(np, T / 64), max_blocks);
    hipLaunchKernelGGL(k_refine<64>, dim3(g), dim3(T), sm, st, P, R, pa, pb, large, 1, rp, out, CAP_LARGE, ovf, ctl);
    GSJ_LAUNCHED();
  }
  GSJ_HIP(hipEventRecord(c.ev[2], st));
  OverflowCtl* h = c.h_scal.as<OverflowCtl>(2);
  GSJ_HIP(hipMemcpyAsync(h, ctl, sizeof(OverflowCtl), hipMemcpyDeviceToHost, st));
  GSJ_HIP(hipStreamSynchronize(st));
  const uint32_t nover = h->count, need = h->need;
  if (stats) {
    float ms_s = 0, ms_l = 0;
    GSJ_HIP(hipEventElapsedTime(&ms_s, c.ev[0], c.ev[1]));
    GSJ_HIP(hipEventElapsedTime(&ms_l, c.ev[1], c.ev[2]));
    stats->k_small_ms += ms_s; stats->k_large_ms += ms_l;
    stats->n_small += h->n_small; stats->n_large += h->n_large; stats->n_over += nover;
    stats->launches++;
  }
  if (nover) {
    const size_t per = scratch_bytes((int)need, rel);
    const uint64_t budget = 1ull << 30;
    uint64_t waves = std::max<uint64_t>(1, std::min<uint64_t>(nover, budget / std::max<size_t>(per, 1)));
    waves = (waves + 3) / 4 * 4;
    unsigned char* slab = c.s_heavy_edges.as<unsigned char>(per * waves + 64);
    GSJ_HIP(hipEventRecord(c.ev[3], st));
    hipLaunchKernelGGL(k_refine_overflow, dim3((uint32_t)(waves / 4)), dim3(T), 0, st, P, R, pa, pb, ovf, nover, rp, out, (int)need, slab);
    GSJ_LAUNCHED();
    GSJ_HIP(hipEventRecord(c.ev[4], st));
    if (stats) {
      GSJ_HIP(hipEventSynchronize(c.ev[4]));
      float ms_o = 0;
      GSJ_HIP(hipEventElapsedTime(&ms_o, c.ev[3], c.ev[4]));
      stats->k_over_ms += ms_o;
    }
  }
  return nover;
}

}  // namespace gsj
