// Distance joins on lon/lat layers: ST_DWithin(A::geography, B::geography, d) for polygon x polygon and
// point x polygon, and points in eroded polygons (reference: engine/buffer.cuh:185-273 dw_pair / k_dwithin / k_dwithin_heavy, 703-715 k_expand_bb,
// 892-991 dwithin_join_stream, 1004-1049 k_expand_pts / k_pt_dwithin, 1137-1229 points_buffer_join_stream).
//
// Candidates: probe boxes widened by d metres (degrees from the WGS84 radii at the box's worst latitude), rounded
// outward to float, traversed against the build layer's LBVH. Per candidate piece pair: the real edges of both
// pieces inside the widened overlap box (LDS lists, ballot-compacted), the closest distance of every near edge pair
// in a local metric frame at the window's latitude, re-measured as a geography distance (great-circle closest points
// + Vincenty) when it falls inside the frame's error band around d; containment by crossing parity when no boundary
// is near. With report_distance, the distance of every row: 0 when the boundaries cross or one piece contains the
// other, else the minimum geography distance over the edge pairs that could hold the minimum (new output; the
// reference reports NaN). Piece pairs fold to unique parent pairs with the minimum distance.
//
// Work classes as in the refinement: small pairs (both pieces <= 32 stored vertices) on 16-lane groups with 32-edge
// LDS lists, larger pairs on a wave with 128-edge lists, overflowing pairs on a wave with lists in global memory.
// All transcendental math is the portable set of gmath.h, so every threshold decision and distance has the bits of
// the CPU restatement.
#include "pairs.h"
#include "geodesy.h"

namespace gsj {

namespace {

// probe boxes of polygons widened by d metres (buffer.cuh:703-715)
__device__ __forceinline__ float4 expand_poly_box(const DPolys& P, uint32_t i, double d) {
  const double2 o = P.origin[i], e = P.ext[i];
  double dx, dy;
  geo::deg_cover(o.y, o.y + e.y, d, true, dx, dy);
  return make_float4(geo::f_down(o.x - dx), geo::f_down(o.y - dy), geo::f_up(o.x + e.x + dx), geo::f_up(o.y + e.y + dy));
}
__global__ void k_expand_polys(const DPolys P, double d, float4* __restrict__ out) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < P.n) out[i] = expand_poly_box(P, i, d);
}
// only the listed probes (a window's members, distinct ids), at their ids
__global__ void k_expand_polys_list(const DPolys P, double d, const uint32_t* __restrict__ list, uint32_t n, float4* __restrict__ out) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[list[i]] = expand_poly_box(P, list[i], d);
}
// batched windows: the member at position i belongs to request r (mstart[r] <= i < mstart[r + 1]) and is widened by
// that request's distance; boxes by position
struct DistSet {
  uint32_t k;
  double d[MAX_WINDOWS];
};
__global__ void k_expand_members(const DPolys P, const uint32_t* __restrict__ list, uint32_t n, const uint64_t* __restrict__ mstart, const DistSet ds,
                                 float4* __restrict__ out) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  uint32_t r = 0;
  while (r + 1 < ds.k && mstart[r + 1] <= i) r++;
  out[i] = expand_poly_box(P, list[i], ds.d[r]);
}

// probe boxes of points widened by d metres (buffer.cuh:1004-1016)
__global__ void k_expand_points(const DPoints pts, double d, float4* __restrict__ out) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= pts.n) return;
  const double2 p = pts.xy[i];
  double dx, dy;
  geo::deg_cover(p.y, p.y, d, true, dx, dy);
  out[i] = make_float4(geo::f_down(p.x - dx), geo::f_down(p.y - dy), geo::f_up(p.x + dx), geo::f_up(p.y + dy));
}

// real probe vertex of ring r (engine/engine.cuh:322-327); false for an empty ring or one without real edges
__device__ __forceinline__ bool probe_real(const DPolys& S, uint32_t r, double shx, double shy, double& x, double& y) {
  if (S.rvo[r] == S.rvo[r + 1]) return false;
  return probe_vertex(S, r, shx, shy, x, y);
}

// A piece read through its packed record (DPolys::pk; the vertex-list classes' pieces, at most PACK_VCAP stored
// vertices and PACK_RCAP rings): vertices, flag bytes (VF_*) and ring starts of one contiguous record. The functions
// below give exactly what collect / pip_full / probe_real give from the layer's arrays (same edges in the same order,
// same arithmetic), from loads of one record instead of the geometry -> ring offsets -> vertices chain.
struct RecView {
  const double2* V;
  const uint8_t* F;
  const uint8_t* rs;   // ring starts, rs[nr] = nv
  double2 o, e;        // origin and extent (the layer's, copied into the record)
  int nv, nr;
};
__device__ __forceinline__ RecView rec_view(const DPolys& S, uint32_t g) {
  const uint4* rec = S.pk + S.pk_off[g];
  const uint8_t* h = reinterpret_cast<const uint8_t*>(rec + 2);
  RecView v;
  v.o = reinterpret_cast<const double2*>(rec)[0];
  v.e = reinterpret_cast<const double2*>(rec)[1];
  v.nv = h[0];
  v.nr = h[1];
  v.rs = h + 2;
  v.V = reinterpret_cast<const double2*>(rec + 3);
  v.F = reinterpret_cast<const uint8_t*>(rec + 3 + v.nv);
  return v;
}
// whether vertex k starts an edge (k + 1 within its ring)
__device__ __forceinline__ bool rec_edge(const RecView& v, int k) {
  int r = 0;
  while (r + 1 < v.nr && v.rs[r + 1] <= k) r++;
  return k + 1 < v.rs[r + 1];
}
// collect() of both pieces at once: their edges meeting ov, in ring order, ballot-compacted into LA / LB; every load
// of both records (PER vertices, their successors and flag bytes per lane) is issued before the first list write.
// Returns the full counts (more than cap: the list is incomplete and the pair overflows).
template <int G, int PER>
__device__ void collect_rec2(const RecView& A, double ash, double ashy, const RecView& B, double bsh, double bshy, const Box& ov, Ed* LA, Ed* LB,
                             int cap, int gl, int& na, int& nb) {
  double2 pa[PER], qa[PER], pb[PER], qb[PER];
  uint8_t fa[PER], fb[PER];
  bool ea[PER], eb[PER];
#pragma unroll
  for (int j = 0; j < PER; j++) {
    const int k = j * G + gl;
    ea[j] = k < A.nv && rec_edge(A, k);
    eb[j] = k < B.nv && rec_edge(B, k);
    if (ea[j]) { pa[j] = A.V[k]; qa[j] = A.V[k + 1]; fa[j] = A.F[k]; }
    if (eb[j]) { pb[j] = B.V[k]; qb[j] = B.V[k + 1]; fb[j] = B.F[k]; }
  }
  int ca = 0, cb = 0;
#pragma unroll
  for (int j = 0; j < PER; j++) {
    Ed x, y;
    bool va = false, vb = false;
    if (ea[j]) {
      x.x0 = ash + pa[j].x; x.y0 = ashy + pa[j].y; x.x1 = ash + qa[j].x; x.y1 = ashy + qa[j].y;
      x.syn = (fa[j] & VF_SYN) ? 1 : 0;
      x.cut = (fa[j] >> 2) & 3;
      va = !edge_outside(x.x0, x.y0, x.x1, x.y1, ov);
    }
    if (eb[j]) {
      y.x0 = bsh + pb[j].x; y.y0 = bshy + pb[j].y; y.x1 = bsh + qb[j].x; y.y1 = bshy + qb[j].y;
      y.syn = (fb[j] & VF_SYN) ? 1 : 0;
      y.cut = (fb[j] >> 2) & 3;
      vb = !edge_outside(y.x0, y.y0, y.x1, y.y1, ov);
    }
    const uint64_t ma = gballot<G>(va), mb = gballot<G>(vb), below = (1ull << gl) - 1;
    const int posa = ca + __popcll(ma & below), posb = cb + __popcll(mb & below);
    if (va && posa < cap) LA[posa] = x;
    if (vb && posb < cap) LB[posb] = y;
    ca += __popcll(ma);
    cb += __popcll(mb);
  }
  na = ca;
  nb = cb;
}
// pip_full() over the record's edges
template <int G>
__device__ bool pip_rec(const RecView& S, double shx, double shy, double mx, double my, int gl) {
  const double ox = shx - 1.0, oy = my;
  bool par = false;
  for (int k = gl; k < S.nv; k += G) {
    if (!rec_edge(S, k)) continue;
    const double2 p = S.V[k], q = S.V[k + 1];
    const double x0 = shx + p.x, y0 = shy + p.y, x1 = shx + q.x, y1 = shy + q.y;
    if ((y0 > my) == (y1 > my) && y0 != my && y1 != my) continue;
    const bool so = left_of(x0, y0, x1, y1, ox, oy), sc = left_of(x0, y0, x1, y1, mx, my);
    if (so == sc) continue;
    const bool su = orient(ox, oy, mx, my, x0, y0) > 0, sv = orient(ox, oy, mx, my, x1, y1) > 0;
    if (su != sv) par = !par;
  }
  return (__popcll(gballot<G>(par)) & 1) != 0;
}
// probe_real() of ring r of the record (has_syn: the layer has synthetic edges)
__device__ __forceinline__ bool probe_rec(const RecView& S, int r, bool has_syn, double shx, double shy, double& x, double& y) {
  const int s = S.rs[r], e = S.rs[r + 1];
  if (s == e) return false;
  int v = s;
  if (has_syn) {
    while (v + 1 < e && (S.F[v] & VF_SYN)) v++;
    if (v + 1 >= e) return false;
  }
  const double2 p = S.V[v];
  x = shx + p.x;
  y = shy + p.y;
  return true;
}

// The real edges of L[base .. min(base + 64, n)) whose boxes meet [x0, x1] x [y0, y1], one bit per entry. The pair
// tests below run over these bits in list order: the box test alone runs once per entry, and the distance work runs
// together across a wave's lanes, about as many times as the most crowded lane has neighbours.
__device__ __forceinline__ uint64_t near_mask(const Ed* L, int base, int n, double x0, double x1, double y0, double y1) {
  const int m = min(64, n - base);
  uint64_t cand = 0;
  for (int j = 0; j < m; j++) {
    const Ed f = L[base + j];
    if (!(f.syn || fmax(f.x0, f.x1) < x0 || fmin(f.x0, f.x1) > x1 || fmax(f.y0, f.y1) < y0 || fmin(f.y0, f.y1) > y1)) cand |= 1ull << j;
  }
  return cand;
}

// One piece pair on a lane group (buffer.cuh:185-237). Returns 1 / 0, or -1 when a side has more than cap edges in
// the window (need = edges required). With report and a hit, dist = the pair's distance (see the file comment).
// REC: both pieces have packed records with at most PER * G stored vertices, read instead of the layer's arrays.
template <int G, bool REC, int PER = 1>
__device__ int dw_eval(const DPolys& P, const DPolys& R, uint32_t a, uint32_t b, double d, bool report, Ed* LA, Ed* LB, int cap, int gl,
                       double& dist, int& need) {
  RecView VA{}, VB{};
  if constexpr (REC) { VA = rec_view(P, a); VB = rec_view(R, b); }
  const double2 oa = REC ? VA.o : P.origin[a], ob = REC ? VB.o : R.origin[b], ea = REC ? VA.e : P.ext[a], eb = REC ? VB.e : R.ext[b];
  const double sx = ob.x - oa.x, sy = ob.y - oa.y;
  double gx, gy;
  geo::deg_cover(fmin(oa.y, ob.y), fmax(oa.y + ea.y, ob.y + eb.y), d, true, gx, gy);
  const Box ov{fmax(-gx, sx - gx), fmax(-gy, sy - gy), fmin(ea.x + gx, sx + eb.x + gx), fmin(ea.y + gy, sy + eb.y + gy)};
  if (ov.x0 > ov.x1 || ov.y0 > ov.y1) return 0;
  const double ash = oa.x - oa.x, ashy = oa.y - oa.y;   // A in its own frame
  int na, nb;
  if constexpr (REC) collect_rec2<G, PER>(VA, ash, ashy, VB, sx, sy, ov, LA, LB, cap, gl, na, nb);
  else { na = collect<G>(P, a, ash, ashy, ov, LA, cap, gl); nb = collect<G>(R, b, sx, sy, ov, LB, cap, gl); }
  if (na > cap || nb > cap) { need = max(na, nb); return -1; }
  wave_lds_sync();
  const geo::Frame F = geo::frame_of(oa.y + ov.y0, oa.y + ov.y1);   // metric frame of the pair's window
  const double kx = F.kx, ky = F.ky;
  bool hit = false, cross = false;
  double m_up = INFINITY;
  // A's edges G at a time, the group stopping together once the pair is settled: a hit anywhere settles "within" (and,
  // without report, the row); with report, a crossing anywhere settles the distance (0), so the remaining edge pairs
  // cannot change the row either
  for (int i0 = 0; i0 < na; i0 += G) {
    const int i = i0 + gl;
    const bool go = i < na && (report || !hit) && !LA[i].syn;
    if (go) {
      const Ed e = LA[i];
      const double bx0 = fmin(e.x0, e.x1) - gx, bx1 = fmax(e.x0, e.x1) + gx, by0 = fmin(e.y0, e.y1) - gy, by1 = fmax(e.y0, e.y1) + gy;
      const double ux = (e.x1 - e.x0) * kx, uy = (e.y1 - e.y0) * ky, La2 = ux * ux + uy * uy;
      for (int base = 0; base < nb && !cross && (report || !hit); base += 64) {
        uint64_t cand = near_mask(LB, base, nb, bx0, bx1, by0, by1);
        while (cand != 0) {
          const int j = base + __builtin_ctzll(cand);
          cand &= cand - 1;
          const Ed f = LB[j];
          if (seg_isect(e.x0, e.y0, e.x1, e.y1, f.x0, f.y0, f.x1, f.y1)) {   // settles the row, distance 0 included
            hit = cross = true;
            break;
          }
          const double d2 = geo::seg_seg_d2(e.x0 * kx, e.y0 * ky, e.x1 * kx, e.y1 * ky, f.x0 * kx, f.y0 * ky, f.x1 * kx, f.y1 * ky);
          const double vx = (f.x1 - f.x0) * kx, vy = (f.y1 - f.y0) * ky;
          const double band = geo::geog_band(d, F, La2, vx * vx + vy * vy);
          const double dlo = d - band, dhi = d + band;
          if (d2 <= dlo * dlo) hit = true;
          else if (!hit && d2 <= dhi * dhi &&   // (with report, a lane already within d needs no geography distance here)
                   geo::geog_dist(oa.x + e.x0, oa.y + e.y0, oa.x + e.x1, oa.y + e.y1, oa.x + f.x0, oa.y + f.y0, oa.x + f.x1, oa.y + f.y1, false) <= d)
            hit = true;
          if (report) m_up = fmin(m_up, sqrt(d2) + band);
          else if (hit) break;
        }
      }
    }
    if (gany<G>(report ? cross : hit)) break;
  }
  hit = gany<G>(hit);
  cross = gany<G>(cross);
  bool contained = false;
  if (!hit || report) {
    // no boundary within d: within d iff one contains the other (a real vertex of one inside the other piece)
    const double cx = ov.x0 + 0.381966 * (ov.x1 - ov.x0), cy = ov.y0 + 0.414213 * (ov.y1 - ov.y0);
    bool inB_c, inA_c;
    if constexpr (REC) { inB_c = pip_rec<G>(VB, sx, sy, cx, cy, gl); inA_c = pip_rec<G>(VA, ash, ashy, cx, cy, gl); }
    else { inB_c = pip_full<G>(R, b, sx, sy, cx, cy, gl); inA_c = pip_full<G>(P, a, ash, ashy, cx, cy, gl); }
    const ListSrc SA{LA, na}, SB{LB, nb};
    bool h = false;
    if constexpr (REC) {
      for (int r = gl; r < VA.nr && !h; r += G) {
        double x, y;
        if (probe_rec(VA, r, P.synb != nullptr, ash, ashy, x, y) && x >= ov.x0 && x <= ov.x1 && y >= ov.y0 && y <= ov.y1 &&
            inside_from_c(SB, cx, cy, inB_c, x, y))
          h = true;
      }
      for (int r = gl; r < VB.nr && !h; r += G) {
        double x, y;
        if (probe_rec(VB, r, R.synb != nullptr, sx, sy, x, y) && x >= ov.x0 && x <= ov.x1 && y >= ov.y0 && y <= ov.y1 &&
            inside_from_c(SA, cx, cy, inA_c, x, y))
          h = true;
      }
    } else {
      for (uint32_t r = P.gro[a] + gl; r < P.gro[a + 1] && !h; r += G) {
        double x, y;
        if (probe_real(P, r, ash, ashy, x, y) && x >= ov.x0 && x <= ov.x1 && y >= ov.y0 && y <= ov.y1 && inside_from_c(SB, cx, cy, inB_c, x, y)) h = true;
      }
      for (uint32_t r = R.gro[b] + gl; r < R.gro[b + 1] && !h; r += G) {
        double x, y;
        if (probe_real(R, r, sx, sy, x, y) && x >= ov.x0 && x <= ov.x1 && y >= ov.y0 && y <= ov.y1 && inside_from_c(SA, cx, cy, inA_c, x, y)) h = true;
      }
    }
    contained = gany<G>(h);
  }
  const bool within = hit || contained;
  if (report && within) {
    if (cross || contained) dist = 0.0;
    else {
      // minimum geography distance over the edge pairs whose frame distance could hold it
      const double m = group_min<G>(m_up);
      double best = INFINITY;
      for (int i = gl; i < na; i += G) {
        const Ed e = LA[i];
        if (e.syn) continue;
        const double bx0 = fmin(e.x0, e.x1) - gx, bx1 = fmax(e.x0, e.x1) + gx, by0 = fmin(e.y0, e.y1) - gy, by1 = fmax(e.y0, e.y1) + gy;
        const double ux = (e.x1 - e.x0) * kx, uy = (e.y1 - e.y0) * ky, La2 = ux * ux + uy * uy;
        for (int base = 0; base < nb; base += 64) {
          uint64_t cand = near_mask(LB, base, nb, bx0, bx1, by0, by1);
          while (cand != 0) {
            const int j = base + __builtin_ctzll(cand);
            cand &= cand - 1;
            const Ed f = LB[j];
            const double d2 = geo::seg_seg_d2(e.x0 * kx, e.y0 * ky, e.x1 * kx, e.y1 * ky, f.x0 * kx, f.y0 * ky, f.x1 * kx, f.y1 * ky);
            const double vx = (f.x1 - f.x0) * kx, vy = (f.y1 - f.y0) * ky;
            const double band = geo::geog_band(d, F, La2, vx * vx + vy * vy);
            if (sqrt(d2) - band > m) continue;
            best = fmin(best, geo::geog_dist(oa.x + e.x0, oa.y + e.y0, oa.x + e.x1, oa.y + e.y1, oa.x + f.x0, oa.y + f.y0, oa.x + f.x1, oa.y + f.y1,
                                             false));
          }
        }
      }
      dist = group_min<G>(best);
    }
  }
  return within ? 1 : 0;
}

struct DwArgs {
  DPolys P, R;
  const uint32_t* pa;
  const uint32_t* pb;
  double d;
  int same, report;
  uint8_t* hit;
  double* dist;
  // batched requests: candidate p belongs to request r when seg[r] <= p < seg[r + 1], distance ds.d[r]
  const uint64_t* seg;
  DistSet ds;
};
__device__ __forceinline__ double pair_distance(const DwArgs& A, uint32_t p) {
  if (!A.seg) return A.d;
  uint32_t r = 0;
  while (r + 1 < A.ds.k && A.seg[r + 1] <= p) r++;
  return A.ds.d[r];
}

// pieces of one parent are within any distance of each other on a self-join (buffer.cuh:246-249)
__device__ __forceinline__ bool self_piece_pair(const DwArgs& A, uint32_t a, uint32_t b) {
  if (!A.same) return false;
  const uint32_t qa = A.P.parent ? A.P.parent[a] : a, qb = A.R.parent ? A.R.parent[b] : b;
  return qa == qb;
}

// list classes: a G-lane group per pair, CAP-edge LDS lists per side; class `which` (PairClass) of the partition;
// pairs whose lists overflow are queued. REC: the class's pieces have packed records (tiny, small, medium: at most CAP
// stored vertices each), read instead of the layer's arrays.
template <int G, int CAP, bool REC>
__global__ void __launch_bounds__(256) k_dw_lists(const DwArgs A, const uint32_t* __restrict__ list, const uint32_t* __restrict__ counts, int which,
                                                  uint32_t* __restrict__ ovf_list, OverflowCtl* ctl) {
  __shared__ Ed lds[256 / G][2 * CAP];
  const int gpb = blockDim.x / G;
  const int gib = threadIdx.x / G, gl = threadIdx.x % G;
  Ed* LA = lds[gib];
  Ed* LB = LA + CAP;
  const uint32_t n = counts[which];
  const uint32_t* L = list + class_first(counts, which);
  PairStream ps(L, A.pa, A.pb, n, blockIdx.x * gpb + gib, gridDim.x * gpb);
  for (; ps.k < n; ps.advance()) {
    ps.prefetch();
    const uint32_t p = ps.p, a = ps.a, b = ps.b;
    if (self_piece_pair(A, a, b)) {
      if (gl == 0) { A.hit[p] = 1; if (A.dist) A.dist[p] = 0.0; }
      continue;
    }
    double dist = NAN;
    int need = 0;
    const int r = dw_eval<G, REC, CAP / G>(A.P, A.R, a, b, pair_distance(A, p), A.report != 0, LA, LB, CAP, gl, dist, need);
    if (gl == 0) {
      if (r < 0) queue_overflow(p, need, ovf_list, ctl);
      else {
        A.hit[p] = (uint8_t)r;
        if (A.dist) A.dist[p] = r ? dist : NAN;
      }
    }
    wave_lds_sync();
  }
}

// overflow: a wave per pair, lists in a per-wave slab of global memory
__global__ void __launch_bounds__(256) k_dw_overflow(const DwArgs A, const uint32_t* __restrict__ list, uint32_t n, int cap, Ed* __restrict__ slab) {
  const int wpb = blockDim.x / 64;
  const int wib = threadIdx.x / 64, gl = threadIdx.x % 64;
  const uint64_t wave = (uint64_t)blockIdx.x * wpb + wib;
  Ed* LA = slab + wave * 2 * (uint64_t)cap;
  Ed* LB = LA + cap;
  for (uint32_t k = (uint32_t)wave; k < n; k += gridDim.x * wpb) {
    const uint32_t p = list[k];
    double dist = NAN;
    int need = 0;
    const int r = dw_eval<64, false>(A.P, A.R, A.pa[p], A.pb[p], pair_distance(A, p), A.report != 0, LA, LB, cap, gl, dist, need);
    if (gl == 0) {
      A.hit[p] = r > 0 ? 1 : 0;
      if (A.dist) A.dist[p] = r > 0 ? dist : NAN;
    }
    wave_lds_sync();
  }
}

// thread per (point, piece): inside the piece, or within d of its real boundary (buffer.cuh:1018-1049); with report,
// the distance (0 inside, else the minimum geography distance over the edges that could hold it)
__global__ void __launch_bounds__(256) k_pt_dwithin(const DPoints pts, const DPolys R, const uint32_t* __restrict__ pa,
                                                    const uint32_t* __restrict__ pb, uint64_t np, double d, int report, uint8_t* __restrict__ hit,
                                                    double* __restrict__ dist) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= np) return;
  const uint32_t g = pb[i];
  const double2 pt = pts.xy[pa[i]], o = R.origin[g];
  const double px = pt.x - o.x, py = pt.y - o.y;
  const double reach = d / geo::KY_MIN;   // latitude the closest boundary point can be from the point
  const geo::Frame F = geo::frame_of(pt.y - reach, pt.y + reach);
  const double kx = F.kx, ky = F.ky;
  const double ox = (o.x - o.x) - 1.0;
  bool par = false, h = false;
  double m_up = INFINITY;
  const uint32_t r1 = R.gro[g + 1];
  for (uint32_t r = R.gro[g]; r < r1 && (report || !h); r++) {
    const uint64_t s0 = R.rvo[r], e = R.rvo[r + 1];
    double2 v0 = R.xy[s0];
    for (uint64_t j = s0 + 1; j < e && (report || !h); j++) {
      const double2 v1 = R.xy[j];
      const double x0 = v0.x, y0 = v0.y, x1 = v1.x, y1 = v1.y;
      if (!syn_bit(R.synb, j - 1)) {
        const double d2 = geo::pt_seg_d2(px * kx, py * ky, x0 * kx, y0 * ky, x1 * kx, y1 * ky);
        const double ex = (x1 - x0) * kx, ey = (y1 - y0) * ky;
        const double band = geo::geog_band(d, F, ex * ex + ey * ey, 0.0), dlo = d - band, dhi = d + band;
        if (d2 <= dlo * dlo || (d2 <= dhi * dhi && geo::geog_dist(o.x + x0, o.y + y0, o.x + x1, o.y + y1, pt.x, pt.y, pt.x, pt.y, true) <= d)) h = true;
        if (report) m_up = fmin(m_up, sqrt(d2) + band);
      }
      if (!((y0 > py) == (y1 > py) && y0 != py && y1 != py)) {
        const bool so = left_of(x0, y0, x1, y1, ox, py), sp = left_of(x0, y0, x1, y1, px, py);
        if (so != sp) {
          const bool su = orient(ox, py, px, py, x0, y0) > 0, sv = orient(ox, py, px, py, x1, y1) > 0;
          if (su != sv) par = !par;
        }
      }
      v0 = v1;
    }
  }
  const bool within = h || par;
  hit[i] = within ? 1 : 0;
  if (!dist) return;
  if (!within) { dist[i] = NAN; return; }
  if (par) { dist[i] = 0.0; return; }
  double best = INFINITY;
  for (uint32_t r = R.gro[g]; r < r1; r++) {
    const uint64_t s0 = R.rvo[r], e = R.rvo[r + 1];
    for (uint64_t j = s0 + 1; j < e; j++) {
      if (syn_bit(R.synb, j - 1)) continue;
      const double2 v0 = R.xy[j - 1], v1 = R.xy[j];
      const double d2 = geo::pt_seg_d2(px * kx, py * ky, v0.x * kx, v0.y * ky, v1.x * kx, v1.y * ky);
      const double ex = (v1.x - v0.x) * kx, ey = (v1.y - v0.y) * ky;
      const double band = geo::geog_band(d, F, ex * ex + ey * ey, 0.0);
      if (sqrt(d2) - band > m_up) continue;
      best = fmin(best, geo::geog_dist(o.x + v0.x, o.y + v0.y, o.x + v1.x, o.y + v1.y, pt.x, pt.y, pt.x, pt.y, true));
    }
  }
  dist[i] = best;
}

// ---- points in eroded polygons (buffer.cuh:1053-1131) ---------------------------------------------------------------
constexpr double FILLET_QUANTUM = gm::PI_HI / 16.0;   // GEOS: 8 quadrant segments

// w (metres from vertex v) strictly inside v's region between unit directions e -> f: convex, the disk outside the
// interior wedge; reflex, the GEOS fillet polygon (buffer.cuh:1056-1078)
__device__ bool vertex_excl(double ex, double ey, double fx, double fy, double r, double wx, double wy, double eps) {
  const double d2 = wx * wx + wy * wy;
  if (d2 >= r * r) return false;
  const double cr = ex * fy - ey * fx, dt = ex * fx + ey * fy;
  if (cr > 0) {
    const double ri = r - eps, s1 = wy * ex - wx * ey, s2 = wy * fx - wx * fy;
    return d2 < ri * ri && !(s1 > eps && s2 > eps);
  }
  if (cr == 0 && dt >= 0) return false;   // straight
  const double th = cr == 0 ? gm::PI_HI : gm::atan2(-cr, dt);
  int m = (int)(th / FILLET_QUANTUM + 0.5);
  if (m < 1) m = 1;
  const double inc = th / m;
  double si, ci;
  gm::sincos(inc, si, ci);
  const double nkx = -ey, nky = ex, nsx = -fy, nsy = fx;
  if (nkx * wy - nky * wx > eps || nsx * wy - nsy * wx < -eps) return false;
  const double chh = sqrt(0.5 * (1.0 + ci)), shh = sqrt(fmax(0.0, 0.5 * (1.0 - ci)));
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

// thread per (point, piece) widened by r: 1 = inside the piece, 2 = within r of its real boundary (excluded),
// 0 = outside (buffer.cuh:1080-1128)
__global__ void __launch_bounds__(256) k_pt_erode(const DPoints pts, const DPolys R, const uint32_t* __restrict__ pa, const uint32_t* __restrict__ pb,
                                                  uint64_t np, double r_nom, uint8_t* __restrict__ bits) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= np) return;
  const uint32_t g = pb[i];
  const double2 pt = pts.xy[pa[i]], o = R.origin[g], ext = R.ext[g];
  const double r = r_nom / geo::utm_k(o.x + 0.5 * ext.x, o.y + 0.5 * ext.y, pt.x, pt.y);
  double kx, ky;
  geo::m_per_deg(pt.y, kx, ky);
  const double eps = 1e-9;
  const double shx = o.x - pt.x, shy = o.y - pt.y, ox = shx * kx - 1.0;   // frame: metres relative to the point
  auto vx = [&](uint64_t j, double& x, double& y) { const double2 q = R.xy[j]; x = (shx + q.x) * kx; y = (shy + q.y) * ky; };
  bool par = false, ex = false;
  for (uint32_t rr = R.gro[g]; rr < R.gro[g + 1] && !ex; rr++) {
    const uint64_t s0 = R.rvo[rr], e = R.rvo[rr + 1];
    if (e - s0 < 2) continue;
    // previous real edge direction (region of the vertex that starts the current edge); the ring closes
    bool pv = false;
    double pux = 0, puy = 0;
    {
      const uint64_t j = e - 2;
      if (!syn_bit(R.synb, j)) {
        double a, b, c, dd;
        vx(j, a, b);
        vx(j + 1, c, dd);
        const double L = sqrt((c - a) * (c - a) + (dd - b) * (dd - b));
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
      const double L = sqrt((x1 - x0) * (x1 - x0) + (y1 - y0) * (y1 - y0));
      const bool sj = syn_bit(R.synb, j) != 0;
      if (sj || L == 0) {
        if (sj) pv = false;
        x0 = x1; y0 = y1;
        continue;
      }
      const double ux = (x1 - x0) / L, uy = (y1 - y0) / L;
      if (r > 0) {
        const double t = -x0 * ux - y0 * uy, s = -y0 * ux + x0 * uy;   // the point (0, 0) relative to the edge
        if (fabs(s) < r - eps && t > -eps && t < L + eps) ex = true;
        else if (pv && vertex_excl(pux, puy, ux, uy, r, -x0, -y0, eps)) ex = true;
      }
      pv = true; pux = ux; puy = uy;
      x0 = x1; y0 = y1;
    }
  }
  bits[i] = ex ? 2 : (par ? 1 : 0);   // the scan stops early only once the point is known to be excluded
}

__global__ void k_nonzero_u32(const uint8_t* __restrict__ b, uint64_t n, uint32_t* __restrict__ f) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) f[i] = b[i] ? 1u : 0u;
}
__global__ void k_one_u32(const uint8_t* __restrict__ b, uint64_t n, uint32_t* __restrict__ f) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) f[i] = b[i] == 1 ? 1u : 0u;
}
// (point, parent) keys of the pairs with nonzero bits
__global__ void k_pe_keys(const uint32_t* __restrict__ Rpar, const uint32_t* __restrict__ pa, const uint32_t* __restrict__ pb,
                          const uint8_t* __restrict__ bits, const uint32_t* __restrict__ f, const uint64_t* __restrict__ off, uint64_t n,
                          uint64_t* __restrict__ keys, uint32_t* __restrict__ vals) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n || !f[i]) return;
  keys[off[i]] = ((uint64_t)pa[i] << 32) | (Rpar ? Rpar[pb[i]] : pb[i]);
  vals[off[i]] = bits[i];
}
// OR of the bits per key; a row where the OR is 1 (inside some piece, near no real boundary)
__global__ void k_pe_fold(const uint64_t* __restrict__ keys, const uint32_t* __restrict__ vals, uint64_t m, uint32_t* __restrict__ keep) {
  const uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= m) return;
  if (j > 0 && keys[j] == keys[j - 1]) { keep[j] = 0; return; }
  uint32_t b = 0;
  for (uint64_t k = j; k < m && keys[k] == keys[j]; k++) b |= vals[k];
  keep[j] = b == 1 ? 1u : 0u;
}
__global__ void k_pe_emit(const uint64_t* __restrict__ keys, const uint32_t* __restrict__ keep, const uint64_t* __restrict__ off, uint64_t m,
                          uint32_t* __restrict__ orow) {
  const uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= m || !keep[j]) return;
  orow[2 * off[j]] = (uint32_t)(keys[j] >> 32);
  orow[2 * off[j] + 1] = (uint32_t)keys[j];
}
__global__ void k_pe_direct(const uint32_t* __restrict__ pa, const uint32_t* __restrict__ pb, const uint32_t* __restrict__ f,
                            const uint64_t* __restrict__ off, uint64_t n, uint32_t* __restrict__ orow) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n || !f[i]) return;
  orow[2 * off[i]] = pa[i];
  orow[2 * off[i] + 1] = pb[i];
}

// ---- compaction: unique (parent a, parent b) rows of the hit pairs, minimum distance -----------------------------
__global__ void k_hit_u32(const uint8_t* __restrict__ hit, uint64_t n, uint32_t* __restrict__ f) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) f[i] = hit[i] ? 1u : 0u;
}
__global__ void k_dw_keys(const uint32_t* __restrict__ Ppar, const uint32_t* __restrict__ Rpar, const uint32_t* __restrict__ pa,
                          const uint32_t* __restrict__ pb, const uint32_t* __restrict__ f, const uint64_t* __restrict__ off, uint64_t n,
                          uint64_t* __restrict__ keys, uint32_t* __restrict__ vals) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n || !f[i]) return;
  const uint64_t a = Ppar ? Ppar[pa[i]] : pa[i], b = Rpar ? Rpar[pb[i]] : pb[i];
  keys[off[i]] = (a << 32) | b;
  vals[off[i]] = (uint32_t)i;
}
// layers without parents: every candidate is its own parent pair, rows in candidate order
__global__ void k_dw_direct(const uint32_t* __restrict__ pa, const uint32_t* __restrict__ pb, const uint32_t* __restrict__ f,
                            const uint64_t* __restrict__ off, uint64_t n, const double* __restrict__ dist, uint32_t* __restrict__ orow,
                            double* __restrict__ odist) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n || !f[i]) return;
  const uint64_t o = off[i];
  orow[2 * o] = pa[i];
  orow[2 * o + 1] = pb[i];
  if (odist) odist[o] = dist[i];
}
__global__ void k_key_heads(const uint64_t* __restrict__ keys, uint64_t m, uint32_t* __restrict__ head) {
  const uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j < m) head[j] = (j == 0 || keys[j] != keys[j - 1]) ? 1u : 0u;
}
__global__ void k_dw_fold(const uint64_t* __restrict__ keys, const uint32_t* __restrict__ vals, const uint32_t* __restrict__ head,
                          const uint64_t* __restrict__ hoff, uint64_t m, const double* __restrict__ dist, uint32_t* __restrict__ orow,
                          double* __restrict__ odist) {
  const uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= m || !head[j]) return;
  const uint64_t o = hoff[j];
  orow[2 * o] = (uint32_t)(keys[j] >> 32);
  orow[2 * o + 1] = (uint32_t)keys[j];
  if (!dist) return;
  double v = INFINITY;
  for (uint64_t k = j; k < m && keys[k] == keys[j]; k++) v = fmin(v, dist[vals[k]]);
  odist[o] = v;
}

// rows of one sub-batch on the device; returns the row count
uint64_t dw_compact(Context& c, const uint32_t* Ppar, const uint32_t* Rpar, const uint32_t* pa, const uint32_t* pb, uint64_t np, const uint8_t* hit,
                    const double* dist, uint32_t* orow, double* odist, const uint64_t** direct_offs = nullptr) {
  hipStream_t st = c.stream;
  uint32_t* f = c.s_tmp.as<uint32_t>(np + 16);
  uint64_t* off = c.s_offs.as<uint64_t>(np + 16);
  hipLaunchKernelGGL(k_hit_u32, dim3(div_up(np, 256)), dim3(256), 0, st, hit, np, f);
  GSJ_LAUNCHED();
  const uint64_t nh = scan_u32_to_u64(c, f, off, np);
  if (!Ppar && !Rpar) {   // a probe meets each leaf once, so the candidates are distinct parent pairs
    if (direct_offs) *direct_offs = off;
    if (nh) {
      hipLaunchKernelGGL(k_dw_direct, dim3(div_up(np, 256)), dim3(256), 0, st, pa, pb, f, off, np, dist, orow, odist);
      GSJ_LAUNCHED();
    }
    return nh;
  }
  GSJ_REQUIRE(!direct_offs, GSJ_E_INTERNAL, "candidate-ordered rows need layers without parents");
  if (nh == 0) return 0;
  uint64_t* keys = c.s_keys.as<uint64_t>(nh + 16);
  uint64_t* keys2 = c.s_keys2.as<uint64_t>(nh + 16);
  uint32_t* vals = c.s_vals.as<uint32_t>(nh + 16);
  uint32_t* vals2 = c.s_vals2.as<uint32_t>(nh + 16);
  hipLaunchKernelGGL(k_dw_keys, dim3(div_up(np, 256)), dim3(256), 0, st, Ppar, Rpar, pa, pb, f, off, np, keys, vals);
  GSJ_LAUNCHED();
  radix_sort_pairs(c, keys, keys2, vals, vals2, nh, 0, 64);
  uint32_t* head = c.s_seg.as<uint32_t>(nh + 16);
  uint64_t* hoff = c.s_tmp2.as<uint64_t>(nh + 16);
  hipLaunchKernelGGL(k_key_heads, dim3(div_up(nh, 256)), dim3(256), 0, st, keys, nh, head);
  GSJ_LAUNCHED();
  const uint64_t nrows = scan_u32_to_u64(c, head, hoff, nh);
  hipLaunchKernelGGL(k_dw_fold, dim3(div_up(nh, 256)), dim3(256), 0, st, keys, vals, head, hoff, nh, dist, orow, odist);
  GSJ_LAUNCHED();
  return nrows;
}

// deliver one sub-batch's rows through the callback
void dw_deliver(Context& c, const uint32_t* orow, const double* odist, uint64_t nrows, gsj_join_stats& s, gsj_dist_cb cb, void* user) {
  if (nrows == 0) return;
  WallTimer t4;
  uint32_t* hrow = c.h_pairs.as<uint32_t>(2 * nrows);
  double* hd = odist ? c.h_area.as<double>(nrows) : nullptr;
  GSJ_HIP(hipMemcpyAsync(hrow, orow, nrows * 8, hipMemcpyDeviceToHost, c.stream));
  if (hd) GSJ_HIP(hipMemcpyAsync(hd, odist, nrows * 8, hipMemcpyDeviceToHost, c.stream));
  GSJ_HIP(hipStreamSynchronize(c.stream));
  s.copy_ms += t4.ms();
  if (cb && cb(user, hrow, hd, (size_t)nrows)) throw Error(GSJ_E_CANCEL, "distance join cancelled by the row callback");
}

}  // namespace

// refine one sub-batch's candidates into A.hit / A.dist (list classes, then the overflow queue)
void dw_refine(Context& c, const DwArgs& A, uint64_t np, gsj_join_stats& s) {
  hipStream_t stream = c.stream;
  const LaunchShape& sh = c.shape(np);
  const PairClasses pc = partition_pairs(c, A.P, A.R, A.pa, A.pb, np, sh.mcap, true, sh.tcap, sh.dw_size_sort);
  uint32_t* ovf = c.s_heavy.as<uint32_t>(np + 64);
  const uint32_t max_blocks = (uint32_t)c.cu_count * sh.dw_bpc;
  // the classes are independent (each appends its overflowing pairs to the queue with atomics): with sh.fork the
  // mixed and large classes run on the side stream beside the others
  hipStream_t sl = stream;
  if (sh.fork) {
    sl = c.side;
    GSJ_HIP(hipEventRecord(c.ev[10], stream));
    GSJ_HIP(hipStreamWaitEvent(sl, c.ev[10], 0));
  }
  GSJ_HIP(hipEventRecord(c.ev[5], sl));
  hipLaunchKernelGGL((k_dw_lists<32, 64, false>), dim3(std::min<uint32_t>(div_up(np, 8), max_blocks)), dim3(256), 0, sl, A, pc.list, pc.counts, (int)CL_MIXED,
                     ovf, pc.ctl);
  GSJ_LAUNCHED();
  GSJ_HIP(hipEventRecord(c.ev[7], sl));
  hipLaunchKernelGGL((k_dw_lists<64, 128, false>), dim3(std::min<uint32_t>(div_up(np, 4), max_blocks)), dim3(256), 0, sl, A, pc.list, pc.counts, (int)CL_LARGE,
                     ovf, pc.ctl);
  GSJ_LAUNCHED();
  GSJ_HIP(hipEventRecord(c.ev[2], sl));
  GSJ_HIP(hipEventRecord(c.ev[0], stream));
  hipLaunchKernelGGL((k_dw_lists<8, 16, true>), dim3(std::min<uint32_t>(div_up(np, 32), max_blocks)), dim3(256), 0, stream, A, pc.list, pc.counts, (int)CL_TINY,
                     ovf, pc.ctl);
  GSJ_LAUNCHED();
  GSJ_HIP(hipEventRecord(c.ev[8], stream));
  hipLaunchKernelGGL((k_dw_lists<16, 32, true>), dim3(std::min<uint32_t>(div_up(np, 16), max_blocks)), dim3(256), 0, stream, A, pc.list, pc.counts, (int)CL_SMALL,
                     ovf, pc.ctl);
  GSJ_LAUNCHED();
  GSJ_HIP(hipEventRecord(c.ev[1], stream));
  hipLaunchKernelGGL((k_dw_lists<32, 64, true>), dim3(std::min<uint32_t>(div_up(np, 8), max_blocks)), dim3(256), 0, stream, A, pc.list, pc.counts, (int)CL_MEDIUM,
                     ovf, pc.ctl);
  GSJ_LAUNCHED();
  GSJ_HIP(hipEventRecord(c.ev[9], stream));
  if (sh.fork) {
    GSJ_HIP(hipEventRecord(c.ev[11], sl));
    GSJ_HIP(hipStreamWaitEvent(stream, c.ev[11], 0));
  }
  struct Host { uint32_t counts[8]; OverflowCtl ctl; };
  Host* h = c.h_scal.as<Host>(2);
  GSJ_HIP(hipMemcpyAsync(h->counts, pc.counts, 32, hipMemcpyDeviceToHost, stream));
  GSJ_HIP(hipMemcpyAsync(&h->ctl, pc.ctl, sizeof(OverflowCtl), hipMemcpyDeviceToHost, stream));
  GSJ_HIP(hipStreamSynchronize(stream));
  float ms_t = 0, ms_s = 0, ms_m = 0, ms_x = 0, ms_l = 0;
  GSJ_HIP(hipEventElapsedTime(&ms_t, c.ev[0], c.ev[8]));
  GSJ_HIP(hipEventElapsedTime(&ms_s, c.ev[8], c.ev[1]));
  GSJ_HIP(hipEventElapsedTime(&ms_m, c.ev[1], c.ev[9]));
  GSJ_HIP(hipEventElapsedTime(&ms_x, c.ev[5], c.ev[7]));
  GSJ_HIP(hipEventElapsedTime(&ms_l, c.ev[7], c.ev[2]));
  s.k_tiny_ms += ms_t; s.k_small_ms += ms_s; s.k_medium_ms += ms_m; s.k_mixed_ms += ms_x; s.k_large_ms += ms_l;
  s.n_tiny += h->counts[CL_TINY]; s.n_small += h->counts[CL_SMALL]; s.n_medium += h->counts[CL_MEDIUM]; s.n_mixed += h->counts[CL_MIXED];
  s.n_large += h->counts[CL_LARGE];
  s.launches++;
  const uint32_t nover = h->ctl.count;
  if (nover) {
    const int cap = (int)h->ctl.need;
    const size_t per = 2 * (size_t)cap * sizeof(Ed);
    uint64_t waves = std::max<uint64_t>(1, std::min<uint64_t>(nover, (1ull << 30) / per));
    waves = (waves + 3) / 4 * 4;
    Ed* slab = c.s_heavy_edges.as<Ed>(2 * (uint64_t)cap * waves + 16);
    GSJ_HIP(hipEventRecord(c.ev[3], stream));
    hipLaunchKernelGGL(k_dw_overflow, dim3((uint32_t)(waves / 4)), dim3(256), 0, stream, A, ovf, nover, cap, slab);
    GSJ_LAUNCHED();
    GSJ_HIP(hipEventRecord(c.ev[4], stream));
    GSJ_HIP(hipEventSynchronize(c.ev[4]));
    float ms_o = 0;
    GSJ_HIP(hipEventElapsedTime(&ms_o, c.ev[3], c.ev[4]));
    s.k_over_ms += ms_o;
    s.n_over += nover;
    s.heavy += nover;
  }
}

void dwithin_run(Context& c, const Layer& PL, const Layer& RL, double d, bool same, bool report, const gsj_join_opts& o, gsj_dist_cb cb, void* user,
                 gsj_join_stats* st) {
  GSJ_REQUIRE(!PL.points && !RL.points, GSJ_E_ARG, "dwithin: both layers must be polygon layers (points: gsj_points_dwithin)");
  GSJ_REQUIRE(RL.has_bvh, GSJ_E_ARG, "dwithin: build layer has no BVH (call gsj_bvh_build)");
  GSJ_REQUIRE(d >= 0 && d < 1e7, GSJ_E_ARG, "dwithin: distance must be in [0, 1e7) metres");
  gsj_join_stats s{};
  if (PL.n == 0 || RL.n == 0) { if (st) *st = s; return; }
  hipStream_t stream = c.stream;
  // search boxes widened by d, for the probes that will be walked only (a window's members)
  const MembersFn widen = [&](const uint32_t* list, uint32_t n, const uint64_t*) -> const float4* {
    if (d == 0 || n == 0) return nullptr;
    float4* e = c.s_qbb.as<float4>(PL.n + 16);
    if (list) hipLaunchKernelGGL(k_expand_polys_list, dim3(div_up(n, 256)), dim3(256), 0, stream, PL.poly, d, list, n, e);
    else hipLaunchKernelGGL(k_expand_polys, dim3(div_up(PL.n, 256)), dim3(256), 0, stream, PL.poly, d, e);
    GSJ_LAUNCHED();
    return e;
  };
  for_each_candidate_batch(c, probes_of(PL), RL.bvh, o.upper != 0, o, s, [&](uint32_t* pa, uint32_t* pb, uint64_t np) {
    WallTimer t2;
    DwArgs A{PL.poly, RL.poly, pa, pb, d, same ? 1 : 0, report ? 1 : 0, c.s_flag.as<uint8_t>(np + 16), report ? c.s_area.as<double>(np + 16) : nullptr,
             nullptr, DistSet{}};
    dw_refine(c, A, np, s);
    GSJ_HIP(hipStreamSynchronize(stream));
    s.refine_ms += t2.ms();
    WallTimer t3;
    uint32_t* orow = c.s_out_a.as<uint32_t>(2 * np + 16);
    double* odist = report ? c.s_out_area.as<double>(np + 16) : nullptr;
    const uint64_t nrows = dw_compact(c, PL.poly.parent, RL.poly.parent, pa, pb, np, A.hit, A.dist, orow, odist);
    GSJ_HIP(hipStreamSynchronize(stream));
    s.compact_ms += t3.ms();
    s.hits += nrows;
    dw_deliver(c, orow, odist, nrows, s, cb, user);
  }, widen);
  if (st) *st = s;
}

bool dwithin_batchable(const Layer& P, const Layer& R, const gsj_join_opts& o) {
  return o.has_window && !o.cancel && !P.points && !R.points && R.has_bvh && P.n && R.n && !P.poly.parent && !R.poly.parent &&
         probes_batchable(probes_of(P));
}

void dwithin_batch_run(Context& c, const Layer& PL, const Layer& RL, bool same, bool report, std::vector<BatchDist>& reqs) {
  const uint32_t k = (uint32_t)reqs.size();
  GSJ_REQUIRE(k >= 1 && k <= MAX_WINDOWS, GSJ_E_ARG, "batch size out of range");
  const gsj_join_opts& o = *reqs[0].o;
  WindowSet ws;
  DistSet ds{};
  ws.k = ds.k = k;
  for (uint32_t r = 0; r < k; r++) {
    const gsj_join_opts& q = *reqs[r].o;
    GSJ_REQUIRE(dwithin_batchable(PL, RL, q) && q.upper == o.upper && q.batch_q == o.batch_q && q.max_pairs == o.max_pairs, GSJ_E_ARG,
                "request cannot join this batch");
    GSJ_REQUIRE(reqs[r].d >= 0 && reqs[r].d < 1e7, GSJ_E_ARG, "dwithin: distance must be in [0, 1e7) metres");
    for (int j = 0; j < 4; j++) ws.w[r][j] = q.window[j];
    ds.d[r] = reqs[r].d;
  }
  hipStream_t stream = c.stream;
  const MembersFn widen = [&](const uint32_t* list, uint32_t n, const uint64_t* mstart) -> const float4* {
    float4* e = c.s_qbb.as<float4>((uint64_t)n + 16);
    hipLaunchKernelGGL(k_expand_members, dim3(div_up(n, 256)), dim3(256), 0, stream, PL.poly, list, n, mstart, ds, e);
    GSJ_LAUNCHED();
    return e;
  };
  gsj_join_stats s{};
  std::vector<uint64_t> cand(k, 0), hits(k, 0), rlo(k + 1);
  for_each_candidate_multi(c, probes_of(PL), RL.bvh, o.upper != 0, o, ws, s, [&](uint32_t* pa, uint32_t* pb, uint64_t np, const SegBounds& sb) {
    for (uint32_t r = 0; r < k; r++) cand[r] += sb.h[r + 1] - sb.h[r];
    WallTimer t2;
    DwArgs A{PL.poly, RL.poly, pa, pb, 0.0, same ? 1 : 0, report ? 1 : 0, c.s_flag.as<uint8_t>(np + 16), report ? c.s_area.as<double>(np + 16) : nullptr,
             sb.d, ds};
    dw_refine(c, A, np, s);
    GSJ_HIP(hipStreamSynchronize(stream));
    s.refine_ms += t2.ms();
    WallTimer t3;
    uint32_t* orow = c.s_out_a.as<uint32_t>(2 * np + 16);
    double* odist = report ? c.s_out_area.as<double>(np + 16) : nullptr;
    const uint64_t* offs = nullptr;
    const uint64_t nrows = dw_compact(c, nullptr, nullptr, pa, pb, np, A.hit, A.dist, orow, odist, &offs);
    seg_rows(c, sb, offs, np, nrows, rlo.data());
    s.compact_ms += t3.ms();
    if (nrows == 0) return;
    WallTimer t4;
    uint32_t* hrow = c.h_pairs.as<uint32_t>(2 * nrows);
    double* hd = odist ? c.h_area.as<double>(nrows) : nullptr;
    GSJ_HIP(hipMemcpyAsync(hrow, orow, nrows * 8, hipMemcpyDeviceToHost, stream));
    if (hd) GSJ_HIP(hipMemcpyAsync(hd, odist, nrows * 8, hipMemcpyDeviceToHost, stream));
    GSJ_HIP(hipStreamSynchronize(stream));
    s.copy_ms += t4.ms();
    for (uint32_t r = 0; r < k; r++) {
      const uint64_t a = rlo[r], n = rlo[r + 1] - rlo[r];
      hits[r] += n;
      BatchDist& q = reqs[r];
      if (n == 0 || !q.cb || q.rc != GSJ_OK) continue;
      if (q.cb(q.user, hrow + 2 * a, hd ? hd + a : nullptr, (size_t)n)) q.rc = GSJ_E_CANCEL;
    }
  }, widen);
  for (uint32_t r = 0; r < k; r++) {
    if (!reqs[r].st) continue;
    gsj_join_stats& st = *reqs[r].st;
    st = s;
    st.candidates = cand[r];
    st.hits = hits[r];
  }
}

void points_dwithin_run(Context& c, const Layer& PL, const Layer& RL, double d, bool report, const gsj_join_opts& o, gsj_dist_cb cb, void* user,
                        gsj_join_stats* st) {
  GSJ_REQUIRE(PL.points && !RL.points, GSJ_E_ARG, "points_dwithin: needs a point layer and a polygon layer");
  GSJ_REQUIRE(RL.has_bvh, GSJ_E_ARG, "points_dwithin: polygon layer has no BVH");
  GSJ_REQUIRE(d >= 0 && d < 1e7, GSJ_E_ARG, "points_dwithin: distance must be in [0, 1e7) metres");
  gsj_join_stats s{};
  if (PL.n == 0 || RL.n == 0) { if (st) *st = s; return; }
  hipStream_t stream = c.stream;
  const float4* qbb = PL.pts.bb;
  if (d > 0) {
    float4* e = c.s_qbb.as<float4>(PL.n + 16);
    hipLaunchKernelGGL(k_expand_points, dim3(div_up(PL.n, 256)), dim3(256), 0, stream, PL.pts, d, e);
    GSJ_LAUNCHED();
    qbb = e;
  }
  for_each_candidate_batch(c, probes_of(PL, qbb), RL.bvh, false, o, s, [&](uint32_t* pa, uint32_t* pb, uint64_t np) {
    WallTimer t2;
    uint8_t* hit = c.s_flag.as<uint8_t>(np + 16);
    double* dist = report ? c.s_area.as<double>(np + 16) : nullptr;
    GSJ_HIP(hipEventRecord(c.ev[0], stream));
    hipLaunchKernelGGL(k_pt_dwithin, dim3(div_up(np, 256)), dim3(256), 0, stream, PL.pts, RL.poly, pa, pb, np, d, report ? 1 : 0, hit, dist);
    GSJ_LAUNCHED();
    GSJ_HIP(hipEventRecord(c.ev[1], stream));
    GSJ_HIP(hipStreamSynchronize(stream));
    float ms = 0;
    GSJ_HIP(hipEventElapsedTime(&ms, c.ev[0], c.ev[1]));
    s.k_small_ms += ms;
    s.n_small += np;
    s.launches++;
    s.refine_ms += t2.ms();
    WallTimer t3;
    uint32_t* orow = c.s_out_a.as<uint32_t>(2 * np + 16);
    double* odist = report ? c.s_out_area.as<double>(np + 16) : nullptr;
    const uint64_t nrows = dw_compact(c, nullptr, RL.poly.parent, pa, pb, np, hit, dist, orow, odist);
    GSJ_HIP(hipStreamSynchronize(stream));
    s.compact_ms += t3.ms();
    s.hits += nrows;
    dw_deliver(c, orow, odist, nrows, s, cb, user);
  });
  if (st) *st = s;
}

void points_erode_run(Context& c, const Layer& PL, const Layer& RL, double r, const gsj_join_opts& o, gsj_dist_cb cb, void* user, gsj_join_stats* st) {
  GSJ_REQUIRE(PL.points && !RL.points, GSJ_E_ARG, "points_intersects_buffered: needs a point layer and a polygon layer");
  GSJ_REQUIRE(RL.has_bvh, GSJ_E_ARG, "points_intersects_buffered: polygon layer has no BVH");
  GSJ_REQUIRE(r >= 0 && r < 1e5, GSJ_E_ARG, "points_intersects_buffered: radius must be in [0, 1e5) metres");
  gsj_join_stats s{};
  if (PL.n == 0 || RL.n == 0) { if (st) *st = s; return; }
  hipStream_t stream = c.stream;
  // every piece whose boundary may be within r of the point (buffer.cuh:1145-1152)
  const double grow = r * 1.002;
  const float4* qbb = PL.pts.bb;
  if (grow > 0) {
    float4* e = c.s_qbb.as<float4>(PL.n + 16);
    hipLaunchKernelGGL(k_expand_points, dim3(div_up(PL.n, 256)), dim3(256), 0, stream, PL.pts, grow, e);
    GSJ_LAUNCHED();
    qbb = e;
  }
  for_each_candidate_batch(c, probes_of(PL, qbb), RL.bvh, false, o, s, [&](uint32_t* pa, uint32_t* pb, uint64_t np) {
    WallTimer t2;
    uint8_t* bits = c.s_flag.as<uint8_t>(np + 16);
    GSJ_HIP(hipEventRecord(c.ev[0], stream));
    hipLaunchKernelGGL(k_pt_erode, dim3(div_up(np, 256)), dim3(256), 0, stream, PL.pts, RL.poly, pa, pb, np, r, bits);
    GSJ_LAUNCHED();
    GSJ_HIP(hipEventRecord(c.ev[1], stream));
    GSJ_HIP(hipStreamSynchronize(stream));
    float ms = 0;
    GSJ_HIP(hipEventElapsedTime(&ms, c.ev[0], c.ev[1]));
    s.k_small_ms += ms;
    s.n_small += np;
    s.launches++;
    s.refine_ms += t2.ms();
    WallTimer t3;
    uint32_t* f = c.s_tmp.as<uint32_t>(np + 16);
    uint64_t* off = c.s_offs.as<uint64_t>(np + 16);
    uint32_t* orow = c.s_out_a.as<uint32_t>(2 * np + 16);
    uint64_t nrows = 0;
    if (!RL.poly.parent) {
      hipLaunchKernelGGL(k_one_u32, dim3(div_up(np, 256)), dim3(256), 0, stream, bits, np, f);
      GSJ_LAUNCHED();
      nrows = scan_u32_to_u64(c, f, off, np);
      hipLaunchKernelGGL(k_pe_direct, dim3(div_up(np, 256)), dim3(256), 0, stream, pa, pb, f, off, np, orow);
      GSJ_LAUNCHED();
    } else {
      hipLaunchKernelGGL(k_nonzero_u32, dim3(div_up(np, 256)), dim3(256), 0, stream, bits, np, f);
      GSJ_LAUNCHED();
      const uint64_t nz = scan_u32_to_u64(c, f, off, np);
      if (nz) {
        uint64_t* keys = c.s_keys.as<uint64_t>(nz + 16);
        uint64_t* keys2 = c.s_keys2.as<uint64_t>(nz + 16);
        uint32_t* vals = c.s_vals.as<uint32_t>(nz + 16);
        uint32_t* vals2 = c.s_vals2.as<uint32_t>(nz + 16);
        hipLaunchKernelGGL(k_pe_keys, dim3(div_up(np, 256)), dim3(256), 0, stream, RL.poly.parent, pa, pb, bits, f, off, np, keys, vals);
        GSJ_LAUNCHED();
        radix_sort_pairs(c, keys, keys2, vals, vals2, nz, 0, 64);
        uint32_t* keep = c.s_seg.as<uint32_t>(nz + 16);
        uint64_t* koff = c.s_tmp2.as<uint64_t>(nz + 16);
        hipLaunchKernelGGL(k_pe_fold, dim3(div_up(nz, 256)), dim3(256), 0, stream, keys, vals, nz, keep);
        GSJ_LAUNCHED();
        nrows = scan_u32_to_u64(c, keep, koff, nz);
        hipLaunchKernelGGL(k_pe_emit, dim3(div_up(nz, 256)), dim3(256), 0, stream, keys, keep, koff, nz, orow);
        GSJ_LAUNCHED();
      }
    }
    GSJ_HIP(hipStreamSynchronize(stream));
    s.compact_ms += t3.ms();
    s.hits += nrows;
    dw_deliver(c, orow, nullptr, nrows, s, cb, user);
  });
  if (st) *st = s;
}

}  // namespace gsj
