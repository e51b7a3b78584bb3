// Device building blocks shared by the pair kernels (polygon refinement, distance joins): edge records, lane-group
// ballots, ballot-compacted edge lists of a geometry inside a box, crossing-parity containment (reference:
// engine/engine.cuh:265-412), the overflow queue, and the work-class partition of candidate pairs.
#pragma once
#include "engine.h"

namespace gsj {

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

// ------------------------------------------------------------------------------------------- edge sources
// Sources enumerate "the other polygon's edges meeting the overlap box" in ring order.
struct ListSrc {   // compacted list (large / overflow classes)
  const Ed* L;
  int n;
  __device__ __forceinline__ int size() const { return n; }
  __device__ __forceinline__ bool get(int k, Ed& e) const { e = L[k]; return true; }
};

// inside(x) = inside(c) xor parity of segment c->x across the local edges (engine/engine.cuh:400-412)
template <class Src>
__device__ bool inside_from_c(const Src& L, double cx, double cy, bool in_c, double x, double y) {
  bool in = in_c;
  const double sx0 = fmin(cx, x), sx1 = fmax(cx, x), sy0 = fmin(cy, y), sy1 = fmax(cy, y);
  const int n = L.size();
  for (int k = 0; k < n; k++) {
    Ed e;
    if (!L.get(k, e)) continue;
    if (fmax(e.x0, e.x1) < sx0 || fmin(e.x0, e.x1) > sx1 || fmax(e.y0, e.y1) < sy0 || fmin(e.y0, e.y1) > sy1) continue;
    const bool sc = left_of(e.x0, e.y0, e.x1, e.y1, cx, cy), sx = left_of(e.x0, e.y0, e.x1, e.y1, x, y);
    if (sc == sx) continue;
    const bool su = orient(cx, cy, x, y, e.x0, e.y0) > 0, sv = orient(cx, cy, x, y, e.x1, e.y1) > 0;
    if (su != sv) in = !in;
  }
  return in;
}

// ------------------------------------------------------------------------------------------- edge chunks
// Edges (i, i+1) of a geometry are walked in runs: a ring's edges split at chunk boundaries (chunk k of a chunked
// geometry holds i in [v0 + CHUNK k, v0 + CHUNK (k+1))), so a run lies in one chunk and one ring. A run whose chunk box,
// shifted by (shx, shy), misses the box of interest is skipped whole. Rounding is monotonic, so a chunk box that misses
// the box in floating point holds only vertices whose shifted coordinates miss it too: every edge it covers is one the
// caller would reject, and skipping it changes no result.
__device__ __forceinline__ bool chunk_misses(const float4 c, double shx, double shy, const Box& o) {
  return shx + (double)c.z < o.x0 || shx + (double)c.x > o.x1 || shy + (double)c.w < o.y0 || shy + (double)c.y > o.y1;
}
// the chunk's vertices all strictly above or strictly below the horizontal line y = my (shifted by shy)
__device__ __forceinline__ bool chunk_off_row(const float4 c, double shy, double my) { return shy + (double)c.y > my || shy + (double)c.w < my; }

// calls run(i0, i1) for the runs of edges i in [i0, i1) of geometry g, in ring and edge order; skip(chunk box) decides
// whether a chunked run can be skipped. A G-lane group walks together (group-uniform arguments): the skip decisions
// of 64 chunks at a time are taken across its lanes (lane l: chunks w + l, w + l + G, ...; their box loads in flight
// together) and read back from the ballot as the walk reaches them, so a polygon of thousands of edges costs a few
// rounds of box loads instead of one load latency per chunk. The runs and their order are those of a chunk-by-chunk
// walk.
template <int G, class Skip, class Run>
__device__ __forceinline__ void walk_edge_runs(const DPolys& S, uint32_t g, int gl, const Skip& skip, const Run& run) {
  const uint32_t r0 = S.gro[g], r1 = S.gro[g + 1];
  const uint64_t v0 = S.rvo[r0];
  const uint32_t cf = S.cfirst ? S.cfirst[g] : NO_CHUNKS;
  // chunks of the geometry (as the layer builds them: chunk k exists while v0 + CHUNK k + 1 < the geometry's end)
  const uint64_t nch = cf != NO_CHUNKS ? (S.rvo[r1] - v0 - 1 + CHUNK - 1) / CHUNK : 0;
  uint64_t wb = 0, keep = 0;   // decisions of chunks [wb, wb + 64): bit j set = chunk wb + j is walked
  bool have = false;
  for (uint32_t r = r0; r < r1; r++) {
    const uint64_t s = S.rvo[r], e = S.rvo[r + 1];
    for (uint64_t i0 = s; i0 + 1 < e;) {
      uint64_t i1 = e - 1;
      if (cf != NO_CHUNKS) {
        const uint64_t k = (i0 - v0) / CHUNK;
        i1 = min(i1, v0 + (k + 1) * CHUNK);
        if (!have || k >= wb + 64) {   // the walk's chunk index only grows
          wb = k;
          have = true;
          keep = 0;
#pragma unroll
          for (int j = 0; j < 64 / G; j++) {
            const uint64_t c = wb + (uint64_t)(j * G + gl);
            const bool walk = c < nch && !skip(S.cbb[cf + c]);
            keep |= gballot<G>(walk) << (j * G);
          }
        }
        if (!((keep >> (k - wb)) & 1)) {
          // straight to the next walked chunk of the window (or the window's end), or the ring's end: the chunks in
          // between are skipped ones
          const uint32_t o = (uint32_t)(k - wb) + 1;
          const uint64_t rest = o < 64 ? keep >> o : 0;
          const uint64_t kn = rest ? k + 1 + (uint64_t)__builtin_ctzll(rest) : wb + 64;
          i0 = min(e - 1, v0 + kn * CHUNK);
          continue;
        }
      }
      run(s, e, i0, i1);
      i0 = i1;
    }
  }
}

// ------------------------------------------------------------------------------------------- whole geometries
// edges of geometry g meeting box ov, in ring order, compacted by group ballot (engine/engine.cuh:358-377);
// returns the full count (may exceed cap: then the list is incomplete and the pair overflows)
template <int G>
__device__ int collect(const DPolys& S, uint32_t g, double shx, double shy, const Box& ov, Ed* out, int cap, int gl) {
  int cnt = 0;
  walk_edge_runs<G>(S, g, gl, [&](const float4 c) { return chunk_misses(c, shx, shy, ov); }, [&](uint64_t, uint64_t, uint64_t i0, uint64_t i1) {
    for (uint64_t base = i0; base < i1; base += G) {
      const uint64_t i = base + gl;
      bool v = false;
      Ed ed;
      if (i < i1) {
        const double2 p = S.xy[i], q = S.xy[i + 1];
        ed.x0 = shx + p.x; ed.y0 = shy + p.y; ed.x1 = shx + q.x; ed.y1 = shy + q.y;
        const uint8_t f = S.vflag ? S.vflag[i] : 0;
        ed.syn = (f & VF_SYN) ? 1 : 0;
        ed.cut = (f >> 2) & 3;   // bit 0 VF_CUT0, bit 1 VF_CUT1
        v = !edge_outside(ed.x0, ed.y0, ed.x1, ed.y1, ov);
      }
      const uint64_t m = gballot<G>(v);
      const int pos = cnt + __popcll(m & ((1ull << gl) - 1));
      if (v && pos < cap) out[pos] = ed;
      cnt += __popcll(m);
    }
  });
  return cnt;
}

// inside(m) over every edge of geometry g, read from global memory (engine/engine.cuh:382-397)
template <int G>
__device__ bool pip_full(const DPolys& S, uint32_t g, double shx, double shy, double mx, double my, int gl) {
  const double ox = shx - 1.0, oy = my;
  bool par = false;
  walk_edge_runs<G>(S, g, gl, [&](const float4 c) { return chunk_off_row(c, shy, my); }, [&](uint64_t, uint64_t, uint64_t i0, uint64_t i1) {
    for (uint64_t i = i0 + gl; i < i1; i += G) {
      const double2 p = S.xy[i], q = S.xy[i + 1];
      const double x0 = shx + p.x, y0 = shy + p.y, x1 = shx + q.x, y1 = shy + q.y;
      if ((y0 > my) == (y1 > my) && y0 != my && y1 != my) continue;
      const bool so = left_of(x0, y0, x1, y1, ox, oy), sc = left_of(x0, y0, x1, y1, mx, my);
      if (so == sc) continue;
      const bool su = orient(ox, oy, mx, my, x0, y0) > 0, sv = orient(ox, oy, mx, my, x1, y1) > 0;
      if (su != sv) par = !par;
    }
  });
  return (__popcll(gballot<G>(par)) & 1) != 0;
}

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

// A group's walk over a pair list (positions first, first + stride, ...). List position -> pair index -> the pair's two
// geometry ids are dependent loads; a pair's index is loaded two iterations ahead and its geometry ids one iteration
// ahead, so both arrive while the group works on earlier pairs. The prefetch loads are issued before the current
// pair's geometry loads, so waiting for those never stalls on a prefetch that was not already due.
struct PairStream {
  const uint32_t* cl;
  const uint32_t* pa;
  const uint32_t* pb;
  uint32_t n, stride, k, p, a, b, pn, an, bn, pnn;
  __device__ __forceinline__ PairStream(const uint32_t* cl_, const uint32_t* pa_, const uint32_t* pb_, uint32_t n_, uint32_t first, uint32_t stride_)
      : cl(cl_), pa(pa_), pb(pb_), n(n_), stride(stride_), k(first), p(0), a(0), b(0), pn(0), an(0), bn(0), pnn(0) {
    if (k < n) { p = cl[k]; a = pa[p]; b = pb[p]; }
    if ((uint64_t)k + stride < n) pn = cl[k + stride];
  }
  // the loads for the next two iterations: call once at the top of each iteration (k < n)
  __device__ __forceinline__ void prefetch() {
    if ((uint64_t)k + stride < n) { an = pa[pn]; bn = pb[pn]; }
    if ((uint64_t)k + 2ull * stride < n) pnn = cl[k + 2 * stride];
  }
  __device__ __forceinline__ void advance() {
    k += stride;
    p = pn; a = an; b = bn; pn = pnn;
  }
};

struct OverflowCtl { uint32_t count; uint32_t need; uint32_t pad0, pad1; };

// hand pair p to the overflow kernel (lists of up to `need` edges per side); called by one lane
__device__ __forceinline__ void queue_overflow(uint32_t p, int need, uint32_t* ovf_list, OverflowCtl* ctl) {
  const uint32_t slot = atomicAdd(&ctl->count, 1u);
  ovf_list[slot] = p;
  atomicMax(&ctl->need, (uint32_t)need);
}


// ------------------------------------------------------------------------------------------- work classes
constexpr int TCAP = 16;             // tiny class: stored vertices per polygon (rings within RCAP)
constexpr int VCAP = 32, RCAP = 8;   // small class: stored vertices and rings per polygon
constexpr int MCAP = 64;             // medium class: stored vertices per polygon (rings within RCAP)
// the classes, in list order
enum PairClass : int { CL_TINY = 0, CL_SMALL = 1, CL_MEDIUM = 2, CL_MIXED = 3, CL_LARGE = 4, CL_COUNT = 5 };

// Candidate pairs split by work class, on the device: list[0 .. counts[0]) the tiny pairs (both polygons within tcap
// / RCAP; none when tcap is 0), then counts[1] small pairs (both within VCAP / RCAP), counts[2] medium pairs (both
// within mcap / RCAP; none when mcap is 0), counts[3] mixed pairs (with `mixed`: one polygon within VCAP vertices),
// then the counts[4] others; within a class by mean vertex count, then candidate order; ctl cleared for the overflow
// queue. counts[8 ..) is left zeroed for the callers' queues.
// Uses the context's s_keys, s_keys2, s_vals, s_vals2, s_tmp2 and s_misc scratch (the list lives in one of the
// s_vals buffers until their next use).
struct PairClasses {
  uint32_t* list;
  uint32_t* counts;
  OverflowCtl* ctl;
};
// size_sort: within a class, pairs of similar size side by side, in buckets 2^(size_sort-1) times the class's base
// width (0: candidate order, a probe's pairs together)
PairClasses partition_pairs(Context& c, const DPolys& P, const DPolys& R, const uint32_t* pa, const uint32_t* pb, uint64_t np, uint32_t mcap,
                            bool mixed, uint32_t tcap, uint32_t size_sort = 1);
// first list position of class cl
__host__ __device__ __forceinline__ uint32_t class_first(const uint32_t* counts, int cl) {
  uint32_t f = 0;
  for (int w = 0; w < cl; w++) f += counts[w];
  return f;
}

// group-wide minimum (order-independent, so deterministic)
template <int G>
__device__ __forceinline__ double group_min(double v) {
#pragma unroll
  for (int o = G / 2; o > 0; o >>= 1) v = fmin(v, __shfl_xor(v, o, G));
  return v;
}

}  // namespace gsj
