// Points in polygons (reference: k_pip_pairs / join_points_stream, engine/engine.cuh:1079-1178) and
// ST_ClusterIntersecting by lock-free union-find over the intersects rows (engine/engine.cuh:1183-1234).
#include "engine.h"
#include <algorithm>
#include <chrono>

namespace gsj {

namespace {

__device__ __forceinline__ int syn_bit(const uint32_t* synb, uint64_t i) { return synb ? (int)((synb[i >> 5] >> (i & 31)) & 1u) : 0; }

// thread per (point, piece): 2 = on a real edge, 1 = interior (crossing parity of a horizontal segment from a point
// left of the piece's bbox), 0 = outside. Coordinates relative to the piece's origin (exact, as everywhere).
__global__ void __launch_bounds__(256) k_pip(const DPoints pts, const DPolys R, const uint32_t* __restrict__ pa, const uint32_t* __restrict__ pb,
                                             uint64_t np, uint8_t* __restrict__ hit) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= np) return;
  const uint32_t g = pb[i];
  const double2 pt = pts.xy[pa[i]], o = R.origin[g];
  const double px = pt.x - o.x, py = pt.y - o.y;
  const double ox = (o.x - o.x) - 1.0;
  bool par = false, on = false;
  const uint32_t r1 = R.gro[g + 1];
  const uint64_t vg = R.rvo[R.gro[g]];
  const uint32_t cf = R.cfirst ? R.cfirst[g] : NO_CHUNKS;
  for (uint32_t r = R.gro[g]; r < r1 && !on; r++) {
    const uint64_t s0 = R.rvo[r], e = R.rvo[r + 1];
    // edges (j-1, j) taken an edge chunk at a time; a chunk whose vertices all lie strictly above or strictly below
    // the row y = py holds only edges the parity test skips
    for (uint64_t a = s0; a + 1 < e && !on;) {
      uint64_t b = e - 1;
      if (cf != NO_CHUNKS) {
        const uint64_t k = (a - vg) / CHUNK;
        b = min(b, vg + (k + 1) * CHUNK);
        const float4 c = R.cbb[cf + k];
        if ((double)c.y > py || (double)c.w < py) { a = b; continue; }
      }
      double2 v0 = R.xy[a];
      for (uint64_t j = a + 1; j <= b; j++) {
        const double2 v1 = R.xy[j];
        const double x0 = v0.x, y0 = v0.y, x1 = v1.x, y1 = v1.y;
        if (!((y0 > py) == (y1 > py) && y0 != py && y1 != py)) {
          if (orient(x0, y0, x1, y1, px, py) == 0 && !syn_bit(R.synb, j - 1) && onseg(x0, y0, x1, y1, px, py)) { on = true; break; }
          const bool so = left_of(x0, y0, x1, y1, ox, py), sp = left_of(x0, y0, x1, y1, px, py);
          if (so != sp) {
            const bool su = orient(ox, py, px, py, x0, y0) > 0, sv = orient(ox, py, px, py, x1, y1) > 0;
            if (su != sv) par = !par;
          }
        }
        v0 = v1;
      }
      a = b;
    }
  }
  hit[i] = on ? 2 : (par ? 1 : 0);
}

__global__ void k_hit_flags(const uint8_t* __restrict__ hit, uint64_t n, uint32_t* __restrict__ f) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) f[i] = hit[i] != 0;
}

// direct rows (no subdivided parents) or (point, parent) keys for the fold
__global__ void k_pip_emit(const DPolys R, const uint32_t* __restrict__ pa, const uint32_t* __restrict__ pb, const uint8_t* __restrict__ hit,
                           const uint32_t* __restrict__ f, const uint64_t* __restrict__ offs, uint64_t n, uint32_t* __restrict__ orow,
                           uint8_t* __restrict__ owhere, uint64_t* __restrict__ keys, uint32_t* __restrict__ vals) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n || !f[i]) return;
  const uint64_t o = offs[i];
  const uint32_t b = R.parent ? R.parent[pb[i]] : pb[i];
  if (keys) { keys[o] = ((uint64_t)pa[i] << 32) | b; vals[o] = hit[i]; return; }
  orow[2 * o] = pa[i]; orow[2 * o + 1] = b; owhere[o] = hit[i];
}

__global__ void k_heads64(const uint64_t* __restrict__ keys, uint64_t m, uint32_t* __restrict__ head) {
  const uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j < m) head[j] = (j == 0 || keys[j] != keys[j - 1]) ? 1u : 0u;
}

// boundary wins over interior: max of the piece results per (point, parent)
__global__ void k_pip_fold(const uint64_t* __restrict__ keys, const uint32_t* __restrict__ vals, const uint32_t* __restrict__ head,
                           const uint64_t* __restrict__ hoff, uint64_t m, uint32_t* __restrict__ orow, uint8_t* __restrict__ owhere) {
  const uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= m || !head[j]) return;
  uint32_t w = 0;
  for (uint64_t k = j; k < m && keys[k] == keys[j]; k++) w = max(w, vals[k]);
  const uint64_t o = hoff[j];
  orow[2 * o] = (uint32_t)(keys[j] >> 32); orow[2 * o + 1] = (uint32_t)keys[j]; owhere[o] = (uint8_t)w;
}

// ---- union-find (root = smallest index of the component) ---------------------------------------------------
__device__ __forceinline__ uint32_t uf_load(uint32_t* p) { return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT); }

// find with path halving; every store replaces a parent by one of its ancestors, so concurrent finds stay valid
__device__ uint32_t uf_find(uint32_t* parent, uint32_t x) {
  for (;;) {
    const uint32_t p = uf_load(parent + x);
    if (p == x) return x;
    const uint32_t gp = uf_load(parent + p);
    if (gp != p) __hip_atomic_store(parent + x, gp, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    x = p;
  }
}

__global__ void k_uf_init(uint32_t* parent, uint32_t n) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) parent[i] = i;
}

// hook the larger root under the smaller one; retry when another thread re-hooked it first
__global__ void k_uf_union(const uint32_t* __restrict__ rows, uint64_t n, uint32_t* parent) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint32_t a = rows[2 * i], b = rows[2 * i + 1];
  if (a == b) return;
  for (;;) {
    uint32_t ra = uf_find(parent, a), rb = uf_find(parent, b);
    if (ra == rb) return;
    if (ra > rb) { const uint32_t t = ra; ra = rb; rb = t; }
    uint32_t expect = rb;
    if (__hip_atomic_compare_exchange_strong(parent + rb, &expect, ra, __ATOMIC_RELAXED, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) return;
  }
}

// windowed clusters: only the members' entries of parent[] are initialised and ever read
__global__ void k_uf_init_list(uint32_t* parent, const uint32_t* __restrict__ list, uint32_t n) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) parent[list[i]] = list[i];
}
// rows whose two parents both meet the window (the probe side always does; the same closed test in double as the
// membership), hooked as k_uf_union does
__global__ void k_uf_union_window(const uint32_t* __restrict__ rows, uint64_t n, const float4* __restrict__ bb, double x0, double y0, double x1,
                                  double y1, uint32_t* parent) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint32_t a = rows[2 * i], b = rows[2 * i + 1];
  if (a == b) return;
  const float4 q = bb[b];
  if (!((double)q.x <= x1 && (double)q.z >= x0 && (double)q.y <= y1 && (double)q.w >= y0)) return;
  for (;;) {
    uint32_t ra = uf_find(parent, a), rb = uf_find(parent, b);
    if (ra == rb) return;
    if (ra > rb) { const uint32_t t = ra; ra = rb; rb = t; }
    uint32_t expect = rb;
    if (__hip_atomic_compare_exchange_strong(parent + rb, &expect, ra, __ATOMIC_RELAXED, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) return;
  }
}
__global__ void k_uf_roots_list(uint32_t* parent, const uint32_t* __restrict__ list, uint32_t n, uint32_t* __restrict__ rows) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint32_t g = list[i];
  rows[2 * i] = g;
  rows[2 * i + 1] = uf_find(parent, g);
}

// roots into a separate array (writing them into parent[] would race with other threads' path-halving stores)
__global__ void k_uf_roots(uint32_t* parent, uint32_t n, uint32_t* root) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) root[i] = uf_find(parent, i);
}

}  // namespace

void points_join_run(Context& c, const Layer& PL, const Layer& RL, const gsj_join_opts& o, gsj_points_cb cb, void* user, gsj_join_stats* st) {
  GSJ_REQUIRE(PL.points && !RL.points, GSJ_E_ARG, "points_in_polygons: needs a point layer and a polygon layer");
  GSJ_REQUIRE(RL.has_bvh, GSJ_E_ARG, "points_in_polygons: polygon layer has no BVH");
  gsj_join_stats s{};
  if (PL.n == 0 || RL.n == 0) { if (st) *st = s; return; }
  hipStream_t stream = c.stream;
  for_each_candidate_batch(c, probes_of(PL), RL.bvh, false, o, s, [&](uint32_t* pa, uint32_t* pb, uint64_t np) {
      WallTimer t2;
      uint8_t* hit = c.s_flag.as<uint8_t>(np + 16);
      GSJ_HIP(hipEventRecord(c.ev[0], stream));
      hipLaunchKernelGGL(k_pip, dim3(div_up(np, 256)), dim3(256), 0, stream, PL.pts, RL.poly, pa, pb, np, hit);
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
      uint64_t* hoffs = c.s_offs.as<uint64_t>(np + 16);
      hipLaunchKernelGGL(k_hit_flags, dim3(div_up(np, 256)), dim3(256), 0, stream, hit, np, f);
      GSJ_LAUNCHED();
      const uint64_t nh = scan_u32_to_u64(c, f, hoffs, np);
      uint32_t* orow = c.s_out_a.as<uint32_t>(2 * np + 16);
      uint8_t* owhere = c.s_out_flag.as<uint8_t>(np + 16);
      uint64_t nrows = nh;
      if (!RL.poly.parent) {
        hipLaunchKernelGGL(k_pip_emit, dim3(div_up(np, 256)), dim3(256), 0, stream, RL.poly, pa, pb, hit, f, hoffs, np, orow, owhere, nullptr, nullptr);
        GSJ_LAUNCHED();
      } else if (nh) {
        uint64_t* keys = c.s_keys.as<uint64_t>(nh);
        uint64_t* keys2 = c.s_keys2.as<uint64_t>(nh);
        uint32_t* vals = c.s_vals.as<uint32_t>(nh);
        uint32_t* vals2 = c.s_vals2.as<uint32_t>(nh);
        hipLaunchKernelGGL(k_pip_emit, dim3(div_up(np, 256)), dim3(256), 0, stream, RL.poly, pa, pb, hit, f, hoffs, np, orow, owhere, keys, vals);
        GSJ_LAUNCHED();
        radix_sort_pairs(c, keys, keys2, vals, vals2, nh, 0, 64);
        uint32_t* head = c.s_seg.as<uint32_t>(nh + 16);
        uint64_t* hoff2 = c.s_tmp2.as<uint64_t>(nh + 16);
        hipLaunchKernelGGL(k_heads64, dim3(div_up(nh, 256)), dim3(256), 0, stream, keys, nh, head);
        GSJ_LAUNCHED();
        nrows = scan_u32_to_u64(c, head, hoff2, nh);
        hipLaunchKernelGGL(k_pip_fold, dim3(div_up(nh, 256)), dim3(256), 0, stream, keys, vals, head, hoff2, nh, orow, owhere);
        GSJ_LAUNCHED();
      }
      GSJ_HIP(hipStreamSynchronize(stream));
      s.compact_ms += t3.ms();
      s.hits += nrows;
      if (!nrows) return;
      WallTimer t4;
      uint32_t* hrow = c.h_pairs.as<uint32_t>(2 * nrows);
      uint8_t* hw = c.h_where.as<uint8_t>(nrows);
      GSJ_HIP(hipMemcpyAsync(hrow, orow, nrows * 8, hipMemcpyDeviceToHost, stream));
      GSJ_HIP(hipMemcpyAsync(hw, owhere, nrows, hipMemcpyDeviceToHost, stream));
      GSJ_HIP(hipStreamSynchronize(stream));
      s.copy_ms += t4.ms();
      if (cb && cb(user, hrow, hw, (size_t)nrows)) throw Error(GSJ_E_CANCEL, "points join cancelled by the row callback");
  });
  if (st) *st = s;
}

void cluster_run(Context& c, const Layer& PL, const gsj_join_opts& o, uint32_t* root, uint64_t nroot, gsj_join_stats* st) {
  GSJ_REQUIRE(!PL.points, GSJ_E_ARG, "cluster: needs a polygon layer");
  GSJ_REQUIRE(PL.has_bvh, GSJ_E_ARG, "cluster: layer has no BVH");
  const uint32_t n = PL.nparent;
  GSJ_REQUIRE(root && nroot >= n, GSJ_E_ARG, "cluster: root array shorter than the number of parent geometries");
  hipStream_t stream = c.stream;
  uint32_t* parent = c.s_uf.as<uint32_t>((uint64_t)n + 16);
  uint32_t* droot = c.s_root.as<uint32_t>((uint64_t)n + 16);
  if (n) hipLaunchKernelGGL(k_uf_init, dim3(div_up(n, 256)), dim3(256), 0, stream, parent, n);
  GSJ_LAUNCHED();
  GSJ_REQUIRE(!o.has_window, GSJ_E_ARG, "cluster takes no query window");
  gsj_join_opts jo = o;
  jo.want_area = 0; jo.want_relate = 0; jo.upper = 1;   // intersects is symmetric: each unordered pair once
  JoinHooks hooks;
  hooks.device_only = true;
  hooks.on_device = [&](const uint32_t* rows, const double*, uint64_t nr) {
    hipLaunchKernelGGL(k_uf_union, dim3(div_up(nr, 256)), dim3(256), 0, stream, rows, nr, parent);
    GSJ_LAUNCHED();
  };
  join_run(c, PL, PL, jo, nullptr, nullptr, st, &hooks);
  if (n) {
    hipLaunchKernelGGL(k_uf_roots, dim3(div_up(n, 256)), dim3(256), 0, stream, parent, n, droot);
    GSJ_LAUNCHED();
    GSJ_HIP(hipMemcpyAsync(root, droot, (size_t)n * 4, hipMemcpyDeviceToHost, stream));
  }
  GSJ_HIP(hipStreamSynchronize(stream));
}

void cluster_rows_run(Context& c, const Layer& PL, const gsj_join_opts& o, gsj_rows_cb cb, void* user, gsj_join_stats* st) {
  GSJ_REQUIRE(!PL.points, GSJ_E_ARG, "cluster: needs a polygon layer");
  GSJ_REQUIRE(PL.has_bvh, GSJ_E_ARG, "cluster: layer has no BVH");
  const uint32_t n = PL.nparent;
  if (!o.has_window) {   // every parent, in id order
    std::vector<uint32_t> root(std::max<uint32_t>(n, 1));
    cluster_run(c, PL, o, root.data(), n, st);
    const uint32_t chunk = 1u << 20;
    std::vector<uint32_t> rows(2 * (size_t)std::min(n, chunk));
    for (uint32_t g0 = 0; g0 < n; g0 += chunk) {
      const uint32_t m = std::min(chunk, n - g0);
      for (uint32_t i = 0; i < m; i++) { rows[2 * i] = g0 + i; rows[2 * i + 1] = root[g0 + i]; }
      if (cb && cb(user, rows.data(), nullptr, nullptr, m)) throw Error(GSJ_E_CANCEL, "cluster cancelled by the row callback");
    }
    return;
  }
  GSJ_REQUIRE(PL.h_parent.empty(), GSJ_E_ARG, "a windowed cluster needs a layer without subdivided parents");
  GSJ_REQUIRE(o.window[0] <= o.window[2] && o.window[1] <= o.window[3], GSJ_E_ARG, "window must have x0 <= x1 and y0 <= y1");
  hipStream_t stream = c.stream;
  uint32_t* parent = c.s_uf.as<uint32_t>((uint64_t)n + 16);
  const uint32_t* mlist = nullptr;
  uint32_t nm = 0;
  gsj_join_opts jo = o;
  jo.want_area = 0; jo.want_relate = 0; jo.upper = 1;   // intersects is symmetric: each unordered pair once
  JoinHooks hooks;
  hooks.device_only = true;
  hooks.on_members = [&](const uint32_t* list, uint32_t m, const uint64_t*) -> const float4* {
    mlist = list;
    nm = m;
    if (m) {
      hipLaunchKernelGGL(k_uf_init_list, dim3(div_up(m, 256)), dim3(256), 0, stream, parent, list, m);
      GSJ_LAUNCHED();
    }
    return nullptr;
  };
  hooks.on_device = [&](const uint32_t* rows, const double*, uint64_t nr) {
    hipLaunchKernelGGL(k_uf_union_window, dim3(div_up(nr, 256)), dim3(256), 0, stream, rows, nr, PL.poly.bb, o.window[0], o.window[1], o.window[2],
                       o.window[3], parent);
    GSJ_LAUNCHED();
  };
  gsj_join_stats s{};
  join_run(c, PL, PL, jo, nullptr, nullptr, &s, &hooks);
  if (nm) {
    uint32_t* drows = c.s_root.as<uint32_t>(2 * (uint64_t)nm + 16);
    hipLaunchKernelGGL(k_uf_roots_list, dim3(div_up(nm, 256)), dim3(256), 0, stream, parent, mlist, nm, drows);
    GSJ_LAUNCHED();
    std::vector<uint64_t> rows(nm);   // (parent, root) packed, sorted by parent
    uint32_t* h = c.h_pairs.as<uint32_t>(2 * (uint64_t)nm);
    GSJ_HIP(hipMemcpyAsync(h, drows, (size_t)nm * 8, hipMemcpyDeviceToHost, stream));
    GSJ_HIP(hipStreamSynchronize(stream));
    for (uint32_t i = 0; i < nm; i++) rows[i] = ((uint64_t)h[2 * i] << 32) | h[2 * i + 1];
    std::sort(rows.begin(), rows.end());
    for (uint32_t i = 0; i < nm; i++) { h[2 * i] = (uint32_t)(rows[i] >> 32); h[2 * i + 1] = (uint32_t)rows[i]; }
    if (cb && cb(user, h, nullptr, nullptr, nm)) throw Error(GSJ_E_CANCEL, "cluster cancelled by the row callback");
  }
  GSJ_HIP(hipStreamSynchronize(stream));
  if (st) *st = s;
}

}  // namespace gsj
