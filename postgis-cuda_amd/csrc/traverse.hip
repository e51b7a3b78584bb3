// Candidate search (reference: k_traverse, engine/engine.cuh:233-259): every leaf whose float box overlaps the
// probe's box (closed test), restricted to leaf >= probe for self-joins.
//
// Stackless: nodes are in depth-first preorder, so a lane walks forward through the node array, descending into
// the next node on overlap and jumping to the node's skip link otherwise. No per-lane stack (which would live in
// scratch memory on CDNA), one 16-byte + one 8-byte load per visited node, monotonically increasing addresses.
// Two passes with the same walk: count, then write at the scanned offsets (SoA pair buffers).
#include "engine.h"

namespace gsj {

namespace {

template <bool WRITE>
__global__ void __launch_bounds__(256) k_traverse(const float4* __restrict__ qbb, const uint32_t* __restrict__ qidx, uint32_t qbase, uint32_t nq,
                                                  const BNode* __restrict__ nodes, uint32_t nnodes, int upper, uint32_t* __restrict__ counts,
                                                  const uint64_t* __restrict__ offs, uint32_t* __restrict__ out_a, uint32_t* __restrict__ out_b) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nq) return;
  const uint32_t q = qidx ? qidx[qbase + t] : qbase + t;
  const float4 b = qbb[q];
  uint64_t o = WRITE ? offs[t] : 0;
  uint32_t cnt = 0;
  uint32_t i = 0;
  // every step moves strictly forward (left child = i + 1, skip > i), so the walk ends within nnodes steps
  while (i < nnodes) {
    const float4 c = nodes[i].box;
    const bool ov = !(c.x > b.z || c.z < b.x || c.y > b.w || c.w < b.y);
    const uint32_t skip = (uint32_t)nodes[i].skip;
    uint32_t next;
    if (!ov) next = skip;
    else {
      const int32_t prim = nodes[i].prim;
      if (prim < 0) next = i + 1;
      else {
        if (!upper || (uint32_t)prim >= q) {
          if (WRITE) { out_a[o + cnt] = q; out_b[o + cnt] = (uint32_t)prim; }
          cnt++;
        }
        next = skip;
      }
    }
    if (next <= i) break;
    i = next;
  }
  if (!WRITE) counts[t] = cnt;
}

// window membership of the probe at every walk position: its own (outward float) box meets the closed window,
// compared in double
__global__ void k_win_flags(const float4* __restrict__ bb, const uint32_t* __restrict__ order, uint32_t n, double x0, double y0, double x1, double y1,
                            uint32_t* __restrict__ f) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float4 b = bb[order ? order[i] : i];
  f[i] = ((double)b.x <= x1 && (double)b.z >= x0 && (double)b.y <= y1 && (double)b.w >= y0) ? 1u : 0u;
}

// a parent is in when any of its pieces is; then all its pieces are (thread per parent, pieces pfirst[p] .. pfirst[p+1])
__global__ void k_win_parents(const uint32_t* __restrict__ pfirst, uint32_t nparent, uint32_t* __restrict__ f) {
  const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= nparent) return;
  const uint32_t a = pfirst[p], b = pfirst[p + 1];
  if (b - a < 2) return;
  uint32_t any = 0;
  for (uint32_t i = a; i < b; i++) any |= f[i];
  if (any)
    for (uint32_t i = a; i < b; i++) f[i] = 1u;
}

__global__ void k_win_scatter(const uint32_t* __restrict__ f, const uint64_t* __restrict__ off, const uint32_t* __restrict__ order, uint32_t n,
                              uint32_t* __restrict__ out) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n && f[i]) out[off[i]] = order ? order[i] : i;
}

// tiles whose box meets the window (a probe meeting it lies in such a tile)
__global__ void k_win_tile_flags(const float4* __restrict__ tiles, uint32_t nt, double x0, double y0, double x1, double y1, uint32_t* __restrict__ f) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nt) return;
  const float4 b = tiles[t];
  f[t] = ((double)b.x <= x1 && (double)b.z >= x0 && (double)b.y <= y1 && (double)b.w >= y0) ? 1u : 0u;
}
__global__ void k_win_tile_list(const uint32_t* __restrict__ f, const uint64_t* __restrict__ off, uint32_t nt, uint32_t* __restrict__ out) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < nt && f[t]) out[off[t]] = t;
}
// membership of the probes of the selected tiles: slot k * PROBE_TILE + j holds walk position tl[k] * PROBE_TILE + j
__global__ void k_win_tile_probes(const float4* __restrict__ bb, const uint32_t* __restrict__ order, uint32_t n, const uint32_t* __restrict__ tl,
                                  uint64_t nslots, double x0, double y0, double x1, double y1, uint32_t* __restrict__ f) {
  const uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nslots) return;
  const uint64_t pos = (uint64_t)tl[k / PROBE_TILE] * PROBE_TILE + k % PROBE_TILE;
  bool in = false;
  if (pos < n) {
    const float4 b = bb[order[pos]];
    in = (double)b.x <= x1 && (double)b.z >= x0 && (double)b.y <= y1 && (double)b.w >= y0;
  }
  f[k] = in ? 1u : 0u;
}
__global__ void k_win_tile_scatter(const uint32_t* __restrict__ f, const uint64_t* __restrict__ off, const uint32_t* __restrict__ order,
                                   const uint32_t* __restrict__ tl, uint64_t nslots, uint32_t* __restrict__ out) {
  const uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k < nslots && f[k]) out[off[k]] = order[(uint64_t)tl[k / PROBE_TILE] * PROBE_TILE + k % PROBE_TILE];
}

}  // namespace

Probes probes_of(const Layer& L, const float4* qbb) {
  Probes p;
  p.own_bb = L.points ? L.pts.bb : L.poly.bb;
  p.qbb = qbb ? qbb : p.own_bb;
  p.n = (uint32_t)L.n;
  if (L.points) { p.order = L.pts.order; p.tiles = L.pts.tiles; }
  else if (L.h_parent.empty() && L.has_bvh) { p.order = L.bvh.order; p.tiles = L.bvh.tiles; }   // the tree's leaf order
  if (!L.points && !L.h_parent.empty()) {
    p.h_parent = &L.h_parent;
    p.pfirst = L.poly.pfirst;
    p.nparent = L.nparent;
  }
  return p;
}

void traverse(Context& c, const float4* qbb, const uint32_t* qidx, uint32_t qbase, uint32_t nq, const DBvh& t, bool upper, uint32_t* counts,
              const uint64_t* offs, uint32_t* out_a, uint32_t* out_b) {
  if (nq == 0) return;
  if (t.nnodes == 0) {
    if (!offs) GSJ_HIP(hipMemsetAsync(counts, 0, (size_t)nq * 4, c.stream));
    return;
  }
  if (!offs)
    hipLaunchKernelGGL(k_traverse<false>, dim3(div_up(nq, 256)), dim3(256), 0, c.stream, qbb, qidx, qbase, nq, t.nodes, t.nnodes, upper ? 1 : 0,
                       counts, nullptr, nullptr, nullptr);
  else
    hipLaunchKernelGGL(k_traverse<true>, dim3(div_up(nq, 256)), dim3(256), 0, c.stream, qbb, qidx, qbase, nq, t.nodes, t.nnodes, upper ? 1 : 0,
                       nullptr, offs, out_a, out_b);
  GSJ_LAUNCHED();
}

void for_each_candidate_batch(Context& c, const Probes& pr, const DBvh& t, bool upper, const gsj_join_opts& o, gsj_join_stats& s,
                              const PairsFn& fn) {
  hipStream_t stream = c.stream;
  const uint32_t batch_q = o.batch_q ? o.batch_q : (1u << 22);
  const uint64_t max_pairs = o.max_pairs ? o.max_pairs : (128ull << 20);
  const float4* qbb = pr.qbb;
  static const std::vector<uint32_t> none;
  const std::vector<uint32_t>& hp = pr.h_parent ? *pr.h_parent : none;
  // probe positions 0 .. nQ: probe ids directly, through the walk order, or through the window's member list (in walk
  // order); a subdivided layer walks in id order (its parents' pieces are contiguous)
  uint32_t nQ = pr.n;
  const uint32_t* qidx = pr.h_parent && !pr.h_parent->empty() ? nullptr : pr.order;
  std::vector<uint32_t> hq;
  if (o.has_window && pr.n) {
    const double* w = o.window;
    GSJ_REQUIRE(w[0] <= w[2] && w[1] <= w[3], GSJ_E_ARG, "window must have x0 <= x1 and y0 <= y1");
    WallTimer tw;
    uint32_t* list = nullptr;
    if (qidx && pr.tiles) {
      // through the tile boxes of the walk order: only the tiles meeting the window are scanned, and their members
      // come out in walk order
      const uint32_t nt = (pr.n + PROBE_TILE - 1) / PROBE_TILE;
      uint32_t* tf = c.s_win.as<uint32_t>((uint64_t)nt + 16);
      uint64_t* toff = c.s_win_off.as<uint64_t>((uint64_t)nt + 16);
      hipLaunchKernelGGL(k_win_tile_flags, dim3(div_up(nt, 256)), dim3(256), 0, stream, pr.tiles, nt, w[0], w[1], w[2], w[3], tf);
      GSJ_LAUNCHED();
      const uint32_t nsel = (uint32_t)scan_u32_to_u64(c, tf, toff, nt);
      uint32_t* tl = c.s_qidx.as<uint32_t>((uint64_t)nsel + 16);
      hipLaunchKernelGGL(k_win_tile_list, dim3(div_up(nt, 256)), dim3(256), 0, stream, tf, toff, nt, tl);
      GSJ_LAUNCHED();
      const uint64_t nslots = (uint64_t)nsel * PROBE_TILE;
      nQ = 0;
      if (nslots) {
        uint32_t* f = c.s_win.as<uint32_t>(nslots + 16);
        uint64_t* off = c.s_win_off.as<uint64_t>(nslots + 16);
        hipLaunchKernelGGL(k_win_tile_probes, dim3(div_up(nslots, 256)), dim3(256), 0, stream, pr.own_bb, qidx, pr.n, tl, nslots, w[0], w[1], w[2],
                           w[3], f);
        GSJ_LAUNCHED();
        nQ = (uint32_t)scan_u32_to_u64(c, f, off, nslots);
        list = c.s_qlist.as<uint32_t>((uint64_t)nQ + 16);
        hipLaunchKernelGGL(k_win_tile_scatter, dim3(div_up(nslots, 256)), dim3(256), 0, stream, f, off, qidx, tl, nslots, list);
        GSJ_LAUNCHED();
      }
    } else {
      uint32_t* f = c.s_win.as<uint32_t>((uint64_t)pr.n + 16);
      hipLaunchKernelGGL(k_win_flags, dim3(div_up(pr.n, 256)), dim3(256), 0, stream, pr.own_bb, qidx, pr.n, w[0], w[1], w[2], w[3], f);
      GSJ_LAUNCHED();
      if (pr.pfirst && pr.nparent) {
        hipLaunchKernelGGL(k_win_parents, dim3(div_up(pr.nparent, 256)), dim3(256), 0, stream, pr.pfirst, pr.nparent, f);
        GSJ_LAUNCHED();
      }
      uint64_t* off = c.s_win_off.as<uint64_t>((uint64_t)pr.n + 16);
      nQ = (uint32_t)scan_u32_to_u64(c, f, off, pr.n);
      // members written behind the flags (the flags are read by the scatter itself, so they need their own region)
      list = c.s_qlist.as<uint32_t>((uint64_t)nQ + 16);
      hipLaunchKernelGGL(k_win_scatter, dim3(div_up(pr.n, 256)), dim3(256), 0, stream, f, off, qidx, pr.n, list);
      GSJ_LAUNCHED();
    }
    hq.resize(nQ);
    if (nQ) GSJ_HIP(hipMemcpyAsync(hq.data(), list, (size_t)nQ * 4, hipMemcpyDeviceToHost, stream));
    GSJ_HIP(hipStreamSynchronize(stream));
    qidx = list;
    s.traverse_ms += tw.ms();
  }
  auto pid = [&](uint32_t pos) { return hp[qidx ? hq[pos] : pos]; };   // parent of the probe at a position
  std::vector<uint32_t> hcounts;
  for (uint32_t q0 = 0, qn; q0 < nQ; q0 = qn) {
    qn = (uint32_t)std::min<uint64_t>((uint64_t)q0 + batch_q, nQ);
    if (!hp.empty()) while (qn < nQ && pid(qn) == pid(qn - 1)) qn++;
    if (o.cancel && *o.cancel) throw Error(GSJ_E_CANCEL, "cancelled");
    const uint32_t nq = qn - q0;
    WallTimer tt;
    uint32_t* counts = c.s_counts.as<uint32_t>(nq + 16);
    traverse(c, qbb, qidx, q0, nq, t, upper, counts, nullptr, nullptr, nullptr);
    uint64_t* offs = c.s_qoffs.as<uint64_t>((uint64_t)nq + 16);   // probe offsets (s_offs belongs to compaction)
    const uint64_t total = scan_u32_to_u64(c, counts, offs, nq);
    s.traverse_ms += tt.ms();
    // sub-batches of at most max_pairs candidates (a probe parent's pieces stay together)
    std::vector<std::pair<uint32_t, uint32_t>> subs;
    if (total <= max_pairs) subs.push_back({0, nq});
    else {
      hcounts.resize(nq);
      GSJ_HIP(hipMemcpyAsync(hcounts.data(), counts, (size_t)nq * 4, hipMemcpyDeviceToHost, stream));
      GSJ_HIP(hipStreamSynchronize(stream));
      for (uint32_t s0 = 0; s0 < nq;) {
        uint64_t np = 0;
        uint32_t s1 = s0;
        while (s1 < nq && (np + hcounts[s1] <= max_pairs || s1 == s0 || (!hp.empty() && pid(q0 + s1) == pid(q0 + s1 - 1)))) np += hcounts[s1++];
        GSJ_REQUIRE(np < (1ull << 32), GSJ_E_LIMIT, "a probe batch has >= 2^32 candidates (lower max_pairs)");
        subs.push_back({s0, s1});
        s0 = s1;
      }
    }
    for (auto [s0, s1] : subs) {
      if (o.cancel && *o.cancel) throw Error(GSJ_E_CANCEL, "cancelled");
      WallTimer t1;
      uint64_t np = total;
      if (subs.size() > 1) np = scan_u32_to_u64(c, counts + s0, offs, s1 - s0);
      s.batches++;
      s.candidates += np;
      if (np == 0) continue;
      GSJ_REQUIRE(np < (1ull << 32), GSJ_E_LIMIT, "a sub-batch has >= 2^32 candidates (lower max_pairs)");
      uint32_t* pa = c.s_pairs_a.as<uint32_t>(np + 16);
      uint32_t* pb = c.s_pairs_b.as<uint32_t>(np + 16);
      traverse(c, qbb, qidx, q0 + s0, s1 - s0, t, upper, nullptr, offs, pa, pb);
      GSJ_HIP(hipStreamSynchronize(stream));
      s.traverse_ms += t1.ms();
      fn(pa, pb, np);
    }
  }
}

}  // namespace gsj
