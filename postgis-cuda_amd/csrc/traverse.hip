// Candidate search (reference: k_traverse, engine/engine.cuh:233-259): every leaf whose float box overlaps the
// probe's box (closed test), restricted to leaf >= probe for self-joins.
//
// Stackless: nodes are in depth-first preorder, so a lane walks forward through the node array, descending into
// the next node on overlap and jumping to the node's skip link otherwise. No per-lane stack (which would live in
// scratch memory on CDNA), one 16-byte + one 8-byte load per visited node, monotonically increasing addresses.
// One walk per probe in the common case: it counts the probe's candidates and keeps the first SLOT_K of them in the
// probe's slots; after the scan, a gather writes them at their offsets (SoA pair buffers), and only the probes with
// more candidates than slots walk again, writing at their offsets.
#include "engine.h"

namespace gsj {

namespace {

constexpr uint32_t SLOT_K = 16;   // candidates kept per probe by the counting walk

// WRITE false: count the probe's candidates into counts[t] and keep the first SLOT_K in slots[t * SLOT_K ..]
// (probes with more are counted in *nover); WRITE true: the probes with counts[t] > SLOT_K write all theirs at offs[t]
template <bool WRITE>
__global__ void __launch_bounds__(256) k_traverse(const float4* __restrict__ qbb, const uint32_t* __restrict__ qidx, uint32_t qbase, uint32_t nq,
                                                  const BNode* __restrict__ nodes, uint32_t nnodes, int upper, uint32_t* __restrict__ counts,
                                                  const uint64_t* __restrict__ offs, uint32_t* __restrict__ out_a, uint32_t* __restrict__ out_b,
                                                  int box_by_pos, uint32_t* __restrict__ slots, uint32_t* __restrict__ nover) {
  const uint32_t t = xcd_block() * blockDim.x + threadIdx.x;
  if (t >= nq) return;
  if (WRITE && counts[t] <= SLOT_K) return;
  const uint32_t q = qidx ? qidx[qbase + t] : qbase + t;
  const float4 b = qbb[box_by_pos ? qbase + t : q];
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
          else if (cnt < SLOT_K) slots[(uint64_t)t * SLOT_K + cnt] = (uint32_t)prim;
          cnt++;
        }
        next = skip;
      }
    }
    if (next <= i) break;
    i = next;
  }
  if (!WRITE) {
    counts[t] = cnt;
    if (cnt > SLOT_K) atomicAdd(nover, 1u);
  }
}

__global__ void k_zero_u32(uint32_t* p, uint64_t n) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0u;
}

// the kept candidates of the probes with at most SLOT_K, at their offsets
__global__ void k_slot_gather(const uint32_t* __restrict__ qidx, uint32_t qbase, uint32_t nq, const uint32_t* __restrict__ counts,
                              const uint64_t* __restrict__ offs, const uint32_t* __restrict__ slots, uint32_t* __restrict__ out_a,
                              uint32_t* __restrict__ out_b) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nq) return;
  const uint32_t n = counts[t];
  if (n == 0 || n > SLOT_K) return;
  const uint32_t q = qidx ? qidx[qbase + t] : qbase + t;
  const uint64_t o = offs[t];
  for (uint32_t j = 0; j < n; j++) {
    out_a[o + j] = q;
    out_b[o + j] = slots[(uint64_t)t * SLOT_K + j];
  }
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

__device__ __forceinline__ bool box_meets(const float4 b, const double* w) {
  return (double)b.x <= w[2] && (double)b.z >= w[0] && (double)b.y <= w[3] && (double)b.w >= w[1];
}

// Windows over the tile boxes. Entry e = r * nt + t stands for tile t of window r; a probe meeting window r lies in
// a tile whose box meets it.
__global__ void k_win_tile_flags(const float4* __restrict__ tiles, uint32_t nt, const WindowSet ws, uint32_t* __restrict__ f) {
  const uint32_t e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= ws.k * nt) return;
  const uint32_t r = e / nt;
  f[e] = box_meets(tiles[e - r * nt], ws.w[r]) ? 1u : 0u;
}
__global__ void k_win_tile_list(const uint32_t* __restrict__ f, const uint64_t* __restrict__ off, uint32_t ne, uint32_t* __restrict__ out) {
  const uint32_t e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e < ne && f[e]) out[off[e]] = e;
}
// first selected tile of each window: tstart[r] = selected entries before r * nt (tstart[k] = all of them)
__global__ void k_win_tstart(const uint64_t* __restrict__ off, uint32_t nt, uint32_t k, uint64_t nsel, uint64_t* __restrict__ tstart) {
  const uint32_t r = threadIdx.x;
  if (r <= k) tstart[r] = r < k ? off[(uint64_t)r * nt] : nsel;
}
// membership of the probes of the selected tiles: slot s holds walk position t * PROBE_TILE + s % PROBE_TILE of the
// entry tl[s / PROBE_TILE] = r * nt + t, tested against window r
__global__ void k_win_tile_probes(const float4* __restrict__ bb, const uint32_t* __restrict__ order, uint32_t n, uint32_t nt,
                                  const uint32_t* __restrict__ tl, uint64_t nslots, const WindowSet ws, uint32_t* __restrict__ f) {
  const uint64_t sl = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (sl >= nslots) return;
  const uint32_t e = tl[sl / PROBE_TILE], r = e / nt;
  const uint64_t pos = (uint64_t)(e - r * nt) * PROBE_TILE + sl % PROBE_TILE;
  f[sl] = (pos < n && box_meets(bb[order[pos]], ws.w[r])) ? 1u : 0u;
}
__global__ void k_win_tile_scatter(const uint32_t* __restrict__ f, const uint64_t* __restrict__ off, const uint32_t* __restrict__ order, uint32_t nt,
                                   const uint32_t* __restrict__ tl, uint64_t nslots, uint32_t* __restrict__ out) {
  const uint64_t sl = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (sl >= nslots || !f[sl]) return;
  const uint32_t e = tl[sl / PROBE_TILE];
  out[off[sl]] = order[(uint64_t)(e % nt) * PROBE_TILE + sl % PROBE_TILE];
}
// first member position of each window: mstart[r] = members before slot tstart[r] * PROBE_TILE (mstart[k] = all)
__global__ void k_win_mstart(const uint64_t* __restrict__ tstart, uint32_t k, const uint64_t* __restrict__ off, uint64_t nslots, uint64_t nmem,
                             uint64_t* __restrict__ mstart) {
  const uint32_t r = threadIdx.x;
  if (r > k) return;
  const uint64_t sl = tstart[r] * PROBE_TILE;
  mstart[r] = (r < k && sl < nslots) ? off[sl] : nmem;
}
// candidate boundaries of the windows inside the sub-batch of walk positions [base, base + nsub) with candidate
// offsets offs[0 .. nsub) and np candidates in all
__global__ void k_seg_bounds(const uint64_t* __restrict__ mstart, uint32_t k, uint64_t base, uint64_t nsub, const uint64_t* __restrict__ offs, uint64_t np,
                             uint64_t* __restrict__ lo) {
  const uint32_t r = threadIdx.x;
  if (r > k) return;
  const uint64_t m = mstart[r];
  lo[r] = m <= base ? 0 : (m >= base + nsub ? np : offs[m - base]);
}

// Members of the k windows through the tile boxes: their probe ids (window after window, each in walk order) in
// *list and the member boundaries mstart[0 .. k] on the device; returns the member count.
uint32_t window_members_tiled(Context& c, const Probes& pr, const WindowSet& ws, uint32_t*& list, uint64_t*& mstart) {
  hipStream_t stream = c.stream;
  const uint32_t nt = (pr.n + PROBE_TILE - 1) / PROBE_TILE;
  GSJ_REQUIRE(ws.k >= 1 && ws.k <= MAX_WINDOWS, GSJ_E_ARG, "window count out of range");
  GSJ_REQUIRE((uint64_t)ws.k * nt < (1ull << 32), GSJ_E_LIMIT, "too many windows for the layer's tile count");
  for (uint32_t r = 0; r < ws.k; r++)
    GSJ_REQUIRE(ws.w[r][0] <= ws.w[r][2] && ws.w[r][1] <= ws.w[r][3], GSJ_E_ARG, "window must have x0 <= x1 and y0 <= y1");
  const uint32_t ne = ws.k * nt;
  uint32_t* tf = c.s_wtile.as<uint32_t>((uint64_t)ne + 16);
  uint64_t* toff = c.s_wtoff.as<uint64_t>((uint64_t)ne + 16);
  uint64_t* tstart = wseg(c, 0);
  mstart = wseg(c, 2);
  hipLaunchKernelGGL(k_win_tile_flags, dim3(div_up(ne, 256)), dim3(256), 0, stream, pr.tiles, nt, ws, tf);
  GSJ_LAUNCHED();
  const uint64_t nsel = scan_u32_to_u64(c, tf, toff, ne);
  hipLaunchKernelGGL(k_win_tstart, dim3(1), dim3(64), 0, stream, toff, nt, ws.k, nsel, tstart);
  GSJ_LAUNCHED();
  uint32_t* tl = c.s_qidx.as<uint32_t>(nsel + 16);
  hipLaunchKernelGGL(k_win_tile_list, dim3(div_up(ne, 256)), dim3(256), 0, stream, tf, toff, ne, tl);
  GSJ_LAUNCHED();
  const uint64_t nslots = nsel * PROBE_TILE;
  uint64_t nmem = 0;
  list = nullptr;
  if (nslots) {
    uint32_t* f = c.s_win.as<uint32_t>(nslots + 16);
    uint64_t* off = c.s_win_off.as<uint64_t>(nslots + 16);
    hipLaunchKernelGGL(k_win_tile_probes, dim3(div_up(nslots, 256)), dim3(256), 0, stream, pr.own_bb, pr.order, pr.n, nt, tl, nslots, ws, f);
    GSJ_LAUNCHED();
    nmem = scan_u32_to_u64(c, f, off, nslots);
    GSJ_REQUIRE(nmem < (1ull << 32), GSJ_E_LIMIT, "window members exceed 2^32");
    list = c.s_qlist.as<uint32_t>(nmem + 16);
    hipLaunchKernelGGL(k_win_tile_scatter, dim3(div_up(nslots, 256)), dim3(256), 0, stream, f, off, pr.order, nt, tl, nslots, list);
    GSJ_LAUNCHED();
    hipLaunchKernelGGL(k_win_mstart, dim3(1), dim3(64), 0, stream, tstart, ws.k, off, nslots, nmem, mstart);
    GSJ_LAUNCHED();
  } else {
    hipLaunchKernelGGL(k_win_mstart, dim3(1), dim3(64), 0, stream, tstart, ws.k, nullptr, (uint64_t)0, (uint64_t)0, mstart);
    GSJ_LAUNCHED();
  }
  return (uint32_t)nmem;
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

void traverse_count(Context& c, const float4* qbb, const uint32_t* qidx, uint32_t qbase, uint32_t nq, const DBvh& t, bool upper, uint32_t* counts,
                    uint32_t* slots, uint32_t* nover, bool box_by_pos) {
  if (nq == 0) return;
  if (t.nnodes == 0) {
    hipLaunchKernelGGL(k_zero_u32, dim3(div_up(nq, 256)), dim3(256), 0, c.stream, counts, (uint64_t)nq);
    GSJ_LAUNCHED();
    return;
  }
  const uint32_t tb = c.shape(nq).trav_block;   // one thread per probe; the bucket is chosen by the probe count
  hipLaunchKernelGGL(k_traverse<false>, dim3(div_up(nq, tb)), dim3(tb), 0, c.stream, qbb, qidx, qbase, nq, t.nodes, t.nnodes, upper ? 1 : 0, counts,
                     nullptr, nullptr, nullptr, box_by_pos ? 1 : 0, slots, nover);
  GSJ_LAUNCHED();
}

void traverse_write(Context& c, const float4* qbb, const uint32_t* qidx, uint32_t qbase, uint32_t nq, const DBvh& t, bool upper,
                    const uint32_t* counts, const uint64_t* offs, const uint32_t* slots, bool any_over, uint32_t* out_a, uint32_t* out_b,
                    bool box_by_pos) {
  if (nq == 0 || t.nnodes == 0) return;
  hipLaunchKernelGGL(k_slot_gather, dim3(div_up(nq, 256)), dim3(256), 0, c.stream, qidx, qbase, nq, counts, offs, slots, out_a, out_b);
  GSJ_LAUNCHED();
  if (any_over) {
    const uint32_t tb = c.shape(nq).trav_block;
    hipLaunchKernelGGL(k_traverse<true>, dim3(div_up(nq, tb)), dim3(tb), 0, c.stream, qbb, qidx, qbase, nq, t.nodes, t.nnodes, upper ? 1 : 0,
                       const_cast<uint32_t*>(counts), offs, out_a, out_b, box_by_pos ? 1 : 0, nullptr, nullptr);
    GSJ_LAUNCHED();
  }
}

namespace {

// The probe batches and candidate sub-batches over probe positions 0 .. nQ (probe ids directly, or qidx[pos]); with
// mstart (k windows' member boundaries on the device), every sub-batch also carries the windows' candidate
// boundaries.
void walk_batches(Context& c, const Probes& pr, const DBvh& t, bool upper, const gsj_join_opts& o, gsj_join_stats& s, const uint32_t* qidx,
                  uint32_t nQ, const std::vector<uint32_t>& hq, const uint64_t* mstart, uint32_t k, const float4* qbb, bool box_by_pos,
                  const SegPairsFn& fn) {
  hipStream_t stream = c.stream;
  const uint32_t batch_q = o.batch_q ? o.batch_q : (1u << 22);
  const uint64_t max_pairs = o.max_pairs ? o.max_pairs : (128ull << 20);
  static const std::vector<uint32_t> none;
  const std::vector<uint32_t>& hp = pr.h_parent ? *pr.h_parent : none;
  auto pid = [&](uint32_t pos) { return hp[qidx ? hq[pos] : pos]; };   // parent of the probe at a position
  std::vector<uint32_t> hcounts;
  uint64_t* dlo = mstart ? wseg(c, 1) : nullptr;
  uint64_t* hlo = mstart ? c.h_seg.as<uint64_t>(MAX_WINDOWS + 1) : nullptr;
  for (uint32_t q0 = 0, qn; q0 < nQ; q0 = qn) {
    qn = (uint32_t)std::min<uint64_t>((uint64_t)q0 + batch_q, nQ);
    if (!hp.empty()) while (qn < nQ && pid(qn) == pid(qn - 1)) qn++;
    if (o.cancel && *o.cancel) throw Error(GSJ_E_CANCEL, "cancelled");
    const uint32_t nq = qn - q0;
    WallTimer tt;
    uint32_t* counts = c.s_counts.as<uint32_t>(nq + 16);
    uint32_t* slots = c.s_slots.as<uint32_t>((uint64_t)nq * SLOT_K + 16);
    uint32_t* nover = c.s_slots_ctl.as<uint32_t>(16);
    uint32_t* hover = c.h_slots_ctl.as<uint32_t>(4);
    hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(64), 0, stream, nover, (uint64_t)1);
    GSJ_LAUNCHED();
    GSJ_HIP(hipEventRecord(c.ev[12], stream));
    traverse_count(c, qbb, qidx, q0, nq, t, upper, counts, slots, nover, box_by_pos);
    GSJ_HIP(hipEventRecord(c.ev[13], stream));
    GSJ_HIP(hipMemcpyAsync(hover, nover, 4, hipMemcpyDeviceToHost, stream));
    uint64_t* offs = c.s_qoffs.as<uint64_t>((uint64_t)nq + 16);   // probe offsets (s_offs belongs to compaction)
    const uint64_t total = scan_u32_to_u64(c, counts, offs, nq);   // synchronises: *hover is there too
    const bool any_over = *hover != 0;
    {
      float ms = 0;
      GSJ_HIP(hipEventElapsedTime(&ms, c.ev[12], c.ev[13]));
      s.k_traverse_ms += ms;
    }
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
      GSJ_HIP(hipEventRecord(c.ev[14], stream));
      traverse_write(c, qbb, qidx, q0 + s0, s1 - s0, t, upper, counts + s0, offs, slots + (uint64_t)s0 * SLOT_K, any_over, pa, pb, box_by_pos);
      GSJ_HIP(hipEventRecord(c.ev[15], stream));
      SegBounds sb;
      if (mstart) {
        hipLaunchKernelGGL(k_seg_bounds, dim3(1), dim3(64), 0, stream, mstart, k, (uint64_t)(q0 + s0), (uint64_t)(s1 - s0), offs, np, dlo);
        GSJ_LAUNCHED();
        GSJ_HIP(hipMemcpyAsync(hlo, dlo, (size_t)(k + 1) * 8, hipMemcpyDeviceToHost, stream));
        sb.k = k; sb.h = hlo; sb.d = dlo;
        GSJ_HIP(hipStreamSynchronize(stream));   // the bounds are read on the host
      }
      // otherwise no wait here: the pair work follows in stream order
      s.traverse_ms += t1.ms();
      fn(pa, pb, np, sb);
      {
        float ms = 0;
        GSJ_HIP(hipEventSynchronize(c.ev[15]));   // done already when the pair work read anything back
        GSJ_HIP(hipEventElapsedTime(&ms, c.ev[14], c.ev[15]));
        s.k_traverse_ms += ms;
      }
    }
  }
}

}  // namespace

bool probes_batchable(const Probes& pr) { return pr.order && pr.tiles && !pr.h_parent; }

void for_each_candidate_batch(Context& c, const Probes& pr, const DBvh& t, bool upper, const gsj_join_opts& o, gsj_join_stats& s,
                              const PairsFn& fn, const MembersFn& on_members) {
  hipStream_t stream = c.stream;
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
      WindowSet ws;
      ws.k = 1;
      for (int j = 0; j < 4; j++) ws.w[0][j] = w[j];
      uint64_t* mstart = nullptr;
      nQ = window_members_tiled(c, pr, ws, list, mstart);
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
    if (pr.h_parent && !pr.h_parent->empty()) {   // parents of the members, for keeping their pieces together
      hq.resize(nQ);
      if (nQ) GSJ_HIP(hipMemcpyAsync(hq.data(), list, (size_t)nQ * 4, hipMemcpyDeviceToHost, stream));
    }
    GSJ_HIP(hipStreamSynchronize(stream));
    qidx = list;
    s.traverse_ms += tw.ms();
  }
  const float4* qbb = pr.qbb;
  if (on_members) {
    const float4* b = on_members(o.has_window ? qidx : nullptr, nQ, nullptr);
    if (b) qbb = b;
  }
  walk_batches(c, pr, t, upper, o, s, qidx, nQ, hq, nullptr, 0, qbb, false,
               [&](uint32_t* pa, uint32_t* pb, uint64_t np, const SegBounds&) { fn(pa, pb, np); });
}

void for_each_candidate_multi(Context& c, const Probes& pr, const DBvh& t, bool upper, const gsj_join_opts& o, const WindowSet& ws,
                              gsj_join_stats& s, const SegPairsFn& fn, const MembersFn& on_members) {
  GSJ_REQUIRE(probes_batchable(pr), GSJ_E_ARG, "batched windows need probes with tile boxes and no parents");
  if (pr.n == 0) return;
  WallTimer tw;
  uint32_t* list = nullptr;
  uint64_t* mstart = nullptr;
  const uint32_t nQ = window_members_tiled(c, pr, ws, list, mstart);
  GSJ_HIP(hipStreamSynchronize(c.stream));
  s.traverse_ms += tw.ms();
  static const std::vector<uint32_t> none;
  const float4* mbb = on_members && nQ ? on_members(list, nQ, mstart) : nullptr;   // boxes by member position
  walk_batches(c, pr, t, upper, o, s, list, nQ, none, mstart, ws.k, mbb ? mbb : pr.qbb, mbb != nullptr, fn);
}

namespace {
// rows of the batched requests: rows of candidates [lo[r], lo[r + 1]) are [rlo[r], rlo[r + 1]), from the exclusive
// scan offs of the candidates' kept flags (an order-preserving compaction)
__global__ void k_row_bounds(const uint64_t* __restrict__ lo, uint32_t k, const uint64_t* __restrict__ offs, uint64_t np, uint64_t nrows,
                             uint64_t* __restrict__ rlo) {
  const uint32_t r = threadIdx.x;
  if (r > k) return;
  const uint64_t i = lo[r];
  rlo[r] = i < np ? offs[i] : nrows;
}
}  // namespace

uint64_t* wseg(Context& c, int region) { return c.s_wseg.as<uint64_t>(8 * (MAX_WINDOWS + 1)) + (size_t)region * (MAX_WINDOWS + 1); }

void seg_bounds_dev(Context& c, const uint64_t* lo, uint32_t k, const uint64_t* offs, uint64_t np, uint64_t nrows, uint64_t* out) {
  hipLaunchKernelGGL(k_row_bounds, dim3(1), dim3(64), 0, c.stream, lo, k, offs, np, nrows, out);
  GSJ_LAUNCHED();
}

void seg_rows(Context& c, const SegBounds& sb, const uint64_t* offs, uint64_t np, uint64_t nrows, uint64_t* h_rlo) {
  uint64_t* d = wseg(c, 3);
  seg_bounds_dev(c, sb.d, sb.k, offs, np, nrows, d);
  GSJ_HIP(hipMemcpyAsync(h_rlo, d, (size_t)(sb.k + 1) * 8, hipMemcpyDeviceToHost, c.stream));
  GSJ_HIP(hipStreamSynchronize(c.stream));
}

}  // namespace gsj
