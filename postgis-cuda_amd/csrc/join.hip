// Join driver: probe layer P against build layer R's LBVH (reference: join_stream, engine/engine.cuh:895-1061).
//
// Per probe batch (<= batch_q probes, a subdivided parent's pieces never split across batches): count pass, scan,
// write pass, refinement, compaction, piece -> parent aggregation, DE-9IM codes; then either a device-side consumer
// (cluster) or one copy of the rows into pinned host memory and the row callback.
//
// Compaction is split by what a candidate needs: a pair whose two parents are single pieces maps to its parent pair
// directly (kept iff its flag is 1). Only pairs touching a subdivided parent are sorted and folded, first by piece
// pair and then (stable) by parent pair, so every parent's terms are added in the same order as in the CPU
// restatement: a subdivided layer's aggregated areas are bit-identical too.
#include "engine.h"
#include <algorithm>
#include <chrono>

namespace gsj {

namespace {

__device__ __forceinline__ uint32_t parent_of(const DPolys& S, uint32_t g) { return S.parent ? S.parent[g] : g; }
__device__ __forceinline__ bool multi_piece(const DPolys& S, uint32_t p) { return S.pfirst && (S.pfirst[p + 1] - S.pfirst[p] > 1); }

// flags: single-piece rows (kept iff flag 1) and multi-piece candidates (kept iff flag != 0)
__global__ void k_row_flags(const DPolys P, const DPolys R, const uint32_t* __restrict__ pa, const uint32_t* __restrict__ pb,
                            const uint8_t* __restrict__ isect, uint64_t n, int no_agg, uint32_t* __restrict__ fs, uint32_t* __restrict__ fm) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint8_t f = isect[i];
  if (no_agg) { fs[i] = f != 0; fm[i] = 0; return; }
  const bool multi = multi_piece(P, parent_of(P, pa[i])) || multi_piece(R, parent_of(R, pb[i]));
  fs[i] = !multi && f == 1;
  fm[i] = multi && f != 0;
}

__global__ void k_emit_single(const DPolys P, const DPolys R, const uint32_t* __restrict__ pa, const uint32_t* __restrict__ pb,
                              const uint32_t* __restrict__ fs, const uint64_t* __restrict__ offs, uint64_t n, int no_agg, const double* __restrict__ area,
                              const double4* __restrict__ ev, const uint32_t* __restrict__ evb, uint32_t* __restrict__ orow, double* __restrict__ oarea,
                              uint32_t* __restrict__ ocode) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n || !fs[i]) return;
  const uint64_t o = offs[i];
  const uint32_t a = no_agg ? pa[i] : parent_of(P, pa[i]), b = no_agg ? pb[i] : parent_of(R, pb[i]);
  orow[2 * o] = a; orow[2 * o + 1] = b;
  if (oarea) oarea[o] = area[i];
  if (ocode) {
    const double4 e = ev[i];
    ocode[o] = de9im_code(e.x, e.y, e.z, e.w, evb[i], P.blen[parent_of(P, pa[i])], R.blen[parent_of(R, pb[i])]);
  }
}

__global__ void k_multi_keys(const uint32_t* __restrict__ pa, const uint32_t* __restrict__ pb, const uint32_t* __restrict__ fm,
                             const uint64_t* __restrict__ offs, uint64_t n, uint64_t* __restrict__ keys, uint32_t* __restrict__ vals) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n || !fm[i]) return;
  const uint64_t o = offs[i];
  keys[o] = ((uint64_t)pa[i] << 32) | pb[i];
  vals[o] = (uint32_t)i;
}

__global__ void k_parent_keys(const DPolys P, const DPolys R, const uint32_t* __restrict__ pa, const uint32_t* __restrict__ pb,
                              const uint32_t* __restrict__ vals, uint64_t m, uint64_t* __restrict__ keys) {
  const uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= m) return;
  const uint32_t i = vals[j];
  keys[j] = ((uint64_t)parent_of(P, pa[i]) << 32) | parent_of(R, pb[i]);
}

__global__ void k_heads(const uint64_t* __restrict__ keys, uint64_t m, uint32_t* __restrict__ head) {
  const uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j < m) head[j] = (j == 0 || keys[j] != keys[j - 1]) ? 1u : 0u;
}

__global__ void k_seg_starts(const uint32_t* __restrict__ head, const uint64_t* __restrict__ hoff, uint64_t m, uint64_t* __restrict__ start) {
  const uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j < m && head[j]) start[hoff[j]] = j;
}

// one thread per parent pair: terms added in (piece a, piece b) order; flag = min over its piece pairs
__global__ void k_seg_fold(const uint64_t* __restrict__ keys, const uint32_t* __restrict__ vals, const uint64_t* __restrict__ start, uint64_t nseg,
                           uint64_t m, const uint8_t* __restrict__ isect, const double* __restrict__ area, const double4* __restrict__ ev,
                           const uint32_t* __restrict__ evb, uint32_t* __restrict__ keep, double* __restrict__ sarea, double4* __restrict__ sev,
                           uint32_t* __restrict__ sevb) {
  const uint64_t s = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= nseg) return;
  const uint64_t j0 = start[s], j1 = s + 1 < nseg ? start[s + 1] : m;
  double ar = 0, e0 = 0, e1 = 0, e2 = 0, e3 = 0;
  uint32_t bits = 0;
  uint8_t f = 255;
  for (uint64_t j = j0; j < j1; j++) {
    const uint32_t i = vals[j];
    if (area) ar += area[i];
    if (ev) { const double4 e = ev[i]; e0 += e.x; e1 += e.y; e2 += e.z; e3 += e.w; bits |= evb[i]; }
    f = min(f, isect[i]);
  }
  keep[s] = f == 1;
  sarea[s] = ar;
  if (ev) { sev[s] = make_double4(e0, e1, e2, e3); sevb[s] = bits; }
}

__global__ void k_emit_multi(const DPolys P, const DPolys R, const uint64_t* __restrict__ keys, const uint64_t* __restrict__ start,
                             const uint32_t* __restrict__ keep, const uint64_t* __restrict__ koff, uint64_t nseg, uint64_t base,
                             const double* __restrict__ sarea, const double4* __restrict__ sev, const uint32_t* __restrict__ sevb,
                             uint32_t* __restrict__ orow, double* __restrict__ oarea, uint32_t* __restrict__ ocode, int want_area) {
  const uint64_t s = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= nseg || !keep[s]) return;
  const uint64_t o = base + koff[s];
  const uint64_t k = keys[start[s]];
  const uint32_t a = (uint32_t)(k >> 32), b = (uint32_t)k;
  orow[2 * o] = a; orow[2 * o + 1] = b;
  if (want_area) oarea[o] = sarea[s];
  if (ocode) { const double4 e = sev[s]; ocode[o] = de9im_code(e.x, e.y, e.z, e.w, sevb[s], P.blen[a], R.blen[b]); }
}

inline int bits_for(uint64_t maxval) { int b = 1; while (b < 64 && (maxval >> b)) b++; return b; }

}  // namespace

// Compact the refined candidates of one sub-batch into parent-pair rows on the device; returns the row count.
static uint64_t compact_rows(Context& c, const Layer& PL, const Layer& RL, const uint32_t* pa, const uint32_t* pb, uint64_t np, const RefineParams& rp,
                             const RefineOut& ro, bool no_agg, uint32_t* orow, double* oarea, uint32_t* ocode, const uint64_t** single_offs = nullptr) {
  hipStream_t st = c.stream;
  const DPolys& P = PL.poly;
  const DPolys& R = RL.poly;
  uint32_t* fs = c.s_tmp.as<uint32_t>(2 * np + 16);
  uint32_t* fm = fs + np;
  uint64_t* offs = c.s_offs.as<uint64_t>(np + 16);
  hipLaunchKernelGGL(k_row_flags, dim3(div_up(np, 256)), dim3(256), 0, st, P, R, pa, pb, ro.isect, np, no_agg ? 1 : 0, fs, fm);
  GSJ_LAUNCHED();
  const uint64_t nsingle = scan_u32_to_u64(c, fs, offs, np);
  if (single_offs) *single_offs = offs;
  hipLaunchKernelGGL(k_emit_single, dim3(div_up(np, 256)), dim3(256), 0, st, P, R, pa, pb, fs, offs, np, no_agg ? 1 : 0, ro.area, ro.ev, ro.evb,
                     orow, rp.want_area ? oarea : nullptr, rp.want_relate ? ocode : nullptr);
  GSJ_LAUNCHED();
  const bool any_multi = !no_agg && (P.pfirst || R.pfirst);
  if (!any_multi) return nsingle;
  const uint64_t m = scan_u32_to_u64(c, fm, offs, np);
  if (m == 0) return nsingle;
  uint64_t* keys = c.s_keys.as<uint64_t>(m);
  uint64_t* keys2 = c.s_keys2.as<uint64_t>(m);
  uint32_t* vals = c.s_vals.as<uint32_t>(m);
  uint32_t* vals2 = c.s_vals2.as<uint32_t>(m);
  hipLaunchKernelGGL(k_multi_keys, dim3(div_up(np, 256)), dim3(256), 0, st, pa, pb, fm, offs, np, keys, vals);
  GSJ_LAUNCHED();
  // piece-pair order, then stable parent-pair order
  radix_sort_pairs(c, keys, keys2, vals, vals2, m, 0, 64);
  hipLaunchKernelGGL(k_parent_keys, dim3(div_up(m, 256)), dim3(256), 0, st, P, R, pa, pb, vals, m, keys);
  GSJ_LAUNCHED();
  radix_sort_pairs(c, keys, keys2, vals, vals2, m, 0, 32 + bits_for(std::max<uint64_t>(PL.nparent, 1)));
  uint32_t* head = c.s_seg.as<uint32_t>(3 * m + 16);
  uint32_t* keep = head + m;
  uint64_t* hoff = c.s_tmp2.as<uint64_t>(2 * m + 16);
  hipLaunchKernelGGL(k_heads, dim3(div_up(m, 256)), dim3(256), 0, st, keys, m, head);
  GSJ_LAUNCHED();
  const uint64_t nseg = scan_u32_to_u64(c, head, hoff, m);
  uint64_t* start = hoff + m;
  hipLaunchKernelGGL(k_seg_starts, dim3(div_up(m, 256)), dim3(256), 0, st, head, hoff, m, start);
  GSJ_LAUNCHED();
  double* sarea = c.s_out_flag.as<double>(nseg * 6 + 16);
  double4* sev = reinterpret_cast<double4*>(sarea + ((nseg + 3) & ~uint64_t(3)));
  uint32_t* sevb = reinterpret_cast<uint32_t*>(sev + nseg);
  const bool rel = rp.want_relate != 0;
  hipLaunchKernelGGL(k_seg_fold, dim3(div_up(nseg, 256)), dim3(256), 0, st, keys, vals, start, nseg, m, ro.isect, rp.want_area ? ro.area : nullptr,
                     rel ? ro.ev : nullptr, rel ? ro.evb : nullptr, keep, sarea, sev, sevb);
  GSJ_LAUNCHED();
  uint64_t* koff = c.s_offs.as<uint64_t>(std::max<uint64_t>(np, nseg) + 16);   // offs no longer needed
  const uint64_t nkeep = scan_u32_to_u64(c, keep, koff, nseg);
  hipLaunchKernelGGL(k_emit_multi, dim3(div_up(nseg, 256)), dim3(256), 0, st, P, R, keys, start, keep, koff, nseg, nsingle, sarea, sev, sevb, orow,
                     oarea, rel ? ocode : nullptr, rp.want_area);
  GSJ_LAUNCHED();
  return nsingle + nkeep;
}

// Sub-batches of at most this many candidates whose rows need no piece -> parent folding compact and copy their rows
// with one host wait (compact_copy_small)
constexpr uint64_t SMALL_COPY_PAIRS = 1u << 20;

// compact_rows and the row copy of such a sub-batch: the copy takes np rows (the candidate count, known on the host, is
// an upper bound of the row count) together with the device's row count, so the count is not read back first: one
// host wait instead of two. Returns the row count; the rows are in the context's pinned buffers.
static uint64_t compact_copy_small(Context& c, const Layer& PL, const Layer& RL, const uint32_t* pa, const uint32_t* pb, uint64_t np,
                                   const RefineParams& rp, const RefineOut& ro, bool no_agg, uint32_t* orow, double* oarea, uint32_t* ocode,
                                   uint32_t*& hrow, double*& harea, uint32_t*& hcode) {
  hipStream_t st = c.stream;
  const DPolys& P = PL.poly;
  const DPolys& R = RL.poly;
  uint32_t* fs = c.s_tmp.as<uint32_t>(2 * np + 16);
  uint32_t* fm = fs + np;
  uint64_t* offs = c.s_offs.as<uint64_t>(np + 16);
  uint64_t* d_total = c.s_misc.as<uint64_t>(4);   // the scan primitives' scalar
  hipLaunchKernelGGL(k_row_flags, dim3(div_up(np, 256)), dim3(256), 0, st, P, R, pa, pb, ro.isect, np, no_agg ? 1 : 0, fs, fm);
  GSJ_LAUNCHED();
  scan_u32_to_u64_async(c, fs, offs, np, d_total);
  hipLaunchKernelGGL(k_emit_single, dim3(div_up(np, 256)), dim3(256), 0, st, P, R, pa, pb, fs, offs, np, no_agg ? 1 : 0, ro.area, ro.ev, ro.evb,
                     orow, rp.want_area ? oarea : nullptr, rp.want_relate ? ocode : nullptr);
  GSJ_LAUNCHED();
  uint64_t* h_total = c.h_scal.as<uint64_t>(4);
  hrow = c.h_pairs.as<uint32_t>(2 * np);
  harea = rp.want_area ? c.h_area.as<double>(np) : nullptr;
  hcode = rp.want_relate ? c.h_code.as<uint32_t>(np) : nullptr;
  GSJ_HIP(hipMemcpyAsync(h_total, d_total, 8, hipMemcpyDeviceToHost, st));
  GSJ_HIP(hipMemcpyAsync(hrow, orow, np * 8, hipMemcpyDeviceToHost, st));
  if (harea) GSJ_HIP(hipMemcpyAsync(harea, oarea, np * 8, hipMemcpyDeviceToHost, st));
  if (hcode) GSJ_HIP(hipMemcpyAsync(hcode, ocode, np * 4, hipMemcpyDeviceToHost, st));
  GSJ_HIP(hipStreamSynchronize(st));
  return h_total[0];
}

void join_run(Context& c, const Layer& PL, const Layer& RL, const gsj_join_opts& o, gsj_rows_cb cb, void* user, gsj_join_stats* st,
              const JoinHooks* hooks) {
  GSJ_REQUIRE(!PL.points && !RL.points, GSJ_E_ARG, "join: both layers must be polygon layers");
  GSJ_REQUIRE(RL.has_bvh, GSJ_E_ARG, "join: build layer has no BVH (call gsj_bvh_build)");
  gsj_join_stats s{};
  if (PL.n == 0 || RL.n == 0) { if (st) *st = s; return; }
  hipStream_t stream = c.stream;
  RefineParams rp;
  rp.want_area = o.want_area ? 1 : 0;
  rp.want_relate = o.want_relate ? 1 : 0;
  rp.snap_half_ulp = o.snap_half_ulp ? 1 : 0;
  const MembersFn none;
  for_each_candidate_batch(c, probes_of(PL), RL.bvh, o.upper != 0, o, s, [&](uint32_t* pa, uint32_t* pb, uint64_t np) {
    if (hooks && hooks->prefilter) {
      WallTimer tp;
      np = hooks->prefilter(pa, pb, np);
      s.compact_ms += tp.ms();
      if (np == 0) return;
    }
    WallTimer t2;
    RefineOut ro;
    ro.isect = c.s_flag.as<uint8_t>(np + 16);
    ro.area = rp.want_area ? c.s_area.as<double>(np + 16) : nullptr;
    ro.ev = rp.want_relate ? c.s_ev.as<double4>(np + 16) : nullptr;
    ro.evb = rp.want_relate ? c.s_evb.as<uint32_t>(np + 16) : nullptr;
    // no host waits between refinement, compaction and the row copy beyond the ones that read sizes back: the
    // stream orders the work (refine_ms and compact_ms are host times up to those reads)
    s.heavy += refine_pairs(c, PL.poly, RL.poly, pa, pb, np, rp, ro, &s);
    s.refine_ms += t2.ms();
    WallTimer t3;
    uint32_t* orow = c.s_out_a.as<uint32_t>(2 * np + 16);
    double* oarea = c.s_out_area.as<double>(np + 16);
    uint32_t* ocode = c.s_out_code.as<uint32_t>(np + 16);
    if (!hooks && np <= SMALL_COPY_PAIRS && (o.no_agg || (!PL.poly.pfirst && !RL.poly.pfirst))) {
      uint32_t* hrow = nullptr;
      double* harea = nullptr;
      uint32_t* hcode = nullptr;
      const uint64_t nrows = compact_copy_small(c, PL, RL, pa, pb, np, rp, ro, o.no_agg != 0, orow, oarea, ocode, hrow, harea, hcode);
      s.compact_ms += t3.ms();
      s.hits += nrows;
      if (nrows && cb && cb(user, hrow, harea, hcode, (size_t)nrows)) throw Error(GSJ_E_CANCEL, "join cancelled by the row callback");
      return;
    }
    const uint64_t nrows = compact_rows(c, PL, RL, pa, pb, np, rp, ro, o.no_agg != 0, orow, oarea, ocode);
    s.compact_ms += t3.ms();
    s.hits += nrows;
    if (hooks && hooks->on_device && nrows) hooks->on_device(orow, rp.want_area ? oarea : nullptr, nrows);
    if ((hooks && hooks->device_only) || nrows == 0) return;
    WallTimer t4;
    uint32_t* hrow = c.h_pairs.as<uint32_t>(2 * nrows);
    double* harea = rp.want_area ? c.h_area.as<double>(nrows) : nullptr;
    uint32_t* hcode = rp.want_relate ? c.h_code.as<uint32_t>(nrows) : nullptr;
    GSJ_HIP(hipMemcpyAsync(hrow, orow, nrows * 8, hipMemcpyDeviceToHost, stream));
    if (harea) GSJ_HIP(hipMemcpyAsync(harea, oarea, nrows * 8, hipMemcpyDeviceToHost, stream));
    if (hcode) GSJ_HIP(hipMemcpyAsync(hcode, ocode, nrows * 4, hipMemcpyDeviceToHost, stream));
    GSJ_HIP(hipStreamSynchronize(stream));
    s.copy_ms += t4.ms();
    if (cb && cb(user, hrow, harea, hcode, (size_t)nrows)) throw Error(GSJ_E_CANCEL, "join cancelled by the row callback");
  }, hooks ? hooks->on_members : none);
  if (st) *st = s;
}

uint64_t compact_single_rows(Context& c, const Layer& PL, const Layer& RL, const uint32_t* pa, const uint32_t* pb, uint64_t np, const RefineParams& rp,
                             const RefineOut& ro, uint32_t* orow, double* oarea, const uint64_t** offs) {
  GSJ_REQUIRE(!PL.poly.pfirst && !RL.poly.pfirst, GSJ_E_INTERNAL, "candidate-ordered rows need layers without parents");
  return compact_rows(c, PL, RL, pa, pb, np, rp, ro, false, orow, oarea, nullptr, offs);
}

bool join_batchable(const Layer& P, const Layer& R, const gsj_join_opts& o) {
  // rows map one to one onto candidates (no piece -> parent folding), so a request's rows are a contiguous run
  return o.has_window && !o.cancel && !P.points && !R.points && R.has_bvh && P.n && R.n && probes_batchable(probes_of(P)) &&
         (o.no_agg || (!P.poly.pfirst && !R.poly.pfirst));
}

bool join_compatible(const gsj_join_opts& a, const gsj_join_opts& b) {
  return a.want_area == b.want_area && a.want_relate == b.want_relate && a.upper == b.upper && a.snap_half_ulp == b.snap_half_ulp &&
         a.no_agg == b.no_agg && a.batch_q == b.batch_q && a.max_pairs == b.max_pairs;
}

void join_batch_run(Context& c, const Layer& PL, const Layer& RL, std::vector<BatchJoin>& reqs) {
  const uint32_t k = (uint32_t)reqs.size();
  GSJ_REQUIRE(k >= 1 && k <= MAX_WINDOWS, GSJ_E_ARG, "batch size out of range");
  const gsj_join_opts& o = *reqs[0].o;
  WindowSet ws;
  ws.k = k;
  for (uint32_t r = 0; r < k; r++) {
    GSJ_REQUIRE(join_batchable(PL, RL, *reqs[r].o) && join_compatible(o, *reqs[r].o), GSJ_E_ARG, "request cannot join this batch");
    for (int j = 0; j < 4; j++) ws.w[r][j] = reqs[r].o->window[j];
  }
  hipStream_t stream = c.stream;
  RefineParams rp;
  rp.want_area = o.want_area ? 1 : 0;
  rp.want_relate = o.want_relate ? 1 : 0;
  rp.snap_half_ulp = o.snap_half_ulp ? 1 : 0;
  gsj_join_stats s{};
  std::vector<uint64_t> cand(k, 0), hits(k, 0);
  for_each_candidate_multi(c, probes_of(PL), RL.bvh, o.upper != 0, o, ws, s, [&](uint32_t* pa, uint32_t* pb, uint64_t np, const SegBounds& sb) {
    for (uint32_t r = 0; r < k; r++) cand[r] += sb.h[r + 1] - sb.h[r];
    WallTimer t2;
    RefineOut ro;
    ro.isect = c.s_flag.as<uint8_t>(np + 16);
    ro.area = rp.want_area ? c.s_area.as<double>(np + 16) : nullptr;
    ro.ev = rp.want_relate ? c.s_ev.as<double4>(np + 16) : nullptr;
    ro.evb = rp.want_relate ? c.s_evb.as<uint32_t>(np + 16) : nullptr;
    s.heavy += refine_pairs(c, PL.poly, RL.poly, pa, pb, np, rp, ro, &s);
    s.refine_ms += t2.ms();
    WallTimer t3;
    uint32_t* orow = c.s_out_a.as<uint32_t>(2 * np + 16);
    double* oarea = c.s_out_area.as<double>(np + 16);
    uint32_t* ocode = c.s_out_code.as<uint32_t>(np + 16);
    const uint64_t* offs = nullptr;
    const uint64_t nrows = compact_rows(c, PL, RL, pa, pb, np, rp, ro, o.no_agg != 0, orow, oarea, ocode, &offs);
    std::vector<uint64_t> rlo(k + 1);
    seg_rows(c, sb, offs, np, nrows, rlo.data());
    s.compact_ms += t3.ms();
    if (nrows == 0) return;
    WallTimer t4;
    uint32_t* hrow = c.h_pairs.as<uint32_t>(2 * nrows);
    double* harea = rp.want_area ? c.h_area.as<double>(nrows) : nullptr;
    uint32_t* hcode = rp.want_relate ? c.h_code.as<uint32_t>(nrows) : nullptr;
    GSJ_HIP(hipMemcpyAsync(hrow, orow, nrows * 8, hipMemcpyDeviceToHost, stream));
    if (harea) GSJ_HIP(hipMemcpyAsync(harea, oarea, nrows * 8, hipMemcpyDeviceToHost, stream));
    if (hcode) GSJ_HIP(hipMemcpyAsync(hcode, ocode, nrows * 4, hipMemcpyDeviceToHost, stream));
    GSJ_HIP(hipStreamSynchronize(stream));
    s.copy_ms += t4.ms();
    for (uint32_t r = 0; r < k; r++) {
      const uint64_t a = rlo[r], n = rlo[r + 1] - rlo[r];
      hits[r] += n;
      BatchJoin& q = reqs[r];
      if (n == 0 || !q.cb || q.rc != GSJ_OK) continue;
      if (q.cb(q.user, hrow + 2 * a, harea ? harea + a : nullptr, hcode ? hcode + a : nullptr, (size_t)n)) q.rc = GSJ_E_CANCEL;
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
