// Layer preparation on the device: exact identical-geometry dedup (the restatement of dedup_set, engine/dedup.h:51-124,
// that hostprep.cpp also implements on the host; both give the same distinct set, member lists and order).
//
// A thread per geometry hashes its ring structure and coordinate bits (the host's 64-bit content hash); a stable
// radix sort by hash puts each hash group in index order; a thread per group finds every member's representative
// (the first earlier member with the same rings and bit-identical coordinates); a scan over the representatives in
// index order numbers the distinct geometries by first occurrence; a stable sort by distinct number gives the member
// lists in index order; the distinct geometries' rings and vertices are compacted by two more scans.
#include "engine.h"
#include "hostprep.h"
#include <cstring>
#include <memory>

namespace gsj {

namespace {

__global__ void k_dd_hash(const uint64_t* __restrict__ gro, const uint64_t* __restrict__ rvo, const double* __restrict__ xy, uint64_t n,
                          uint64_t* __restrict__ h, uint32_t* __restrict__ idx) {
  const uint64_t g = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= n) return;
  uint64_t v = 0x9E3779B97F4A7C15ull ^ (gro[g + 1] - gro[g]);
  for (uint64_t r = gro[g]; r < gro[g + 1]; r++) {
    v = (v ^ (rvo[r + 1] - rvo[r])) * 0xBF58476D1CE4E5B9ull;
    const unsigned long long* w = reinterpret_cast<const unsigned long long*>(xy);
    for (uint64_t k = 2 * rvo[r]; k < 2 * rvo[r + 1]; k++) {
      v = (v ^ w[k]) * 0x94D049BB133111EBull;
      v ^= v >> 31;
    }
  }
  h[g] = v;
  idx[g] = (uint32_t)g;
}

__device__ bool dd_same(const uint64_t* gro, const uint64_t* rvo, const double* xy, uint64_t a, uint64_t b) {
  const uint64_t nra = gro[a + 1] - gro[a];
  if (nra != gro[b + 1] - gro[b]) return false;
  for (uint64_t k = 0; k < nra; k++)
    if (rvo[gro[a] + k + 1] - rvo[gro[a] + k] != rvo[gro[b] + k + 1] - rvo[gro[b] + k]) return false;
  const unsigned long long* w = reinterpret_cast<const unsigned long long*>(xy);
  const uint64_t va = 2 * rvo[gro[a]], vb = 2 * rvo[gro[b]], nw = 2 * (rvo[gro[a + 1]] - rvo[gro[a]]);
  for (uint64_t k = 0; k < nw; k++)
    if (w[va + k] != w[vb + k]) return false;   // bit-identical coordinates
  return true;
}

__global__ void k_dd_heads(const uint64_t* __restrict__ h, uint64_t n, uint32_t* __restrict__ head) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) head[i] = (i == 0 || h[i] != h[i - 1]) ? 1u : 0u;
}
__global__ void k_dd_starts(const uint32_t* __restrict__ head, const uint64_t* __restrict__ off, uint64_t n, uint64_t* __restrict__ start) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n && head[i]) start[off[i]] = i;
}

// a thread per hash group (sorted positions start[s] .. start[s+1]), members in index order: rep = the first earlier
// representative with the same geometry, else itself
__global__ void k_dd_group(const uint64_t* __restrict__ gro, const uint64_t* __restrict__ rvo, const double* __restrict__ xy,
                           const uint32_t* __restrict__ ord, const uint64_t* __restrict__ start, uint64_t ngroups, uint64_t n, uint32_t* __restrict__ rep) {
  const uint64_t s = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= ngroups) return;
  const uint64_t i0 = start[s], i1 = s + 1 < ngroups ? start[s + 1] : n;
  for (uint64_t k = i0; k < i1; k++) {
    const uint32_t g = ord[k];
    uint32_t r = g;
    for (uint64_t q = i0; q < k; q++) {
      const uint32_t o = ord[q];
      if (rep[o] == o && dd_same(gro, rvo, xy, o, g)) { r = o; break; }
    }
    rep[g] = r;
  }
}

__global__ void k_dd_isrep(const uint32_t* __restrict__ rep, uint64_t n, uint32_t* __restrict__ f) {
  const uint64_t g = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g < n) f[g] = rep[g] == (uint32_t)g ? 1u : 0u;
}
// first[u] = the representative numbered u; key[g] = the number of g's representative (member list key)
__global__ void k_dd_number(const uint32_t* __restrict__ rep, const uint64_t* __restrict__ uid, uint64_t n, uint32_t* __restrict__ first,
                            uint64_t* __restrict__ key, uint32_t* __restrict__ val) {
  const uint64_t g = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= n) return;
  if (rep[g] == (uint32_t)g) first[uid[g]] = (uint32_t)g;
  key[g] = uid[rep[g]];
  val[g] = (uint32_t)g;
}
// member offsets: moff[u] = first sorted position with key u (keys 0 .. un-1 all present), moff[un] = n
__global__ void k_dd_moff(const uint64_t* __restrict__ key, uint64_t n, uint64_t un, uint64_t* __restrict__ moff) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n && (i == 0 || key[i] != key[i - 1])) moff[key[i]] = i;
  if (i == 0) moff[un] = n;
}

// distinct geometries' rings: counts, then per output ring its source ring and vertex count
__global__ void k_dd_ring_counts(const uint64_t* __restrict__ gro, const uint32_t* __restrict__ first, uint64_t un, uint32_t* __restrict__ nring) {
  const uint64_t u = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (u < un) nring[u] = (uint32_t)(gro[first[u] + 1] - gro[first[u]]);
}
__global__ void k_dd_rings(const uint64_t* __restrict__ gro, const uint64_t* __restrict__ rvo, const uint32_t* __restrict__ first,
                           const uint64_t* __restrict__ ugro, uint64_t un, uint64_t* __restrict__ src, uint32_t* __restrict__ nvert) {
  const uint64_t u = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (u >= un) return;
  const uint64_t g = first[u], o = ugro[u];
  for (uint64_t r = gro[g]; r < gro[g + 1]; r++) {
    src[o + (r - gro[g])] = r;
    nvert[o + (r - gro[g])] = (uint32_t)(rvo[r + 1] - rvo[r]);
  }
}
__global__ void k_dd_copy(const uint64_t* __restrict__ rvo, const double2* __restrict__ xy, const uint64_t* __restrict__ src,
                          const uint64_t* __restrict__ urvo, uint64_t nur, double2* __restrict__ uxy) {
  const uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nur) return;
  const uint64_t r = src[k], a = rvo[r], m = rvo[r + 1] - a, o = urvo[k];
  for (uint64_t j = 0; j < m; j++) uxy[o + j] = xy[a + j];
}

template <class T>
T* dev_copy(DevBuf& b, const T* h, uint64_t count, hipStream_t st) {
  T* d = b.as<T>(count + 1);
  if (count) GSJ_HIP(hipMemcpyAsync(d, h, count * sizeof(T), hipMemcpyHostToDevice, st));
  return d;
}

}  // namespace

HostSet dedup_device(Context& c, const uint64_t* gro, const uint64_t* rvo, const double* xy, uint64_t n, std::vector<uint64_t>& moff,
                     std::vector<uint32_t>& midx) {
  GSJ_REQUIRE(n < (1ull << 32), GSJ_E_LIMIT, "dedup: more than 2^32 geometries");
  hipStream_t st = c.stream;
  HostSet u;
  u.gro.push_back(0);
  u.rvo.push_back(0);
  moff.assign(1, 0);
  midx.clear();
  if (n == 0) return u;
  const uint64_t nr = gro[n], nv = rvo[nr];
  DevBuf b_gro, b_rvo, b_xy, b_h, b_h2, b_idx, b_idx2, b_head, b_off, b_start, b_rep, b_uid, b_first, b_moff, b_nring, b_ugro, b_src, b_nvert,
      b_urvo, b_uxy;
  const uint64_t* dgro = dev_copy(b_gro, gro, n + 1, st);
  const uint64_t* drvo = dev_copy(b_rvo, rvo, nr + 1, st);
  const double* dxy = dev_copy(b_xy, xy, 2 * nv, st);
  uint64_t* h = b_h.as<uint64_t>(n);
  uint64_t* h2 = b_h2.as<uint64_t>(n);
  uint32_t* ord = b_idx.as<uint32_t>(n);
  uint32_t* ord2 = b_idx2.as<uint32_t>(n);
  hipLaunchKernelGGL(k_dd_hash, dim3(div_up(n, 256)), dim3(256), 0, st, dgro, drvo, dxy, n, h, ord);
  GSJ_LAUNCHED();
  radix_sort_pairs(c, h, h2, ord, ord2, n, 0, 64);   // stable: a hash group in index order
  uint32_t* head = b_head.as<uint32_t>(n + 16);
  uint64_t* off = b_off.as<uint64_t>(n + 16);
  hipLaunchKernelGGL(k_dd_heads, dim3(div_up(n, 256)), dim3(256), 0, st, h, n, head);
  GSJ_LAUNCHED();
  const uint64_t ngroups = scan_u32_to_u64(c, head, off, n);
  uint64_t* start = b_start.as<uint64_t>(ngroups + 1);
  hipLaunchKernelGGL(k_dd_starts, dim3(div_up(n, 256)), dim3(256), 0, st, head, off, n, start);
  GSJ_LAUNCHED();
  uint32_t* rep = b_rep.as<uint32_t>(n);
  hipLaunchKernelGGL(k_dd_group, dim3(div_up(ngroups, 256)), dim3(256), 0, st, dgro, drvo, dxy, ord, start, ngroups, n, rep);
  GSJ_LAUNCHED();
  uint32_t* isrep = head;   // reused
  uint64_t* uid = b_uid.as<uint64_t>(n + 16);
  hipLaunchKernelGGL(k_dd_isrep, dim3(div_up(n, 256)), dim3(256), 0, st, rep, n, isrep);
  GSJ_LAUNCHED();
  const uint64_t un = scan_u32_to_u64(c, isrep, uid, n);
  uint32_t* first = b_first.as<uint32_t>(un);
  uint64_t* key = h;     // reused: member-list keys and the sorted members
  uint64_t* key2 = h2;
  uint32_t* val = ord;
  uint32_t* val2 = ord2;
  hipLaunchKernelGGL(k_dd_number, dim3(div_up(n, 256)), dim3(256), 0, st, rep, uid, n, first, key, val);
  GSJ_LAUNCHED();
  radix_sort_pairs(c, key, key2, val, val2, n, 0, 32);   // stable: members in index order
  uint64_t* dmoff = b_moff.as<uint64_t>(un + 1);
  hipLaunchKernelGGL(k_dd_moff, dim3(div_up(n, 256)), dim3(256), 0, st, key, n, un, dmoff);
  GSJ_LAUNCHED();
  // the distinct geometries
  uint32_t* nring = b_nring.as<uint32_t>(un + 16);
  uint64_t* ugro = b_ugro.as<uint64_t>(un + 16);
  hipLaunchKernelGGL(k_dd_ring_counts, dim3(div_up(un, 256)), dim3(256), 0, st, dgro, first, un, nring);
  GSJ_LAUNCHED();
  const uint64_t nur = scan_u32_to_u64(c, nring, ugro, un);
  uint64_t* src = b_src.as<uint64_t>(nur + 1);
  uint32_t* nvert = b_nvert.as<uint32_t>(nur + 16);
  uint64_t* urvo = b_urvo.as<uint64_t>(nur + 16);
  hipLaunchKernelGGL(k_dd_rings, dim3(div_up(un, 256)), dim3(256), 0, st, dgro, drvo, first, ugro, un, src, nvert);
  GSJ_LAUNCHED();
  const uint64_t unv = scan_u32_to_u64(c, nvert, urvo, nur);
  double2* uxy = b_uxy.as<double2>(unv + 1);
  hipLaunchKernelGGL(k_dd_copy, dim3(div_up(nur, 256)), dim3(256), 0, st, drvo, reinterpret_cast<const double2*>(dxy), src, urvo, nur, uxy);
  GSJ_LAUNCHED();
  u.gro.resize(un + 1);
  u.rvo.resize(nur + 1);
  u.xy.resize(2 * unv);
  moff.resize(un + 1);
  midx.resize(n);
  GSJ_HIP(hipMemcpyAsync(u.gro.data(), ugro, un * 8, hipMemcpyDeviceToHost, st));
  GSJ_HIP(hipMemcpyAsync(u.rvo.data(), urvo, nur * 8, hipMemcpyDeviceToHost, st));
  if (unv) GSJ_HIP(hipMemcpyAsync(u.xy.data(), uxy, unv * 16, hipMemcpyDeviceToHost, st));
  GSJ_HIP(hipMemcpyAsync(moff.data(), dmoff, (un + 1) * 8, hipMemcpyDeviceToHost, st));
  GSJ_HIP(hipMemcpyAsync(midx.data(), val, n * 4, hipMemcpyDeviceToHost, st));
  GSJ_HIP(hipStreamSynchronize(st));
  u.gro[un] = nur;
  u.rvo[nur] = unv;
  return u;
}

}  // namespace gsj

// ------------------------------------------------------------------------------------------- subdivision
// ST_Subdivide analogue on the device (restating subdivide_set, engine/subdivide.h:14-129, as hostprep.cpp does on the
// host; the same pieces, bit for bit, in the same order). Level-synchronous: every piece of the current level (level
// 0: the input geometries) is either final (at most maxv stored vertices, depth 24, or a point-sized bbox) or cut at
// the middle of its bbox's longer side; each of its rings is clipped against both halves by a thread (a counting
// pass, a scan, and a writing pass that repeats the same arithmetic), and the halves with rings become the next
// level's pieces. A piece's path (lower half 0, upper 1, one bit per level from the top) orders the final pieces of a
// geometry depth-first, lower half first, as the host's stack does.
namespace gsj {

namespace {

constexpr int SD_DEPTH = 24;

struct SdLevel {   // pieces and rings of one level; the vertex pool of level 0 is the input itself
  uint32_t nitems = 0;
  uint64_t nrings = 0;
  const uint32_t* it_g = nullptr;
  const uint32_t* it_path = nullptr;
  const uint32_t* it_r0 = nullptr;
  const uint32_t* it_nr = nullptr;
  const uint64_t* rg_off = nullptr;   // first vertex of each ring in the pool
  const uint32_t* rg_n = nullptr;     // stored vertices of each ring
  const double2* pool = nullptr;
  const uint8_t* syn = nullptr;       // null: no edge lies on a cut
  int depth = 0;
};

// Sutherland-Hodgman clip of one ring against the closed half-plane (axis coordinate <= m for the lower half, >= m for
// the upper), with the host's repeated-point merge, closing, and drop of rings with < 4 stored vertices or zero area.
// WRITE false returns the output's vertex count (0: dropped); WRITE true also stores it at out / osyn.
template <bool WRITE>
__device__ uint32_t sd_clip(const double2* __restrict__ in, const uint8_t* __restrict__ isyn, uint32_t nverts, int axis, double m, bool lower,
                            double2* __restrict__ out, uint8_t* __restrict__ osyn) {
  uint32_t k = 0;
  double fx = 0, fy = 0, lx = 0, ly = 0, a2 = 0;
  auto emit = [&](double x, double y, uint8_t s) {
    if (k >= 1 && lx == x && ly == y) {   // repeated point: the later edge flag wins
      if (WRITE) osyn[k - 1] = s;
      return;
    }
    if (k == 0) { fx = x; fy = y; }
    else a2 += lx * y - x * ly;
    if (WRITE) { out[k] = make_double2(x, y); osyn[k] = s; }
    lx = x; ly = y;
    k++;
  };
  for (uint32_t i = 0; i + 1 < nverts; i++) {
    const double2 a = in[i], b = in[i + 1];
    const double va = axis == 0 ? a.x : a.y, vb = axis == 0 ? b.x : b.y;
    const bool ka = lower ? va <= m : va >= m, kb = lower ? vb <= m : vb >= m;
    const uint8_t s = isyn ? isyn[i] : 0;
    double cx, cy;
    if (axis == 0) { cx = m; cy = a.y + (m - a.x) * (b.y - a.y) / (b.x - a.x); }
    else { cy = m; cx = a.x + (m - a.y) * (b.x - a.x) / (b.y - a.y); }
    if (ka) emit(a.x, a.y, s);
    if (ka && !kb) emit(cx, cy, 1);         // leaving: the next edge runs along the cut
    else if (!ka && kb) emit(cx, cy, s);    // entering
  }
  if (k < 3) return 0;
  if (fx != lx || fy != ly) {
    a2 += lx * fy - fx * ly;
    if (WRITE) { out[k] = make_double2(fx, fy); osyn[k] = 0; }
    k++;
  } else if (WRITE) {
    osyn[k - 1] = 0;
  }
  if (k < 4 || a2 == 0.0) return 0;
  return k;
}

// per piece: final (0) or cut along axis (1: x, 2: y) at m
__global__ void k_sd_decide(SdLevel L, int maxv, uint8_t* __restrict__ cut, double* __restrict__ mid, uint32_t* __restrict__ fin) {
  const uint32_t it = blockIdx.x * blockDim.x + threadIdx.x;
  if (it >= L.nitems) return;
  uint64_t nv = 0;
  double x0 = 1e300, y0 = 1e300, x1 = -1e300, y1 = -1e300;
  for (uint32_t r = L.it_r0[it]; r < L.it_r0[it] + L.it_nr[it]; r++) {
    nv += L.rg_n[r];
    for (uint64_t v = L.rg_off[r]; v < L.rg_off[r] + L.rg_n[r]; v++) {
      const double2 p = L.pool[v];
      x0 = fmin(x0, p.x); x1 = fmax(x1, p.x); y0 = fmin(y0, p.y); y1 = fmax(y1, p.y);
    }
  }
  const bool final_ = nv <= (uint64_t)maxv || L.depth >= SD_DEPTH || (x1 - x0 <= 0 && y1 - y0 <= 0);
  fin[it] = final_ ? 1u : 0u;
  cut[it] = final_ ? 0 : ((x1 - x0) >= (y1 - y0) ? 1 : 2);
  mid[it] = final_ ? 0.0 : ((x1 - x0) >= (y1 - y0) ? 0.5 * (x0 + x1) : 0.5 * (y0 + y1));
}

// ring slots of the cut pieces, grouped by child: piece it's rings [r0, r0 + nr) go to slots 2 r0 + side nr + (r - r0)
__global__ void k_sd_count(SdLevel L, const uint32_t* __restrict__ ring_it, const uint8_t* __restrict__ cut, const double* __restrict__ mid,
                           uint32_t* __restrict__ cnt) {
  const uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= 2 * L.nrings) return;
  const uint64_t r = j >> 1;
  const int side = (int)(j & 1);
  const uint32_t it = ring_it[r];
  const uint64_t slot = 2 * (uint64_t)L.it_r0[it] + (uint64_t)side * L.it_nr[it] + (r - L.it_r0[it]);
  uint32_t k = 0;
  if (cut[it]) k = sd_clip<false>(L.pool + L.rg_off[r], L.syn ? L.syn + L.rg_off[r] : nullptr, L.rg_n[r], cut[it] - 1, mid[it], side == 0, nullptr, nullptr);
  cnt[slot] = k;
}
__global__ void k_sd_write(SdLevel L, const uint32_t* __restrict__ ring_it, const uint8_t* __restrict__ cut, const double* __restrict__ mid,
                           const uint32_t* __restrict__ cnt, const uint64_t* __restrict__ voff, double2* __restrict__ pool, uint8_t* __restrict__ syn) {
  const uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= 2 * L.nrings) return;
  const uint64_t r = j >> 1;
  const int side = (int)(j & 1);
  const uint32_t it = ring_it[r];
  const uint64_t slot = 2 * (uint64_t)L.it_r0[it] + (uint64_t)side * L.it_nr[it] + (r - L.it_r0[it]);
  if (cnt[slot] == 0) return;
  const uint64_t o = voff[slot];
  sd_clip<true>(L.pool + L.rg_off[r], L.syn ? L.syn + L.rg_off[r] : nullptr, L.rg_n[r], cut[it] - 1, mid[it], side == 0, pool + o, syn + o);
}
__global__ void k_sd_ring_items(SdLevel L, uint32_t* __restrict__ ring_it) {
  const uint32_t it = blockIdx.x * blockDim.x + threadIdx.x;
  if (it >= L.nitems) return;
  for (uint32_t r = L.it_r0[it]; r < L.it_r0[it] + L.it_nr[it]; r++) ring_it[r] = it;
}
__global__ void k_nonzero(const uint32_t* __restrict__ cnt, uint64_t n, uint32_t* __restrict__ f) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) f[i] = cnt[i] ? 1u : 0u;
}
// children: child c = 2 it + side has the non-empty slots of its slot range; ring index of its first one
__global__ void k_sd_children(SdLevel L, const uint8_t* __restrict__ cut, const uint32_t* __restrict__ cnt, const uint64_t* __restrict__ rix,
                              uint32_t* __restrict__ cnr, uint32_t* __restrict__ cr0) {
  const uint64_t c = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= 2 * (uint64_t)L.nitems) return;
  const uint32_t it = (uint32_t)(c >> 1), side = (uint32_t)(c & 1);
  uint32_t m = 0;
  if (cut[it]) {
    const uint64_t s0 = 2 * (uint64_t)L.it_r0[it] + (uint64_t)side * L.it_nr[it];
    for (uint32_t q = 0; q < L.it_nr[it]; q++) m += cnt[s0 + q] ? 1u : 0u;
    cr0[c] = (uint32_t)rix[s0];   // rings of a child are consecutive non-empty slots
  }
  cnr[c] = m;
}
__global__ void k_sd_next(SdLevel L, const uint32_t* __restrict__ cnr, const uint32_t* __restrict__ cr0, const uint64_t* __restrict__ cix,
                          uint32_t* __restrict__ g, uint32_t* __restrict__ path, uint32_t* __restrict__ r0, uint32_t* __restrict__ nr) {
  const uint64_t c = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= 2 * (uint64_t)L.nitems || cnr[c] == 0) return;
  const uint32_t it = (uint32_t)(c >> 1), side = (uint32_t)(c & 1);
  const uint64_t o = cix[c];
  g[o] = L.it_g[it];
  path[o] = L.it_path[it] | (side << (SD_DEPTH - 1 - L.depth));
  r0[o] = cr0[c];
  nr[o] = cnr[c];
}
__global__ void k_sd_ring_arrays(uint64_t n, const uint32_t* __restrict__ cnt, const uint64_t* __restrict__ rix, const uint64_t* __restrict__ voff,
                                 uint64_t* __restrict__ off, uint32_t* __restrict__ rn) {
  const uint64_t s = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n || cnt[s] == 0) return;
  off[rix[s]] = voff[s];
  rn[rix[s]] = cnt[s];
}

// final pieces of a level, appended to the piece table: key (geometry, path), level, first ring, ring count
__global__ void k_sd_finals(SdLevel L, const uint32_t* __restrict__ fin, const uint64_t* __restrict__ foff, uint64_t base, uint32_t lvl,
                            uint64_t* __restrict__ key, uint32_t* __restrict__ plev, uint32_t* __restrict__ pr0, uint32_t* __restrict__ pnr) {
  const uint32_t it = blockIdx.x * blockDim.x + threadIdx.x;
  if (it >= L.nitems || !fin[it]) return;
  const uint64_t o = base + foff[it];
  key[o] = ((uint64_t)L.it_g[it] << SD_DEPTH) | L.it_path[it];
  plev[o] = lvl;
  pr0[o] = L.it_r0[it];
  pnr[o] = L.it_nr[it];
}

struct SdLevelPtrs {   // ring tables and pools of every level, for the final gather
  const uint64_t* rg_off[SD_DEPTH + 2];
  const uint32_t* rg_n[SD_DEPTH + 2];
  const double2* pool[SD_DEPTH + 2];
  const uint8_t* syn[SD_DEPTH + 2];
};
__global__ void k_sd_out_nr(const uint32_t* __restrict__ order, const uint32_t* __restrict__ pnr, uint64_t np, uint32_t* __restrict__ onr) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < np) onr[i] = pnr[order[i]];
}
__global__ void k_sd_out_rings(const uint32_t* __restrict__ order, const uint32_t* __restrict__ plev, const uint32_t* __restrict__ pr0,
                               const uint32_t* __restrict__ pnr, const uint64_t* __restrict__ ogro, uint64_t np, SdLevelPtrs lv,
                               uint32_t* __restrict__ rlev, uint32_t* __restrict__ rsrc, uint32_t* __restrict__ rn) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= np) return;
  const uint32_t p = order[i], l = plev[p];
  for (uint32_t q = 0; q < pnr[p]; q++) {
    const uint64_t o = ogro[i] + q;
    rlev[o] = l;
    rsrc[o] = pr0[p] + q;
    rn[o] = lv.rg_n[l][pr0[p] + q];
  }
}
__global__ void k_sd_out_copy(const uint32_t* __restrict__ rlev, const uint32_t* __restrict__ rsrc, const uint64_t* __restrict__ orvo, uint64_t nr,
                              SdLevelPtrs lv, double2* __restrict__ oxy, uint8_t* __restrict__ osyn) {
  const uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nr) return;
  const uint32_t l = rlev[k], r = rsrc[k];
  const uint64_t a = lv.rg_off[l][r], n = lv.rg_n[l][r], o = orvo[k];
  for (uint64_t j = 0; j < n; j++) {
    oxy[o + j] = lv.pool[l][a + j];
    osyn[o + j] = lv.syn[l] ? lv.syn[l][a + j] : 0;
  }
}
__global__ void k_sd_parent(const uint64_t* __restrict__ key, uint64_t np, uint32_t* __restrict__ parent) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < np) parent[i] = (uint32_t)(key[i] >> SD_DEPTH);
}
__global__ void k_iota_u32(uint32_t* p, uint64_t n) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = (uint32_t)i;
}
__global__ void k_sd_level0(const uint64_t* __restrict__ gro, const uint64_t* __restrict__ rvo, uint64_t n, uint64_t nr, uint32_t* __restrict__ g,
                            uint32_t* __restrict__ path, uint32_t* __restrict__ r0, uint32_t* __restrict__ rnr, uint64_t* __restrict__ off,
                            uint32_t* __restrict__ rn) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) { g[i] = (uint32_t)i; path[i] = 0; r0[i] = (uint32_t)gro[i]; rnr[i] = (uint32_t)(gro[i + 1] - gro[i]); }
  if (i < nr) { off[i] = rvo[i]; rn[i] = (uint32_t)(rvo[i + 1] - rvo[i]); }
}

struct SdBufs {   // device memory of one level
  DevBuf g, path, r0, nr, off, rn, pool, syn;
};

}  // namespace

HostSet subdivide_device(Context& c, const uint64_t* gro, const uint64_t* rvo, const double* xy, uint64_t n, int maxv) {
  GSJ_REQUIRE(maxv >= 4, GSJ_E_ARG, "subdivide: maxv must be >= 4");
  hipStream_t st = c.stream;
  HostSet o;
  o.has_parent = true;
  o.gro.push_back(0);
  o.rvo.push_back(0);
  if (n == 0) return o;
  const uint64_t nr0 = gro[n], nv0 = rvo[nr0];
  GSJ_REQUIRE(n < (1ull << 32) && nr0 < (1ull << 32), GSJ_E_LIMIT, "subdivide: more than 2^32 geometries or rings");
  DevBuf b_gro, b_rvo, b_xy;
  const uint64_t* dgro = dev_copy(b_gro, gro, n + 1, st);
  const uint64_t* drvo = dev_copy(b_rvo, rvo, nr0 + 1, st);
  const double2* dxy = reinterpret_cast<const double2*>(dev_copy(b_xy, xy, 2 * nv0, st));
  std::vector<std::unique_ptr<SdBufs>> bufs;
  std::vector<SdLevel> levels;
  // level 0: the input geometries
  bufs.push_back(std::make_unique<SdBufs>());
  {
    SdBufs& B = *bufs.back();
    SdLevel L;
    uint32_t* g = B.g.as<uint32_t>(n); uint32_t* path = B.path.as<uint32_t>(n); uint32_t* r0 = B.r0.as<uint32_t>(n); uint32_t* rnr = B.nr.as<uint32_t>(n);
    uint64_t* off = B.off.as<uint64_t>(nr0); uint32_t* rn = B.rn.as<uint32_t>(nr0);
    hipLaunchKernelGGL(k_sd_level0, dim3(div_up(std::max(n, nr0), 256)), dim3(256), 0, st, dgro, drvo, n, nr0, g, path, r0, rnr, off, rn);
    GSJ_LAUNCHED();
    L.nitems = (uint32_t)n; L.nrings = nr0; L.it_g = g; L.it_path = path; L.it_r0 = r0; L.it_nr = rnr; L.rg_off = off; L.rg_n = rn; L.pool = dxy;
    L.syn = nullptr; L.depth = 0;
    levels.push_back(L);
  }
  DevBuf b_cut, b_mid, b_fin, b_foff, b_ringit, b_cnt, b_f, b_rix, b_voff, b_cnr, b_cr0, b_cix;
  DevBuf b_key, b_key2, b_plev, b_pr0, b_pnr;
  uint64_t npieces = 0, cap = 0;
  uint64_t* key = nullptr; uint32_t* plev = nullptr; uint32_t* pr0 = nullptr; uint32_t* pnr = nullptr;
  for (int depth = 0;; depth++) {
    SdLevel L = levels.back();
    uint8_t* cut = b_cut.as<uint8_t>(L.nitems + 16);
    double* mid = b_mid.as<double>(L.nitems + 16);
    uint32_t* fin = b_fin.as<uint32_t>(L.nitems + 16);
    uint64_t* foff = b_foff.as<uint64_t>(L.nitems + 16);
    hipLaunchKernelGGL(k_sd_decide, dim3(div_up(L.nitems, 256)), dim3(256), 0, st, L, maxv, cut, mid, fin);
    GSJ_LAUNCHED();
    const uint64_t nfin = scan_u32_to_u64(c, fin, foff, L.nitems);
    if (npieces + nfin > cap) {   // grow the piece table (contents kept)
      const uint64_t ncap = std::max<uint64_t>(2 * cap, npieces + nfin + 1024);
      DevBuf nk, nl, n0, nn;
      uint64_t* k2 = nk.as<uint64_t>(ncap); uint32_t* l2 = nl.as<uint32_t>(ncap); uint32_t* r2 = n0.as<uint32_t>(ncap); uint32_t* c2 = nn.as<uint32_t>(ncap);
      if (npieces) {
        GSJ_HIP(hipMemcpyAsync(k2, key, npieces * 8, hipMemcpyDeviceToDevice, st));
        GSJ_HIP(hipMemcpyAsync(l2, plev, npieces * 4, hipMemcpyDeviceToDevice, st));
        GSJ_HIP(hipMemcpyAsync(r2, pr0, npieces * 4, hipMemcpyDeviceToDevice, st));
        GSJ_HIP(hipMemcpyAsync(c2, pnr, npieces * 4, hipMemcpyDeviceToDevice, st));
        GSJ_HIP(hipStreamSynchronize(st));
      }
      for (auto [a, b] : {std::pair<DevBuf*, DevBuf*>{&b_key, &nk}, {&b_plev, &nl}, {&b_pr0, &n0}, {&b_pnr, &nn}}) {
        std::swap(a->p, b->p);
        std::swap(a->cap, b->cap);
      }
      key = k2; plev = l2; pr0 = r2; pnr = c2; cap = ncap;
    }
    hipLaunchKernelGGL(k_sd_finals, dim3(div_up(L.nitems, 256)), dim3(256), 0, st, L, fin, foff, npieces, (uint32_t)depth, key, plev, pr0, pnr);
    GSJ_LAUNCHED();
    npieces += nfin;
    if (nfin == L.nitems) break;   // nothing is cut (depth 24 finalises every piece)
    // clip every ring of the cut pieces against both halves
    uint32_t* ring_it = b_ringit.as<uint32_t>(L.nrings + 16);
    hipLaunchKernelGGL(k_sd_ring_items, dim3(div_up(L.nitems, 256)), dim3(256), 0, st, L, ring_it);
    GSJ_LAUNCHED();
    const uint64_t nslots = 2 * L.nrings;
    uint32_t* cnt = b_cnt.as<uint32_t>(nslots + 16);
    hipLaunchKernelGGL(k_sd_count, dim3(div_up(nslots, 256)), dim3(256), 0, st, L, ring_it, cut, mid, cnt);
    GSJ_LAUNCHED();
    uint32_t* f = b_f.as<uint32_t>(nslots + 16);
    uint64_t* rix = b_rix.as<uint64_t>(nslots + 16);
    uint64_t* voff = b_voff.as<uint64_t>(nslots + 16);
    hipLaunchKernelGGL(k_nonzero, dim3(div_up(nslots, 256)), dim3(256), 0, st, cnt, nslots, f);
    GSJ_LAUNCHED();
    const uint64_t nrings = scan_u32_to_u64(c, f, rix, nslots);
    const uint64_t nverts = scan_u32_to_u64(c, cnt, voff, nslots);
    uint32_t* cnr = b_cnr.as<uint32_t>(2 * (uint64_t)L.nitems + 16);
    uint32_t* cr0 = b_cr0.as<uint32_t>(2 * (uint64_t)L.nitems + 16);
    uint64_t* cix = b_cix.as<uint64_t>(2 * (uint64_t)L.nitems + 16);
    hipLaunchKernelGGL(k_sd_children, dim3(div_up(2 * (uint64_t)L.nitems, 256)), dim3(256), 0, st, L, cut, cnt, rix, cnr, cr0);
    GSJ_LAUNCHED();
    hipLaunchKernelGGL(k_nonzero, dim3(div_up(2 * (uint64_t)L.nitems, 256)), dim3(256), 0, st, cnr, 2 * (uint64_t)L.nitems, f);
    GSJ_LAUNCHED();
    const uint64_t nchild = scan_u32_to_u64(c, f, cix, 2 * (uint64_t)L.nitems);
    if (nchild == 0) break;
    GSJ_REQUIRE(depth + 1 <= SD_DEPTH, GSJ_E_INTERNAL, "subdivide: deeper than 24 levels");
    bufs.push_back(std::make_unique<SdBufs>());
    SdBufs& B = *bufs.back();
    SdLevel N;
    uint32_t* g = B.g.as<uint32_t>(nchild); uint32_t* path = B.path.as<uint32_t>(nchild); uint32_t* r0 = B.r0.as<uint32_t>(nchild);
    uint32_t* nr = B.nr.as<uint32_t>(nchild);
    uint64_t* off = B.off.as<uint64_t>(nrings + 1); uint32_t* rn = B.rn.as<uint32_t>(nrings + 1);
    double2* pool = B.pool.as<double2>(nverts + 1); uint8_t* syn = B.syn.as<uint8_t>(nverts + 1);
    hipLaunchKernelGGL(k_sd_next, dim3(div_up(2 * (uint64_t)L.nitems, 256)), dim3(256), 0, st, L, cnr, cr0, cix, g, path, r0, nr);
    hipLaunchKernelGGL(k_sd_ring_arrays, dim3(div_up(nslots, 256)), dim3(256), 0, st, nslots, cnt, rix, voff, off, rn);
    hipLaunchKernelGGL(k_sd_write, dim3(div_up(nslots, 256)), dim3(256), 0, st, L, ring_it, cut, mid, cnt, voff, pool, syn);
    GSJ_LAUNCHED();
    N.nitems = (uint32_t)nchild; N.nrings = nrings; N.it_g = g; N.it_path = path; N.it_r0 = r0; N.it_nr = nr; N.rg_off = off; N.rg_n = rn;
    N.pool = pool; N.syn = syn; N.depth = depth + 1;
    levels.push_back(N);
  }
  if (npieces == 0) return o;   // every ring vanished in the cuts
  // the final pieces in (geometry, path) order: depth-first, lower half first
  DevBuf b_ord, b_ord2;
  uint32_t* ord = b_ord.as<uint32_t>(npieces + 1);
  uint32_t* ord2 = b_ord2.as<uint32_t>(npieces + 1);
  uint64_t* key2 = b_key2.as<uint64_t>(npieces + 1);
  hipLaunchKernelGGL(k_iota_u32, dim3(div_up(npieces, 256)), dim3(256), 0, st, ord, npieces);
  GSJ_LAUNCHED();
  radix_sort_pairs(c, key, key2, ord, ord2, npieces, 0, 64);
  SdLevelPtrs lv{};
  for (size_t l = 0; l < levels.size(); l++) { lv.rg_off[l] = levels[l].rg_off; lv.rg_n[l] = levels[l].rg_n; lv.pool[l] = levels[l].pool; lv.syn[l] = levels[l].syn; }
  DevBuf b_onr, b_ogro, b_rlev, b_rsrc, b_rn, b_orvo, b_oxy, b_osyn, b_par;
  uint32_t* onr = b_onr.as<uint32_t>(npieces + 16);
  uint64_t* ogro = b_ogro.as<uint64_t>(npieces + 16);
  hipLaunchKernelGGL(k_sd_out_nr, dim3(div_up(npieces, 256)), dim3(256), 0, st, ord, pnr, npieces, onr);
  GSJ_LAUNCHED();
  const uint64_t onrings = scan_u32_to_u64(c, onr, ogro, npieces);
  GSJ_REQUIRE(onrings < (1ull << 32), GSJ_E_LIMIT, "subdivide: more than 2^32 output rings");
  uint32_t* rlev = b_rlev.as<uint32_t>(onrings + 1);
  uint32_t* rsrc = b_rsrc.as<uint32_t>(onrings + 1);
  uint32_t* rn = b_rn.as<uint32_t>(onrings + 16);
  uint64_t* orvo = b_orvo.as<uint64_t>(onrings + 16);
  hipLaunchKernelGGL(k_sd_out_rings, dim3(div_up(npieces, 256)), dim3(256), 0, st, ord, plev, pr0, pnr, ogro, npieces, lv, rlev, rsrc, rn);
  GSJ_LAUNCHED();
  const uint64_t onv = scan_u32_to_u64(c, rn, orvo, onrings);
  double2* oxy = b_oxy.as<double2>(onv + 1);
  uint8_t* osyn = b_osyn.as<uint8_t>(onv + 1);
  uint32_t* par = b_par.as<uint32_t>(npieces + 1);
  if (onrings) hipLaunchKernelGGL(k_sd_out_copy, dim3(div_up(onrings, 256)), dim3(256), 0, st, rlev, rsrc, orvo, onrings, lv, oxy, osyn);
  hipLaunchKernelGGL(k_sd_parent, dim3(div_up(npieces, 256)), dim3(256), 0, st, key, npieces, par);
  GSJ_LAUNCHED();
  o.gro.resize(npieces + 1);
  o.rvo.resize(onrings + 1);
  o.xy.resize(2 * onv);
  o.syn.resize(onv);
  o.parent.resize(npieces);
  GSJ_HIP(hipMemcpyAsync(o.gro.data(), ogro, npieces * 8, hipMemcpyDeviceToHost, st));
  GSJ_HIP(hipMemcpyAsync(o.rvo.data(), orvo, onrings * 8, hipMemcpyDeviceToHost, st));
  if (onv) {
    GSJ_HIP(hipMemcpyAsync(o.xy.data(), oxy, onv * 16, hipMemcpyDeviceToHost, st));
    GSJ_HIP(hipMemcpyAsync(o.syn.data(), osyn, onv, hipMemcpyDeviceToHost, st));
  }
  GSJ_HIP(hipMemcpyAsync(o.parent.data(), par, npieces * 4, hipMemcpyDeviceToHost, st));
  GSJ_HIP(hipStreamSynchronize(st));
  o.gro[npieces] = onrings;
  o.rvo[onrings] = onv;
  return o;
}

}  // namespace gsj
