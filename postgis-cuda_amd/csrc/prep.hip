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
