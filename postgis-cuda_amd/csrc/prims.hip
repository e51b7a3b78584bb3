// Device primitives written for CDNA (wave64): exclusive scan and a stable LSD radix sort of (u64 key, u32 value).
//
// Scan: reduce-then-scan over 2048-element tiles (256 threads x 8 items, one u64 running sum per thread); inputs of
// up to 16384 elements are scanned by one block in one launch (a request's small scans cost one launch, not three).
// Radix sort: 8-bit digits, 4096-element tiles (256 threads x 16 items). Per pass one histogram kernel, one scan of
// the digit-major histogram, one scatter kernel. Ranks inside a tile come from a wave-level "match" (8 ballots give
// the mask of lanes holding the same digit) plus per-wave digit counts in LDS, so each element's destination is
// computed without per-element LDS atomics and the order of equal keys is preserved. Digits that are identical in
// every key are skipped (found from the OR / AND of all keys).
#include "engine.h"

namespace gsj {

namespace {

constexpr int SCAN_T = 256, SCAN_ITEMS = 8, SCAN_TILE = SCAN_T * SCAN_ITEMS;
constexpr int RS_T = 256, RS_ITEMS = 16, RS_TILE = RS_T * RS_ITEMS, RS_WAVES = RS_T / 64;

__device__ __forceinline__ uint64_t wave_incl_scan(uint64_t v) {
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    uint64_t t = __shfl_up(v, o, 64);
    if (lane >= o) v += t;
  }
  return v;
}

// block-wide exclusive scan of one value per thread (blockDim == SCAN_T); returns exclusive prefix, *total = sum
__device__ __forceinline__ uint64_t block_excl_scan(uint64_t v, uint64_t* total) {
  __shared__ uint64_t wsum[SCAN_T / 64];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  uint64_t inc = wave_incl_scan(v);
  if (lane == 63) wsum[w] = inc;
  __syncthreads();
  uint64_t base = 0, tot = 0;
#pragma unroll
  for (int k = 0; k < SCAN_T / 64; k++) {
    if (k < w) base += wsum[k];
    tot += wsum[k];
  }
  __syncthreads();
  *total = tot;
  return base + inc - v;
}

template <class TI>
__global__ void __launch_bounds__(SCAN_T) k_tile_sums(const TI* in, uint64_t n, uint64_t* sums) {
  const uint64_t base = (uint64_t)blockIdx.x * SCAN_TILE;
  uint64_t s = 0;
#pragma unroll
  for (int k = 0; k < SCAN_ITEMS; k++) {
    uint64_t i = base + (uint64_t)k * SCAN_T + threadIdx.x;
    if (i < n) s += (uint64_t)in[i];
  }
  uint64_t tot;
  (void)block_excl_scan(s, &tot);
  if (threadIdx.x == 0) sums[blockIdx.x] = tot;
}

// single block: exclusive scan of nb tile sums in place; grand total to *total
__global__ void __launch_bounds__(SCAN_T) k_scan_sums(uint64_t* sums, uint64_t nb, uint64_t* total) {
  uint64_t carry = 0;
  for (uint64_t base = 0; base < nb; base += SCAN_T) {
    uint64_t i = base + threadIdx.x;
    uint64_t v = i < nb ? sums[i] : 0;
    uint64_t tot;
    uint64_t ex = block_excl_scan(v, &tot);
    if (i < nb) sums[i] = carry + ex;
    carry += tot;
  }
  if (threadIdx.x == 0 && total) *total = carry;
}

template <class TI, class TO>
__global__ void __launch_bounds__(SCAN_T) k_tile_scan(const TI* in, TO* out, uint64_t n, const uint64_t* offs) {
  const uint64_t base = (uint64_t)blockIdx.x * SCAN_TILE + (uint64_t)threadIdx.x * SCAN_ITEMS;   // contiguous per thread
  uint64_t v[SCAN_ITEMS];
  uint64_t s = 0;
#pragma unroll
  for (int k = 0; k < SCAN_ITEMS; k++) {
    uint64_t i = base + k;
    v[k] = i < n ? (uint64_t)in[i] : 0;
    s += v[k];
  }
  uint64_t tot;
  uint64_t ex = block_excl_scan(s, &tot) + offs[blockIdx.x];
#pragma unroll
  for (int k = 0; k < SCAN_ITEMS; k++) {
    uint64_t i = base + k;
    if (i < n) out[i] = (TO)ex;
    ex += v[k];
  }
}

// one block: exclusive scan of n <= SS_MAX elements (thread t holds elements [SS_ITEMS t, SS_ITEMS (t + 1))); every
// element is read before the barrier and written after it, so in and out may be the same array
constexpr int SS_T = 1024, SS_ITEMS = 16, SS_MAX = SS_T * SS_ITEMS;
template <class TI, class TO>
__global__ void __launch_bounds__(SS_T) k_scan_small(const TI* in, TO* out, uint64_t n, uint64_t* total) {
  __shared__ uint64_t wsum[SS_T / 64];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const uint64_t base = (uint64_t)threadIdx.x * SS_ITEMS;
  uint64_t v[SS_ITEMS];
  uint64_t s = 0;
#pragma unroll
  for (int k = 0; k < SS_ITEMS; k++) {
    v[k] = base + k < n ? (uint64_t)in[base + k] : 0;
    s += v[k];
  }
  const uint64_t inc = wave_incl_scan(s);
  if (lane == 63) wsum[w] = inc;
  __syncthreads();
  uint64_t ex = inc - s, tot = 0;
#pragma unroll
  for (int k = 0; k < SS_T / 64; k++) {
    if (k < w) ex += wsum[k];
    tot += wsum[k];
  }
#pragma unroll
  for (int k = 0; k < SS_ITEMS; k++) {
    if (base + k < n) out[base + k] = (TO)ex;
    ex += v[k];
  }
  if (threadIdx.x == 0 && total) *total = tot;
}

// tile sums for the contiguous-per-thread layout used by k_tile_scan (same tile boundaries, any order inside)
template <class TI>
void scan_impl(Context& c, const TI* in, void* out, bool out64, uint64_t n, uint64_t* d_total) {
  if (n == 0) {
    if (d_total) GSJ_HIP(hipMemsetAsync(d_total, 0, 8, c.stream));
    return;
  }
  if (n <= (uint64_t)SS_MAX) {
    if (out64)
      hipLaunchKernelGGL((k_scan_small<TI, uint64_t>), dim3(1), dim3(SS_T), 0, c.stream, in, (uint64_t*)out, n, d_total);
    else
      hipLaunchKernelGGL((k_scan_small<TI, uint32_t>), dim3(1), dim3(SS_T), 0, c.stream, in, (uint32_t*)out, n, d_total);
    GSJ_LAUNCHED();
    return;
  }
  const uint64_t nb = (n + SCAN_TILE - 1) / SCAN_TILE;
  uint64_t* sums = c.s_scal.as<uint64_t>(nb + 8);
  hipLaunchKernelGGL(k_tile_sums<TI>, dim3((uint32_t)nb), dim3(SCAN_T), 0, c.stream, in, n, sums);
  GSJ_LAUNCHED();
  hipLaunchKernelGGL(k_scan_sums, dim3(1), dim3(SCAN_T), 0, c.stream, sums, nb, d_total);
  GSJ_LAUNCHED();
  if (out64)
    hipLaunchKernelGGL((k_tile_scan<TI, uint64_t>), dim3((uint32_t)nb), dim3(SCAN_T), 0, c.stream, in, (uint64_t*)out, n, sums);
  else
    hipLaunchKernelGGL((k_tile_scan<TI, uint32_t>), dim3((uint32_t)nb), dim3(SCAN_T), 0, c.stream, in, (uint32_t*)out, n, sums);
  GSJ_LAUNCHED();
}

// ------------------------------------------------------------------------------------------- radix sort
__device__ __forceinline__ uint64_t lanemask_lt() {
  const int lane = threadIdx.x & 63;
  return lane == 0 ? 0ull : (~0ull >> (64 - lane));
}

// lanes whose 8-bit digit equals mine (valid lanes only); `valid` lanes carry d in [0, 256)
__device__ __forceinline__ uint64_t match_digit(uint32_t d, bool valid) {
  uint64_t m = __ballot(valid);
#pragma unroll
  for (int b = 0; b < 8; b++) {
    const bool bit = (d >> b) & 1u;
    const uint64_t bal = __ballot(valid && bit);
    m &= bit ? bal : ~bal;
  }
  return valid ? m : 0ull;
}

__global__ void __launch_bounds__(RS_T) k_rs_hist(const uint64_t* keys, uint64_t n, int shift, uint32_t ntiles, uint32_t* ghist) {
  __shared__ uint32_t hist[256];
  hist[threadIdx.x] = 0;
  __syncthreads();
  const uint64_t base = (uint64_t)blockIdx.x * RS_TILE;
  for (int j = 0; j < RS_ITEMS; j++) {
    const uint64_t i = base + (uint64_t)j * RS_T + threadIdx.x;
    const bool valid = i < n;
    const uint32_t d = valid ? (uint32_t)(keys[i] >> shift) & 255u : 0u;
    const uint64_t peers = match_digit(d, valid);
    if (valid && (peers & lanemask_lt()) == 0) atomicAdd(&hist[d], (uint32_t)__popcll(peers));
  }
  __syncthreads();
  ghist[(uint64_t)threadIdx.x * ntiles + blockIdx.x] = hist[threadIdx.x];
}

__global__ void __launch_bounds__(RS_T) k_rs_scatter(const uint64_t* keys, const uint32_t* vals, uint64_t n, int shift, uint32_t ntiles,
                                                     const uint32_t* goff, uint64_t* keys_out, uint32_t* vals_out) {
  __shared__ uint32_t off[256];
  __shared__ uint32_t wcnt[RS_WAVES][256];
  const int t = threadIdx.x, w = t >> 6;
  off[t] = goff[(uint64_t)t * ntiles + blockIdx.x];
#pragma unroll
  for (int k = 0; k < RS_WAVES; k++) wcnt[k][t] = 0;
  __syncthreads();
  const uint64_t base = (uint64_t)blockIdx.x * RS_TILE;
  for (int j = 0; j < RS_ITEMS; j++) {
    const uint64_t i = base + (uint64_t)j * RS_T + t;
    const bool valid = i < n;
    uint64_t key = 0; uint32_t val = 0;
    if (valid) { key = keys[i]; val = vals[i]; }
    const uint32_t d = valid ? (uint32_t)(key >> shift) & 255u : 0u;
    const uint64_t peers = match_digit(d, valid);
    const uint32_t rank = (uint32_t)__popcll(peers & lanemask_lt());
    if (valid && rank == 0) wcnt[w][d] = (uint32_t)__popcll(peers);
    __syncthreads();
    if (valid) {
      uint32_t pos = off[d] + rank;
      for (int k = 0; k < w; k++) pos += wcnt[k][d];
      keys_out[pos] = key;
      vals_out[pos] = val;
    }
    __syncthreads();
    uint32_t add = 0;
#pragma unroll
    for (int k = 0; k < RS_WAVES; k++) { add += wcnt[k][t]; wcnt[k][t] = 0; }
    off[t] += add;
    __syncthreads();
  }
}

__global__ void k_key_bits(const uint64_t* keys, uint64_t n, unsigned long long* acc) {
  uint64_t o = 0, a = ~0ull;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) { o |= keys[i]; a &= keys[i]; }
#pragma unroll
  for (int s = 32; s > 0; s >>= 1) { o |= __shfl_xor(o, s, 64); a &= __shfl_xor(a, s, 64); }
  if ((threadIdx.x & 63) == 0) { atomicOr(&acc[0], (unsigned long long)o); atomicAnd(&acc[1], (unsigned long long)a); }
}

__global__ void k_init_bits(unsigned long long* acc) { acc[0] = 0ull; acc[1] = ~0ull; }

}  // namespace

uint64_t scan_u32_to_u64(Context& c, const uint32_t* in, uint64_t* out, uint64_t n) {
  uint64_t* d_total = c.s_misc.as<uint64_t>(4);
  scan_impl<uint32_t>(c, in, out, true, n, d_total);
  uint64_t* h = c.h_scal.as<uint64_t>(4);
  GSJ_HIP(hipMemcpyAsync(h, d_total, 8, hipMemcpyDeviceToHost, c.stream));
  GSJ_HIP(hipStreamSynchronize(c.stream));
  return h[0];
}

void scan_u32_to_u64_async(Context& c, const uint32_t* in, uint64_t* out, uint64_t n, uint64_t* d_total) {
  scan_impl<uint32_t>(c, in, out, true, n, d_total);
}

void key_bits(Context& c, const uint64_t* keys, uint64_t n, uint64_t& or_all, uint64_t& and_all) {
  unsigned long long* acc = c.s_misc.as<unsigned long long>(4);
  hipLaunchKernelGGL(k_init_bits, dim3(1), dim3(1), 0, c.stream, acc);
  GSJ_LAUNCHED();
  if (n) {
    uint32_t g = (uint32_t)std::min<uint64_t>((n + 255) / 256, 4096);
    hipLaunchKernelGGL(k_key_bits, dim3(g), dim3(256), 0, c.stream, keys, n, acc);
    GSJ_LAUNCHED();
  }
  uint64_t* h = c.h_scal.as<uint64_t>(4);
  GSJ_HIP(hipMemcpyAsync(h, acc, 16, hipMemcpyDeviceToHost, c.stream));
  GSJ_HIP(hipStreamSynchronize(c.stream));
  or_all = h[0]; and_all = n ? h[1] : 0;
}

void radix_sort_pairs(Context& c, uint64_t*& keys, uint64_t*& keys_alt, uint32_t*& vals, uint32_t*& vals_alt, uint64_t n, int bit0, int bit1,
                      bool probe) {
  if (n <= 1) return;
  GSJ_REQUIRE(n < (1ull << 32), GSJ_E_LIMIT, "radix sort: more than 2^32 elements");
  uint64_t varying = ~0ull;
  if (probe) {
    uint64_t o, a;
    key_bits(c, keys, n, o, a);
    varying = o ^ a;
  }
  const uint32_t ntiles = (uint32_t)((n + RS_TILE - 1) / RS_TILE);
  uint32_t* ghist = c.s_tmp2.as<uint32_t>((size_t)256 * ntiles + 8);
  for (int shift = bit0 & ~7; shift < bit1; shift += 8) {
    if (((varying >> shift) & 255u) == 0) continue;
    hipLaunchKernelGGL(k_rs_hist, dim3(ntiles), dim3(RS_T), 0, c.stream, keys, n, shift, ntiles, ghist);
    GSJ_LAUNCHED();
    scan_impl<uint32_t>(c, ghist, ghist, false, (uint64_t)256 * ntiles, nullptr);
    hipLaunchKernelGGL(k_rs_scatter, dim3(ntiles), dim3(RS_T), 0, c.stream, keys, vals, n, shift, ntiles, ghist, keys_alt, vals_alt);
    GSJ_LAUNCHED();
    std::swap(keys, keys_alt);
    std::swap(vals, vals_alt);
  }
}

}  // namespace gsj
