// LBVH over the layer's outward-rounded float boxes (reference: build_bvh and its kernels, engine/engine.cuh:99-228).
//
// MI355X design:
//  * 64-bit Morton keys (32 bits per axis, quantised in FP64) instead of 32-bit, sorted by the engine's own LSD radix
//    sort (prims.hip);
//  * Karras (2012) hierarchy emission with ties broken by index; internal boxes from each node's leaf range over a
//    pyramid of aligned block unions (no bottom-up walk: its device-scope acquire/release counters made the L2 of
//    every XCD write back, 6.6 ms at 2.4M leaves);
//  * the finished tree is re-emitted in depth-first preorder: the left child of an internal node is the next node
//    and each node carries the index of the first node after its subtree. Traversal is then a stackless forward
//    scan (no per-lane stack in scratch memory) that touches nodes in increasing address order. The preorder index
//    of a node with leaf range [lo, hi] is 2*lo + (number of left turns on its root path), and its subtree holds
//    2*(hi - lo) + 1 nodes, so both links come from one walk up the parent pointers.
#include "engine.h"

namespace gsj {

namespace {

__device__ __forceinline__ int32_t ord_f(float f) { int32_t i = __float_as_int(f); return i >= 0 ? i : i ^ 0x7FFFFFFF; }
__device__ __forceinline__ float unord_f(int32_t i) { return __int_as_float(i >= 0 ? i : i ^ 0x7FFFFFFF); }

__global__ void k_bounds_init(int32_t* b) {
  b[0] = ord_f(INFINITY); b[1] = ord_f(INFINITY); b[2] = ord_f(-INFINITY); b[3] = ord_f(-INFINITY);
}

// min / max of box centres; one atomic per wave
__global__ void k_centre_bounds(const float4* bb, uint32_t n, int32_t* b) {
  float x0 = INFINITY, y0 = INFINITY, x1 = -INFINITY, y1 = -INFINITY;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const float4 q = bb[i];
    const float cx = 0.5f * (q.x + q.z), cy = 0.5f * (q.y + q.w);
    x0 = fminf(x0, cx); y0 = fminf(y0, cy); x1 = fmaxf(x1, cx); y1 = fmaxf(y1, cy);
  }
#pragma unroll
  for (int s = 32; s > 0; s >>= 1) {
    x0 = fminf(x0, __shfl_xor(x0, s, 64)); y0 = fminf(y0, __shfl_xor(y0, s, 64));
    x1 = fmaxf(x1, __shfl_xor(x1, s, 64)); y1 = fmaxf(y1, __shfl_xor(y1, s, 64));
  }
  if ((threadIdx.x & 63) == 0) {
    atomicMin(&b[0], ord_f(x0)); atomicMin(&b[1], ord_f(y0)); atomicMax(&b[2], ord_f(x1)); atomicMax(&b[3], ord_f(y1));
  }
}

__device__ __forceinline__ uint64_t spread32(uint64_t v) {
  v &= 0xFFFFFFFFull;
  v = (v | (v << 16)) & 0x0000FFFF0000FFFFull;
  v = (v | (v << 8)) & 0x00FF00FF00FF00FFull;
  v = (v | (v << 4)) & 0x0F0F0F0F0F0F0F0Full;
  v = (v | (v << 2)) & 0x3333333333333333ull;
  v = (v | (v << 1)) & 0x5555555555555555ull;
  return v;
}

__global__ void k_morton64(const float4* bb, uint32_t n, const int32_t* b, uint64_t* keys, uint32_t* vals) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double mx = unord_f(b[0]), my = unord_f(b[1]), Mx = unord_f(b[2]), My = unord_f(b[3]);
  const float4 q = bb[i];
  const double cx = 0.5f * (q.x + q.z), cy = 0.5f * (q.y + q.w);
  const double sx = Mx > mx ? 4294967295.0 / (Mx - mx) : 0.0, sy = My > my ? 4294967295.0 / (My - my) : 0.0;
  const double fx = fmin(fmax((cx - mx) * sx, 0.0), 4294967295.0), fy = fmin(fmax((cy - my) * sy, 0.0), 4294967295.0);
  keys[i] = (spread32((uint64_t)fx) << 1) | spread32((uint64_t)fy);
  vals[i] = i;
}

// common-prefix length of sorted keys i and j (ties broken by index); -1 outside [0, n)
__device__ __forceinline__ int kdelta(const uint64_t* k, int n, int i, int j) {
  if (j < 0 || j >= n) return -1;
  const uint64_t a = k[i], b = k[j];
  if (a == b) return 64 + __clz((uint32_t)(i ^ j));
  return __clzll((long long)(a ^ b));
}

// internal node i (Karras 2012, "Maximizing parallelism in the construction of BVHs, octrees, and k-d trees");
// leaves are numbered n-1 .. 2n-2
__global__ void k_karras(const uint64_t* keys, int n, int32_t* left, int32_t* right, int32_t* parent, uint32_t* lo, uint32_t* hi) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n - 1) return;
  const int d = (kdelta(keys, n, i, i + 1) - kdelta(keys, n, i, i - 1)) >= 0 ? 1 : -1;
  const int dmin = kdelta(keys, n, i, i - d);
  int lmax = 2;
  while (kdelta(keys, n, i, i + lmax * d) > dmin) lmax *= 2;
  int l = 0;
  for (int t = lmax >> 1; t >= 1; t >>= 1)
    if (kdelta(keys, n, i, i + (l + t) * d) > dmin) l += t;
  const int j = i + l * d;
  const int dnode = kdelta(keys, n, i, j);
  int s = 0, t = l;
  do {
    t = (t + 1) >> 1;
    if (kdelta(keys, n, i, i + (s + t) * d) > dnode) s += t;
  } while (t > 1);
  const int gamma = i + s * d + (d < 0 ? d : 0);
  const int a = min(i, j), b = max(i, j);
  const int lc = (a == gamma) ? (n - 1 + gamma) : gamma;
  const int rc = (b == gamma + 1) ? (n - 1 + gamma + 1) : gamma + 1;
  left[i] = lc; right[i] = rc;
  parent[lc] = i; parent[rc] = i;
  lo[i] = (uint32_t)a; hi[i] = (uint32_t)b;
}

__global__ void k_leaf_boxes(const float4* bb, const uint32_t* order, uint32_t n, float4* box) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) box[n - 1 + i] = bb[order[i]];
}

// Internal node boxes without a bottom-up walk: node x covers the leaves lo[x] .. hi[x] (Karras), so its box is the
// union of their boxes, read from a pyramid of aligned blocks (level k entry j: the union of leaves [j 2^k, (j+1) 2^k))
// as at most two blocks per level. The same leaves give the same min / max, so the boxes equal those of a refit
// through the children; every node is independent (no arrival counters, no device-scope atomics, no cross-XCD
// coherence traffic).
constexpr int PYR_LEVELS = 34;
struct Pyramid {
  const float4* lev[PYR_LEVELS];
  uint64_t size[PYR_LEVELS];
  int n;
};

__device__ __forceinline__ float4 box_union(float4 a, float4 b) {
  return make_float4(fminf(a.x, b.x), fminf(a.y, b.y), fmaxf(a.z, b.z), fmaxf(a.w, b.w));
}

// one pyramid level from the one below (an odd last entry is carried up alone)
__global__ void k_box_level(const float4* __restrict__ below, uint64_t nb, float4* __restrict__ above, uint64_t na) {
  const uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= na) return;
  const float4 a = below[2 * j];
  above[j] = 2 * j + 1 < nb ? box_union(a, below[2 * j + 1]) : a;
}

__global__ void k_range_boxes(int n, const uint32_t* __restrict__ lo, const uint32_t* __restrict__ hi, Pyramid P, float4* __restrict__ box) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  if (x >= n - 1) return;
  uint64_t l = lo[x], r = (uint64_t)hi[x] + 1;   // leaves [l, r) at level 0
  float4 u = make_float4(INFINITY, INFINITY, -INFINITY, -INFINITY);
  for (int k = 0; l < r && k < P.n; k++) {
    if (l & 1) u = box_union(u, P.lev[k][l++]);
    if (r & 1) u = box_union(u, P.lev[k][--r]);
    l >>= 1;
    r >>= 1;
  }
  box[x] = u;
}

// preorder position and skip link of every node; scatter into the final node array
__global__ void k_preorder(int n, const int32_t* left, const int32_t* parent, const uint32_t* lo, const uint32_t* hi, const float4* box,
                           const uint32_t* order, BNode* nodes, uint32_t* err) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  const int nn = 2 * n - 1;
  if (x >= nn) return;
  uint32_t lefts = 0;
  int cur = x, guard = 0;
  for (;;) {
    const int p = parent[cur];
    if (p < 0) break;
    if (left[p] == cur) lefts++;
    cur = p;
    if (++guard > 200) { atomicOr(err, 1u); return; }
  }
  const bool leaf = x >= n - 1;
  const uint32_t a = leaf ? (uint32_t)(x - (n - 1)) : lo[x], b = leaf ? a : hi[x];
  const uint32_t pre = 2 * a + lefts;
  const uint32_t skip = pre + 2 * (b - a) + 1;
  if (pre >= (uint32_t)nn || skip > (uint32_t)nn || b < a) { atomicOr(err, 2u); return; }
  BNode nd;
  nd.box = box[x];
  nd.skip = (int32_t)skip;
  nd.prim = leaf ? (int32_t)order[a] : -1;
  nd.pad0 = nd.pad1 = 0;
  nodes[pre] = nd;
}

__global__ void k_fill_i32(int32_t* p, uint64_t n, int32_t v) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}
__global__ void k_fill_u32(uint32_t* p, uint64_t n, uint32_t v) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

}  // namespace

__global__ void k_tile_boxes(const float4* __restrict__ bb, const uint32_t* __restrict__ order, uint32_t n, float4* __restrict__ tiles) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  const uint64_t p0 = (uint64_t)t * PROBE_TILE;
  if (p0 >= n) return;
  const uint64_t p1 = min<uint64_t>(p0 + PROBE_TILE, n);
  float4 u = bb[order[p0]];
  for (uint64_t p = p0 + 1; p < p1; p++) {
    const float4 b = bb[order[p]];
    u.x = fminf(u.x, b.x); u.y = fminf(u.y, b.y); u.z = fmaxf(u.z, b.z); u.w = fmaxf(u.w, b.w);
  }
  tiles[t] = u;
}

void tile_boxes(Context& c, const float4* bb, const uint32_t* order, uint32_t n, float4* tiles) {
  const uint32_t nt = (n + PROBE_TILE - 1) / PROBE_TILE;
  if (nt == 0) return;
  hipLaunchKernelGGL(k_tile_boxes, dim3(div_up(nt, 256)), dim3(256), 0, c.stream, bb, order, n, tiles);
  GSJ_LAUNCHED();
}

void morton_order(Context& c, const float4* bb, uint32_t n, uint32_t* order) {
  if (n == 0) return;
  hipStream_t st = c.stream;
  DevBuf t_keys, t_keys2, t_vals, t_vals2, t_misc;
  uint64_t* keys = t_keys.as<uint64_t>(n);
  uint64_t* keys2 = t_keys2.as<uint64_t>(n);
  uint32_t* vals = t_vals.as<uint32_t>(n);
  uint32_t* vals2 = t_vals2.as<uint32_t>(n);
  int32_t* bounds = t_misc.as<int32_t>(16);
  hipLaunchKernelGGL(k_bounds_init, dim3(1), dim3(1), 0, st, bounds);
  hipLaunchKernelGGL(k_centre_bounds, dim3(std::min<uint32_t>(div_up(n, 256), 2048)), dim3(256), 0, st, bb, n, bounds);
  hipLaunchKernelGGL(k_morton64, dim3(div_up(n, 256)), dim3(256), 0, st, bb, n, bounds, keys, vals);
  GSJ_LAUNCHED();
  radix_sort_pairs(c, keys, keys2, vals, vals2, n, 0, 64);
  GSJ_HIP(hipMemcpyAsync(order, vals, (size_t)n * 4, hipMemcpyDeviceToDevice, st));
  GSJ_HIP(hipStreamSynchronize(st));   // the temporaries are freed on return
}

void bvh_build(Layer& L) {
  Context& c = *L.ctx;
  hipStream_t st = c.stream;
  if (L.bvh_mem) { GSJ_HIP(hipFree(L.bvh_mem)); L.bvh_mem = nullptr; L.bvh_bytes = 0; }
  const uint32_t n = (uint32_t)L.n;
  const float4* bb = L.points ? L.pts.bb : L.poly.bb;
  L.bvh = DBvh();
  L.bvh.n = n;
  if (n == 0) { L.has_bvh = true; return; }
  const uint32_t nn = 2 * n - 1;
  const size_t tiles_off = (((size_t)nn * sizeof(BNode) + 255) & ~size_t(255)) + (((size_t)n * 4 + 255) & ~size_t(255));
  L.bvh_bytes = tiles_off + (size_t)((n + PROBE_TILE - 1) / PROBE_TILE) * sizeof(float4) + 256;
  GSJ_HIP(hipMalloc(&L.bvh_mem, L.bvh_bytes));
  BNode* nodes = static_cast<BNode*>(L.bvh_mem);
  uint32_t* order = reinterpret_cast<uint32_t*>(static_cast<char*>(L.bvh_mem) + (((size_t)nn * sizeof(BNode) + 255) & ~size_t(255)));
  // temporaries
  DevBuf t_keys, t_keys2, t_vals, t_vals2, t_tree, t_box, t_misc, t_pyr;
  uint64_t* keys = t_keys.as<uint64_t>(n);
  uint64_t* keys2 = t_keys2.as<uint64_t>(n);
  uint32_t* vals = t_vals.as<uint32_t>(n);
  uint32_t* vals2 = t_vals2.as<uint32_t>(n);
  int32_t* tree = t_tree.as<int32_t>((size_t)n * 2 + (size_t)nn + (size_t)n * 3 + 16);
  int32_t* left = tree;
  int32_t* right = left + n;
  int32_t* parent = right + n;
  uint32_t* lo = reinterpret_cast<uint32_t*>(parent + nn);
  uint32_t* hi = lo + n;
  float4* box = t_box.as<float4>(nn);
  int32_t* bounds = t_misc.as<int32_t>(16);
  uint32_t* err = reinterpret_cast<uint32_t*>(bounds + 8);
  GSJ_HIP(hipEventRecord(c.ev[12], st));
  hipLaunchKernelGGL(k_bounds_init, dim3(1), dim3(1), 0, st, bounds);
  hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, st, err, (uint64_t)1, 0u);
  hipLaunchKernelGGL(k_centre_bounds, dim3(std::min<uint32_t>(div_up(n, 256), 2048)), dim3(256), 0, st, bb, n, bounds);
  hipLaunchKernelGGL(k_morton64, dim3(div_up(n, 256)), dim3(256), 0, st, bb, n, bounds, keys, vals);
  GSJ_LAUNCHED();
  radix_sort_pairs(c, keys, keys2, vals, vals2, n, 0, 64);
  GSJ_HIP(hipMemcpyAsync(order, vals, (size_t)n * 4, hipMemcpyDeviceToDevice, st));
  hipLaunchKernelGGL(k_fill_i32, dim3(div_up(nn, 256)), dim3(256), 0, st, parent, (uint64_t)nn, -1);
  hipLaunchKernelGGL(k_leaf_boxes, dim3(div_up(n, 256)), dim3(256), 0, st, bb, order, n, box);
  if (n > 1) {
    hipLaunchKernelGGL(k_karras, dim3(div_up(n - 1, 256)), dim3(256), 0, st, keys, (int)n, left, right, parent, lo, hi);
    // the block pyramid over the leaf boxes (level 0 is box[n - 1 ..], leaves in Morton order), then every internal
    // node from its leaf range
    Pyramid pyr{};
    pyr.lev[0] = box + (n - 1);
    pyr.size[0] = n;
    uint64_t total = 0;
    for (uint64_t s = n; s > 1; s = (s + 1) / 2) total += (s + 1) / 2;
    float4* up = t_pyr.as<float4>(total + 1);
    int k = 0;
    for (uint64_t s = n; s > 1 && k + 1 < PYR_LEVELS; k++) {
      const uint64_t na = (s + 1) / 2;
      hipLaunchKernelGGL(k_box_level, dim3(div_up(na, 256)), dim3(256), 0, st, pyr.lev[k], s, up, na);
      pyr.lev[k + 1] = up;
      pyr.size[k + 1] = na;
      up += na;
      s = na;
    }
    pyr.n = k + 1;
    GSJ_LAUNCHED();
    hipLaunchKernelGGL(k_range_boxes, dim3(div_up(n - 1, 256)), dim3(256), 0, st, (int)n, lo, hi, pyr, box);
  }
  hipLaunchKernelGGL(k_preorder, dim3(div_up(nn, 256)), dim3(256), 0, st, (int)n, left, parent, lo, hi, box, order, nodes, err);
  GSJ_LAUNCHED();
  float4* tiles = reinterpret_cast<float4*>(static_cast<char*>(L.bvh_mem) + tiles_off);
  tile_boxes(c, bb, order, n, tiles);
  GSJ_HIP(hipEventRecord(c.ev[13], st));
  uint32_t* herr = c.h_scal.as<uint32_t>(4);
  GSJ_HIP(hipMemcpyAsync(herr, err, 4, hipMemcpyDeviceToHost, st));
  GSJ_HIP(hipStreamSynchronize(st));
  GSJ_REQUIRE(herr[0] == 0, GSJ_E_INTERNAL, "bvh build: tree deeper than expected");
  float ms = 0;
  GSJ_HIP(hipEventElapsedTime(&ms, c.ev[12], c.ev[13]));
  L.build_ms = ms;
  L.bvh.nnodes = nn;
  L.bvh.nodes = nodes;
  L.bvh.order = order;
  L.bvh.tiles = tiles;
  L.h_tiles.resize((n + PROBE_TILE - 1) / PROBE_TILE);
  GSJ_HIP(hipMemcpyAsync(L.h_tiles.data(), tiles, L.h_tiles.size() * sizeof(float4), hipMemcpyDeviceToHost, c.stream));
  GSJ_HIP(hipStreamSynchronize(st));
  L.has_bvh = true;
}

}  // namespace gsj
