// Resident layers: upload of the flat .gset arrays and on-device preparation (reference: upload_set,
// engine/engine.cuh:42-89; upload_points, engine/engine.cuh:1068-1076).
//
// One allocation per layer holds every array (64-byte aligned slices). The host sends the raw world coordinates;
// a kernel computes per geometry its FP64 bbox (origin = min corner, extent), the outward-rounded float box, the
// edge count, the real boundary length, and then rewrites the vertices in place as v - origin. Origin-relative
// coordinates are what makes shared vertices bit-identical in every pair frame (see geom.h).
#include "engine.h"
#include <algorithm>
#include <cstring>

namespace gsj {

namespace {

struct Slicer {   // carves aligned slices out of one allocation (first pass sizes, second pass assigns)
  size_t off = 0;
  char* base = nullptr;
  template <class T> T* take(size_t count) {
    off = (off + 63) & ~size_t(63);
    T* p = base ? reinterpret_cast<T*>(base + off) : nullptr;
    off += count * sizeof(T);
    return p;
  }
};

__device__ __forceinline__ int syn_of(const uint8_t* syn, uint64_t v) { return syn ? syn[v] : 0; }

// thread per geometry: bbox, float box, edge count, boundary length; then shift the vertices
__global__ void k_poly_prep(uint32_t n, const uint32_t* gro, const uint64_t* rvo, double2* xy, const uint8_t* syn, double2* origin, double2* ext,
                            float4* bb, uint32_t* ginfo, double* glen) {
  const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= n) return;
  const uint32_t r0 = gro[g], r1 = gro[g + 1];
  const uint64_t v0 = rvo[r0], v1 = rvo[r1];
  double x0 = 1e300, y0 = 1e300, x1 = -1e300, y1 = -1e300;
  for (uint64_t v = v0; v < v1; v++) {
    const double2 p = xy[v];
    x0 = fmin(x0, p.x); x1 = fmax(x1, p.x); y0 = fmin(y0, p.y); y1 = fmax(y1, p.y);
  }
  origin[g] = make_double2(x0, y0);
  ext[g] = make_double2(x1 - x0, y1 - y0);
  // largest float <= lower corner, smallest float >= upper corner (engine/engine.cuh:34-35, 59)
  bb[g] = make_float4(__double2float_rd(x0), __double2float_rd(y0), __double2float_ru(x1), __double2float_ru(y1));
  ginfo[g] = (uint32_t)min<uint64_t>(v1 - v0, 0xFFFFu) | (uint32_t)min<uint32_t>(r1 - r0, 0xFFFFu) << 16;
  double L = 0;
  for (uint32_t r = r0; r < r1; r++)
    for (uint64_t v = rvo[r]; v + 1 < rvo[r + 1]; v++)
      if (!syn_of(syn, v)) { const double2 p = xy[v], q = xy[v + 1]; L += hypot(q.x - p.x, q.y - p.y); }
  glen[g] = L;
  for (uint64_t v = v0; v < v1; v++) { double2 p = xy[v]; xy[v] = make_double2(p.x - x0, p.y - y0); }
}

// packed record of every piece with a unit offset (after k_poly_prep and the flags): origin, extent, header (stored
// vertices, rings, ring starts relative to the first vertex), the origin-relative vertices, their VF_* flag bytes
__global__ void k_pack(uint32_t n, const uint32_t* __restrict__ gro, const uint64_t* __restrict__ rvo, const double2* __restrict__ xy,
                       const uint8_t* __restrict__ vflag, const double2* __restrict__ origin, const double2* __restrict__ ext,
                       const uint32_t* __restrict__ pk_off, uint4* __restrict__ pk) {
  const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= n || pk_off[g] == NO_PACK) return;
  uint4* rec = pk + pk_off[g];
  const uint32_t r0 = gro[g], nr = gro[g + 1] - r0;
  const uint64_t v0 = rvo[r0];
  const uint32_t nv = (uint32_t)(rvo[r0 + nr] - v0);
  reinterpret_cast<double2*>(rec)[0] = origin[g];
  reinterpret_cast<double2*>(rec)[1] = ext[g];
  uint8_t h[16] = {};
  h[0] = (uint8_t)nv;
  h[1] = (uint8_t)nr;
  for (uint32_t r = 0; r <= nr; r++) h[2 + r] = (uint8_t)(rvo[r0 + r] - v0);
  uint4 hw;
  __builtin_memcpy(&hw, h, 16);
  rec[2] = hw;
  double2* V = reinterpret_cast<double2*>(rec + 3);
  for (uint32_t k = 0; k < nv; k++) V[k] = xy[v0 + k];
  uint8_t* F = reinterpret_cast<uint8_t*>(rec + 3 + nv);
  for (uint32_t k = 0; k < (nv + 15) / 16 * 16; k++) F[k] = (k < nv && vflag) ? vflag[v0 + k] : 0;
}

__global__ void k_synb(uint64_t nv, const uint8_t* syn, uint32_t* synb) {
  const uint64_t w = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (w >= (nv + 31) / 32) return;
  uint32_t bits = 0;
  for (int k = 0; k < 32; k++) { uint64_t v = w * 32 + k; if (v < nv && syn[v]) bits |= 1u << k; }
  synb[w] = bits;
}

// VF_* flags of every vertex from the synthetic-edge bits, a thread per ring (the same neighbour rule as the pair
// kernels' edge collection, engine/engine.cuh:358-377)
__global__ void k_vflag(uint32_t nr, const uint64_t* rvo, const uint32_t* synb, uint8_t* vf) {
  const uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= nr) return;
  const uint64_t s = rvo[r], e = rvo[r + 1];
  auto bit = [&](uint64_t i) { return (synb[i >> 5] >> (i & 31)) & 1u; };
  for (uint64_t i = s; i < e; i++) {
    uint8_t f = 0;
    if (i + 1 < e) {
      if (bit(i)) f |= VF_SYN;
      if (bit(i > s ? i - 1 : e - 2)) f |= VF_CUT0;
      if (bit(i + 2 < e ? i + 1 : s)) f |= VF_CUT1;
    }
    vf[i] = f;
  }
}

// first piece of each parent (parents sorted; parents without pieces get an empty range)
__global__ void k_pfirst(uint32_t n, uint32_t nparent, const uint32_t* parent, uint32_t* pf) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint32_t p = parent[i], lo = i ? parent[i - 1] + 1 : 0;
  for (uint32_t q = lo; q <= p; q++) pf[q] = i;
  if (i == n - 1) for (uint32_t q = p + 1; q <= nparent; q++) pf[q] = n;
}

// boundary length per parent, summed over its pieces in piece order (engine/engine.cuh:80)
__global__ void k_blen(uint32_t nparent, const uint32_t* pf, const double* glen, double* blen) {
  const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= nparent) return;
  double s = 0;
  for (uint32_t g = pf[p]; g < pf[p + 1]; g++) s += glen[g];
  blen[p] = s;
}

// outward float boxes of the edge chunks (after the vertices were shifted to their origin); thread per geometry
__global__ void k_chunk_boxes(uint32_t n, const uint32_t* gro, const uint64_t* rvo, const double2* xy, const uint32_t* cfirst, float4* cbb) {
  const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= n || cfirst[g] == NO_CHUNKS) return;
  const uint64_t v0 = rvo[gro[g]], v1 = rvo[gro[g + 1]];
  for (uint64_t k = 0; v0 + k * CHUNK + 1 < v1; k++) {
    const uint64_t a = v0 + k * CHUNK, b = min(a + CHUNK, v1 - 1);   // vertices a..b
    double x0 = 1e300, y0 = 1e300, x1 = -1e300, y1 = -1e300;
    for (uint64_t v = a; v <= b; v++) {
      const double2 p = xy[v];
      x0 = fmin(x0, p.x); x1 = fmax(x1, p.x); y0 = fmin(y0, p.y); y1 = fmax(y1, p.y);
    }
    cbb[cfirst[g] + k] = make_float4(__double2float_rd(x0), __double2float_rd(y0), __double2float_ru(x1), __double2float_ru(y1));
  }
}

__global__ void k_point_prep(uint32_t n, const double2* xy, float4* bb) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double2 p = xy[i];
  bb[i] = make_float4(__double2float_rd(p.x), __double2float_rd(p.y), __double2float_ru(p.x), __double2float_ru(p.y));
}

}  // namespace

Layer::~Layer() {
  if (mem) (void)hipFree(mem);
  if (bvh_mem) (void)hipFree(bvh_mem);
}

void layer_upload_polys(Layer& L, const uint64_t* gro, const uint64_t* rvo, const double* xy, uint64_t n, uint64_t nr, uint64_t nv,
                        const uint32_t* parent, const uint8_t* syn) {
  GSJ_REQUIRE(n < (1ull << 31), GSJ_E_LIMIT, "layer too large: geometries >= 2^31");
  GSJ_REQUIRE(nr < (1ull << 32) - 1, GSJ_E_LIMIT, "layer too large: rings >= 2^32");
  GSJ_REQUIRE(gro && rvo && (xy || nv == 0), GSJ_E_ARG, "null layer array");
  GSJ_REQUIRE(gro[0] == 0 && gro[n] == nr && rvo[0] == 0 && rvo[nr] == nv, GSJ_E_ARG, "inconsistent gro/rvo offsets");
  for (uint64_t g = 0; g < n; g++) GSJ_REQUIRE(gro[g + 1] >= gro[g], GSJ_E_ARG, "gro not non-decreasing");
  for (uint64_t r = 0; r < nr; r++) GSJ_REQUIRE(rvo[r + 1] >= rvo[r], GSJ_E_ARG, "rvo not non-decreasing");
  Context& c = *L.ctx;
  L.points = false; L.n = n; L.nr = nr; L.nv = nv;
  uint32_t nparent = (uint32_t)n;
  if (parent && n) {
    for (uint64_t g = 1; g < n; g++) GSJ_REQUIRE(parent[g] >= parent[g - 1], GSJ_E_ARG, "parent[] must be non-decreasing");
    nparent = parent[n - 1] + 1;
    L.h_parent.assign(parent, parent + n);
  }
  L.nparent = nparent;
  // edge chunks of the geometries with more than CHUNK edges (vertex pairs)
  std::vector<uint32_t> hcf(n);
  uint64_t nchunks = 0;
  for (uint64_t g = 0; g < n; g++) {
    const uint64_t nvg = rvo[gro[g + 1]] - rvo[gro[g]];
    if (nvg > CHUNK + 1) {
      hcf[g] = (uint32_t)nchunks;
      nchunks += (nvg - 1 + CHUNK - 1) / CHUNK;
    } else {
      hcf[g] = NO_CHUNKS;
    }
  }
  GSJ_REQUIRE(nchunks < NO_CHUNKS, GSJ_E_LIMIT, "layer too large: edge chunks >= 2^32");
  // packed records (see DPolys::pk): offsets in 16-byte units
  std::vector<uint32_t> hpk(n);
  uint64_t npk = 0;
  for (uint64_t g = 0; g < n; g++) {
    const uint64_t nrg = gro[g + 1] - gro[g], nvg = rvo[gro[g + 1]] - rvo[gro[g]];
    if (nvg <= PACK_VCAP && nrg <= PACK_RCAP) {
      hpk[g] = (uint32_t)npk;
      npk += pack_units(nvg);
      GSJ_REQUIRE(npk < NO_PACK, GSJ_E_LIMIT, "layer too large: packed records >= 2^32 units of 16 bytes");
    } else {
      hpk[g] = NO_PACK;
    }
  }
  // sizes (pass 1) and assignment (pass 2)
  auto carve = [&](Slicer& s, DPolys& d, uint32_t*& gro32, uint8_t*& syn8, double*& glen) {
    d.origin = s.take<double2>(n); d.ext = s.take<double2>(n); gro32 = s.take<uint32_t>(n + 1); d.rvo = s.take<uint64_t>(nr + 1);
    d.xy = s.take<double2>(nv); d.bb = s.take<float4>(n); d.ginfo = s.take<uint32_t>(n);
    d.parent = parent ? s.take<uint32_t>(n) : nullptr;
    d.synb = syn ? s.take<uint32_t>((nv + 31) / 32) : nullptr;
    d.blen = s.take<double>(nparent);
    d.pfirst = parent ? s.take<uint32_t>((uint64_t)nparent + 1) : nullptr;
    syn8 = syn ? s.take<uint8_t>(nv) : nullptr;
    glen = parent ? s.take<double>(n) : nullptr;
    d.cfirst = nchunks ? s.take<uint32_t>(n) : nullptr;
    d.cbb = nchunks ? s.take<float4>(nchunks) : nullptr;
    d.pk_off = s.take<uint32_t>(n);
    d.pk = s.take<uint4>(npk + 1);
  };
  Slicer s1; DPolys d; uint32_t* gro32; uint8_t* syn8; double* glen;
  carve(s1, d, gro32, syn8, glen);
  L.mem_bytes = s1.off + 64;
  GSJ_HIP(hipMalloc(&L.mem, L.mem_bytes));
  Slicer s2; s2.base = static_cast<char*>(L.mem);
  carve(s2, d, gro32, syn8, glen);
  d.n = (uint32_t)n; d.nr = (uint32_t)nr; d.nv = nv; d.nparent = nparent; d.gro = gro32;
  std::vector<uint32_t> g32(gro, gro + n + 1);   // ring indices fit 32 bits (checked above)
  hipStream_t st = c.stream;
  GSJ_HIP(hipMemcpyAsync(gro32, g32.data(), (n + 1) * 4, hipMemcpyHostToDevice, st));
  GSJ_HIP(hipMemcpyAsync((void*)d.rvo, rvo, (nr + 1) * 8, hipMemcpyHostToDevice, st));
  if (nv) GSJ_HIP(hipMemcpyAsync((void*)d.xy, xy, nv * 16, hipMemcpyHostToDevice, st));
  if (parent) GSJ_HIP(hipMemcpyAsync((void*)d.parent, parent, n * 4, hipMemcpyHostToDevice, st));
  if (syn) GSJ_HIP(hipMemcpyAsync(syn8, syn, nv, hipMemcpyHostToDevice, st));
  if (nchunks) GSJ_HIP(hipMemcpyAsync((void*)d.cfirst, hcf.data(), n * 4, hipMemcpyHostToDevice, st));
  if (n) GSJ_HIP(hipMemcpyAsync((void*)d.pk_off, hpk.data(), n * 4, hipMemcpyHostToDevice, st));
  if (n) {
    // without parents the boundary length of a geometry is its parent's
    double* glen_out = parent ? glen : const_cast<double*>(d.blen);
    hipLaunchKernelGGL(k_poly_prep, dim3(div_up(n, 256)), dim3(256), 0, st, (uint32_t)n, gro32, d.rvo, const_cast<double2*>(d.xy), syn8,
                       const_cast<double2*>(d.origin), const_cast<double2*>(d.ext), const_cast<float4*>(d.bb), const_cast<uint32_t*>(d.ginfo),
                       glen_out);
    GSJ_LAUNCHED();
    if (nchunks) {
      hipLaunchKernelGGL(k_chunk_boxes, dim3(div_up(n, 256)), dim3(256), 0, st, (uint32_t)n, gro32, d.rvo, d.xy, d.cfirst, const_cast<float4*>(d.cbb));
      GSJ_LAUNCHED();
    }
    if (syn && nv) {
      hipLaunchKernelGGL(k_synb, dim3(div_up((nv + 31) / 32, 256)), dim3(256), 0, st, nv, syn8, const_cast<uint32_t*>(d.synb));
      GSJ_LAUNCHED();
      // the uploaded 0/1 bytes are read by k_poly_prep and k_synb above (same stream); then they become the flags
      hipLaunchKernelGGL(k_vflag, dim3(div_up(nr, 256)), dim3(256), 0, st, (uint32_t)nr, d.rvo, d.synb, syn8);
      GSJ_LAUNCHED();
      d.vflag = syn8;
    }
    hipLaunchKernelGGL(k_pack, dim3(div_up(n, 256)), dim3(256), 0, st, (uint32_t)n, gro32, d.rvo, d.xy, d.vflag, d.origin, d.ext, d.pk_off,
                       const_cast<uint4*>(d.pk));
    GSJ_LAUNCHED();
    if (parent) {
      hipLaunchKernelGGL(k_pfirst, dim3(div_up(n, 256)), dim3(256), 0, st, (uint32_t)n, nparent, d.parent, const_cast<uint32_t*>(d.pfirst));
      GSJ_LAUNCHED();
      hipLaunchKernelGGL(k_blen, dim3(div_up(nparent, 256)), dim3(256), 0, st, nparent, d.pfirst, glen, const_cast<double*>(d.blen));
      GSJ_LAUNCHED();
    }
  }
  GSJ_HIP(hipStreamSynchronize(st));
  L.poly = d;
}

void layer_upload_points(Layer& L, const double* xy, uint64_t n) {
  GSJ_REQUIRE(n < (1ull << 31), GSJ_E_LIMIT, "point layer too large: points >= 2^31");
  GSJ_REQUIRE(xy || n == 0, GSJ_E_ARG, "null point array");
  Context& c = *L.ctx;
  L.points = true; L.n = n; L.nparent = (uint32_t)n;
  const uint32_t nt = (uint32_t)((n + PROBE_TILE - 1) / PROBE_TILE);
  Slicer s1; s1.take<double2>(n); s1.take<float4>(n); s1.take<uint32_t>(n); s1.take<float4>(nt);
  L.mem_bytes = s1.off + 64;
  GSJ_HIP(hipMalloc(&L.mem, L.mem_bytes));
  Slicer s2; s2.base = static_cast<char*>(L.mem);
  double2* dxy = s2.take<double2>(n);
  float4* dbb = s2.take<float4>(n);
  uint32_t* order = s2.take<uint32_t>(n);
  float4* tiles = s2.take<float4>(nt);
  if (n) {
    GSJ_HIP(hipMemcpyAsync(dxy, xy, n * 16, hipMemcpyHostToDevice, c.stream));
    hipLaunchKernelGGL(k_point_prep, dim3(div_up(n, 256)), dim3(256), 0, c.stream, (uint32_t)n, dxy, dbb);
    GSJ_LAUNCHED();
    // probes walk the points in Morton order: neighbouring queries visit the same tree nodes and polygons
    morton_order(c, dbb, (uint32_t)n, order);
    tile_boxes(c, dbb, order, (uint32_t)n, tiles);
  }
  GSJ_HIP(hipStreamSynchronize(c.stream));
  L.pts.n = (uint32_t)n; L.pts.xy = dxy; L.pts.bb = dbb; L.pts.order = n ? order : nullptr; L.pts.tiles = n ? tiles : nullptr;
  L.h_tiles.resize(n ? nt : 0);
  if (n) {
    GSJ_HIP(hipMemcpyAsync(L.h_tiles.data(), tiles, nt * sizeof(float4), hipMemcpyDeviceToHost, c.stream));
    GSJ_HIP(hipStreamSynchronize(c.stream));
  }
}

}  // namespace gsj
