// The layer loader's parse of a binary COPY stream (daemon/loader.h:46-170) with the EWKB work on the device: the
// host frames the rows (2 + 4 + 4 bytes per row, sequential), then a thread per row walks each value's headers twice
// (counts, then ring records) and a thread per ring decodes its coordinates, takes the compensated signed area and
// writes the ring in the orientation the rules ask for (shells counter-clockwise, holes clockwise). The reference
// parses on a CPU thread per COPY connection.
#include "capi_internal.h"
#include "engine.h"
#include "hostprep.h"
#include <cstring>

namespace gsj {

namespace {

constexpr int MAX_NEST = 32;   // nesting of collections accepted (the reference recurses without a bound)

// row outcomes; R_NO_RINGS (polygonal, but no ring of 4 or more points) is empty in a polygon layer and a wrong-kind
// row in a point layer once that layer's first point was read (loader.h:158-161 test the kind first)
enum : uint8_t { R_NULL = 0, R_BAD = 1, R_EMPTY = 2, R_POLY = 3, R_POINT = 4, R_NO_RINGS = 5 };

__device__ __forceinline__ uint32_t ld32(const uint8_t* p, bool little) {
  uint32_t v;
  __builtin_memcpy(&v, p, 4);
  return little ? v : __builtin_bswap32(v);
}
__device__ __forceinline__ double ld64(const uint8_t* p, bool little) {
  uint64_t v;
  __builtin_memcpy(&v, p, 8);
  if (!little) v = __builtin_bswap64(v);
  return __longlong_as_double((long long)v);
}

struct RingRec {        // a kept ring: where its points start, how many, their layout, and where they go
  uint64_t src, dst;
  uint32_t np;
  uint8_t dims, little, shell, pad;
};

// One walk over the value at b[0, len): false when it is malformed or unsupported (the parse's failure cases,
// loader.h:74-123). EMIT false: counts the kept rings (>= 4 points) and their points; EMIT true: writes their
// records from ring slot r0 and vertex slot v0 on. kind: R_POLY once a polygon header is read, R_POINT for a top-level
// point other than POINT EMPTY (px, py: its coordinates), else R_EMPTY.
template <bool EMIT>
__device__ bool walk(const uint8_t* b, uint32_t len, uint8_t& kind, uint32_t& nr, uint32_t& nv, double& px, double& py, RingRec* rings,
                     uint64_t* rvo, uint64_t r0, uint64_t v0) {
  uint32_t left[MAX_NEST];
  int depth = 0;
  left[0] = 1;
  uint32_t at = 0;
  kind = R_EMPTY;
  nr = 0;
  nv = 0;
  for (;;) {
    if (left[depth] == 0) {
      if (depth == 0) return true;
      depth--;
      continue;
    }
    left[depth]--;
    if (len - at < 5) return false;
    const bool little = b[at] == 1;
    uint32_t type = ld32(b + at + 1, little);
    at += 5;
    bool z = (type & 0x80000000u) != 0, m = (type & 0x40000000u) != 0;
    if (type & 0x20000000u) {
      if (len - at < 4) return false;
      at += 4;
    }
    uint32_t base = type & 0x0FFFFFFFu;
    if (base >= 1000) {
      const uint32_t iso = base / 1000;
      base %= 1000;
      z = z || iso == 1 || iso == 3;
      m = m || iso == 2 || iso == 3;
    }
    const uint32_t dims = 2 + (z ? 1 : 0) + (m ? 1 : 0);
    if (base == 1) {
      if (len - at < 8 * dims) return false;
      const double x = ld64(b + at, little), y = ld64(b + at + 8, little);
      at += 8 * dims;
      if (isnan(x) || isnan(y)) continue;   // POINT EMPTY
      if (depth > 0) return false;          // a point inside a collection: not a polygonal layer row
      kind = R_POINT;
      px = x;
      py = y;
      continue;
    }
    if (base == 3) {
      if (len - at < 4) return false;
      const uint32_t rings_here = ld32(b + at, little);
      at += 4;
      for (uint32_t k = 0; k < rings_here; k++) {
        if (len - at < 4) return false;
        const uint32_t np = ld32(b + at, little);
        at += 4;
        if ((uint64_t)(len - at) < (uint64_t)np * 8 * dims) return false;
        if (np >= 4) {
          if (EMIT) {
            RingRec& R = rings[r0 + nr];
            R.src = at;   // relative to the value; the ring kernel adds the value's offset
            R.dst = v0 + nv;
            R.np = np;
            R.dims = (uint8_t)dims;
            R.little = little ? 1 : 0;
            R.shell = k == 0 ? 1 : 0;
            rvo[r0 + nr] = v0 + nv;
          }
          nr++;
          nv += np;
        }
        at += np * 8 * dims;
      }
      kind = R_POLY;
      continue;
    }
    if (base == 6 || base == 7) {
      if (len - at < 4) return false;
      const uint32_t ng = ld32(b + at, little);
      at += 4;
      if (depth + 1 == MAX_NEST) return false;
      left[++depth] = ng;
      continue;
    }
    return false;
  }
}

// pass 1: a thread per row: outcome class, kept rings and points
__global__ void k_ewkb_scan(const uint8_t* __restrict__ buf, const uint64_t* __restrict__ at, const int32_t* __restrict__ len, uint64_t n,
                            uint8_t* __restrict__ cls, uint32_t* __restrict__ nr, uint32_t* __restrict__ nv, uint64_t* __restrict__ first) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  uint8_t c = R_NULL;
  uint32_t r = 0, v = 0;
  if (len[i] >= 0) {
    uint8_t kind;
    double px, py;
    if (!walk<false>(buf + at[i], (uint32_t)len[i], kind, r, v, px, py, nullptr, nullptr, 0, 0)) c = R_BAD;
    else if (kind == R_POINT) c = R_POINT;
    else if (kind == R_POLY) c = r > 0 ? R_POLY : R_NO_RINGS;
    else c = R_EMPTY;
  }
  cls[i] = c;
  nr[i] = r;
  nv[i] = v;
  if (c == R_POLY || c == R_POINT) atomicMin((unsigned long long*)first, (unsigned long long)i);
}

// the layer's kind: the one given (an earlier piece of the stream decided it: 1 polygons, 2 points), else that of the
// first accepted row (R_EMPTY when there is none)
__global__ void k_ewkb_kind(const uint8_t* __restrict__ cls, uint64_t n, const uint64_t* __restrict__ first, int given, uint8_t* __restrict__ kind) {
  *kind = given == 1 ? R_POLY : given == 2 ? R_POINT : (*first < n ? cls[*first] : R_EMPTY);
}

// rows of the layer's kind keep their counts; rows of the other kind (all after the first accepted row) are bad rows,
// and so are polygonal rows without rings after the first point of a point layer (loader.h:158-161)
__global__ void k_ewkb_keep(uint8_t* __restrict__ cls, uint32_t* __restrict__ nr, uint32_t* __restrict__ nv, uint32_t* __restrict__ keep, uint64_t n,
                            const uint64_t* __restrict__ first, int given, const uint8_t* __restrict__ kindp, uint32_t* __restrict__ hist) {
  __shared__ uint32_t h[8];
  if (threadIdx.x < 8) h[threadIdx.x] = 0;
  __syncthreads();
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) {
    const uint8_t kind = *kindp;
    uint8_t c = cls[i];
    if ((c == R_POLY || c == R_POINT) && c != kind) c = R_BAD;
    if (c == R_NO_RINGS) c = (kind == R_POINT && (given != 0 || i > *first)) ? R_BAD : R_EMPTY;
    cls[i] = c;
    const bool k = c == R_POLY || c == R_POINT;
    keep[i] = k ? 1u : 0u;
    if (!k) { nr[i] = 0; nv[i] = 0; }
    atomicAdd(&h[c], 1u);
  }
  __syncthreads();
  if (threadIdx.x < 8 && h[threadIdx.x]) atomicAdd(&hist[threadIdx.x], h[threadIdx.x]);   // one add per class and block
}

// pass 2: ring records and offsets of the kept polygon rows, the row of each output geometry
__global__ void k_ewkb_rings(const uint8_t* __restrict__ buf, const uint64_t* __restrict__ at, const int32_t* __restrict__ len, uint64_t n,
                             const uint8_t* __restrict__ cls, const uint64_t* __restrict__ gix, const uint64_t* __restrict__ roff,
                             const uint64_t* __restrict__ voff, uint64_t* __restrict__ gro, uint64_t* __restrict__ rvo, RingRec* __restrict__ rings,
                             uint64_t* __restrict__ row) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n || cls[i] != R_POLY) return;
  uint8_t kind;
  uint32_t r, v;
  double px, py;
  walk<true>(buf + at[i], (uint32_t)len[i], kind, r, v, px, py, rings, rvo, roff[i], voff[i]);
  for (uint32_t k = 0; k < r; k++) rings[roff[i] + k].src += at[i];
  gro[gix[i]] = roff[i];
  row[gix[i]] = i;
}

// the kept point rows: coordinates and rows
__global__ void k_ewkb_points(const uint8_t* __restrict__ buf, const uint64_t* __restrict__ at, const int32_t* __restrict__ len, uint64_t n,
                              const uint8_t* __restrict__ cls, const uint64_t* __restrict__ gix, double2* __restrict__ pxy, uint64_t* __restrict__ row) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n || cls[i] != R_POINT) return;
  uint8_t kind;
  uint32_t r, v;
  double px = 0, py = 0;
  walk<false>(buf + at[i], (uint32_t)len[i], kind, r, v, px, py, nullptr, nullptr, 0, 0);
  pxy[gix[i]] = make_double2(px, py);
  row[gix[i]] = i;
}

// a thread per kept ring: twice its signed area (fan from the first point over points 1 .. np-2 in coordinates
// relative to it, each cross product as two rounded products and their fma residues, Neumaier-compensated in that
// order, loader.h:56-70), then its points in the orientation asked for (reversed when the sign disagrees)
__global__ void k_ewkb_coords(const uint8_t* __restrict__ buf, const RingRec* __restrict__ rings, uint64_t nr, double2* __restrict__ xy) {
  const uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nr) return;
  const RingRec R = rings[k];
  const bool little = R.little != 0;
  const uint64_t stride = 8ull * R.dims;
  const uint8_t* p = buf + R.src;
  const double ox = ld64(p, little), oy = ld64(p + 8, little);
  double sum = 0.0, comp = 0.0;
  double ax = ld64(p + stride, little) - ox, ay = ld64(p + stride + 8, little) - oy;
  for (uint32_t i = 1; i + 2 < R.np; i++) {
    const uint8_t* q = p + (uint64_t)(i + 1) * stride;
    const double bx = ld64(q, little) - ox, by = ld64(q + 8, little) - oy;
    const double pp = ax * by, qq = ay * bx;
    const double terms[4] = {pp, -qq, fma(ax, by, -pp), -fma(ay, bx, -qq)};
    for (int t = 0; t < 4; t++) {
      const double v = terms[t], s = sum + v;
      comp += fabs(sum) >= fabs(v) ? (sum - s) + v : (v - s) + sum;
      sum = s;
    }
    ax = bx;
    ay = by;
  }
  const bool ccw = sum + comp > 0;
  const bool flip = ccw != (R.shell != 0);
  double2* out = xy + R.dst;
  for (uint32_t i = 0; i < R.np; i++) {
    const uint8_t* q = p + (uint64_t)i * stride;
    out[flip ? R.np - 1 - i : i] = make_double2(ld64(q, little), ld64(q + 8, little));
  }
}

__global__ void k_ewkb_init(uint64_t* first, uint32_t* hist) {
  *first = ~0ull;
  for (int c = 0; c < 8; c++) hist[c] = 0;
}

uint32_t be32(const uint8_t* p) {
  uint32_t v;
  std::memcpy(&v, p, 4);
  return __builtin_bswap32(v);
}

}  // namespace

CopyFrames copy_frames(const uint8_t* buf, uint64_t len, bool header) {
  CopyFrames f;
  uint64_t pos = 0;
  if (header) {
    static const uint8_t sig[11] = {'P', 'G', 'C', 'O', 'P', 'Y', '\n', 0xFF, '\r', '\n', 0};
    if (len < 19) return f;
    GSJ_REQUIRE(std::memcmp(buf, sig, 11) == 0, GSJ_E_ARG, "malformed COPY stream: no PGCOPY signature");
    const uint64_t ext = be32(buf + 15);
    if (len < 19 + ext) return f;
    pos = 19 + ext;
  }
  for (;;) {
    f.consumed = pos;
    if (len - pos < 2) break;
    const int16_t nf = (int16_t)(uint16_t)((buf[pos] << 8) | buf[pos + 1]);
    if (nf == -1) {
      f.done = true;
      f.consumed = pos + 2;
      break;
    }
    uint64_t q = pos + 2;
    if (len - q < 4) break;
    const int32_t lid = (int32_t)be32(buf + q);
    q += 4;
    if (lid > 0 && len - q < (uint64_t)lid) break;
    const uint64_t id_at = q;
    q += lid > 0 ? (uint64_t)lid : 0;
    if (len - q < 4) break;
    const int32_t lg = (int32_t)be32(buf + q);
    q += 4;
    if (lg > 0 && len - q < (uint64_t)lg) break;
    f.geom_at.push_back(q);
    f.geom_len.push_back(lid < 0 || lg < 0 ? -1 : lg);   // a NULL id or geometry: a NULL row
    f.id_at.push_back(id_at);
    f.id_len.push_back(lid < 0 ? 0 : (uint32_t)lid);
    q += lg > 0 ? (uint64_t)lg : 0;
    pos = q;
  }
  return f;
}

CopyLayer copy_parse_device(Context& c, const uint8_t* buf, uint64_t len, bool header, int given) {
  GSJ_REQUIRE(given >= 0 && given <= 2, GSJ_E_ARG, "copy parse: layer kind must be 0 (undecided), 1 (polygons) or 2 (points)");
  CopyLayer out;
  const CopyFrames f = copy_frames(buf, len, header);
  out.consumed = f.consumed;
  out.done = f.done;
  const uint64_t n = f.geom_at.size();
  out.rows = n;
  out.polys.gro.assign(1, 0);
  out.polys.rvo.assign(1, 0);
  out.id_off.assign(1, 0);   // id_off[0] exists whatever returns below (the C ABI copies id_off[0 .. accepted])
  if (n == 0) return out;
  GSJ_REQUIRE(n < (1ull << 32), GSJ_E_LIMIT, "copy parse: more than 2^32 rows in one call");
  hipStream_t st = c.stream;
  DevBuf b_buf, b_at, b_len, b_cls, b_nr, b_nv, b_keep, b_gix, b_roff, b_voff, b_misc, b_gro, b_rvo, b_ring, b_row, b_xy;
  const uint64_t span = f.consumed;   // the rows' bytes (the header and any incomplete row are not needed)
  uint8_t* dbuf = b_buf.as<uint8_t>(span + 16);
  GSJ_HIP(hipMemcpyAsync(dbuf, buf, span, hipMemcpyHostToDevice, st));
  uint64_t* at = b_at.as<uint64_t>(n);
  int32_t* ln = b_len.as<int32_t>(n);
  GSJ_HIP(hipMemcpyAsync(at, f.geom_at.data(), n * 8, hipMemcpyHostToDevice, st));
  GSJ_HIP(hipMemcpyAsync(ln, f.geom_len.data(), n * 4, hipMemcpyHostToDevice, st));
  uint8_t* cls = b_cls.as<uint8_t>(n);
  uint32_t* nr = b_nr.as<uint32_t>(n + 1);
  uint32_t* nv = b_nv.as<uint32_t>(n + 1);
  uint32_t* keep = b_keep.as<uint32_t>(n + 1);
  char* misc = b_misc.as<char>(128);
  uint64_t* first = reinterpret_cast<uint64_t*>(misc);
  uint32_t* hist = reinterpret_cast<uint32_t*>(misc + 64);
  const uint32_t blocks = (uint32_t)div_up(n, 256);
  hipLaunchKernelGGL(k_ewkb_init, dim3(1), dim3(1), 0, st, first, hist);
  GSJ_LAUNCHED();
  GSJ_HIP(hipEventRecord(c.ev[0], st));
  hipLaunchKernelGGL(k_ewkb_scan, dim3(blocks), dim3(256), 0, st, dbuf, at, ln, n, cls, nr, nv, first);
  GSJ_LAUNCHED();
  uint8_t* kind = reinterpret_cast<uint8_t*>(misc + 8);
  hipLaunchKernelGGL(k_ewkb_kind, dim3(1), dim3(1), 0, st, cls, n, first, given, kind);
  GSJ_LAUNCHED();
  hipLaunchKernelGGL(k_ewkb_keep, dim3(blocks), dim3(256), 0, st, cls, nr, nv, keep, n, first, given, kind, hist);
  GSJ_LAUNCHED();
  GSJ_HIP(hipEventRecord(c.ev[1], st));
  uint32_t* hh = c.h_scal.as<uint32_t>(8);
  GSJ_HIP(hipMemcpyAsync(hh, hist, 32, hipMemcpyDeviceToHost, st));
  GSJ_HIP(hipStreamSynchronize(st));
  float ms = 0;
  GSJ_HIP(hipEventElapsedTime(&ms, c.ev[0], c.ev[1]));
  out.scan_ms = ms;
  uint32_t h_hist[8];
  std::memcpy(h_hist, hh, sizeof h_hist);
  out.nulls = h_hist[R_NULL];
  out.bad_type = h_hist[R_BAD];
  out.empty = h_hist[R_EMPTY];
  uint64_t* gix = b_gix.as<uint64_t>(n + 1);
  const uint64_t ng = scan_u32_to_u64(c, keep, gix, n);
  if (ng == 0) return out;
  uint64_t* row = b_row.as<uint64_t>(ng);
  out.row.resize(ng);
  out.points_layer = given == 2 || (given == 0 && h_hist[R_POINT] > 0);
  if (out.points_layer) {
    out.npoints = ng;
    double2* pxy = b_xy.as<double2>(ng);
    GSJ_HIP(hipEventRecord(c.ev[2], st));
    hipLaunchKernelGGL(k_ewkb_points, dim3(blocks), dim3(256), 0, st, dbuf, at, ln, n, cls, gix, pxy, row);
    GSJ_LAUNCHED();
    out.pxy.resize(2 * ng);
    GSJ_HIP(hipEventRecord(c.ev[3], st));
    GSJ_HIP(hipMemcpyAsync(out.pxy.data(), pxy, ng * 16, hipMemcpyDeviceToHost, st));
  } else {
    out.npolys = ng;
    uint64_t* roff = b_roff.as<uint64_t>(n + 1);
    uint64_t* voff = b_voff.as<uint64_t>(n + 1);
    const uint64_t nrt = scan_u32_to_u64(c, nr, roff, n);
    const uint64_t nvt = scan_u32_to_u64(c, nv, voff, n);
    uint64_t* gro = b_gro.as<uint64_t>(ng + 1);
    uint64_t* rvo = b_rvo.as<uint64_t>(nrt + 1);
    RingRec* rings = b_ring.as<RingRec>(nrt);
    double2* xy = b_xy.as<double2>(nvt);
    GSJ_HIP(hipEventRecord(c.ev[2], st));
    hipLaunchKernelGGL(k_ewkb_rings, dim3(blocks), dim3(256), 0, st, dbuf, at, ln, n, cls, gix, roff, voff, gro, rvo, rings, row);
    GSJ_LAUNCHED();
    hipLaunchKernelGGL(k_ewkb_coords, dim3((uint32_t)div_up(nrt, 256)), dim3(256), 0, st, dbuf, rings, nrt, xy);
    GSJ_LAUNCHED();
    GSJ_HIP(hipEventRecord(c.ev[3], st));
    out.polys.gro.resize(ng + 1);
    out.polys.rvo.resize(nrt + 1);
    out.polys.xy.resize(2 * nvt);
    GSJ_HIP(hipMemcpyAsync(out.polys.gro.data(), gro, ng * 8, hipMemcpyDeviceToHost, st));
    GSJ_HIP(hipMemcpyAsync(out.polys.rvo.data(), rvo, nrt * 8, hipMemcpyDeviceToHost, st));
    GSJ_HIP(hipMemcpyAsync(out.polys.xy.data(), xy, nvt * 16, hipMemcpyDeviceToHost, st));
    out.polys.gro[ng] = nrt;
    out.polys.rvo[nrt] = nvt;
  }
  GSJ_HIP(hipMemcpyAsync(out.row.data(), row, ng * 8, hipMemcpyDeviceToHost, st));
  GSJ_HIP(hipStreamSynchronize(st));
  GSJ_HIP(hipEventElapsedTime(&ms, c.ev[2], c.ev[3]));
  out.emit_ms = ms;
  // the accepted rows' ids, in row order
  out.id_off.reserve(ng + 1);
  for (uint64_t g = 0; g < ng; g++) {
    const uint64_t r = out.row[g];
    out.ids.append(reinterpret_cast<const char*>(buf + f.id_at[r]), f.id_len[r]);
    out.id_off.push_back(out.ids.size());
  }
  return out;
}

}  // namespace gsj
