// Host-side layer preparation: .gset/.pset files, ST_Subdivide analogue, exact dedup, and the deterministic
// synthetic layer generators. Plain C++17 (compiled by the host compiler with -ffp-contract=off: the clip
// arithmetic must round exactly as written so that cut points are reproducible).
#include "hostprep.h"
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <numeric>
#include <stdexcept>
#include <string>

namespace gsj {

// ------------------------------------------------------------------------------------------- files
// Malformed contents (header, sizes, offsets) throw std::invalid_argument (GSJ_E_ARG at the C ABI); failing file
// operations throw std::runtime_error (GSJ_E_IO).
// .gset: u32 'GSET', u32 version 1, u64 n, nr, nv, u64 gro[n+1], u64 rvo[nr+1], f64 xy[2 nv] (little endian)
// (reference format: engine/gset.h:13-31, scripts/dump_gset.py:5-12)
namespace {
// size of an open file in bytes
uint64_t file_size(FILE* f, const std::string& path) {
  if (std::fseek(f, 0, SEEK_END) != 0) throw std::runtime_error("cannot seek " + path);
  const long sz = std::ftell(f);
  if (sz < 0 || std::fseek(f, 0, SEEK_SET) != 0) throw std::runtime_error("cannot seek " + path);
  return (uint64_t)sz;
}
}  // namespace

void check_offsets(const uint64_t* gro, const uint64_t* rvo, uint64_t n, uint64_t nr_max, uint64_t nv_max) {
  if (!gro || !rvo) throw std::invalid_argument("null gro or rvo");
  if (gro[0] != 0) throw std::invalid_argument("gro[0] must be 0");
  for (uint64_t g = 0; g < n; g++)
    if (gro[g + 1] < gro[g]) throw std::invalid_argument("gro must be non-decreasing (geometry " + std::to_string(g) + ")");
  const uint64_t nr = gro[n];
  if (nr > nr_max) throw std::invalid_argument("gro[n] exceeds the ring count");
  if (rvo[0] != 0) throw std::invalid_argument("rvo[0] must be 0");
  for (uint64_t r = 0; r < nr; r++)
    if (rvo[r + 1] < rvo[r]) throw std::invalid_argument("rvo must be non-decreasing (ring " + std::to_string(r) + ")");
  if (rvo[nr] > nv_max) throw std::invalid_argument("rvo[nr] exceeds the vertex count");
}

HostSet read_gset(const std::string& path) {
  FILE* f = std::fopen(path.c_str(), "rb");
  if (!f) throw std::runtime_error("cannot open " + path);
  HostSet s;
  try {
    const uint64_t size = file_size(f, path);
    char magic[4];
    uint32_t ver = 0;
    uint64_t hdr[3];
    const bool ok = std::fread(magic, 1, 4, f) == 4 && std::memcmp(magic, "GSET", 4) == 0 && std::fread(&ver, 4, 1, f) == 1 &&
                    std::fread(hdr, 8, 3, f) == 3;
    if (!ok) throw std::invalid_argument("bad gset header in " + path);
    const uint64_t n = hdr[0], nr = hdr[1], nv = hdr[2];
    // the three arrays must fill the rest of the file exactly: bound each count by the file size before any
    // arithmetic on it, so that n + 1, nr + 1 and 2 nv cannot overflow and no allocation exceeds the file
    const uint64_t body = size - 32;
    if (size < 32 || n >= body / 8 || nr >= body / 8 || nv > body / 16 || (n + 1) * 8 + (nr + 1) * 8 + nv * 16 != body)
      throw std::invalid_argument("gset header (n " + std::to_string(n) + ", nr " + std::to_string(nr) + ", nv " + std::to_string(nv) +
                               ") does not match the file size " + std::to_string(size) + " of " + path);
    s.gro.resize(n + 1); s.rvo.resize(nr + 1); s.xy.resize(2 * nv);
    if (std::fread(s.gro.data(), 8, n + 1, f) != n + 1 || std::fread(s.rvo.data(), 8, nr + 1, f) != nr + 1 ||
        std::fread(s.xy.data(), 8, 2 * nv, f) != 2 * nv)
      throw std::runtime_error("short read in " + path);
    std::fclose(f);
    f = nullptr;
    if (s.gro[n] != nr || s.rvo[nr] != nv) throw std::invalid_argument("inconsistent offsets in " + path);
    try {
      check_offsets(s.gro.data(), s.rvo.data(), n, nr, nv);
    } catch (const std::invalid_argument& e) {
      throw std::invalid_argument(std::string(e.what()) + " in " + path);
    }
  } catch (...) {
    if (f) std::fclose(f);
    throw;
  }
  return s;
}

void write_gset(const std::string& path, const uint64_t* gro, const uint64_t* rvo, const double* xy, uint64_t n, uint64_t nr, uint64_t nv) {
  FILE* f = std::fopen(path.c_str(), "wb");
  if (!f) throw std::runtime_error("cannot create " + path);
  const uint32_t ver = 1;
  const uint64_t hdr[3] = {n, nr, nv};
  bool ok = std::fwrite("GSET", 1, 4, f) == 4 && std::fwrite(&ver, 4, 1, f) == 1 && std::fwrite(hdr, 8, 3, f) == 3 &&
            std::fwrite(gro, 8, n + 1, f) == n + 1 && std::fwrite(rvo, 8, nr + 1, f) == nr + 1 && (nv == 0 || std::fwrite(xy, 8, 2 * nv, f) == 2 * nv);
  ok = (std::fclose(f) == 0) && ok;
  if (!ok) throw std::runtime_error("write failed: " + path);
}

// .pset: u32 'PSET', u32 version 1, u64 n, f64 xy[2n] (engine/gset.h:33-45, scripts/dump_pset.py:3)
std::vector<double> read_pset(const std::string& path) {
  FILE* f = std::fopen(path.c_str(), "rb");
  if (!f) throw std::runtime_error("cannot open " + path);
  std::vector<double> xy;
  try {
    const uint64_t size = file_size(f, path);
    char magic[4];
    uint32_t ver;
    uint64_t n = 0;
    const bool ok = std::fread(magic, 1, 4, f) == 4 && std::memcmp(magic, "PSET", 4) == 0 && std::fread(&ver, 4, 1, f) == 1 &&
                    std::fread(&n, 8, 1, f) == 1;
    if (!ok) throw std::invalid_argument("bad pset header in " + path);
    if (size < 16 || n > (size - 16) / 16 || n * 16 != size - 16)
      throw std::invalid_argument("pset header (n " + std::to_string(n) + ") does not match the file size " + std::to_string(size) + " of " + path);
    xy.resize(2 * n);
    if (std::fread(xy.data(), 8, 2 * n, f) != 2 * n) throw std::runtime_error("short read in " + path);
  } catch (...) {
    std::fclose(f);
    throw;
  }
  std::fclose(f);
  return xy;
}

void write_pset(const std::string& path, const double* xy, uint64_t n) {
  FILE* f = std::fopen(path.c_str(), "wb");
  if (!f) throw std::runtime_error("cannot create " + path);
  const uint32_t ver = 1;
  bool ok = std::fwrite("PSET", 1, 4, f) == 4 && std::fwrite(&ver, 4, 1, f) == 1 && std::fwrite(&n, 8, 1, f) == 1 && (n == 0 || std::fwrite(xy, 8, 2 * n, f) == 2 * n);
  ok = (std::fclose(f) == 0) && ok;
  if (!ok) throw std::runtime_error("write failed: " + path);
}

// ------------------------------------------------------------------------------------------- subdivide
// Recursive halving at the middle of the longer bbox side (depth <= 24) until a piece has at most maxv stored
// vertices; every ring is clipped against the half-plane (Sutherland-Hodgman). Edges running along a cut are
// marked synthetic. Pieces come out in depth-first order, lower half first. Same cut arithmetic as the reference
// (engine/subdivide.h:14-67) so cut points are bit-identical: x = m, y = py + (m - px)(qy - py)/(qx - px).
namespace {

struct CRing { std::vector<double> xy; std::vector<uint8_t> syn; };

void clip_half(const CRing& in, int axis, double m, bool lower, CRing& out) {
  out.xy.clear(); out.syn.clear();
  const size_t nverts = in.xy.size() / 2;
  auto keep = [&](size_t i) { const double v = in.xy[2 * i + axis]; return lower ? v <= m : v >= m; };
  auto emit = [&](double x, double y, uint8_t s) {
    const size_t k = out.xy.size();
    if (k >= 2 && out.xy[k - 2] == x && out.xy[k - 1] == y) { out.syn.back() = s; return; }   // repeated point
    out.xy.push_back(x); out.xy.push_back(y); out.syn.push_back(s);
  };
  for (size_t i = 0; i + 1 < nverts; i++) {
    const double ax = in.xy[2 * i], ay = in.xy[2 * i + 1], bx = in.xy[2 * i + 2], by = in.xy[2 * i + 3];
    const bool ka = keep(i), kb = keep(i + 1);
    const uint8_t s = in.syn[i];
    double cx, cy;
    if (axis == 0) { cx = m; cy = ay + (m - ax) * (by - ay) / (bx - ax); }
    else { cy = m; cx = ax + (m - ay) * (bx - ax) / (by - ay); }
    if (ka) emit(ax, ay, s);
    if (ka && !kb) emit(cx, cy, 1);          // leaving: the next edge runs along the cut
    else if (!ka && kb) emit(cx, cy, s);     // entering
  }
  size_t k = out.xy.size() / 2;
  if (k < 3) { out.xy.clear(); out.syn.clear(); return; }
  if (out.xy[0] != out.xy[2 * k - 2] || out.xy[1] != out.xy[2 * k - 1]) {
    const double x0 = out.xy[0], y0 = out.xy[1];
    out.xy.push_back(x0); out.xy.push_back(y0); out.syn.push_back(0);
  } else {
    out.syn.back() = 0;
  }
  k = out.xy.size() / 2;
  double a2 = 0;
  for (size_t i = 0; i + 1 < k; i++) a2 += out.xy[2 * i] * out.xy[2 * i + 3] - out.xy[2 * i + 2] * out.xy[2 * i + 1];
  if (k < 4 || a2 == 0.0) { out.xy.clear(); out.syn.clear(); }
}

struct Work { std::vector<CRing> rings; int depth; };

void split_geometry(std::vector<CRing> rings, int maxv, std::vector<std::vector<CRing>>& pieces) {
  std::vector<Work> stack;
  stack.push_back({std::move(rings), 0});
  while (!stack.empty()) {
    Work w = std::move(stack.back());
    stack.pop_back();
    size_t nv = 0;
    for (const CRing& r : w.rings) nv += r.xy.size() / 2;
    if (nv == 0) continue;
    double x0 = 1e300, y0 = 1e300, x1 = -1e300, y1 = -1e300;
    for (const CRing& r : w.rings)
      for (size_t i = 0; i < r.xy.size(); i += 2) {
        x0 = std::min(x0, r.xy[i]); x1 = std::max(x1, r.xy[i]); y0 = std::min(y0, r.xy[i + 1]); y1 = std::max(y1, r.xy[i + 1]);
      }
    if ((int64_t)nv <= maxv || w.depth >= 24 || (x1 - x0 <= 0 && y1 - y0 <= 0)) { pieces.push_back(std::move(w.rings)); continue; }
    const int axis = (x1 - x0) >= (y1 - y0) ? 0 : 1;
    const double m = axis == 0 ? 0.5 * (x0 + x1) : 0.5 * (y0 + y1);
    Work halves[2];
    for (int side = 0; side < 2; side++) {
      halves[side].depth = w.depth + 1;
      CRing c;
      for (const CRing& r : w.rings) { clip_half(r, axis, m, side == 0, c); if (!c.xy.empty()) halves[side].rings.push_back(c); }
    }
    // depth-first, lower half first
    if (!halves[1].rings.empty()) stack.push_back(std::move(halves[1]));
    if (!halves[0].rings.empty()) stack.push_back(std::move(halves[0]));
  }
}

}  // namespace

HostSet subdivide(const uint64_t* gro, const uint64_t* rvo, const double* xy, uint64_t n, int maxv) {
  HostSet o;
  o.gro.push_back(0); o.rvo.push_back(0);
  o.has_parent = true;
  std::vector<std::vector<CRing>> pieces;
  for (uint64_t g = 0; g < n; g++) {
    const uint64_t r0 = gro[g], r1 = gro[g + 1];
    const uint64_t nv = rvo[r1] - rvo[r0];
    if ((int64_t)nv <= maxv) {
      for (uint64_t r = r0; r < r1; r++) {
        o.xy.insert(o.xy.end(), xy + 2 * rvo[r], xy + 2 * rvo[r + 1]);
        o.syn.insert(o.syn.end(), rvo[r + 1] - rvo[r], 0);
        o.rvo.push_back(o.xy.size() / 2);
      }
      o.gro.push_back(o.rvo.size() - 1);
      o.parent.push_back((uint32_t)g);
      continue;
    }
    std::vector<CRing> rings;
    for (uint64_t r = r0; r < r1; r++) {
      CRing c;
      c.xy.assign(xy + 2 * rvo[r], xy + 2 * rvo[r + 1]);
      c.syn.assign(rvo[r + 1] - rvo[r], 0);
      rings.push_back(std::move(c));
    }
    pieces.clear();
    split_geometry(std::move(rings), maxv, pieces);
    for (auto& p : pieces) {
      for (auto& r : p) {
        o.xy.insert(o.xy.end(), r.xy.begin(), r.xy.end());
        o.syn.insert(o.syn.end(), r.syn.begin(), r.syn.end());
        o.rvo.push_back(o.xy.size() / 2);
      }
      o.gro.push_back(o.rvo.size() - 1);
      o.parent.push_back((uint32_t)g);
    }
  }
  return o;
}

// ------------------------------------------------------------------------------------------- dedup
// Identical geometries (same ring structure, bit-identical coordinates) stored once, in order of first occurrence;
// member lists expand results back to rows (reference: engine/dedup.h:51-124). Grouping by a 64-bit content hash
// with full comparison inside a hash group.
namespace {
uint64_t content_hash(const uint64_t* gro, const uint64_t* rvo, const double* xy, uint64_t g) {
  uint64_t h = 0x9E3779B97F4A7C15ull ^ (gro[g + 1] - gro[g]);
  for (uint64_t r = gro[g]; r < gro[g + 1]; r++) {
    h = (h ^ (rvo[r + 1] - rvo[r])) * 0xBF58476D1CE4E5B9ull;
    for (uint64_t v = 2 * rvo[r]; v < 2 * rvo[r + 1]; v++) {
      uint64_t w;
      std::memcpy(&w, xy + v, 8);
      h = (h ^ w) * 0x94D049BB133111EBull;
      h ^= h >> 31;
    }
  }
  return h;
}
bool same_geometry(const uint64_t* gro, const uint64_t* rvo, const double* xy, uint64_t a, uint64_t b) {
  const uint64_t nra = gro[a + 1] - gro[a];
  if (nra != gro[b + 1] - gro[b]) return false;
  for (uint64_t k = 0; k < nra; k++)
    if (rvo[gro[a] + k + 1] - rvo[gro[a] + k] != rvo[gro[b] + k + 1] - rvo[gro[b] + k]) return false;
  const uint64_t va = rvo[gro[a]], vb = rvo[gro[b]], nv = rvo[gro[a + 1]] - va;
  return std::memcmp(xy + 2 * va, xy + 2 * vb, nv * 16) == 0;
}
}  // namespace

HostSet dedup(const uint64_t* gro, const uint64_t* rvo, const double* xy, uint64_t n, std::vector<uint64_t>& moff, std::vector<uint32_t>& midx) {
  std::vector<uint64_t> h(n);
  for (uint64_t g = 0; g < n; g++) h[g] = content_hash(gro, rvo, xy, g);
  std::vector<uint32_t> ord(n);
  std::iota(ord.begin(), ord.end(), 0u);
  std::sort(ord.begin(), ord.end(), [&](uint32_t a, uint32_t b) { return h[a] != h[b] ? h[a] < h[b] : a < b; });
  std::vector<uint32_t> rep(n);
  for (uint64_t i = 0; i < n;) {
    uint64_t j = i;
    while (j < n && h[ord[j]] == h[ord[i]]) j++;
    for (uint64_t k = i; k < j; k++) {   // representatives of this hash group seen so far, in index order
      const uint32_t g = ord[k];
      rep[g] = g;
      for (uint64_t q = i; q < k; q++)
        if (rep[ord[q]] == ord[q] && same_geometry(gro, rvo, xy, ord[q], g)) { rep[g] = ord[q]; break; }
    }
    i = j;
  }
  std::vector<uint32_t> uid(n, UINT32_MAX), first;
  for (uint64_t g = 0; g < n; g++) if (rep[g] == g) { uid[g] = (uint32_t)first.size(); first.push_back((uint32_t)g); }
  const uint64_t un = first.size();
  moff.assign(un + 1, 0);
  for (uint64_t g = 0; g < n; g++) moff[uid[rep[g]] + 1]++;
  for (uint64_t i = 0; i < un; i++) moff[i + 1] += moff[i];
  midx.resize(n);
  std::vector<uint64_t> fill(moff.begin(), moff.end() - 1);
  for (uint64_t g = 0; g < n; g++) midx[fill[uid[rep[g]]]++] = (uint32_t)g;
  HostSet u;
  u.gro.push_back(0); u.rvo.push_back(0);
  for (uint32_t g : first) {
    for (uint64_t r = gro[g]; r < gro[g + 1]; r++) {
      u.xy.insert(u.xy.end(), xy + 2 * rvo[r], xy + 2 * rvo[r + 1]);
      u.rvo.push_back(u.xy.size() / 2);
    }
    u.gro.push_back(u.rvo.size() - 1);
  }
  return u;
}

// ------------------------------------------------------------------------------------------- generators
namespace {

// counter-based RNG: a pure function of (seed, stream, index), so any geometry is reproducible on its own
inline uint64_t mix64(uint64_t z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
struct Rng {
  uint64_t key;
  Rng(uint64_t seed, uint64_t stream) : key(mix64(mix64(seed) ^ (stream * 0xD1B54A32D192ED03ull))) {}
  uint64_t bits(uint64_t i) const { return mix64(key ^ mix64(i)); }
  double uni(uint64_t i) const { return (double)(bits(i) >> 11) * (1.0 / 9007199254740992.0); }   // [0, 1)
  double normal(uint64_t i) const {   // Box-Muller from two counters
    const double u1 = 1.0 - uni(2 * i), u2 = uni(2 * i + 1);
    return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
  }
};

struct Spec {
  std::map<std::string, std::string> kv;
  explicit Spec(const char* s) {
    std::string str = s ? s : "";
    size_t p = 0;
    while (p < str.size()) {
      size_t e = str.find(';', p);
      if (e == std::string::npos) e = str.size();
      std::string item = str.substr(p, e - p);
      size_t eq = item.find('=');
      if (eq != std::string::npos) kv[item.substr(0, eq)] = item.substr(eq + 1);
      p = e + 1;
    }
  }
  double num(const char* k, double d) const {
    auto it = kv.find(k);
    if (it == kv.end()) return d;
    char* end;
    double v = std::strtod(it->second.c_str(), &end);
    if (*end) throw std::invalid_argument(std::string("generator spec: bad number for ") + k);
    return v;
  }
};

struct Builder {
  HostSet s;
  Builder() { s.gro.push_back(0); s.rvo.push_back(0); }
  void vertex(double x, double y) { s.xy.push_back(x); s.xy.push_back(y); }
  void close_ring(size_t first_vertex) {
    const double x = s.xy[2 * first_vertex], y = s.xy[2 * first_vertex + 1];
    vertex(x, y);
    s.rvo.push_back(s.xy.size() / 2);
  }
  void end_geometry() { s.gro.push_back(s.rvo.size() - 1); }
};

// config 1 parcels: one CCW star polygon per grid cell, k ~ U{kmin..kmax} vertices at evenly spaced angles from a
// random phase, radius rfrac * cell * U(rmin, rmax) drawn per vertex
HostSet gen_parcels_grid(const Spec& sp) {
  const int nx = (int)sp.num("nx", 100), ny = (int)sp.num("ny", 100);
  const double x0 = sp.num("x0", -118.5), y0 = sp.num("y0", 34.0), cell = sp.num("cell", 0.001);
  const int kmin = (int)sp.num("kmin", 4), kmax = (int)sp.num("kmax", 12);
  const double rfrac = sp.num("rfrac", 0.45), rmin = sp.num("rmin", 0.8), rmax = sp.num("rmax", 1.0);
  const uint64_t seed = (uint64_t)sp.num("seed", 0);
  if (nx <= 0 || ny <= 0 || kmin < 3 || kmax < kmin) throw std::invalid_argument("parcels_grid: bad spec");
  Rng rng(seed, 1);
  const double two_pi = 6.283185307179586;
  // two passes over the cells (vertex counts, then vertices at their offsets); every cell draws from its own
  // counter range, so the result does not depend on the number of threads
  const uint64_t n = (uint64_t)nx * ny;
  HostSet s;
  s.gro.resize(n + 1);
  s.rvo.resize(n + 1);
  std::vector<int> kk(n);
#pragma omp parallel for schedule(static)
  for (int64_t id = 0; id < (int64_t)n; id++) kk[id] = kmin + (int)(rng.uni((uint64_t)id * 64) * (kmax - kmin + 1));
  s.gro[0] = 0; s.rvo[0] = 0;
  for (uint64_t id = 0; id < n; id++) { s.gro[id + 1] = id + 1; s.rvo[id + 1] = s.rvo[id] + kk[id] + 1; }
  s.xy.resize(2 * s.rvo[n]);
#pragma omp parallel for schedule(static)
  for (int64_t id = 0; id < (int64_t)n; id++) {
    const int i = (int)(id % nx), j = (int)(id / nx);
    const uint64_t base = (uint64_t)id * 64;
    const int k = kk[id];
    const double phase = rng.uni(base + 1) * two_pi / k;
    const double cx = x0 + (i + 0.5) * cell, cy = y0 + (j + 0.5) * cell;
    double* out = &s.xy[2 * s.rvo[id]];
    for (int v = 0; v < k; v++) {
      const double ang = phase + two_pi * v / k;
      const double r = rfrac * cell * (rmin + (rmax - rmin) * rng.uni(base + 2 + v));
      out[2 * v] = cx + r * std::cos(ang);
      out[2 * v + 1] = cy + r * std::sin(ang);
    }
    out[2 * k] = out[0];
    out[2 * k + 1] = out[1];
  }
  return s;
}

// config 1 regions: kx x ky tiles over [x0, x0 + nx*cell] x [y0, y0 + ny*cell]; every tile side has `edge_segs`
// segments; interior boundaries are jittered perpendicular to the side by N(0, sigma) clamped to 4 sigma (the same
// polyline in both neighbours: bit-identical shared vertices), the outer boundary and the tile corners are not
HostSet gen_regions_grid(const Spec& sp) {
  const int nx = (int)sp.num("nx", 100), ny = (int)sp.num("ny", 100);
  const double x0 = sp.num("x0", -118.5), y0 = sp.num("y0", 34.0), cell = sp.num("cell", 0.001);
  const int kx = (int)sp.num("kx", 4), ky = (int)sp.num("ky", 4), segs = (int)sp.num("edge_segs", 500);
  const double sigma = sp.num("sigma", 1e-5);
  const uint64_t seed = (uint64_t)sp.num("seed", 0);
  if (kx <= 0 || ky <= 0 || segs <= 0) throw std::invalid_argument("regions_grid: bad spec");
  Rng rng(seed, 2);
  const double W = nx * cell / kx, H = ny * cell / ky;
  auto X = [&](int i) { return x0 + W * i; };
  auto Y = [&](int j) { return y0 + H * j; };
  auto jit = [&](uint64_t id) { double z = rng.normal(id); z = std::max(-4.0, std::min(4.0, z)); return sigma * z; };
  // vertical line i (x = X(i)) between rows j and j+1; interior points k = 1..segs-1, bottom to top
  auto vline = [&](int i, int j, int k, double& x, double& y) {
    y = Y(j) + (Y(j + 1) - Y(j)) * k / segs;
    x = X(i) + ((i > 0 && i < kx) ? jit(((uint64_t)(i * (ky + 1) + j) * segs + k) * 2) : 0.0);
  };
  auto hline = [&](int j, int i, int k, double& x, double& y) {
    x = X(i) + (X(i + 1) - X(i)) * k / segs;
    y = Y(j) + ((j > 0 && j < ky) ? jit(((uint64_t)(j * (kx + 1) + i) * segs + k) * 2 + 1) : 0.0);
  };
  Builder b;
  for (int ty = 0; ty < ky; ty++)
    for (int tx = 0; tx < kx; tx++) {
      const size_t first = b.s.xy.size() / 2;
      double x, y;
      b.vertex(X(tx), Y(ty));
      for (int k = 1; k < segs; k++) { hline(ty, tx, k, x, y); b.vertex(x, y); }            // bottom, left to right
      b.vertex(X(tx + 1), Y(ty));
      for (int k = 1; k < segs; k++) { vline(tx + 1, ty, k, x, y); b.vertex(x, y); }        // right, upward
      b.vertex(X(tx + 1), Y(ty + 1));
      for (int k = segs - 1; k >= 1; k--) { hline(ty + 1, tx, k, x, y); b.vertex(x, y); }   // top, right to left
      b.vertex(X(tx), Y(ty + 1));
      for (int k = segs - 1; k >= 1; k--) { vline(tx, ty, k, x, y); b.vertex(x, y); }       // left, downward
      b.close_ring(first);
      b.end_geometry();
    }
  return std::move(b.s);
}

// config 2 parcels: cells of a jittered lattice. Each interior side between two cells is SHARED (both polygons use
// the exact lattice segment), OVERLAP (one side's owner bulges 1-5% of a cell into its neighbour, the neighbour
// keeps the exact segment) or a GAP (both pull back 0.5-2% of a cell, with vertices along the side). A corner is
// the exact lattice point when one of its two sides is SHARED/OVERLAP, else it is inset. Vertex counts:
// clamp(round(lognormal(ln median, sigma)), kmin, kmax), the surplus over corners and bulges placed on GAP sides.
HostSet gen_parcels_mesh(const Spec& sp) {
  const int nx = (int)sp.num("nx", 1549), ny = (int)sp.num("ny", 1549);
  const double x0 = sp.num("x0", -95.5), y0 = sp.num("y0", 38.5);
  const double w = sp.num("w", 1.0 / 1549), h = sp.num("h", 1.0 / 1549);
  const double jitter = sp.num("jitter", 0.12), p_share = sp.num("share", 0.3), p_over = sp.num("overlap", 0.05);
  const double med = sp.num("median", 7.0), sigma = sp.num("sigma", 1.0);
  const int kmin = (int)sp.num("kmin", 4), kmax = (int)sp.num("kmax", 5000);
  const uint64_t seed = (uint64_t)sp.num("seed", 1);
  if (nx <= 0 || ny <= 0 || kmin < 4 || kmax < kmin) throw std::invalid_argument("parcels_mesh: bad spec");
  const Rng r_lat(seed, 10), r_side(seed, 11), r_poly(seed, 12), r_pick(seed, 13), r_pos(seed, 14);
  const double unit = std::min(w, h);
  auto lattice = [&](int i, int j, double& x, double& y) {
    x = x0 + w * i; y = y0 + h * j;
    if (i > 0 && i < nx && j > 0 && j < ny) {
      const uint64_t id = (uint64_t)j * (nx + 1) + i;
      x += w * jitter * (2.0 * r_lat.uni(2 * id) - 1.0);
      y += h * jitter * (2.0 * r_lat.uni(2 * id + 1) - 1.0);
    }
  };
  enum { GAP = 0, SHARED = 1, OVER_LO = 2, OVER_HI = 3 };   // OVER_LO: the lower/left cell owns the bulge
  // side ids: horizontal H(i, j) between lattice (i, j)-(i+1, j); vertical V(i, j) between (i, j)-(i, j+1)
  auto side_state = [&](bool horiz, int i, int j) -> int {
    if (horiz ? (j == 0 || j == ny) : (i == 0 || i == nx)) return GAP;   // outer boundary
    const uint64_t id = horiz ? ((uint64_t)j * nx + i) * 2 : ((uint64_t)j * (nx + 1) + i) * 2 + 1;
    const double u = r_side.uni(2 * id);
    if (u < p_share) return SHARED;
    if (u < p_share + p_over) return r_side.uni(2 * id + 1) < 0.5 ? OVER_LO : OVER_HI;
    return GAP;
  };
  auto side_depth = [&](bool horiz, int i, int j) {
    const uint64_t id = horiz ? ((uint64_t)j * nx + i) * 2 : ((uint64_t)j * (nx + 1) + i) * 2 + 1;
    return unit * (0.01 + 0.04 * r_side.uni(2 * id + 1000000007ull));
  };
  // per cell: side kinds, bulge depths, corners and the vertex budget (everything the vertex count depends on)
  struct Plan {
    int kind[4];   // 0 gap, 1 straight exact, 2 bulge (this cell owns the overlap)
    double bdepth[4], cx[4], cy[4], px[4], py[4];
    long extra;
    int ngap, nverts;   // nverts: stored vertices including the closing one
  };
  auto plan = [&](int i, int j, Plan& P) {
    lattice(i, j, P.cx[0], P.cy[0]); lattice(i + 1, j, P.cx[1], P.cy[1]); lattice(i + 1, j + 1, P.cx[2], P.cy[2]); lattice(i, j + 1, P.cx[3], P.cy[3]);
    // sides in CCW order: bottom, right, top, left; this cell is the lower/left neighbour of its top/right sides
    const int st[4] = {side_state(true, i, j), side_state(false, i + 1, j), side_state(true, i, j + 1), side_state(false, i, j)};
    const bool lower_of[4] = {false, true, true, false};
    for (int s = 0; s < 4; s++) {
      P.bdepth[s] = 0;
      if (st[s] == GAP) P.kind[s] = 0;
      else if (st[s] == SHARED) P.kind[s] = 1;
      else {
        const bool owner = (st[s] == OVER_LO) == lower_of[s];
        P.kind[s] = owner ? 2 : 1;
        if (owner) {
          const int si = s == 1 ? i + 1 : i, sj = s == 2 ? j + 1 : j;
          P.bdepth[s] = side_depth(s == 0 || s == 2, si, sj);
        }
      }
    }
    const double mcx = 0.25 * (P.cx[0] + P.cx[1] + P.cx[2] + P.cx[3]), mcy = 0.25 * (P.cy[0] + P.cy[1] + P.cy[2] + P.cy[3]);
    for (int k = 0; k < 4; k++) {   // polygon corners
      const bool exact = P.kind[(k + 3) & 3] != 0 || P.kind[k] != 0;   // sides before and after corner k
      P.px[k] = exact ? P.cx[k] : P.cx[k] + 0.06 * (mcx - P.cx[k]);
      P.py[k] = exact ? P.cy[k] : P.cy[k] + 0.06 * (mcy - P.cy[k]);
    }
    const uint64_t id = (uint64_t)j * nx + i;
    const double z = r_poly.normal(id);
    long kv = std::lround(std::exp(std::log(med) + sigma * z));
    kv = std::max<long>(kmin, std::min<long>(kmax, kv));
    int base_v = 4;
    P.ngap = 0;
    for (int s = 0; s < 4; s++) { if (P.kind[s] == 2) base_v += 2; if (P.kind[s] == 0) P.ngap++; }
    P.extra = std::max<long>(0, kv - base_v);
    P.nverts = base_v + (P.ngap > 0 ? (int)P.extra : 0) + 1;
  };
  // two passes over the cells (vertex counts, then vertices at their offsets); every cell draws from its own
  // counter ranges, so the result does not depend on the number of threads
  const uint64_t n = (uint64_t)nx * ny;
  HostSet out;
  out.gro.resize(n + 1);
  out.rvo.resize(n + 1);
  std::vector<int> nvc(n);
#pragma omp parallel for schedule(static)
  for (int64_t id = 0; id < (int64_t)n; id++) {
    Plan P;
    plan((int)(id % nx), (int)(id / nx), P);
    nvc[id] = P.nverts;
  }
  out.gro[0] = 0; out.rvo[0] = 0;
  for (uint64_t id = 0; id < n; id++) { out.gro[id + 1] = id + 1; out.rvo[id + 1] = out.rvo[id] + nvc[id]; }
  std::vector<int>().swap(nvc);
  out.xy.resize(2 * out.rvo[n]);
#pragma omp parallel
  {
    int per_side[4];
    std::vector<double> ts;
#pragma omp for schedule(static)
    for (int64_t id = 0; id < (int64_t)n; id++) {
      Plan P;
      plan((int)(id % nx), (int)(id / nx), P);
      double* o = &out.xy[2 * out.rvo[id]];
      int nv = 0;
      auto vertex = [&](double x, double y) { o[2 * nv] = x; o[2 * nv + 1] = y; nv++; };
      for (int s = 0; s < 4; s++) per_side[s] = 0;
      if (P.ngap > 0)
        for (long e = 0; e < P.extra; e++) {
          int pick = (int)(r_pick.uni(((uint64_t)id << 21) + (uint64_t)e) * P.ngap);
          for (int s = 0; s < 4; s++) if (P.kind[s] == 0 && pick-- == 0) { per_side[s]++; break; }
        }
      uint64_t ctr = (uint64_t)id << 21;   // per-polygon counters for vertex positions
      for (int s = 0; s < 4; s++) {
        const int e = (s + 1) & 3;
        vertex(P.px[s], P.py[s]);
        // side direction on the lattice and its inward (left) unit normal, scaled to the cell
        const double dx = P.cx[e] - P.cx[s], dy = P.cy[e] - P.cy[s], len = std::sqrt(dx * dx + dy * dy);
        const double nxu = -dy / len, nyu = dx / len;
        if (P.kind[s] == 2) {
          const double t0 = 0.3, t1 = 0.7;
          vertex(P.cx[s] + t0 * dx - P.bdepth[s] * nxu, P.cy[s] + t0 * dy - P.bdepth[s] * nyu);
          vertex(P.cx[s] + t1 * dx - P.bdepth[s] * nxu, P.cy[s] + t1 * dy - P.bdepth[s] * nyu);
        } else if (P.kind[s] == 0 && per_side[s] > 0) {
          ts.resize(per_side[s]);
          for (int k = 0; k < per_side[s]; k++) ts[k] = 0.1 + 0.8 * r_pos.uni(ctr++);
          std::sort(ts.begin(), ts.end());
          for (int k = 0; k < per_side[s]; k++) {
            const double d = unit * (0.005 + 0.015 * r_pos.uni(ctr++));
            vertex(P.cx[s] + ts[k] * dx + d * nxu, P.cy[s] + ts[k] * dy + d * nyu);
          }
        }
      }
      vertex(o[0], o[1]);   // closing vertex
    }
  }
  return out;
}

}  // namespace

HostSet generate(const std::string& kind, const char* spec) {
  Spec sp(spec);
  if (kind == "parcels_grid") return gen_parcels_grid(sp);
  if (kind == "regions_grid") return gen_regions_grid(sp);
  if (kind == "parcels_mesh") return gen_parcels_mesh(sp);
  throw std::invalid_argument("unknown generator kind: " + kind);
}

// WGS84 metres per degree of longitude (kx) and latitude (ky) at latitude lat (radii of curvature)
static void metres_per_degree(double lat, double& kx, double& ky) {
  const double a = 6378137.0, f = 1.0 / 298.257223563, e2 = f * (2.0 - f), rad = 3.141592653589793 / 180.0;
  const double s = std::sin(lat * rad), c = std::cos(lat * rad);
  const double w2 = 1.0 - e2 * s * s, wq = std::sqrt(w2);
  kx = a / wq * c * rad;
  ky = a * (1.0 - e2) / (w2 * wq) * rad;
}

// config 3 points: n points, a fraction clustered in Gaussian clusters (sigma in metres), the rest uniform
std::vector<double> generate_points(const char* spec) {
  Spec sp(spec);
  const uint64_t n = (uint64_t)sp.num("n", 5000000);
  const double x0 = sp.num("x0", -95.5), y0 = sp.num("y0", 38.5), W = sp.num("w", 1.0), H = sp.num("h", 1.0);
  const uint64_t nc = (uint64_t)sp.num("clusters", 1000);
  const double sig_m = sp.num("sigma_m", 50.0), frac = sp.num("frac_clustered", 0.5);
  const uint64_t seed = (uint64_t)sp.num("seed", 2);
  const Rng r(seed, 20), rc(seed, 21);
  std::vector<double> xy(2 * n);
  const uint64_t ncl = nc ? (uint64_t)std::llround(frac * (double)n) : 0;
#pragma omp parallel for schedule(static)
  for (int64_t ii = 0; ii < (int64_t)n; ii++) {   // every point draws from its own counters
    const uint64_t i = (uint64_t)ii;
    if (i < n - ncl) {
      xy[2 * i] = x0 + W * r.uni(2 * i);
      xy[2 * i + 1] = y0 + H * r.uni(2 * i + 1);
    } else {
      const uint64_t c = (i - (n - ncl)) % nc;
      const double ccx = x0 + W * rc.uni(2 * c), ccy = y0 + H * rc.uni(2 * c + 1);
      double kx, ky;
      metres_per_degree(ccy, kx, ky);
      xy[2 * i] = ccx + sig_m / kx * r.normal(2 * i);
      xy[2 * i + 1] = ccy + sig_m / ky * r.normal(2 * i + 1);
    }
  }
  return xy;
}

// ------------------------------------------------------------------------------------------- binary COPY writer
namespace {
void put_be16(std::string& s, int16_t v) {
  const uint16_t u = (uint16_t)v;
  s.push_back((char)(u >> 8));
  s.push_back((char)(u & 0xFF));
}
void put_be32(std::string& s, int32_t v) {
  const uint32_t u = (uint32_t)v;
  for (int sh = 24; sh >= 0; sh -= 8) s.push_back((char)((u >> sh) & 0xFF));
}
void put_le32(std::string& s, uint32_t v) { s.append(reinterpret_cast<const char*>(&v), 4); }
void put_f64(std::string& s, double v) { s.append(reinterpret_cast<const char*>(&v), 8); }

// sign of a ring's area (shoelace relative to its first vertex): true for counter-clockwise
bool ring_ccw(const double* xy, uint64_t n) {
  double s = 0.0;
  for (uint64_t i = 1; i + 1 < n; i++) s += (xy[2 * i] - xy[0]) * (xy[2 * i + 3] - xy[1]) - (xy[2 * i + 1] - xy[1]) * (xy[2 * i + 2] - xy[0]);
  return s > 0;
}

// one geometry as little-endian EWKB: a POLYGON when its first ring is its only counter-clockwise one, else a
// MULTIPOLYGON whose polygons start at every counter-clockwise ring (clockwise rings are holes of the polygon before)
void put_geometry(std::string& s, const uint64_t* rvo, const double* xy, uint64_t r0, uint64_t r1, int srid) {
  std::vector<uint64_t> starts;
  for (uint64_t r = r0; r < r1; r++)
    if (r == r0 || ring_ccw(xy + 2 * rvo[r], rvo[r + 1] - rvo[r])) starts.push_back(r);
  starts.push_back(r1);
  auto head = [&](uint32_t type, bool with_srid) {
    s.push_back(1);
    put_le32(s, type | (with_srid ? 0x20000000u : 0u));
    if (with_srid) put_le32(s, (uint32_t)srid);
  };
  auto polygon = [&](uint64_t a, uint64_t b, bool with_srid) {
    head(3, with_srid);
    put_le32(s, (uint32_t)(b - a));
    for (uint64_t r = a; r < b; r++) {
      put_le32(s, (uint32_t)(rvo[r + 1] - rvo[r]));
      for (uint64_t v = rvo[r]; v < rvo[r + 1]; v++) {
        put_f64(s, xy[2 * v]);
        put_f64(s, xy[2 * v + 1]);
      }
    }
  };
  const bool with_srid = srid > 0;
  if (starts.size() <= 2) {
    polygon(r0, r1, with_srid);
    return;
  }
  head(6, with_srid);
  put_le32(s, (uint32_t)(starts.size() - 1));
  for (size_t k = 0; k + 1 < starts.size(); k++) polygon(starts[k], starts[k + 1], false);
}
}  // namespace

std::string copy_write(const uint64_t* gro, const uint64_t* rvo, const double* xy, uint64_t n, const char* ids, const uint64_t* id_off, int srid) {
  std::string s("PGCOPY\n\xff\r\n\0", 11);
  const uint64_t nr = n ? gro[n] : 0, nv = n ? rvo[nr] : 0;
  s.reserve(21 + n * 48 + nr * 8 + nv * 16);   // framing and headers per row, a count per ring, 16 bytes per point
  put_be32(s, 0);   // flags
  put_be32(s, 0);   // header extension length
  for (uint64_t g = 0; g < n; g++) {
    put_be16(s, 2);
    std::string id = ids ? std::string(ids + id_off[g], ids + id_off[g + 1]) : std::to_string(g);
    if (id.size() >= (1ull << 31)) throw std::invalid_argument("copy_write: id of 2 GiB or more");
    put_be32(s, (int32_t)id.size());
    s += id;
    const size_t at = s.size();
    put_be32(s, 0);   // geometry length, patched below
    put_geometry(s, rvo, xy, gro[g], gro[g + 1], srid);
    const uint64_t len = s.size() - at - 4;
    if (len >= (1ull << 31)) throw std::invalid_argument("copy_write: geometry value of 2 GiB or more");
    for (int k = 0; k < 4; k++) s[at + k] = (char)((len >> (24 - 8 * k)) & 0xFF);
  }
  put_be16(s, -1);
  return s;
}

}  // namespace gsj
