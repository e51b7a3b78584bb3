// Oracle (test infrastructure, not the product): the layer loader's binary COPY framing and EWKB parse restated on
// the CPU (daemon/loader.h:46-170), the checker of the device parse. A stream of (id text, geom) rows in PGCOPY
// format becomes one layer: polygonal rows one geometry each (all rings of all polygons, shells counter-clockwise and
// holes clockwise by a compensated signed area), point rows points; the first accepted row decides the layer's kind.
#include <cmath>
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

namespace {

struct Bytes {   // a cursor over one value, in the byte order of the geometry being read (loader.h:46-51)
  const uint8_t* at;
  const uint8_t* end;
  bool little = true;
  bool has(uint64_t n) const { return (uint64_t)(end - at) >= n; }
  uint32_t u32() {
    uint32_t v;
    std::memcpy(&v, at, 4);
    at += 4;
    return little ? v : __builtin_bswap32(v);
  }
  double f64() {
    uint64_t v;
    std::memcpy(&v, at, 8);
    at += 8;
    if (!little) v = __builtin_bswap64(v);
    double d;
    std::memcpy(&d, &v, 8);
    return d;
  }
};

#ifndef ORC_MUTANT
#define ORC_MUTANT 0
#endif

// twice the signed area of a closed ring of n points (loader.h:56-70): a fan from the first vertex over vertices
// 1 .. n-2 in coordinates relative to it, each cross product entered as its two rounded products and their fma
// residues into a Neumaier-compensated sum (the order of the four terms fixes the rounding, and with it the sign of
// near-zero rings)
double twice_area(const double* xy, uint64_t n) {
  if (n < 4) return 0.0;
  const double ox = xy[0], oy = xy[1];
#if ORC_MUTANT == 6   // planted mistake: the plain shoelace on absolute coordinates
  double s = 0.0;
  for (uint64_t i = 0; i + 1 < n; i++) s += xy[2 * i] * xy[2 * i + 3] - xy[2 * i + 2] * xy[2 * i + 1];
  (void)ox; (void)oy;
  return s;
#else
  double sum = 0.0, comp = 0.0;
  auto add = [&](double v) {
    const double t = sum + v;
    comp += std::fabs(sum) >= std::fabs(v) ? (sum - t) + v : (v - t) + sum;
    sum = t;
  };
  for (uint64_t i = 1; i + 2 < n; i++) {
    const double ax = xy[2 * i] - ox, ay = xy[2 * i + 1] - oy;
    const double bx = xy[2 * i + 2] - ox, by = xy[2 * i + 3] - oy;
    const double p = ax * by, q = ay * bx;
    add(p);
    add(-q);
    add(std::fma(ax, by, -p));
    add(-std::fma(ay, bx, -q));
  }
  return sum + comp;
#endif
}

struct Layer {   // one stream's output (loader.h:40-45 ChunkOut)
  std::vector<uint64_t> gro{0}, rvo{0}, id_off{0}, row;
  std::vector<double> xy, pxy;
  std::string ids;
  uint64_t rows = 0, nulls = 0, bad_type = 0, empty = 0, polys = 0, points = 0, consumed = 0;
  bool done = false;
  bool have_polys = false, have_points = false;   // the layer's kind, once a row of it was accepted (or given)
  std::string error;
};

enum { K_NONE = 0, K_POLY = 1, K_POINT = 2 };

// one EWKB value (loader.h:72-123): appends its rings (or its point) to L; false on malformed or unsupported input
// (lines, multipoints, a point inside a collection). kind: K_POLY once any polygon is read, K_POINT for a top-level
// point that is not POINT EMPTY.
bool read_geom(Bytes& b, Layer& L, std::vector<double>& ring, uint64_t& rings, int& kind) {
  if (!b.has(5)) return false;
  b.little = *b.at++ == 1;
  const uint32_t type = b.u32();
  bool z = (type & 0x80000000u) != 0, m = (type & 0x40000000u) != 0;
  if (type & 0x20000000u) {   // SRID
    if (!b.has(4)) return false;
    (void)b.u32();
  }
  uint32_t base = type & 0x0FFFFFFFu;
  if (base >= 1000) {         // ISO WKB: 1000s Z, 2000s M, 3000s ZM
    const uint32_t iso = base / 1000;
    base %= 1000;
    z = z || iso == 1 || iso == 3;
    m = m || iso == 2 || iso == 3;
  }
  const uint64_t dims = 2 + (z ? 1 : 0) + (m ? 1 : 0);
  if (base == 1) {
    if (!b.has(8 * dims)) return false;
    const double x = b.f64(), y = b.f64();
    for (uint64_t d = 2; d < dims; d++) (void)b.f64();
    if (std::isnan(x) || std::isnan(y)) return true;
    L.pxy.push_back(x);
    L.pxy.push_back(y);
    kind = K_POINT;
    return true;
  }
  if (base == 3) {
    if (!b.has(4)) return false;
    const uint32_t nr = b.u32();
    for (uint32_t k = 0; k < nr; k++) {
      if (!b.has(4)) return false;
      const uint32_t np = b.u32();
      if (!b.has((uint64_t)np * 8 * dims)) return false;
      ring.resize(2 * (uint64_t)np);
      for (uint32_t i = 0; i < np; i++) {
        ring[2 * i] = b.f64();
        ring[2 * i + 1] = b.f64();
        for (uint64_t d = 2; d < dims; d++) (void)b.f64();
      }
      if (np < 4) continue;   // the ring index k still counts it
#if ORC_MUTANT == 7           // planted mistake: the first kept ring is the shell
      const bool shell = rings == 0;
#else
      const bool shell = k == 0;
#endif
      if ((twice_area(ring.data(), np) > 0) != shell)
        for (uint32_t i = 0, j = np - 1; i < j; i++, j--) {
          std::swap(ring[2 * i], ring[2 * j]);
          std::swap(ring[2 * i + 1], ring[2 * j + 1]);
        }
      L.xy.insert(L.xy.end(), ring.begin(), ring.end());
      L.rvo.push_back(L.xy.size() / 2);
      rings++;
    }
    if (kind == K_NONE) kind = K_POLY;
    return true;
  }
  if (base == 6 || base == 7) {   // multipolygon, geometry collection: polygonal members only
    if (!b.has(4)) return false;
    const uint32_t ng = b.u32();
    for (uint32_t k = 0; k < ng; k++) {
      int sub = K_NONE;
      if (!read_geom(b, L, ring, rings, sub)) return false;
      if (sub == K_POINT) return false;
      if (sub == K_POLY) kind = K_POLY;
    }
    return true;
  }
  return false;
}

uint32_t be32(const uint8_t* p) {
  uint32_t v;
  std::memcpy(&v, p, 4);
  return __builtin_bswap32(v);
}

// the rows of buf[0, len) (loader.h:125-170): with header, buf starts with the PGCOPY signature; rows up to the
// trailer or up to the first incomplete row (consumed = bytes before it, for a caller that appends more)
void read_stream(const uint8_t* buf, uint64_t len, bool header, Layer& L) {
  uint64_t pos = 0;
  if (header) {
    static const uint8_t sig[11] = {'P', 'G', 'C', 'O', 'P', 'Y', '\n', 0xFF, '\r', '\n', 0};
    if (len < 19) return;
    if (std::memcmp(buf, sig, 11) != 0) {
      L.error = "malformed COPY stream";
      return;
    }
    const uint64_t ext = be32(buf + 15);
    if (len < 19 + ext) return;
    pos = 19 + ext;
  }
  std::vector<double> ring;
  for (;;) {
    L.consumed = pos;
    if (len - pos < 2) break;
    const int16_t nf = (int16_t)(uint16_t)((buf[pos] << 8) | buf[pos + 1]);
    if (nf == -1) {
      L.done = true;
      L.consumed = pos + 2;
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
    const uint64_t g_at = q;
    q += lg > 0 ? (uint64_t)lg : 0;
    L.rows++;
    if (lid < 0 || lg < 0) {
      L.nulls++;
    } else {
      Bytes b{buf + g_at, buf + g_at + lg};
      uint64_t rings = 0;
      int kind = K_NONE;
      const size_t xy0 = L.xy.size(), rvo0 = L.rvo.size(), p0 = L.pxy.size();
      const bool ok = read_geom(b, L, ring, rings, kind);
      if (!ok || (L.have_polys && kind == K_POINT) || (L.have_points && kind == K_POLY)) {
        L.xy.resize(xy0);
        L.rvo.resize(rvo0);
        L.pxy.resize(p0);
        L.bad_type++;
      } else if (kind == K_NONE || (kind == K_POLY && rings == 0)) {
        L.xy.resize(xy0);
        L.rvo.resize(rvo0);
        L.empty++;
      } else {
        if (kind == K_POLY) {
          L.gro.push_back(L.rvo.size() - 1);
          L.polys++;
          L.have_polys = true;
        } else {
          L.points++;
          L.have_points = true;
        }
        L.ids.append(reinterpret_cast<const char*>(buf + id_at), (size_t)lid);
        L.id_off.push_back(L.ids.size());
        L.row.push_back(L.rows - 1);
      }
    }
    pos = q;
  }
}

}  // namespace

extern "C" {

// kind: 0 undecided, 1 a polygon layer, 2 a point layer (rows of a stream parsed in pieces: the kind an earlier
// piece decided)
void* orc_copy_parse(const uint8_t* buf, uint64_t len, int header, int kind) {
  Layer* L = new Layer();
  L->have_polys = kind == 1;
  L->have_points = kind == 2;
  read_stream(buf, len, header != 0, *L);
  return L;
}

const char* orc_copy_error(void* h) { return static_cast<Layer*>(h)->error.c_str(); }

// rows, nulls, bad_type, empty, polys, points, consumed, done, n (geometries), nr, nv, npoints
void orc_copy_sizes(void* h, uint64_t* out) {
  const Layer& L = *static_cast<Layer*>(h);
  const uint64_t v[12] = {L.rows, L.nulls, L.bad_type, L.empty, L.polys, L.points, L.consumed, L.done ? 1u : 0u,
                          L.gro.size() - 1, L.rvo.size() - 1, L.xy.size() / 2, L.pxy.size() / 2};
  std::memcpy(out, v, sizeof v);
}

void orc_copy_arrays(void* h, const uint64_t** gro, const uint64_t** rvo, const double** xy, const double** pxy, const uint64_t** id_off,
                     const char** ids, const uint64_t** row) {
  const Layer& L = *static_cast<Layer*>(h);
  *gro = L.gro.data();
  *rvo = L.rvo.data();
  *xy = L.xy.data();
  *pxy = L.pxy.data();
  *id_off = L.id_off.data();
  *ids = L.ids.data();
  *row = L.row.data();
}

void orc_copy_free(void* h) { delete static_cast<Layer*>(h); }

double orc_ring_twice_area(const double* xy, uint64_t n) { return twice_area(xy, n); }

}  // extern "C"
