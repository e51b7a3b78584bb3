// Host-side layer preparation (files, subdivision, dedup, generators). No HIP dependency.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

namespace gsj {

// flat polygon set in the .gset layout (reference HostSet, engine/gset.h:5-11), plus subdivision metadata
struct HostSet {
  std::vector<uint64_t> gro, rvo;   // [n+1], [nr+1]
  std::vector<double> xy;           // [2 nv]
  bool has_parent = false;
  std::vector<uint32_t> parent;     // [n] piece -> source geometry (subdivided sets)
  std::vector<uint8_t> syn;         // [nv] edge v -> v+1 lies on a cut (subdivided sets)
  uint64_t n() const { return gro.empty() ? 0 : gro.size() - 1; }
  uint64_t nr() const { return rvo.empty() ? 0 : rvo.size() - 1; }
  uint64_t nv() const { return xy.size() / 2; }
};

// Offsets of a flat set: gro[0] = 0, gro non-decreasing over n + 1 entries, gro[n] <= nr_max, rvo[0] = 0, rvo
// non-decreasing over gro[n] + 1 entries, rvo[gro[n]] <= nv_max. Throws std::invalid_argument otherwise. Every entry
// point that takes caller arrays (files, subdivide, dedup, their device versions) checks them before reading further.
void check_offsets(const uint64_t* gro, const uint64_t* rvo, uint64_t n, uint64_t nr_max = UINT64_MAX, uint64_t nv_max = UINT64_MAX);

HostSet read_gset(const std::string& path);
void write_gset(const std::string& path, const uint64_t* gro, const uint64_t* rvo, const double* xy, uint64_t n, uint64_t nr, uint64_t nv);
std::vector<double> read_pset(const std::string& path);
void write_pset(const std::string& path, const double* xy, uint64_t n);

HostSet subdivide(const uint64_t* gro, const uint64_t* rvo, const double* xy, uint64_t n, int maxv);
HostSet dedup(const uint64_t* gro, const uint64_t* rvo, const double* xy, uint64_t n, std::vector<uint64_t>& moff, std::vector<uint32_t>& midx);
// the same on the device of c (prep.hip): identical distinct set, member lists and order
struct Context;
HostSet dedup_device(Context& c, const uint64_t* gro, const uint64_t* rvo, const double* xy, uint64_t n, std::vector<uint64_t>& moff,
                     std::vector<uint32_t>& midx);
// subdivide on the device of c (prep.hip): the same pieces, bit for bit, in the same order
HostSet subdivide_device(Context& c, const uint64_t* gro, const uint64_t* rvo, const double* xy, uint64_t n, int maxv);

// Binary COPY streams of (id text, geom) rows, the input of the layer loader (daemon/loader.h:125-170).
// copy_frames: the rows of buf[0, len) up to the trailer or the first incomplete row (consumed = bytes before it);
// with header, buf starts with the PGCOPY signature (std::invalid_argument when it does not).
struct CopyFrames {
  std::vector<uint64_t> geom_at, id_at;
  std::vector<int32_t> geom_len;   // -1: the row's id or geometry is NULL
  std::vector<uint32_t> id_len;
  uint64_t consumed = 0;
  bool done = false;               // the trailer was read
};
CopyFrames copy_frames(const uint8_t* buf, uint64_t len, bool header);
// the parsed layer (ewkb.hip): polygon rows as one geometry each (rings oriented), or point rows; ids and stream rows
// of the accepted rows
struct CopyLayer {
  HostSet polys;
  std::vector<double> pxy;
  bool points_layer = false;
  std::vector<uint64_t> row, id_off;
  std::string ids;
  uint64_t rows = 0, nulls = 0, bad_type = 0, empty = 0, npolys = 0, npoints = 0, consumed = 0;
  bool done = false;
  double scan_ms = 0, emit_ms = 0;   // device time of the row walk (pass 1) and of the ring records + coordinates
};
// kind: the layer kind an earlier piece of the same stream decided (0 none yet, 1 polygons, 2 points)
CopyLayer copy_parse_device(Context& c, const uint8_t* buf, uint64_t len, bool header, int kind);
// a polygon set as a binary COPY stream (little-endian EWKB; SRID flag when srid > 0): POLYGON per geometry, or
// MULTIPOLYGON when it holds several counter-clockwise rings (clockwise rings are holes of the polygon before them);
// ids/id_off: the rows' ids, or NULL for the row numbers
std::string copy_write(const uint64_t* gro, const uint64_t* rvo, const double* xy, uint64_t n, const char* ids, const uint64_t* id_off,
                       int srid);

HostSet generate(const std::string& kind, const char* spec);
std::vector<double> generate_points(const char* spec);

}  // namespace gsj
