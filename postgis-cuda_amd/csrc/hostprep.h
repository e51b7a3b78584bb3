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

HostSet generate(const std::string& kind, const char* spec);
std::vector<double> generate_points(const char* spec);

}  // namespace gsj
