// Launch-shape table: parse, print, pick by launch size, and find the table of a device at context creation.
#include "tune.h"
#include <dlfcn.h>
#include <cstdlib>
#include <fstream>
#include <sstream>
#include <stdexcept>
#include <sys/stat.h>

namespace gsj {

namespace {

struct Key {
  const char* name;
  uint32_t LaunchShape::*field;
  uint32_t lo, hi;
};

// keys and their admissible ranges (the kernels' compile-time capacities bound tcap and mcap)
const Key kKeys[] = {
    {"tcap", &LaunchShape::tcap, 0, 16},          {"mcap", &LaunchShape::mcap, 0, 64},
    {"screen_occ", &LaunchShape::screen_occ, 4, 5}, {"screen_bpc", &LaunchShape::screen_bpc, 1, 64},
    {"edge_bpc", &LaunchShape::edge_bpc, 1, 64},  {"mixed_bpc", &LaunchShape::mixed_bpc, 1, 64},
    {"large_bpc", &LaunchShape::large_bpc, 1, 64}, {"dw_bpc", &LaunchShape::dw_bpc, 1, 64},
    {"trav_block", &LaunchShape::trav_block, 64, 256}, {"ovf_lds", &LaunchShape::ovf_lds, 0, 2},
    {"fork", &LaunchShape::fork, 0, 1},           {"mixed", &LaunchShape::mixed, 0, 1},
    {"large_cap", &LaunchShape::large_cap, 0, 96},  {"size_sort", &LaunchShape::size_sort, 0, 4},
    {"dw_size_sort", &LaunchShape::dw_size_sort, 0, 4},
};

[[noreturn]] void bad(const std::string& source, int line, const std::string& what) {
  throw std::invalid_argument("tuning " + source + " line " + std::to_string(line) + ": " + what);
}

bool parse_u64(const std::string& s, uint64_t& v) {
  if (s.empty() || s.size() > 19) return false;
  v = 0;
  for (char ch : s) {
    if (ch < '0' || ch > '9') return false;
    v = v * 10 + (uint64_t)(ch - '0');
  }
  return true;
}

bool file_exists(const std::string& p) {
  struct stat st;
  return stat(p.c_str(), &st) == 0 && S_ISREG(st.st_mode);
}

std::string read_file(const std::string& p) {
  std::ifstream f(p, std::ios::binary);
  if (!f) throw std::invalid_argument("cannot read tuning file " + p);
  std::ostringstream o;
  o << f.rdbuf();
  return o.str();
}

// directory of the shared object that holds this function (…/lib for the in-tree build)
std::string library_dir() {
  Dl_info info;
  if (!dladdr(reinterpret_cast<void*>(&library_dir), &info) || !info.dli_fname) return "";
  std::string p = info.dli_fname;
  const size_t s = p.rfind('/');
  return s == std::string::npos ? std::string(".") : p.substr(0, s);
}

}  // namespace

const LaunchShape& Tuning::pick(uint64_t np) const {
  int k = 0;
  while (k + 1 < nb && np >= buckets[k + 1].np_min) k++;
  return buckets[k];
}

Tuning Tuning::parse(const std::string& text, const std::string& source) {
  Tuning t;
  t.nb = 0;
  t.source = source;
  std::istringstream in(text);
  std::string raw;
  int line = 0;
  while (std::getline(in, raw)) {
    line++;
    const size_t hash = raw.find('#');
    std::istringstream ws(hash == std::string::npos ? raw : raw.substr(0, hash));
    std::string tok;
    if (!(ws >> tok)) continue;
    if (tok != "bucket") bad(source, line, "expected 'bucket', got '" + tok + "'");
    LaunchShape b;
    bool have_np = false;
    while (ws >> tok) {
      if (tok.rfind("np>=", 0) == 0) {
        if (!parse_u64(tok.substr(4), b.np_min)) bad(source, line, "bad bucket bound '" + tok + "'");
        have_np = true;
        continue;
      }
      const size_t eq = tok.find('=');
      if (eq == std::string::npos) bad(source, line, "expected key=value, got '" + tok + "'");
      const std::string k = tok.substr(0, eq);
      uint64_t v = 0;
      if (!parse_u64(tok.substr(eq + 1), v)) bad(source, line, "bad value in '" + tok + "'");
      const Key* key = nullptr;
      for (const Key& c : kKeys)
        if (k == c.name) key = &c;
      if (!key) bad(source, line, "unknown key '" + k + "'");
      if (v < key->lo || v > key->hi) bad(source, line, k + " outside [" + std::to_string(key->lo) + ", " + std::to_string(key->hi) + "]");
      b.*(key->field) = (uint32_t)v;
    }
    if (!have_np) bad(source, line, "missing np>=N");
    if (b.mcap != 0 && b.mcap <= 32) bad(source, line, "mcap must be 0 or above the small class's 32 vertices");
    if (b.trav_block != 64 && b.trav_block != 128 && b.trav_block != 256) bad(source, line, "trav_block must be 64, 128 or 256");
    if (t.nb == 0 ? b.np_min != 0 : b.np_min <= t.buckets[t.nb - 1].np_min)
      bad(source, line, "buckets must start at np>=0 and ascend");
    if (t.nb == MAX_BUCKETS) bad(source, line, "more than " + std::to_string(MAX_BUCKETS) + " buckets");
    t.buckets[t.nb++] = b;
  }
  if (t.nb == 0) bad(source, line, "no bucket");
  return t;
}

std::string Tuning::text() const {
  std::ostringstream o;
  o << "# source: " << source << "\n";
  for (int i = 0; i < nb; i++) {
    const LaunchShape& b = buckets[i];
    o << "bucket np>=" << b.np_min;
    for (const Key& k : kKeys) o << " " << k.name << "=" << b.*(k.field);
    o << "\n";
  }
  return o.str();
}

Tuning load_tuning(const std::string& arch) {
  if (const char* env = std::getenv("GSJ_TUNING")) {
    if (*env) return Tuning::parse(read_file(env), env);
  }
  const std::string dir = library_dir();
  if (!dir.empty() && !arch.empty()) {
    const std::string p = dir + "/../tuning/" + arch + ".txt";
    if (file_exists(p)) return Tuning::parse(read_file(p), p);
  }
  return Tuning{};
}

}  // namespace gsj
