// Launch shapes of the pair kernels, per size bucket of a launch (number of candidate pairs). The reference fixes its
// shapes in code (engine/engine.cuh:945-951 refinement blocks, engine/buffer.cuh:946-947 distance joins); here they
// come from a table measured on the device by gsj/tune.py and read when a context is created, so a retune needs no
// rebuild. Every shape computes the same rows: the class of a pair only decides which kernel refines it, and every
// class kernel is exact.
#pragma once
#include <cstdint>
#include <string>

namespace gsj {

struct LaunchShape {
  uint64_t np_min = 0;        // the bucket holds launches of at least np_min pairs (up to the next bucket's np_min)
  uint32_t tcap = 16;         // tiny class: both polygons within tcap stored vertices (0: no tiny class; at most 16)
  uint32_t mcap = 64;         // medium class: both within mcap (0: none, such pairs go to the mixed and large kernels)
  uint32_t screen_occ = 5;    // register budget of the screens, as waves per SIMD (4 or 5)
  uint32_t screen_bpc = 16;   // grid caps of the grid-stride pair kernels, in blocks per CU
  uint32_t edge_bpc = 16;
  uint32_t mixed_bpc = 32;
  uint32_t large_bpc = 16;
  uint32_t dw_bpc = 16;       // distance-join list kernels
  uint32_t trav_block = 256;  // traversal block size (64, 128 or 256 threads)
  uint32_t ovf_lds = 1;       // overflow pairs with lists of at most 448 edges keep them in LDS (0: global slabs; 2: as 1, and the
                              // list classes' overflow pairs within 256 edges are refined on their own chain, no host read first)
  uint32_t fork = 1;          // mixed and large classes on a second stream beside the vertex-list classes (0: in line)
  uint32_t mixed = 1;         // mixed class (0: its pairs are large-class pairs)
  uint32_t large_cap = 96;    // large-class LDS list capacity (0: every large pair goes to the overflow kernel)
  uint32_t size_sort = 1;     // join: within a class, pairs ordered by size in buckets 2^(v-1) times the base width (0: candidate order)
  uint32_t dw_size_sort = 0;  // the same for dwithin
};

struct Tuning {
  static constexpr int MAX_BUCKETS = 16;
  LaunchShape buckets[MAX_BUCKETS];   // [0, nb): ascending np_min, the first at 0 (fixed storage: a context built
  int nb = 1;                         // inside the no-allocation entry point allocates nothing for it)
  std::string source = "built-in";
  const LaunchShape& pick(uint64_t np) const;
  // text form: '#' comments; one line per bucket, "bucket np>=N key=value ..." (keys as in LaunchShape; a key left
  // out keeps its built-in value). Throws std::invalid_argument naming the line on malformed input.
  static Tuning parse(const std::string& text, const std::string& source);
  std::string text() const;
};

// the table of a device: the file named by GSJ_TUNING when set, else tuning/<arch>.txt beside the library's lib/
// directory when it exists, else the built-in shapes
Tuning load_tuning(const std::string& arch);

}  // namespace gsj
