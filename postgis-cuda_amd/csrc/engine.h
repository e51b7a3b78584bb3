// Engine internals shared by the .hip translation units: device layer layout, BVH layout, context, primitives.
#pragma once
#include "common.h"
#include "geom.h"
#include "tune.h"
#include <functional>
#include <memory>
#include <string>

namespace gsj {

// ------------------------------------------------------------------------------------------- device layouts
// Polygon layer resident in HBM (SoA). Sizes at 162M parcels x ~10 vertices: xy 26 GB, per-geometry arrays
// ~60 B x n, BVH 32 B x (2n - 1).
struct DPolys {
  uint32_t n = 0, nr = 0;
  uint64_t nv = 0;
  const double2* origin = nullptr;   // [n] bbox min (FP64)
  const double2* ext = nullptr;      // [n] bbox extent
  const uint32_t* gro = nullptr;     // [n+1] geometry -> first ring
  const uint64_t* rvo = nullptr;     // [nr+1] ring -> first vertex
  const double2* xy = nullptr;       // [nv] v - origin of its geometry
  const float4* bb = nullptr;        // [n] outward-rounded float box (minx, miny, maxx, maxy), world coords
  const uint32_t* ginfo = nullptr;   // [n] stored vertices (low 16 bits, saturated) | rings << 16, for work classes
  const uint32_t* parent = nullptr;  // [n] piece -> parent, nullptr if not subdivided
  const uint32_t* synb = nullptr;    // [(nv+31)/32] bit v: edge v -> v+1 was made by subdivision; nullptr if none
  const uint8_t* vflag = nullptr;    // [nv] VF_* bits of vertex v (present exactly when synb is)
  const double* blen = nullptr;      // [nparent] real boundary length per parent
  const uint32_t* pfirst = nullptr;  // [nparent+1] first piece of each parent (subdivided layers only)
  uint32_t nparent = 0;
  // edge chunks of geometries with more than CHUNK edges: chunk k of geometry g covers the vertex pairs
  // (i, i+1), i in [v0 + CHUNK k, v0 + CHUNK k + CHUNK), v0 its first vertex; cbb[cfirst[g] + k] bounds those
  // vertices (origin-relative, rounded outward to float). cfirst[g] == NO_CHUNKS for smaller geometries.
  const uint32_t* cfirst = nullptr;  // [n]
  const float4* cbb = nullptr;       // [nchunks]
  // packed records of the pieces the vertex-list classes can take (at most PACK_VCAP stored vertices and PACK_RCAP
  // rings), in 16-byte units: [origin][extent][header: stored vertices, rings, ring starts][vertices as in xy][one
  // VF_* byte per vertex, padded], so that a pair kernel reads one contiguous record per piece instead of four
  // dependent arrays (geometry -> rings -> vertices, flags)
  const uint4* pk = nullptr;         // [records]
  const uint32_t* pk_off = nullptr;  // [n] first unit of the piece's record, NO_PACK for the others
};
constexpr int CHUNK = 64;
// per-vertex flags of a subdivided layer (ring-aware, so the pair kernels need no neighbour lookups): VF_SYN edge
// v -> v+1 is synthetic; VF_CUT0 / VF_CUT1 the ring's previous / next edge is (wrapping around the closed ring);
// all 0 at a ring's last vertex
constexpr uint8_t VF_SYN = 2, VF_CUT0 = 4, VF_CUT1 = 8;
constexpr uint32_t NO_CHUNKS = 0xFFFFFFFFu;
constexpr uint32_t PACK_VCAP = 64, PACK_RCAP = 8, NO_PACK = 0xFFFFFFFFu;
// 16-byte units of a packed record: origin, extent, header, the vertices, their flag bytes
__host__ __device__ constexpr uint64_t pack_units(uint64_t nv) { return 3 + nv + (nv + 15) / 16; }

constexpr uint32_t PROBE_TILE = 1024;   // probes per tile box (window membership)

struct DPoints {
  uint32_t n = 0;
  const double2* xy = nullptr;   // [n] world coordinates
  const float4* bb = nullptr;    // [n] outward-rounded float point boxes
  const uint32_t* order = nullptr;   // [n] point ids in Morton order of their boxes (the probe walk order)
  const float4* tiles = nullptr;     // [(n + PROBE_TILE - 1) / PROBE_TILE] box of each run of PROBE_TILE walk positions
};

// LBVH node in depth-first (preorder) layout: the left child of an internal node is the next node; `skip` is the
// first node after the subtree (== node count: done); `prim` is the primitive of a leaf, -1 for internal nodes.
struct alignas(16) BNode {
  float4 box;
  int32_t skip;
  int32_t prim;
  int32_t pad0, pad1;
};

struct DBvh {
  uint32_t n = 0;           // primitives
  uint32_t nnodes = 0;      // 2n - 1
  const BNode* nodes = nullptr;
  const uint32_t* order = nullptr;   // [n] primitives in Morton (leaf) order
  const float4* tiles = nullptr;     // [(n + PROBE_TILE - 1) / PROBE_TILE] box of each run of PROBE_TILE leaves in that order
};

// ------------------------------------------------------------------------------------------- host objects
struct Context;

struct Layer {
  Context* ctx = nullptr;
  bool points = false;
  uint64_t n = 0, nr = 0, nv = 0;
  uint32_t nparent = 0;
  // owned device memory
  void* mem = nullptr;          // one allocation holding every layer array
  size_t mem_bytes = 0;
  void* bvh_mem = nullptr;
  size_t bvh_bytes = 0;
  DPolys poly;
  DPoints pts;
  DBvh bvh;
  bool has_bvh = false;
  std::vector<uint32_t> h_parent;   // host copy: keeps a parent's pieces in one probe batch
  std::vector<float4> h_tiles;      // host copy of the probe tile boxes (scratch estimates of windowed requests)
  double build_ms = 0;              // device time of the last BVH build (HIP events around its kernels)
  ~Layer();
};

struct Context {
  int device = 0;
  hipStream_t stream = nullptr;
  hipStream_t side = nullptr;     // second stream (default priority): list classes beside the vertex-list classes
  std::mutex mu;                  // one call at a time per context
  std::string status;             // last status JSON
  // scratch, grow-only, reused across calls
  DevBuf s_counts, s_qoffs, s_offs, s_pairs_a, s_pairs_b, s_flag, s_area, s_ev, s_evb, s_tmp, s_tmp2, s_cls, s_out_a, s_out_b, s_out_area,
      s_out_code, s_out_flag, s_keys, s_keys2, s_vals, s_vals2, s_seg, s_heavy, s_heavy_edges, s_scal, s_qbb, s_uf, s_root, s_misc, s_win, s_qidx,
      s_win_off, s_edgeq, s_qlist, s_wtile, s_wtoff, s_wseg, s_slots, s_slots_ctl;
  // erosion pass 2 (it runs inside the join's on-device hook, beside the join's own scratch)
  DevBuf s_er_keep, s_er_off, s_er_pairs, s_er_flag, s_er_list, s_er_slab, s_er_out, s_er_need, s_er_pf;
  PinBuf h_pairs, h_area, h_code, h_where, h_scal, h_counts, h_seg, h_slots_ctl;
  int cu_count = 256;
  hipEvent_t ev[17] = {};   // kernel timing events (created with the context); ev[10], ev[11]: fork / join of `side`;
                            // ev[12 .. 15]: the traversal passes and the BVH build; ev[16]: the list classes' overflow
  bool borrowed = false;   // stream and events belong to another context (not destroyed with this one)
  Arena* dev_arena = nullptr;   // set: every scratch buffer comes from this arena (the no-allocation entry point)
  Tuning tune;                  // launch shapes of the pair kernels per launch size
  const Tuning* tune_from = nullptr;   // set: launch with another context's shapes instead
  const LaunchShape& shape(uint64_t np) const { return (tune_from ? *tune_from : tune).pick(np); }
  std::string tune_text;        // last text handed out by gsj_ctx_tuning
  template <class F> void for_each_devbuf(F f) {
    for (DevBuf* b : {&s_counts, &s_qoffs, &s_offs, &s_pairs_a, &s_pairs_b, &s_flag, &s_area, &s_ev, &s_evb, &s_tmp, &s_tmp2, &s_cls, &s_out_a,
                      &s_out_b, &s_out_area, &s_out_code, &s_out_flag, &s_keys, &s_keys2, &s_vals, &s_vals2, &s_seg, &s_heavy, &s_heavy_edges,
                      &s_scal, &s_qbb, &s_uf, &s_root, &s_misc, &s_win, &s_qidx, &s_win_off, &s_edgeq, &s_qlist, &s_wtile, &s_wtoff, &s_wseg,
                      &s_slots, &s_slots_ctl, &s_er_keep, &s_er_off, &s_er_pairs, &s_er_flag, &s_er_list, &s_er_slab, &s_er_out, &s_er_need,
                      &s_er_pf})
      f(*b);
  }
  template <class F> void for_each_pinbuf(F f) {
    for (PinBuf* b : {&h_pairs, &h_area, &h_code, &h_where, &h_scal, &h_counts, &h_seg, &h_slots_ctl}) f(*b);
  }
  ~Context();
};

// ------------------------------------------------------------------------------------------- primitives (prims.hip)
// exclusive prefix sum of n u32 values into u64 offsets; returns the total (synchronises)
uint64_t scan_u32_to_u64(Context& c, const uint32_t* in, uint64_t* out, uint64_t n);
// exclusive prefix sum without host sync: total written to *d_total
void scan_u32_to_u64_async(Context& c, const uint32_t* in, uint64_t* out, uint64_t n, uint64_t* d_total);
// stable LSD radix sort of (key, value) by key; only the digits in [bit0, bit1) are sorted; returns pointers to the
// sorted arrays (either the inputs or the alternates). probe: first read back which bits vary (a host synchronisation)
// and skip the digits that do not; without it every digit in the range is sorted and nothing synchronises
void radix_sort_pairs(Context& c, uint64_t*& keys, uint64_t*& keys_alt, uint32_t*& vals, uint32_t*& vals_alt, uint64_t n, int bit0,
                      int bit1, bool probe = true);
// OR and AND of all keys (for choosing digits); synchronises
void key_bits(Context& c, const uint64_t* keys, uint64_t n, uint64_t& or_all, uint64_t& and_all);

// ------------------------------------------------------------------------------------------- layers / bvh
void layer_upload_polys(Layer& L, const uint64_t* gro, const uint64_t* rvo, const double* xy, uint64_t n, uint64_t nr, uint64_t nv,
                        const uint32_t* parent, const uint8_t* syn);
void layer_upload_points(Layer& L, const double* xy, uint64_t n);
void bvh_build(Layer& L);
// ids 0 .. n-1 of the boxes bb in the order of their centres' 64-bit Morton keys (stable), into order[n]
void morton_order(Context& c, const float4* bb, uint32_t n, uint32_t* order);
// tiles[t] = union of the boxes bb[order[pos]] for walk positions pos in [PROBE_TILE t, PROBE_TILE (t + 1)) (async)
void tile_boxes(Context& c, const float4* bb, const uint32_t* order, uint32_t n, float4* tiles);

// ------------------------------------------------------------------------------------------- traversal
// count pass (write == false) or write pass over probe boxes qbb[qbase .. qbase+nq)
// Candidates of the probes at positions qbase .. qbase+nq (probe id qidx[pos], or pos): traverse_count walks the tree
// once per probe, counting into counts[] and keeping the first candidates in slots[] (probes with more are counted in
// *nover); after the scan of the counts into offs, traverse_write writes every candidate at its offset (SoA), walking
// again only for the probes whose candidates did not fit (any_over). box_by_pos: the search box of position p is
// qbb[p] instead of the box of its probe.
void traverse_count(Context& c, const float4* qbb, const uint32_t* qidx, uint32_t qbase, uint32_t nq, const DBvh& t, bool upper, uint32_t* counts,
                    uint32_t* slots, uint32_t* nover, bool box_by_pos);
void traverse_write(Context& c, const float4* qbb, const uint32_t* qidx, uint32_t qbase, uint32_t nq, const DBvh& t, bool upper,
                    const uint32_t* counts, const uint64_t* offs, const uint32_t* slots, bool any_over, uint32_t* out_a, uint32_t* out_b,
                    bool box_by_pos);

// The probe side of a candidate search: traversal boxes qbb[0 .. n) (probe index = box index), the probes' own boxes
// (the window test; differ from qbb when the search box is widened), and the parent structure of a subdivided layer
// (host h_parent keeps a parent's pieces in one batch; device pfirst lets a window take whole parents).
struct Probes {
  const float4* qbb = nullptr;
  const float4* own_bb = nullptr;
  const uint32_t* order = nullptr;   // probe walk order (ids at positions 0 .. n-1), or id order when null
  const float4* tiles = nullptr;     // boxes of PROBE_TILE consecutive walk positions (with order), or null
  uint32_t n = 0;
  const std::vector<uint32_t>* h_parent = nullptr;
  const uint32_t* pfirst = nullptr;
  uint32_t nparent = 0;
};
Probes probes_of(const Layer& L, const float4* qbb = nullptr);

// Candidate pairs of the probes against tree t, in probe batches of at most o.batch_q probes that keep a parent's
// pieces together and sub-batches of at most o.max_pairs candidates (engine/engine.cuh:903-937); with o.has_window
// only the probes of parents meeting the window. fn(pa, pb, np) runs on every non-empty sub-batch with the pairs on
// the device and the stream synchronised. Adds traverse_ms, batches and candidates to s; polls o.cancel.
// on_members(list, n, mstart), when given, runs once before the walk with the probes that will be walked (list ==
// nullptr: all n probes in id order) and may return the search boxes to use instead of pr.qbb (indexed by probe id;
// by member position in for_each_candidate_multi, whose mstart gives the windows' member boundaries).
using PairsFn = std::function<void(uint32_t* pa, uint32_t* pb, uint64_t np)>;
using MembersFn = std::function<const float4*(const uint32_t* list, uint32_t n, const uint64_t* mstart)>;
void for_each_candidate_batch(Context& c, const Probes& pr, const DBvh& t, bool upper, const gsj_join_opts& o, gsj_join_stats& s,
                              const PairsFn& fn, const MembersFn& on_members = nullptr);

// Several windowed requests over the same probes in one candidate search (the batched served path). Window r's
// members come after window r-1's, each in walk order, so request r's candidates are a contiguous run of every
// sub-batch, in the order its own search would produce them. Needs probes with tile boxes (a layer without parents).
constexpr uint32_t MAX_WINDOWS = 32;
struct WindowSet {   // passed to kernels by value
  uint32_t k = 0;
  double w[MAX_WINDOWS][4];
};
// candidate-index boundaries of the k requests inside one sub-batch: request r owns pairs [lo[r], lo[r + 1])
struct SegBounds {
  uint32_t k = 0;
  const uint64_t* h = nullptr;   // host, k + 1 entries
  const uint64_t* d = nullptr;   // device copy
};
using SegPairsFn = std::function<void(uint32_t* pa, uint32_t* pb, uint64_t np, const SegBounds& sb)>;
bool probes_batchable(const Probes& pr);
void for_each_candidate_multi(Context& c, const Probes& pr, const DBvh& t, bool upper, const gsj_join_opts& o, const WindowSet& ws,
                              gsj_join_stats& s, const SegPairsFn& fn, const MembersFn& on_members = nullptr);
// h_rlo[0 .. sb.k] = row boundaries of the requests of sub-batch sb after an order-preserving compaction whose
// exclusive offsets are offs[0 .. np) (nrows rows); synchronises
void seg_rows(Context& c, const SegBounds& sb, const uint64_t* offs, uint64_t np, uint64_t nrows, uint64_t* h_rlo);
// the same on the device: out[r] = boundaries after the compaction of lo[0 .. k] (async)
void seg_bounds_dev(Context& c, const uint64_t* lo, uint32_t k, const uint64_t* offs, uint64_t np, uint64_t nrows, uint64_t* out);
// region i of 8 in the small per-context buffer of window and request boundaries (MAX_WINDOWS + 1 entries each):
// 0 tile starts, 1 sub-batch candidate bounds, 2 member starts, 3 row bounds, 4-7 free for a caller's stages
uint64_t* wseg(Context& c, int region);

// ------------------------------------------------------------------------------------------- refinement
struct RefineParams {   // engine/engine.cuh:517-518
  double tol_rel = 1e-12, eps_par = 1e-7, tol_abs = 1e-11;
  int want_area = 1, want_relate = 0, snap_half_ulp = 0;
};
struct RefineOut {
  uint8_t* isect;   // [np] 0 / 1 real contact or containment / 2 area-only piece overlap
  double* area;     // [np] or nullptr
  double4* ev;      // [np] (A in, A on, B in, B on) or nullptr
  uint32_t* evb;    // [np] evidence bits or nullptr
};
// refine np candidate pairs (pa[i], pb[i]); returns the number of pairs that needed global-memory edge lists.
// Kernel durations and per-class pair counts are added to *st (if given).
uint64_t refine_pairs(Context& c, const DPolys& P, const DPolys& R, const uint32_t* pa, const uint32_t* pb, uint64_t np, const RefineParams& rp,
                      const RefineOut& out, gsj_join_stats* st = nullptr);

// ------------------------------------------------------------------------------------------- join driver
struct JoinHooks {
  // consumer of each finished sub-batch while on the device (rows: 2n interleaved parent ids, area or nullptr, n)
  std::function<void(const uint32_t*, const double*, uint64_t)> on_device;
  bool device_only = false;
  // candidate filter before refinement (JoinOpts::prefilter, engine/engine.cuh:812, 935): compacts the sub-batch's
  // pairs in place on the device and returns how many are kept
  std::function<uint64_t(uint32_t* pa, uint32_t* pb, uint64_t np)> prefilter;
  // the probes about to be walked (for_each_candidate_batch's on_members)
  MembersFn on_members;
};
void join_run(Context& c, const Layer& P, const Layer& R, const gsj_join_opts& o, gsj_rows_cb cb, void* user, gsj_join_stats* st,
              const JoinHooks* hooks);
// windowed joins of the same layers and refinement options in one pass (see for_each_candidate_multi); each
// request's rows, in order, and its candidates and hits are those of its own join_run; timing fields are the batch's
struct BatchJoin {
  const gsj_join_opts* o = nullptr;
  gsj_rows_cb cb = nullptr;
  void* user = nullptr;
  gsj_join_stats* st = nullptr;
  int rc = GSJ_OK;   // GSJ_E_CANCEL when its callback asked to stop
};
bool join_batchable(const Layer& P, const Layer& R, const gsj_join_opts& o);
// compaction of refined candidates of layers without parents into rows in candidate order (area / relate as rp asks);
// *offs = the exclusive offsets of the kept candidates
uint64_t compact_single_rows(Context& c, const Layer& P, const Layer& R, const uint32_t* pa, const uint32_t* pb, uint64_t np, const RefineParams& rp,
                             const RefineOut& ro, uint32_t* orow, double* oarea, const uint64_t** offs);
bool join_compatible(const gsj_join_opts& a, const gsj_join_opts& b);
void join_batch_run(Context& c, const Layer& P, const Layer& R, std::vector<BatchJoin>& reqs);
void points_join_run(Context& c, const Layer& pts, const Layer& R, const gsj_join_opts& o, gsj_points_cb cb, void* user,
                     gsj_join_stats* st);
void cluster_run(Context& c, const Layer& P, const gsj_join_opts& o, uint32_t* root, uint64_t nroot, gsj_join_stats* st);
void cluster_rows_run(Context& c, const Layer& P, const gsj_join_opts& o, gsj_rows_cb cb, void* user, gsj_join_stats* st);

// ------------------------------------------------------------------------------------------- distance joins (dwithin.hip)
void dwithin_run(Context& c, const Layer& P, const Layer& R, double d, bool same, bool report, const gsj_join_opts& o, gsj_dist_cb cb, void* user,
                 gsj_join_stats* st);
void points_dwithin_run(Context& c, const Layer& pts, const Layer& R, double d, bool report, const gsj_join_opts& o, gsj_dist_cb cb, void* user,
                        gsj_join_stats* st);
// windowed dwithin requests of the same layers, `same`, `report`, upper / batch_q / max_pairs, each with its own
// distance, in one pass (as join_batch_run)
struct BatchDist {
  const gsj_join_opts* o = nullptr;
  double d = 0;
  gsj_dist_cb cb = nullptr;
  void* user = nullptr;
  gsj_join_stats* st = nullptr;
  int rc = GSJ_OK;
};
bool dwithin_batchable(const Layer& P, const Layer& R, const gsj_join_opts& o);
void dwithin_batch_run(Context& c, const Layer& P, const Layer& R, bool same, bool report, std::vector<BatchDist>& reqs);
// eroded intersects (erode.hip, dwithin.hip)
// Which pairs of a parent with itself (same layers) the eroded intersects tests: with `listed`, the parents in the
// ascending list[0, n) (none when n is 0); otherwise those whose flag is set (flags: one per parent of P), or every
// one when flags is null. SelfNeedHost is the caller's (host memory), SelfNeed its copy on the device.
struct SelfNeedHost {
  const uint8_t* flags = nullptr;
  const uint32_t* list = nullptr;
  uint64_t n = 0;
  bool listed = false;
};
struct SelfNeed {
  const uint8_t* flags;
  const uint32_t* list;
  uint32_t n;
  int listed;
};
void erode_run(Context& c, const Layer& P, const Layer& R, double ra, double rb, bool same, const SelfNeedHost& need, const gsj_join_opts& o,
               gsj_dist_cb cb, void* user, gsj_join_stats* st);
void points_erode_run(Context& c, const Layer& pts, const Layer& R, double r, const gsj_join_opts& o, gsj_dist_cb cb, void* user, gsj_join_stats* st);
// windowed eroded-intersects requests of the same layers, `same`, upper / batch_q / max_pairs / snap_half_ulp, no
// need_self, each with its own radii, in one pass (as join_batch_run)
struct BatchErode {
  const gsj_join_opts* o = nullptr;
  double ra = 0, rb = 0;
  gsj_dist_cb cb = nullptr;
  void* user = nullptr;
  gsj_join_stats* st = nullptr;
  int rc = GSJ_OK;
};
bool erode_batchable(const Layer& P, const Layer& R, const gsj_join_opts& o, const SelfNeedHost& need);
// no_self: no parent's pair with itself is tested (the requests' empty lists); otherwise every one
void erode_batch_run(Context& c, const Layer& P, const Layer& R, bool same, bool no_self, std::vector<BatchErode>& reqs);

}  // namespace gsj
