/* gsj.h — C ABI of the MI355X spatial join engine (libgsj.so).
 *
 * This is the boundary a host program binds to (the gpujoind daemon in this repository, or any FFI). Every
 * entry point replaces one engine entry point of the reference (landstats-us/postgis-cuda), whose engine is a
 * set of C++ templates in engine/engine.cuh and engine/buffer.cuh; the reference file:line is given per function.
 * All geometry is FP64. Plain pointers and sizes only; no device pointers cross this boundary (callbacks receive
 * host buffers that stay valid until the callback returns).
 *
 * Errors: every int-returning function returns GSJ_OK (0) or a negative GSJ_E_* code and sets a thread-local
 * message readable with gsj_last_error(). Nothing in the library calls exit() (the reference exits on any CUDA
 * error or overflow, engine/common.h:13-14, engine/engine.cuh:43, 927).
 *
 * Threading: one context owns one HIP stream on one device; calls on one context are serialised internally (the
 * reference serialises all GPU work under one mutex, daemon/gpujoind.cu:35). Layers belong to the context that
 * created them.
 */
#ifndef GSJ_H
#define GSJ_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------ status */
#define GSJ_OK 0
#define GSJ_E_ARG -1      /* invalid argument (null, size out of range, wrong layer kind, malformed offsets or file contents) */
#define GSJ_E_HIP -2      /* HIP runtime error (message carries hipGetErrorString) */
#define GSJ_E_NOMEM -3    /* device or host allocation failed */
#define GSJ_E_LIMIT -4    /* a documented limit was exceeded (geometries >= 2^31, rings >= 2^32, ...) */
#define GSJ_E_CANCEL -5   /* the cancel flag was raised; partial results were delivered */
#define GSJ_E_IO -6       /* file could not be opened, read or written */
#define GSJ_E_INTERNAL -7

/* message of the last failing call on this thread ("" if none) */
const char* gsj_last_error(void);
/* library version string */
const char* gsj_version(void);

/* ------------------------------------------------------------------ context */
typedef struct gsj_ctx gsj_ctx;
typedef struct gsj_layer gsj_layer;

/* Create a context on HIP device `device` (its own stream, default priority). Its launch-shape table comes from the
 * file named by the environment variable GSJ_TUNING, else from tuning/<arch>.txt beside the library's directory
 * (postgis-cuda_amd/tuning/gfx950.txt in-tree), else the built-in shapes; a malformed file fails with GSJ_E_ARG. */
int gsj_ctx_create(int device, gsj_ctx** out);
void gsj_ctx_destroy(gsj_ctx* ctx);
/* JSON description of the device and memory (reference op_status, daemon/gpujoind.cu:496-504); the returned
 * string is owned by the context and valid until the next call on it */
const char* gsj_ctx_status_json(gsj_ctx* ctx);
/* Release the context's cached scratch buffers (they are grow-only and reused across calls). */
int gsj_ctx_trim(gsj_ctx* ctx);
/* Launch shapes of the pair kernels per launch size, as text: '#' comments and one line per size bucket,
 * "bucket np>=N key=value ..." with keys tcap (0..16), mcap (0 or 33..64), screen_occ (4, 5), screen_bpc, edge_bpc,
 * mixed_bpc, large_bpc, dw_bpc (blocks per CU, 1..64), trav_block (64, 128, 256) and ovf_lds (1: overflow pairs
 * with lists of at most 448 edges per side keep them in LDS; 0: in global memory; 2: as 1, and the mixed and large
 * classes' overflow pairs with lists within 256 edges are refined right behind them, on their stream, before the
 * host reads the queue) and fork (1: the mixed and large
 * classes run on a second stream of the context beside the other classes; 0: in line), mixed (0: no mixed class)
 * large_cap (0..96: the large class's list capacity; 0 sends every large pair to the overflow kernel) and size_sort
 * and dw_size_sort (join and dwithin, 0..4: v > 0 puts pairs of similar size side by side within a class, in size
 * buckets 2^(v-1) times the class's base width; 0 keeps candidate order); buckets ascend from np>=0.
 * Replaces the reference's fixed launch shapes (engine/engine.cuh:945-951, engine/buffer.cuh:946-947); every shape
 * gives the same rows. set: GSJ_E_ARG on malformed text (the table is unchanged); applies to later calls and to
 * schedulers created afterwards. get: the table in effect; the string is owned by the context, valid until the next
 * call of gsj_ctx_tuning on it. */
int gsj_ctx_set_tuning(gsj_ctx* ctx, const char* text);
const char* gsj_ctx_tuning(gsj_ctx* ctx);

/* ------------------------------------------------------------------ layers
 * Polygon layer from the flat .gset arrays (engine/gset.h:5-11): gro[n+1] geometry -> first ring, rvo[nr+1] ring
 * -> first vertex (rings closed, shells CCW, holes CW), xy[2 nv]. Optional parent[n] (piece -> parent geometry,
 * non-decreasing) and syn[nv] (edge v -> v+1 was created by subdivision) describe a subdivided layer.
 * Replaces upload_set (engine/engine.cuh:42-89): per-geometry origin, exact origin-relative coordinates,
 * outward-rounded float boxes, real boundary length per parent and the synthetic-edge bitset are computed on the
 * device. Limits: n < 2^31, nr < 2^32 (engine/engine.cuh:43). */
int gsj_layer_create_polys(gsj_ctx* ctx, const uint64_t* gro, const uint64_t* rvo, const double* xy, uint64_t n, uint64_t nr,
                           uint64_t nv, const uint32_t* parent, const uint8_t* syn, gsj_layer** out);
/* Point layer (engine/engine.cuh:1068-1076, upload_points): xy[2n]. */
int gsj_layer_create_points(gsj_ctx* ctx, const double* xy, uint64_t n, gsj_layer** out);
void gsj_layer_free(gsj_layer* layer);
/* Build (or rebuild) the layer's LBVH (engine/engine.cuh:197-228, build_bvh): 64-bit Morton keys, on-device
 * radix sort, Karras hierarchy, internal boxes from leaf ranges, nodes stored in depth-first order with skip links. */
int gsj_bvh_build(gsj_layer* layer);

typedef struct {
  uint64_t n, nr, nv;      /* geometries (pieces), rings, vertices */
  uint32_t nparent;        /* distinct parent geometries (n if not subdivided) */
  int is_points;           /* 1 for a point layer */
  int has_bvh;
  uint64_t device_bytes;   /* layer + BVH */
  double bvh_ms;           /* device time of the last BVH build (HIP events around its kernels) */
} gsj_layer_info;
int gsj_layer_info_get(const gsj_layer* layer, gsj_layer_info* info);
/* Copy the prepared device arrays back (for verification): origin[2n], ext[2n], bb[4n] (float), q[2nv]
 * (origin-relative xy), blen[nparent]. Any pointer may be NULL. */
int gsj_layer_export(const gsj_layer* layer, double* origin, double* ext, float* bb, double* q, double* blen);

/* ------------------------------------------------------------------ joins */
typedef struct {
  int want_area;           /* intersection area per row (engine/engine.cuh:517 RefineParams::want_area) */
  int want_relate;         /* DE-9IM code per row (RefineParams::want_relate) */
  int upper;               /* self-join: each unordered pair once, b >= a (JoinOpts::upper, engine.cuh:810) */
  int snap_half_ulp;       /* RefineParams::snap_half_ulp */
  uint32_t batch_q;        /* probe geometries per traversal batch; 0 = default 4M (JoinOpts::batch_q) */
  uint64_t max_pairs;      /* candidate pairs per refine launch; 0 = default 128M (JoinOpts::max_pairs) */
  int no_agg;              /* debug: report piece pairs instead of parent pairs (JoinOpts::no_agg) */
  const volatile int* cancel;   /* polled between batches (JoinOpts::cancel); may be NULL */
  /* Query window (no counterpart in the reference; the windowed serving requests use it): with has_window set,
   * only probe-side parents with a piece whose float box meets the closed rectangle window = {x0, y0, x1, y1}
   * (layer coordinates) are probed, so the rows are exactly the unwindowed rows whose probe parent meets it.
   * Joins, point joins and the distance / eroded joins honour it; gsj_cluster rejects it (gsj_cluster_rows takes it). */
  int has_window;
  double window[4];
} gsj_join_opts;

typedef struct {
  double traverse_ms, refine_ms, compact_ms, copy_ms;   /* JoinStats, engine/engine.cuh:814 */
  uint64_t candidates;     /* candidate pairs (refined pairs) */
  uint64_t hits;           /* result rows delivered (parent pairs) */
  uint64_t heavy;          /* pairs whose local edge lists exceeded on-chip capacity */
  uint64_t batches;
  /* pair kernels, timed with HIP events on the context's stream, summed durations and pairs handled by each. Polygon
   * joins: small = 16-lane class (both polygons <= 32 stored vertices and <= 8 rings), medium = 32-lane class (not
   * small, both <= 64 stored vertices and <= 8 rings); these two classes are screened first (contact, containment;
   * k_screen_ms, every small and medium pair) and k_small_ms / k_medium_ms time their edge-contribution kernels over
   * the n_small_edge / n_medium_edge pairs that need one; large = a wave per pair; over = the global-memory overflow
   * pass; mixed = 16-lane groups with edge lists for pairs of a small polygon (<= 32 vertices) and a larger one.
   * Distance joins: tiny, small, medium, mixed and large list classes. Point joins: the per-pair kernel as small. */
  double k_small_ms, k_large_ms, k_over_ms;
  uint64_t n_small, n_large, n_over;
  uint64_t launches;       /* refinement launches (sub-batches) */
  double k_medium_ms;
  uint64_t n_medium;
  double k_screen_ms;
  uint64_t n_small_edge, n_medium_edge;
  double k_mixed_ms;
  uint64_t n_mixed;
  /* tiny = 8-lane class (both polygons <= 16 stored vertices and <= 8 rings), screened with the small and medium
   * classes; k_tiny_ms times its edge-contribution kernel over n_tiny_edge pairs (distance joins: its list kernel) */
  double k_tiny_ms;
  uint64_t n_tiny, n_tiny_edge;
  double k_traverse_ms;    /* the candidate search's walks (count and write passes), HIP-event time */
} gsj_join_stats;

/* Row callback: n rows; pairs[2i], pairs[2i+1] = (a, b) parent ids (probe side, build side); area[i] or NULL;
 * code[i] (DE-9IM, 2 bits per cell II IB IE BI BB BE EI EB EE, 0=F 1='0' 2='1' 3='2') or NULL. Buffers are valid
 * until the callback returns. Return nonzero to cancel. */
typedef int (*gsj_rows_cb)(void* user, const uint32_t* pairs, const double* area, const uint32_t* code, size_t n);

/* ST_Intersects [+ ST_Area(ST_Intersection)] [+ ST_Relate] of probe layer P against build layer R (R needs a BVH).
 * Replaces join_stream (engine/engine.cuh:895-1061). Rows are distinct parent pairs. */
int gsj_join(gsj_ctx* ctx, const gsj_layer* P, const gsj_layer* R, const gsj_join_opts* opts, gsj_rows_cb cb, void* user,
             gsj_join_stats* stats);
/* gsj_join without any allocation (no counterpart in the reference): for callers that cannot allocate (code under HIP
 * graph capture, tight serving loops, harnesses that forbid it). The same kernels and rows as gsj_join; every scratch
 * buffer is a slice of a fixed 128 MiB device array per GPU (held by each GPU the library runs on until the process
 * ends) and a 16 MiB host staging region per GPU set up by gsj_ctx_create. Limits: max_pairs <= 131072 and
 * batch_q <= 32768 (0 selects these; larger values fail with GSJ_E_LIMIT), no query window (GSJ_E_ARG), and a request
 * whose sub-batch needs more scratch than the array holds fails with GSJ_E_LIMIT. Calls on one GPU share the array and
 * are serialised by a per-GPU lock. */
int gsj_join_noalloc(gsj_ctx* ctx, const gsj_layer* P, const gsj_layer* R, const gsj_join_opts* opts, gsj_rows_cb cb, void* user,
                     gsj_join_stats* stats);

/* Points in polygons, boundary counts (replaces join_points_stream, engine/engine.cuh:1112-1178). Callback rows:
 * pairs (point index, parent polygon), where[i] = 1 interior, 2 boundary. */
typedef int (*gsj_points_cb)(void* user, const uint32_t* pairs, const uint8_t* where, size_t n);
int gsj_points_in_polygons(gsj_ctx* ctx, const gsj_layer* pts, const gsj_layer* R, const gsj_join_opts* opts, gsj_points_cb cb,
                           void* user, gsj_join_stats* stats);

/* ST_ClusterIntersecting over the parents of P (replaces cluster_intersecting, engine/engine.cuh:1216-1234):
 * root[g] = smallest parent index of g's connected component, for g in [0, nparent). P needs a BVH. */
int gsj_cluster(gsj_ctx* ctx, const gsj_layer* P, const gsj_join_opts* opts, uint32_t* root, uint64_t nroot, gsj_join_stats* stats);
/* The same as rows for the served path: pairs[2i] = parent g, pairs[2i+1] = smallest parent index of g's component,
 * in ascending g, through cb (area and code NULL). With opts->has_window: cluster_intersecting of the sub-layer of
 * the parents whose box meets the window (components over the intersecting pairs of two such parents), rows for those
 * parents only; needs a layer without subdivided parents. No counterpart in the reference (cluster_intersecting,
 * engine/engine.cuh:1216-1234, has no window; the windowed serving stream asks for one). */
int gsj_cluster_rows(gsj_ctx* ctx, const gsj_layer* P, const gsj_join_opts* opts, gsj_rows_cb cb, void* user, gsj_join_stats* stats);

/* ------------------------------------------------------------------ distance joins (lon/lat, metres, WGS84)
 * Row callback of the distance joins: pairs[2i], pairs[2i+1] = (a, b) (polygon joins: parent ids of P and R; point
 * joins: point index, parent polygon id); dist[i] = the geography distance in metres when report_distance was set
 * (0 when the geometries intersect or one contains the other), else NULL. Return nonzero to cancel. */
typedef int (*gsj_dist_cb)(void* user, const uint32_t* pairs, const double* dist, size_t n);

/* ST_DWithin(a::geography, b::geography, d) of polygon layers (replaces buffer_join_stream in DWITHIN mode,
 * engine/buffer.cuh:892-991, 994-998; pair test dw_pair, :185-237). P's probe boxes are widened by d metres; R needs
 * a BVH. `same`: P and R are the same layer (pieces of one parent are within any distance of each other; pair with
 * opts->upper = 1 for each unordered pair once). report_distance: new output (the reference reports NaN,
 * daemon/gpujoind.cu:468). Only opts->upper, batch_q, max_pairs and cancel are used. */
int gsj_dwithin(gsj_ctx* ctx, const gsj_layer* P, const gsj_layer* R, double d, int same, int report_distance, const gsj_join_opts* opts,
                gsj_dist_cb cb, void* user, gsj_join_stats* stats);
/* ST_DWithin of points against polygons (replaces points_buffer_join_stream in DWITHIN mode, engine/buffer.cuh:
 * 1137-1229; per pair k_pt_dwithin, :1018-1049): rows (point index, parent polygon id), inside or within d of the
 * real boundary. */
int gsj_points_dwithin(gsj_ctx* ctx, const gsj_layer* pts, const gsj_layer* R, double d, int report_distance, const gsj_join_opts* opts,
                       gsj_dist_cb cb, void* user, gsj_join_stats* stats);

/* ST_Intersects(ST_Buffer(a::geography, -ra)::geometry, ST_Buffer(b::geography, -rb)::geometry) of polygon layers
 * without building buffers (replaces buffer_join_stream in ERODE mode, engine/buffer.cuh:779-889; pair test er_pair,
 * :502-588). ra, rb >= 0 metres (0 = that side not buffered). Pass 1 is the intersects + area join; parent pairs whose
 * overlap can hold a disk of radius min(ra, rb) are tested on their eroded boundaries. `same` / opts->upper as for
 * gsj_dwithin; need_self (host, one flag per parent of P, may be NULL): with `same`, a parent's pair with itself is
 * tested only where its flag is set (daemon/gpujoind.cu:438-443). Rows: parent pairs (dist NULL). */
int gsj_intersects_buffered(gsj_ctx* ctx, const gsj_layer* P, const gsj_layer* R, double ra, double rb, int same, const uint8_t* need_self,
                            const gsj_join_opts* opts, gsj_dist_cb cb, void* user, gsj_join_stats* stats);
/* Points inside ST_Buffer(polygon::geography, -r) (replaces points_buffer_join_stream in ERODE mode,
 * engine/buffer.cuh:1137-1229; per pair k_pt_erode, :1080-1128): rows (point index, parent polygon) where the point is
 * inside some piece and within r of no real boundary. */
int gsj_points_intersects_buffered(gsj_ctx* ctx, const gsj_layer* pts, const gsj_layer* R, double r, const gsj_join_opts* opts, gsj_dist_cb cb,
                                   void* user, gsj_join_stats* stats);

/* ------------------------------------------------------------------ served path: continuous batching
 * No counterpart function in the reference, which runs one request at a time under a global GPU mutex
 * (daemon/gpujoind.cu:35, 189-206): the reference's entry points above, run through a scheduler instead.
 * A scheduler owns `lanes` execution lanes on the context's device (one HIP stream and one scratch pool each) and
 * one worker thread per lane. Submitted requests queue in arrival order; the head of the queue is admitted as soon
 * as a lane is free and its estimated scratch fits the budget (a request runs alone when nothing else is admitted),
 * so requests that arrive while others are on the GPU overlap with them on the device. A request's rows are exactly
 * those of the same call made directly on the context. Callbacks run on the lane's worker thread; one request's
 * callbacks never run concurrently. Layers (of the scheduler's context) are pinned until their requests finish. */
typedef struct gsj_sched gsj_sched;

#define GSJ_REQ_JOIN 1                /* gsj_join: rows_cb */
#define GSJ_REQ_POINTS_IN_POLYGONS 2  /* gsj_points_in_polygons: points_cb */
#define GSJ_REQ_CLUSTER 3             /* gsj_cluster: root / nroot */
#define GSJ_REQ_DWITHIN 4             /* gsj_dwithin: d, same, report_distance, dist_cb */
#define GSJ_REQ_POINTS_DWITHIN 5      /* gsj_points_dwithin: d, report_distance, dist_cb */
#define GSJ_REQ_INTERSECTS_BUFFERED 6 /* gsj_intersects_buffered: ra, rb, same, need_self, dist_cb */
#define GSJ_REQ_POINTS_INTERSECTS_BUFFERED 7  /* gsj_points_intersects_buffered: rb (the polygon radius), dist_cb */
#define GSJ_REQ_CLUSTER_ROWS 8        /* gsj_cluster_rows: rows_cb */

typedef struct {
  int kind;                       /* GSJ_REQ_* */
  const gsj_layer* P;             /* probe layer (points for the point requests) */
  const gsj_layer* R;             /* build layer (needs a BVH); unused by GSJ_REQ_CLUSTER */
  gsj_join_opts opts;             /* cancel: polled by the lane as in the direct call */
  double d;                       /* distance requests: metres */
  int same, report_distance;      /* distance requests */
  double ra, rb;                  /* buffered requests: erosion radii, metres */
  const uint8_t* need_self;       /* buffered requests: see gsj_intersects_buffered (host, must outlive the request) */
  gsj_rows_cb rows_cb;
  gsj_points_cb points_cb;
  gsj_dist_cb dist_cb;
  void* user;
  uint32_t* root;                 /* GSJ_REQ_CLUSTER output, nroot >= parents of P */
  uint64_t nroot;
  /* buffered requests, in place of need_self (which must then be NULL) when need_listed is set: a parent's pair with
     itself is tested only for the parents in need_list[0, n_need), ascending (host, must outlive the request; none
     when n_need is 0). A list costs 4 bytes per listed parent instead of one byte per parent of P, and requests with
     an empty list (or with neither) can share a batch. No counterpart in the reference: it rebuilds its flags for
     every request (daemon/gpujoind.cu:438-443). */
  int need_listed;
  const uint32_t* need_list;
  uint64_t n_need;
  /* called once on the lane's worker thread when the request has finished (its last callback has returned; before
     gsj_wait can return for it), whatever its status; may be NULL. Lets a caller that sends a request's rows on from
     another thread wait for rows and for the end of the request together. No counterpart in the reference, whose
     daemon writes rows from the thread that runs the request (daemon/gpujoind.cu:189-206). */
  void (*done_cb)(void* done_user);
  void* done_user;
} gsj_request;

typedef struct {
  double queue_ms;   /* submit -> admitted to a lane */
  double run_ms;     /* admitted -> finished (device work, copies and callbacks) */
  double total_ms;   /* submit -> finished */
  int lane;
  int batch;         /* requests that ran together in its pass (1 = alone; see gsj_sched_set_batch) */
} gsj_request_times;

/* lanes in [1, 16]; scratch_budget_bytes: admission budget over the estimated scratch of running requests (0 = no
 * limit besides the lanes) */
int gsj_sched_create(gsj_ctx* ctx, int lanes, uint64_t scratch_budget_bytes, gsj_sched** out);
/* waits for every submitted request, then stops the lanes and frees their scratch */
void gsj_sched_destroy(gsj_sched* s);
/* Batching of windowed requests: GSJ_REQ_JOIN, GSJ_REQ_DWITHIN and GSJ_REQ_INTERSECTS_BUFFERED with has_window, the
 * same kind, P and R, no cancel flag, a probe layer with a BVH and neither layer subdivided, and the same options
 * except the window (joins: want_area / want_relate / upper / snap_half_ulp / no_agg / batch_q / max_pairs; dwithin:
 * same / report_distance / upper / batch_q / max_pairs, any d; eroded intersects: same / upper / snap_half_ulp /
 * batch_q / max_pairs, no need_self, any ra / rb). A lane admitting one such request also takes up to
 * max_requests - 1 more from the queue and runs them in one pass. Each request's rows (in order), candidates and hits
 * are those it gets alone. max_requests in [1, 32], default 16; 1 turns batching off. No counterpart in the
 * reference, which serves one request at a time (daemon/gpujoind.cu:35). */
int gsj_sched_set_batch(gsj_sched* s, int max_requests);
/* queue a request; *ticket identifies it for gsj_wait / gsj_poll */
int gsj_submit(gsj_sched* s, const gsj_request* req, uint64_t* ticket);
/* wait for the request and forget the ticket; returns the request's own status (GSJ_OK or GSJ_E_*, message in
 * gsj_last_error()); stats and times may be NULL */
int gsj_wait(gsj_sched* s, uint64_t ticket, gsj_join_stats* stats, gsj_request_times* times);
/* 1 finished, 0 queued or running, GSJ_E_ARG for an unknown ticket */
int gsj_poll(gsj_sched* s, uint64_t ticket);

/* ------------------------------------------------------------------ DE-9IM helpers (engine/engine.cuh:871-888) */
void gsj_de9im_str(uint32_t code, char out[9]);
uint32_t gsj_de9im_transpose(uint32_t code);
/* pattern of 9 chars from T F * 0 1 2 (case-insensitive); returns 1 on match */
int gsj_de9im_match(uint32_t code, const char* pattern);
#define GSJ_DE9IM_PT_INTERIOR (1u | 3u << 12 | 2u << 14 | 3u << 16)   /* 0FFFFF212, engine/engine.cuh:337 */
#define GSJ_DE9IM_PT_BOUNDARY (1u << 2 | 3u << 12 | 2u << 14 | 3u << 16) /* F0FFFF212, engine/engine.cuh:338 */

/* ------------------------------------------------------------------ host-side layer preparation
 * Flat polygon set owned by the library (the reference HostSet, engine/gset.h:5-11). */
typedef struct gsj_hostset gsj_hostset;
void gsj_hostset_free(gsj_hostset* s);
int gsj_hostset_sizes(const gsj_hostset* s, uint64_t* n, uint64_t* nr, uint64_t* nv);
/* pointers into the set (valid until gsj_hostset_free); parent/syn are NULL unless the set is subdivided */
int gsj_hostset_arrays(const gsj_hostset* s, const uint64_t** gro, const uint64_t** rvo, const double** xy, const uint32_t** parent,
                       const uint8_t** syn);

/* .gset / .pset files (engine/gset.h:13-45; layout scripts/dump_gset.py:5-12, scripts/dump_pset.py:3) */
int gsj_gset_read(const char* path, gsj_hostset** out);
int gsj_gset_write(const char* path, const uint64_t* gro, const uint64_t* rvo, const double* xy, uint64_t n, uint64_t nr, uint64_t nv);
int gsj_pset_read(const char* path, double** xy, uint64_t* n);   /* *xy freed with gsj_free */
int gsj_pset_write(const char* path, const double* xy, uint64_t n);
void gsj_free(void* p);

/* Input sets of the four preparation calls below: gro[n+1], rvo[gro[n]+1], xy[2 rvo[gro[n]]] (the lengths follow from
 * the offsets, as in a .gset file). gro and rvo are checked first (start at 0, never decrease: GSJ_E_ARG otherwise);
 * xy must hold the vertex count the offsets name. */
/* ST_Subdivide analogue (engine/subdivide.h:71-129): pieces of at most maxv stored vertices (depth <= 24) */
int gsj_subdivide(const uint64_t* gro, const uint64_t* rvo, const double* xy, uint64_t n, int maxv, gsj_hostset** out);
/* exact dedup of identical geometries (engine/dedup.h:51-124, first-occurrence order): distinct set plus
 * member lists moff[un+1], midx[n] (caller-allocated, n entries suffice for both) */
int gsj_dedup(const uint64_t* gro, const uint64_t* rvo, const double* xy, uint64_t n, gsj_hostset** distinct, uint64_t* moff,
              uint32_t* midx, uint64_t* n_distinct);
/* the same subdivision on ctx's device (level by level: a thread per piece decides and a thread per ring and half
 * clips; host arrays in and out): identical pieces, cut flags, parents and order. */
int gsj_subdivide_device(gsj_ctx* ctx, const uint64_t* gro, const uint64_t* rvo, const double* xy, uint64_t n, int maxv, gsj_hostset** out);
/* the same dedup on ctx's device (hash, stable radix sorts, group compares and compaction on the GPU; host arrays in
 * and out): identical distinct set, member lists and order. Replaces dedup_set (engine/dedup.h:51-124) on the load
 * path, the step the reference runs on host cores. */
int gsj_dedup_device(gsj_ctx* ctx, const uint64_t* gro, const uint64_t* rvo, const double* xy, uint64_t n, gsj_hostset** distinct,
                     uint64_t* moff, uint32_t* midx, uint64_t* n_distinct);

/* The layer loader's parse of a binary COPY stream (daemon/loader.h:125-170: COPY (SELECT id::text, geom ...) TO
 * STDOUT (FORMAT binary), two fields per row, geometry as EWKB), with the EWKB work on ctx's device: the host frames
 * the rows, a thread per row walks each value's headers and a thread per ring decodes its points and orients it.
 * Rows: a polygonal value (POLYGON, MULTIPOLYGON, or a GEOMETRYCOLLECTION of them; either byte order, SRID, Z, M,
 * ISO type codes) becomes one geometry holding the rings of all its polygons with 4 or more points, each polygon's
 * first ring counter-clockwise and the others clockwise by the sign of a compensated area (daemon/loader.h:56-70,
 * 100-102); a POINT a point. The layer's kind is that of its first accepted row, or `kind` when an earlier piece of
 * the same stream decided it (1 polygons, 2 points); rows of the other kind, unsupported types, points inside
 * collections and malformed values are skipped_type, values without rings or POINT EMPTY skipped_empty, NULL ids or
 * geometries skipped_null (the reference's counters). Collections nest at most 32 deep (deeper: skipped_type).
 * header: buf starts with the PGCOPY signature (GSJ_E_ARG otherwise); rows are read up to the trailer or the first
 * incomplete row (stats->consumed bytes: a caller reading a stream in pieces keeps the rest for the next call, with
 * header 0 and the kind returned). Outputs: *polys (gsj_hostset_free) for a polygon layer, else *pts[2 *npts] for a
 * point layer; ids[id_off[k] .. id_off[k+1]) per accepted row in stream order; *pts, *ids, *id_off freed with gsj_free. */
typedef struct {
  uint64_t rows, skipped_null, skipped_type, skipped_empty, polys, points;
  uint64_t consumed;   /* bytes of buf read */
  int done;            /* the trailer was read */
  int kind;            /* the layer's kind after this piece: 0 none yet, 1 polygons, 2 points */
  double scan_ms;      /* device time of the row walk (outcomes, ring and point counts) */
  double emit_ms;      /* device time of the ring records and the oriented coordinates (or the points) */
} gsj_copy_stats;
int gsj_copy_parse_device(gsj_ctx* ctx, const uint8_t* buf, uint64_t len, int header, int kind, gsj_hostset** polys, double** pts,
                          uint64_t* npts, char** ids, uint64_t** id_off, gsj_copy_stats* stats);
/* A polygon set as such a stream (little-endian EWKB; the SRID flag with srid when srid > 0): a POLYGON per geometry,
 * or a MULTIPOLYGON whose polygons start at each counter-clockwise ring; ids/id_off as above, or NULL for the row
 * numbers "0", "1", ...; *out freed with gsj_free. No counterpart in the reference (its input comes from Postgres);
 * it feeds the parse from the synthetic layers. */
int gsj_copy_write(const uint64_t* gro, const uint64_t* rvo, const double* xy, uint64_t n, const char* ids, const uint64_t* id_off, int srid,
                   uint8_t** out, uint64_t* len);

/* Deterministic synthetic layers (counter-based RNG, identical on every run):
 *   kind "parcels_grid": config-1 parcels (one 4..12-vertex star polygon per cell of an nx x ny grid)
 *   kind "regions_grid": config-1 regions (kx x ky tiles, `edge_verts` vertices per side, shared jittered borders)
 *   kind "parcels_mesh": config-2 parcels (jittered lattice, lognormal vertex counts, shared / overlapping edges)
 * `spec` is "key=value;key=value" (keys and defaults in hostprep.cpp, gen_*); returns the polygon set. */
int gsj_generate(const char* kind, const char* spec, gsj_hostset** out);
/* config-3 points: "n=..;seed=..;x0=..;y0=..;w=..;h=..;clusters=..;sigma_m=..;frac_clustered=.." */
int gsj_generate_points(const char* spec, double** xy, uint64_t* n);

#ifdef __cplusplus
}
#endif
#endif /* GSJ_H */
