/* gpujoind wire protocol: the reference daemon's format, kept byte for byte so the unchanged gpu_spatial extension
 * can talk to this daemon (reference: daemon/protocol.h:1-44). Little endian.
 *
 * request : u32 magic GSP_MAGIC, u32 op, u32 nargs (<= 64), nargs x { u32 len (<= 64 MiB), bytes }; args[0] is the
 *           auth token (empty when the daemon runs without --token-file), the op's args follow
 * response: u32 status (GSP_OK | GSP_ERR); GSP_ERR: u32 len, message
 *   LOAD / LOAD_QUERY / DROP / STATUS / LIST: u32 len, JSON
 *   JOIN / CLUSTER / RELATE / DWITHIN: frames { u32 nrows, rows } ended by nrows == 0, then u32 len, JSON stats;
 *   nrows == GSP_FRAME_ERR is followed by u32 len, message (error after the stream started)
 *   join / dwithin row : u16 la, a_id, u16 lb, b_id, f64 (area; NaN when not requested; for DWITHIN NaN, or the
 *                        distance in metres with the "distance" extension below)
 *   cluster row        : u16 la, id, u32 cluster
 *   relate row         : u16 la, a_id, u16 lb, b_id, char[9]
 *
 * ops (args are UTF-8 strings):
 *   LOAD       layer, source, id_col, geom_col, filter, subdivide, dedup
 *   LOAD_QUERY layer, source, subdivide, dedup
 *   DROP layer | LIST | STATUS
 *   JOIN       layer_a, layer_b, want_area, self_pairs ("all" | "distinct")[, "stats"]
 *   CLUSTER    layer
 *   RELATE     layer_a, layer_b, pattern ("" = every intersecting pair), self_pairs
 *   DWITHIN    layer_a, layer_b, mode ("dwithin" | "buffer"), dist_a, dist_b, self_pairs[, "stats" | "distance"]
 *
 * Differences, all additive: there is no Postgres behind this daemon, so LOAD's `source` (LOAD_QUERY's sql) names a
 * layer file or a generator instead of a table or query ("gset:PATH", "pset:PATH", "copy:PATH" (a binary COPY dump of
 * (id text, geom) rows, the stream the reference's loader reads from Postgres), "gen:KIND:SPEC", "genpts:SPEC";
 * id_col, geom_col and filter are accepted and ignored); DWITHIN's optional 7th argument may be
 * "distance", which fills the f64 slot with the geography distance (dwithin mode only). */
#ifndef GSJ_DAEMON_PROTOCOL_H
#define GSJ_DAEMON_PROTOCOL_H

#define GSP_MAGIC 0x52505347u /* "GSPR" */
#define GSP_OK 0u
#define GSP_ERR 1u
#define GSP_FRAME_ERR 0xFFFFFFFFu

enum {
  GSP_OP_LOAD = 1,
  GSP_OP_LOAD_QUERY = 2,
  GSP_OP_DROP = 3,
  GSP_OP_LIST = 4,
  GSP_OP_STATUS = 5,
  GSP_OP_JOIN = 6,
  GSP_OP_CLUSTER = 7,
  GSP_OP_RELATE = 8,
  GSP_OP_DWITHIN = 9
};

#endif
