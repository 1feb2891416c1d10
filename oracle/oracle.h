/*
 * CPU oracle for the evolving-graph hot path (TEST INFRASTRUCTURE ONLY).
 *
 * This library is the checker: a plain C++ restatement of the reference's
 * streaming ingest and incremental vertex-program executors, plus textbook
 * checkers (queue BFS, Dijkstra, min-reachable labels, FP64 PageRank).
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load it.  The product (cytom_amd/, include/cytom.h) never links it.
 *
 * Reference anchors (paths relative to the reference repository):
 *   hash / hashDouble ............ src/include/util/util.h:178-200
 *   RMAT recurrence .............. src/include/edge-inserter/rmat-context.h:19-73
 *   tile routing ................. src/include/core/tile-store.h:88-96, traversal row_first.cc:6-11
 *   insert / delete .............. src/include/tile-manager/tile-manager.h:136-190,
 *                                  src/include/tile-manager/meta-tile-manager.h:112-231
 *   pull / applyTiles ............ src/include/core/algorithm-executor.h:214-505
 *   apply ........................ src/include/core/apply-executor.h:40-115
 *   iteration loop ............... src/include/core/continuous-algorithm-executor.h:49-249
 *   programs ..................... examples/{bfs,cc,sssp,pagerank-delta}.cc
 *
 * Parity pins (what checks each function; see tests/test_oracle.py):
 *   orc_hash64 / orc_hash_double / orc_rmat_edges: the reference's own
 *     hash and rMatIterative text compiled by oracle/ref_build.py into
 *     oracle/_ref (when the reference tree is present), the recursive
 *     form rMat (rmat-context.h:19-37) restated independently, quadrant
 *     frequency statistics, and committed fixtures.
 *   store (insert/delete/degrees): brute-force multiset replay.
 *   BFS / CC / SSSP executors: textbook checkers below, bit-exact.
 *   PageRank-delta executor: FP64 restatement (orc_prd_f64) and FP64
 *     power iteration; hand-computed values on tiny graphs.
 *   stream weights / deletion sampler: definitional (no reference
 *     counterpart) -- pinned by invariants (range, uniqueness, liveness),
 *     otherwise "parity unpinned".
 *   Personalized PageRank-delta (ORC_PPR, no reference counterpart): the
 *     FP64 power iteration orc_ppr_f64 and hand-computed tiny cases.
 */
#ifndef CYTOM_ORACLE_H
#define CYTOM_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { ORC_BFS = 0, ORC_CC = 1, ORC_SSSP = 2, ORC_PRD = 3, ORC_PPR = 4 };

/* ---- a1: hash + RMAT ---------------------------------------------------- */
uint32_t orc_hash32(uint32_t a);
uint64_t orc_hash64(uint64_t a);
double orc_hash_double(uint64_t i);
/* Edges [first, first+n) of the RMAT stream with 2^scale vertices. */
void orc_rmat_edges(uint32_t scale, uint64_t seed, uint64_t first, uint64_t n,
                    uint32_t* src, uint32_t* dst);
/* The recursive form of the same recurrence (rmat-context.h:19-37). */
void orc_rmat_edges_recursive(uint32_t scale, uint64_t seed, uint64_t first,
                              uint64_t n, uint32_t* src, uint32_t* dst);
/* The same edges for an arbitrary list of stream indices. */
void orc_rmat_gather(uint32_t scale, uint64_t seed, const uint64_t* index, uint64_t n,
                     uint32_t* src, uint32_t* dst);
void orc_stream_weights_u32_at(uint64_t seed, const uint64_t* index, uint64_t n, uint32_t* w);
/* Per-edge weights of the stream: u32 in [1,255] and f32 in (0,1]. */
void orc_stream_weights_u32(uint64_t seed, uint64_t first, uint64_t n, uint32_t* w);
void orc_stream_weights_f32(uint64_t seed, uint64_t first, uint64_t n, float* w);

/* ---- stream with deletion sampler -------------------------------------- */
typedef struct orc_stream orc_stream;
orc_stream* orc_stream_create(uint64_t seed);
void orc_stream_destroy(orc_stream* s);
/* One batch: inserts are stream indices [*ins_first, *ins_first+n_ins);
 * deletions are n_del live stream indices written to del_idx.
 * Returns the number of deletions actually drawn (<= live count). */
int64_t orc_stream_next(orc_stream* s, uint64_t n_ins, uint64_t n_del,
                        uint64_t* ins_first, uint64_t* del_idx);
uint64_t orc_stream_live_count(const orc_stream* s);

/* ---- tiled store + executors ------------------------------------------- */
typedef struct orc_graph orc_graph;

typedef struct {
  uint32_t iterations;
  uint32_t converged;
  uint32_t reexecuted; /* 1 if the last update batch re-initialised the program */
  uint32_t pad;
  uint64_t traversed_edges; /* sum over iterations of live out-edges of active vertices */
  uint64_t active_vertices; /* sum over iterations of |frontier| */
} orc_stats;

orc_graph* orc_create(uint64_t n_vertices, uint32_t vertices_per_tile, int weighted,
                      uint32_t max_iterations, double delta, int n_threads);
void orc_destroy(orc_graph* g);
uint64_t orc_num_vertices(const orc_graph* g);
int orc_register(orc_graph* g, int kind, uint32_t source);
/* One EdgeInserter batch (edge-inserter.h:53-194): inserts, deletions, hooks.
 * dw may be NULL (match any weight).  Returns number of deletions found. */
int64_t orc_update(orc_graph* g, const uint32_t* src, const uint32_t* dst,
                   const uint32_t* w, uint64_t n_ins, const uint32_t* dsrc,
                   const uint32_t* ddst, const uint32_t* dw, uint64_t n_del);
int orc_execute(orc_graph* g, int prog, orc_stats* out);
/* Replaces a BFS/SSSP source (n == 1) or the PPR teleport set (duplicates
 * ignored) and re-runs init.  Returns 0 or -22. */
int orc_set_sources(orc_graph* g, int prog, const uint32_t* sources, uint32_t n);
/* BFS/CC/SSSP: uint32[V]; PRD/PPR: float rank[V]. */
int orc_read(const orc_graph* g, int prog, void* out);
int orc_last_reexec(const orc_graph* g, int prog);
uint64_t orc_live_edges(const orc_graph* g, uint32_t* src, uint32_t* dst, uint32_t* w);
void orc_degrees(const orc_graph* g, uint32_t* outdeg, uint32_t* indeg);

/* ---- textbook checkers on an explicit edge list ------------------------ */
/* Threads that build the checkers' adjacency rows (default 1); the searches
 * themselves are sequential, so results do not depend on it. */
void orc_checker_threads(int n_threads);
void orc_check_bfs(uint64_t V, uint64_t E, const uint32_t* src, const uint32_t* dst,
                   uint32_t source, uint32_t* out);
void orc_check_sssp(uint64_t V, uint64_t E, const uint32_t* src, const uint32_t* dst,
                    const uint32_t* w, uint32_t source, uint32_t* out);
void orc_check_cc(uint64_t V, uint64_t E, const uint32_t* src, const uint32_t* dst,
                  uint32_t* out);
/* The same labels with the rows built by n_threads threads (billion-edge
 * graphs); the depth-first sweep stays sequential. */
void orc_check_cc_mt(uint64_t V, uint64_t E, const uint32_t* src, const uint32_t* dst,
                     int n_threads, uint32_t* out);
/* Queue BFS levels (w == NULL) or Dijkstra distances (u32 weights w) from
 * each of n sources on one set of rows; out holds n x V values. */
void orc_check_multi(uint64_t V, uint64_t E, const uint32_t* src, const uint32_t* dst,
                     const uint32_t* w, const uint32_t* sources, uint32_t n, int n_threads,
                     uint32_t* out);
/* FP64 restatement of the PageRank-delta program from scratch (same cap and
 * activation threshold).  Returns the iteration count. */
uint32_t orc_prd_f64(uint64_t V, uint64_t E, const uint32_t* src, const uint32_t* dst,
                     uint32_t max_iterations, double delta, double* rank);
/* The same with the vertex range split over n_threads threads. */
uint32_t orc_prd_f64_mt(uint64_t V, uint64_t E, const uint32_t* src, const uint32_t* dst,
                        uint32_t max_iterations, double delta, int n_threads, double* rank);
/* One FP64 sweep of the PageRank fixed point over FP32 ranks r:
 * out[0] = sum_v |r_v - (1-alpha) - alpha sum_{u->v} r_u/outdeg(u)|,
 * out[1] = sum_v |r_v|. */
void orc_prd_residual(uint64_t V, uint64_t E, const uint32_t* src, const uint32_t* dst,
                      const float* rank, double alpha, int n_threads, double* out);
/* FP64 power iteration of personalized PageRank:
 * r = (1-a) t + a * sum_{u->v} r_u / outdeg(u), t uniform over the distinct
 * teleport vertices; mass at vertices without out-edges is dropped. */
uint32_t orc_ppr_f64(uint64_t V, uint64_t E, const uint32_t* src, const uint32_t* dst,
                     const uint32_t* teleport, uint32_t n_teleport, double alpha, double tol,
                     uint32_t max_iterations, double* rank);
/* FP64 power iteration r = (1-a) + a * sum_{u->v} r_u / outdeg(u). */
uint32_t orc_pagerank_f64(uint64_t V, uint64_t E, const uint32_t* src, const uint32_t* dst,
                          double alpha, double tol, uint32_t max_iterations, double* rank);

#ifdef __cplusplus
}
#endif
#endif
