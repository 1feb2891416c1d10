/*
 * cytom.h -- C-ABI of the MI355X-native evolving-graph engine.
 *
 * One context = one evolving graph held in HBM (tiled edge store, degrees)
 * plus the vertex programs registered on it.  Every entry point returns 0
 * (or a non-negative count) on success and a negative errno-style code on
 * failure; cytom_last_error() then describes the failure and the context
 * stays usable.  A context must be driven from one host thread; all device
 * work runs on the context's own HIP stream.
 *
 * Each entry point names the reference interface it replaces (paths are
 * relative to the reference repository).
 */
#ifndef CYTOM_H
#define CYTOM_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CYTOM_ABI_VERSION 1

/* Built-in vertex programs (examples/{bfs,cc,sssp,pagerank-delta}.cc). */
enum cytom_program_kind {
  CYTOM_BFS = 0,  /* examples/bfs.cc:18-230 */
  CYTOM_CC = 1,   /* examples/cc.cc:18-214 */
  CYTOM_SSSP = 2, /* examples/sssp.cc:18-228 (u32 weights and distances) */
  CYTOM_PRD = 3,  /* examples/pagerank-delta.cc:30-308 */
  /* Personalized PageRank-delta (no reference counterpart): the PRD program
   * with the teleport term (1-alpha) replaced by (1-alpha)*t_v, t uniform
   * over the teleport set; starts from rank = delta = (1-alpha)*t with the
   * teleport set active and re-initialises after every update batch. */
  CYTOM_PPR = 4,
};

/* Error codes (negative return values). */
#define CYTOM_EINVAL (-22)
#define CYTOM_ENOMEM (-12)
#define CYTOM_ERANGE (-34)
#define CYTOM_EDEVICE (-5)

typedef struct cytom_ctx cytom_ctx;

/* Engine configuration: replaces the option set parsed by parseOption and
 * normalised by Config::normalizeConfig (src/include/main/execute.h:45-362,
 * src/lib/util/config.cc:10-47). */
typedef struct {
  uint64_t n_vertices;        /* rounded up to a power of two (>= 64) */
  uint32_t vertices_per_tile; /* power of two in [64, 16384]; 0 = default */
  uint32_t max_iterations;    /* algorithm_iterations; 0 = 100 (run.py:119) */
  double delta;               /* algorithm_config.delta (PRD threshold) */
  uint64_t edge_capacity;     /* initial edge-pool slots; 0 = automatic */
  uint32_t weighted;          /* keep a u32 weight per edge (SSSP) */
  uint32_t device;            /* HIP device ordinal */
} cytom_config;

/* Boolean statics of the Algorithm concept (examples/bfs.cc:24-30) in
 * cytom_program_desc.flags, plus CYTOM_STATICS_GIVEN when the caller states
 * the statics for cytom_program_register to check. */
#define CYTOM_STATICS_GIVEN (1u << 0)
#define CYTOM_NEEDS_SOURCES_ACTIVE (1u << 1) /* tile-manager.h:405-427 */
#define CYTOM_NEEDS_TARGETS_ACTIVE (1u << 2) /* tile-manager.h:429-451 */
#define CYTOM_ASYNC (1u << 3)
#define CYTOM_USE_CHANGED (1u << 4)          /* apply-executor.h:65, algorithm-executor.h:496 */

/* First-iteration rule of a program's apply. */
enum cytom_first_iteration {
  CYTOM_FIRST_ITERATION_NONE = 0,
  /* GLOBAL_FIRST_ITERATION (pagerank-delta.cc:146-152): in the first
   * iteration after init_vertices every vertex applies
   * delta = (1 - ALPHA) + ALPHA * delta_sum, rank += delta, delta -= 1. */
  CYTOM_FIRST_ITERATION_PRD_ALL = 1,
};

/* Program descriptor: the constructor argument and the statics of the
 * Algorithm concept (examples/bfs.cc:20-34).  kind and source are required;
 * the statics are optional (flags = 0 skips them): cytom_program_statics
 * fills them for a kind, and with CYTOM_STATICS_GIVEN registration checks
 * them against the kind.  neutral_element, vertex_bytes, n_fields,
 * first_iteration and the NEEDS_SOURCES_ACTIVE, NEEDS_TARGETS_ACTIVE and
 * ASYNC bits must equal the kind's, since they change the values a run
 * produces; USE_CHANGED may take either value (the engine's apply visits the
 * vertices a gather or hook changed, which gives the values of a sweep over
 * all vertices: only those can have next != current). */
typedef struct {
  uint32_t kind;             /* enum cytom_program_kind */
  uint32_t source;           /* algo_start_id for BFS/SSSP (execute.h:312) */
  uint32_t flags;            /* CYTOM_STATICS_GIVEN | CYTOM_NEEDS_SOURCES_ACTIVE | ... */
  uint32_t vertex_bytes;     /* sizeof(VertexType) in the results layout: 4, or 12 */
  uint32_t first_iteration;  /* enum cytom_first_iteration */
  uint32_t n_fields;         /* fields of VertexType: 1, or 3 (delta_sum, rank, delta) */
  double neutral_element[3]; /* neutral_element, field by field (unused fields 0) */
} cytom_program_desc;

/* Per-run statistics of cytom_execute
 * (continuous-algorithm-executor.h:224-247 counters). */
typedef struct {
  uint32_t iterations;
  uint32_t converged;
  uint32_t reexecuted;       /* last batch re-initialised the program */
  uint32_t pad;
  uint64_t traversed_edges;  /* sum over iterations of live out-edges of active vertices */
  uint64_t active_vertices;  /* sum over iterations of |frontier| */
  uint64_t slots_scanned;    /* edge slots read by the gather kernels */
  double t_compute_ms;       /* device time of the run (HIP events) */
} cytom_run_stats;

/* Per-batch statistics of cytom_update_batch*. */
typedef struct {
  uint64_t inserted;
  uint64_t deleted;          /* deletions that matched a live edge */
  uint64_t missing;          /* deletions with no live copy */
  uint64_t live_edges;
  uint32_t tiles_relocated;
  uint32_t tiles_compacted;
  uint32_t tiles_sorted;     /* tiles rewritten in source order (sort-compaction) */
  uint32_t host_syncs;       /* host synchronisations the call made */
  double t_ingest_ms;        /* device time of the batch (HIP events) */
} cytom_batch_stats;

/* ---- context ------------------------------------------------------------ */

/* Replaces execute<VertexType,Algorithm,is_weighted> set-up: VertexStore::init,
 * TileStore::init, MetaTileStore::init (execute.h:534-652). */
int cytom_create(const cytom_config* cfg, cytom_ctx** out);
/* Replaces the teardown at the end of execute (execute.h:755-771). */
int cytom_destroy(cytom_ctx* ctx);
const char* cytom_last_error(const cytom_ctx* ctx);
/* Effective geometry after normalisation (config.cc:10-47). */
int cytom_geometry(const cytom_ctx* ctx, uint64_t* n_vertices, uint32_t* vertices_per_tile,
                   uint32_t* tiles_per_dim);

/* ---- algorithm store ---------------------------------------------------- */

/* Registers a vertex program on the graph and runs its init_vertices
 * (the main_<name> export + execute<> instantiation, examples/bfs.cc:235-238;
 * ContinuousAlgorithmExecutor::initAlgorithm, continuous-algorithm-executor.h:342). */
/* The statics of kind, as the reference's algorithm states them
 * (examples/bfs.cc:20-32, cc.cc:20-32, sssp.cc:20-32,
 * pagerank-delta.cc:47-65), with CYTOM_STATICS_GIVEN set and source 0. */
int cytom_program_statics(uint32_t kind, cytom_program_desc* out);
int cytom_program_register(cytom_ctx* ctx, const cytom_program_desc* desc, int* prog_id);
/* Re-runs init_vertices (EdgeInserter re-execution, edge-inserter.h:164-167). */
int cytom_program_init(cytom_ctx* ctx, int prog_id);
/* Replaces the program's source (BFS/SSSP: n == 1, algo_start_id,
 * execute.h:312) or teleport set (PPR: n >= 1 vertices, duplicates ignored)
 * and re-runs init_vertices. */
int cytom_program_set_sources(cytom_ctx* ctx, int prog_id, const uint32_t* sources, uint32_t n);

/* ---- update ingestion --------------------------------------------------- */

/* One EdgeInserter batch (edge-inserter.h:53-194 with
 * MetaTileManager::run, meta-tile-manager.h:156-274): n_ins insertions then
 * n_del deletions, then per program either re-initialisation or the
 * edgeChanged(INSERT) hooks.  Host pointers; w/del_w may be NULL.  With
 * del_w NULL a deletion matches any weight. */
int cytom_update_batch(cytom_ctx* ctx, const uint32_t* src, const uint32_t* dst,
                       const uint32_t* w, uint64_t n_ins, const uint32_t* del_src,
                       const uint32_t* del_dst, const uint32_t* del_w, uint64_t n_del,
                       cytom_batch_stats* stats);
/* The same with device pointers already resident in HBM. */
int cytom_update_batch_device(cytom_ctx* ctx, const uint32_t* d_src, const uint32_t* d_dst,
                              const uint32_t* d_w, uint64_t n_ins, const uint32_t* d_del_src,
                              const uint32_t* d_del_dst, const uint32_t* d_del_w,
                              uint64_t n_del, cytom_batch_stats* stats);
/* The same batch from host records in the reference's binary edge layouts:
 * record_bytes 8 is file_edge_t {int32 src, int32 tgt} / edge_compact_t
 * {u32 src, u32 tgt} (util/datatypes.h:41-44, 55-58; the binary generator's
 * input, jobs.h:52-75), 12 is edge_compact_weighted_t {u32 src, u32 tgt,
 * float weight} (util/datatypes.h:60-64; the weighted edge logs,
 * edges-writer.h:150-160).  Deletions use the same layout.  Records are
 * copied through pinned staging buffers owned by the context (the caller's
 * pages may be an mmap of the file).  8-byte records insert default_weight
 * into a weighted graph; 12-byte weights must be integral in [1, 2^24].
 * Deletions without weights match any weight. */
int cytom_update_batch_records(cytom_ctx* ctx, const void* records, uint32_t record_bytes,
                               uint64_t n_ins, uint32_t default_weight,
                               const void* del_records, uint64_t n_del,
                               cytom_batch_stats* stats);

/* ---- execution ---------------------------------------------------------- */

/* ContinuousAlgorithmExecutor::executeAlgorithm
 * (continuous-algorithm-executor.h:49-249): runs the program until
 * convergence or max_iterations. */
int cytom_execute(cytom_ctx* ctx, int prog_id, cytom_run_stats* stats);
/* Runs of several programs on the current graph, as cytom_execute of each in
 * turn (the same values), but concurrently, each on a stream of its own
 * (cytom_set_concurrent_runs), with one host synchronisation per iteration
 * group of all of them: the jobs of one process that share a graph
 * (jobs.h:33-140) run as one device-side mix.  run_stats: n_progs entries or
 * NULL. */
int cytom_execute_many(cytom_ctx* ctx, const int* prog_ids, uint32_t n_progs,
                       cytom_run_stats* run_stats);
/* One step of the continuous loop, edge-inserter.h:129-184: the update batch
 * (device pointers, as cytom_update_batch_device), then a run of each of the
 * n_progs listed programs (distinct ids), with the re-initialisations the
 * batch requires.  The values equal those of cytom_update_batch_device
 * followed by cytom_execute of each program in turn; the batch and every program's first iteration group are enqueued before the
 * host reads anything, so a step whose programs converge within the
 * previous run's iteration count plus 1/16 synchronises once, and otherwise
 * once more per further group (batch_stats->host_syncs).  A refused batch
 * (CYTOM_EINVAL) changes nothing and runs no program.  run_stats: n_progs
 * entries or NULL. */
int cytom_step(cytom_ctx* ctx, const uint32_t* d_src, const uint32_t* d_dst, const uint32_t* d_w,
               uint64_t n_ins, const uint32_t* d_del_src, const uint32_t* d_del_dst,
               const uint32_t* d_del_w, uint64_t n_del, const int* prog_ids, uint32_t n_progs,
               cytom_batch_stats* batch_stats, cytom_run_stats* run_stats);
/* Step graphs (no reference counterpart: the reference's EdgeInserter
 * launches nothing).  With enable != 0, cytom_step records its enqueue phase
 * into HIP graphs (one for the update batch and the work lists, one for each
 * listed program's first iteration group) and replays them on later steps
 * whose arguments, programs, device buffers, geometry and tuning match; a
 * phase that differs in any of them (another batch size, a grown scratch
 * buffer, a repacked pool) records another graph, and at most 32 are kept
 * (least recently used evicted).  First iteration groups get an even length
 * so that the buffer parity they leave repeats from batch to batch (a
 * launch past convergence changes nothing).  Values and statistics
 * equal those of direct launches; with profiling on, steps launch directly.
 * 0 disables graphs and frees them. */
int cytom_set_step_graphs(cytom_ctx* ctx, int enable);
/* With enable != 0 (the default) cytom_step runs its programs concurrently
 * after the update batch, each on a stream of its own (at most 4), so one
 * program's launches fill the gaps and tails of another's; 0 runs them one
 * after the other on the context's stream.  Values are the same either way. */
int cytom_set_concurrent_runs(cytom_ctx* ctx, int enable);
/* Graphs recorded and replayed by the context's steps, and graphs kept. */
int cytom_step_graph_stats(cytom_ctx* ctx, uint64_t* captures, uint64_t* replays,
                           uint32_t* cached);
/* Copies the program's values (the `next` array written by ResultsWriter,
 * results-writer.h:54-67): BFS/CC/SSSP uint32[V], PRD/PPR float rank[V]. */
int cytom_read_values(cytom_ctx* ctx, int prog_id, void* out, size_t bytes);
/* The program's vertex values in the reference's results_<n> file layout
 * (ResultsWriter::run, results-writer.h:46-81, raw VertexType[V]): BFS/CC
 * uint32, SSSP float (FLT_MAX unreached, examples/sssp.cc:21), PRD/PPR
 * {float delta_sum, rank, delta} (examples/pagerank-delta.cc:30-33).  V is
 * the context's vertex count (cytom_geometry).  Returns the bytes written,
 * or with out == NULL the bytes needed; negative on error. */
int64_t cytom_read_results(cytom_ctx* ctx, int prog_id, void* out, size_t bytes);

/* ---- batched queries ---------------------------------------------------- */

/* One query of a batch: BFS or SSSP from one source, or personalized
 * PageRank-delta from a teleport set (duplicates ignored). */
typedef struct {
  uint32_t kind;            /* CYTOM_BFS, CYTOM_SSSP or CYTOM_PPR */
  uint32_t n_sources;       /* 1 for BFS/SSSP; teleport set size for PPR */
  const uint32_t* sources;  /* host array */
} cytom_query;

/* Runs nq queries against the current live graph as one batch (the
 * reference composes several jobs per process, jobs.h:33-140, and fuses two
 * PageRank-delta instances, pagerank-delta-merge.cc:151-252).  Each query's
 * values equal a single-query run of the same program: BFS/SSSP exactly,
 * PPR ranks exactly (same arithmetic and iteration cap).  Results stay in the
 * context until the next batch; stats aggregate the batch (traversed_edges
 * counts each shared frontier edge once per query group). */
int cytom_batched_execute(cytom_ctx* ctx, const cytom_query* q, uint32_t nq,
                          cytom_run_stats* stats);
/* Values of query `query` of the last batch: BFS/SSSP uint32[V], PPR
 * float[V] (ResultsWriter layout, results-writer.h:54-67). */
int cytom_read_query(cytom_ctx* ctx, uint32_t query, void* out, size_t bytes);

/* ---- differentiable personalized PageRank ------------------------------- */

/* Weighted personalized PageRank on an explicit edge list (device arrays;
 * no reference counterpart).  P_e = w_e / W_u with W_u the sum of the
 * weights leaving u (rows without out-edges drop their mass); weights must
 * be positive and finite (CYTOM_EINVAL otherwise).
 * r = (1-alpha) s + alpha P^T r, solved by FP32 Jacobi sweeps from
 * (1-alpha) s until the largest change is <= tol or max_iterations sweeps;
 * a non-finite iterate (from non-finite s) fails with CYTOM_ERANGE.
 * Deterministic (fixed summation order).  d_r receives float[V]. */
int cytom_ppr_weighted(uint64_t n_vertices, uint64_t n_edges, const uint32_t* d_src,
                       const uint32_t* d_dst, const float* d_w, const float* d_s, float alpha,
                       float tol, uint32_t max_iterations, float* d_r, uint32_t* iterations);
/* Reverse mode of cytom_ppr_weighted for a loss gradient dL/dr: solves
 * lambda = dL/dr + alpha P lambda the same way and writes
 * dL/dw_e = alpha r_u (lambda_v - sum_e' P_e' lambda_dst(e')) / W_u (d_dL_dw,
 * float[E], may be NULL) and dL/ds = (1-alpha) lambda (d_dL_ds, float[V],
 * may be NULL).  d_r is the forward result. */
int cytom_ppr_adjoint(uint64_t n_vertices, uint64_t n_edges, const uint32_t* d_src,
                      const uint32_t* d_dst, const float* d_w, const float* d_r,
                      const float* d_dL_dr, float alpha, float tol, uint32_t max_iterations,
                      float* d_dL_dw, float* d_dL_ds, uint32_t* iterations);
/* The same operator on a fixed edge list, for repeated calls (no reference
 * counterpart; for training loops that keep a graph and change weights or
 * teleport vectors).  cytom_ppr_graph_create copies the endpoints, checks
 * them against n_vertices and builds the CSR / CSC rows once;
 * cytom_ppr_graph_set_weights recomputes only the transition probabilities
 * (float[E] in the edge order given at creation); forward and adjoint then
 * behave as cytom_ppr_weighted / cytom_ppr_adjoint without rebuilding the
 * rows.  Calls on one graph are serialised on its own stream; device arrays
 * passed in must be ready when the call is made. */
typedef struct cytom_ppr_graph cytom_ppr_graph;
int cytom_ppr_graph_create(uint64_t n_vertices, uint64_t n_edges, const uint32_t* d_src,
                           const uint32_t* d_dst, const float* d_w, cytom_ppr_graph** out);
int cytom_ppr_graph_set_weights(cytom_ppr_graph* g, const float* d_w);
int cytom_ppr_graph_forward(cytom_ppr_graph* g, const float* d_s, float alpha, float tol,
                            uint32_t max_iterations, float* d_r, uint32_t* iterations);
int cytom_ppr_graph_adjoint(cytom_ppr_graph* g, const float* d_r, const float* d_dL_dr,
                            float alpha, float tol, uint32_t max_iterations, float* d_dL_dw,
                            float* d_dL_ds, uint32_t* iterations);
int cytom_ppr_graph_destroy(cytom_ppr_graph* g);
/* cytom_ppr_weighted without any device allocation (no counterpart in the
 * reference; for callers that cannot allocate: code under stream capture,
 * tight serving loops, harnesses that forbid hipMalloc).  Same arguments and
 * results (bit for bit) as cytom_ppr_weighted; every temporary (rows,
 * probabilities, sort and solver buffers) lives in one 64 MiB device array
 * per GPU that the library holds until the process ends, which bounds the
 * graph to roughly 40 B per vertex plus 40 B per edge (for example 2^18
 * vertices and 2^20 edges): larger graphs fail with CYTOM_ERANGE before any
 * work.  Runs on the null stream.  Calls on one GPU share that array, so
 * calls that overlap in time on one GPU (two host threads) race. */
int cytom_ppr_weighted_noalloc(uint64_t n_vertices, uint64_t n_edges, const uint32_t* d_src,
                               const uint32_t* d_dst, const float* d_w, const float* d_s,
                               float alpha, float tol, uint32_t max_iterations, float* d_r,
                               uint32_t* iterations);
/* Reverse mode of a registered personalized-PageRank program (CYTOM_PPR)
 * on the context's current graph (no reference counterpart).  The
 * program's ranks r solve r = (1-alpha) t + alpha P^T r over the live
 * multigraph with P_uv = 1/out_degree(u) per live edge, i.e. every edge
 * weight 1 in the operator of cytom_ppr_weighted.  For d_dL_dr (float[V],
 * device) this solves lambda = dL/dr + alpha P lambda by FP32 Jacobi sweeps
 * over the tile store's live slots in the reverse orientation of the
 * program's gather, from lambda = dL/dr until the largest change is <= tol
 * or max_iterations sweeps (0: 1000), and writes dL/dt = (1-alpha) lambda
 * (d_dL_dt, float[V], may be NULL) and, for every live edge in the order of
 * cytom_export_edges, its endpoints and dL/dw at w = 1,
 * alpha r_u (lambda_v - mu_u) / out_degree(u) with mu_u the mean lambda over
 * u's live out-edges (d_src, d_dst, d_dL_dw: cap entries each, any may be
 * NULL).  r is the program's rank from its last run.  Returns the number of
 * live edges; CYTOM_ERANGE when an edge output is given and cap is below it,
 * CYTOM_EINVAL for a program of another kind. */
int64_t cytom_ppr_program_adjoint(cytom_ctx* ctx, int prog_id, const float* d_dL_dr, float tol,
                                  uint32_t max_iterations, float* d_dL_dt, uint32_t* d_src,
                                  uint32_t* d_dst, float* d_dL_dw, uint64_t cap,
                                  uint32_t* iterations);
/* Message of the last failed cytom_ppr_* call on this thread. */
const char* cytom_ppr_last_error(void);

/* ---- inspection --------------------------------------------------------- */

/* VertexStore degrees (vertex-store.h:64-70). */
int cytom_read_degrees(cytom_ctx* ctx, uint32_t* out_degree, uint32_t* in_degree);
/* Live edges of the store (TileStore::countEdges / edge lists,
 * tile-store.h:174-185).  Returns the live count; writes at most cap edges
 * (arrays may be NULL to query the count). */
int64_t cytom_export_edges(cytom_ctx* ctx, uint32_t* src, uint32_t* dst, uint32_t* w,
                           uint64_t cap);

/* ---- tuning ------------------------------------------------------------- */

/* Launch shape of the gather (the device analogue of the TileDistributor's
 * tile_batching and executor count, tile-distributor.h:113-177,
 * run.py:115-132).  Zero fields keep the current value. */
typedef struct {
  uint32_t chunk_slots;   /* edge slots per work chunk (multiple of 64, <= 1024) */
  uint32_t job_chunks;    /* chunks per column job (<= 2^20; reported 0 = sized from the
                             store: at least 131072 slots, ~8 jobs per CU for large stores) */
  uint32_t blocks_per_cu; /* persistent gather blocks per compute unit */
  uint32_t source_bands;  /* source-row bands of about equal slot counts, one job queue
                             each, served first by every NB-th block (power of two <= 16) */
  uint32_t resort_pct;    /* sort-compact a tile once its unsorted tail reaches this
                             percentage of its sorted prefix; CYTOM_RESORT_OFF disables */
  uint32_t resort_max_slots; /* slots staged per sort-compaction pass (more tiles wait) */
  uint32_t prd_frac_bits; /* fraction bits of the PageRank-delta fixed-point sums, in
                             [16, 60 - ceil(log2 2V)] (the default is the maximum; lower
                             values exist to measure the quantisation).  Set it before
                             PageRank-delta programs are registered or re-initialise them. */
} cytom_tuning;
#define CYTOM_RESORT_OFF 0xFFFFFFFFu
int cytom_set_tuning(cytom_ctx* ctx, const cytom_tuning* t);
int cytom_get_tuning(const cytom_ctx* ctx, cytom_tuning* t);

/* ---- measurement ------------------------------------------------------- */

/* Enables (n >= 1) or disables (0) HIP-event timing of gather launches and
 * resets the accumulators (replaces the PerfEventScoped spans around the
 * pull phase, continuous-algorithm-executor.h:93-97): with n = 1 every launch
 * is timed, with n > 1 every n-th iteration of every run, the first timed
 * iteration rotating from run to run so that every iteration index is sampled
 * alike (each timed launch costs a few microseconds of device time).
 * The profile then accumulates the sampled launches only. */
int cytom_set_profiling(cytom_ctx* ctx, int enable);
/* Accumulated gather-kernel time, launch count and algorithmic bytes of one
 * program kind since profiling was enabled. */
int cytom_kernel_profile(cytom_ctx* ctx, uint32_t kind, double* gather_ms, uint64_t* launches,
                         double* alg_bytes);

/* ---- synthetic update streams ------------------------------------------ */

/* RMAT edges [first, first+n) with 2^scale vertices on the device
 * (RMATContext::getEdge, src/include/edge-inserter/rmat-context.h:39-73);
 * d_w (nullable) receives the stream's u32 weights in [1,255]. */
int cytom_rmat_generate_device(uint32_t scale, uint64_t seed, uint64_t first, uint64_t n,
                               uint32_t* d_src, uint32_t* d_dst, uint32_t* d_w);
/* The same for an arbitrary list of stream indices (device pointer). */
int cytom_rmat_gather_device(uint32_t scale, uint64_t seed, const uint64_t* d_index,
                             uint64_t n, uint32_t* d_src, uint32_t* d_dst, uint32_t* d_w);

/* Deletion sampler of the synthetic stream: inserts are consecutive stream
 * indices; deletions are drawn uniformly without replacement from the live
 * indices (the deterministic replacement of deleteRandomEdge,
 * tile-manager.h:175-190). */
typedef struct cytom_stream cytom_stream;
int cytom_stream_create(uint64_t seed, cytom_stream** out);
int cytom_stream_destroy(cytom_stream* s);
int64_t cytom_stream_next(cytom_stream* s, uint64_t n_ins, uint64_t n_del, uint64_t* ins_first,
                          uint64_t* del_index);

/* ---- device memory helpers for callers without a HIP runtime ----------- */
int cytom_device_alloc(void** ptr, size_t bytes);
int cytom_device_free(void* ptr);
int cytom_memcpy_h2d(void* dst, const void* src, size_t bytes);
int cytom_memcpy_d2h(void* dst, const void* src, size_t bytes);
/* Device-to-device copy (a plain copy kernel: the reference point for HBM
 * traffic counters). */
int cytom_memcpy_d2d(void* dst, const void* src, size_t bytes);
int cytom_device_synchronize(void);
int cytom_device_count(void);

#ifdef __cplusplus
}
#endif
#endif
