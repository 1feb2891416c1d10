// Internal declarations shared by the HIP sources of libcytom_hip.so.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/cytom.h"

namespace cytom {

constexpr uint32_t INF = 0xFFFFFFFFu;
// Edge slot: 4 bytes, the tile-local source in the low half and the
// tile-local target in the high half (the tile supplies the high bits; the
// reference's local_edge_t, datatypes.h:77-92).  Tiles are at most 16,384
// vertices wide, so a deleted slot can hold TOMB (0xFFFF, 0xFFFF).
constexpr uint32_t TOMB = 0xFFFFFFFFu;
constexpr uint64_t MAX_VERTICES = 1ull << 31;

__host__ __device__ inline uint32_t pack_local(uint32_t s, uint32_t d, uint32_t mask) {
  return (s & mask) | ((d & mask) << 16);
}

// Global (src, dst) key of the deletion table.
__host__ __device__ inline uint64_t pack_edge(uint32_t s, uint32_t d) {
  return (uint64_t)s | ((uint64_t)d << 32);
}

// Ligra's 64-bit integer hash (reference util.h:188-196).
__host__ __device__ inline uint64_t hash64(uint64_t a) {
  a = (a + 0x7ed55d166bef7a1dull) + (a << 12);
  a = (a ^ 0xc761c23c510fa2ddull) ^ (a >> 9);
  a = (a + 0x165667b183a9c0e1ull) + (a << 59);
  a = (a + 0xd3a2646cab3487e3ull) ^ (a << 49);
  a = (a + 0xfd7046c5ef9ab54cull) + (a << 3);
  a = (a ^ 0xb55a4f090dd4a67bull) ^ (a >> 32);
  return a;
}

// Stream salts: weights and deletion draws (no reference counterpart).
constexpr uint64_t WEIGHT_SALT = 0x7765696768747321ull;
constexpr uint64_t DELETE_SALT = 0x64656c6574657321ull;

__host__ __device__ inline uint32_t sat_add(uint32_t a, uint32_t b) {
  uint64_t s = (uint64_t)a + b;
  return s >= INF ? INF : (uint32_t)s;
}

// ---------------------------------------------------------------- store view
// Device-resident tiled edge store.  Tile t = y*T + x holds edges with
// src in [x*vpt, (x+1)*vpt) and dst in [y*vpt, (y+1)*vpt) (reference
// tile-store.h:88-96 with the row_first.cc:6-11 traversal order).
struct StoreView {
  uint64_t V;
  uint32_t vpt, T, log2vpt, pad;
  uint64_t n_tiles;
  uint64_t* tile_off;   // first slot of the tile's block in the pool
  uint32_t* tile_cap;   // slots reserved for the tile
  uint32_t* tile_cnt;   // slots used (live + tombstones)
  uint32_t* tile_live;  // live edges
  uint32_t* tile_sorted;  // leading slots in source order (sort-compacted prefix)
  uint32_t* pool;       // local edge pairs
  uint32_t* wpool;      // per-slot weights (nullptr when unweighted)
  uint32_t* outdeg;
  uint32_t* indeg;
};

// PageRank-delta and its personalized form share the gather and apply.
__host__ __device__ inline bool prd_like(int kind) { return kind == CYTOM_PRD || kind == CYTOM_PPR; }

// Control words of an update batch in cytom_ctx::d_misc that refuse it: an
// invalid update (MISC_ERR) or relocations past the pool (MISC_OVERFLOW).  A
// refused batch changes nothing, and program runs enqueued behind it in the
// same step (Program::refuse) do nothing either.
constexpr int MISC_ERR = 1;
constexpr int MISC_OVERFLOW = 8;
__device__ inline bool batch_refused(const unsigned long long* misc) {
  return misc != nullptr && (misc[MISC_ERR] | misc[MISC_OVERFLOW]) != 0;
}

// Internal program kind of the batched queries: two SSSP queries in one run,
// with a (u32, u32) pair of distances per vertex in cur / nxt / fval (each
// array V pairs) and a pair window in the gather.  Never registered in a
// context's program list, so the update hooks never see it.
constexpr int KIND_SSSP_PAIR = 5;
// Internal kind of the batched queries on the 4,096-vertex query tiles: eight
// SSSP queries per vertex (u32 distances, vertex-major).
constexpr int KIND_SSSP_OCT = 6;

// Tile width of the query tiles the eight-query kind runs on (its LDS window
// holds 4,096 targets x 8 u32 = 128 KB).
constexpr uint32_t QUERY_TILE_VPT = 4096;

// Queries held per vertex by a program kind.
__host__ __device__ constexpr uint32_t kind_width(int kind) {
  return kind == KIND_SSSP_PAIR ? 2u : kind == KIND_SSSP_OCT ? 8u : 1u;
}

// Per-iteration device counters.
struct IterStats {
  unsigned long long te;       // live out-edges of the frontier
  unsigned long long count;    // frontier size
  unsigned long long scanned;  // edge slots read by the gather
  unsigned long long updated;  // targets whose value the gather changed
  unsigned long long job_cursor;
  unsigned long long pad[3];
  unsigned long long band_cursor[16];  // next job of each source band's queue
};
constexpr uint32_t MAX_SOURCE_BANDS = 16;

// Per-program device state.
struct Program {
  int kind = 0;
  uint32_t source = 0;
  uint32_t source2 = 0;          // second query's source (KIND_SSSP_PAIR)
  uint32_t* teleport = nullptr;  // PPR teleport set (device, distinct ids)
  uint32_t n_teleport = 0;
  // Sources of the eight-query kind (device): n_qseed vertices, then each
  // one's query.
  uint32_t* qseed = nullptr;
  uint32_t n_qseed = 0;
  uint32_t* cur = nullptr;  // BFS/CC/SSSP current values
  uint32_t* nxt = nullptr;  // BFS/CC/SSSP next values
  uint32_t* fval = nullptr; // BFS/CC/SSSP: cur[v] while v is active, else INF
  float* rank = nullptr;    // PRD next.rank
  float* delta = nullptr;   // PRD current.delta
  long long* dsum = nullptr;  // PRD next.delta_sum, fixed point scaled by 2^prd_shift
  float* contrib = nullptr; // PRD delta/out_degree while v is active, else 0
  uint64_t* act[2] = {nullptr, nullptr};  // active bitmaps (wave-sized words)
  uint8_t* gact[2] = {nullptr, nullptr};  // per source row: 64 group-activity bytes
  uint8_t* changed = nullptr;               // per-vertex changed flags
  // Per-run iteration counters ([max_iterations + 2], device and pinned host
  // copy) and device control words: [0] = the PageRank-delta first-iteration
  // rule applies (set by init, cleared by the first apply), [1] = apply blocks
  // done (the last one clears [0]).
  IterStats* d_iter = nullptr;
  IterStats* h_iter = nullptr;
  uint32_t* dctl = nullptr;
  std::vector<hipEvent_t> gev;  // start and end of every iteration group of a run
  std::vector<hipEvent_t> kev;  // around every gather launch (profiling)
  // The run in flight: iterations launched, groups enqueued, counters
  // examined, the converged iteration once seen (-1 before), and the refusal
  // words of the batch it follows (nullptr outside a step).
  uint32_t launched = 0, groups = 0, checked = 0;
  int conv = -1;
  const unsigned long long* refuse = nullptr;
  // The stream a step runs the program on (nullptr: the context's stream).
  hipStream_t stream = nullptr;
  int cur_act = 0;
  bool fixpoint = false;
  int last_reexec = 0;
  uint32_t last_iters = 0;  // iterations of the previous run (first launch group)
  // Runs opened so far: a sampled profile times iterations i with
  // i % every == runs % every, so every iteration index is sampled alike.
  uint32_t runs = 0;
};


struct GrowBuf {
  void* p = nullptr;
  size_t bytes = 0;
};

struct QueryEngine;  // batched queries (query.hip)

}  // namespace cytom

struct cytom_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
  uint64_t V = 0;
  uint32_t vpt = 0, T = 0, log2vpt = 0;
  uint64_t n_tiles = 0;
  uint32_t max_iter = 100;
  double delta = 0.0;
  bool weighted = false;
  // store
  uint64_t* tile_off = nullptr;
  uint32_t* tile_cap = nullptr;
  uint32_t* tile_cnt = nullptr;
  uint32_t* tile_live = nullptr;
  uint32_t* tile_sorted = nullptr;
  uint32_t* pool = nullptr;
  uint32_t* wpool = nullptr;
  uint64_t pool_size = 0;
  uint32_t* outdeg = nullptr;
  uint32_t* indeg = nullptr;
  uint64_t live_edges = 0;
  // device scalars: [0] pool_top, [1..] scratch counters
  unsigned long long* d_misc = nullptr;
  unsigned long long* h_misc = nullptr;  // pinned: [0, 64) the control words, [64, 112) decisions
  uint32_t* h_pflags = nullptr;  // pinned per-program batch rules: [p] force, [64 + p] found_forces
  unsigned long long* h_scalar = nullptr;  // pinned landing slots of small reads (8)
  // ingest scratch
  cytom::GrowBuf g_keys, g_vals, g_keys2, g_vals2, g_flags, g_scan, g_seg_start, g_seg_tile,
      g_reloc_cap, g_reloc_list, g_ins_slot, g_scan_tmp, g_hist, g_in_src, g_in_dst, g_in_w,
      g_del_src, g_del_dst, g_del_w, g_tab_key, g_tab_head, g_node_next, g_claimed,
      g_tile_flag, g_del_tiles, g_found, g_tile_scan, g_rs_e, g_rs_w, g_rs_off, g_rec, g_del_sig,
      g_pflags;  // per program: tight-deletion flags, decisions, rules
  // pinned staging of host records (two buffers used in turn)
  void* h_stage[2] = {nullptr, nullptr};
  hipEvent_t stage_ev[2] = {nullptr, nullptr};
  bool stage_busy[2] = {false, false};
  int stage_next = 0;
  // execution work lists
  cytom::GrowBuf g_nch, g_cbase, g_chunks, g_colj, g_jbase, g_jobs;
  cytom::GrowBuf g_bands;  // source-band boundaries (bands + 1 source rows) and row slot sums
  cytom::GrowBuf g_grp;    // source groups of every 64-slot pool row
  cytom::GrowBuf g_nrow, g_rowb, g_rows;  // rows per chunk, their scan, the row descriptors
  cytom::GrowBuf g_jobs_raw, g_jsort;     // jobs before the largest-first sort, its scratch
  cytom::GrowBuf g_del_bloom;             // Bloom filter of a batch's deletion keys
  bool jobs_dirty = true;
  uint64_t n_chunks = 0, n_cells = 0, n_jobs = 0;
  uint32_t n_bands = 1;
  uint32_t chunk_slots = 512;   // slots per chunk
  uint32_t job_chunks = 0;      // chunks per column job (0: sized from the store)
  uint32_t source_bands = 1;    // source-row bands, one job queue each (see build_jobs)
  // A tile is sort-compacted once its unsorted tail reaches 256 slots and
  // resort_pct percent of its sorted prefix (CYTOM_RESORT_OFF = never).
  uint32_t resort_pct = 25;
  uint32_t resort_max_slots = 1u << 26;  // staged slots per pass (plus one tile)
  uint64_t tiles_resorted = 0, slots_resorted = 0;
  uint32_t gather_threads = 512;
  uint32_t gather_blocks = 0;   // persistent grid (0 = the resident blocks per CU, at most 2)
  uint32_t gather_per_cu[8] = {};  // resident gather blocks per CU by kind (0 = not asked yet)
  // PageRank-delta sums are exact int64 fixed point with this many fraction
  // bits: 60 - ceil(log2(2V)) keeps any vertex's sum (bounded by 2V) in range.
  int prd_shift = 36;
  int num_cus = 256;
  std::vector<cytom::Program> progs;
  cytom::QueryEngine* qe = nullptr;
  std::string err;
  uint64_t batches = 0;
  // Optional per-kernel timing of the gather launches (HIP events on the
  // context stream), accumulated per program kind.
  bool profile = false;
  uint32_t profile_every = 1;  // time every n-th gather launch of a run (phase rotating per run)
  double gather_ms[5] = {0, 0, 0, 0, 0};
  uint64_t gather_launches[5] = {0, 0, 0, 0, 0};
  double gather_alg_bytes[5] = {0, 0, 0, 0, 0};
  // Step graphs (cytom_set_step_graphs): the enqueue phase of cytom_step
  // recorded as HIP graphs, each keyed by every host value its launches
  // read, replayed while the key holds (least recently used evicted).
  struct StepGraph {
    std::vector<uint64_t> key;
    hipGraphExec_t exec = nullptr;
    uint64_t used = 0;
  };
  bool step_graphs = false;
  // Steps run their programs concurrently, each on a stream of its own
  // (forked from the context's stream after the update batch).
  bool concurrent_runs = true;
  std::vector<hipStream_t> run_streams;
  hipEvent_t fork_ev = nullptr;   // the work lists are built (PageRank-delta runs may start)
  hipEvent_t fork_ev2 = nullptr;  // the row source groups are built too (the other runs)
  // While a step graph is being recorded the enqueue functions leave their
  // timing events out (events recorded inside a graph carry no timestamps);
  // the step records them around the graph launches instead.
  bool capturing = false;
  std::vector<StepGraph> graphs;
  uint64_t graph_captures = 0, graph_replays = 0, graph_clock = 0;

  cytom::StoreView view() const {
    cytom::StoreView s;
    s.V = V;
    s.vpt = vpt;
    s.T = T;
    s.log2vpt = log2vpt;
    s.pad = 0;
    s.n_tiles = n_tiles;
    s.tile_off = tile_off;
    s.tile_cap = tile_cap;
    s.tile_cnt = tile_cnt;
    s.tile_live = tile_live;
    s.tile_sorted = tile_sorted;
    s.pool = pool;
    s.wpool = wpool;
    s.outdeg = outdeg;
    s.indeg = indeg;
    return s;
  }
};

namespace cytom {

// The stream a program's iterations are launched on.
inline hipStream_t prog_stream(const cytom_ctx* c, const Program& P) {
  return P.stream ? P.stream : c->stream;
}

int set_err(cytom_ctx* ctx, int code, const std::string& msg);
int hip_fail(cytom_ctx* ctx, hipError_t e, const char* what);
// Grow-only device buffer; returns 0 or a negative error.
int ensure(cytom_ctx* ctx, GrowBuf& b, size_t bytes);

#define CY_HIP(ctx, call)                              \
  do {                                                 \
    hipError_t e_ = (call);                            \
    if (e_ != hipSuccess) return cytom::hip_fail(ctx, e_, #call); \
  } while (0)

// ------------------------------------------------------------ primitives
// Exclusive scan of n values into out[0..n]; out[n] holds the total.
// tmp must hold scan_tmp_bytes(n) bytes (valid for both widths).
size_t scan_tmp_bytes(uint64_t n);
void scan_exclusive_u32(const uint32_t* in, uint32_t* out, uint64_t n, void* tmp,
                        hipStream_t st);
void scan_exclusive_u64(const uint64_t* in, uint64_t* out, uint64_t n, void* tmp,
                        hipStream_t st);
// The same with the length read from device memory (*n_dev <= n_max) when
// the scan runs; the launch and tmp are sized for n_max.
void scan_exclusive_u64_devn(const uint64_t* in, uint64_t* out, const uint32_t* n_dev,
                             uint64_t n_max, void* tmp, hipStream_t st);
// Fills n words with v (a kernel: usable inside captured graphs, where
// memset nodes are avoided).
void fill_u32(uint32_t* p, uint64_t n, uint32_t v, hipStream_t st);
void fill_u64(uint64_t* p, uint64_t n, uint64_t v, hipStream_t st);
// Several fills in one launch (n words of v each, 8-byte words when wide).
struct FillOp {
  void* p;
  uint64_t n;
  uint64_t v;
  uint32_t wide;
};
constexpr int MAX_FILLS = 8;
struct FillSet {
  FillOp op[MAX_FILLS];
};
void fill_many(const FillOp* ops, int n_ops, hipStream_t st);
// Stable LSD radix sort of (key, value) pairs on the low key_bits bits.
// Result lands in (k_out, v_out); (k_tmp, v_tmp) are scratch of n entries;
// hist must hold radix_hist_bytes(n) bytes.
size_t radix_hist_bytes(uint64_t n);
void radix_sort_pairs(const uint32_t* k_in, const uint32_t* v_in, uint32_t* k_out,
                      uint32_t* v_out, uint32_t* k_tmp, uint32_t* v_tmp, uint64_t n,
                      uint32_t key_bits, void* hist, void* scan_tmp, hipStream_t st);
// The same with the length read from device memory (*n_dev <= n_max) when
// the sort runs; launches, hist and scan_tmp are sized for n_max.
void radix_sort_pairs_devn(const uint32_t* k_in, const uint32_t* v_in, uint32_t* k_out,
                           uint32_t* v_out, uint32_t* k_tmp, uint32_t* v_tmp,
                           const uint32_t* n_dev, uint64_t n_max, uint32_t key_bits, void* hist,
                           void* scan_tmp, hipStream_t st);
void scan_exclusive_u32_devn(const uint32_t* in, uint32_t* out, const uint32_t* n_dev,
                             uint64_t n_max, void* tmp, hipStream_t st);

// ------------------------------------------------------------ stream gen
void launch_rmat(uint32_t scale, uint64_t seed, uint64_t first, const uint64_t* index,
                 uint64_t n, uint32_t* src, uint32_t* dst, uint32_t* w, hipStream_t st);

// ------------------------------------------------------------ store ops
int ingest_batch(cytom_ctx* ctx, const uint32_t* d_src, const uint32_t* d_dst,
                 const uint32_t* d_w, uint64_t n_ins, const uint32_t* d_dsrc,
                 const uint32_t* d_ddst, const uint32_t* d_dw, uint64_t n_del,
                 cytom_batch_stats* st);
// The stages of ingest_batch: ingest_prepare sizes the scratch (host only),
// ingest_enqueue launches the batch with no host reads, and after one
// synchronisation ingest_finish returns 0, CYTOM_EINVAL (refused) or 1 (the
// pool must grow by ingest_need() slots plus the live edges: nothing was
// applied).
int ingest_prepare(cytom_ctx* ctx, uint64_t n_ins, uint64_t n_del);
void ingest_enqueue(cytom_ctx* ctx, const uint32_t* d_src, const uint32_t* d_dst,
                    const uint32_t* d_w, uint64_t n_ins, const uint32_t* d_dsrc,
                    const uint32_t* d_ddst, const uint32_t* d_dw, uint64_t n_del);
int ingest_finish(cytom_ctx* ctx, uint64_t n_ins, uint64_t n_del, cytom_batch_stats* bst);
int64_t export_edges(cytom_ctx* ctx, uint32_t* src, uint32_t* dst, uint32_t* w, uint64_t cap);

// ------------------------------------------------------------ executors
int program_alloc(cytom_ctx* ctx, Program& P);
void program_free(Program& P);
// Re-runs init_vertices; with a gate, only when the batch decided so on the
// device (*gate == 1; see k_init_min).
int program_init(cytom_ctx* ctx, Program& P, const unsigned long long* gate = nullptr);
// Replaces a BFS/SSSP source or the PPR teleport set (host ids, validated).
int program_set_sources(cytom_ctx* ctx, Program& P, const uint32_t* sources, uint32_t n);
int run_program(cytom_ctx* ctx, Program& P, cytom_run_stats* st);
// The stages of a run, for callers that enqueue several programs before one
// host check: run_open resets the run state, run_group enqueues the next
// iteration group (the first one also computes the frontier) and its
// counter readback, run_needs_more (after a synchronisation) tells whether
// another group is needed, and run_finish fills the stats.
int program_events(cytom_ctx* ctx, Program& P);
void run_open(Program& P);
void run_group(cytom_ctx* ctx, Program& P);
// The host-side effects of run_group without its launches (a replayed step
// graph holds them), and the length of the group run_group would enqueue.
void run_group_replayed(const cytom_ctx* ctx, Program& P);
uint32_t run_group_length(const cytom_ctx* ctx, const Program& P);
bool run_needs_more(const cytom_ctx* ctx, Program& P);
int run_finish(cytom_ctx* ctx, Program& P, cytom_run_stats* out);
// Sets the gather kernels' LDS limit and persistent grid for the context.
int configure_gather(cytom_ctx* ctx);
// Rebuilds the work lists of the current store (chunks of chunk_slots slots
// per tile in g_chunks, the column jobs, the row source groups):
// prepare_jobs sizes the buffers by upper bounds (host only), enqueue_jobs
// launches the device-driven build, build_jobs does both.
int build_jobs(cytom_ctx* ctx);
int prepare_jobs(cytom_ctx* ctx);
// with_groups = false leaves out the row source groups (enqueue_row_groups),
// which only the BFS / CC / SSSP gathers read.
void enqueue_jobs(cytom_ctx* ctx, bool with_groups = true);
void enqueue_row_groups(cytom_ctx* ctx);
// The current chunk count (one host round trip).
int64_t jobs_chunk_count(cytom_ctx* ctx);

// ------------------------------------------------------------ program adjoint
// Reverse mode of a registered CYTOM_PPR program on the current graph; the
// live-edge count, or a negative error code.
int64_t ppr_program_adjoint(cytom_ctx* ctx, Program& P, const float* d_g, float tol,
                            uint32_t max_iterations, float* d_ds, uint32_t* d_src,
                            uint32_t* d_dst, float* d_dw, uint64_t cap, uint32_t* iterations);

// ------------------------------------------------------------ batched queries
int batched_execute(cytom_ctx* ctx, const cytom_query* q, uint32_t nq, cytom_run_stats* stats);
int read_query(cytom_ctx* ctx, uint32_t query, void* out, size_t bytes);
void query_free(cytom_ctx* ctx);

// ------------------------------------------------------------ host records, results
// Inserts/deletes records {u32 src, u32 tgt[, f32 weight]} (rec_bytes 8 or 12)
// from host memory through the pinned staging buffers.
int ingest_records(cytom_ctx* ctx, const void* rec, uint32_t rec_bytes, uint64_t n_ins,
                   uint32_t default_w, const void* del_rec, uint64_t n_del,
                   cytom_batch_stats* stats);
// Writes the program's vertex values in the reference's results layout; with
// out == nullptr returns the byte count only.
int64_t read_results(cytom_ctx* ctx, const Program& P, void* out, size_t bytes);
void io_free(cytom_ctx* ctx);
// Re-execution tests of the found deletions (*n_dev of at most n_max)
// against the program's values; sets *flag (device) to 1 when a
// re-initialisation is needed.
void launch_reexec_check(cytom_ctx* ctx, Program& P, const uint32_t* found, uint64_t n_max,
                         const unsigned long long* n_dev, unsigned long long* flag);
// edgeChanged(INSERT) hooks of the batch's surviving inserts and the
// PRD edgeChanged(DELETE) side effect of the found deletions (*n_found_dev
// of at most n_found_max); with a gate, only when the batch decided on
// hooks (*gate == 0).
void launch_hooks(cytom_ctx* ctx, Program& P, const uint32_t* d_src, const uint32_t* d_dst,
                  const uint32_t* d_w, const uint64_t* ins_slot, uint64_t n_ins,
                  const uint32_t* found, uint64_t n_found_max,
                  const unsigned long long* n_found_dev, const unsigned long long* gate);

}  // namespace cytom
