// C-ABI host engine: context life cycle, algorithm store, batch control.
//
// Replaces the job driver execute<VertexType, Algorithm, is_weighted>
// (src/include/main/execute.h:534-771), the EdgeInserter batch loop
// (edge-inserter.h:53-194) and the plugin registration performed by the
// composeJobs pass (src/lib/llvm-pass/ParallizeAlgorithm.cpp:68-154):
// programs are registered through cytom_program_register instead of LLVM
// bitcode surgery.
#include <algorithm>
#include <cstring>
#include <functional>
#include <new>

#include "internal.h"
#include "tuning_table.h"

namespace cytom {

int set_err(cytom_ctx* ctx, int code, const std::string& msg) {
  if (ctx) ctx->err = msg;
  return code;
}

int hip_fail(cytom_ctx* ctx, hipError_t e, const char* what) {
  std::string m = std::string(what) + ": " + hipGetErrorString(e);
  if (ctx) ctx->err = m;
  return e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation ? CYTOM_ENOMEM : CYTOM_EDEVICE;
}

int ensure(cytom_ctx* ctx, GrowBuf& b, size_t bytes) {
  if (bytes <= b.bytes) return 0;
  if (b.p) {
    (void)hipStreamSynchronize(ctx->stream);
    (void)hipFree(b.p);
    b.p = nullptr;
    b.bytes = 0;
  }
  // Doubling keeps re-allocations (a device-wide synchronisation each) to a
  // few per buffer over a stream whose graph keeps growing.
  const size_t want = std::max<size_t>(bytes, 65536) * 2;
  if (hipMalloc(&b.p, want) != hipSuccess) {
    b.p = nullptr;
    return set_err(ctx, CYTOM_ENOMEM, "scratch allocation failed");
  }
  b.bytes = want;
  return 0;
}

}  // namespace cytom

using namespace cytom;

namespace {

void free_buf(GrowBuf& b) {
  if (b.p) (void)hipFree(b.p);
  b.p = nullptr;
  b.bytes = 0;
}

// Every grow-only scratch buffer of a context.
std::vector<GrowBuf*> grow_bufs(cytom_ctx* c) {
  return {&c->g_keys,      &c->g_vals,       &c->g_keys2,    &c->g_vals2,     &c->g_flags,
          &c->g_scan,      &c->g_seg_start,  &c->g_seg_tile, &c->g_reloc_cap, &c->g_reloc_list,
          &c->g_ins_slot,  &c->g_scan_tmp,   &c->g_hist,     &c->g_in_src,    &c->g_in_dst,
          &c->g_in_w,      &c->g_del_src,    &c->g_del_dst,  &c->g_del_w,     &c->g_tab_key,
          &c->g_tab_head,  &c->g_node_next,  &c->g_claimed,  &c->g_tile_flag, &c->g_del_tiles,
          &c->g_found,     &c->g_tile_scan,  &c->g_nch,      &c->g_cbase,     &c->g_chunks,
          &c->g_colj,      &c->g_jbase,      &c->g_jobs,     &c->g_bands,     &c->g_grp,
          &c->g_rs_e,      &c->g_rs_w,       &c->g_rs_off,   &c->g_pflags,    &c->g_rec,
          &c->g_del_sig,   &c->g_nrow,       &c->g_rowb,     &c->g_rows,      &c->g_jobs_raw,
          &c->g_jsort,     &c->g_del_bloom};
}

void free_graphs(cytom_ctx* c) {
  for (auto& g : c->graphs)
    if (g.exec) (void)hipGraphExecDestroy(g.exec);
  c->graphs.clear();
}

void free_ctx(cytom_ctx* c) {
  if (!c) return;
  (void)hipSetDevice(c->device);
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  free_graphs(c);
  for (hipStream_t st : c->run_streams) (void)hipStreamDestroy(st);
  if (c->fork_ev) (void)hipEventDestroy(c->fork_ev);
  if (c->fork_ev2) (void)hipEventDestroy(c->fork_ev2);
  for (auto& P : c->progs) program_free(P);
  query_free(c);
  io_free(c);
  void* ps[] = {c->tile_off, c->tile_cap, c->tile_cnt, c->tile_live, c->tile_sorted,
                c->pool,     c->wpool,    c->outdeg,   c->indeg,    c->d_misc};
  for (void* p : ps)
    if (p) (void)hipFree(p);
  if (c->h_misc) (void)hipHostFree(c->h_misc);
  if (c->h_scalar) (void)hipHostFree(c->h_scalar);
  if (c->h_pflags) (void)hipHostFree(c->h_pflags);
  for (GrowBuf* b : grow_bufs(c)) free_buf(*b);
  for (auto& e : c->ev)
    if (e) (void)hipEventDestroy(e);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
}

// Fraction bits of the PageRank-delta fixed-point sums that keep any
// vertex's sum (at most 2V in magnitude) inside int64: 60 - ceil(log2 2V).
int max_prd_shift(uint64_t V) {
  int lg = 0;
  while ((1ull << lg) < V) ++lg;
  return 60 - (lg + 1);
}

#define CY_CHECK_CTX(ctx)                                   \
  do {                                                      \
    if (!(ctx)) return CYTOM_EINVAL;                        \
    if (hipSetDevice((ctx)->device) != hipSuccess)          \
      return set_err(ctx, CYTOM_EDEVICE, "hipSetDevice failed"); \
  } while (0)

}  // namespace

struct cytom_stream {
  uint64_t hash_del = 0;
  uint64_t next_index = 0;
  uint64_t batch = 0;
  std::vector<uint64_t> live;
};

extern "C" {

namespace {

// The autotuned row for a context of V vertices (see tuning_table.h).
const TuneRow* tune_row(uint64_t V, uint64_t edge_capacity) {
  uint32_t lv = 0;
  while ((1ull << lv) < V) ++lv;
  const uint64_t ef = edge_capacity ? std::max<uint64_t>(edge_capacity / V, 1) : 16;
  uint64_t best_gap = ~0ull;
  for (const TuneRow& r : TUNE_TABLE) {
    const uint64_t gap = r.edge_factor > ef ? r.edge_factor - ef : ef - r.edge_factor;
    best_gap = std::min(best_gap, gap);
  }
  const TuneRow* below = nullptr;
  const TuneRow* above = nullptr;
  for (const TuneRow& r : TUNE_TABLE) {
    const uint64_t gap = r.edge_factor > ef ? r.edge_factor - ef : ef - r.edge_factor;
    if (gap != best_gap) continue;
    if (r.log2v <= lv && (!below || r.log2v > below->log2v)) below = &r;
    if (r.log2v > lv && (!above || r.log2v < above->log2v)) above = &r;
  }
  return below ? below : above;
}

}  // namespace

int cytom_create(const cytom_config* cfg, cytom_ctx** out) {
  if (!cfg || !out) return CYTOM_EINVAL;
  *out = nullptr;
  if (cfg->n_vertices == 0 || cfg->n_vertices > MAX_VERTICES) return CYTOM_EINVAL;
  uint64_t V = 64;
  while (V < cfg->n_vertices) V <<= 1;
  uint32_t vpt = cfg->vertices_per_tile;
  const TuneRow* tune = nullptr;
  if (vpt == 0) {
    tune = tune_row(V, cfg->edge_capacity);
    vpt = tune->vpt;
    // A row measured on a smaller graph: wider tiles until the tile count
    // stays within 2^26 (tile metadata and work lists grow with it).
    while (vpt < 16384 && vpt < V && (V / vpt) * (V / vpt) > (1ull << 26)) vpt <<= 1;
  }
  // 16-bit local ids in the edge slots and an 8-byte LDS column window.
  if (vpt < 64 || vpt > 16384 || (vpt & (vpt - 1))) return CYTOM_EINVAL;
  if (vpt > V) vpt = (uint32_t)V;
  if ((V / vpt) * (V / vpt) > (1ull << 30)) return CYTOM_EINVAL;  // tile ids stay 32-bit keys
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return CYTOM_EDEVICE;
  if ((int)cfg->device >= ndev) return CYTOM_EINVAL;
  auto* c = new (std::nothrow) cytom_ctx();
  if (!c) return CYTOM_ENOMEM;
  c->device = (int)cfg->device;
  c->V = V;
  c->vpt = vpt;
  c->T = (uint32_t)(V / vpt);
  c->log2vpt = 0;
  while ((1u << c->log2vpt) < vpt) ++c->log2vpt;
  c->n_tiles = (uint64_t)c->T * c->T;
  c->prd_shift = max_prd_shift(V);
  c->max_iter = cfg->max_iterations ? cfg->max_iterations : 100;
  if (tune) {
    c->chunk_slots = tune->chunk_slots;
    c->source_bands = tune->source_bands;
  }
  c->delta = cfg->delta;
  c->weighted = cfg->weighted != 0;
  auto fail = [&](int code) {
    free_ctx(c);
    return code;
  };
  if (hipSetDevice(c->device) != hipSuccess) return fail(CYTOM_EDEVICE);
  if (hipStreamCreate(&c->stream) != hipSuccess) return fail(CYTOM_EDEVICE);
  for (auto& e : c->ev)
    if (hipEventCreate(&e) != hipSuccess) return fail(CYTOM_EDEVICE);
  int cus = 0;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, c->device) == hipSuccess &&
      cus > 0)
    c->num_cus = cus;
  const uint64_t nt = c->n_tiles;
  bool ok = hipMalloc(&c->tile_off, nt * 8) == hipSuccess &&
            hipMalloc(&c->tile_cap, nt * 4) == hipSuccess &&
            hipMalloc(&c->tile_cnt, nt * 4) == hipSuccess &&
            hipMalloc(&c->tile_live, nt * 4) == hipSuccess &&
            hipMalloc(&c->tile_sorted, nt * 4) == hipSuccess &&
            hipMalloc(&c->outdeg, V * 4) == hipSuccess &&
            hipMalloc(&c->indeg, V * 4) == hipSuccess &&
            hipMalloc(&c->d_misc, 64 * 8) == hipSuccess &&
            hipHostMalloc(&c->h_misc, 128 * 8, hipHostMallocDefault) == hipSuccess &&
            hipHostMalloc(&c->h_pflags, 128 * 4, hipHostMallocDefault) == hipSuccess &&
            hipHostMalloc(&c->h_scalar, 8 * 8, hipHostMallocDefault) == hipSuccess &&
            true;
  if (!ok) return fail(CYTOM_ENOMEM);
  c->pool_size = cfg->edge_capacity ? cfg->edge_capacity : (1ull << 22);
  if (hipMalloc(&c->pool, c->pool_size * 4) != hipSuccess) return fail(CYTOM_ENOMEM);
  if (c->weighted && hipMalloc(&c->wpool, c->pool_size * 4) != hipSuccess)
    return fail(CYTOM_ENOMEM);
  ok = hipMemsetAsync(c->tile_off, 0, nt * 8, c->stream) == hipSuccess &&
       hipMemsetAsync(c->tile_cap, 0, nt * 4, c->stream) == hipSuccess &&
       hipMemsetAsync(c->tile_cnt, 0, nt * 4, c->stream) == hipSuccess &&
       hipMemsetAsync(c->tile_live, 0, nt * 4, c->stream) == hipSuccess &&
       hipMemsetAsync(c->tile_sorted, 0, nt * 4, c->stream) == hipSuccess &&
       hipMemsetAsync(c->outdeg, 0, V * 4, c->stream) == hipSuccess &&
       hipMemsetAsync(c->indeg, 0, V * 4, c->stream) == hipSuccess &&
       hipMemsetAsync(c->d_misc, 0, 64 * 8, c->stream) == hipSuccess &&
       hipStreamSynchronize(c->stream) == hipSuccess;
  if (!ok) return fail(CYTOM_EDEVICE);
  if (configure_gather(c)) return fail(CYTOM_EDEVICE);
  *out = c;
  return 0;
}

int cytom_destroy(cytom_ctx* ctx) {
  if (!ctx) return CYTOM_EINVAL;
  free_ctx(ctx);
  return 0;
}

const char* cytom_last_error(const cytom_ctx* ctx) {
  if (!ctx) return "null context";
  return ctx->err.c_str();
}

int cytom_geometry(const cytom_ctx* ctx, uint64_t* n_vertices, uint32_t* vertices_per_tile,
                   uint32_t* tiles_per_dim) {
  if (!ctx) return CYTOM_EINVAL;
  if (n_vertices) *n_vertices = ctx->V;
  if (vertices_per_tile) *vertices_per_tile = ctx->vpt;
  if (tiles_per_dim) *tiles_per_dim = ctx->T;
  return 0;
}

int cytom_program_statics(uint32_t kind, cytom_program_desc* out) {
  if (!out || kind > CYTOM_PPR) return CYTOM_EINVAL;
  *out = cytom_program_desc{};
  out->kind = kind;
  out->flags = CYTOM_STATICS_GIVEN | CYTOM_NEEDS_SOURCES_ACTIVE;
  if (kind != CYTOM_SSSP) out->flags |= CYTOM_USE_CHANGED;  // sssp.cc:30 states false
  if (prd_like((int)kind)) {
    // pagerank-delta.cc:54-57: {delta_sum, rank, delta} = {0, 0, 0}
    out->vertex_bytes = 12;
    out->n_fields = 3;
    out->first_iteration =
        kind == CYTOM_PRD ? CYTOM_FIRST_ITERATION_PRD_ALL : CYTOM_FIRST_ITERATION_NONE;
  } else {
    out->vertex_bytes = 4;
    out->n_fields = 1;
    out->first_iteration = CYTOM_FIRST_ITERATION_NONE;
    // bfs.cc:22, cc.cc:22: UINT32_MAX; sssp.cc:22: FLT_MAX
    out->neutral_element[0] = kind == CYTOM_SSSP ? 3.4028234663852886e38 : 4294967295.0;
  }
  return 0;
}

int cytom_program_register(cytom_ctx* ctx, const cytom_program_desc* desc, int* prog_id) {
  CY_CHECK_CTX(ctx);
  if (!desc || !prog_id) return set_err(ctx, CYTOM_EINVAL, "null descriptor");
  if (desc->kind > CYTOM_PPR) return set_err(ctx, CYTOM_EINVAL, "unknown program kind");
  if (desc->flags & CYTOM_STATICS_GIVEN) {
    cytom_program_desc k;
    cytom_program_statics(desc->kind, &k);
    const uint32_t must = CYTOM_NEEDS_SOURCES_ACTIVE | CYTOM_NEEDS_TARGETS_ACTIVE | CYTOM_ASYNC;
    if ((desc->flags & must) != (k.flags & must))
      return set_err(ctx, CYTOM_EINVAL,
                     "statics: the engine's kernels need NEEDS_SOURCES_ACTIVE set and "
                     "NEEDS_TARGETS_ACTIVE and ASYNC clear");
    if (desc->vertex_bytes != k.vertex_bytes || desc->n_fields != k.n_fields)
      return set_err(ctx, CYTOM_EINVAL, "statics: VertexType differs from the kind's");
    if (desc->first_iteration != k.first_iteration)
      return set_err(ctx, CYTOM_EINVAL, "statics: first-iteration rule differs from the kind's");
    for (int f = 0; f < 3; ++f)
      if (desc->neutral_element[f] != k.neutral_element[f])
        return set_err(ctx, CYTOM_EINVAL, "statics: neutral_element differs from the kind's");
  }
  if (desc->kind == CYTOM_SSSP && !ctx->weighted)
    return set_err(ctx, CYTOM_EINVAL, "SSSP needs a weighted graph");
  if ((desc->kind == CYTOM_BFS || desc->kind == CYTOM_SSSP || desc->kind == CYTOM_PPR) &&
      desc->source >= ctx->V)
    return set_err(ctx, CYTOM_EINVAL, "source vertex out of range");
  if (ctx->progs.size() >= 48) return set_err(ctx, CYTOM_ERANGE, "too many programs");
  Program P;
  P.kind = (int)desc->kind;
  P.source = desc->source;
  int rc = program_alloc(ctx, P);
  if (rc) return rc;
  rc = P.kind == CYTOM_PPR ? program_set_sources(ctx, P, &desc->source, 1) : program_init(ctx, P);
  if (rc) {
    program_free(P);
    return rc;
  }
  ctx->progs.push_back(P);
  *prog_id = (int)ctx->progs.size() - 1;
  if (hipStreamSynchronize(ctx->stream) != hipSuccess)
    return set_err(ctx, CYTOM_EDEVICE, "program init failed");
  return 0;
}

int cytom_program_init(cytom_ctx* ctx, int prog_id) {
  CY_CHECK_CTX(ctx);
  if (prog_id < 0 || prog_id >= (int)ctx->progs.size())
    return set_err(ctx, CYTOM_EINVAL, "bad program id");
  int rc = program_init(ctx, ctx->progs[prog_id]);
  if (rc) return rc;
  CY_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return 0;
}

int cytom_program_set_sources(cytom_ctx* ctx, int prog_id, const uint32_t* sources, uint32_t n) {
  CY_CHECK_CTX(ctx);
  if (prog_id < 0 || prog_id >= (int)ctx->progs.size())
    return set_err(ctx, CYTOM_EINVAL, "bad program id");
  int rc = program_set_sources(ctx, ctx->progs[prog_id], sources, n);
  if (rc) return rc;
  CY_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return 0;
}

int cytom_update_batch_device(cytom_ctx* ctx, const uint32_t* d_src, const uint32_t* d_dst,
                              const uint32_t* d_w, uint64_t n_ins, const uint32_t* d_del_src,
                              const uint32_t* d_del_dst, const uint32_t* d_del_w,
                              uint64_t n_del, cytom_batch_stats* stats) {
  CY_CHECK_CTX(ctx);
  if ((n_ins && (!d_src || !d_dst)) || (n_del && (!d_del_src || !d_del_dst)))
    return set_err(ctx, CYTOM_EINVAL, "null update arrays");
  return ingest_batch(ctx, d_src, d_dst, d_w, n_ins, d_del_src, d_del_dst, d_del_w, n_del, stats);
}

int cytom_update_batch(cytom_ctx* ctx, const uint32_t* src, const uint32_t* dst,
                       const uint32_t* w, uint64_t n_ins, const uint32_t* del_src,
                       const uint32_t* del_dst, const uint32_t* del_w, uint64_t n_del,
                       cytom_batch_stats* stats) {
  CY_CHECK_CTX(ctx);
  if ((n_ins && (!src || !dst)) || (n_del && (!del_src || !del_dst)))
    return set_err(ctx, CYTOM_EINVAL, "null update arrays");
  int rc;
  const uint64_t ni = std::max<uint64_t>(n_ins, 1), nd = std::max<uint64_t>(n_del, 1);
  if ((rc = ensure(ctx, ctx->g_in_src, ni * 4)) || (rc = ensure(ctx, ctx->g_in_dst, ni * 4)) ||
      (rc = ensure(ctx, ctx->g_in_w, ni * 4)) || (rc = ensure(ctx, ctx->g_del_src, nd * 4)) ||
      (rc = ensure(ctx, ctx->g_del_dst, nd * 4)) || (rc = ensure(ctx, ctx->g_del_w, nd * 4)))
    return rc;
  hipStream_t st = ctx->stream;
  uint32_t* ds = static_cast<uint32_t*>(ctx->g_in_src.p);
  uint32_t* dd = static_cast<uint32_t*>(ctx->g_in_dst.p);
  uint32_t* dw = w ? static_cast<uint32_t*>(ctx->g_in_w.p) : nullptr;
  uint32_t* xs = static_cast<uint32_t*>(ctx->g_del_src.p);
  uint32_t* xd = static_cast<uint32_t*>(ctx->g_del_dst.p);
  uint32_t* xw = del_w ? static_cast<uint32_t*>(ctx->g_del_w.p) : nullptr;
  if (n_ins) {
    CY_HIP(ctx, hipMemcpyAsync(ds, src, n_ins * 4, hipMemcpyHostToDevice, st));
    CY_HIP(ctx, hipMemcpyAsync(dd, dst, n_ins * 4, hipMemcpyHostToDevice, st));
    if (w) CY_HIP(ctx, hipMemcpyAsync(dw, w, n_ins * 4, hipMemcpyHostToDevice, st));
  }
  if (n_del) {
    CY_HIP(ctx, hipMemcpyAsync(xs, del_src, n_del * 4, hipMemcpyHostToDevice, st));
    CY_HIP(ctx, hipMemcpyAsync(xd, del_dst, n_del * 4, hipMemcpyHostToDevice, st));
    if (del_w) CY_HIP(ctx, hipMemcpyAsync(xw, del_w, n_del * 4, hipMemcpyHostToDevice, st));
  }
  return ingest_batch(ctx, n_ins ? ds : nullptr, n_ins ? dd : nullptr, n_ins ? dw : nullptr,
                      n_ins, n_del ? xs : nullptr, n_del ? xd : nullptr, n_del ? xw : nullptr,
                      n_del, stats);
}

int cytom_update_batch_records(cytom_ctx* ctx, const void* records, uint32_t record_bytes,
                               uint64_t n_ins, uint32_t default_weight,
                               const void* del_records, uint64_t n_del,
                               cytom_batch_stats* stats) {
  CY_CHECK_CTX(ctx);
  return ingest_records(ctx, records, record_bytes, n_ins, default_weight, del_records, n_del,
                        stats);
}

int cytom_execute(cytom_ctx* ctx, int prog_id, cytom_run_stats* stats) {
  CY_CHECK_CTX(ctx);
  if (prog_id < 0 || prog_id >= (int)ctx->progs.size())
    return set_err(ctx, CYTOM_EINVAL, "bad program id");
  return run_program(ctx, ctx->progs[prog_id], stats);
}

namespace {

// Keys of step graphs: every host value an enqueue phase passes to a launch
// or a copy.  Equal keys record identical graphs, so a graph is replayed only
// while its key holds.  key_ctx covers the context (geometry, tuning, the
// store's and the scratch buffers' addresses and sizes), key_prog one
// program (its buffers, sources and buffer parity).
void key_ctx(cytom_ctx* c, std::vector<uint64_t>& k) {
  uint64_t dbits;
  memcpy(&dbits, &c->delta, 8);
  for (uint64_t v : {c->V, (uint64_t)c->vpt, (uint64_t)c->T, (uint64_t)c->log2vpt, c->n_tiles,
                     (uint64_t)c->max_iter, dbits, (uint64_t)c->weighted, c->pool_size,
                     c->n_chunks, c->n_cells, c->n_jobs, (uint64_t)c->n_bands,
                     (uint64_t)c->chunk_slots, (uint64_t)c->job_chunks, (uint64_t)c->source_bands,
                     (uint64_t)c->resort_pct, (uint64_t)c->resort_max_slots,
                     (uint64_t)c->gather_threads, (uint64_t)c->gather_blocks,
                     (uint64_t)(int64_t)c->prd_shift, (uint64_t)c->num_cus,
                     (uint64_t)c->progs.size()})
    k.push_back(v);
  for (const void* p : {(const void*)c->stream, (const void*)c->tile_off,
                        (const void*)c->tile_cap, (const void*)c->tile_cnt,
                        (const void*)c->tile_live, (const void*)c->tile_sorted,
                        (const void*)c->pool, (const void*)c->wpool, (const void*)c->outdeg,
                        (const void*)c->indeg, (const void*)c->d_misc, (const void*)c->h_misc,
                        (const void*)c->h_pflags})
    k.push_back((uint64_t)(uintptr_t)p);
  for (GrowBuf* b : grow_bufs(c)) {
    k.push_back((uint64_t)(uintptr_t)b->p);
    k.push_back(b->bytes);
  }
}

void key_prog(const Program& P, std::vector<uint64_t>& k) {
  for (uint64_t v : {(uint64_t)P.kind, (uint64_t)P.source, (uint64_t)P.source2,
                     (uint64_t)P.n_teleport, (uint64_t)P.n_qseed, (uint64_t)P.cur_act})
    k.push_back(v);
  for (const void* p :
       {(const void*)P.teleport, (const void*)P.qseed, (const void*)P.cur, (const void*)P.nxt,
        (const void*)P.fval, (const void*)P.rank, (const void*)P.delta, (const void*)P.dsum,
        (const void*)P.contrib, (const void*)P.act[0], (const void*)P.act[1],
        (const void*)P.gact[0], (const void*)P.gact[1], (const void*)P.changed,
        (const void*)P.d_iter, (const void*)P.h_iter, (const void*)P.dctl})
    k.push_back((uint64_t)(uintptr_t)p);
}

constexpr size_t MAX_STEP_GRAPHS = 32;

// Launches one enqueue phase through a step graph: replays the graph of an
// equal key and applies the host-side effects the recorded phase had
// (replayed), or records the phase (enqueue) into a new graph and launches
// it.  A failed recording restores the host state the phase changed
// (save / restore) and launches the phase directly.
int launch_recorded(cytom_ctx* ctx, hipStream_t st, const std::vector<uint64_t>& key,
                    const std::function<void()>& enqueue, const std::function<void()>& replayed,
                    const std::function<void()>& save, const std::function<void()>& restore) {
  for (auto& g : ctx->graphs) {
    if (g.key != key) continue;
    CY_HIP(ctx, hipGraphLaunch(g.exec, st));
    replayed();
    g.used = ++ctx->graph_clock;
    ++ctx->graph_replays;
    return 0;
  }
  save();
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  bool ok = hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal) == hipSuccess;
  if (ok) {
    ctx->capturing = true;
    enqueue();
    ctx->capturing = false;
    ok = hipStreamEndCapture(st, &graph) == hipSuccess && graph != nullptr;
    ok = ok && hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0) == hipSuccess;
    if (graph) (void)hipGraphDestroy(graph);
  }
  (void)hipGetLastError();
  if (!ok) {
    if (exec) (void)hipGraphExecDestroy(exec);
    restore();
    enqueue();
    return 0;
  }
  if (ctx->graphs.size() >= MAX_STEP_GRAPHS) {
    auto lru = std::min_element(ctx->graphs.begin(), ctx->graphs.end(),
                                [](const cytom_ctx::StepGraph& a, const cytom_ctx::StepGraph& b) {
                                  return a.used < b.used;
                                });
    (void)hipGraphExecDestroy(lru->exec);
    ctx->graphs.erase(lru);
  }
  cytom_ctx::StepGraph g;
  g.key = key;
  g.exec = exec;
  g.used = ++ctx->graph_clock;
  ctx->graphs.push_back(std::move(g));
  ++ctx->graph_captures;
  CY_HIP(ctx, hipGraphLaunch(exec, st));
  return 0;
}

// The enqueue phase of a step through step graphs: one graph for the batch
// and the work lists, one per listed program for its first iteration group,
// with the timing events recorded between the graph launches.
int step_graphs(cytom_ctx* ctx, const uint32_t* d_src, const uint32_t* d_dst,
                const uint32_t* d_w, uint64_t n_ins, const uint32_t* d_del_src,
                const uint32_t* d_del_dst, const uint32_t* d_del_w, uint64_t n_del,
                const int* prog_ids, const std::vector<Program*>& ps,
                const std::function<void()>& fork) {
  hipStream_t st = ctx->stream;
  int rc;
  std::vector<uint64_t> key = {1};  // the batch phase
  for (const void* a : {(const void*)d_src, (const void*)d_dst, (const void*)d_w,
                        (const void*)d_del_src, (const void*)d_del_dst, (const void*)d_del_w})
    key.push_back((uint64_t)(uintptr_t)a);
  key.push_back(n_ins);
  key.push_back(n_del);
  key_ctx(ctx, key);
  for (const Program& P : ctx->progs) key_prog(P, key);
  bool dirty = false;
  CY_HIP(ctx, hipEventRecord(ctx->ev[0], st));
  rc = launch_recorded(
      ctx, st, key,
      [&] {
        ingest_enqueue(ctx, d_src, d_dst, d_w, n_ins, d_del_src, d_del_dst, d_del_w, n_del);
        enqueue_jobs(ctx);
      },
      [&] { ctx->jobs_dirty = false; }, [&] { dirty = ctx->jobs_dirty; },
      [&] { ctx->jobs_dirty = dirty; });
  if (rc) return rc;
  CY_HIP(ctx, hipEventRecord(ctx->ev[1], st));
  fork();
  for (size_t i = 0; i < ps.size(); ++i) {
    Program& P = *ps[i];
    P.refuse = ctx->d_misc;
    run_open(P);
    key.assign({2, (uint64_t)prog_ids[i], run_group_length(ctx, P)});  // a first group
    key_ctx(ctx, key);
    key_prog(P, key);
    int cur_act = P.cur_act;
    const hipStream_t ps_st = prog_stream(ctx, P);
    CY_HIP(ctx, hipEventRecord(P.gev[0], ps_st));
    rc = launch_recorded(
        ctx, ps_st, key, [&] { run_group(ctx, P); }, [&] { run_group_replayed(ctx, P); },
        [&] { cur_act = P.cur_act; },
        [&] {
          P.cur_act = cur_act;
          run_open(P);
        });
    P.refuse = nullptr;
    if (rc) return rc;
    CY_HIP(ctx, hipEventRecord(P.gev[1], ps_st));
  }
  return 0;
}

constexpr size_t MAX_RUN_STREAMS = 4;

// The streams of a concurrent step (created once per context).
int run_streams(cytom_ctx* ctx, size_t n) {
  while (ctx->run_streams.size() < n) {
    hipStream_t s = nullptr;
    CY_HIP(ctx, hipStreamCreate(&s));
    ctx->run_streams.push_back(s);
  }
  if (!ctx->fork_ev) CY_HIP(ctx, hipEventCreateWithFlags(&ctx->fork_ev, hipEventDisableTiming));
  if (!ctx->fork_ev2) CY_HIP(ctx, hipEventCreateWithFlags(&ctx->fork_ev2, hipEventDisableTiming));
  return 0;
}

}  // namespace

// The listed programs of a call (ids in range, each at most once).
static int program_list(cytom_ctx* ctx, const int* prog_ids, uint32_t n_progs,
                 std::vector<Program*>& ps) {
  if (n_progs && !prog_ids) return set_err(ctx, CYTOM_EINVAL, "null program list");
  for (uint32_t k = 0; k < n_progs; ++k) {
    const int id = prog_ids[k];
    if (id < 0 || id >= (int)ctx->progs.size()) return set_err(ctx, CYTOM_EINVAL, "bad program id");
    for (uint32_t j = 0; j < k; ++j)
      if (prog_ids[j] == id) return set_err(ctx, CYTOM_EINVAL, "program listed twice");
    ps.push_back(&ctx->progs[id]);
  }
  return 0;
}

// edge-inserter.h:129-184: one batch, then (re-initialisation where the
// batch requires it and) every registered algorithm's run.  The batch, the
// work lists and the first iteration group of every program are enqueued
// before the host checks anything; programs that need more iterations get
// one further group each per synchronisation.
int cytom_step(cytom_ctx* ctx, const uint32_t* d_src, const uint32_t* d_dst, const uint32_t* d_w,
               uint64_t n_ins, const uint32_t* d_del_src, const uint32_t* d_del_dst,
               const uint32_t* d_del_w, uint64_t n_del, const int* prog_ids, uint32_t n_progs,
               cytom_batch_stats* batch_stats, cytom_run_stats* run_stats) {
  CY_CHECK_CTX(ctx);
  if ((n_ins && (!d_src || !d_dst)) || (n_del && (!d_del_src || !d_del_dst)))
    return set_err(ctx, CYTOM_EINVAL, "null update arrays");
  if (ctx->weighted && n_ins && !d_w)
    return set_err(ctx, CYTOM_EINVAL, "weighted graph needs insertion weights");
  std::vector<Program*> ps;
  int rc;
  if ((rc = program_list(ctx, prog_ids, n_progs, ps))) return rc;
  if ((rc = ingest_prepare(ctx, n_ins, n_del)) || (rc = prepare_jobs(ctx))) return rc;
  for (Program* P : ps)
    if ((rc = program_events(ctx, *P))) return rc;
  // Programs run concurrently after the batch, each on a stream of its own
  // (their state is disjoint; the store and the work lists are read-only
  // during runs).
  const size_t ns = ctx->concurrent_runs && ps.size() > 1 ? std::min(ps.size(), MAX_RUN_STREAMS)
                                                          : 0;
  if (ns && (rc = run_streams(ctx, ns))) return rc;
  struct StreamReset {
    std::vector<Program*>& ps;
    ~StreamReset() {
      for (Program* P : ps) P->stream = nullptr;
    }
  } reset{ps};
  // Each program's stream waits for ev (recorded on the context's stream)
  // when prd_like(kind) == prd.
  auto fork_part = [&](hipEvent_t ev, bool prd) {
    (void)hipEventRecord(ev, ctx->stream);
    for (size_t k = 0; k < ps.size(); ++k) {
      ps[k]->stream = ctx->run_streams[k % ns];
      if (prd_like(ps[k]->kind) == prd) (void)hipStreamWaitEvent(ps[k]->stream, ev, 0);
    }
  };
  auto fork = [&]() {
    if (!ns) return;
    fork_part(ctx->fork_ev, true);
    fork_part(ctx->fork_ev, false);
  };
  auto sync_all = [&]() -> int {
    CY_HIP(ctx, hipStreamSynchronize(ctx->stream));
    for (size_t i = 0; i < ns; ++i) CY_HIP(ctx, hipStreamSynchronize(ctx->run_streams[i]));
    return 0;
  };
  std::vector<int> parity(ps.size());
  for (size_t k = 0; k < ps.size(); ++k) {
    parity[k] = ps[k]->cur_act;
    run_open(*ps[k]);
  }
  // The enqueue phase: no host reads, no allocation (so it can be recorded).
  auto enqueue = [&]() {
    ingest_enqueue(ctx, d_src, d_dst, d_w, n_ins, d_del_src, d_del_dst, d_del_w, n_del);
    if (ns) {
      // PageRank-delta runs never read the row source groups: they start as
      // soon as the job lists are built, the other runs after the groups.
      enqueue_jobs(ctx, false);
      fork_part(ctx->fork_ev, true);
      enqueue_row_groups(ctx);
      fork_part(ctx->fork_ev2, false);
    } else {
      enqueue_jobs(ctx);
    }
    for (Program* P : ps) {
      P->refuse = ctx->d_misc;
      run_open(*P);
      run_group(ctx, *P);
      P->refuse = nullptr;
    }
  };
  if (ctx->step_graphs && !ctx->profile) {
    if ((rc = step_graphs(ctx, d_src, d_dst, d_w, n_ins, d_del_src, d_del_dst, d_del_w, n_del,
                          prog_ids, ps, fork)))
      return rc;
  } else {
    enqueue();
  }
  uint32_t syncs = 1;
  if ((rc = sync_all())) return rc;
  CY_HIP(ctx, hipGetLastError());
  rc = ingest_finish(ctx, n_ins, n_del, batch_stats);
  if (rc != 0) {
    // The batch was refused or needs a larger pool: the runs did nothing but
    // flip the host's buffer parity.
    for (size_t k = 0; k < ps.size(); ++k) {
      ps[k]->cur_act = parity[k];
      ps[k]->stream = nullptr;
    }
    if (rc < 0) return rc;
    cytom_batch_stats bs{};
    if ((rc = ingest_batch(ctx, d_src, d_dst, d_w, n_ins, d_del_src, d_del_dst, d_del_w, n_del,
                           &bs)))
      return rc;
    syncs += bs.host_syncs;
    for (size_t k = 0; k < ps.size(); ++k) {
      if ((rc = run_program(ctx, *ps[k], run_stats ? &run_stats[k] : nullptr))) return rc;
      syncs += ps[k]->groups;  // one synchronisation per group
    }
    bs.host_syncs = syncs;
    if (batch_stats) *batch_stats = bs;
    return 0;
  }
  std::vector<char> more(ps.size());
  bool any = false;
  for (size_t k = 0; k < ps.size(); ++k) any |= (more[k] = run_needs_more(ctx, *ps[k])) != 0;
  while (any) {
    for (size_t k = 0; k < ps.size(); ++k)
      if (more[k]) run_group(ctx, *ps[k]);
    ++syncs;
    if ((rc = sync_all())) return rc;
    any = false;
    for (size_t k = 0; k < ps.size(); ++k)
      if (more[k]) any |= (more[k] = run_needs_more(ctx, *ps[k])) != 0;
  }
  for (size_t k = 0; k < ps.size(); ++k)
    if ((rc = run_finish(ctx, *ps[k], run_stats ? &run_stats[k] : nullptr))) return rc;
  if (batch_stats) batch_stats->host_syncs = syncs;
  return 0;
}

int64_t cytom_ppr_program_adjoint(cytom_ctx* ctx, int prog_id, const float* d_dL_dr, float tol,
                                  uint32_t max_iterations, float* d_dL_dt, uint32_t* d_src,
                                  uint32_t* d_dst, float* d_dL_dw, uint64_t cap,
                                  uint32_t* iterations) {
  CY_CHECK_CTX(ctx);
  if (prog_id < 0 || prog_id >= (int)ctx->progs.size())
    return set_err(ctx, CYTOM_EINVAL, "bad program id");
  return ppr_program_adjoint(ctx, ctx->progs[prog_id], d_dL_dr, tol, max_iterations, d_dL_dt, d_src,
                             d_dst, d_dL_dw, cap, iterations);
}

int cytom_batched_execute(cytom_ctx* ctx, const cytom_query* q, uint32_t nq,
                          cytom_run_stats* stats) {
  CY_CHECK_CTX(ctx);
  if (nq && !q) return set_err(ctx, CYTOM_EINVAL, "null query array");
  return batched_execute(ctx, q, nq, stats);
}

int cytom_read_query(cytom_ctx* ctx, uint32_t query, void* out, size_t bytes) {
  CY_CHECK_CTX(ctx);
  return read_query(ctx, query, out, bytes);
}

int cytom_read_values(cytom_ctx* ctx, int prog_id, void* out, size_t bytes) {
  CY_CHECK_CTX(ctx);
  if (prog_id < 0 || prog_id >= (int)ctx->progs.size())
    return set_err(ctx, CYTOM_EINVAL, "bad program id");
  if (!out || bytes < ctx->V * 4) return set_err(ctx, CYTOM_EINVAL, "output buffer too small");
  const Program& P = ctx->progs[prog_id];
  const void* src = prd_like(P.kind) ? (const void*)P.rank : (const void*)P.cur;
  CY_HIP(ctx, hipStreamSynchronize(ctx->stream));
  CY_HIP(ctx, hipMemcpy(out, src, ctx->V * 4, hipMemcpyDeviceToHost));
  return 0;
}

int64_t cytom_read_results(cytom_ctx* ctx, int prog_id, void* out, size_t bytes) {
  CY_CHECK_CTX(ctx);
  if (prog_id < 0 || prog_id >= (int)ctx->progs.size())
    return set_err(ctx, CYTOM_EINVAL, "bad program id");
  return read_results(ctx, ctx->progs[prog_id], out, bytes);
}

int cytom_read_degrees(cytom_ctx* ctx, uint32_t* out_degree, uint32_t* in_degree) {
  CY_CHECK_CTX(ctx);
  CY_HIP(ctx, hipStreamSynchronize(ctx->stream));
  if (out_degree) CY_HIP(ctx, hipMemcpy(out_degree, ctx->outdeg, ctx->V * 4, hipMemcpyDeviceToHost));
  if (in_degree) CY_HIP(ctx, hipMemcpy(in_degree, ctx->indeg, ctx->V * 4, hipMemcpyDeviceToHost));
  return 0;
}

int64_t cytom_export_edges(cytom_ctx* ctx, uint32_t* src, uint32_t* dst, uint32_t* w,
                           uint64_t cap) {
  CY_CHECK_CTX(ctx);
  if (src && !dst) return set_err(ctx, CYTOM_EINVAL, "dst array missing");
  return export_edges(ctx, src, dst, w, cap);
}

int cytom_set_tuning(cytom_ctx* ctx, const cytom_tuning* t) {
  CY_CHECK_CTX(ctx);
  if (!t) return set_err(ctx, CYTOM_EINVAL, "null tuning");
  if (t->chunk_slots && (t->chunk_slots % 64 || t->chunk_slots > 1024))
    return set_err(ctx, CYTOM_EINVAL, "chunk_slots must be a multiple of 64 up to 1024");
  if (t->job_chunks > (1u << 20)) return set_err(ctx, CYTOM_EINVAL, "job_chunks above 2^20");
  if (t->blocks_per_cu > 16) return set_err(ctx, CYTOM_EINVAL, "blocks_per_cu above 16");
  if (t->source_bands && (t->source_bands & (t->source_bands - 1)))
    return set_err(ctx, CYTOM_EINVAL, "source_bands must be a power of two");
  if (t->source_bands > MAX_SOURCE_BANDS)
    return set_err(ctx, CYTOM_EINVAL, "source_bands above 16");
  if (t->prd_frac_bits && (t->prd_frac_bits < 16 || (int)t->prd_frac_bits > max_prd_shift(ctx->V)))
    return set_err(ctx, CYTOM_EINVAL, "prd_frac_bits outside [16, 60 - ceil(log2 2V)]");
  if (t->prd_frac_bits) ctx->prd_shift = (int)t->prd_frac_bits;
  if (t->chunk_slots) ctx->chunk_slots = t->chunk_slots;
  if (t->job_chunks) ctx->job_chunks = t->job_chunks;
  if (t->blocks_per_cu) ctx->gather_blocks = (uint32_t)ctx->num_cus * t->blocks_per_cu;
  if (t->source_bands) ctx->source_bands = t->source_bands;
  if (t->resort_pct) ctx->resort_pct = t->resort_pct;
  if (t->resort_max_slots) ctx->resort_max_slots = t->resort_max_slots;
  ctx->jobs_dirty = true;
  return 0;
}

int cytom_get_tuning(const cytom_ctx* ctx, cytom_tuning* t) {
  if (!ctx || !t) return CYTOM_EINVAL;
  t->chunk_slots = ctx->chunk_slots;
  t->job_chunks = ctx->job_chunks;
  t->blocks_per_cu = ctx->gather_blocks / (uint32_t)ctx->num_cus;
  t->source_bands = std::min(ctx->source_bands, ctx->T);
  t->resort_pct = ctx->resort_pct;
  t->resort_max_slots = ctx->resort_max_slots;
  t->prd_frac_bits = (uint32_t)ctx->prd_shift;
  return 0;
}

int cytom_set_profiling(cytom_ctx* ctx, int enable) {
  CY_CHECK_CTX(ctx);
  if (enable < 0) return set_err(ctx, CYTOM_EINVAL, "profiling takes 0 or a sampling period >= 1");
  ctx->profile = enable != 0;
  ctx->profile_every = enable > 0 ? (uint32_t)enable : 1u;
  for (auto& P : ctx->progs) {
    int rc = program_events(ctx, P);
    if (rc) return rc;
  }
  for (int k = 0; k < 5; ++k) {
    ctx->gather_ms[k] = 0;
    ctx->gather_launches[k] = 0;
    ctx->gather_alg_bytes[k] = 0;
  }
  return 0;
}

int cytom_set_step_graphs(cytom_ctx* ctx, int enable) {
  CY_CHECK_CTX(ctx);
  ctx->step_graphs = enable != 0;
  if (!ctx->step_graphs) {
    CY_HIP(ctx, hipStreamSynchronize(ctx->stream));
    free_graphs(ctx);
  }
  return 0;
}

// The runs of several programs on the current graph (no update batch): as
// cytom_execute of each in turn, but run concurrently, one stream each, with
// one host synchronisation per iteration group of all of them.
int cytom_execute_many(cytom_ctx* ctx, const int* prog_ids, uint32_t n_progs,
                       cytom_run_stats* run_stats) {
  CY_CHECK_CTX(ctx);
  std::vector<Program*> ps;
  int rc;
  if ((rc = program_list(ctx, prog_ids, n_progs, ps))) return rc;
  if (ctx->jobs_dirty && (rc = build_jobs(ctx))) return rc;
  for (Program* P : ps)
    if ((rc = program_events(ctx, *P))) return rc;
  const size_t ns = ctx->concurrent_runs && ps.size() > 1 ? std::min(ps.size(), MAX_RUN_STREAMS)
                                                          : 0;
  if (ns && (rc = run_streams(ctx, ns))) return rc;
  struct StreamReset {
    std::vector<Program*>& ps;
    ~StreamReset() {
      for (Program* P : ps) P->stream = nullptr;
    }
  } reset{ps};
  if (ns) {
    CY_HIP(ctx, hipEventRecord(ctx->fork_ev, ctx->stream));
    for (size_t k = 0; k < ps.size(); ++k) {
      ps[k]->stream = ctx->run_streams[k % ns];
      CY_HIP(ctx, hipStreamWaitEvent(ps[k]->stream, ctx->fork_ev, 0));
    }
  }
  std::vector<char> more(ps.size(), 1);
  for (Program* P : ps) run_open(*P);
  for (bool any = !ps.empty(); any;) {
    for (size_t k = 0; k < ps.size(); ++k)
      if (more[k]) run_group(ctx, *ps[k]);
    CY_HIP(ctx, hipStreamSynchronize(ctx->stream));
    for (size_t i = 0; i < ns; ++i) CY_HIP(ctx, hipStreamSynchronize(ctx->run_streams[i]));
    any = false;
    for (size_t k = 0; k < ps.size(); ++k)
      if (more[k]) any |= (more[k] = run_needs_more(ctx, *ps[k])) != 0;
  }
  for (size_t k = 0; k < ps.size(); ++k)
    if ((rc = run_finish(ctx, *ps[k], run_stats ? &run_stats[k] : nullptr))) return rc;
  return 0;
}

int cytom_set_concurrent_runs(cytom_ctx* ctx, int enable) {
  CY_CHECK_CTX(ctx);
  ctx->concurrent_runs = enable != 0;
  return 0;
}

int cytom_step_graph_stats(cytom_ctx* ctx, uint64_t* captures, uint64_t* replays,
                           uint32_t* cached) {
  CY_CHECK_CTX(ctx);
  if (captures) *captures = ctx->graph_captures;
  if (replays) *replays = ctx->graph_replays;
  if (cached) *cached = (uint32_t)ctx->graphs.size();
  return 0;
}

int cytom_kernel_profile(cytom_ctx* ctx, uint32_t kind, double* gather_ms, uint64_t* launches,
                         double* alg_bytes) {
  if (!ctx || kind > CYTOM_PPR) return CYTOM_EINVAL;
  if (gather_ms) *gather_ms = ctx->gather_ms[kind];
  if (launches) *launches = ctx->gather_launches[kind];
  if (alg_bytes) *alg_bytes = ctx->gather_alg_bytes[kind];
  return 0;
}

int cytom_rmat_generate_device(uint32_t scale, uint64_t seed, uint64_t first, uint64_t n,
                               uint32_t* d_src, uint32_t* d_dst, uint32_t* d_w) {
  if (scale > 31 || (n && (!d_src || !d_dst))) return CYTOM_EINVAL;
  launch_rmat(scale, seed, first, nullptr, n, d_src, d_dst, d_w, nullptr);
  if (hipStreamSynchronize(nullptr) != hipSuccess) return CYTOM_EDEVICE;
  return hipGetLastError() == hipSuccess ? 0 : CYTOM_EDEVICE;
}

int cytom_rmat_gather_device(uint32_t scale, uint64_t seed, const uint64_t* d_index, uint64_t n,
                             uint32_t* d_src, uint32_t* d_dst, uint32_t* d_w) {
  if (scale > 31 || (n && (!d_src || !d_dst || !d_index))) return CYTOM_EINVAL;
  launch_rmat(scale, seed, 0, d_index, n, d_src, d_dst, d_w, nullptr);
  if (hipStreamSynchronize(nullptr) != hipSuccess) return CYTOM_EDEVICE;
  return hipGetLastError() == hipSuccess ? 0 : CYTOM_EDEVICE;
}

int cytom_stream_create(uint64_t seed, cytom_stream** out) {
  if (!out) return CYTOM_EINVAL;
  auto* s = new (std::nothrow) cytom_stream();
  if (!s) return CYTOM_ENOMEM;
  s->hash_del = hash64(seed ^ DELETE_SALT);
  *out = s;
  return 0;
}

int cytom_stream_destroy(cytom_stream* s) {
  delete s;
  return 0;
}

// Inserts take the next n_ins stream indices; then n_del live indices are
// removed by sequential swap-remove sampling: draw j picks position
// hash(hd ^ (batch << 32 | j)) mod |live|.
namespace {

// Position of a deletion draw among n live indices: SplitMix64's output
// function (Steele, Lea and Flood, OOPSLA 2014) as a full-avalanche mix of
// the draw's key, scaled to [0, n) by the high half of a 64 x 64-bit product.
// (Ligra's hash64 taken modulo n correlated the positions of different
// seeds: the drawn indices were visibly non-uniform.)
uint64_t mix_position(uint64_t z, uint64_t n) {
  z += 0x9e3779b97f4a7c15ull;
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  z ^= z >> 31;
  return (uint64_t)(((unsigned __int128)z * n) >> 64);
}

}  // namespace

int64_t cytom_stream_next(cytom_stream* s, uint64_t n_ins, uint64_t n_del, uint64_t* ins_first,
                          uint64_t* del_index) {
  if (!s || !ins_first || (n_del && !del_index)) return CYTOM_EINVAL;
  *ins_first = s->next_index;
  try {
    // Geometric growth: reserving the exact new size would copy the whole
    // live list on every batch (16 GB per batch late in a 2^31-edge stream).
    const size_t need = s->live.size() + n_ins;
    if (need > s->live.capacity()) s->live.reserve(std::max(need, 2 * s->live.capacity()));
  } catch (...) {
    return CYTOM_ENOMEM;
  }
  for (uint64_t j = 0; j < n_ins; ++j) s->live.push_back(s->next_index + j);
  s->next_index += n_ins;
  const uint64_t m = std::min<uint64_t>(n_del, s->live.size());
  for (uint64_t j = 0; j < m; ++j) {
    const uint64_t n_live = s->live.size();
    const uint64_t pos = mix_position(s->hash_del ^ ((s->batch << 32) | j), n_live);
    del_index[j] = s->live[pos];
    s->live[pos] = s->live.back();
    s->live.pop_back();
  }
  s->batch += 1;
  return (int64_t)m;
}

int cytom_device_alloc(void** ptr, size_t bytes) {
  if (!ptr) return CYTOM_EINVAL;
  return hipMalloc(ptr, std::max<size_t>(bytes, 1)) == hipSuccess ? 0 : CYTOM_ENOMEM;
}
int cytom_device_free(void* ptr) { return hipFree(ptr) == hipSuccess ? 0 : CYTOM_EDEVICE; }
int cytom_memcpy_h2d(void* dst, const void* src, size_t bytes) {
  return hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice) == hipSuccess ? 0 : CYTOM_EDEVICE;
}
int cytom_memcpy_d2h(void* dst, const void* src, size_t bytes) {
  return hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost) == hipSuccess ? 0 : CYTOM_EDEVICE;
}
int cytom_memcpy_d2d(void* dst, const void* src, size_t bytes) {
  return hipMemcpy(dst, src, bytes, hipMemcpyDeviceToDevice) == hipSuccess ? 0 : CYTOM_EDEVICE;
}
int cytom_device_synchronize(void) {
  return hipDeviceSynchronize() == hipSuccess ? 0 : CYTOM_EDEVICE;
}
int cytom_device_count(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) return 0;
  return n;
}

}  // extern "C"
