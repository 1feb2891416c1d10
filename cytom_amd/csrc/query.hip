// Batched queries: many BFS / SSSP / personalized-PageRank queries against
// one snapshot of the live graph, run as a few query groups instead of one
// program run per query.
//
// The reference runs several jobs side by side through its multi-job
// composition (src/include/main/jobs.h:33-140) and a fused two-instance
// PageRank-delta (examples/pagerank-delta-merge.cc:151-252); it has no
// multi-query kernel.  Here a batch works on a CSR snapshot (out-edges by
// source, built once from the tile store) with frontier-driven push kernels
// whose per-vertex state holds a whole query group:
//   * BFS: 64 queries per group as bits of one u64 per vertex (visited,
//     frontier, next), levels kept as u16 per (vertex, query);
//   * SSSP: 16 queries per group, u32 distances stored vertex-major
//     (64 B per vertex), relaxed with atomicMin (any relaxation order
//     reaches the same fixpoint, so the result is exact);
//   * PPR: 16 queries per group, the PageRank-delta iteration of the
//     personalized program with the same FP32 per-vertex arithmetic and
//     exact fixed-point sums as the single-query program, so each query's
//     ranks equal its single-query run.
// Sparse frontiers push along out-edges (CSR); once a frontier's out-edges
// exceed E / PULL_DIV the targets pull along in-edges (CSC) instead, which
// needs no atomics for BFS and SSSP and keeps PPR sums in registers.
// SSSP and PPR groups on graphs above VECTOR_MAX_EDGES run on the tile engine
// instead: SSSP eight queries per run on a 4,096-vertex-tile copy of the
// snapshot (run_tile_sssp), PPR one query per run (run_seq_group).
// The frontier's out-edges are split evenly over the wavefronts (a lane
// finds its first edge by binary search over the frontier's degree prefix
// sum), so hub vertices do not serialise the expansion.
#include <algorithm>
#include <cstring>
#include <vector>

#include "devutil.h"
#include "internal.h"

namespace cytom {

// Device-side control of a BFS group's levels: every kernel of a level reads
// the frontier size and the push/pull choice here, so the host enqueues
// several levels between synchronisations instead of reading counts back
// after each one.
struct BfsCtl {
  uint32_t n;      // frontier size of the next level (0: the group is done)
  uint32_t next;   // next-list length (the expansion's append counter)
  uint32_t mode;   // this level: 0 nothing, 1 push along out-edges, 2 pull along in-edges
  uint32_t level;  // levels run so far
  unsigned long long te, active, iters;  // run statistics
};

struct QueryEngine {
  // CSR snapshot
  uint64_t snap_batch = ~0ull;  // ctx->batches when the snapshot was taken
  // The snapshot's edges in a context of 4,096-vertex query tiles, for the
  // multi-query SSSP and PPR kinds (built on first use per snapshot).
  cytom_ctx* qt = nullptr;
  uint64_t qt_batch = ~0ull;
  uint64_t E = 0;
  uint64_t* row_off = nullptr;  // V + 1
  uint32_t* col = nullptr;      // E
  uint32_t* wt = nullptr;       // E (weighted graphs)
  uint64_t* in_off = nullptr;   // V + 1: in-edges by target (CSC)
  uint32_t* in_src = nullptr;   // E
  uint32_t* in_w = nullptr;     // E (weighted graphs)
  uint32_t* heavy = nullptr;    // targets with more than HUB in-edges
  uint32_t n_heavy = 0;
  // in-edge segments of the heavy targets (at most SEG edges each)
  uint32_t* seg_v = nullptr;
  uint64_t* seg_b = nullptr;  // [2 * n_seg]: begin, end
  uint32_t n_seg = 0;
  // frontier work space
  uint32_t* flist = nullptr;  // V
  uint32_t* nlist = nullptr;  // V
  uint32_t* glist = nullptr;  // V
  uint64_t* fdeg = nullptr;   // V + 1
  uint64_t* fpre = nullptr;   // V + 1
  uint32_t* mark = nullptr;   // V
  unsigned int* counter = nullptr;  // [0] next-list length
  void* scan_tmp = nullptr;
  // BFS groups run concurrently, each on a lane of its own: the group's
  // visited / frontier / next masks, frontier lists, degree prefix, control
  // block (device, and its pinned host copy) and stream (allocated by the
  // first BFS groups).
  struct BfsLane {
    uint64_t *vis = nullptr, *front = nullptr, *nxt = nullptr;
    uint32_t *src = nullptr, *flist = nullptr, *nlist = nullptr;
    uint64_t *fdeg = nullptr, *fpre = nullptr;
    void* scan_tmp = nullptr;
    BfsCtl* c = nullptr;
    BfsCtl* h = nullptr;
    hipStream_t st = nullptr;
  };
  std::vector<BfsLane> blanes;
  hipEvent_t bfork = nullptr;  // the snapshot is built (the lanes start after it)
  // results: per query (kind, group, lane)
  struct Q {
    int kind;
    uint32_t group, lane;
  };
  std::vector<Q> queries;
  std::vector<void*> group_buf;  // BFS: u16 levels [V][64]; SSSP: u32 [V][16]; PPR: f32 rank [V][16]
  std::vector<int> group_kind;
};

namespace {

constexpr int QT = 256;          // threads per block
constexpr uint32_t EPL = 8;      // edges per lane (scalar) or per lane group (vector)
constexpr uint32_t BFS_W = 64;   // queries per BFS group (bits of a u64)
constexpr uint32_t VEC_W = 16;   // queries per SSSP / PPR group (a power of two <= 32)
constexpr uint32_t VEC_MASK = VEC_W - 1;
constexpr uint32_t VEC_SLOTS = 64 / VEC_W;  // edges per wavefront step in the vector kernels
constexpr uint64_t VEC_BITS = (1ull << VEC_W) - 1;
constexpr uint16_t LEVEL_NONE = 0xFFFF;
// In-degree above which a whole block pulls a vertex.
constexpr uint64_t HUB = 256;
// Heavy targets' in-edges are pulled in segments of this many edges, one
// block each, combined with atomics.
constexpr uint64_t SEG = 2048;
// A frontier pulls (dense) once its out-edges exceed E / PULL_DIV.
constexpr uint64_t PULL_DIV = 16;

inline uint32_t blocks_for(uint64_t n, uint32_t per_block, uint32_t cap = 65536) {
  uint64_t b = (n + per_block - 1) / per_block;
  return (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(b, cap));
}

// --------------------------------------------------------------- snapshot
__global__ void k_widen(const uint32_t* __restrict__ in, uint64_t* __restrict__ out, uint64_t n) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x)
    out[i] = in[i];
}

// Snapshot build: the live slots of every work-list chunk are compacted into
// (src, dst[, w]) arrays (one wavefront per chunk, ballot ranks), then a
// stable radix sort of (source, export index) gives the CSR order and one of
// (target, export index) the CSC order.  No per-vertex cursors: hub vertices
// cost no atomic contention.
__global__ void k_chunk_live(const uint64_t* __restrict__ chunks, uint64_t n_chunks, StoreView S,
                             uint32_t CH, uint64_t* __restrict__ live) {
  const uint32_t lane = threadIdx.x & 63;
  for (uint64_t c = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6; c < n_chunks;
       c += ((uint64_t)gridDim.x * blockDim.x) >> 6) {
    const uint64_t ch = chunks[c];
    const uint32_t t = (uint32_t)(ch >> 32), s0 = (uint32_t)ch;
    const uint32_t n = min(CH, S.tile_cnt[t] - s0);
    const uint64_t off = S.tile_off[t] + s0;
    uint32_t k = 0;
    for (uint32_t i = lane; i < n; i += 64) k += S.pool[off + i] != TOMB ? 1u : 0u;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) k += __shfl_xor(k, o, 64);
    if (lane == 0) live[c] = k;
  }
}

__global__ void k_chunk_export(const uint64_t* __restrict__ chunks, uint64_t n_chunks,
                               StoreView S, uint32_t CH, const uint64_t* __restrict__ base,
                               uint32_t* __restrict__ src, uint32_t* __restrict__ dst,
                               uint32_t* __restrict__ w, uint32_t* __restrict__ idx) {
  const uint32_t lane = threadIdx.x & 63;
  for (uint64_t c = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6; c < n_chunks;
       c += ((uint64_t)gridDim.x * blockDim.x) >> 6) {
    const uint64_t ch = chunks[c];
    const uint32_t t = (uint32_t)(ch >> 32), s0 = (uint32_t)ch;
    const uint32_t n = min(CH, S.tile_cnt[t] - s0);
    const uint64_t off = S.tile_off[t] + s0;
    const uint32_t xb = (t % S.T) << S.log2vpt, yb = (t / S.T) << S.log2vpt;
    uint64_t p = base[c];
    for (uint32_t i0 = 0; i0 < n; i0 += 64) {
      const uint32_t i = i0 + lane;
      const uint32_t e = i < n ? S.pool[off + i] : TOMB;
      const bool live = e != TOMB;
      const uint64_t m = __ballot(live);
      if (live) {
        const uint64_t q = p + __popcll(m & ((1ull << lane) - 1ull));
        src[q] = xb + (e & 0xFFFFu);
        dst[q] = yb + (e >> 16);
        if (w) w[q] = S.wpool[off + i];
        idx[q] = (uint32_t)q;
      }
      p += __popcll(m);
    }
  }
}

// out[i] = in[perm[i]] (and the weights alongside).
__global__ void k_permute(const uint32_t* __restrict__ perm, uint64_t n,
                          const uint32_t* __restrict__ in, uint32_t* __restrict__ out,
                          const uint32_t* __restrict__ win, uint32_t* __restrict__ wout) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x) {
    const uint32_t j = perm[i];
    out[i] = in[j];
    if (wout) wout[i] = win[j];
  }
}

__global__ void k_heavy_list(const uint32_t* __restrict__ indeg, uint64_t V,
                             uint32_t* __restrict__ list, unsigned int* __restrict__ counter) {
  for (uint64_t v = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; v < V;
       v += (uint64_t)gridDim.x * blockDim.x)
    if (indeg[v] > HUB) list[atomicAdd(counter, 1u)] = (uint32_t)v;
}

// In-edge segments of the heavy targets: counts, then (after a scan) the
// segments [b, min(b + SEG, end)) of each target's in-edge range.
__global__ void k_seg_count(const uint32_t* __restrict__ heavy, uint32_t n,
                            const uint64_t* __restrict__ in_off, uint64_t* __restrict__ nseg) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const uint32_t v = heavy[i];
    nseg[i] = (in_off[v + 1] - in_off[v] + SEG - 1) / SEG;
  }
}

__global__ void k_seg_emit(const uint32_t* __restrict__ heavy, uint32_t n,
                           const uint64_t* __restrict__ in_off, const uint64_t* __restrict__ sbase,
                           uint32_t* __restrict__ seg_v, uint64_t* __restrict__ seg_b) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const uint32_t v = heavy[i];
    const uint64_t e1 = in_off[v + 1];
    uint64_t k = sbase[i];
    for (uint64_t e = in_off[v]; e < e1; e += SEG, ++k) {
      seg_v[k] = v;
      seg_b[2 * k] = e;
      seg_b[2 * k + 1] = min(e + SEG, e1);
    }
  }
}

// ------------------------------------------------------------- frontier
__global__ void k_frontier_deg(const uint32_t* __restrict__ flist, uint32_t n,
                               const uint32_t* __restrict__ outdeg, uint64_t* __restrict__ fdeg) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    fdeg[i] = outdeg[flist[i]];
}

// Frontier entry holding edge e: fpre[i] <= e < fpre[i + 1].
__device__ inline uint32_t find_entry(const uint64_t* fpre, uint32_t n, uint64_t e) {
  uint32_t lo = 0, hi = n;
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) >> 1;
    if (fpre[mid] <= e) lo = mid;
    else hi = mid;
  }
  return lo;
}

// Frontier out-degrees of a BFS level (the frontier size from the control
// block).
__global__ void k_bfs_deg(const uint32_t* __restrict__ flist, const BfsCtl* __restrict__ c,
                          const uint32_t* __restrict__ outdeg, uint64_t* __restrict__ fdeg) {
  const uint32_t n = c->n;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    fdeg[i] = outdeg[flist[i]];
}

// Level start (one thread): a frontier whose out-edges (fpre[n]) exceed
// E / PULL_DIV pulls along in-edges, a smaller one pushes; statistics; the
// next-list counter restarts.
__global__ void k_bfs_level(BfsCtl* c, const uint64_t* __restrict__ fpre, uint64_t E) {
  const uint32_t n = c->n;
  c->next = 0;
  if (n == 0) {
    c->mode = 0;
    return;
  }
  const uint64_t te = fpre[n];
  c->mode = te * PULL_DIV > E ? 2u : (te ? 1u : 0u);
  c->level += 1;
  c->te += te;
  c->active += n;
  c->iters += 1;
}

// Level end (one thread): the next list becomes the frontier.
__global__ void k_bfs_level_end(BfsCtl* c) { c->n = c->next; }

// Edge layout of the expansion: wavefront w owns edges
// [w*64*EPL, (w+1)*64*EPL) of the frontier's concatenated out-edges, lane l
// takes edges l, l+64, ... of that range (coalesced column reads).
__global__ void __launch_bounds__(QT) k_bfs_expand(
    const uint32_t* __restrict__ flist, const uint64_t* __restrict__ fpre, BfsCtl* c,
    const uint64_t* __restrict__ row_off, const uint32_t* __restrict__ col,
    const uint64_t* __restrict__ front, const uint64_t* __restrict__ vis,
    unsigned long long* __restrict__ nxt, uint32_t* __restrict__ nlist) {
  if (c->mode != 1u) return;  // uniform
  const uint32_t n = c->n;
  const uint64_t TE = fpre[n];
  const uint32_t lane = threadIdx.x & 63;
  const uint64_t wave = ((uint64_t)blockIdx.x * QT + threadIdx.x) >> 6;
  const uint64_t nwaves = ((uint64_t)gridDim.x * QT) >> 6;
  for (uint64_t b = wave * 64 * EPL; b < TE; b += nwaves * 64 * EPL) {
    uint64_t e = b + lane;
    if (e >= TE) continue;
    uint32_t i = find_entry(fpre, n, e);
    const uint64_t end = min(b + 64 * EPL, TE);
    for (; e < end; e += 64) {
      while (fpre[i + 1] <= e) ++i;
      const uint32_t u = flist[i];
      const uint32_t v = col[row_off[u] + (e - fpre[i])];
      const uint64_t m = front[u] & ~vis[v];
      if (m) {
        const unsigned long long old = atomicOr(&nxt[v], (unsigned long long)m);
        if (old == 0ull) nlist[atomicAdd(&c->next, 1u)] = v;
      }
    }
  }
}

// New levels of the vertices reached this round; the reached set becomes
// the next frontier.  front[] of the old frontier was cleared before.
__global__ void k_bfs_apply(const uint32_t* __restrict__ nlist, const BfsCtl* __restrict__ c,
                            uint64_t* __restrict__ nxt, uint64_t* __restrict__ vis,
                            uint64_t* __restrict__ front, uint16_t* __restrict__ lev) {
  const uint32_t n = c->next;
  const uint16_t level = (uint16_t)c->level;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const uint32_t v = nlist[i];
    const uint64_t nb = nxt[v] & ~vis[v];
    nxt[v] = 0;
    vis[v] |= nb;
    front[v] = nb;
    uint64_t m = nb;
    while (m) {
      const int q = __ffsll((unsigned long long)m) - 1;
      m &= m - 1;
      lev[(uint64_t)v * BFS_W + q] = level;
    }
  }
}

__global__ void k_clear_front(const uint32_t* __restrict__ flist, const BfsCtl* __restrict__ c,
                              uint64_t* __restrict__ front) {
  const uint32_t n = c->n;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    front[flist[i]] = 0;
}

__global__ void k_bfs_seed(const uint32_t* __restrict__ src, uint32_t nq, uint64_t* __restrict__ vis,
                           uint64_t* __restrict__ front, uint16_t* __restrict__ lev) {
  for (uint32_t q = blockIdx.x * blockDim.x + threadIdx.x; q < nq; q += gridDim.x * blockDim.x) {
    const uint32_t s = src[q];
    atomicOr(reinterpret_cast<unsigned long long*>(&vis[s]), 1ull << q);
    atomicOr(reinterpret_cast<unsigned long long*>(&front[s]), 1ull << q);
    lev[(uint64_t)s * BFS_W + q] = 0;
  }
}

// ------------------------------------------------------------------ SSSP
// 16 lanes per edge (one per query of the group), 4 edges per wavefront
// step.  sssp.cc:90-99 relaxation with u32 saturating distances.
__global__ void __launch_bounds__(QT) k_sssp_expand(
    const uint32_t* __restrict__ flist, const uint64_t* __restrict__ fpre, uint32_t n,
    const uint64_t* __restrict__ row_off, const uint32_t* __restrict__ col,
    const uint32_t* __restrict__ wt, uint32_t* __restrict__ dist, uint32_t* __restrict__ mark,
    uint32_t* __restrict__ nlist, unsigned int* __restrict__ counter) {
  const uint64_t TE = fpre[n];
  const uint32_t lane = threadIdx.x & 63, q = lane & VEC_MASK, slot = lane / VEC_W;
  const uint64_t wave = ((uint64_t)blockIdx.x * QT + threadIdx.x) >> 6;
  const uint64_t nwaves = ((uint64_t)gridDim.x * QT) >> 6;
  for (uint64_t b = wave * VEC_SLOTS * EPL; b < TE; b += nwaves * VEC_SLOTS * EPL) {
    uint64_t e = b + slot;
    const uint64_t end = min(b + VEC_SLOTS * EPL, TE);
    if (e >= end) continue;
    uint32_t i = find_entry(fpre, n, e);
    for (; e < end; e += VEC_SLOTS) {
      while (fpre[i + 1] <= e) ++i;
      const uint32_t u = flist[i];
      const uint64_t p = row_off[u] + (e - fpre[i]);
      const uint32_t v = col[p];
      const uint32_t du = dist[(uint64_t)u * VEC_W + q];
      bool improved = false;
      if (du != INF) {
        const uint32_t c = sat_add(du, wt[p]);
        uint32_t* dv = &dist[(uint64_t)v * VEC_W + q];
        if (c < *dv) improved = c < atomicMin(dv, c);
      }
      // One lane of the edge's 16 queues the target when any query improved.
      const uint64_t any = __ballot(improved);
      if (q == 0 && ((any >> (slot * VEC_W)) & VEC_BITS)) {
        if (atomicCAS(&mark[v], 0u, 1u) == 0u) nlist[atomicAdd(counter, 1u)] = v;
      }
    }
  }
}

__global__ void k_unmark(const uint32_t* __restrict__ list, uint32_t n, uint32_t* __restrict__ mark) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    mark[list[i]] = 0;
}

__global__ void k_fill_u32(uint32_t* __restrict__ p, uint64_t n, uint32_t v) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x)
    p[i] = v;
}

__global__ void k_sssp_seed(const uint32_t* __restrict__ src, uint32_t nq,
                            uint32_t* __restrict__ dist) {
  for (uint32_t q = blockIdx.x * blockDim.x + threadIdx.x; q < nq; q += gridDim.x * blockDim.x)
    dist[(uint64_t)src[q] * VEC_W + q] = 0;
}

// ------------------------------------------------------------------- PPR
// Frontier values of a group: contrib[v][q] = delta/out_degree while query
// q is active at v, else 0 (the single-query program's contribution array),
// as floats; each edge adds round(contrib * 2^S) to its target's exact
// fixed-point sum (pagerank-delta.cc:127-132).
__device__ inline long long ppr_fixed(float c, float qscale) {
  return c != 0.f ? __float2ll_rn(c * qscale) : 0ll;
}

__global__ void __launch_bounds__(QT) k_ppr_expand(
    const uint32_t* __restrict__ flist, const uint64_t* __restrict__ fpre, uint32_t n,
    const uint64_t* __restrict__ row_off, const uint32_t* __restrict__ col,
    const float* __restrict__ contrib, long long* __restrict__ dsum, uint32_t* __restrict__ mark,
    uint32_t* __restrict__ nlist, unsigned int* __restrict__ counter, float qscale) {
  const uint64_t TE = fpre[n];
  const uint32_t lane = threadIdx.x & 63, q = lane & VEC_MASK, slot = lane / VEC_W;
  const uint64_t wave = ((uint64_t)blockIdx.x * QT + threadIdx.x) >> 6;
  const uint64_t nwaves = ((uint64_t)gridDim.x * QT) >> 6;
  for (uint64_t b = wave * VEC_SLOTS * EPL; b < TE; b += nwaves * VEC_SLOTS * EPL) {
    uint64_t e = b + slot;
    const uint64_t end = min(b + VEC_SLOTS * EPL, TE);
    if (e >= end) continue;
    uint32_t i = find_entry(fpre, n, e);
    for (; e < end; e += VEC_SLOTS) {
      while (fpre[i + 1] <= e) ++i;
      const uint32_t u = flist[i];
      const uint32_t v = col[row_off[u] + (e - fpre[i])];
      const long long c = ppr_fixed(contrib[(uint64_t)u * VEC_W + q], qscale);
      if (c != 0)
        atomicAdd(reinterpret_cast<unsigned long long*>(&dsum[(uint64_t)v * VEC_W + q]),
                  (unsigned long long)c);
      const uint64_t any = __ballot(c != 0);
      if (q == 0 && ((any >> (slot * VEC_W)) & VEC_BITS)) {
        if (atomicCAS(&mark[v], 0u, 1u) == 0u) nlist[atomicAdd(counter, 1u)] = v;
      }
    }
  }
}

// pagerank-delta.cc:142-174 (not the first-iteration rule) for one
// (vertex, query) whose fixed-point sum is s_fx: applied only where a sum
// arrived; writes the next contribution and returns whether the query stays
// active at the vertex.
__device__ inline bool ppr_apply_one(uint64_t k, long long s_fx, uint32_t od, double thr,
                                     double qinv, float* rank, float* contrib_next) {
  if (s_fx == 0) return false;
  const float s = (float)((double)s_fx * qinv);
  const float d = static_cast<float>(0.85 * (double)s);
  const float r = rank[k] + d;
  rank[k] = r;
  const bool a = (double)fabsf(d) > thr * (double)r;
  if (a) contrib_next[k] = od ? d / (float)od : 0.f;
  return a;
}

// Apply over the vertices the push reached; the active ones form the next
// frontier.  contrib_next was cleared for the whole group before.
__global__ void k_ppr_apply(const uint32_t* __restrict__ nlist, uint32_t n, double thr, double qinv,
                            const uint32_t* __restrict__ outdeg, float* __restrict__ rank,
                            float* __restrict__ contrib_next, long long* __restrict__ dsum,
                            uint32_t* __restrict__ mark, uint32_t* __restrict__ flist,
                            unsigned int* __restrict__ counter) {
  const uint32_t q = threadIdx.x & VEC_MASK;
  const uint64_t nthr = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; t < (uint64_t)n * VEC_W;
       t += nthr) {
    const uint32_t v = nlist[t / VEC_W];
    const uint64_t k = (uint64_t)v * VEC_W + q;
    const long long s_fx = dsum[k];
    dsum[k] = 0;
    const bool a = ppr_apply_one(k, s_fx, outdeg[v], thr, qinv, rank, contrib_next);
    // 16 consecutive lanes hold one vertex.
    const uint64_t bits = (__ballot(a) >> (threadIdx.x & (63 & ~VEC_MASK))) & VEC_BITS;
    if (q == 0) {
      mark[v] = 0;
      if (bits) flist[atomicAdd(counter, 1u)] = v;
    }
  }
}

// Pull form for dense frontiers: one 16-lane group per target sums its
// in-edges' contributions in registers (exact, so equal to the push sums),
// eight in-edges per step so their loads overlap, and applies; targets
// above HUB in-edges are left to k_ppr_pull_heavy.
constexpr int PULL_ILP = 8;

__global__ void __launch_bounds__(QT) k_ppr_pull(
    const uint64_t* __restrict__ in_off, const uint32_t* __restrict__ in_src, uint64_t V,
    const uint32_t* __restrict__ outdeg, const float* __restrict__ contrib,
    float* __restrict__ contrib_next, float* __restrict__ rank, double thr, double qinv,
    float qscale, uint32_t* __restrict__ flist, unsigned int* __restrict__ counter) {
  const uint32_t q = threadIdx.x & VEC_MASK;
  const uint64_t ng = ((uint64_t)gridDim.x * blockDim.x) / VEC_W;
  for (uint64_t v = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) / VEC_W; v < V; v += ng) {
    const uint64_t e0 = in_off[v], e1 = in_off[v + 1];
    if (e1 - e0 > HUB) continue;  // uniform per group
    long long acc = 0;
    for (uint64_t e = e0; e < e1; e += PULL_ILP) {
      uint32_t u[PULL_ILP];
#pragma unroll
      for (int k = 0; k < PULL_ILP; ++k) u[k] = in_src[min(e + k, e1 - 1)];
      float c[PULL_ILP];
#pragma unroll
      for (int k = 0; k < PULL_ILP; ++k) c[k] = contrib[(uint64_t)u[k] * VEC_W + q];
#pragma unroll
      for (int k = 0; k < PULL_ILP; ++k)
        if (e + k < e1) acc += ppr_fixed(c[k], qscale);
    }
    const bool a = ppr_apply_one(v * VEC_W + q, acc, outdeg[v], thr, qinv, rank, contrib_next);
    const uint64_t bits = (__ballot(a) >> (threadIdx.x & (63 & ~VEC_MASK))) & VEC_BITS;
    if (q == 0 && bits) flist[atomicAdd(counter, 1u)] = (uint32_t)v;
  }
}

__global__ void k_clear_contrib(const uint32_t* __restrict__ list, uint32_t n,
                                float* __restrict__ contrib) {
  for (uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; t < (uint64_t)n * VEC_W;
       t += (uint64_t)gridDim.x * blockDim.x)
    contrib[(uint64_t)list[t / VEC_W] * VEC_W + (t & VEC_MASK)] = 0.f;
}

// Teleport sets: rank = delta = (1-alpha)/|T| (|T| distinct), active, so
// contrib = delta / out_degree.
__global__ void k_ppr_seed(const uint32_t* __restrict__ members, const uint32_t* __restrict__ qof,
                           const float* __restrict__ r0, uint32_t n,
                           const uint32_t* __restrict__ outdeg, float* __restrict__ rank,
                           float* __restrict__ contrib) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const uint32_t v = members[i], q = qof[i];
    rank[(uint64_t)v * VEC_W + q] = r0[q];
    const uint32_t od = outdeg[v];
    contrib[(uint64_t)v * VEC_W + q] = od ? r0[q] / (float)od : 0.f;
  }
}

// ------------------------------------------------------------ pull forms
// BFS: one lane per target ORs its in-neighbours' frontier masks, eight at
// a time, and stops once every query still missing the target reached it.
__global__ void k_bfs_pull(const uint64_t* __restrict__ in_off, const uint32_t* __restrict__ in_src,
                           uint64_t V, uint64_t qmask, const uint64_t* __restrict__ front,
                           const uint64_t* __restrict__ vis, uint64_t* __restrict__ nxt,
                           uint32_t* __restrict__ nlist, BfsCtl* c) {
  if (c->mode != 2u) return;  // uniform
  for (uint64_t v = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; v < V;
       v += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t e0 = in_off[v], e1 = in_off[v + 1];
    if (e1 - e0 > HUB) continue;
    const uint64_t want = qmask & ~vis[v];
    if (!want) continue;
    uint64_t m = 0;
    for (uint64_t e = e0; e < e1 && (m & want) != want; e += PULL_ILP) {
      uint32_t u[PULL_ILP];
#pragma unroll
      for (int k = 0; k < PULL_ILP; ++k) u[k] = in_src[min(e + k, e1 - 1)];
#pragma unroll
      for (int k = 0; k < PULL_ILP; ++k) m |= front[u[k]];  // clamped repeats add nothing
    }
    m &= want;
    if (m) {
      nxt[v] = m;
      nlist[atomicAdd(&c->next, 1u)] = (uint32_t)v;
    }
  }
}

// SSSP: one 16-lane group per target takes the minimum over its in-edges
// (sssp.cc:90-99), eight at a time, and lowers its distances in place.
__global__ void __launch_bounds__(QT) k_sssp_pull(
    const uint64_t* __restrict__ in_off, const uint32_t* __restrict__ in_src,
    const uint32_t* __restrict__ in_w, uint64_t V, uint32_t* __restrict__ dist,
    uint32_t* __restrict__ nlist, unsigned int* __restrict__ counter) {
  const uint32_t q = threadIdx.x & VEC_MASK;
  const uint64_t ng = ((uint64_t)gridDim.x * blockDim.x) / VEC_W;
  for (uint64_t v = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) / VEC_W; v < V; v += ng) {
    const uint64_t e0 = in_off[v], e1 = in_off[v + 1];
    if (e1 - e0 > HUB) continue;
    uint32_t best = INF;
    for (uint64_t e = e0; e < e1; e += PULL_ILP) {
      uint32_t u[PULL_ILP], w[PULL_ILP];
#pragma unroll
      for (int k = 0; k < PULL_ILP; ++k) {
        const uint64_t x = min(e + k, e1 - 1);
        u[k] = in_src[x];
        w[k] = in_w[x];
      }
      uint32_t du[PULL_ILP];
#pragma unroll
      for (int k = 0; k < PULL_ILP; ++k) du[k] = dist[(uint64_t)u[k] * VEC_W + q];
#pragma unroll
      for (int k = 0; k < PULL_ILP; ++k)
        if (du[k] != INF) best = min(best, sat_add(du[k], w[k]));  // clamped repeats are harmless
    }
    uint32_t* dv = &dist[v * VEC_W + q];
    const bool improved = best < *dv;
    if (improved) *dv = best;
    const uint64_t bits = (__ballot(improved) >> (threadIdx.x & (63 & ~VEC_MASK))) & VEC_BITS;
    if (q == 0 && bits) nlist[atomicAdd(counter, 1u)] = (uint32_t)v;
  }
}

// Heavy targets: one block per in-edge segment (16 edge lanes x 16 queries
// for SSSP / PPR, 256 edge lanes for BFS), reduced in LDS, then combined
// into the target with one atomic per query.
__global__ void __launch_bounds__(QT) k_bfs_pull_seg(
    const uint32_t* __restrict__ seg_v, const uint64_t* __restrict__ seg_b, uint32_t n_seg,
    const uint32_t* __restrict__ in_src, uint64_t qmask, const uint64_t* __restrict__ front,
    const uint64_t* __restrict__ vis, unsigned long long* __restrict__ nxt,
    uint32_t* __restrict__ nlist, BfsCtl* c) {
  __shared__ unsigned long long m_s;
  if (c->mode != 2u) return;  // uniform
  for (uint32_t sg = blockIdx.x; sg < n_seg; sg += gridDim.x) {
    const uint32_t v = seg_v[sg];
    const uint64_t want = qmask & ~vis[v];
    if (!want) continue;  // uniform
    if (threadIdx.x == 0) m_s = 0;
    __syncthreads();
    const uint64_t e0 = seg_b[2 * sg], e1 = seg_b[2 * sg + 1];
    uint64_t m = 0;
    for (uint64_t e = e0 + threadIdx.x; e < e1; e += QT * 4) {
      uint32_t u[4];
#pragma unroll
      for (int k = 0; k < 4; ++k) u[k] = in_src[min(e + k * QT, e1 - 1)];
#pragma unroll
      for (int k = 0; k < 4; ++k) m |= front[u[k]];
    }
    m &= want;
    if (m) atomicOr(&m_s, (unsigned long long)m);
    __syncthreads();
    if (threadIdx.x == 0 && m_s) {
      if (atomicOr(&nxt[v], m_s) == 0ull) nlist[atomicAdd(&c->next, 1u)] = v;
    }
    __syncthreads();
  }
}

__global__ void __launch_bounds__(QT) k_sssp_pull_seg(
    const uint32_t* __restrict__ seg_v, const uint64_t* __restrict__ seg_b, uint32_t n_seg,
    const uint32_t* __restrict__ in_src, const uint32_t* __restrict__ in_w,
    uint32_t* __restrict__ dist, uint32_t* __restrict__ mark, uint32_t* __restrict__ nlist,
    unsigned int* __restrict__ counter) {
  __shared__ uint32_t best_s[VEC_W];
  __shared__ uint32_t any_s;
  const uint32_t q = threadIdx.x & VEC_MASK, el = threadIdx.x / VEC_W;
  constexpr uint32_t LANES = QT / VEC_W;
  for (uint32_t sg = blockIdx.x; sg < n_seg; sg += gridDim.x) {
    const uint32_t v = seg_v[sg];
    if (threadIdx.x < VEC_W) best_s[threadIdx.x] = INF;
    if (threadIdx.x == 0) any_s = 0;
    __syncthreads();
    const uint64_t e0 = seg_b[2 * sg], e1 = seg_b[2 * sg + 1];
    uint32_t best = INF;
    for (uint64_t e = e0 + el; e < e1; e += LANES * 4) {
      uint32_t u[4], w[4];
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        const uint64_t x = min(e + k * LANES, e1 - 1);
        u[k] = in_src[x];
        w[k] = in_w[x];
      }
      uint32_t du[4];
#pragma unroll
      for (int k = 0; k < 4; ++k) du[k] = dist[(uint64_t)u[k] * VEC_W + q];
#pragma unroll
      for (int k = 0; k < 4; ++k)
        if (du[k] != INF) best = min(best, sat_add(du[k], w[k]));
    }
    if (best != INF) atomicMin(&best_s[q], best);
    __syncthreads();
    if (threadIdx.x < VEC_W) {
      const uint32_t b = best_s[threadIdx.x];
      if (b != INF && b < atomicMin(&dist[(uint64_t)v * VEC_W + threadIdx.x], b)) any_s = 1;
    }
    __syncthreads();
    if (threadIdx.x == 0 && any_s && atomicCAS(&mark[v], 0u, 1u) == 0u)
      nlist[atomicAdd(counter, 1u)] = v;
    __syncthreads();
  }
}

__global__ void __launch_bounds__(QT) k_ppr_pull_seg(
    const uint32_t* __restrict__ seg_v, const uint64_t* __restrict__ seg_b, uint32_t n_seg,
    const uint32_t* __restrict__ in_src, const float* __restrict__ contrib,
    long long* __restrict__ dsum, float qscale) {
  __shared__ long long part[QT];
  const uint32_t q = threadIdx.x & VEC_MASK, el = threadIdx.x / VEC_W;
  constexpr uint32_t LANES = QT / VEC_W;
  for (uint32_t sg = blockIdx.x; sg < n_seg; sg += gridDim.x) {
    const uint32_t v = seg_v[sg];
    const uint64_t e0 = seg_b[2 * sg], e1 = seg_b[2 * sg + 1];
    long long acc = 0;
    for (uint64_t e = e0 + el; e < e1; e += LANES * 4) {
      uint32_t u[4];
#pragma unroll
      for (int k = 0; k < 4; ++k) u[k] = in_src[min(e + k * LANES, e1 - 1)];
      float c[4];
#pragma unroll
      for (int k = 0; k < 4; ++k) c[k] = contrib[(uint64_t)u[k] * VEC_W + q];
#pragma unroll
      for (int k = 0; k < 4; ++k)
        if (e + k * LANES < e1) acc += ppr_fixed(c[k], qscale);
    }
    part[threadIdx.x] = acc;
    __syncthreads();
    if (threadIdx.x < VEC_W) {
      long long s_fx = 0;
      for (uint32_t k = 0; k < LANES; ++k) s_fx += part[k * VEC_W + q];
      if (s_fx != 0)
        atomicAdd(reinterpret_cast<unsigned long long*>(&dsum[(uint64_t)v * VEC_W + q]),
                  (unsigned long long)s_fx);
    }
    __syncthreads();
  }
}

// ----------------------------------------------------------- extraction
__global__ void k_extract_lev(const uint16_t* __restrict__ lev, uint32_t q, uint64_t V,
                              uint32_t* __restrict__ out) {
  for (uint64_t v = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; v < V;
       v += (uint64_t)gridDim.x * blockDim.x) {
    const uint16_t l = lev[v * BFS_W + q];
    out[v] = l == LEVEL_NONE ? INF : (uint32_t)l;
  }
}

__global__ void k_extract_vec(const uint32_t* __restrict__ vec, uint32_t q, uint64_t V,
                              uint32_t* __restrict__ out) {
  for (uint64_t v = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; v < V;
       v += (uint64_t)gridDim.x * blockDim.x)
    out[v] = vec[v * VEC_W + q];
}

}  // namespace

// ================================================================= host
namespace {

template <typename T>
int dev_alloc(cytom_ctx* ctx, T** p, size_t bytes) {
  if (hipMalloc((void**)p, std::max<size_t>(bytes, 4)) != hipSuccess) {
    *p = nullptr;
    return set_err(ctx, CYTOM_ENOMEM, "query engine allocation failed");
  }
  return 0;
}

void free_results(QueryEngine* qe) {
  for (void* p : qe->group_buf)
    if (p) (void)hipFree(p);
  qe->group_buf.clear();
  qe->group_kind.clear();
  qe->queries.clear();
}

void free_query_tiles(QueryEngine* qe) {
  if (qe->qt) (void)cytom_destroy(qe->qt);
  qe->qt = nullptr;
  qe->qt_batch = ~0ull;
}

void free_snapshot(QueryEngine* qe) {
  free_query_tiles(qe);
  void* ps[] = {qe->row_off, qe->col,     qe->wt,   qe->in_off,  qe->in_src,  qe->in_w,
                qe->heavy,   qe->flist,   qe->nlist, qe->glist, qe->fdeg,   qe->fpre,    qe->mark,
                qe->counter, qe->scan_tmp, qe->seg_v, qe->seg_b};
  for (void* p : ps)
    if (p) (void)hipFree(p);
  for (auto& b : qe->blanes) {
    for (void* p : {(void*)b.vis, (void*)b.front, (void*)b.nxt, (void*)b.src, (void*)b.flist,
                    (void*)b.nlist, (void*)b.fdeg, (void*)b.fpre, b.scan_tmp, (void*)b.c})
      if (p) (void)hipFree(p);
    if (b.h) (void)hipHostFree(b.h);
    if (b.st) (void)hipStreamDestroy(b.st);
  }
  qe->blanes.clear();
  if (qe->bfork) (void)hipEventDestroy(qe->bfork);
  qe->bfork = nullptr;
  qe->row_off = qe->in_off = nullptr;
  qe->col = qe->wt = qe->in_src = qe->in_w = qe->heavy = qe->flist = qe->nlist = nullptr;
  qe->glist = nullptr;
  qe->n_heavy = 0;
  qe->seg_v = nullptr;
  qe->seg_b = nullptr;
  qe->n_seg = 0;
  qe->fdeg = qe->fpre = nullptr;
  qe->mark = nullptr;
  qe->counter = nullptr;
  qe->scan_tmp = nullptr;
  qe->snap_batch = ~0ull;
}

// CSR (col, wt) and CSC (in_src, in_w) rows of the live edges by export and
// two stable radix sorts; temporaries are freed before returning.
int fill_rows(cytom_ctx* ctx, QueryEngine* qe, uint64_t E) {
  hipStream_t st = ctx->stream;
  int rc;
  if (ctx->jobs_dirty && (rc = build_jobs(ctx))) return rc;
  if (E > 0xFFFFFFFFull) return set_err(ctx, CYTOM_ERANGE, "snapshot above 2^32 edges");
  const int64_t ncr = jobs_chunk_count(ctx);
  if (ncr < 0) return (int)ncr;
  const uint64_t nc = (uint64_t)ncr;
  uint32_t bits = 0;
  while ((1ull << bits) < ctx->V) ++bits;
  uint64_t *live = nullptr, *base = nullptr;
  uint32_t *src = nullptr, *dst = nullptr, *w = nullptr, *idx = nullptr, *ko = nullptr,
           *vo = nullptr, *kt = nullptr, *vt = nullptr;
  void *hist = nullptr, *tmp = nullptr;
  auto done = [&](int r) {
    (void)hipStreamSynchronize(st);
    void* ps[] = {live, base, src, dst, w, idx, ko, vo, kt, vt, hist, tmp};
    for (void* p : ps)
      if (p) (void)hipFree(p);
    return r;
  };
  const uint64_t n = std::max<uint64_t>(E, 1);
  if ((rc = dev_alloc(ctx, &live, (nc + 1) * 8)) || (rc = dev_alloc(ctx, &base, (nc + 1) * 8)) ||
      (rc = dev_alloc(ctx, &src, n * 4)) || (rc = dev_alloc(ctx, &dst, n * 4)) ||
      (ctx->wpool && (rc = dev_alloc(ctx, &w, n * 4))) || (rc = dev_alloc(ctx, &idx, n * 4)) ||
      (rc = dev_alloc(ctx, &ko, n * 4)) || (rc = dev_alloc(ctx, &vo, n * 4)) ||
      (rc = dev_alloc(ctx, &kt, n * 4)) || (rc = dev_alloc(ctx, &vt, n * 4)) ||
      (rc = dev_alloc(ctx, (char**)&hist, radix_hist_bytes(n))) ||
      (rc = dev_alloc(ctx, (char**)&tmp,
                      scan_tmp_bytes(std::max<uint64_t>({n, nc + 1, 256 * ((n + 4095) / 4096) + 1})) +
                          256)))
    return done(rc);
  const auto* chunks = static_cast<const uint64_t*>(ctx->g_chunks.p);
  if (nc) {
    k_chunk_live<<<blocks_for(nc, 4, 16384), 256, 0, st>>>(chunks, nc, ctx->view(),
                                                          ctx->chunk_slots, live);
    scan_exclusive_u64(live, base, nc, tmp, st);
    k_chunk_export<<<blocks_for(nc, 4, 16384), 256, 0, st>>>(chunks, nc, ctx->view(),
                                                            ctx->chunk_slots, base, src, dst, w,
                                                            idx);
  }
  if (E) {
    radix_sort_pairs(src, idx, ko, vo, kt, vt, E, bits, hist, tmp, st);
    k_permute<<<blocks_for(E, 256, 65536), 256, 0, st>>>(vo, E, dst, qe->col, w, qe->wt);
    radix_sort_pairs(dst, idx, ko, vo, kt, vt, E, bits, hist, tmp, st);
    k_permute<<<blocks_for(E, 256, 65536), 256, 0, st>>>(vo, E, src, qe->in_src, w, qe->in_w);
  }
  return done(0);
}

// CSR snapshot of the live graph (rebuilt after every update batch).
int snapshot(cytom_ctx* ctx, QueryEngine* qe) {
  if (qe->snap_batch == ctx->batches && qe->row_off) return 0;
  hipStream_t st = ctx->stream;
  const uint64_t V = ctx->V, E = ctx->live_edges;
  int rc;
  free_snapshot(qe);
  if ((rc = dev_alloc(ctx, &qe->row_off, (V + 1) * 8)) || (rc = dev_alloc(ctx, &qe->col, E * 4)) ||
      (ctx->wpool && (rc = dev_alloc(ctx, &qe->wt, E * 4))) ||
      (rc = dev_alloc(ctx, &qe->in_off, (V + 1) * 8)) || (rc = dev_alloc(ctx, &qe->in_src, E * 4)) ||
      (ctx->wpool && (rc = dev_alloc(ctx, &qe->in_w, E * 4))) ||
      (rc = dev_alloc(ctx, &qe->heavy, V * 4)) ||
      (rc = dev_alloc(ctx, &qe->flist, V * 4)) || (rc = dev_alloc(ctx, &qe->nlist, V * 4)) ||
      (rc = dev_alloc(ctx, &qe->glist, V * 4)) ||
      (rc = dev_alloc(ctx, &qe->fdeg, (V + 1) * 8)) ||
      (rc = dev_alloc(ctx, &qe->fpre, (V + 1) * 8)) || (rc = dev_alloc(ctx, &qe->mark, V * 4)) ||
      (rc = dev_alloc(ctx, &qe->counter, 256)) ||
      (rc = dev_alloc(ctx, (char**)&qe->scan_tmp, scan_tmp_bytes(V + 1) + 256)))
    return rc;
  // row_off / in_off = exclusive scans of the out- / in-degrees.
  k_widen<<<blocks_for(V, 256), 256, 0, st>>>(ctx->outdeg, qe->fdeg, V);
  scan_exclusive_u64(qe->fdeg, qe->row_off, V, qe->scan_tmp, st);
  k_widen<<<blocks_for(V, 256), 256, 0, st>>>(ctx->indeg, qe->fdeg, V);
  scan_exclusive_u64(qe->fdeg, qe->in_off, V, qe->scan_tmp, st);
  if ((rc = fill_rows(ctx, qe, E))) return rc;
  CY_HIP(ctx, hipMemsetAsync(qe->counter, 0, 4, st));
  k_heavy_list<<<blocks_for(V, 256, 16384), 256, 0, st>>>(ctx->indeg, V, qe->heavy, qe->counter);
  uint64_t total = 0, total_in = 0;
  CY_HIP(ctx, hipMemcpyAsync(&total, qe->row_off + V, 8, hipMemcpyDeviceToHost, st));
  CY_HIP(ctx, hipMemcpyAsync(&total_in, qe->in_off + V, 8, hipMemcpyDeviceToHost, st));
  CY_HIP(ctx, hipMemcpyAsync(&qe->n_heavy, qe->counter, 4, hipMemcpyDeviceToHost, st));
  CY_HIP(ctx, hipMemsetAsync(qe->mark, 0, V * 4, st));
  CY_HIP(ctx, hipStreamSynchronize(st));
  if (total != E || total_in != E)
    return set_err(ctx, CYTOM_EDEVICE, "degree sums differ from live edges");
  // Segments of the heavy targets' in-edge ranges, built on the device (the
  // frontier scratch arrays fdeg / fpre are free here).  Their order does not
  // matter: the pull kernels combine segments with OR, min or exact sums.
  const uint32_t nh = qe->n_heavy;
  uint64_t n_seg = 0;
  if (nh) {
    k_seg_count<<<blocks_for(nh, 256), 256, 0, st>>>(qe->heavy, nh, qe->in_off, qe->fdeg);
    scan_exclusive_u64(qe->fdeg, qe->fpre, nh, qe->scan_tmp, st);
    CY_HIP(ctx, hipMemcpyAsync(ctx->h_scalar + 4, qe->fpre + nh, 8, hipMemcpyDeviceToHost, st));
    CY_HIP(ctx, hipStreamSynchronize(st));
    n_seg = ctx->h_scalar[4];
  }
  if (n_seg > 0xFFFFFFFFull) return set_err(ctx, CYTOM_ERANGE, "too many in-edge segments");
  qe->n_seg = (uint32_t)n_seg;
  if ((rc = dev_alloc(ctx, &qe->seg_v, n_seg * 4)) || (rc = dev_alloc(ctx, &qe->seg_b, n_seg * 16)))
    return rc;
  if (n_seg)
    k_seg_emit<<<blocks_for(nh, 256), 256, 0, st>>>(qe->heavy, nh, qe->in_off, qe->fpre, qe->seg_v,
                                                    qe->seg_b);
  qe->E = E;
  qe->snap_batch = ctx->batches;
  return 0;
}

// Degree prefix of the current frontier; returns its out-edge count.
int frontier_prefix(cytom_ctx* ctx, QueryEngine* qe, uint32_t n, uint64_t* te) {
  hipStream_t st = ctx->stream;
  k_frontier_deg<<<blocks_for(n, 256), 256, 0, st>>>(qe->flist, n, ctx->outdeg, qe->fdeg);
  scan_exclusive_u64(qe->fdeg, qe->fpre, n, qe->scan_tmp, st);
  CY_HIP(ctx, hipMemcpyAsync(ctx->h_scalar + 2, qe->fpre + n, 8, hipMemcpyDeviceToHost, st));
  CY_HIP(ctx, hipStreamSynchronize(st));
  *te = ctx->h_scalar[2];
  return 0;
}

uint32_t expand_blocks(cytom_ctx* ctx, uint64_t te, uint32_t edges_per_wave) {
  const uint64_t waves = (te + edges_per_wave - 1) / edges_per_wave;
  const uint64_t cap = (uint64_t)ctx->num_cus * 16;
  return (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>((waves + 3) / 4, cap));
}

int read_counter(cytom_ctx* ctx, QueryEngine* qe, uint32_t* n) {
  ctx->h_scalar[3] = 0;
  CY_HIP(ctx, hipMemcpyAsync(ctx->h_scalar + 3, qe->counter, 4, hipMemcpyDeviceToHost,
                             ctx->stream));
  CY_HIP(ctx, hipStreamSynchronize(ctx->stream));
  *n = (uint32_t)ctx->h_scalar[3];
  return 0;
}

// Unique sources of a group as the first frontier.
int seed_frontier(cytom_ctx* ctx, QueryEngine* qe, std::vector<uint32_t> v, uint32_t* n) {
  std::sort(v.begin(), v.end());
  v.erase(std::unique(v.begin(), v.end()), v.end());
  CY_HIP(ctx, hipMemcpyAsync(qe->flist, v.data(), v.size() * 4, hipMemcpyHostToDevice, ctx->stream));
  CY_HIP(ctx, hipStreamSynchronize(ctx->stream));
  *n = (uint32_t)v.size();
  return 0;
}

// Levels are enqueued in rounds without host round trips: each kernel reads
// the frontier size and the push/pull choice from the control block, and a
// level whose frontier is empty does nothing.  The host reads the control
// block after each round (8 levels, then 4) to see whether the group is done.
constexpr size_t BFS_LANES = 4;

int bfs_lanes(cytom_ctx* ctx, QueryEngine* qe, size_t n) {
  const uint64_t V = ctx->V;
  int rc;
  while (qe->blanes.size() < n) {
    qe->blanes.emplace_back();  // listed first: free_snapshot frees a partial lane
    QueryEngine::BfsLane& b = qe->blanes.back();
    if ((rc = dev_alloc(ctx, &b.vis, V * 8)) || (rc = dev_alloc(ctx, &b.front, V * 8)) ||
        (rc = dev_alloc(ctx, &b.nxt, V * 8)) || (rc = dev_alloc(ctx, &b.src, BFS_W * 4)) ||
        (rc = dev_alloc(ctx, &b.flist, V * 4)) || (rc = dev_alloc(ctx, &b.nlist, V * 4)) ||
        (rc = dev_alloc(ctx, &b.fdeg, (V + 1) * 8)) || (rc = dev_alloc(ctx, &b.fpre, (V + 1) * 8)) ||
        (rc = dev_alloc(ctx, (char**)&b.scan_tmp, scan_tmp_bytes(V + 1) + 256)) ||
        (rc = dev_alloc(ctx, &b.c, sizeof(BfsCtl))))
      return rc;
    if (hipHostMalloc((void**)&b.h, sizeof(BfsCtl), hipHostMallocDefault) != hipSuccess ||
        hipStreamCreate(&b.st) != hipSuccess)
      return set_err(ctx, CYTOM_ENOMEM, "BFS lane allocation failed");
  }
  if (!qe->bfork && hipEventCreateWithFlags(&qe->bfork, hipEventDisableTiming) != hipSuccess)
    return set_err(ctx, CYTOM_EDEVICE, "BFS lane event creation failed");
  return 0;
}

// One BFS group (up to 64 queries as the bits of one u64 per vertex) in
// flight on a lane.
struct BfsRun {
  QueryEngine::BfsLane* L = nullptr;
  uint16_t* lev = nullptr;
  uint64_t qmask = 0;
  uint32_t enqueued = 0;  // levels enqueued
  std::vector<uint32_t> src, seeds;
};

int bfs_start(cytom_ctx* ctx, BfsRun& r) {
  QueryEngine::BfsLane& L = *r.L;
  const hipStream_t st = L.st;
  const uint64_t V = ctx->V;
  r.seeds = r.src;
  std::sort(r.seeds.begin(), r.seeds.end());
  r.seeds.erase(std::unique(r.seeds.begin(), r.seeds.end()), r.seeds.end());
  L.h[0] = BfsCtl{(uint32_t)r.seeds.size(), 0, 0, 0, 0, 0, 0};
  if (hipMemsetAsync(L.vis, 0, V * 8, st) || hipMemsetAsync(L.front, 0, V * 8, st) ||
      hipMemsetAsync(L.nxt, 0, V * 8, st) || hipMemsetAsync(r.lev, 0xFF, V * BFS_W * 2, st) ||
      hipMemcpyAsync(L.src, r.src.data(), r.src.size() * 4, hipMemcpyHostToDevice, st) ||
      hipMemcpyAsync(L.flist, r.seeds.data(), r.seeds.size() * 4, hipMemcpyHostToDevice, st) ||
      hipMemcpyAsync(L.c, L.h, sizeof(BfsCtl), hipMemcpyHostToDevice, st))
    return set_err(ctx, CYTOM_EDEVICE, "BFS group set-up failed");
  k_bfs_seed<<<1, 64, 0, st>>>(L.src, (uint32_t)r.src.size(), L.vis, L.front, r.lev);
  r.qmask = r.src.size() >= 64 ? ~0ull : ((1ull << r.src.size()) - 1);
  r.enqueued = 0;
  return 0;
}

// Enqueues the next round of levels (8 first, then 4) and the copy of the
// control block the host reads after the round.  Levels past the last one
// find an empty frontier and do nothing.
int bfs_round(cytom_ctx* ctx, QueryEngine* qe, BfsRun& r) {
  QueryEngine::BfsLane& L = *r.L;
  const hipStream_t st = L.st;
  const uint64_t V = ctx->V;
  const uint32_t round = r.enqueued ? 4 : 8;
  if (r.enqueued + round + 1 >= LEVEL_NONE)
    return set_err(ctx, CYTOM_ERANGE, "BFS depth above 65534");
  const uint32_t fixed = (uint32_t)ctx->num_cus * 4;  // grid of the per-frontier kernels
  uint32_t* lists[2] = {L.flist, L.nlist};
  BfsCtl* c = L.c;
  for (uint32_t k = 0; k < round; ++k, ++r.enqueued) {
    uint32_t* fl = lists[r.enqueued & 1];
    uint32_t* nl = lists[(r.enqueued + 1) & 1];
    k_bfs_deg<<<fixed, 256, 0, st>>>(fl, c, ctx->outdeg, L.fdeg);
    scan_exclusive_u64_devn(L.fdeg, L.fpre, &c->n, V, L.scan_tmp, st);
    k_bfs_level<<<1, 1, 0, st>>>(c, L.fpre, qe->E);
    k_bfs_pull<<<blocks_for(V, 256, 16384), 256, 0, st>>>(qe->in_off, qe->in_src, V, r.qmask,
                                                           L.front, L.vis, L.nxt, nl, c);
    if (qe->n_seg)
      k_bfs_pull_seg<<<std::min<uint32_t>(qe->n_seg, 16384), QT, 0, st>>>(
          qe->seg_v, qe->seg_b, qe->n_seg, qe->in_src, r.qmask, L.front, L.vis,
          reinterpret_cast<unsigned long long*>(L.nxt), nl, c);
    k_bfs_expand<<<(uint32_t)ctx->num_cus * 16, QT, 0, st>>>(
        fl, L.fpre, c, qe->row_off, qe->col, L.front, L.vis,
        reinterpret_cast<unsigned long long*>(L.nxt), nl);
    k_clear_front<<<fixed, 256, 0, st>>>(fl, c, L.front);
    k_bfs_apply<<<fixed, 256, 0, st>>>(nl, c, L.nxt, L.vis, L.front, r.lev);
    k_bfs_level_end<<<1, 1, 0, st>>>(c);
  }
  CY_HIP(ctx, hipMemcpyAsync(L.h, c, sizeof(BfsCtl), hipMemcpyDeviceToHost, st));
  return 0;
}

// BFS groups, up to BFS_LANES at a time, each on its own stream: every
// round enqueues the next levels of every group in flight, then waits for
// all of them and reads their control blocks; a lane whose group is done
// starts the next group.
int run_bfs_groups(cytom_ctx* ctx, QueryEngine* qe, std::vector<std::vector<uint32_t>>& srcs,
                   const std::vector<uint16_t*>& levs, cytom_run_stats* acc) {
  const size_t n = srcs.size();
  if (!n) return 0;
  const size_t nl = std::min(n, BFS_LANES);
  int rc;
  if ((rc = bfs_lanes(ctx, qe, nl))) return rc;
  CY_HIP(ctx, hipEventRecord(qe->bfork, ctx->stream));
  for (size_t l = 0; l < nl; ++l) CY_HIP(ctx, hipStreamWaitEvent(qe->blanes[l].st, qe->bfork, 0));
  std::vector<BfsRun> runs(nl);
  std::vector<char> busy(nl, 0);
  auto sync_lanes = [&]() -> int {
    for (size_t l = 0; l < nl; ++l) CY_HIP(ctx, hipStreamSynchronize(qe->blanes[l].st));
    return 0;
  };
  auto fail = [&](int r) {
    (void)sync_lanes();
    return r;
  };
  size_t next = 0;
  auto start = [&](size_t l) -> int {
    runs[l].L = &qe->blanes[l];
    runs[l].lev = levs[next];
    runs[l].src = std::move(srcs[next]);
    ++next;
    busy[l] = 1;
    return bfs_start(ctx, runs[l]);
  };
  for (size_t l = 0; l < nl; ++l)
    if ((rc = start(l))) return fail(rc);
  for (bool any = true; any;) {
    for (size_t l = 0; l < nl; ++l)
      if (busy[l] && (rc = bfs_round(ctx, qe, runs[l]))) return fail(rc);
    if ((rc = sync_lanes())) return rc;
    any = false;
    for (size_t l = 0; l < nl; ++l) {
      if (!busy[l]) continue;
      const BfsCtl h = *runs[l].L->h;
      if (h.n == 0) {
        acc->traversed_edges += h.te;
        acc->active_vertices += h.active;
        acc->iterations += h.iters;
        busy[l] = 0;
        if (next < n && (rc = start(l))) return fail(rc);
      }
      any |= busy[l] != 0;
    }
  }
  return 0;
}

int run_sssp_group(cytom_ctx* ctx, QueryEngine* qe, const std::vector<uint32_t>& src,
                   uint32_t* dist, cytom_run_stats* acc) {
  hipStream_t st = ctx->stream;
  const uint64_t V = ctx->V;
  int rc;
  uint32_t* dsrc = nullptr;
  if ((rc = dev_alloc(ctx, &dsrc, src.size() * 4))) return rc;
  auto done = [&](int r) {
    (void)hipStreamSynchronize(st);
    (void)hipFree(dsrc);
    return r;
  };
  k_fill_u32<<<blocks_for(V * VEC_W, 256, 16384), 256, 0, st>>>(dist, V * VEC_W, INF);
  if (hipMemcpyAsync(dsrc, src.data(), src.size() * 4, hipMemcpyHostToDevice, st))
    return done(set_err(ctx, CYTOM_EDEVICE, "SSSP group set-up failed"));
  k_sssp_seed<<<1, 64, 0, st>>>(dsrc, (uint32_t)src.size(), dist);
  uint32_t n = 0;
  if ((rc = seed_frontier(ctx, qe, src, &n))) return done(rc);
  while (n > 0) {
    uint64_t te = 0;
    if ((rc = frontier_prefix(ctx, qe, n, &te))) return done(rc);
    acc->traversed_edges += te;
    acc->active_vertices += n;
    CY_HIP(ctx, hipMemsetAsync(qe->counter, 0, 4, st));
    const bool pull = te * PULL_DIV > qe->E;
    if (pull) {
      k_sssp_pull<<<blocks_for(V * VEC_W, 256, 16384), 256, 0, st>>>(
          qe->in_off, qe->in_src, qe->in_w, V, dist, qe->nlist, qe->counter);
      if (qe->n_seg)
        k_sssp_pull_seg<<<std::min<uint32_t>(qe->n_seg, 16384), QT, 0, st>>>(
            qe->seg_v, qe->seg_b, qe->n_seg, qe->in_src, qe->in_w, dist, qe->mark, qe->nlist,
            qe->counter);
    } else if (te) {
      k_sssp_expand<<<expand_blocks(ctx, te, VEC_SLOTS * EPL), QT, 0, st>>>(
          qe->flist, qe->fpre, n, qe->row_off, qe->col, qe->wt, dist, qe->mark, qe->nlist,
          qe->counter);
    }
    uint32_t m = 0;
    if ((rc = read_counter(ctx, qe, &m))) return done(rc);
    if (m) k_unmark<<<blocks_for(m, 256), 256, 0, st>>>(qe->nlist, m, qe->mark);
    std::swap(qe->flist, qe->nlist);
    n = m;
    acc->iterations += 1;
  }
  return done(0);
}

int run_ppr_group(cytom_ctx* ctx, QueryEngine* qe, const std::vector<std::vector<uint32_t>>& sets,
                  float* rank, cytom_run_stats* acc) {
  hipStream_t st = ctx->stream;
  const uint64_t V = ctx->V;
  int rc;
  float* contrib[2] = {nullptr, nullptr};
  long long* dsum = nullptr;
  uint32_t *members = nullptr, *qof = nullptr;
  float* r0 = nullptr;
  std::vector<uint32_t> hm, hq, all;
  std::vector<float> hr0(VEC_W, 0.f);
  for (uint32_t q = 0; q < sets.size(); ++q) {
    std::vector<uint32_t> s = sets[q];
    std::sort(s.begin(), s.end());
    s.erase(std::unique(s.begin(), s.end()), s.end());
    hr0[q] = (float)((1. - 0.85) * (1.0 / (double)s.size()));
    for (uint32_t v : s) {
      hm.push_back(v);
      hq.push_back(q);
      all.push_back(v);
    }
  }
  auto release = [&]() {
    (void)hipFree(contrib[0]), (void)hipFree(contrib[1]), (void)hipFree(dsum);
    (void)hipFree(members), (void)hipFree(qof), (void)hipFree(r0);
  };
  const size_t vec_bytes = V * VEC_W * 4;
  if ((rc = dev_alloc(ctx, &contrib[0], vec_bytes)) || (rc = dev_alloc(ctx, &contrib[1], vec_bytes)) ||
      (rc = dev_alloc(ctx, &dsum, V * VEC_W * 8)) ||
      (rc = dev_alloc(ctx, &members, hm.size() * 4)) || (rc = dev_alloc(ctx, &qof, hq.size() * 4)) ||
      (rc = dev_alloc(ctx, &r0, VEC_W * 4))) {
    release();
    return rc;
  }
  auto done = [&](int r) {
    (void)hipStreamSynchronize(st);
    release();
    return r;
  };
  if (hipMemsetAsync(rank, 0, vec_bytes, st) || hipMemsetAsync(contrib[0], 0, vec_bytes, st) ||
      hipMemsetAsync(contrib[1], 0, vec_bytes, st) || hipMemsetAsync(dsum, 0, V * VEC_W * 8, st) ||
      hipMemcpyAsync(members, hm.data(), hm.size() * 4, hipMemcpyHostToDevice, st) ||
      hipMemcpyAsync(qof, hq.data(), hq.size() * 4, hipMemcpyHostToDevice, st) ||
      hipMemcpyAsync(r0, hr0.data(), VEC_W * 4, hipMemcpyHostToDevice, st))
    return done(set_err(ctx, CYTOM_EDEVICE, "PPR group set-up failed"));
  k_ppr_seed<<<blocks_for(hm.size(), 256), 256, 0, st>>>(members, qof, r0, (uint32_t)hm.size(),
                                                         ctx->outdeg, rank, contrib[0]);
  uint32_t n = 0;
  if ((rc = seed_frontier(ctx, qe, all, &n))) return done(rc);
  const float qscale = ldexpf(1.f, ctx->prd_shift);
  const double qinv = ldexp(1.0, -ctx->prd_shift);
  uint32_t it = 0;
  int c = 0;  // contrib[c]: the frontier's contributions; contrib[c ^ 1] is all zero
  while (n > 0 && it < ctx->max_iter) {
    uint64_t te = 0;
    if ((rc = frontier_prefix(ctx, qe, n, &te))) return done(rc);
    acc->traversed_edges += te;
    acc->active_vertices += n;
    CY_HIP(ctx, hipMemsetAsync(qe->counter, 0, 4, st));
    uint32_t next = 0;
    if (te * PULL_DIV > qe->E) {
      k_ppr_pull<<<blocks_for(V * VEC_W, 256, 16384), 256, 0, st>>>(
          qe->in_off, qe->in_src, V, ctx->outdeg, contrib[c], contrib[c ^ 1], rank, ctx->delta,
          qinv, qscale, qe->glist, qe->counter);
      if (qe->n_seg) {
        k_ppr_pull_seg<<<std::min<uint32_t>(qe->n_seg, 16384), QT, 0, st>>>(
            qe->seg_v, qe->seg_b, qe->n_seg, qe->in_src, contrib[c], dsum, qscale);
        k_ppr_apply<<<blocks_for((uint64_t)qe->n_heavy * VEC_W, 256), 256, 0, st>>>(
            qe->heavy, qe->n_heavy, ctx->delta, qinv, ctx->outdeg, rank, contrib[c ^ 1], dsum,
            qe->mark, qe->glist, qe->counter);
      }
      if ((rc = read_counter(ctx, qe, &next))) return done(rc);
    } else {
      if (te)
        k_ppr_expand<<<expand_blocks(ctx, te, VEC_SLOTS * EPL), QT, 0, st>>>(
            qe->flist, qe->fpre, n, qe->row_off, qe->col, contrib[c], dsum, qe->mark, qe->nlist,
            qe->counter, qscale);
      uint32_t m = 0;
      if ((rc = read_counter(ctx, qe, &m))) return done(rc);
      CY_HIP(ctx, hipMemsetAsync(qe->counter, 0, 4, st));
      if (m)
        k_ppr_apply<<<blocks_for((uint64_t)m * VEC_W, 256), 256, 0, st>>>(
            qe->nlist, m, ctx->delta, qinv, ctx->outdeg, rank, contrib[c ^ 1], dsum, qe->mark,
            qe->glist, qe->counter);
      if ((rc = read_counter(ctx, qe, &next))) return done(rc);
    }
    // The frontier's contributions are spent: zero them for reuse.
    k_clear_contrib<<<blocks_for((uint64_t)n * VEC_W, 256), 256, 0, st>>>(qe->flist, n, contrib[c]);
    std::swap(qe->flist, qe->glist);
    n = next;
    c ^= 1;
    ++it;
    acc->iterations += 1;
  }
  if (n > 0) acc->converged = 0;  // stopped at the iteration cap
  return done(0);
}

// Sources of the CSR rows: one wavefront per row writes its row id over the
// row's entries.
__global__ void k_row_sources(const uint64_t* __restrict__ row_off, uint64_t V,
                              uint32_t* __restrict__ src) {
  const uint32_t lane = threadIdx.x & 63;
  for (uint64_t u = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6; u < V;
       u += ((uint64_t)gridDim.x * blockDim.x) >> 6)
    for (uint64_t e = row_off[u] + lane; e < row_off[u + 1]; e += 64) src[e] = (uint32_t)u;
}

// Column q of a vertex-major [V][W] array into a V-entry result row.
template <uint32_t W, typename T>
__global__ void k_split_vec(const T* __restrict__ x, uint64_t V, uint32_t m,
                            T* __restrict__ out) {
  for (uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; t < V * m;
       t += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t q = t / V, v = t - q * V;
    out[q * V + v] = x[v * W + q];
  }
}

// The query-tile context of the current snapshot: the snapshot's edges
// ingested as one batch into 4,096-vertex tiles (the eight-query kind's LDS
// window holds 4,096 targets x 8 distances), with the main context's
// iteration cap so every query runs as its single run.
int query_tiles(cytom_ctx* ctx, QueryEngine* qe) {
  if (qe->qt && qe->qt_batch == ctx->batches) return 0;
  free_query_tiles(qe);
  const uint64_t V = ctx->V, E = qe->E;
  uint32_t* src = nullptr;
  int rc;
  if ((rc = dev_alloc(ctx, &src, E * 4))) return rc;
  k_row_sources<<<blocks_for(V * 64, 256), 256, 0, ctx->stream>>>(qe->row_off, V, src);
  if (hipStreamSynchronize(ctx->stream) != hipSuccess) {
    (void)hipFree(src);
    return set_err(ctx, CYTOM_EDEVICE, "query tile sources failed");
  }
  cytom_config cfg{};
  cfg.n_vertices = V;
  cfg.vertices_per_tile = QUERY_TILE_VPT;
  cfg.max_iterations = ctx->max_iter;
  cfg.delta = ctx->delta;
  const uint64_t T = std::max<uint64_t>(V / QUERY_TILE_VPT, 1);
  // room for every tile's first block (1.5 x its edges + 64 slots)
  cfg.edge_capacity = 2 * E + 64 * std::min<uint64_t>(T * T, E) + 4096;
  cfg.weighted = ctx->weighted ? 1 : 0;
  cfg.device = (uint32_t)ctx->device;
  cytom_ctx* qt = nullptr;
  if ((rc = cytom_create(&cfg, &qt))) {
    (void)hipFree(src);
    return set_err(ctx, rc, "query tile context creation failed");
  }
  qt->prd_shift = ctx->prd_shift;
  qt->resort_pct = CYTOM_RESORT_OFF;  // the batch sort already orders each tile by source
  if (E) rc = ingest_batch(qt, src, qe->col, ctx->weighted ? qe->wt : nullptr, E, nullptr,
                          nullptr, nullptr, 0, nullptr);
  (void)hipFree(src);
  if (rc) {
    const std::string msg = "query tile ingest failed: " + qt->err;
    (void)cytom_destroy(qt);
    return set_err(ctx, rc, msg);
  }
  qe->qt = qt;
  qe->qt_batch = ctx->batches;
  return 0;
}

void accumulate(cytom_run_stats* acc, const cytom_run_stats& rs) {
  acc->iterations += rs.iterations;
  acc->traversed_edges += rs.traversed_edges;
  acc->active_vertices += rs.active_vertices;
  acc->slots_scanned += rs.slots_scanned;
  if (!rs.converged) acc->converged = 0;
}

// SSSP queries on the query tiles, eight per run (KIND_SSSP_OCT): one pass
// over the slots and weights per iteration serves all eight, and each query
// reaches its own fixpoint (the single run's distances).
int run_tile_sssp(cytom_ctx* ctx, QueryEngine* qe, const cytom_query* q, const uint32_t* idx,
                  uint32_t n, uint32_t* buf, cytom_run_stats* acc) {
  cytom_ctx* qt = qe->qt;
  Program P;
  P.kind = KIND_SSSP_OCT;
  int rc = program_alloc(qt, P);
  if (rc) return set_err(ctx, rc, qt->err);
  if (hipMalloc((void**)&P.qseed, 3 * 8 * 4) != hipSuccess) {
    program_free(P);
    return set_err(ctx, CYTOM_ENOMEM, "query seed allocation failed");
  }
  for (uint32_t k = 0; k < n && !rc; k += 8) {
    const uint32_t m = std::min<uint32_t>(8, n - k);
    std::vector<uint32_t> seed(2 * m + 8, 0u);
    for (uint32_t j = 0; j < m; ++j) {
      seed[j] = q[idx[k + j]].sources[0];
      seed[m + j] = j;
    }
    P.n_qseed = m;
    cytom_run_stats rs{};
    if (hipMemcpy(P.qseed, seed.data(), seed.size() * 4, hipMemcpyHostToDevice) != hipSuccess) {
      rc = set_err(ctx, CYTOM_EDEVICE, "query seed copy failed");
    } else if ((rc = program_init(qt, P)) || (rc = run_program(qt, P, &rs))) {
      rc = set_err(ctx, rc, qt->err);
    } else {
      accumulate(acc, rs);
      k_split_vec<8, uint32_t><<<blocks_for(ctx->V * m, 256, 16384), 256, 0, qt->stream>>>(
          P.cur, ctx->V, m, buf + (uint64_t)k * ctx->V);
    }
  }
  (void)hipStreamSynchronize(qt->stream);
  program_free(P);
  return rc;
}

// SSSP and PPR on large graphs: the vector groups trade their shared edge
// reads for 16x wider random vertex-state reads, which loses to the tile
// engine's LDS-windowed gather once the graph outgrows the caches.  Such
// groups run the tile-engine program once per query instead (same values by
// construction), each result copied into the group's buffer.
constexpr uint64_t VECTOR_MAX_EDGES = 1ull << 24;
// Up to this many edges SSSP groups run on the query tiles, eight queries per
// run (the one-batch ingest of the snapshot needs ~40 B of scratch per edge);
// beyond, two per run on the context's own tiles.
constexpr uint64_t QUERY_TILES_MAX_EDGES = 1ull << 30;
constexpr int SEQ_ROUTE = 8;  // group_kind flag: one V-sized result per query

int run_seq_group(cytom_ctx* ctx, int kind, const cytom_query* q, const uint32_t* idx, uint32_t n,
                  uint32_t* buf, cytom_run_stats* acc) {
  Program P;
  P.kind = kind;
  P.source = q[idx[0]].sources[0];
  int rc = program_alloc(ctx, P);
  if (rc) return rc;
  for (uint32_t k = 0; k < n && !rc; ++k) {
    const cytom_query& x = q[idx[k]];
    cytom_run_stats rs{};
    if ((rc = program_set_sources(ctx, P, x.sources, x.n_sources))) break;
    if ((rc = run_program(ctx, P, &rs))) break;
    acc->iterations += rs.iterations;
    acc->traversed_edges += rs.traversed_edges;
    acc->active_vertices += rs.active_vertices;
    acc->slots_scanned += rs.slots_scanned;
    if (!rs.converged) acc->converged = 0;
    const void* src = prd_like(kind) ? (const void*)P.rank : (const void*)P.cur;
    if (hipMemcpyAsync(buf + (uint64_t)k * ctx->V, src, ctx->V * 4, hipMemcpyDeviceToDevice,
                       ctx->stream) != hipSuccess)
      rc = set_err(ctx, CYTOM_EDEVICE, "query result copy failed");
  }
  (void)hipStreamSynchronize(ctx->stream);
  program_free(P);
  return rc;
}

// Distances of a pair run into two result rows.
__global__ void k_split_pair(const uint2* __restrict__ cur, uint64_t V, uint32_t* __restrict__ a,
                             uint32_t* __restrict__ b) {
  for (uint64_t v = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; v < V;
       v += (uint64_t)gridDim.x * blockDim.x) {
    const uint2 x = cur[v];
    a[v] = x.x;
    b[v] = x.y;
  }
}

// SSSP queries of the tile-engine route, two per run (KIND_SSSP_PAIR): one
// pass over the slots and weights per iteration serves both, and each
// query's distances are its own fixpoint, the same as a single run.  An odd
// last query runs alone.
int run_seq_pairs(cytom_ctx* ctx, const cytom_query* q, const uint32_t* idx, uint32_t n,
                  uint32_t* buf, cytom_run_stats* acc) {
  int rc = 0;
  if (n >= 2) {
    Program P;
    P.kind = KIND_SSSP_PAIR;
    if ((rc = program_alloc(ctx, P))) return rc;
    for (uint32_t k = 0; k + 1 < n && !rc; k += 2) {
      P.source = q[idx[k]].sources[0];
      P.source2 = q[idx[k + 1]].sources[0];
      cytom_run_stats rs{};
      if ((rc = program_init(ctx, P)) || (rc = run_program(ctx, P, &rs))) break;
      acc->iterations += rs.iterations;
      acc->traversed_edges += rs.traversed_edges;
      acc->active_vertices += rs.active_vertices;
      acc->slots_scanned += rs.slots_scanned;
      if (!rs.converged) acc->converged = 0;
      k_split_pair<<<blocks_for(ctx->V, 256, 16384), 256, 0, ctx->stream>>>(
          reinterpret_cast<const uint2*>(P.cur), ctx->V, buf + (uint64_t)k * ctx->V,
          buf + (uint64_t)(k + 1) * ctx->V);
    }
    (void)hipStreamSynchronize(ctx->stream);
    program_free(P);
  }
  if (!rc && (n & 1))
    rc = run_seq_group(ctx, CYTOM_SSSP, q, idx + n - 1, 1, buf + (uint64_t)(n - 1) * ctx->V, acc);
  return rc;
}

}  // namespace

void query_free(cytom_ctx* ctx) {
  if (!ctx->qe) return;
  free_results(ctx->qe);
  free_snapshot(ctx->qe);
  delete ctx->qe;
  ctx->qe = nullptr;
}

int batched_execute(cytom_ctx* ctx, const cytom_query* q, uint32_t nq, cytom_run_stats* stats) {
  for (uint32_t i = 0; i < nq; ++i) {
    const cytom_query& x = q[i];
    if (x.kind != CYTOM_BFS && x.kind != CYTOM_SSSP && x.kind != CYTOM_PPR)
      return set_err(ctx, CYTOM_EINVAL, "batched queries are BFS, SSSP or PPR");
    if (x.kind == CYTOM_SSSP && !ctx->weighted)
      return set_err(ctx, CYTOM_EINVAL, "SSSP needs a weighted graph");
    if (!x.sources || x.n_sources == 0) return set_err(ctx, CYTOM_EINVAL, "query without sources");
    if (x.kind != CYTOM_PPR && x.n_sources != 1)
      return set_err(ctx, CYTOM_EINVAL, "BFS and SSSP queries take one source");
    for (uint32_t k = 0; k < x.n_sources; ++k)
      if (x.sources[k] >= ctx->V) return set_err(ctx, CYTOM_EINVAL, "source vertex out of range");
  }
  if (!ctx->qe) ctx->qe = new (std::nothrow) QueryEngine();
  if (!ctx->qe) return set_err(ctx, CYTOM_ENOMEM, "query engine allocation failed");
  QueryEngine* qe = ctx->qe;
  free_results(qe);
  int rc;
  if ((rc = snapshot(ctx, qe))) return rc;
  hipStream_t st = ctx->stream;
  CY_HIP(ctx, hipEventRecord(ctx->ev[0], st));
  cytom_run_stats acc{};
  acc.converged = 1;
  qe->queries.assign(nq, QueryEngine::Q{0, 0, 0});
  const uint64_t V = ctx->V;
  // Group queries of one kind in submission order.
  for (int kind : {CYTOM_BFS, CYTOM_SSSP, CYTOM_PPR}) {
    const uint32_t width = kind == CYTOM_BFS ? BFS_W : VEC_W;
    std::vector<uint32_t> idx;
    for (uint32_t i = 0; i < nq; ++i)
      if ((int)q[i].kind == kind) idx.push_back(i);
    const bool seq = kind != CYTOM_BFS && qe->E > VECTOR_MAX_EDGES;
    std::vector<std::vector<uint32_t>> bfs_src;  // BFS groups, run together below
    std::vector<uint16_t*> bfs_lev;
    for (size_t g0 = 0; g0 < idx.size(); g0 += width) {
      const size_t g1 = std::min(idx.size(), g0 + width);
      const uint32_t gid = (uint32_t)qe->group_buf.size();
      const size_t bytes = kind == CYTOM_BFS ? V * BFS_W * 2 : V * VEC_W * 4;
      void* buf = nullptr;
      if (hipMalloc(&buf, bytes) != hipSuccess)
        return set_err(ctx, CYTOM_ENOMEM, "query result allocation failed");
      qe->group_buf.push_back(buf);
      qe->group_kind.push_back(seq ? kind | SEQ_ROUTE : kind);
      for (size_t k = g0; k < g1; ++k) qe->queries[idx[k]] = QueryEngine::Q{kind, gid, (uint32_t)(k - g0)};
      const bool tiles = seq && kind == CYTOM_SSSP && qe->E <= QUERY_TILES_MAX_EDGES;
      if (tiles && (rc = query_tiles(ctx, qe))) return rc;
      if (tiles) {
        rc = run_tile_sssp(ctx, qe, q, idx.data() + g0, (uint32_t)(g1 - g0),
                           static_cast<uint32_t*>(buf), &acc);
      } else if (seq && kind == CYTOM_SSSP) {
        rc = run_seq_pairs(ctx, q, idx.data() + g0, (uint32_t)(g1 - g0),
                           static_cast<uint32_t*>(buf), &acc);
      } else if (seq) {
        rc = run_seq_group(ctx, kind, q, idx.data() + g0, (uint32_t)(g1 - g0),
                           static_cast<uint32_t*>(buf), &acc);
      } else if (kind == CYTOM_BFS) {
        bfs_src.emplace_back();
        for (size_t k = g0; k < g1; ++k) bfs_src.back().push_back(q[idx[k]].sources[0]);
        bfs_lev.push_back(static_cast<uint16_t*>(buf));
      } else if (kind == CYTOM_SSSP) {
        std::vector<uint32_t> src;
        for (size_t k = g0; k < g1; ++k) src.push_back(q[idx[k]].sources[0]);
        rc = run_sssp_group(ctx, qe, src, static_cast<uint32_t*>(buf), &acc);
      } else {
        std::vector<std::vector<uint32_t>> sets;
        for (size_t k = g0; k < g1; ++k)
          sets.emplace_back(q[idx[k]].sources, q[idx[k]].sources + q[idx[k]].n_sources);
        rc = run_ppr_group(ctx, qe, sets, static_cast<float*>(buf), &acc);
      }
      if (rc) return rc;
    }
    if (!bfs_src.empty() && (rc = run_bfs_groups(ctx, qe, bfs_src, bfs_lev, &acc))) return rc;
  }
  CY_HIP(ctx, hipEventRecord(ctx->ev[1], st));
  CY_HIP(ctx, hipEventSynchronize(ctx->ev[1]));
  float ms = 0.f;
  CY_HIP(ctx, hipEventElapsedTime(&ms, ctx->ev[0], ctx->ev[1]));
  acc.t_compute_ms = ms;
  acc.reexecuted = 0;
  if (stats) *stats = acc;
  return 0;
}

int read_query(cytom_ctx* ctx, uint32_t query, void* out, size_t bytes) {
  QueryEngine* qe = ctx->qe;
  if (!qe || query >= qe->queries.size())
    return set_err(ctx, CYTOM_EINVAL, "no such query in the last batch");
  const uint64_t V = ctx->V;
  if (!out || bytes < V * 4) return set_err(ctx, CYTOM_EINVAL, "output buffer too small");
  const QueryEngine::Q& x = qe->queries[query];
  hipStream_t st = ctx->stream;
  const void* buf = qe->group_buf[x.group];
  if (qe->group_kind[x.group] & SEQ_ROUTE) {
    const hipError_t e =
        hipMemcpy(out, static_cast<const uint32_t*>(buf) + (uint64_t)x.lane * V, V * 4,
                  hipMemcpyDeviceToHost);
    return e == hipSuccess ? 0 : hip_fail(ctx, e, "read_query");
  }
  uint32_t* tmp = nullptr;
  CY_HIP(ctx, hipMalloc(&tmp, V * 4));
  if (x.kind == CYTOM_BFS)
    k_extract_lev<<<blocks_for(V, 256, 16384), 256, 0, st>>>(static_cast<const uint16_t*>(buf),
                                                             x.lane, V, tmp);
  else
    k_extract_vec<<<blocks_for(V, 256, 16384), 256, 0, st>>>(static_cast<const uint32_t*>(buf),
                                                             x.lane, V, tmp);
  hipError_t e = hipMemcpyAsync(out, tmp, V * 4, hipMemcpyDeviceToHost, st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  (void)hipFree(tmp);
  if (e != hipSuccess) return hip_fail(ctx, e, "read_query");
  return 0;
}

}  // namespace cytom
