// Incremental vertex-program executors on the device.
//
// Replaces ContinuousAlgorithmExecutor::executeAlgorithm
// (continuous-algorithm-executor.h:49-249), the pull workers
// (AlgorithmExecutor, algorithm-executor.h:110-505), the TileDistributor
// (tile-distributor.h:54-177) and the ApplyExecutor (apply-executor.h:40-115)
// for the BFS, CC, SSSP and PageRank-delta programs
// (examples/{bfs,cc,sssp,pagerank-delta}.cc).
//
// One iteration = gather + apply:
//  * gather: a persistent grid takes "column jobs" from a device work queue.
//    A job is a run of edge chunks of one target column y (tiles (x, y) for
//    all x); the block keeps the column's vertex values in LDS (the
//    reference's per-tile-row local array, algorithm-executor.h:160-200),
//    reads packed edges coalesced, loads the source's frontier value (its
//    value while active, else the neutral element, so the active-bitmap
//    test costs no extra load) and folds candidates with LDS atomics (min
//    for BFS/CC/SSSP, add for PageRank-delta).  The window is then folded into the global next
//    array with one atomic per touched vertex (applyTiles,
//    algorithm-executor.h:431-505), marking `changed`.
//  * apply: one thread per vertex applies the program to changed vertices;
//    the next active bitmap is built a wavefront word at a time by ballot,
//    and the frontier size, its live out-edge count and row activity are
//    reduced on the fly.
#include <algorithm>

#include <hip/hip_ext.h>

#include "devutil.h"
#include "internal.h"

namespace cytom {
namespace {

constexpr int GATHER_THREADS = 1024;
constexpr int APPLY_THREADS = 1024;
constexpr int GATHER_ROWS = 8;  // 64-slot rows per wavefront step
// Rows per step of a kind: the eight-query SSSP kind holds 8 distances per
// row in flight, so it takes 4 rows per step to stay within its registers.
template <int KIND>
constexpr int RPS = KIND == KIND_SSSP_OCT ? 4 : GATHER_ROWS;

inline uint32_t grid_for(uint64_t n, uint32_t threads, uint32_t cap = 65536) {
  uint64_t b = (n + threads - 1) / threads;
  if (b == 0) b = 1;
  return (uint32_t)std::min<uint64_t>(b, cap);
}

// ------------------------------------------------------ source groups
// A source row (the vpt sources of one tile column x) is cut into 64 groups
// of vpt / 64 consecutive sources.  Activity is kept as one byte per group
// (gact[x * 64 + g] = 1 when a vertex of the group is active; plain stores,
// so the apply needs no atomics), and every 64-slot row of the edge pool
// carries the 64-bit mask of the groups its sources fall in (grp); the
// gather reads a row only when the row's mask meets its source row's
// active groups, so a sparse frontier skips most rows of an active source
// row (TileManager::isActive, tile-manager.h:403-455, at group granularity).
__device__ inline uint64_t word_groups(uint64_t word, uint64_t v0, uint32_t log2vpt) {
  const uint32_t gshift = log2vpt - 6;  // log2 of the group width
  const uint32_t g0 = (uint32_t)((v0 & ((1ull << log2vpt) - 1)) >> gshift);
  if (gshift >= 6) return word ? 1ull << g0 : 0ull;  // the word lies in one group
  const uint32_t gs = 1u << gshift;
  const uint64_t gmask = (1ull << gs) - 1;
  uint64_t m = 0;
  for (uint32_t k = 0; k < 64 / gs; ++k)
    if ((word >> (k * gs)) & gmask) m |= 1ull << (g0 + k);
  return m;
}

// Marks the groups of one active-bitmap word (vertices v0 .. v0 + 63).
__device__ inline void mark_groups(uint8_t* __restrict__ gact, uint64_t word, uint64_t v0,
                                   uint32_t log2vpt) {
  if (!word) return;
  uint64_t m = word_groups(word, v0, log2vpt);
  uint8_t* row = gact + ((v0 >> log2vpt) << 6);
  while (m) {
    row[__ffsll((unsigned long long)m) - 1] = 1;
    m &= m - 1;
  }
}

// The active groups of source row x as a 64-bit mask.  Each 8-byte word
// holds 8 group bytes of value 0 or 1; multiplying by sum_j 2^(56-7j) moves
// byte j's bit to bit 56+j (all other partial products land below bit 56 or
// above bit 63, at distinct positions, so nothing carries into the top byte).
__device__ inline uint64_t group_mask(const uint8_t* __restrict__ gact, uint32_t x) {
  const uint64_t* p = reinterpret_cast<const uint64_t*>(gact + ((uint64_t)x << 6));
  uint64_t m = 0;
#pragma unroll
  for (int i = 0; i < 8; ++i) m |= ((p[i] * 0x0102040810204080ull) >> 56) << (8 * i);
  return m;
}

// ---------------------------------------------------------------- init
// bfs.cc:170-197, cc.cc:171-189, sssp.cc:171-195 (+ source set to 0).
// A gate (nullable) holds the batch's decision for a program, written on the
// device by the update batch: 1 = re-initialise, 0 = run the insert hooks,
// 2 = neither (the batch was refused).  Init kernels run when the gate is
// null or 1, hook kernels when it is null or 0.
__device__ inline bool gate_is(const unsigned long long* gate, unsigned long long v) {
  return gate == nullptr || *gate == v;
}

__global__ void k_init_min(int kind, uint64_t V, uint32_t source, uint32_t* __restrict__ cur,
                           uint32_t* __restrict__ nxt, uint64_t* __restrict__ act,
                           uint8_t* __restrict__ changed, const unsigned long long* gate) {
  if (!gate_is(gate, 1)) return;
  for (uint64_t v = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; v < V;
       v += (uint64_t)gridDim.x * blockDim.x) {
    uint32_t val = kind == CYTOM_CC ? (uint32_t)v : (v == source ? 0u : INF);
    cur[v] = val;
    nxt[v] = val;
    changed[v] = 0;
    if ((v & 63) == 0) {
      const uint64_t w = v >> 6;
      uint64_t word = 0;
      if (kind == CYTOM_CC) word = ~0ull;
      else if ((source >> 6) == w) word = 1ull << (source & 63);
      act[w] = word;
    }
  }
}

// Two SSSP queries (KIND_SSSP_PAIR): distance pairs, each source at 0 in its
// own component, both sources active.
__global__ void k_init_pair(uint64_t V, uint32_t s0, uint32_t s1, uint2* __restrict__ cur,
                            uint2* __restrict__ nxt, uint64_t* __restrict__ act,
                            uint8_t* __restrict__ changed, const unsigned long long* gate) {
  if (!gate_is(gate, 1)) return;
  for (uint64_t v = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; v < V;
       v += (uint64_t)gridDim.x * blockDim.x) {
    const uint2 val = make_uint2(v == s0 ? 0u : INF, v == s1 ? 0u : INF);
    cur[v] = val;
    nxt[v] = val;
    changed[v] = 0;
    if ((v & 63) == 0) {
      const uint64_t w = v >> 6;
      uint64_t word = 0;
      if ((s0 >> 6) == w) word |= 1ull << (s0 & 63);
      if ((s1 >> 6) == w) word |= 1ull << (s1 & 63);
      act[w] = word;
    }
  }
}

// pagerank-delta.cc:246-270 (rank 0, delta 1, all active); for PPR every
// vertex starts at rank = delta = 0, inactive, before k_init_ppr_set.
__global__ void k_init_prd(uint64_t V, int personalized, float* __restrict__ rank,
                           float* __restrict__ delta, long long* __restrict__ dsum,
                           uint64_t* __restrict__ act, uint8_t* __restrict__ changed,
                           uint32_t* __restrict__ dctl, const unsigned long long* gate) {
  if (!gate_is(gate, 1)) return;
  // the first-iteration apply rule (pagerank-delta.cc:148-156) for the next run
  if (blockIdx.x == 0 && threadIdx.x == 0) dctl[0] = personalized ? 0u : 1u;
  for (uint64_t v = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; v < V;
       v += (uint64_t)gridDim.x * blockDim.x) {
    rank[v] = 0.f;
    delta[v] = personalized ? 0.f : 1.f;
    dsum[v] = 0;
    changed[v] = 0;
    if ((v & 63) == 0) act[v >> 6] = personalized ? 0ull : ~0ull;
  }
}

// PPR teleport set: rank = delta = (1-alpha)/|T|, active.
__global__ void k_init_ppr_set(const uint32_t* __restrict__ set, uint32_t n, float* __restrict__ rank,
                               float* __restrict__ delta, uint64_t* __restrict__ act,
                               const unsigned long long* gate) {
  if (!gate_is(gate, 1)) return;
  const float r0 = (float)((1. - 0.85) * (1.0 / (double)n));
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const uint32_t v = set[i];
    rank[v] = r0;
    delta[v] = r0;
    atomicOr(reinterpret_cast<unsigned long long*>(&act[v >> 6]), 1ull << (v & 63));
  }
}

// W SSSP queries per vertex (the batched queries' eight-query kind): every
// distance INF, nothing active.
template <uint32_t W>
__global__ void k_init_vec(uint64_t V, uint32_t* __restrict__ cur, uint32_t* __restrict__ nxt,
                           uint64_t* __restrict__ act, uint8_t* __restrict__ changed) {
  for (uint64_t v = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; v < V;
       v += (uint64_t)gridDim.x * blockDim.x) {
#pragma unroll
    for (uint32_t q = 0; q < W; ++q) {
      cur[v * W + q] = INF;
      nxt[v * W + q] = INF;
    }
    changed[v] = 0;
    if ((v & 63) == 0) act[v >> 6] = 0;
  }
}

// Sources: seed[i] is the source of query seed[n + i] (distance 0, active).
template <uint32_t W>
__global__ void k_seed_vec(const uint32_t* __restrict__ seed, uint32_t n,
                           uint32_t* __restrict__ cur, uint32_t* __restrict__ nxt,
                           uint64_t* __restrict__ act) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const uint32_t v = seed[i], q = seed[n + i];
    cur[(uint64_t)v * W + q] = 0;
    nxt[(uint64_t)v * W + q] = 0;
    atomicOr(reinterpret_cast<unsigned long long*>(&act[v >> 6]), 1ull << (v & 63));
  }
}

// ------------------------------------------------------------- frontier
// Frontier size, live out-edges and row activity of the active bitmap
// (TileManager::isActive over full words); PRD contributions of active
// vertices.
__global__ void __launch_bounds__(APPLY_THREADS) k_frontier(
    int kind, uint64_t V, uint32_t log2vpt, const uint64_t* __restrict__ act,
    const uint32_t* __restrict__ outdeg, uint8_t* __restrict__ gact,
    const float* __restrict__ delta, float* __restrict__ contrib,
    const uint32_t* __restrict__ cur, uint32_t* __restrict__ fval, IterStats* it,
    const unsigned long long* refuse) {
  if (batch_refused(refuse)) return;  // the run follows a refused batch: it does nothing
  const int lane = threadIdx.x & 63;
  const uint64_t nwords = V >> 6;
  // Each wavefront takes FW words per step and loads them all before using
  // any (several memory round trips in flight per wavefront).
  constexpr int FW = 4;
  const uint64_t stride = (((uint64_t)gridDim.x * blockDim.x) >> 6) * FW;
  unsigned long long te = 0, cnt = 0;
  for (uint64_t w0 = (((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6) * FW; w0 < nwords;
       w0 += stride) {
    uint64_t word[FW];
#pragma unroll
    for (int k = 0; k < FW; ++k) word[k] = act[min(w0 + k, nwords - 1)];
    uint32_t od[FW];
    uint32_t cv[FW];
#pragma unroll
    for (int k = 0; k < FW; ++k) {
      const uint64_t v = (min(w0 + k, nwords - 1) << 6) | lane;
      od[k] = outdeg[v];
      cv[k] = (kind == CYTOM_BFS || kind == CYTOM_CC || kind == CYTOM_SSSP) ? cur[v] : 0u;
    }
#pragma unroll
    for (int k = 0; k < FW; ++k) {
      const uint64_t w = w0 + k;
      if (w >= nwords) break;  // uniform
      const uint64_t v = (w << 6) | lane;
      const bool a = (word[k] >> lane) & 1ull;
      if (a) {
        te += od[k];
        ++cnt;
      }
      // Frontier values: what the gather reads per edge instead of the bitmap.
      if (kind == KIND_SSSP_OCT) {
        for (uint32_t q = 0; q < 8; ++q) fval[v * 8 + q] = a ? cur[v * 8 + q] : INF;
      } else if (prd_like(kind)) {
        contrib[v] = (a && od[k]) ? delta[v] / (float)od[k] : 0.f;
      } else if (kind == KIND_SSSP_PAIR) {
        reinterpret_cast<uint2*>(fval)[v] =
            a ? reinterpret_cast<const uint2*>(cur)[v] : make_uint2(INF, INF);
      } else {
        fval[v] = a ? cv[k] : INF;
      }
      if (lane == 0) mark_groups(gact, word[k], v, log2vpt);
    }
  }
  te = block_sum<APPLY_THREADS, unsigned long long>(te);
  cnt = block_sum<APPLY_THREADS, unsigned long long>(cnt);
  if (threadIdx.x == 0) {
    if (te) atomicAdd(&it->te, te);
    if (cnt) atomicAdd(&it->count, cnt);
  }
}

// ------------------------------------------------------------ work lists
__global__ void k_nch(const uint32_t* __restrict__ cnt, uint64_t n_tiles, uint32_t CH,
                      uint32_t* __restrict__ nch) {
  for (uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n_tiles;
       t += (uint64_t)gridDim.x * blockDim.x)
    nch[t] = (cnt[t] + CH - 1) / CH;
}

// One wavefront per tile, its lanes over the tile's chunks (a heavy RMAT
// tile holds thousands).
__global__ void k_emit_chunks(const uint32_t* __restrict__ nch, const uint32_t* __restrict__ cbase,
                              uint64_t n_tiles, uint32_t CH, uint64_t* __restrict__ chunks) {
  const uint32_t lane = threadIdx.x & 63;
  for (uint64_t t = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6; t < n_tiles;
       t += ((uint64_t)gridDim.x * blockDim.x) >> 6) {
    const uint32_t n = nch[t], b = cbase[t];
    for (uint32_t c = lane; c < n; c += 64) chunks[b + c] = ((uint64_t)t << 32) | (uint64_t)(c * CH);
  }
}

// Group mask of every 64-slot row of the store (tile blocks start at
// multiples of 64 slots, so row r of a tile at slot offset o is grp[o/64 +
// r]).  One wavefront per work chunk; tombstones and slots past the tile's
// count add nothing.
__global__ void k_row_groups(const uint64_t* __restrict__ chunks,
                             const uint32_t* __restrict__ n_chunks_dev, StoreView S, uint32_t CH,
                             uint64_t* __restrict__ grp) {
  const uint32_t lane = threadIdx.x & 63;
  const uint64_t n_chunks = *n_chunks_dev;
  const uint32_t gshift = S.log2vpt - 6;
  for (uint64_t c = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6; c < n_chunks;
       c += ((uint64_t)gridDim.x * blockDim.x) >> 6) {
    const uint64_t ch = chunks[c];
    const uint32_t t = (uint32_t)(ch >> 32), s0 = (uint32_t)ch;
    const uint32_t n = min(CH, S.tile_cnt[t] - s0);
    const uint64_t off = S.tile_off[t] + s0;
    for (uint32_t r = 0; r * 64 < n; ++r) {
      const uint32_t i = r * 64 + lane;
      uint64_t m = 0;
      if (i < n) {
        const uint32_t e = S.pool[off + i];
        if (e != TOMB) m = 1ull << ((e & 0xFFFFu) >> gshift);
      }
#pragma unroll
      for (int o = 32; o > 0; o >>= 1) m |= __shfl_xor(m, o, 64);
      if (lane == 0) grp[(off >> 6) + r] = m;
    }
  }
}

// Row descriptors of every chunk in chunk order (a chunk of n slots owns
// ceil(n/64) rows; rowb, the exclusive scan of those counts, gives a job's
// rows as rowb[first chunk] .. rowb[end chunk]).  A gather over a dense
// frontier copies them as they are instead of rebuilding them per pass.
__global__ void k_chunk_rows(const uint64_t* __restrict__ chunks,
                             const uint32_t* __restrict__ n_chunks_dev,
                             const uint32_t* __restrict__ cnt, uint32_t CH,
                             uint32_t* __restrict__ nrow) {
  const uint32_t n_chunks = *n_chunks_dev;
  for (uint32_t c = blockIdx.x * blockDim.x + threadIdx.x; c < n_chunks;
       c += gridDim.x * blockDim.x) {
    const uint64_t ch = chunks[c];
    nrow[c] = (min(CH, cnt[ch >> 32] - (uint32_t)ch) + 63) >> 6;
  }
}

// Jobs are cut from cells: cell (b, y) holds the chunks of column y whose
// source rows fall in band b (source rows [xb[b], xb[b+1])).  Cells are
// numbered band-major, so the jobs of one band form one contiguous queue.
// Band boundaries split the source rows into parts of about equal slot
// counts; each band's queue is served first by the blocks whose number is
// the band modulo the band count.  Workgroups are dispatched round-robin
// over the XCDs, so with 8 bands each XCD works on one band: the source
// values its jobs read (1/8 of them) stay in that XCD's L2.  A block whose
// queue is empty takes jobs from the other bands.
__device__ inline void cell_range(const uint32_t* cbase, const uint32_t* xb, uint32_t T,
                                  uint32_t cell, uint32_t* y, uint32_t* c_begin,
                                  uint32_t* c_end) {
  const uint32_t b = cell / T;
  *y = cell % T;
  const uint64_t t0 = (uint64_t)*y * T;
  *c_begin = cbase[t0 + xb[b]];
  *c_end = cbase[t0 + xb[b + 1]];
}

// Slots per source row (summed over the columns).
__global__ void k_row_slots(const uint32_t* __restrict__ cnt, uint32_t T,
                            unsigned long long* __restrict__ rs) {
  // one wavefront per source row, its lanes over the row's tiles
  const uint32_t lane = threadIdx.x & 63;
  for (uint64_t x = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6; x < T;
       x += ((uint64_t)gridDim.x * blockDim.x) >> 6) {
    unsigned long long s = 0;
    for (uint32_t y = lane; y < T; y += 64) s += cnt[(uint64_t)y * T + x];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
    if (lane == 0) rs[x] = s;
  }
}

// Band b starts at the first source row whose slot prefix reaches
// b / NB of the total (bands may be empty).  One thread: T is at most 2^17.
__global__ void k_band_bounds(const unsigned long long* __restrict__ rs, uint32_t T, uint32_t NB,
                              uint32_t* __restrict__ xb) {
  if (blockIdx.x != 0 || threadIdx.x != 0) return;
  unsigned long long tot = 0;
  for (uint32_t x = 0; x < T; ++x) tot += rs[x];
  xb[0] = 0;
  uint32_t b = 1;
  unsigned long long pre = 0;
  for (uint32_t x = 0; x < T && b < NB; ++x) {
    while (b < NB && pre * NB >= tot * b) xb[b++] = x;
    pre += rs[x];
  }
  while (b < NB) xb[b++] = T;
  xb[NB] = T;
}

// Chunks per job: job_chunks when set, else max(2^17 slots, n_chunks / (8 CUs)).
__global__ void k_job_length(const uint32_t* __restrict__ n_chunks, uint32_t job_chunks,
                             uint32_t CH, uint32_t cus, uint32_t* __restrict__ jc) {
  if (blockIdx.x != 0 || threadIdx.x != 0) return;
  *jc = job_chunks ? job_chunks : max((1u << 17) / CH, *n_chunks / (cus * 8u));
}

__global__ void k_col_jobs(const uint32_t* __restrict__ cbase, const uint32_t* __restrict__ xb,
                           uint32_t T, uint32_t NB, const uint32_t* __restrict__ jc,
                           uint32_t* __restrict__ colj) {
  const uint32_t JC = *jc;
  for (uint32_t c = blockIdx.x * blockDim.x + threadIdx.x; c < T * NB;
       c += gridDim.x * blockDim.x) {
    uint32_t y, c0, c1;
    cell_range(cbase, xb, T, c, &y, &c0, &c1);
    colj[c] = (c1 - c0 + JC - 1) / JC;
  }
}

__global__ void k_emit_jobs(const uint32_t* __restrict__ cbase, const uint32_t* __restrict__ xb,
                            const uint32_t* __restrict__ colj, const uint32_t* __restrict__ jbase,
                            uint32_t T, uint32_t NB, const uint32_t* __restrict__ jc,
                            uint4* __restrict__ jobs) {
  const uint32_t JC = *jc;
  for (uint32_t c = blockIdx.x * blockDim.x + threadIdx.x; c < T * NB;
       c += gridDim.x * blockDim.x) {
    uint32_t y, c_begin, c_end;
    cell_range(cbase, xb, T, c, &y, &c_begin, &c_end);
    const uint32_t nj = colj[c], b = jbase[c];
    for (uint32_t j = 0; j < nj; ++j) {
      const uint32_t c0 = c_begin + j * JC;
      const uint32_t c1 = min(c0 + JC, c_end);
      jobs[b + j] = make_uint4(c0, c1, y, 0u);
    }
  }
}

// Largest jobs first within each band's queue (longest-processing-time
// order): a block that takes a job late takes a short one, so the blocks of
// a launch finish close together.  Sort key: the band (bits 16+), then the
// job's row count, descending, saturated at 2^16 - 1 rows (bits 0-15); the
// sort is stable.
__global__ void k_job_keys(const uint4* __restrict__ jobs, const uint32_t* __restrict__ jbase,
                           uint32_t T, uint32_t NB, const uint32_t* __restrict__ rowb,
                           uint32_t* __restrict__ key, uint32_t* __restrict__ idx) {
  const uint32_t n = jbase[T * NB];
  for (uint32_t j = blockIdx.x * blockDim.x + threadIdx.x; j < n; j += gridDim.x * blockDim.x) {
    uint32_t b = 0;
    while (b + 1 < NB && jbase[(b + 1) * T] <= j) ++b;
    const uint4 jb = jobs[j];
    const uint32_t rows = min(rowb[jb.y] - rowb[jb.x], 0xFFFFu);
    key[j] = (b << 16) | (0xFFFFu - rows);
    idx[j] = j;
  }
}

__global__ void k_permute_jobs(const uint4* __restrict__ in, const uint32_t* __restrict__ idx,
                               const uint32_t* __restrict__ n_dev, uint4* __restrict__ out) {
  const uint32_t n = *n_dev;
  for (uint32_t j = blockIdx.x * blockDim.x + threadIdx.x; j < n; j += gridDim.x * blockDim.x)
    out[j] = in[idx[j]];
}

// --------------------------------------------------------------- gather
template <int KIND>
struct GatherArgs {
  const uint4* jobs;
  const uint32_t* jbase;  // first job of every cell (cells band-major), cells + 1 entries
  uint32_t nbands;
  const uint64_t* chunks;
  const uint32_t* rowb;        // first row of every chunk in rows (chunks + 1 entries)
  const uint64_t* rows;        // row descriptors of every chunk in chunk order
  StoreView S;
  const uint8_t* gact;         // active groups, 64 bytes per source row
  uint64_t* gact_clear;        // the next iteration's group bytes, cleared here (T * 8 words)
  const uint64_t* grp;         // source groups of every 64-slot pool row
  const uint64_t* act;         // this iteration's active bitmap
  const uint32_t* fval;
  uint32_t* nxt;
  const float* contrib;
  long long* dsum;
  uint8_t* changed;
  IterStats* it;
  uint32_t CH;
  float qscale;  // 2^prd_shift
};

// Row descriptors of one pass of a job held in LDS.
constexpr uint32_t MAX_JOB_ROWS = 2048;

// The active chunks of a job form one list of 64-slot rows: a chunk of n
// slots owns ceil(n/64) rows, so a tile with a handful of edges costs one
// row, not a whole chunk.  A row descriptor packs the pool offset of its
// first slot (bits 22-61), the tile's source column x (bits 6-21) and the
// number of valid slots minus one (bits 0-5).  A wavefront reads the
// descriptors of GATHER_ROWS consecutive rows at once (LDS broadcasts), then
// issues their edge loads back to back; lanes past a row's end load the
// row's first slot and are masked off in vmask (bit u = row u is valid for
// this lane).
__device__ inline uint64_t row_desc(uint64_t off, uint32_t x, uint32_t nvalid) {
  return (off << 22) | ((uint64_t)x << 6) | (uint64_t)(nvalid - 1);
}

// The row descriptors of every chunk, at rowb[c] onwards (one thread per chunk).
__global__ void k_emit_rows(const uint64_t* __restrict__ chunks,
                            const uint32_t* __restrict__ n_chunks_dev, StoreView S, uint32_t CH,
                            const uint32_t* __restrict__ rowb, uint64_t* __restrict__ rows) {
  const uint32_t n_chunks = *n_chunks_dev;
  for (uint32_t c = blockIdx.x * blockDim.x + threadIdx.x; c < n_chunks;
       c += gridDim.x * blockDim.x) {
    const uint64_t ch = chunks[c];
    const uint32_t t = (uint32_t)(ch >> 32), s0 = (uint32_t)ch;
    const uint32_t n = min(CH, S.tile_cnt[t] - s0), x = t % S.T;
    const uint64_t off = S.tile_off[t] + s0;
    uint64_t* out = rows + rowb[c];
    for (uint32_t k = 0; 64 * k < n; ++k) out[k] = row_desc(off + 64 * k, x, min(64u, n - 64 * k));
  }
}

template <int KIND>
__device__ inline void load_rows(const GatherArgs<KIND>& A, const uint64_t* rows, uint32_t R,
                                 uint32_t r0, uint32_t (&e)[RPS<KIND>],
                                 uint32_t (&wt)[RPS<KIND>], uint32_t (&xb)[RPS<KIND>],
                                 uint32_t& vmask) {
  const uint32_t lane = threadIdx.x & 63;
  uint64_t d[RPS<KIND>];
#pragma unroll
  for (int u = 0; u < RPS<KIND>; ++u) d[u] = rows[min(r0 + u, R - 1)];
  vmask = 0;
#pragma unroll
  for (int u = 0; u < RPS<KIND>; ++u) {
    const bool v = r0 + u < R && lane <= (uint32_t)(d[u] & 63u);
    const uint64_t idx = (d[u] >> 22) + (v ? lane : 0u);
    // Slots are read once per launch: non-temporal loads keep them from
    // evicting the source values the L2 should hold.
    e[u] = __builtin_nontemporal_load(&A.S.pool[idx]);
    if (KIND == CYTOM_SSSP || KIND == KIND_SSSP_PAIR || KIND == KIND_SSSP_OCT)
      wt[u] = __builtin_nontemporal_load(&A.S.wpool[idx]);
    xb[u] = ((uint32_t)(d[u] >> 6) & 0xFFFFu) << A.S.log2vpt;
    vmask |= (v ? 1u : 0u) << u;
  }
}

// One wavefront's share of a column job, software-pipelined: while the
// source frontier values and LDS atomics of one row block are in flight,
// the edge loads of its next block are already issued.  A source's frontier
// value is INF (0 for PageRank-delta) while it is inactive, so the
// active-bitmap test costs no extra load.
template <int KIND>
__device__ inline void gather_rows(const GatherArgs<KIND>& A, uint32_t* win,
                                   unsigned long long* winq, const uint64_t* rows, uint32_t R,
                                   bool sparse) {
  const uint32_t step = (GATHER_THREADS / 64) * RPS<KIND>;
  uint32_t r0 = (threadIdx.x >> 6) * RPS<KIND>;
  if (r0 >= R) return;
  uint32_t e[RPS<KIND>], wt[RPS<KIND>], xb[RPS<KIND>], vm;
  load_rows<KIND>(A, rows, R, r0, e, wt, xb, vm);
  for (;;) {
    const uint32_t rn = r0 + step;
    const bool more = rn < R;
    uint32_t s[RPS<KIND>];
    bool ok[RPS<KIND>];
#pragma unroll
    for (int u = 0; u < RPS<KIND>; ++u) {
      ok[u] = ((vm >> u) & 1u) && e[u] != TOMB;
      s[u] = ok[u] ? xb[u] + (e[u] & 0xFFFFu) : 0u;
    }
    float c[RPS<KIND>];
    uint32_t val[RPS<KIND>];
    uint2 val2[RPS<KIND>];
    uint4 v8[RPS<KIND>][2];
    if (KIND != CYTOM_PRD && sparse) {
      // Sparse frontier: the active bitmap (V/8 bytes, held by the L2) says
      // which sources are active, so inactive sources' values, scattered over
      // 4 B per vertex, are not fetched.
#pragma unroll
      for (int u = 0; u < RPS<KIND>; ++u)
        ok[u] = ok[u] && ((A.act[s[u] >> 6] >> (s[u] & 63u)) & 1ull);
    }
    if (KIND == CYTOM_PRD) {
#pragma unroll
      for (int u = 0; u < RPS<KIND>; ++u) c[u] = A.contrib[s[u]];
    } else if (KIND == KIND_SSSP_OCT) {
#pragma unroll
      for (int u = 0; u < RPS<KIND>; ++u) {
        if (!sparse || ok[u]) {
          v8[u][0] = reinterpret_cast<const uint4*>(A.fval)[2 * (uint64_t)s[u]];
          v8[u][1] = reinterpret_cast<const uint4*>(A.fval)[2 * (uint64_t)s[u] + 1];
        } else {
          v8[u][0] = v8[u][1] = make_uint4(INF, INF, INF, INF);
        }
      }
    } else if (KIND == KIND_SSSP_PAIR) {
      if (sparse) {
#pragma unroll
        for (int u = 0; u < RPS<KIND>; ++u)
          val2[u] = ok[u] ? reinterpret_cast<const uint2*>(A.fval)[s[u]] : make_uint2(INF, INF);
      } else {
#pragma unroll
        for (int u = 0; u < RPS<KIND>; ++u)
          val2[u] = reinterpret_cast<const uint2*>(A.fval)[s[u]];
      }
    } else {
      if (sparse) {
#pragma unroll
        for (int u = 0; u < RPS<KIND>; ++u) val[u] = ok[u] ? A.fval[s[u]] : INF;
      } else {
#pragma unroll
        for (int u = 0; u < RPS<KIND>; ++u) val[u] = A.fval[s[u]];
      }
    }
    // Issued unconditionally (past the end the clamped descriptors point at
    // the last row, a valid address): a conditional block here makes the
    // compiler wait for these loads before the atomics below.
    uint32_t en[RPS<KIND>], wtn[RPS<KIND>], xbn[RPS<KIND>], vmn = 0;
    load_rows<KIND>(A, rows, R, rn, en, wtn, xbn, vmn);
    if (KIND == CYTOM_PRD) {
#pragma unroll
      for (int u = 0; u < RPS<KIND>; ++u)
        if (ok[u] && c[u] != 0.f)  // pagerank-delta.cc:127-132, as exact fixed point
          atomicAdd(&winq[e[u] >> 16], (unsigned long long)__float2ll_rn(c[u] * A.qscale));
    } else if (KIND == KIND_SSSP_OCT) {
#pragma unroll
      for (int u = 0; u < RPS<KIND>; ++u) {  // sssp.cc:90-99 per query
        const uint32_t dq[8] = {v8[u][0].x, v8[u][0].y, v8[u][0].z, v8[u][0].w,
                                v8[u][1].x, v8[u][1].y, v8[u][1].z, v8[u][1].w};
#pragma unroll
        for (int q = 0; q < 8; ++q) {
          const uint32_t cand = sat_add(dq[q], wt[u]);
          if (ok[u] && cand != INF) atomicMin(&win[8 * (e[u] >> 16) + q], cand);
        }
      }
    } else if (KIND == KIND_SSSP_PAIR) {
#pragma unroll
      for (int u = 0; u < RPS<KIND>; ++u) {  // sssp.cc:90-99 for each query of the pair
        const uint32_t cx = sat_add(val2[u].x, wt[u]), cy = sat_add(val2[u].y, wt[u]);
        if (ok[u] && cx != INF) atomicMin(&win[2 * (e[u] >> 16)], cx);
        if (ok[u] && cy != INF) atomicMin(&win[2 * (e[u] >> 16) + 1], cy);
      }
    } else {
#pragma unroll
      for (int u = 0; u < RPS<KIND>; ++u) {
        uint32_t cand;
        if (KIND == CYTOM_BFS) cand = val[u] == INF ? INF : val[u] + 1;  // bfs.cc:83-91
        else if (KIND == CYTOM_CC) cand = val[u];                         // cc.cc:81-91
        else cand = sat_add(val[u], wt[u]);                               // sssp.cc:90-99
        if (ok[u] && cand != INF) atomicMin(&win[e[u] >> 16], cand);
      }
    }
    if (!more) break;
#pragma unroll
    for (int u = 0; u < RPS<KIND>; ++u) {
      e[u] = en[u];
      wt[u] = wtn[u];
      xb[u] = xbn[u];
    }
    vm = vmn;
    r0 = rn;
  }
}

// A frontier of at least V / DENSE_DIV active vertices gathers whole jobs.
constexpr uint64_t DENSE_DIV = 8;

template <int KIND>
__device__ inline void clear_window(uint32_t* win, unsigned long long* winq, uint32_t vpt) {
  if (KIND == CYTOM_PRD) {
    for (uint32_t i = threadIdx.x; i < vpt; i += GATHER_THREADS) winq[i] = 0ull;
  } else {
    const uint32_t n = (KIND == KIND_SSSP_OCT ? 8u : KIND == KIND_SSSP_PAIR ? 2u : 1u) * vpt;
    for (uint32_t i = threadIdx.x; i < n; i += GATHER_THREADS) win[i] = INF;
  }
}

// Folds a job's column window into the column's values (targets base ..
// base + vpt): int64 sums into dsum, minima into nxt, flagging the changed
// targets.  Other jobs of the same column flush concurrently, hence atomics.
template <int KIND>
__device__ inline void flush_window(const GatherArgs<KIND>& A, const uint32_t* win,
                                    const unsigned long long* winq, uint64_t base, uint32_t vpt) {
  for (uint32_t i = threadIdx.x; i < vpt; i += GATHER_THREADS) {
    if (KIND == CYTOM_PRD) {
      const unsigned long long q = winq[i];
      if (q != 0ull) {
        atomicAdd(reinterpret_cast<unsigned long long*>(&A.dsum[base + i]), q);
        A.changed[base + i] = 1;
      }
    } else if (KIND == KIND_SSSP_OCT || KIND == KIND_SSSP_PAIR) {
      constexpr uint32_t W = KIND == KIND_SSSP_OCT ? 8u : 2u;
      bool ch = false;
#pragma unroll
      for (uint32_t k = 0; k < W; ++k) {
        const uint32_t val = win[W * i + k];
        uint32_t* dst = &A.nxt[W * (base + i) + k];
        if (val != INF && val < *dst) ch |= val < atomicMin(dst, val);
      }
      if (ch) A.changed[base + i] = 1;
    } else {
      const uint32_t val = win[i];
      if (val != INF && val < A.nxt[base + i]) {
        const uint32_t old = atomicMin(&A.nxt[base + i], val);
        if (val < old) A.changed[base + i] = 1;
      }
    }
  }
}

template <int KIND>
__global__ void __launch_bounds__(GATHER_THREADS) k_gather(GatherArgs<KIND> A) {
  // The column window: u32 values for BFS/CC/SSSP, int64 fixed-point sums
  // for PageRank-delta (LDS integer atomics run an order of magnitude faster
  // than LDS float atomics on gfx950, and integer sums are exact).
  extern __shared__ unsigned long long lds_window[];
  uint32_t* win = reinterpret_cast<uint32_t*>(lds_window);
  unsigned long long* winq = lds_window;
  __shared__ uint64_t s_rows[MAX_JOB_ROWS];
  __shared__ uint32_t s_job;
  const uint32_t vpt = A.S.vpt, T = A.S.T;
  // Clear next iteration's row activity (written by the apply kernel).
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < 8 * T; i += gridDim.x * blockDim.x)
    A.gact_clear[i] = 0;
  if (A.it->count == 0) return;  // empty frontier (uniform)
  // A frontier whose live out-edges are below V/64 tests the active bitmap
  // before fetching a source value (measured: only when active slots are this
  // rare does skipping the value loads outweigh the extra bitmap load).
  const bool sparse = A.it->te * 64 < A.S.V;
  const uint32_t NB = A.nbands;
  const uint32_t home = blockIdx.x % NB;
  // A job's chunks are taken in passes of at most PASS chunks (one thread
  // each, at most MAX_JOB_ROWS rows); the window lives across the passes and
  // is flushed once per job.
  const uint32_t PASS = min((uint32_t)GATHER_THREADS, (MAX_JOB_ROWS * 64u) / A.CH);
  // A dense frontier (PageRank-delta's in almost every iteration) takes
  // every row of its jobs: the rows are copied from the prebuilt table, with
  // no per-pass chunk decode or scan.  Rows of inactive sources cost their
  // slot loads only (their source values are 0 / INF).
  const bool dense = A.it->count * DENSE_DIV >= A.S.V;
  unsigned long long scanned = 0;
  for (;;) {
    __syncthreads();  // every thread has read the previous s_job
    if (threadIdx.x == 0) {
      // own band's queue first, then the others in turn
      uint32_t j = UINT32_MAX;
      for (uint32_t k = 0; k < NB; ++k) {
        const uint32_t b = (home + k) % NB;
        const uint32_t lo = A.jbase[b * T], hi = A.jbase[(b + 1) * T];
        if (lo >= hi) continue;
        const uint32_t c = (uint32_t)atomicAdd(&A.it->band_cursor[b], 1ull);
        if (lo + c < hi) {
          j = lo + c;
          break;
        }
      }
      s_job = j;
    }
    __syncthreads();
    const uint32_t j = s_job;
    if (j == UINT32_MAX) break;  // uniform: every thread read s_job after the barrier
    const uint4 job = A.jobs[j];
    const uint32_t y = job.z;
    bool filled = false;  // uniform: the window holds this job's values
    if (dense) {
      const uint32_t rb0 = A.rowb[job.x], rb1 = A.rowb[job.y];
      filled = rb1 > rb0;
      if (filled) clear_window<KIND>(win, winq, vpt);
      for (uint32_t r = rb0; r < rb1; r += MAX_JOB_ROWS) {
        const uint32_t R = min(MAX_JOB_ROWS, rb1 - r);
        for (uint32_t i = threadIdx.x; i < R; i += GATHER_THREADS) {
          const uint64_t d = A.rows[r + i];
          s_rows[i] = d;
          scanned += (d & 63u) + 1;
        }
        __syncthreads();
        gather_rows<KIND>(A, win, winq, s_rows, R, false);
        __syncthreads();
      }
    } else {
      for (uint32_t cs = job.x; cs < job.y; cs += PASS) {
        const uint32_t nc = min(PASS, job.y - cs);
        // Row table of the pass's rows whose sources meet an active group of
        // their source row.  One scan packs the rows (bits 24+) and slots
        // (bits 0-23) before each chunk.
        uint32_t n = 0, x = 0, keep = 0, nrows = 0, nslots = 0;
        uint64_t off = 0;
        if (threadIdx.x < nc) {
          const uint64_t ch = A.chunks[cs + threadIdx.x];
          const uint32_t t = (uint32_t)(ch >> 32), s0 = (uint32_t)ch;
          x = t - y * T;
          const uint64_t m = group_mask(A.gact, x);
          if (m) {
            off = A.S.tile_off[t] + s0;
            n = min(A.CH, A.S.tile_cnt[t] - s0);
            const uint32_t nr = (n + 63) >> 6;
            // PageRank-delta frontiers are dense in almost every iteration: its
            // rows are taken whole once any group of their source row is active.
            const bool whole = KIND == CYTOM_PRD || m == ~0ull;
            for (uint32_t k = 0; k < nr; ++k) {
              if (whole || (m & A.grp[(off >> 6) + k])) {
                keep |= 1u << k;
                ++nrows;
                nslots += min(64u, n - 64 * k);
              }
            }
          }
        }
        const uint64_t packed = ((uint64_t)nrows << 24) | (uint64_t)nslots;
        uint64_t tot;
        const uint64_t ex = block_scan_excl<GATHER_THREADS, uint64_t>(packed, &tot);
        const uint32_t R = (uint32_t)(tot >> 24);
        if (R == 0) continue;  // uniform; the scan's barriers ordered s_rows reuse
        for (uint32_t k = 0, r = (uint32_t)(ex >> 24); keep; ++k, keep >>= 1)
          if (keep & 1u) s_rows[r++] = row_desc(off + 64 * k, x, min(64u, n - 64 * k));
        scanned += tot & 0xFFFFFFull;
        if (!filled) {
          clear_window<KIND>(win, winq, vpt);
          filled = true;
        }
        __syncthreads();
        gather_rows<KIND>(A, win, winq, s_rows, R, sparse);
        __syncthreads();
      }
    }
    if (filled) flush_window<KIND>(A, win, winq, (uint64_t)y * vpt, vpt);
  }
  // The sparse path's scans leave the block's count in every thread; the
  // dense path counts per thread.
  if (dense) scanned = block_sum<GATHER_THREADS, uint64_t>(scanned);
  if (threadIdx.x == 0 && scanned) atomicAdd(&A.it->scanned, scanned);
}

// ---------------------------------------------------------------- apply
// Persistent grid: each wavefront walks bitmap words (64 consecutive
// vertices, one per lane), builds the next active word with a ballot and
// keeps per-lane counters; one reduction and three atomics per block.
__device__ inline void apply_tail(unsigned long long te, unsigned long long cnt,
                                  unsigned long long upd, IterStats* nit) {
  te = block_sum<APPLY_THREADS, unsigned long long>(te);
  cnt = block_sum<APPLY_THREADS, unsigned long long>(cnt);
  upd = block_sum<APPLY_THREADS, unsigned long long>(upd);
  if (threadIdx.x == 0) {
    if (te) atomicAdd(&nit->te, te);
    if (cnt) atomicAdd(&nit->count, cnt);
    if (upd) atomicAdd(&(nit - 1)->updated, upd);
  }
}

// An iteration whose frontier (nit - 1) is empty gathered nothing and set no
// changed flag (hooks set a changed flag only together with an active bit),
// so its apply only has to clear the next active bitmap.  Uniform per grid.
__device__ inline bool empty_frontier(const IterStats* nit, uint64_t* __restrict__ act_next,
                                      uint64_t nwords) {
  if ((nit - 1)->count != 0) return false;
  for (uint64_t w = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; w < nwords;
       w += (uint64_t)gridDim.x * blockDim.x)
    act_next[w] = 0;
  return true;
}

// Apply kernels: each wavefront takes APPLY_WORDS bitmap words (64
// consecutive vertices each, one per lane) per step and issues all of their
// loads before using any, so a wave keeps several memory round trips in
// flight; the next active words are built with ballots.
constexpr int APPLY_WORDS = 4;

// bfs.cc:101-113, cc.cc:101-111, sssp.cc:101-113: copy next into current
// and activate on change.
template <int KIND>
__global__ void __launch_bounds__(APPLY_THREADS) k_apply_min(
    uint64_t V, uint32_t log2vpt, uint32_t* __restrict__ cur, const uint32_t* __restrict__ nxt,
    uint32_t* __restrict__ fval, uint8_t* __restrict__ changed,
    const uint64_t* __restrict__ act_cur, uint64_t* __restrict__ act_next,
    uint8_t* __restrict__ gact_next, const uint32_t* __restrict__ outdeg, IterStats* nit,
    const unsigned long long* refuse) {
  if (batch_refused(refuse)) return;
  const int lane = threadIdx.x & 63;
  const uint64_t nwords = V >> 6;
  if (empty_frontier(nit, act_next, nwords)) return;
  const uint64_t wave = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const uint64_t stride = (((uint64_t)gridDim.x * blockDim.x) >> 6) * APPLY_WORDS;
  unsigned long long te = 0, cnt = 0, upd = 0;
  for (uint64_t w0 = wave * APPLY_WORDS; w0 < nwords; w0 += stride) {
    uint64_t aw[APPLY_WORDS];
    uint8_t ch[APPLY_WORDS];
#pragma unroll
    for (int k = 0; k < APPLY_WORDS; ++k) {
      const uint64_t w = min(w0 + k, nwords - 1);
      aw[k] = act_cur[w];
      ch[k] = changed[(w << 6) | lane];
    }
    uint32_t n[APPLY_WORDS], c[APPLY_WORDS], od[APPLY_WORDS];
#pragma unroll
    for (int k = 0; k < APPLY_WORDS; ++k) {
      const uint64_t v = ((w0 + k) << 6) | lane;
      n[k] = c[k] = od[k] = 0;
      if (ch[k] && w0 + k < nwords) {
        n[k] = nxt[v];
        c[k] = cur[v];
        od[k] = outdeg[v];
      }
    }
#pragma unroll
    for (int k = 0; k < APPLY_WORDS; ++k) {
      const uint64_t w = w0 + k;
      if (w >= nwords) break;  // uniform
      const uint64_t v = (w << 6) | lane;
      const bool was = (aw[k] >> lane) & 1ull;
      bool a = false;
      if (ch[k]) {
        changed[v] = 0;
        ++upd;
        if (n[k] != c[k]) {
          cur[v] = n[k];
          fval[v] = n[k];
          a = true;
          te += od[k];
          ++cnt;
        }
      }
      if (was && !a) fval[v] = INF;
      const uint64_t word = __ballot(a);
      if (lane == 0) {
        act_next[w] = word;
        mark_groups(gact_next, word, v, log2vpt);
      }
    }
  }
  apply_tail(te, cnt, upd, nit);
}

// Apply of a pair run: a vertex is active when either distance dropped; its
// frontier pair carries the dropped components and INF for the others.
__global__ void __launch_bounds__(APPLY_THREADS) k_apply_pair(
    uint64_t V, uint32_t log2vpt, uint2* __restrict__ cur, const uint2* __restrict__ nxt,
    uint2* __restrict__ fval, uint8_t* __restrict__ changed,
    const uint64_t* __restrict__ act_cur, uint64_t* __restrict__ act_next,
    uint8_t* __restrict__ gact_next, const uint32_t* __restrict__ outdeg, IterStats* nit,
    const unsigned long long* refuse) {
  if (batch_refused(refuse)) return;
  const int lane = threadIdx.x & 63;
  const uint64_t nwords = V >> 6;
  if (empty_frontier(nit, act_next, nwords)) return;
  const uint64_t wave = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const uint64_t stride = ((uint64_t)gridDim.x * blockDim.x) >> 6;
  unsigned long long te = 0, cnt = 0, upd = 0;
  for (uint64_t w = wave; w < nwords; w += stride) {
    const uint64_t v = (w << 6) | lane;
    const bool was = (act_cur[w] >> lane) & 1ull;
    bool a = false;
    if (changed[v]) {
      changed[v] = 0;
      ++upd;
      const uint2 n = nxt[v], c = cur[v];
      if (n.x != c.x || n.y != c.y) {
        cur[v] = n;
        fval[v] = make_uint2(n.x != c.x ? n.x : INF, n.y != c.y ? n.y : INF);
        a = true;
        te += outdeg[v];
        ++cnt;
      }
    }
    if (was && !a) fval[v] = make_uint2(INF, INF);
    const uint64_t word = __ballot(a);
    if (lane == 0) {
      act_next[w] = word;
      mark_groups(gact_next, word, v, log2vpt);
    }
  }
  apply_tail(te, cnt, upd, nit);
}

// Apply of a W-query SSSP run (the pair rule, W wide): a vertex is active
// when any of its distances dropped; its frontier values carry the dropped
// distances and INF for the others.
template <uint32_t W>
__global__ void __launch_bounds__(APPLY_THREADS) k_apply_vec_min(
    uint64_t V, uint32_t log2vpt, uint32_t* __restrict__ cur, const uint32_t* __restrict__ nxt,
    uint32_t* __restrict__ fval, uint8_t* __restrict__ changed,
    const uint64_t* __restrict__ act_cur, uint64_t* __restrict__ act_next,
    uint8_t* __restrict__ gact_next, const uint32_t* __restrict__ outdeg, IterStats* nit) {
  const int lane = threadIdx.x & 63;
  const uint64_t nwords = V >> 6;
  if (empty_frontier(nit, act_next, nwords)) return;
  const uint64_t wave = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const uint64_t stride = ((uint64_t)gridDim.x * blockDim.x) >> 6;
  unsigned long long te = 0, cnt = 0, upd = 0;
  for (uint64_t w = wave; w < nwords; w += stride) {
    const uint64_t v = (w << 6) | lane;
    const bool was = (act_cur[w] >> lane) & 1ull;
    bool a = false;
    if (changed[v]) {
      changed[v] = 0;
      ++upd;
#pragma unroll
      for (uint32_t q = 0; q < W; ++q) {
        const uint64_t k = v * W + q;
        const uint32_t n = nxt[k], c = cur[k];
        const bool dq = n != c;
        if (dq) cur[k] = n;
        fval[k] = dq ? n : INF;
        a = a || dq;
      }
      if (a) {
        te += outdeg[v];
        ++cnt;
      }
    } else if (was) {
#pragma unroll
      for (uint32_t q = 0; q < W; ++q) fval[v * W + q] = INF;
    }
    const uint64_t word = __ballot(a);
    if (lane == 0) {
      act_next[w] = word;
      mark_groups(gact_next, word, v, log2vpt);
    }
  }
  apply_tail(te, cnt, upd, nit);
}

// pagerank-delta.cc:142-174 (FP32 state, double intermediates as in the
// reference: ALPHA and the threshold are doubles).  Loads are unconditional:
// most vertices carry a sum while PageRank-delta runs.
__global__ void __launch_bounds__(APPLY_THREADS) k_apply_prd(
    uint64_t V, uint32_t log2vpt, uint32_t* __restrict__ dctl, double thr, double qinv,
    float* __restrict__ rank,
    float* __restrict__ delta, long long* __restrict__ dsum, float* __restrict__ contrib,
    uint8_t* __restrict__ changed, const uint64_t* __restrict__ act_cur,
    uint64_t* __restrict__ act_next, uint8_t* __restrict__ gact_next,
    const uint32_t* __restrict__ outdeg, IterStats* nit, const unsigned long long* refuse) {
  if (batch_refused(refuse)) return;
  const int lane = threadIdx.x & 63;
  const uint64_t nwords = V >> 6;
  // The first-iteration rule holds for this launch when init set dctl[0];
  // every block reads it here and the last block to finish clears it.
  const bool first = dctl[0] != 0;
  __syncthreads();
  if (first) {
    if (threadIdx.x == 0) {
      __threadfence();
      if (atomicAdd(&dctl[1], 1u) == gridDim.x - 1) {
        dctl[1] = 0;
        dctl[0] = 0;
      }
    }
  }
  if (!first && empty_frontier(nit, act_next, nwords)) return;
  const uint64_t wave = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const uint64_t stride = (((uint64_t)gridDim.x * blockDim.x) >> 6) * APPLY_WORDS;
  unsigned long long te = 0, cnt = 0, upd = 0;
  for (uint64_t w0 = wave * APPLY_WORDS; w0 < nwords; w0 += stride) {
    uint64_t aw[APPLY_WORDS];
    uint8_t ch[APPLY_WORDS];
    long long ds[APPLY_WORDS];
    float rk[APPLY_WORDS];
    uint32_t od[APPLY_WORDS];
#pragma unroll
    for (int k = 0; k < APPLY_WORDS; ++k) {
      const uint64_t w = min(w0 + k, nwords - 1);
      const uint64_t v = (w << 6) | lane;
      aw[k] = act_cur[w];
      ch[k] = changed[v];
      ds[k] = dsum[v];
      rk[k] = rank[v];
      od[k] = outdeg[v];
    }
#pragma unroll
    for (int k = 0; k < APPLY_WORDS; ++k) {
      const uint64_t w = w0 + k;
      if (w >= nwords) break;  // uniform
      const uint64_t v = (w << 6) | lane;
      const bool was = (aw[k] >> lane) & 1ull;
      bool a = false;
      if (first || ch[k]) {
        if (ch[k]) {
          changed[v] = 0;
          ++upd;
        }
        const float s = (float)((double)ds[k] * qinv);
        float d = 0.f;
        float r = rk[k];
        bool applied = true;
        if (first) {
          d = static_cast<float>((1. - 0.85) + (0.85 * (double)s));
          r = r + d;
          d = (float)((double)d - 1.0);
        } else if (s == 0.f) {
          delta[v] = 0.f;
          applied = false;
        } else {
          d = static_cast<float>(0.85 * (double)s);
          r = r + d;
        }
        if (applied) {
          rank[v] = r;
          delta[v] = d;
          dsum[v] = 0;
          if ((double)fabsf(d) > thr * (double)r) {
            a = true;
            te += od[k];
            ++cnt;
            contrib[v] = od[k] ? d / (float)od[k] : 0.f;
          }
        }
      }
      if (was && !a) contrib[v] = 0.f;
      const uint64_t word = __ballot(a);
      if (lane == 0) {
        act_next[w] = word;
        mark_groups(gact_next, word, v, log2vpt);
      }
    }
  }
  apply_tail(te, cnt, upd, nit);
}

// ------------------------------------------------------ deletion tests
__global__ void k_reexec_min(int kind, const uint32_t* __restrict__ found, uint64_t n,
                             const unsigned long long* __restrict__ n_dev,
                             const uint32_t* __restrict__ cur, unsigned long long* flag) {
  n = min(n, (uint64_t)*n_dev);
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x) {
    const uint32_t s = found[3 * i], d = found[3 * i + 1], w = found[3 * i + 2];
    const uint32_t u = cur[s], v = cur[d];
    bool tight;
    if (kind == CYTOM_BFS) tight = u != INF && v == u + 1;
    else if (kind == CYTOM_SSSP) tight = u != INF && v == sat_add(u, w);
    else tight = u == v && v != d;
    if (tight) *flag = 1ull;
  }
}

// pagerank-delta.cc:190-206
__global__ void k_reexec_prd(const uint32_t* __restrict__ found, uint64_t n,
                             const unsigned long long* __restrict__ n_dev,
                             const float* __restrict__ rank, const uint32_t* __restrict__ outdeg,
                             double thr, unsigned long long* flag) {
  n = min(n, (uint64_t)*n_dev);
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x) {
    const uint32_t s = found[3 * i], d = found[3 * i + 1];
    const float dl = rank[s] / (float)outdeg[s];
    if ((double)dl > thr * (double)rank[d]) *flag = 1ull;
  }
}

// --------------------------------------------------------------- hooks
// edgeChanged(INSERT): bfs.cc:115-131, cc.cc:113-131, sssp.cc:115-133.
__global__ void k_hooks_min(int kind, const uint32_t* __restrict__ src,
                            const uint32_t* __restrict__ dst, const uint32_t* __restrict__ w,
                            const uint64_t* __restrict__ ins_slot, uint64_t n,
                            const uint32_t* __restrict__ pool, uint32_t* __restrict__ cur,
                            uint32_t* __restrict__ nxt, uint64_t* __restrict__ act,
                            uint8_t* __restrict__ changed, const unsigned long long* gate) {
  if (!gate_is(gate, 0)) return;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x) {
    if (pool[ins_slot[i]] == TOMB) continue;  // deleted within the batch
    const uint32_t s = src[i], d = dst[i];
    const uint32_t sv = nxt[s];
    if (sv == INF) continue;
    uint32_t cand;
    if (kind == CYTOM_BFS) cand = sv + 1;
    else if (kind == CYTOM_CC) cand = sv;
    else cand = sat_add(sv, w[i]);
    if (cand >= nxt[d]) continue;
    const uint32_t old = atomicMin(&nxt[d], cand);
    if (cand < old) {
      atomicMin(&cur[d], cand);
      atomicOr(reinterpret_cast<unsigned long long*>(&act[d >> 6]), 1ull << (d & 63));
      changed[d] = 1;
    }
  }
}

// pagerank-delta.cc:176-189 (INSERT) and 190-196 (DELETE side effect).
__global__ void k_hooks_prd(const uint32_t* __restrict__ src, const uint32_t* __restrict__ dst,
                            const uint64_t* __restrict__ ins_slot, uint64_t n,
                            const uint32_t* __restrict__ found, uint64_t n_found,
                            const unsigned long long* __restrict__ n_found_dev,
                            const uint32_t* __restrict__ pool, const float* __restrict__ rank,
                            const uint32_t* __restrict__ outdeg, long long* __restrict__ dsum,
                            uint64_t* __restrict__ act, uint8_t* __restrict__ changed, double thr,
                            float qscale, const unsigned long long* gate) {
  if (!gate_is(gate, 0)) return;
  if (n_found_dev) n_found = min(n_found, (uint64_t)*n_found_dev);
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n + n_found;
       i += (uint64_t)gridDim.x * blockDim.x) {
    if (i < n) {
      if (pool[ins_slot[i]] == TOMB) continue;
      const uint32_t s = src[i], d = dst[i];
      const float dl = rank[s] / (float)outdeg[s];
      atomicAdd(reinterpret_cast<unsigned long long*>(&dsum[d]),
                (unsigned long long)__float2ll_rn(dl * qscale));
      if ((double)dl > thr * (double)rank[d]) {
        atomicOr(reinterpret_cast<unsigned long long*>(&act[s >> 6]), 1ull << (s & 63));
        changed[s] = 1;
      }
    } else {
      const uint64_t k = i - n;
      const uint32_t s = found[3 * k], d = found[3 * k + 1];
      const float dl = rank[s] / (float)outdeg[s];
      atomicAdd(reinterpret_cast<unsigned long long*>(&dsum[d]),
                (unsigned long long)__float2ll_rn(-dl * qscale));
    }
  }
}

// The LDS column window of a kind: one u32 (BFS/CC/SSSP) or int64 sum
// (PageRank-delta) per target and query.
size_t gather_lds_bytes(int kind, uint32_t vpt) {
  const size_t per = kind == CYTOM_PRD ? 8 : 4;
  return (size_t)vpt * per * kind_width(kind);
}

// Persistent grid: as many blocks per CU as the LDS window allows (at most
// two), unless tuned.
template <int KIND>
uint32_t gather_grid(cytom_ctx* ctx, size_t lds) {
  if (ctx->gather_blocks) return ctx->gather_blocks;
  // The blocks one CU holds at once with this window (LDS and registers both
  // limit it), at most two: a persistent grid larger than what is resident
  // only adds blocks that start after the job queue has drained.  Asked once
  // per context and kind (the window is fixed by the context's tile width).
  uint32_t& per_cu = ctx->gather_per_cu[KIND];
  if (!per_cu) {
    int nb = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_gather<KIND>, GATHER_THREADS, lds) !=
            hipSuccess ||
        nb < 1)
      nb = 1;
    per_cu = (uint32_t)std::min(nb, 2);
  }
  return (uint32_t)ctx->num_cus * per_cu;
}

template <int KIND>
void launch_gather(cytom_ctx* ctx, Program& P, int it, hipEvent_t ev0, hipEvent_t ev1) {
  GatherArgs<KIND> A;
  A.jobs = static_cast<const uint4*>(ctx->g_jobs.p);
  A.jbase = static_cast<const uint32_t*>(ctx->g_jbase.p);
  A.nbands = ctx->n_bands;
  A.chunks = static_cast<const uint64_t*>(ctx->g_chunks.p);
  A.rowb = static_cast<const uint32_t*>(ctx->g_rowb.p);
  A.rows = static_cast<const uint64_t*>(ctx->g_rows.p);
  A.S = ctx->view();
  A.gact = P.gact[P.cur_act];
  A.act = P.act[P.cur_act];
  A.gact_clear = reinterpret_cast<uint64_t*>(P.gact[P.cur_act ^ 1]);
  A.grp = static_cast<const uint64_t*>(ctx->g_grp.p);
  A.fval = P.fval;
  A.nxt = P.nxt;
  A.contrib = P.contrib;
  A.dsum = P.dsum;
  A.qscale = ldexpf(1.f, ctx->prd_shift);
  A.changed = P.changed;
  A.it = P.d_iter + it;
  A.CH = ctx->chunk_slots;
  const size_t lds = gather_lds_bytes(KIND, ctx->vpt);
  if (ev0) {
    // Profiling: the events are part of the dispatch, so they time the
    // kernel itself (separate event records would add the dispatch gaps).
    hipExtLaunchKernelGGL(k_gather<KIND>, dim3(gather_grid<KIND>(ctx, lds)), dim3(GATHER_THREADS),
                          (uint32_t)lds, prog_stream(ctx, P), ev0, ev1, 0u, A);
  } else {
    k_gather<KIND><<<gather_grid<KIND>(ctx, lds), GATHER_THREADS, lds, prog_stream(ctx, P)>>>(A);
  }
}

}  // namespace

// Work lists of the current store, device-driven: every count stays on the
// device (the kernels read it there), buffers are sized by upper bounds
// (chunks <= pool_size / CH + n_tiles, jobs <= chunks / minimum job length +
// cells), so the lists can be rebuilt without a host round trip.
int prepare_jobs(cytom_ctx* ctx) {
  int rc;
  const uint64_t nt = ctx->n_tiles;
  const uint32_t NB = std::min(std::min(ctx->source_bands, ctx->T), MAX_SOURCE_BANDS);
  const uint64_t cells = (uint64_t)ctx->T * NB;
  const uint64_t max_chunks = ctx->pool_size / ctx->chunk_slots + nt + 1;
  const uint64_t jc_min = ctx->job_chunks ? ctx->job_chunks : (1u << 17) / ctx->chunk_slots;
  const uint64_t max_jobs = max_chunks / jc_min + cells + 1;
  // Rows: every chunk's slots in 64-slot rows, one partial row at most each.
  const uint64_t max_rows = ctx->pool_size / 64 + max_chunks;
  const uint64_t scan_n = std::max<uint64_t>(std::max<uint64_t>(nt, cells), max_chunks) + 1;
  // the job sort: keys and indices in, out and between passes, its histograms
  const uint64_t jw = (max_jobs * 4 + 255) / 256 * 256;
  const uint64_t rs_scan = (radix_hist_bytes(max_jobs) / 4) + 1;
  if ((rc = ensure(ctx, ctx->g_nch, nt * 4)) || (rc = ensure(ctx, ctx->g_cbase, (nt + 1) * 4)) ||
      (rc = ensure(ctx, ctx->g_colj, cells * 4)) ||
      (rc = ensure(ctx, ctx->g_jbase, (cells + 1) * 4)) ||
      (rc = ensure(ctx, ctx->g_nrow, max_chunks * 4)) ||
      (rc = ensure(ctx, ctx->g_rowb, (max_chunks + 1) * 4)) ||
      (rc = ensure(ctx, ctx->g_rows, max_rows * 8)) ||
      (rc = ensure(ctx, ctx->g_scan_tmp,
                   scan_tmp_bytes(std::max<uint64_t>(scan_n, rs_scan)) + 256)) ||
      (rc = ensure(ctx, ctx->g_jobs_raw, max_jobs * 16)) ||
      (rc = ensure(ctx, ctx->g_jsort, 6 * jw + radix_hist_bytes(max_jobs))) ||
      (rc = ensure(ctx, ctx->g_bands, 256 + (uint64_t)ctx->T * 8)) ||
      (rc = ensure(ctx, ctx->g_chunks, max_chunks * 8)) ||
      (rc = ensure(ctx, ctx->g_grp, (ctx->pool_size / 64 + 1) * 8)) ||
      (rc = ensure(ctx, ctx->g_jobs, max_jobs * 16)))
    return rc;
  ctx->n_chunks = max_chunks;
  ctx->n_cells = cells;
  ctx->n_bands = NB;
  ctx->n_jobs = max_jobs;
  return 0;
}

int build_jobs(cytom_ctx* ctx) {
  int rc;
  if ((rc = prepare_jobs(ctx))) return rc;
  enqueue_jobs(ctx);
  CY_HIP(ctx, hipGetLastError());
  return 0;
}

void enqueue_row_groups(cytom_ctx* ctx) {
  k_row_groups<<<grid_for(ctx->n_chunks * 64, 256), 256, 0, ctx->stream>>>(
      static_cast<const uint64_t*>(ctx->g_chunks.p),
      static_cast<const uint32_t*>(ctx->g_cbase.p) + ctx->n_tiles, ctx->view(), ctx->chunk_slots,
      static_cast<uint64_t*>(ctx->g_grp.p));
}

void enqueue_jobs(cytom_ctx* ctx, bool with_groups) {
  hipStream_t st = ctx->stream;
  const uint64_t nt = ctx->n_tiles;
  const uint32_t NB = ctx->n_bands;
  const uint64_t cells = ctx->n_cells;
  auto* nch = static_cast<uint32_t*>(ctx->g_nch.p);
  auto* cbase = static_cast<uint32_t*>(ctx->g_cbase.p);
  auto* colj = static_cast<uint32_t*>(ctx->g_colj.p);
  auto* jbase = static_cast<uint32_t*>(ctx->g_jbase.p);
  auto* xb = static_cast<uint32_t*>(ctx->g_bands.p);
  uint32_t* jc = xb + 32;  // the job length chosen on the device
  auto* rs = reinterpret_cast<unsigned long long*>(static_cast<char*>(ctx->g_bands.p) + 256);
  auto* chunks = static_cast<uint64_t*>(ctx->g_chunks.p);
  if (NB > 1) {
    k_row_slots<<<grid_for((uint64_t)ctx->T * 64, 256), 256, 0, st>>>(ctx->tile_cnt, ctx->T, rs);
    k_band_bounds<<<1, 64, 0, st>>>(rs, ctx->T, NB, xb);
  } else {
    const FillOp fills[] = {{xb, 1, 0, 0}, {xb + 1, 1, ctx->T, 0}};  // one band: all source rows
    fill_many(fills, 2, st);
  }
  k_nch<<<grid_for(nt, 256), 256, 0, st>>>(ctx->tile_cnt, nt, ctx->chunk_slots, nch);
  scan_exclusive_u32(nch, cbase, nt, ctx->g_scan_tmp.p, st);
  k_emit_chunks<<<grid_for(nt * 64, 256), 256, 0, st>>>(nch, cbase, nt, ctx->chunk_slots, chunks);
  // Source groups of every 64-slot row, recomputed after each batch (the
  // batch moved, added and tombstoned slots).
  if (with_groups) enqueue_row_groups(ctx);
  // The row table of the dense gathers.
  auto* nrow = static_cast<uint32_t*>(ctx->g_nrow.p);
  auto* rowb = static_cast<uint32_t*>(ctx->g_rowb.p);
  k_chunk_rows<<<grid_for(ctx->n_chunks, 256), 256, 0, st>>>(chunks, cbase + nt, ctx->tile_cnt,
                                                            ctx->chunk_slots, nrow);
  scan_exclusive_u32_devn(nrow, rowb, cbase + nt, ctx->n_chunks, ctx->g_scan_tmp.p, st);
  k_emit_rows<<<grid_for(ctx->n_chunks, 256), 256, 0, st>>>(
      chunks, cbase + nt, ctx->view(), ctx->chunk_slots, rowb, static_cast<uint64_t*>(ctx->g_rows.p));
  // Jobs of job_chunks chunks, or sized from the store: at least 2^17 slots,
  // and about eight jobs per CU when the store is large (fewer column flushes).
  k_job_length<<<1, 64, 0, st>>>(cbase + nt, ctx->job_chunks, ctx->chunk_slots,
                                 (uint32_t)ctx->num_cus, jc);
  k_col_jobs<<<grid_for(cells, 256), 256, 0, st>>>(cbase, xb, ctx->T, NB, jc, colj);
  scan_exclusive_u32(colj, jbase, cells, ctx->g_scan_tmp.p, st);
  auto* raw = static_cast<uint4*>(ctx->g_jobs_raw.p);
  k_emit_jobs<<<grid_for(cells, 256), 256, 0, st>>>(cbase, xb, colj, jbase, ctx->T, NB, jc, raw);
  // Largest first within each band.
  const uint64_t mj = ctx->n_jobs;
  uint32_t band_bits = 0;
  while ((1u << band_bits) < NB) ++band_bits;
  const uint64_t jw = (mj * 4 + 255) / 256 * 256;
  char* js = static_cast<char*>(ctx->g_jsort.p);
  auto* key = reinterpret_cast<uint32_t*>(js);
  auto* idx = reinterpret_cast<uint32_t*>(js + jw);
  auto* key_out = reinterpret_cast<uint32_t*>(js + 2 * jw);
  auto* idx_out = reinterpret_cast<uint32_t*>(js + 3 * jw);
  k_job_keys<<<grid_for(mj, 256), 256, 0, st>>>(raw, jbase, ctx->T, NB, rowb, key, idx);
  radix_sort_pairs_devn(key, idx, key_out, idx_out, reinterpret_cast<uint32_t*>(js + 4 * jw),
                        reinterpret_cast<uint32_t*>(js + 5 * jw), jbase + cells, mj,
                        16 + band_bits,
                        js + 6 * jw, ctx->g_scan_tmp.p, st);
  k_permute_jobs<<<grid_for(mj, 256), 256, 0, st>>>(raw, idx_out, jbase + cells,
                                                    static_cast<uint4*>(ctx->g_jobs.p));
  ctx->jobs_dirty = false;
}

int64_t jobs_chunk_count(cytom_ctx* ctx) {
  uint32_t n = 0;
  CY_HIP(ctx, hipMemcpyAsync(&n, static_cast<uint32_t*>(ctx->g_cbase.p) + ctx->n_tiles, 4,
                             hipMemcpyDeviceToHost, ctx->stream));
  CY_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return n;
}

int program_alloc(cytom_ctx* ctx, Program& P) {
  const uint64_t V = ctx->V;
  auto alloc = [&](void** p, size_t b) -> bool { return hipMalloc(p, b) == hipSuccess; };
  bool ok = true;
  const size_t W = kind_width(P.kind);  // queries per vertex
  if (prd_like(P.kind)) {
    ok = ok && alloc((void**)&P.rank, V * 4 * W) && alloc((void**)&P.delta, V * 4 * W) &&
         alloc((void**)&P.dsum, V * 8 * W) && alloc((void**)&P.contrib, V * 4 * W);
  } else {
    const size_t vb = 4 * W;  // one u32 per vertex and query
    ok = ok && alloc((void**)&P.cur, V * vb) && alloc((void**)&P.nxt, V * vb) &&
         alloc((void**)&P.fval, V * vb);
  }
  for (int i = 0; i < 2 && ok; ++i)
    ok = alloc((void**)&P.act[i], (V / 64) * 8) && alloc((void**)&P.gact[i], (size_t)ctx->T * 64);
  ok = ok && alloc((void**)&P.changed, V);
  const size_t nit = sizeof(IterStats) * (ctx->max_iter + 2);
  ok = ok && alloc((void**)&P.d_iter, nit) && alloc((void**)&P.dctl, 64) &&
       hipHostMalloc((void**)&P.h_iter, nit, hipHostMallocDefault) == hipSuccess;
  ok = ok && hipMemsetAsync(P.dctl, 0, 64, ctx->stream) == hipSuccess &&
       program_events(ctx, P) == 0;
  if (!ok) {
    program_free(P);
    return set_err(ctx, CYTOM_ENOMEM, "program state allocation failed");
  }
  return 0;
}

void program_free(Program& P) {
  void* ps[] = {P.cur,     P.nxt,     P.fval,      P.rank,      P.delta,   P.dsum,
                P.contrib, P.act[0],  P.act[1],    P.gact[0], P.gact[1], P.changed,
                P.teleport, P.qseed};
  for (void* p : ps)
    if (p) (void)hipFree(p);
  if (P.d_iter) (void)hipFree(P.d_iter);
  if (P.dctl) (void)hipFree(P.dctl);
  if (P.h_iter) (void)hipHostFree(P.h_iter);
  for (auto& e : P.gev) (void)hipEventDestroy(e);
  for (auto& e : P.kev) (void)hipEventDestroy(e);
  P.d_iter = nullptr;
  P.dctl = nullptr;
  P.h_iter = nullptr;
  P.gev.clear();
  P.kev.clear();
  P.cur = P.nxt = P.fval = nullptr;
  P.rank = P.delta = P.contrib = nullptr;
  P.dsum = nullptr;
  P.act[0] = P.act[1] = nullptr;
  P.gact[0] = P.gact[1] = nullptr;
  P.changed = nullptr;
  P.teleport = nullptr;
  P.n_teleport = 0;
  P.qseed = nullptr;
  P.n_qseed = 0;
}

int program_init(cytom_ctx* ctx, Program& P, const unsigned long long* gate) {
  hipStream_t st = ctx->stream;
  // Init writes the current buffers (act[cur_act]), so the host's buffer
  // parity stays valid whether or not a gated init runs.
  uint64_t* act = P.act[P.cur_act];
  if (!gate) P.fixpoint = false;
  if (prd_like(P.kind)) {
    const int ppr = P.kind == CYTOM_PPR ? 1 : 0;
    k_init_prd<<<grid_for(ctx->V, 256), 256, 0, st>>>(ctx->V, ppr, P.rank, P.delta, P.dsum, act,
                                                      P.changed, P.dctl, gate);
    if (ppr && P.n_teleport)
      k_init_ppr_set<<<grid_for(P.n_teleport, 256), 256, 0, st>>>(P.teleport, P.n_teleport,
                                                                  P.rank, P.delta, act, gate);
  } else if (P.kind == KIND_SSSP_OCT) {
    k_init_vec<8><<<grid_for(ctx->V, 256), 256, 0, st>>>(ctx->V, P.cur, P.nxt, act, P.changed);
    if (P.n_qseed)
      k_seed_vec<8><<<grid_for(P.n_qseed, 256), 256, 0, st>>>(P.qseed, P.n_qseed, P.cur, P.nxt,
                                                            act);
  } else if (P.kind == KIND_SSSP_PAIR) {
    k_init_pair<<<grid_for(ctx->V, 256), 256, 0, st>>>(
        ctx->V, P.source, P.source2, reinterpret_cast<uint2*>(P.cur),
        reinterpret_cast<uint2*>(P.nxt), act, P.changed, gate);
  } else {
    k_init_min<<<grid_for(ctx->V, 256), 256, 0, st>>>(P.kind, ctx->V, P.source, P.cur, P.nxt,
                                                      act, P.changed, gate);
  }
  CY_HIP(ctx, hipGetLastError());
  return 0;
}

int program_set_sources(cytom_ctx* ctx, Program& P, const uint32_t* sources, uint32_t n) {
  if (!sources || n == 0) return set_err(ctx, CYTOM_EINVAL, "no source vertices");
  for (uint32_t i = 0; i < n; ++i)
    if (sources[i] >= ctx->V) return set_err(ctx, CYTOM_EINVAL, "source vertex out of range");
  if (P.kind == CYTOM_BFS || P.kind == CYTOM_SSSP) {
    if (n != 1) return set_err(ctx, CYTOM_EINVAL, "BFS and SSSP take one source");
    P.source = sources[0];
  } else if (P.kind == CYTOM_PPR) {
    std::vector<uint32_t> set(sources, sources + n);
    std::sort(set.begin(), set.end());
    set.erase(std::unique(set.begin(), set.end()), set.end());
    uint32_t* d = nullptr;
    CY_HIP(ctx, hipStreamSynchronize(ctx->stream));
    CY_HIP(ctx, hipMalloc(&d, set.size() * 4));
    if (hipMemcpy(d, set.data(), set.size() * 4, hipMemcpyHostToDevice) != hipSuccess) {
      (void)hipFree(d);
      return set_err(ctx, CYTOM_EDEVICE, "teleport set copy failed");
    }
    if (P.teleport) (void)hipFree(P.teleport);
    P.teleport = d;
    P.n_teleport = (uint32_t)set.size();
  } else {
    return set_err(ctx, CYTOM_EINVAL, "CC and PageRank-delta take no sources");
  }
  return program_init(ctx, P);
}

void launch_reexec_check(cytom_ctx* ctx, Program& P, const uint32_t* found, uint64_t n_max,
                         const unsigned long long* n_dev, unsigned long long* flag) {
  if (prd_like(P.kind))
    k_reexec_prd<<<grid_for(n_max, 256), 256, 0, ctx->stream>>>(found, n_max, n_dev, P.rank,
                                                                ctx->outdeg, ctx->delta, flag);
  else
    k_reexec_min<<<grid_for(n_max, 256), 256, 0, ctx->stream>>>(P.kind, found, n_max, n_dev,
                                                                P.cur, flag);
}

void launch_hooks(cytom_ctx* ctx, Program& P, const uint32_t* d_src, const uint32_t* d_dst,
                  const uint32_t* d_w, const uint64_t* ins_slot, uint64_t n_ins,
                  const uint32_t* found, uint64_t n_found_max,
                  const unsigned long long* n_found_dev, const unsigned long long* gate) {
  hipStream_t st = ctx->stream;
  if (prd_like(P.kind)) {
    if (n_ins + n_found_max)
      k_hooks_prd<<<grid_for(n_ins + n_found_max, 256), 256, 0, st>>>(
          d_src, d_dst, ins_slot, n_ins, found, n_found_max, n_found_dev, ctx->pool, P.rank,
          ctx->outdeg, P.dsum, P.act[P.cur_act], P.changed, ctx->delta,
          ldexpf(1.f, ctx->prd_shift), gate);
  } else if (n_ins) {
    k_hooks_min<<<grid_for(n_ins, 256), 256, 0, st>>>(P.kind, d_src, d_dst, d_w, ins_slot, n_ins,
                                                      ctx->pool, P.cur, P.nxt, P.act[P.cur_act],
                                                      P.changed, gate);
  }
}

namespace {

inline uint32_t apply_blocks_for(const cytom_ctx* ctx) {
  // One 1024-thread block per CU walks the bitmap words: measured faster than
  // more, smaller blocks (fewer block tails and end-of-block reductions).
  const uint64_t words = ctx->V >> 6;
  return (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>((uint64_t)ctx->num_cus,
                                                             (words + 3) / 4));
}

}  // namespace

void run_iters(cytom_ctx* ctx, Program& P, uint32_t lo, uint32_t n) {
  hipStream_t st = prog_stream(ctx, P);
  const uint32_t ab = apply_blocks_for(ctx);
  const bool prof = ctx->profile && P.kev.size() >= 2 * (size_t)ctx->max_iter;
  for (uint32_t i = lo; i < lo + n; ++i) {
    const bool timed = prof && i % ctx->profile_every == P.runs % ctx->profile_every;
    hipEvent_t e0 = timed ? P.kev[2 * i] : nullptr, e1 = timed ? P.kev[2 * i + 1] : nullptr;
    switch (P.kind) {
      case CYTOM_BFS: launch_gather<CYTOM_BFS>(ctx, P, i, e0, e1); break;
      case CYTOM_CC: launch_gather<CYTOM_CC>(ctx, P, i, e0, e1); break;
      case CYTOM_SSSP: launch_gather<CYTOM_SSSP>(ctx, P, i, e0, e1); break;
      case KIND_SSSP_PAIR: launch_gather<KIND_SSSP_PAIR>(ctx, P, i, e0, e1); break;
      case KIND_SSSP_OCT: launch_gather<KIND_SSSP_OCT>(ctx, P, i, e0, e1); break;
      default: launch_gather<CYTOM_PRD>(ctx, P, i, e0, e1); break;
    }
    const int nx = P.cur_act ^ 1;
    IterStats* nit = P.d_iter + i + 1;
    if (prd_like(P.kind)) {
      k_apply_prd<<<ab, APPLY_THREADS, 0, st>>>(
          ctx->V, ctx->log2vpt, P.dctl, ctx->delta, ldexp(1.0, -ctx->prd_shift), P.rank, P.delta,
          P.dsum, P.contrib, P.changed, P.act[P.cur_act], P.act[nx], P.gact[nx], ctx->outdeg, nit, P.refuse);
    } else if (P.kind == CYTOM_BFS) {
      k_apply_min<CYTOM_BFS><<<ab, APPLY_THREADS, 0, st>>>(
          ctx->V, ctx->log2vpt, P.cur, P.nxt, P.fval, P.changed, P.act[P.cur_act], P.act[nx],
          P.gact[nx], ctx->outdeg, nit, P.refuse);
    } else if (P.kind == CYTOM_CC) {
      k_apply_min<CYTOM_CC><<<ab, APPLY_THREADS, 0, st>>>(
          ctx->V, ctx->log2vpt, P.cur, P.nxt, P.fval, P.changed, P.act[P.cur_act], P.act[nx],
          P.gact[nx], ctx->outdeg, nit, P.refuse);
    } else if (P.kind == KIND_SSSP_OCT) {
      k_apply_vec_min<8><<<ab, APPLY_THREADS, 0, st>>>(ctx->V, ctx->log2vpt, P.cur, P.nxt, P.fval,
                                                       P.changed, P.act[P.cur_act], P.act[nx],
                                                       P.gact[nx], ctx->outdeg, nit);
    } else if (P.kind == KIND_SSSP_PAIR) {
      k_apply_pair<<<ab, APPLY_THREADS, 0, st>>>(
          ctx->V, ctx->log2vpt, reinterpret_cast<uint2*>(P.cur),
          reinterpret_cast<const uint2*>(P.nxt), reinterpret_cast<uint2*>(P.fval), P.changed,
          P.act[P.cur_act], P.act[nx], P.gact[nx], ctx->outdeg, nit, P.refuse);
    } else {
      k_apply_min<CYTOM_SSSP><<<ab, APPLY_THREADS, 0, st>>>(
          ctx->V, ctx->log2vpt, P.cur, P.nxt, P.fval, P.changed, P.act[P.cur_act], P.act[nx],
          P.gact[nx], ctx->outdeg, nit, P.refuse);
    }
    P.cur_act = nx;
  }
}

void run_open(Program& P) {
  ++P.runs;
  P.launched = 0;
  P.groups = 0;
  P.checked = 0;
  P.conv = -1;
}

namespace {

// Iterations of the next group of a run: the first is as long as the
// program's previous run plus 1/16 (PageRank-delta runs take a similar number
// of iterations batch after batch, and launches past convergence are no-ops:
// the gather returns at once and the apply only clears the next bitmap),
// later ones 4, 8, then 16.
uint32_t group_length(const cytom_ctx* ctx, const Program& P) {
  uint32_t n;
  if (P.groups == 0)
    n = std::max<uint32_t>(P.last_iters + std::max<uint32_t>(P.last_iters / 16, 2), 2);
  else
    n = std::min<uint32_t>(4u << std::min<uint32_t>(P.groups - 1, 2), 16);
  // With step graphs the first group has an even length, so the buffer
  // parity it leaves (every later group is even too) repeats from batch to
  // batch, and so does the recorded graph of each length.
  if (ctx->step_graphs) n = (n + 1) & ~1u;
  return std::min(n, ctx->max_iter - P.launched);
}

// Upper bound of the groups of one run (first group of at least one
// iteration, then 4, 8 and 16 at a time).
size_t max_groups(const cytom_ctx* ctx) { return ctx->max_iter / 16 + 6; }

}  // namespace

void run_group(cytom_ctx* ctx, Program& P) {
  hipStream_t st = prog_stream(ctx, P);
  const uint32_t n = group_length(ctx, P);
  const uint32_t g = P.groups++;
  if (g == 0) {
    const FillOp fills[] = {{P.d_iter, sizeof(IterStats) / 8 * (uint64_t)(ctx->max_iter + 2), 0, 1},
                            {P.gact[P.cur_act], (uint64_t)ctx->T * 8, 0, 1}};
    fill_many(fills, 2, st);
  }
  if (!ctx->capturing) (void)hipEventRecord(P.gev[2 * g], st);
  if (g == 0)
    k_frontier<<<apply_blocks_for(ctx), APPLY_THREADS, 0, st>>>(
        P.kind, ctx->V, ctx->log2vpt, P.act[P.cur_act], ctx->outdeg, P.gact[P.cur_act], P.delta,
        P.contrib, P.cur, P.fval, P.d_iter, P.refuse);
  run_iters(ctx, P, P.launched, n);
  if (!ctx->capturing) (void)hipEventRecord(P.gev[2 * g + 1], st);
  // counters of the frontier (first group) and of every iteration launched
  const uint32_t lo = g == 0 ? 0 : P.launched + 1;
  (void)hipMemcpyAsync(&P.h_iter[lo], &P.d_iter[lo], sizeof(IterStats) * (P.launched + n + 1 - lo),
                       hipMemcpyDeviceToHost, st);
  P.launched += n;
}

uint32_t run_group_length(const cytom_ctx* ctx, const Program& P) {
  return group_length(ctx, P);
}

void run_group_replayed(const cytom_ctx* ctx, Program& P) {
  const uint32_t n = group_length(ctx, P);
  ++P.groups;
  P.cur_act ^= (int)(n & 1u);  // run_iters flips the parity once per iteration
  P.launched += n;
}

bool run_needs_more(const cytom_ctx* ctx, Program& P) {
  for (uint32_t k = P.checked + 1; P.conv < 0 && k <= P.launched; ++k)
    if (P.h_iter[k].count == 0) P.conv = (int)k;
  P.checked = P.launched;
  return P.conv < 0 && P.launched < ctx->max_iter;
}

int run_finish(cytom_ctx* ctx, Program& P, cytom_run_stats* out) {
  CY_HIP(ctx, hipGetLastError());
  const bool converged = P.conv >= 0;
  const uint32_t it = converged ? (uint32_t)P.conv : P.launched;
  P.last_iters = it;
  P.fixpoint = converged;
  const uint32_t maxit = ctx->max_iter;
  if (ctx->profile && P.kev.size() >= 2 * (size_t)maxit && P.kind <= CYTOM_PPR) {
    // Algorithmic bytes of one gather launch: 8 B per traversed packed edge
    // (+4 B weight for SSSP), the frontier's vertex state (4 B, or the 12 B
    // PageRank-delta vertex), a read-modify-write of every updated target and
    // one pass over the active bitmap.
    const double eb = P.kind == CYTOM_SSSP ? 12.0 : 8.0;
    const double sv = prd_like(P.kind) ? 12.0 : 4.0;
    for (uint32_t i = P.runs % ctx->profile_every; i < it; i += ctx->profile_every) {
      float ms = 0.f;
      CY_HIP(ctx, hipEventElapsedTime(&ms, P.kev[2 * i], P.kev[2 * i + 1]));
      const IterStats& s = P.h_iter[i];
      ctx->gather_ms[P.kind] += ms;
      ctx->gather_launches[P.kind] += 1;
      ctx->gather_alg_bytes[P.kind] += (double)s.te * eb + (double)s.count * sv +
                                       (double)s.updated * 2.0 * sv + (double)ctx->V / 8.0;
    }
  }
  if (out) {
    unsigned long long te = 0, act = 0, sc = 0;
    for (uint32_t i = 0; i < it; ++i) {
      te += P.h_iter[i].te;
      act += P.h_iter[i].count;
      sc += P.h_iter[i].scanned;
    }
    // device time of the run: its groups (other programs' groups may lie
    // between them in a step)
    double ms = 0.0;
    for (uint32_t g = 0; g < P.groups; ++g) {
      float gm = 0.f;
      CY_HIP(ctx, hipEventElapsedTime(&gm, P.gev[2 * g], P.gev[2 * g + 1]));
      ms += gm;
    }
    out->iterations = it;
    out->converged = converged ? 1 : 0;
    out->reexecuted = (uint32_t)P.last_reexec;
    out->pad = 0;
    out->traversed_edges = te;
    out->active_vertices = act;
    out->slots_scanned = sc;
    out->t_compute_ms = ms;
  }
  return 0;
}

int run_program(cytom_ctx* ctx, Program& P, cytom_run_stats* out) {
  int rc;
  if (ctx->jobs_dirty && (rc = build_jobs(ctx))) return rc;
  if ((rc = program_events(ctx, P))) return rc;
  run_open(P);
  do {
    run_group(ctx, P);
    CY_HIP(ctx, hipStreamSynchronize(ctx->stream));
  } while (run_needs_more(ctx, P));
  return run_finish(ctx, P, out);
}

// Events of a program's runs: the iteration groups' bounds, and with
// profiling one pair around every gather launch.
int program_events(cytom_ctx* ctx, Program& P) {
  while (P.gev.size() < 2 * max_groups(ctx)) {
    hipEvent_t e;
    CY_HIP(ctx, hipEventCreate(&e));
    P.gev.push_back(e);
  }
  if (!ctx->profile) return 0;
  while (P.kev.size() < 2 * (size_t)ctx->max_iter) {
    hipEvent_t e;
    CY_HIP(ctx, hipEventCreate(&e));
    P.kev.push_back(e);
  }
  return 0;
}

int configure_gather(cytom_ctx* ctx) {
  // The limits are per kernel, not per context: each is set to the largest
  // window any context launches it with (tiles of up to 16,384 vertices; the
  // multi-query kinds run on 4,096-vertex query tiles), so contexts of
  // different tile widths coexist.
  struct {
    const void* fn;
    int kind;
    uint32_t vpt;
  } ks[] = {{(const void*)k_gather<CYTOM_BFS>, CYTOM_BFS, 16384},
            {(const void*)k_gather<CYTOM_CC>, CYTOM_CC, 16384},
            {(const void*)k_gather<CYTOM_SSSP>, CYTOM_SSSP, 16384},
            {(const void*)k_gather<CYTOM_PRD>, CYTOM_PRD, 16384},
            {(const void*)k_gather<KIND_SSSP_PAIR>, KIND_SSSP_PAIR, 16384},
            {(const void*)k_gather<KIND_SSSP_OCT>, KIND_SSSP_OCT, QUERY_TILE_VPT}};
  for (const auto& k : ks)
    CY_HIP(ctx, hipFuncSetAttribute(k.fn, hipFuncAttributeMaxDynamicSharedMemorySize,
                                    (int)gather_lds_bytes(k.kind, k.vpt)));
  ctx->gather_blocks = 0;
  return 0;
}

}  // namespace cytom
