// Device-resident tiled edge store: batched insertion and deletion.
//
// Replaces, for one update batch, the provider routing
// (i-edge-provider.h:81-155), MetaTileManager::insertEdge / the deletion
// loop (meta-tile-manager.h:112-231), TileManager::insertEdge /
// deleteRandomEdge / compactStorage (tile-manager.h:136-291) and
// TileStore::doGlobalCompaction (tile-store.h:144-172).
//
// Layout in HBM: a tile index (offset, capacity, used slots, live edges per
// tile) and one pool of 4-byte edges (tile-local src in the low half,
// tile-local dst in the high half) plus a parallel u32 weight pool.  Each
// tile owns a contiguous block of the pool with slack at its end.
//
// Insertion: the batch is radix-sorted by tile key, runs of equal keys are
// the per-tile segments; tiles whose slack is too small are relocated to a
// fresh block (dropping tombstones on the way); then every segment is
// scattered into its tile's slack.  Deletion: the requested edges go into
// an open-addressing table keyed by (src, dst); the tiles that hold a
// request are scanned once and each live slot claims one matching request
// (a CAS on the request), the claimed slot becomes a tombstone.  Tiles left
// with many tombstones are compacted in place.
#include <algorithm>

#include "devutil.h"
#include "internal.h"

namespace cytom {
namespace {

constexpr uint64_t EMPTY64 = ~0ull;
constexpr uint32_t NONE32 = 0xFFFFFFFFu;

// Control words of a batch in d_misc (all on the device: the batch runs
// without host round trips and the host reads them once at the end).
enum Misc : int {
  M_POOL_TOP = 0,
  M_ERR = MISC_ERR,  // bit 0: invalid insertion, bit 1: invalid deletion
  M_NEED = 2,      // slots the relocated tiles need
  M_NRELOC = 3,
  M_RELOC_CUR = 4,
  M_NFOUND = 5,    // deletions that matched a live edge
  M_NCOMPACT = 6,
  M_RS_CUT = 7,    // first selected tile (by rank) left for a later sort-compaction
  M_OVERFLOW = MISC_OVERFLOW,  // the relocations do not fit the pool: the batch is not applied
  M_RS_NPROC = 9,  // tiles sort-compacted by this batch
  M_RS_NSLOTS = 10,
  M_POOL_SIZE = 11,
};

// A batch with an invalid update, or one whose relocations overflow the
// pool, modifies nothing: every kernel that writes the store, the degrees or
// a program checks this first (uniform per launch).
__device__ inline bool batch_dead(const unsigned long long* misc) { return batch_refused(misc); }

__host__ __device__ inline uint32_t grow_cap(uint64_t n) {
  uint64_t c = n + n / 2 + 64;
  c = (c + 63) & ~63ull;
  return c > 0xFFFFFFC0ull ? 0xFFFFFFC0u : (uint32_t)c;
}

__device__ inline uint32_t tile_of(uint32_t s, uint32_t d, uint32_t lv, uint32_t T) {
  return (d >> lv) * T + (s >> lv);
}

// ---- validation and keys
// Sort key of an insert: its tile id in the high bits and the top `sb` bits
// of its tile-local source below, so every tile's segment of the batch lands
// ordered by source (neighbouring gather lanes then share source lines).
__global__ void k_keys(const uint32_t* __restrict__ src, const uint32_t* __restrict__ dst,
                       const uint32_t* __restrict__ w, uint64_t n, uint64_t V, uint32_t lv,
                       uint32_t T, uint32_t sb, int need_w, uint32_t* __restrict__ keys,
                       uint32_t* __restrict__ vals, unsigned long long* misc) {
  const uint32_t smask = (1u << lv) - 1u;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x) {
    const uint32_t s = src[i], d = dst[i];
    const bool bad = s >= V || d >= V || (need_w && w[i] == 0u);
    if (bad) atomicOr(&misc[M_ERR], 1ull);
    keys[i] = bad ? 0u : ((tile_of(s, d, lv, T) << sb) | ((s & smask) >> (lv - sb)));
    vals[i] = (uint32_t)i;
  }
}

__global__ void k_del_validate(const uint32_t* __restrict__ src, const uint32_t* __restrict__ dst,
                               uint64_t n, uint64_t V, unsigned long long* misc) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x) {
    if (src[i] >= V || dst[i] >= V) atomicOr(&misc[M_ERR], 2ull);
  }
}

// ---- segments of the sorted batch
__global__ void k_heads(const uint32_t* __restrict__ keys, uint64_t n, uint32_t sb,
                        uint32_t* __restrict__ f) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x)
    f[i] = (i == 0 || (keys[i] >> sb) != (keys[i - 1] >> sb)) ? 1u : 0u;
}

__global__ void k_seg_emit(const uint32_t* __restrict__ keys, const uint32_t* __restrict__ f,
                           const uint32_t* __restrict__ sx, uint64_t n, uint32_t sb,
                           uint32_t* __restrict__ seg_start, uint32_t* __restrict__ seg_tile) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x) {
    if (f[i]) {
      seg_start[sx[i]] = (uint32_t)i;
      seg_tile[sx[i]] = keys[i] >> sb;
    }
    if (i == n - 1) seg_start[sx[n]] = (uint32_t)n;
  }
}

// ---- slack planning: which tiles must move to a bigger block
__global__ void k_plan(const uint32_t* __restrict__ nseg_p, const uint32_t* __restrict__ seg_start,
                       const uint32_t* __restrict__ seg_tile, StoreView S,
                       uint32_t* __restrict__ reloc_cap, unsigned long long* misc) {
  const uint32_t nseg = *nseg_p;
  for (uint32_t s = blockIdx.x * blockDim.x + threadIdx.x; s < nseg; s += gridDim.x * blockDim.x) {
    const uint32_t len = seg_start[s + 1] - seg_start[s];
    const uint32_t t = seg_tile[s];
    uint32_t nc = 0;
    if ((uint64_t)S.tile_cnt[t] + len > S.tile_cap[t]) {
      nc = grow_cap((uint64_t)S.tile_live[t] + len);
      atomicAdd(&misc[M_NEED], (unsigned long long)nc);
      atomicAdd(&misc[M_NRELOC], 1ull);
    }
    reloc_cap[s] = nc;
  }
}

// misc[M_OVERFLOW] := the relocations of this batch do not fit the pool.
__global__ void k_check_pool(unsigned long long* misc) {
  if (blockIdx.x == 0 && threadIdx.x == 0 && !misc[M_ERR])
    misc[M_OVERFLOW] = misc[M_POOL_TOP] + misc[M_NEED] > misc[M_POOL_SIZE] ? 1ull : 0ull;
}

__global__ void k_reloc_commit(const uint32_t* __restrict__ nseg_p,
                               const uint32_t* __restrict__ seg_tile,
                               const uint32_t* __restrict__ reloc_cap,
                               unsigned long long* __restrict__ reloc_list,
                               unsigned long long* misc) {
  if (batch_dead(misc)) return;
  const uint32_t nseg = *nseg_p;
  for (uint32_t s = blockIdx.x * blockDim.x + threadIdx.x; s < nseg; s += gridDim.x * blockDim.x) {
    const uint32_t nc = reloc_cap[s];
    if (!nc) continue;
    const unsigned long long off = atomicAdd(&misc[M_POOL_TOP], (unsigned long long)nc);
    const unsigned long long k = atomicAdd(&misc[M_RELOC_CUR], 1ull);
    reloc_list[3 * k + 0] = seg_tile[s];
    reloc_list[3 * k + 1] = off;
    reloc_list[3 * k + 2] = nc;
  }
}

// Stable copy of a tile's live slots to dst_off, 1024 slots per step.
// *live_sorted receives the live slots among the first `sorted` ones: the
// copy keeps their order, so they stay the tile's sorted prefix.
constexpr int CP_THREADS = 256;
constexpr int CP_ITEMS = 4;
__device__ inline uint32_t copy_live(const uint32_t* __restrict__ pool_in,
                                     const uint32_t* __restrict__ w_in, uint64_t src_off,
                                     uint32_t cnt, uint32_t* __restrict__ pool_out,
                                     uint32_t* __restrict__ w_out, uint64_t dst_off,
                                     uint32_t sorted, uint32_t* live_sorted) {
  uint32_t written = 0, nls = 0;
  for (uint32_t base = 0; base < cnt; base += CP_THREADS * CP_ITEMS) {
    uint32_t e[CP_ITEMS];
    uint32_t ww[CP_ITEMS];
    uint32_t nlive = 0;
#pragma unroll
    for (int k = 0; k < CP_ITEMS; ++k) {
      const uint32_t i = base + threadIdx.x * CP_ITEMS + k;
      e[k] = i < cnt ? pool_in[src_off + i] : TOMB;
      ww[k] = (w_in && i < cnt) ? w_in[src_off + i] : 0u;
      nlive += e[k] != TOMB;
      nls += (e[k] != TOMB && i < sorted) ? 1u : 0u;
    }
    uint32_t tot;
    uint32_t p = block_scan_excl<CP_THREADS, uint32_t>(nlive, &tot);
#pragma unroll
    for (int k = 0; k < CP_ITEMS; ++k) {
      if (e[k] != TOMB) {
        pool_out[dst_off + written + p] = e[k];
        if (w_out) w_out[dst_off + written + p] = ww[k];
        ++p;
      }
    }
    written += tot;
    __syncthreads();
  }
  *live_sorted = block_sum<CP_THREADS, uint32_t>(nls);
  return written;
}

__global__ void __launch_bounds__(CP_THREADS) k_relocate(const unsigned long long* __restrict__ rl,
                                                         const unsigned long long* misc,
                                                         StoreView S) {
  if (batch_dead(misc)) return;
  const uint32_t nreloc = (uint32_t)misc[M_RELOC_CUR];
  for (uint32_t b = blockIdx.x; b < nreloc; b += gridDim.x) {
    const uint32_t t = (uint32_t)rl[3 * b];
    const uint64_t noff = rl[3 * b + 1];
    const uint32_t ncap = (uint32_t)rl[3 * b + 2];
    uint32_t ls;
    const uint32_t live = copy_live(S.pool, S.wpool, S.tile_off[t], S.tile_cnt[t], S.pool, S.wpool,
                                    noff, S.tile_sorted[t], &ls);
    __syncthreads();
    if (threadIdx.x == 0) {
      S.tile_off[t] = noff;
      S.tile_cap[t] = ncap;
      S.tile_cnt[t] = live;
      S.tile_sorted[t] = ls;
    }
  }
}

// Global repack into a new pool (TileStore::doGlobalCompaction analogue).
__global__ void k_repack_caps(StoreView S, uint64_t* __restrict__ ncap) {
  for (uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; t < S.n_tiles;
       t += (uint64_t)gridDim.x * blockDim.x) {
    const uint32_t live = S.tile_live[t];
    ncap[t] = live ? grow_cap(live) : 0u;
  }
}

__global__ void __launch_bounds__(CP_THREADS) k_repack(StoreView S, const uint64_t* __restrict__ noff,
                                                       uint32_t* __restrict__ npool,
                                                       uint32_t* __restrict__ nwpool) {
  for (uint64_t t = blockIdx.x; t < S.n_tiles; t += gridDim.x) {
    const uint32_t cnt = S.tile_cnt[t];
    uint32_t live = 0, ls = 0;
    if (cnt)
      live = copy_live(S.pool, S.wpool, S.tile_off[t], cnt, npool, nwpool, noff[t],
                       S.tile_sorted[t], &ls);
    __syncthreads();
    if (threadIdx.x == 0) {
      S.tile_off[t] = noff[t];
      S.tile_cap[t] = (uint32_t)(noff[t + 1] - noff[t]);
      S.tile_cnt[t] = live;
      S.tile_sorted[t] = ls;
    }
  }
}

// ---- segmented scatter into slack
__global__ void k_scatter(const uint32_t* __restrict__ keys, const uint32_t* __restrict__ vals,
                          uint64_t n, uint32_t sb, const uint32_t* __restrict__ f,
                          const uint32_t* __restrict__ sx, const uint32_t* __restrict__ seg_start,
                          const uint32_t* __restrict__ src, const uint32_t* __restrict__ dst,
                          const uint32_t* __restrict__ w, StoreView S,
                          uint64_t* __restrict__ ins_slot, const unsigned long long* misc) {
  if (batch_dead(misc)) return;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x) {
    const uint32_t seg = sx[i] + f[i] - 1u;
    const uint32_t pos = (uint32_t)i - seg_start[seg];
    const uint32_t t = keys[i] >> sb;
    const uint32_t v = vals[i];
    const uint32_t s = src[v], d = dst[v];
    const uint64_t slot = S.tile_off[t] + S.tile_cnt[t] + pos;
    S.pool[slot] = pack_local(s, d, S.vpt - 1);
    if (S.wpool) S.wpool[slot] = w ? w[v] : 0u;
    ins_slot[v] = slot;
    atomicAdd(&S.outdeg[s], 1u);
    atomicAdd(&S.indeg[d], 1u);
  }
}

__global__ void k_seg_commit(const uint32_t* __restrict__ nseg_p,
                             const uint32_t* __restrict__ seg_start,
                             const uint32_t* __restrict__ seg_tile, StoreView S,
                             const unsigned long long* misc) {
  if (batch_dead(misc)) return;
  const uint32_t nseg = *nseg_p;
  for (uint32_t s = blockIdx.x * blockDim.x + threadIdx.x; s < nseg; s += gridDim.x * blockDim.x) {
    const uint32_t len = seg_start[s + 1] - seg_start[s];
    const uint32_t t = seg_tile[s];
    S.tile_cnt[t] += len;
    S.tile_live[t] += len;
  }
}

// ---- deletions
// Open-addressing table over (src, dst) keys; a multiplicative hash keeps
// the per-slot probe cheap.
__device__ inline uint64_t tab_hash(uint64_t key, uint32_t log_h) {
  return (key * 0x9E3779B97F4A7C15ull) >> (64 - log_h);
}

// Per-tile 256-bit signature of the batch's deletion keys: a slot whose
// signature bit is clear cannot match, and skips the table probe (tiles
// hold one or two requests on average, so almost every slot does).
__device__ inline uint32_t sig_bit(uint64_t key) {
  return (uint32_t)((key * 0xC2B2AE3D27D4EB4Full) >> 56);
}

// Blocked Bloom filter of the batch's deletion keys: two bits of one 64-bit
// word per key, about one key per word, so a slot of a heavy tile (whose
// 256-bit signature is saturated) reads one word instead of walking a
// linear-probe chain: only the slots whose two bits are both set (the
// matches and ~0.2 % false positives) probe the table.
__device__ inline uint64_t bloom_mask(uint64_t key) {
  const uint64_t h = key * 0xD6E8FEB86659FD93ull;
  return (1ull << (h >> 58)) | (1ull << ((h >> 52) & 63u));
}
__device__ inline uint64_t bloom_word(uint64_t key, uint32_t log_b) {
  return (key * 0x9FB21C651E98DF25ull) >> (64 - log_b);
}

__global__ void k_del_build(const uint32_t* __restrict__ src, const uint32_t* __restrict__ dst,
                            uint64_t m, unsigned long long* __restrict__ tab_key,
                            uint32_t* __restrict__ tab_head, uint32_t log_h,
                            uint32_t* __restrict__ node_next, uint32_t* __restrict__ tile_flag,
                            unsigned long long* __restrict__ tile_sig,
                            unsigned long long* __restrict__ bloom, uint32_t log_b, uint32_t lv,
                            uint32_t T) {
  const uint64_t hmask = (1ull << log_h) - 1;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < m;
       i += (uint64_t)gridDim.x * blockDim.x) {
    const uint32_t s = src[i], d = dst[i];
    const unsigned long long key = pack_edge(s, d);
    uint64_t h = tab_hash(key, log_h);
    for (;;) {
      const unsigned long long prev = atomicCAS(&tab_key[h], (unsigned long long)EMPTY64, key);
      if (prev == EMPTY64 || prev == key) break;
      h = (h + 1) & hmask;
    }
    node_next[i] = atomicExch(&tab_head[h], (uint32_t)i);
    const uint32_t t = tile_of(s, d, lv, T), b = sig_bit(key);
    tile_flag[t] = 1u;
    atomicOr(&tile_sig[4 * (uint64_t)t + (b >> 6)], 1ull << (b & 63u));
    atomicOr(&bloom[bloom_word(key, log_b)], (unsigned long long)bloom_mask(key));
  }
}

__global__ void k_flag_emit(const uint32_t* __restrict__ flag, const uint32_t* __restrict__ fx,
                            uint64_t n, uint32_t* __restrict__ list) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x)
    if (flag[i]) list[fx[i]] = (uint32_t)i;
}

// Chunks of the flagged tiles: DEL_CH slots each, so that one heavy tile is
// shared by many wavefronts.
constexpr uint32_t DEL_CH = 2048;

__global__ void k_del_nch(const uint32_t* __restrict__ flag, const uint32_t* __restrict__ cnt,
                          uint64_t n_tiles, uint32_t* __restrict__ nch) {
  for (uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n_tiles;
       t += (uint64_t)gridDim.x * blockDim.x)
    nch[t] = flag[t] ? (cnt[t] + DEL_CH - 1) / DEL_CH : 0u;
}

// One wavefront per tile, its lanes over the tile's chunks.
__global__ void k_del_chunks(const uint32_t* __restrict__ nch, const uint32_t* __restrict__ cbase,
                             uint64_t n_tiles, uint64_t* __restrict__ chunks) {
  const uint32_t lane = threadIdx.x & 63;
  for (uint64_t t = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6; t < n_tiles;
       t += ((uint64_t)gridDim.x * blockDim.x) >> 6) {
    const uint32_t n = nch[t], b = cbase[t];
    for (uint32_t c = lane; c < n; c += 64)
      chunks[b + c] = ((uint64_t)t << 32) | (uint64_t)(c * DEL_CH);
  }
}

// Each live slot of a chunk probes the table; a hit walks the request list
// of its (src, dst) and claims the first unclaimed request with a matching
// weight (any weight when dw is null).  The claimed slot becomes a tombstone.
constexpr int DM_UNROLL = 8;
__global__ void __launch_bounds__(256) k_del_match(const uint64_t* __restrict__ chunks,
                                                   const uint32_t* __restrict__ nchunks_p,
                                                   StoreView S,
                                                   const unsigned long long* __restrict__ tab_key,
                                                   const uint32_t* __restrict__ tab_head,
                                                   uint32_t log_h,
                                                   const uint32_t* __restrict__ node_next,
                                                   const uint32_t* __restrict__ dw,
                                                   const unsigned long long* __restrict__ tile_sig,
                                                   const unsigned long long* __restrict__ bloom,
                                                   uint32_t log_b,
                                                   unsigned long long* __restrict__ claimed,
                                                   const unsigned long long* misc) {
  if (batch_dead(misc)) return;
  const uint64_t hmask = (1ull << log_h) - 1;
  const uint32_t nchunks = *nchunks_p;
  const int lane = threadIdx.x & 63;
  const uint32_t nwaves = gridDim.x * (blockDim.x >> 6);
  for (uint32_t c = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6); c < nchunks;
       c += nwaves) {
    const uint64_t ch = chunks[c];
    const uint32_t t = (uint32_t)(ch >> 32), s0 = (uint32_t)ch;
    const uint32_t xb = (t % S.T) * S.vpt, yb = (t / S.T) * S.vpt;
    const uint64_t off = S.tile_off[t] + s0;
    const uint32_t n = min(DEL_CH, S.tile_cnt[t] - s0);
    const unsigned long long sg[4] = {tile_sig[4 * (uint64_t)t], tile_sig[4 * (uint64_t)t + 1],
                                      tile_sig[4 * (uint64_t)t + 2],
                                      tile_sig[4 * (uint64_t)t + 3]};
    for (uint32_t base = 0; base < n; base += 64 * DM_UNROLL) {
      uint32_t le[DM_UNROLL];
#pragma unroll
      for (int u = 0; u < DM_UNROLL; ++u) le[u] = S.pool[off + min(base + u * 64 + lane, n - 1)];
      uint64_t e[DM_UNROLL], h[DM_UNROLL];
      unsigned long long k[DM_UNROLL];
      bool cand[DM_UNROLL];
#pragma unroll
      for (int u = 0; u < DM_UNROLL; ++u) {
        e[u] = pack_edge(xb + (le[u] & 0xFFFFu), yb + (le[u] >> 16));
        const uint32_t b = sig_bit(e[u]);
        cand[u] = (sg[b >> 6] >> (b & 63u)) & 1ull;
        const uint64_t bm = bloom_mask(e[u]);
        cand[u] = cand[u] && (bloom[bloom_word(e[u], log_b)] & bm) == bm;
      }
#pragma unroll
      for (int u = 0; u < DM_UNROLL; ++u) {
        h[u] = tab_hash(e[u], log_h);
        k[u] = cand[u] ? tab_key[h[u]] : EMPTY64;
      }
#pragma unroll
      for (int u = 0; u < DM_UNROLL; ++u) {
        const uint32_t i = base + u * 64 + lane;
        if (i >= n || le[u] == TOMB || k[u] == EMPTY64) continue;
        uint64_t hh = h[u];
        unsigned long long kk = k[u];
        while (kk != e[u] && kk != EMPTY64) {
          hh = (hh + 1) & hmask;
          kk = tab_key[hh];
        }
        if (kk == EMPTY64) continue;
        const uint64_t slot = off + i;
        const uint32_t sw = S.wpool ? S.wpool[slot] : 0u;
        for (uint32_t r = tab_head[hh]; r != NONE32; r = node_next[r]) {
          if (dw && dw[r] != sw) continue;
          if (atomicCAS(&claimed[r], (unsigned long long)EMPTY64, (unsigned long long)slot) ==
              EMPTY64) {
            S.pool[slot] = TOMB;
            break;
          }
        }
      }
    }
  }
}

__global__ void k_del_finish(const uint32_t* __restrict__ src, const uint32_t* __restrict__ dst,
                             uint64_t m, const unsigned long long* __restrict__ claimed,
                             StoreView S, uint32_t* __restrict__ found, unsigned long long* misc) {
  if (batch_dead(misc)) return;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < m;
       i += (uint64_t)gridDim.x * blockDim.x) {
    const unsigned long long c = claimed[i];
    if (c == EMPTY64) continue;
    const uint32_t s = src[i], d = dst[i];
    atomicSub(&S.tile_live[tile_of(s, d, S.log2vpt, S.T)], 1u);
    const unsigned long long k = atomicAdd(&misc[M_NFOUND], 1ull);
    found[3 * k + 0] = s;
    found[3 * k + 1] = d;
    found[3 * k + 2] = S.wpool ? S.wpool[c] : 0u;
  }
}

// Degrees of the found deletions (MetaTileManager::deleteEdge,
// meta-tile-manager.h:127-140).  Runs after the re-execution checks and hooks:
// the reference evaluates edgeChanged(DELETE) while the deleted edge still
// counts in its source's out-degree (its random deletions,
// meta-tile-manager.h:200-228, leave the degrees alone, tile-manager.h:175-189),
// so rank / out_degree never divides by zero there.
__global__ void k_del_degrees(const uint32_t* __restrict__ found, uint64_t n,
                              uint32_t* __restrict__ outdeg, uint32_t* __restrict__ indeg,
                              const unsigned long long* misc) {
  if (batch_dead(misc)) return;
  n = min(n, (uint64_t)misc[M_NFOUND]);
  for (uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n;
       k += (uint64_t)gridDim.x * blockDim.x) {
    atomicSub(&outdeg[found[3 * k + 0]], 1u);
    atomicSub(&indeg[found[3 * k + 1]], 1u);
  }
}

// In-place compaction of tiles with many tombstones (compactStorage
// analogue that keeps multiplicity and only drops deleted slots).
__global__ void __launch_bounds__(CP_THREADS) k_compact(const uint32_t* __restrict__ tiles,
                                                        const uint32_t* __restrict__ ntiles_p,
                                                        StoreView S, unsigned long long* misc) {
  if (batch_dead(misc)) return;
  const uint32_t ntiles = *ntiles_p;
  for (uint32_t b = blockIdx.x; b < ntiles; b += gridDim.x) {
    const uint32_t t = tiles[b];
    const uint32_t cnt = S.tile_cnt[t], live = S.tile_live[t];
    const uint32_t dead = cnt - live;
    if (dead < 64 || (uint64_t)dead * 4 < cnt) continue;  // uniform per block
    const uint64_t off = S.tile_off[t];
    uint32_t ls;
    const uint32_t kept =
        copy_live(S.pool, S.wpool, off, cnt, S.pool, S.wpool, off, S.tile_sorted[t], &ls);
    __syncthreads();
    if (threadIdx.x == 0) {
      S.tile_cnt[t] = kept;
      S.tile_sorted[t] = ls;
      atomicAdd(&misc[M_NCOMPACT], 1ull);
    }
  }
}

// ---- sort-compaction: rewrite a tile's live edges in source order
// (the layout of a freshly bulk-loaded tile), so the gather's source-value
// loads of one wavefront fall on few cache lines.  Appends leave an unsorted
// tail behind the sorted prefix; a tile is rewritten once the tail is long
// enough, which bounds the rewrite work per inserted edge by a geometric sum.
constexpr uint32_t RESORT_MIN_TAIL = 256;

// A tile larger than the slot budget cap is never selected (it would stage
// past the buffers and block every later pass).
__global__ void k_resort_flag(StoreView S, uint32_t pct, uint64_t cap, uint32_t* __restrict__ flag,
                              uint64_t* __restrict__ size) {
  for (uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; t < S.n_tiles;
       t += (uint64_t)gridDim.x * blockDim.x) {
    const uint32_t cnt = S.tile_cnt[t], s = S.tile_sorted[t], tail = cnt - s;
    const bool f = tail >= RESORT_MIN_TAIL && (uint64_t)tail * 100 >= (uint64_t)s * pct &&
                   cnt <= cap;
    flag[t] = f ? 1u : 0u;
    size[t] = f ? cnt : 0u;
  }
}

// Selected tiles past the rank budget of this pass, or whose slots would end
// past the slot budget: the smallest such rank is the cut (ranks and offsets
// grow together).
__global__ void k_resort_cut(const uint32_t* __restrict__ flag, const uint32_t* __restrict__ rank,
                             const uint64_t* __restrict__ off, const uint64_t* __restrict__ size,
                             uint64_t n_tiles, uint32_t rmax, uint64_t cap,
                             unsigned long long* misc) {
  for (uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n_tiles;
       t += (uint64_t)gridDim.x * blockDim.x)
    if (flag[t] && (rank[t] >= rmax || off[t] + size[t] > cap))
      atomicMin(&misc[M_RS_CUT], (unsigned long long)rank[t]);
}

// Tiles and slots of this pass: the selected tiles before the cut.
__global__ void k_resort_count(const uint32_t* __restrict__ n_sel_p,
                               const uint64_t* __restrict__ off,
                               const uint32_t* __restrict__ list, uint64_t n_tiles,
                               unsigned long long* misc) {
  if (blockIdx.x != 0 || threadIdx.x != 0) return;
  const uint32_t n_sel = *n_sel_p;
  const uint64_t cut = misc[M_RS_CUT];
  const uint32_t nproc = cut < n_sel ? (uint32_t)cut : n_sel;
  misc[M_RS_NPROC] = nproc;
  misc[M_RS_NSLOTS] = nproc < n_sel ? off[list[nproc]] : off[n_tiles];
}

// The staging copy of the selected tiles is cut into RS_BLOCK-slot pieces,
// one per block of k_resort_gather; bfirst[b] = the selected tile holding
// piece b's first slot (tiles lie back to back in list order).  Hot tiles
// span many pieces, small ones share a piece.
constexpr uint32_t RS_BLOCK = 1024;
constexpr uint32_t RS_TILES = RS_BLOCK / RESORT_MIN_TAIL + 2;  // tiles one piece can touch

__global__ void k_resort_pieces(const uint32_t* __restrict__ list,
                                const unsigned long long* __restrict__ misc,
                                const uint64_t* __restrict__ off,
                                const uint32_t* __restrict__ tile_cnt,
                                uint32_t* __restrict__ bfirst) {
  const uint32_t nproc = (uint32_t)misc[M_RS_NPROC];
  for (uint32_t r = blockIdx.x * blockDim.x + threadIdx.x; r < nproc;
       r += gridDim.x * blockDim.x) {
    const uint32_t t = list[r];
    const uint64_t b0 = off[t], b1 = b0 + tile_cnt[t];
    for (uint64_t b = (b0 + RS_BLOCK - 1) / RS_BLOCK; b * RS_BLOCK < b1; ++b) bfirst[b] = r;
  }
}

// Source bits of a sort-compaction key: the tile-local source's top 8 bits,
// so a tile's prefix ends up ordered by 64-vertex source groups (the lines of
// source values a row reads are the same as with exact order), and with the
// pass's tile rank and the tombstone bit the key fits 24 bits: 3 radix passes.
__host__ __device__ inline uint32_t rs_sbits(uint32_t lv) { return lv < 8 ? lv : 8u; }

// Keys (rank, tombstone bit, source group) and staging copies of one piece
// of selected slots per block.
__global__ void __launch_bounds__(256) k_resort_gather(StoreView S, const uint32_t* __restrict__ list,
                                                       const unsigned long long* __restrict__ misc,
                                                       const uint64_t* __restrict__ off,
                                                       const uint32_t* __restrict__ bfirst,
                                                       uint32_t* __restrict__ keys,
                                                       uint32_t* __restrict__ vals,
                                                       uint32_t* __restrict__ e_st,
                                                       uint32_t* __restrict__ w_st) {
  __shared__ uint32_t s_rank[RS_TILES], s_base[RS_TILES];
  __shared__ uint64_t s_src[RS_TILES];
  __shared__ uint32_t s_n;
  const uint32_t lv = S.log2vpt, sb = rs_sbits(lv);
  const uint32_t nproc = (uint32_t)misc[M_RS_NPROC];
  const uint32_t nslots = (uint32_t)misc[M_RS_NSLOTS];
  for (uint32_t b = blockIdx.x; (uint64_t)b * RS_BLOCK < nslots; b += gridDim.x) {
    const uint32_t k0 = b * RS_BLOCK, k1 = min(k0 + RS_BLOCK, nslots);
    if (threadIdx.x < RS_TILES) {
      const uint32_t r = bfirst[b] + threadIdx.x;
      const bool in = r < nproc && (threadIdx.x == 0 || off[list[r]] < k1);
      if (in) {
        const uint32_t t = list[r];
        s_rank[threadIdx.x] = r;
        s_base[threadIdx.x] = (uint32_t)off[t];
        s_src[threadIdx.x] = S.tile_off[t];
      }
      // the tiles of a piece are consecutive in the list: count them
      const unsigned long long m = __ballot(in);
      if (threadIdx.x == 0) s_n = (uint32_t)__popcll(m);
    }
    __syncthreads();
    const uint32_t n = s_n;
    uint32_t e[RS_BLOCK / 256], wv[RS_BLOCK / 256], j[RS_BLOCK / 256];
#pragma unroll
    for (int u = 0; u < (int)(RS_BLOCK / 256); ++u) {
      const uint32_t k = min(k0 + u * 256 + threadIdx.x, k1 - 1);
      uint32_t jj = 0;
      for (uint32_t q = 1; q < n; ++q) jj = s_base[q] <= k ? q : jj;
      j[u] = jj;
      const uint64_t slot = s_src[jj] + (k - s_base[jj]);
      e[u] = S.pool[slot];
      wv[u] = S.wpool ? S.wpool[slot] : 0u;
    }
#pragma unroll
    for (int u = 0; u < (int)(RS_BLOCK / 256); ++u) {
      const uint32_t k = k0 + u * 256 + threadIdx.x;
      if (k >= k1) continue;
      const uint32_t low = e[u] == TOMB ? (1u << sb) : ((e[u] & 0xFFFFu) >> (lv - sb));
      keys[k] = (s_rank[j[u]] << (sb + 1)) | low;
      vals[k] = k;
      e_st[k] = e[u];
      if (S.wpool) w_st[k] = wv[u];
    }
    __syncthreads();
  }
}

// Sorted position k of a tile lands at slot (k - first position of the
// tile); tombstones sort behind the live edges and are dropped.
__global__ void k_resort_scatter(StoreView S, const uint32_t* __restrict__ list,
                                 const uint64_t* __restrict__ off,
                                 const uint32_t* __restrict__ skeys,
                                 const uint32_t* __restrict__ svals,
                                 const unsigned long long* __restrict__ misc,
                                 const uint32_t* __restrict__ e_st,
                                 const uint32_t* __restrict__ w_st) {
  if (batch_dead(misc)) return;
  const uint32_t sb = rs_sbits(S.log2vpt);
  const uint32_t n = (uint32_t)misc[M_RS_NSLOTS];
  for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) {
    const uint32_t key = skeys[k];
    if (key & (1u << sb)) continue;
    const uint32_t t = list[key >> (sb + 1)];
    const uint64_t slot = S.tile_off[t] + (k - (uint32_t)off[t]);
    const uint32_t j = svals[k];
    S.pool[slot] = e_st[j];
    if (S.wpool) S.wpool[slot] = w_st[j];
  }
}

__global__ void k_resort_commit(StoreView S, const uint32_t* __restrict__ list,
                                const unsigned long long* __restrict__ misc) {
  if (batch_dead(misc)) return;
  const uint32_t nproc = (uint32_t)misc[M_RS_NPROC];
  for (uint32_t r = blockIdx.x * blockDim.x + threadIdx.x; r < nproc; r += gridDim.x * blockDim.x) {
    const uint32_t t = list[r];
    const uint32_t live = S.tile_live[t];
    S.tile_cnt[t] = live;
    S.tile_sorted[t] = live;
  }
}

inline uint32_t grid_for(uint64_t n, uint32_t threads, uint32_t cap = 65536) {
  uint64_t b = (n + threads - 1) / threads;
  if (b == 0) b = 1;
  return (uint32_t)std::min<uint64_t>(b, cap);
}

// Repack every tile into a new pool of at least min_slots slots.
int repack(cytom_ctx* ctx, uint64_t min_slots) {
  hipStream_t st = ctx->stream;
  int rc;
  if ((rc = ensure(ctx, ctx->g_tile_scan, (ctx->n_tiles + 1) * 16 + 256))) return rc;
  uint64_t* ncap = static_cast<uint64_t*>(ctx->g_tile_scan.p);
  uint64_t* noff = ncap + ctx->n_tiles + 1;
  if ((rc = ensure(ctx, ctx->g_scan_tmp, scan_tmp_bytes(ctx->n_tiles) + 256))) return rc;
  StoreView S = ctx->view();
  k_repack_caps<<<grid_for(ctx->n_tiles, 256), 256, 0, st>>>(S, ncap);
  scan_exclusive_u64(ncap, noff, ctx->n_tiles, ctx->g_scan_tmp.p, st);
  uint64_t total = 0;
  CY_HIP(ctx, hipMemcpyAsync(&ctx->h_misc[0], noff + ctx->n_tiles, 8, hipMemcpyDeviceToHost, st));
  CY_HIP(ctx, hipStreamSynchronize(st));
  total = ctx->h_misc[0];
  uint64_t new_size = std::max<uint64_t>(min_slots, total + total / 2 + (1u << 20));
  uint32_t* npool = nullptr;
  uint32_t* nw = nullptr;
  if (hipMalloc(&npool, new_size * 4) != hipSuccess)
    return set_err(ctx, CYTOM_ENOMEM, "edge pool allocation failed");
  if (ctx->weighted && hipMalloc(&nw, new_size * 4) != hipSuccess) {
    (void)hipFree(npool);
    return set_err(ctx, CYTOM_ENOMEM, "weight pool allocation failed");
  }
  k_repack<<<grid_for(ctx->n_tiles, 1, 8192), CP_THREADS, 0, st>>>(S, noff, npool, nw);
  ctx->h_misc[0] = total;
  CY_HIP(ctx, hipMemcpyAsync(&ctx->d_misc[M_POOL_TOP], &ctx->h_misc[0], 8, hipMemcpyHostToDevice, st));
  CY_HIP(ctx, hipStreamSynchronize(st));
  (void)hipFree(ctx->pool);
  if (ctx->wpool) (void)hipFree(ctx->wpool);
  ctx->pool = npool;
  ctx->wpool = nw;
  ctx->pool_size = new_size;
  ctx->jobs_dirty = true;
  return 0;
}

// Slots staged per sort-compaction pass: the budget, bounded by the pool.
uint64_t resort_cap(const cytom_ctx* ctx) {
  return std::min<uint64_t>(ctx->resort_max_slots, ctx->pool_size);
}

int resort_prepare(cytom_ctx* ctx) {
  if (ctx->resort_pct == CYTOM_RESORT_OFF) return 0;
  const uint64_t nt = ctx->n_tiles, cap = std::max<uint64_t>(resort_cap(ctx), 1);
  int rc;
  if ((rc = ensure(ctx, ctx->g_tile_flag, nt * 4)) ||
      (rc = ensure(ctx, ctx->g_tile_scan, (nt + 1) * 16 + 256)) ||
      (rc = ensure(ctx, ctx->g_rs_off, (2 * nt + 1) * 8)) ||
      (rc = ensure(ctx, ctx->g_del_tiles, nt * 4)) || (rc = ensure(ctx, ctx->g_keys, cap * 4)) ||
      (rc = ensure(ctx, ctx->g_vals, cap * 4)) || (rc = ensure(ctx, ctx->g_keys2, cap * 4)) ||
      (rc = ensure(ctx, ctx->g_vals2, cap * 4)) || (rc = ensure(ctx, ctx->g_flags, cap * 4)) ||
      (rc = ensure(ctx, ctx->g_scan, cap * 4)) || (rc = ensure(ctx, ctx->g_rs_e, cap * 4)) ||
      (rc = ensure(ctx, ctx->g_rs_w, ctx->wpool ? cap * 4 : 4)) ||
      (rc = ensure(ctx, ctx->g_hist, radix_hist_bytes(cap))) ||
      (rc = ensure(ctx, ctx->g_scan_tmp,
                   scan_tmp_bytes(std::max<uint64_t>({cap, 256 * ((cap + 4095) / 4096) + 1,
                                                      nt + 1})) +
                       256)))
    return rc;
  return 0;
}

// Sort-compaction of the tiles k_resort_flag selects, at most the slot
// budget per pass (tiles past it wait for a later batch), device-driven.
void resort_enqueue(cytom_ctx* ctx) {
  if (ctx->resort_pct == CYTOM_RESORT_OFF) return;
  hipStream_t st = ctx->stream;
  const uint64_t nt = ctx->n_tiles, cap = std::max<uint64_t>(resort_cap(ctx), 1);
  const uint32_t lv = ctx->log2vpt;
  auto* flag = static_cast<uint32_t*>(ctx->g_tile_flag.p);
  auto* rank = static_cast<uint32_t*>(ctx->g_tile_scan.p);
  auto* size = static_cast<uint64_t*>(ctx->g_rs_off.p);
  uint64_t* off = size + nt;
  auto* list = static_cast<uint32_t*>(ctx->g_del_tiles.p);
  unsigned long long* misc = ctx->d_misc;
  k_resort_flag<<<grid_for(nt, 256), 256, 0, st>>>(ctx->view(), ctx->resort_pct, cap, flag,
                                                   size);
  scan_exclusive_u32(flag, rank, nt, ctx->g_scan_tmp.p, st);
  scan_exclusive_u64(size, off, nt, ctx->g_scan_tmp.p, st);
  k_flag_emit<<<grid_for(nt, 256), 256, 0, st>>>(flag, rank, nt, list);
  // Keys are (rank, tombstone bit, source group bits) in 24 bits.
  const uint32_t sb = rs_sbits(lv);
  const uint32_t rmax = 1u << (23 - sb);
  fill_u64(reinterpret_cast<uint64_t*>(misc + M_RS_CUT), 1, ~0ull, st);
  k_resort_cut<<<grid_for(nt, 256), 256, 0, st>>>(flag, rank, off, size, nt, rmax, cap, misc);
  k_resort_count<<<1, 64, 0, st>>>(rank + nt, off, list, nt, misc);
  const uint64_t max_proc = std::min<uint64_t>(nt, rmax);
  uint32_t rank_bits = 0;
  while ((1ull << rank_bits) < max_proc) ++rank_bits;
  const uint64_t pieces = (cap + RS_BLOCK - 1) / RS_BLOCK;
  auto* keys = static_cast<uint32_t*>(ctx->g_keys.p);
  auto* vals = static_cast<uint32_t*>(ctx->g_vals.p);
  auto* skeys = static_cast<uint32_t*>(ctx->g_keys2.p);
  auto* svals = static_cast<uint32_t*>(ctx->g_vals2.p);
  auto* e_st = static_cast<uint32_t*>(ctx->g_rs_e.p);
  auto* w_st = static_cast<uint32_t*>(ctx->g_rs_w.p);
  auto* bfirst = static_cast<uint32_t*>(ctx->g_flags.p);  // free until the sort below
  // The pass's counts are 32-bit values in the low half of their control
  // words: the sort reads them through 32-bit pointers.
  const auto* nslots_dev = reinterpret_cast<const uint32_t*>(misc + M_RS_NSLOTS);
  k_resort_pieces<<<grid_for(max_proc, 256), 256, 0, st>>>(list, misc, off, ctx->tile_cnt, bfirst);
  k_resort_gather<<<(uint32_t)std::min<uint64_t>(pieces, 16384), 256, 0, st>>>(
      ctx->view(), list, misc, off, bfirst, keys, vals, e_st, w_st);
  radix_sort_pairs_devn(keys, vals, skeys, svals, static_cast<uint32_t*>(ctx->g_flags.p),
                        static_cast<uint32_t*>(ctx->g_scan.p), nslots_dev, cap,
                        rank_bits + sb + 1, ctx->g_hist.p, ctx->g_scan_tmp.p, st);
  k_resort_scatter<<<grid_for(cap, 256), 256, 0, st>>>(ctx->view(), list, off, skeys, svals, misc,
                                                       e_st, w_st);
  k_resort_commit<<<grid_for(max_proc, 256), 256, 0, st>>>(ctx->view(), list, misc);
}

// Per program: the batch's decision, written to gate[p] = 1 (re-initialise),
// 0 (insert hooks) or 2 (the batch was refused: neither).  force: the host's
// rule alone decides re-initialisation (PPR with inserts); found_forces: any
// found deletion re-initialises (BFS/CC/SSSP whose values are not a
// converged fixpoint, PPR); otherwise a tight deletion (flag[p]) does.
__global__ void k_decide(const unsigned long long* __restrict__ misc,
                         const unsigned long long* __restrict__ flag,
                         const uint32_t* __restrict__ force,
                         const uint32_t* __restrict__ found_forces, uint32_t np,
                         unsigned long long* __restrict__ gate) {
  const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= np) return;
  if (batch_dead(misc)) {
    gate[p] = 2;
    return;
  }
  const bool found = misc[M_NFOUND] != 0;
  gate[p] = (force[p] || (found && (flag[p] || found_forces[p]))) ? 1ull : 0ull;
}

}  // namespace

// Sizes every scratch buffer of a batch of n_ins insertions and n_del
// deletions (host only: may allocate, so it runs before any capture).
// Bytes of the deletion Bloom filter for n requests (2^ceil(log2 n) words, at least 64).
uint64_t bloom_bytes(uint64_t n) {
  uint64_t w = 64;
  while (w < n) w <<= 1;
  return w * 8;
}

int ingest_prepare(cytom_ctx* ctx, uint64_t n_ins, uint64_t n_del) {
  int rc;
  if (n_ins > 0xFFFFFFFEull || n_del > 0xFFFFFFFEull)
    return set_err(ctx, CYTOM_ERANGE, "batch larger than 2^32-2 updates");
  const uint64_t nmax = std::max<uint64_t>(n_ins, 1), dmax = std::max<uint64_t>(n_del, 1);
  uint64_t H = 64;
  while (H < 2 * dmax) H <<= 1;
  const uint64_t max_del_chunks = ctx->pool_size / DEL_CH + ctx->n_tiles + 1;
  const uint64_t np = std::max<uint64_t>(ctx->progs.size(), 1);
  if ((rc = ensure(ctx, ctx->g_keys, nmax * 4)) || (rc = ensure(ctx, ctx->g_vals, nmax * 4)) ||
      (rc = ensure(ctx, ctx->g_keys2, nmax * 4)) || (rc = ensure(ctx, ctx->g_vals2, nmax * 4)) ||
      (rc = ensure(ctx, ctx->g_flags, nmax * 4)) ||
      (rc = ensure(ctx, ctx->g_scan, (nmax + 1) * 4)) ||
      (rc = ensure(ctx, ctx->g_seg_start, (nmax + 1) * 4)) ||
      (rc = ensure(ctx, ctx->g_seg_tile, nmax * 4)) ||
      (rc = ensure(ctx, ctx->g_reloc_cap, nmax * 4)) ||
      (rc = ensure(ctx, ctx->g_reloc_list, nmax * 24)) ||
      (rc = ensure(ctx, ctx->g_ins_slot, nmax * 8)) ||
      (rc = ensure(ctx, ctx->g_hist, radix_hist_bytes(nmax))) ||
      (rc = ensure(ctx, ctx->g_scan_tmp,
                   scan_tmp_bytes(std::max<uint64_t>({nmax, 256 * ((nmax + 4095) / 4096) + 1,
                                                      ctx->n_tiles + 1})) +
                       256)) ||
      (rc = ensure(ctx, ctx->g_tab_key, H * 8)) || (rc = ensure(ctx, ctx->g_tab_head, H * 4)) ||
      (rc = ensure(ctx, ctx->g_node_next, dmax * 4)) ||
      (rc = ensure(ctx, ctx->g_claimed, dmax * 8)) ||
      (rc = ensure(ctx, ctx->g_tile_flag, ctx->n_tiles * 4)) ||
      (rc = ensure(ctx, ctx->g_del_sig, ctx->n_tiles * 32)) ||
      (rc = ensure(ctx, ctx->g_del_bloom, bloom_bytes(dmax))) ||
      (rc = ensure(ctx, ctx->g_tile_scan, (ctx->n_tiles + 1) * 16 + 256)) ||
      (rc = ensure(ctx, ctx->g_del_tiles, ctx->n_tiles * 4)) ||
      (rc = ensure(ctx, ctx->g_found, dmax * 12)) ||
      (rc = ensure(ctx, ctx->g_nch, ctx->n_tiles * 4)) ||
      (rc = ensure(ctx, ctx->g_cbase, (ctx->n_tiles + 1) * 4)) ||
      (rc = ensure(ctx, ctx->g_chunks, max_del_chunks * 8)) ||
      (rc = ensure(ctx, ctx->g_pflags, np * 24)) || (rc = resort_prepare(ctx)))
    return rc;
  // The programs' host-side rules of this batch.
  auto* hp = reinterpret_cast<uint32_t*>(ctx->h_pflags);
  for (size_t p = 0; p < ctx->progs.size(); ++p) {
    const Program& P = ctx->progs[p];
    hp[p] = (P.kind == CYTOM_PPR && n_ins) ? 1u : 0u;                        // force
    hp[64 + p] = ((!prd_like(P.kind) && !P.fixpoint) || P.kind == CYTOM_PPR) ? 1u : 0u;
  }
  return 0;
}

// Launches one batch (no host reads; the counts the host used to read stay
// on the device).  The pinned rules written by ingest_prepare are copied in
// first, so a captured graph picks up each replay's rules.
void ingest_enqueue(cytom_ctx* ctx, const uint32_t* d_src, const uint32_t* d_dst,
                    const uint32_t* d_w, uint64_t n_ins, const uint32_t* d_dsrc,
                    const uint32_t* d_ddst, const uint32_t* d_dw, uint64_t n_del) {
  hipStream_t st = ctx->stream;
  const uint32_t lv = ctx->log2vpt, T = ctx->T;
  const uint32_t tile_bits = [&] {
    uint32_t b = 0;
    while ((1ull << b) < ctx->n_tiles) ++b;
    return b;
  }();
  // Source bits that fit beside the tile id in a 32-bit key, cut to a whole
  // number of 8-bit radix passes while the source groups they order stay
  // within 64 vertices (two lines of PageRank-delta source values).
  uint32_t sb = std::min<uint32_t>(lv, 32 - tile_bits);
  const uint32_t cut = (tile_bits + sb) % 8;
  if (cut && sb >= cut && lv - (sb - cut) <= 6) sb -= cut;
  const uint32_t key_bits = tile_bits + sb;
  uint32_t* keys = static_cast<uint32_t*>(ctx->g_keys.p);
  uint32_t* vals = static_cast<uint32_t*>(ctx->g_vals.p);
  uint32_t* skeys = static_cast<uint32_t*>(ctx->g_keys2.p);
  uint32_t* svals = static_cast<uint32_t*>(ctx->g_vals2.p);
  uint32_t* flags = static_cast<uint32_t*>(ctx->g_flags.p);
  uint32_t* sx = static_cast<uint32_t*>(ctx->g_scan.p);
  uint32_t* seg_start = static_cast<uint32_t*>(ctx->g_seg_start.p);
  uint32_t* seg_tile = static_cast<uint32_t*>(ctx->g_seg_tile.p);
  uint32_t* reloc_cap = static_cast<uint32_t*>(ctx->g_reloc_cap.p);
  unsigned long long* reloc_list = static_cast<unsigned long long*>(ctx->g_reloc_list.p);
  uint64_t* ins_slot = static_cast<uint64_t*>(ctx->g_ins_slot.p);
  unsigned long long* misc = ctx->d_misc;
  const size_t np = ctx->progs.size();
  auto* pf = static_cast<unsigned long long*>(ctx->g_pflags.p);  // [np] tight flags
  unsigned long long* gate = pf + np;                             // [np] decisions
  auto* rules = reinterpret_cast<uint32_t*>(gate + np);           // [np] force, [np] found_forces

  ctx->jobs_dirty = true;  // the store changes (and the deletion pass reuses the list buffers)
  if (!ctx->capturing) (void)hipEventRecord(ctx->ev[0], st);
  // misc[1..] := 0 (the pool top in misc[0] persists), the pool size, rules
  {
    const FillOp fills[] = {{misc + 1, M_POOL_SIZE - 1, 0, 1},
                            {misc + M_POOL_SIZE, 1, ctx->pool_size, 1},
                            {misc + M_POOL_SIZE + 1, 63 - M_POOL_SIZE, 0, 1},
                            {pf, np, 0, 1}};
    fill_many(fills, 4, st);
  }
  if (np) {
    (void)hipMemcpyAsync(rules, ctx->h_pflags, np * 4, hipMemcpyHostToDevice, st);
    (void)hipMemcpyAsync(rules + np, reinterpret_cast<uint32_t*>(ctx->h_pflags) + 64, np * 4,
                         hipMemcpyHostToDevice, st);
  }
  // 1. validation (a refused batch modifies nothing)
  if (n_ins)
    k_keys<<<grid_for(n_ins, 256), 256, 0, st>>>(d_src, d_dst, d_w, n_ins, ctx->V, lv, T, sb,
                                                 ctx->weighted ? 1 : 0, keys, vals, misc);
  if (n_del)
    k_del_validate<<<grid_for(n_del, 256), 256, 0, st>>>(d_dsrc, d_ddst, n_del, ctx->V, misc);
  // 2. insertions
  if (n_ins) {
    // flags / sx are free until the segment pass: they serve as the sort's
    // ping-pong buffers.
    radix_sort_pairs(keys, vals, skeys, svals, flags, sx, n_ins, key_bits, ctx->g_hist.p,
                     ctx->g_scan_tmp.p, st);
    k_heads<<<grid_for(n_ins, 256), 256, 0, st>>>(skeys, n_ins, sb, flags);
    scan_exclusive_u32(flags, sx, n_ins, ctx->g_scan_tmp.p, st);
    k_seg_emit<<<grid_for(n_ins, 256), 256, 0, st>>>(skeys, flags, sx, n_ins, sb, seg_start,
                                                     seg_tile);
    const uint32_t* nseg_p = sx + n_ins;
    k_plan<<<grid_for(n_ins, 256), 256, 0, st>>>(nseg_p, seg_start, seg_tile, ctx->view(),
                                                 reloc_cap, misc);
    k_check_pool<<<1, 64, 0, st>>>(misc);
    k_reloc_commit<<<grid_for(n_ins, 256), 256, 0, st>>>(nseg_p, seg_tile, reloc_cap,
                                                         reloc_list, misc);
    k_relocate<<<(uint32_t)std::min<uint64_t>(n_ins, 8192), CP_THREADS, 0, st>>>(
        reloc_list, misc, ctx->view());
    k_scatter<<<grid_for(n_ins, 256), 256, 0, st>>>(skeys, svals, n_ins, sb, flags, sx, seg_start,
                                                    d_src, d_dst, d_w, ctx->view(), ins_slot,
                                                    misc);
    k_seg_commit<<<grid_for(n_ins, 256), 256, 0, st>>>(nseg_p, seg_start, seg_tile, ctx->view(),
                                                       misc);
  }
  // 3. deletions
  uint32_t* found = static_cast<uint32_t*>(ctx->g_found.p);
  uint32_t* del_tiles = static_cast<uint32_t*>(ctx->g_del_tiles.p);
  uint32_t* n_del_tiles_p = nullptr;
  if (n_del) {
    uint64_t H = 64;
    uint32_t log_h = 6;
    while (H < 2 * n_del) {
      H <<= 1;
      ++log_h;
    }
    auto* tab_key = static_cast<unsigned long long*>(ctx->g_tab_key.p);
    auto* tab_head = static_cast<uint32_t*>(ctx->g_tab_head.p);
    auto* node_next = static_cast<uint32_t*>(ctx->g_node_next.p);
    auto* claimed = static_cast<unsigned long long*>(ctx->g_claimed.p);
    auto* tile_flag = static_cast<uint32_t*>(ctx->g_tile_flag.p);
    auto* tscan = static_cast<uint32_t*>(ctx->g_tile_scan.p);
    auto* dnch = static_cast<uint32_t*>(ctx->g_nch.p);
    auto* dcbase = static_cast<uint32_t*>(ctx->g_cbase.p);
    auto* dchunks = static_cast<uint64_t*>(ctx->g_chunks.p);
    auto* tile_sig = static_cast<unsigned long long*>(ctx->g_del_sig.p);
    auto* bloom = static_cast<unsigned long long*>(ctx->g_del_bloom.p);
    uint32_t log_b = 6;  // about one key per filter word
    while ((1ull << log_b) < n_del) ++log_b;
    const FillOp fills[] = {{bloom, 1ull << log_b, 0, 1},
                            {tab_key, H, ~0ull, 1},
                            {tab_head, H, 0xFFFFFFFFu, 0},
                            {claimed, n_del, ~0ull, 1},
                            {tile_flag, ctx->n_tiles, 0, 0},
                            {tile_sig, ctx->n_tiles * 4, 0, 1}};
    fill_many(fills, 6, st);
    k_del_build<<<grid_for(n_del, 256), 256, 0, st>>>(d_dsrc, d_ddst, n_del, tab_key, tab_head,
                                                      log_h, node_next, tile_flag, tile_sig,
                                                      bloom, log_b, lv, T);
    scan_exclusive_u32(tile_flag, tscan, ctx->n_tiles, ctx->g_scan_tmp.p, st);
    k_flag_emit<<<grid_for(ctx->n_tiles, 256), 256, 0, st>>>(tile_flag, tscan, ctx->n_tiles,
                                                            del_tiles);
    n_del_tiles_p = tscan + ctx->n_tiles;
    k_del_nch<<<grid_for(ctx->n_tiles, 256), 256, 0, st>>>(tile_flag, ctx->tile_cnt, ctx->n_tiles,
                                                           dnch);
    scan_exclusive_u32(dnch, dcbase, ctx->n_tiles, ctx->g_scan_tmp.p, st);
    k_del_chunks<<<grid_for(ctx->n_tiles * 64, 256), 256, 0, st>>>(dnch, dcbase, ctx->n_tiles,
                                                                   dchunks);
    k_del_match<<<(uint32_t)ctx->num_cus * 8, 256, 0, st>>>(
        dchunks, dcbase + ctx->n_tiles, ctx->view(), tab_key, tab_head, log_h, node_next,
        ctx->weighted ? d_dw : nullptr, tile_sig, bloom, log_b, claimed, misc);
    k_del_finish<<<grid_for(n_del, 256), 256, 0, st>>>(d_dsrc, d_ddst, n_del, claimed,
                                                       ctx->view(), found, misc);
  }
  // 4. per program: re-execution test, the decision, then init or hooks
  const unsigned long long* n_found_dev = misc + M_NFOUND;
  for (size_t p = 0; p < np; ++p)
    if (n_del) launch_reexec_check(ctx, ctx->progs[p], found, n_del, n_found_dev, pf + p);
  if (np)
    k_decide<<<grid_for(np, 64), 64, 0, st>>>(misc, pf, rules, rules + np, (uint32_t)np, gate);
  for (size_t p = 0; p < np; ++p) {
    Program& P = ctx->progs[p];
    program_init(ctx, P, gate + p);
    if (n_ins || n_del)
      launch_hooks(ctx, P, d_src, d_dst, d_w, ins_slot, n_ins, found, n_del, n_found_dev,
                   gate + p);
  }
  if (n_del)
    k_del_degrees<<<grid_for(n_del, 256), 256, 0, st>>>(found, n_del, ctx->outdeg, ctx->indeg,
                                                        misc);
  // 5. compaction of tiles with many tombstones, then sort-compaction of
  // tiles with long unsorted tails
  if (n_del) k_compact<<<4096, CP_THREADS, 0, st>>>(del_tiles, n_del_tiles_p, ctx->view(), misc);
  resort_enqueue(ctx);
  if (!ctx->capturing) (void)hipEventRecord(ctx->ev[1], st);
  // the control words and decisions, for ingest_finish
  (void)hipMemcpyAsync(ctx->h_misc, misc, 64 * 8, hipMemcpyDeviceToHost, st);
  if (np) (void)hipMemcpyAsync(ctx->h_misc + 64, gate, np * 8, hipMemcpyDeviceToHost, st);
}

// After the batch's synchronisation: 0, CYTOM_EINVAL (refused, nothing
// changed) or 1 (the pool is too small: nothing changed; repack and run the
// batch again).
int ingest_finish(cytom_ctx* ctx, uint64_t n_ins, uint64_t n_del, cytom_batch_stats* bst) {
  const unsigned long long* hm = ctx->h_misc;
  if (hm[M_ERR])
    return set_err(ctx, CYTOM_EINVAL,
                   (hm[M_ERR] & 1) ? "insertion with out-of-range vertex or zero weight"
                                   : "deletion with out-of-range vertex");
  if (hm[M_OVERFLOW]) return 1;
  const uint64_t n_found = hm[M_NFOUND];
  for (size_t p = 0; p < ctx->progs.size(); ++p) {
    Program& P = ctx->progs[p];
    const bool reexec = hm[64 + p] == 1;
    P.last_reexec = reexec ? 1 : 0;
    if (reexec || n_ins || n_found) P.fixpoint = false;
  }
  ctx->live_edges += n_ins;
  ctx->live_edges -= n_found;
  ctx->tiles_resorted += hm[M_RS_NPROC];
  ctx->slots_resorted += hm[M_RS_NSLOTS];
  ctx->batches += 1;
  if (bst) {
    float ms = 0.f;
    CY_HIP(ctx, hipEventElapsedTime(&ms, ctx->ev[0], ctx->ev[1]));
    bst->inserted = n_ins;
    bst->deleted = n_found;
    bst->missing = n_del - n_found;
    bst->live_edges = ctx->live_edges;
    bst->tiles_relocated = (uint32_t)hm[M_NRELOC];
    bst->tiles_compacted = (uint32_t)hm[M_NCOMPACT];
    bst->tiles_sorted = (uint32_t)hm[M_RS_NPROC];
    bst->host_syncs = 1;
    bst->t_ingest_ms = ms;
  }
  return 0;
}

// The slot budget a repack must leave for this batch's relocations.
uint64_t ingest_need(const cytom_ctx* ctx) { return ctx->h_misc[M_NEED]; }

int ingest_batch(cytom_ctx* ctx, const uint32_t* d_src, const uint32_t* d_dst,
                 const uint32_t* d_w, uint64_t n_ins, const uint32_t* d_dsrc,
                 const uint32_t* d_ddst, const uint32_t* d_dw, uint64_t n_del,
                 cytom_batch_stats* bst) {
  if (ctx->weighted && n_ins && !d_w)
    return set_err(ctx, CYTOM_EINVAL, "weighted graph needs insertion weights");
  for (int attempt = 0;; ++attempt) {
    int rc = ingest_prepare(ctx, n_ins, n_del);
    if (rc) return rc;
    ingest_enqueue(ctx, d_src, d_dst, d_w, n_ins, d_dsrc, d_ddst, d_dw, n_del);
    CY_HIP(ctx, hipStreamSynchronize(ctx->stream));
    CY_HIP(ctx, hipGetLastError());
    rc = ingest_finish(ctx, n_ins, n_del, bst);
    if (bst) bst->host_syncs = (uint32_t)attempt + 1;
    if (rc <= 0) return rc;
    if (attempt > 0) return set_err(ctx, CYTOM_ENOMEM, "edge pool exhausted after repack");
    if ((rc = repack(ctx, std::max<uint64_t>(2 * ctx->pool_size,
                                             2 * (ctx->live_edges + n_ins) + ingest_need(ctx)))))
      return rc;
  }
}

int64_t export_edges(cytom_ctx* ctx, uint32_t* src, uint32_t* dst, uint32_t* w, uint64_t cap) {
  if (!src) return (int64_t)ctx->live_edges;
  hipStream_t st = ctx->stream;
  std::vector<uint64_t> off(ctx->n_tiles);
  std::vector<uint32_t> cnt(ctx->n_tiles);
  CY_HIP(ctx, hipMemcpyAsync(off.data(), ctx->tile_off, ctx->n_tiles * 8, hipMemcpyDeviceToHost, st));
  CY_HIP(ctx, hipMemcpyAsync(cnt.data(), ctx->tile_cnt, ctx->n_tiles * 4, hipMemcpyDeviceToHost, st));
  CY_HIP(ctx, hipMemcpyAsync(ctx->h_misc, ctx->d_misc, 8, hipMemcpyDeviceToHost, st));
  CY_HIP(ctx, hipStreamSynchronize(st));
  const uint64_t top = ctx->h_misc[0];
  std::vector<uint32_t> pool(top);
  std::vector<uint32_t> wp(ctx->wpool ? top : 0);
  if (top) {
    CY_HIP(ctx, hipMemcpy(pool.data(), ctx->pool, top * 4, hipMemcpyDeviceToHost));
    if (ctx->wpool) CY_HIP(ctx, hipMemcpy(wp.data(), ctx->wpool, top * 4, hipMemcpyDeviceToHost));
  }
  uint64_t n = 0;
  for (uint64_t t = 0; t < ctx->n_tiles; ++t) {
    for (uint32_t i = 0; i < cnt[t]; ++i) {
      const uint32_t e = pool[off[t] + i];
      if (e == TOMB) continue;
      if (n < cap) {
        src[n] = (uint32_t)((t % ctx->T) * ctx->vpt) + (e & 0xFFFFu);
        dst[n] = (uint32_t)((t / ctx->T) * ctx->vpt) + (e >> 16);
        if (w) w[n] = ctx->wpool ? wp[off[t] + i] : 0u;
      }
      ++n;
    }
  }
  return (int64_t)n;
}

}  // namespace cytom
