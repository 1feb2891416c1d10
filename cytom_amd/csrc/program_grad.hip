// Reverse mode of a registered personalized-PageRank program on the
// context's evolving graph (no reference counterpart).
//
// The program's ranks solve r = (1-alpha) t + alpha P^T r on the live
// multigraph with P_uv = 1 / out_degree(u) per live edge (every weight 1).
// For a loss gradient g = dL/dr the adjoint lambda = g + alpha P lambda is
// solved here by FP32 Jacobi sweeps over the tile store's live slots, each
// slot read in the reverse orientation of the program's gather: the
// program's gather moves rank from a slot's source to its target, the
// adjoint moves lambda from the target back to the source (a double
// accumulator per source).  Then
//   dL/dt    = (1-alpha) lambda
//   dL/dw_uv = alpha r_u (lambda_v - mu_u) / out_degree(u),
//   mu_u     = sum over u's live out-edges of lambda_dst / out_degree(u),
// the derivative with respect to an edge weight at w = 1 (the weighted
// operator P_uv = w_uv / W_u of the explicit-edge-list form).  Edges are
// written in the order cytom_export_edges lists them: tiles in order, slots
// in order, tombstones skipped.
#include <algorithm>
#include <cstring>
#include <vector>

#include "internal.h"

namespace cytom {
namespace {

constexpr float PPR_ALPHA = 0.85f;

inline uint32_t blocks_of(uint64_t n, uint32_t per_block, uint32_t cap = 16384) {
  const uint64_t b = (n + per_block - 1) / per_block;
  return (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(b, cap));
}

// A sweep is skipped once an earlier one changed lambda by at most tol
// (sweeps are enqueued in groups between host checks).  delta[k] holds the
// bits of sweep k's largest change (non-negative floats order as their bits);
// a skipped sweep leaves its entry 0.
__device__ inline bool converged_before(const uint32_t* delta, uint32_t k, float tol) {
  return k > 0 && __uint_as_float(delta[k - 1]) <= tol;
}

// acc[u] += lambda[v] over every live slot (u, v), one wavefront per work
// chunk (the gather's chunk list: tile id in the high word, first slot low).
__global__ void k_pa_accumulate(const uint64_t* __restrict__ chunks,
                                const uint32_t* __restrict__ n_chunks_dev, StoreView S,
                                uint32_t CH, const float* __restrict__ lam,
                                double* __restrict__ acc, const uint32_t* __restrict__ delta,
                                uint32_t k, float tol) {
  if (converged_before(delta, k, tol)) return;
  const uint32_t lane = threadIdx.x & 63;
  const uint64_t n_chunks = *n_chunks_dev;
  for (uint64_t c = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6; c < n_chunks;
       c += ((uint64_t)gridDim.x * blockDim.x) >> 6) {
    const uint64_t ch = chunks[c];
    const uint32_t t = (uint32_t)(ch >> 32), s0 = (uint32_t)ch;
    const uint32_t n = min(CH, S.tile_cnt[t] - s0);
    const uint64_t off = S.tile_off[t] + s0;
    const uint32_t xb = (uint32_t)((t % S.T) << S.log2vpt);
    const uint32_t yb = (uint32_t)((t / S.T) << S.log2vpt);
    for (uint32_t i = lane; i < n; i += 64) {
      const uint32_t e = S.pool[off + i];
      if (e == TOMB) continue;
      atomicAdd(&acc[xb + (e & 0xFFFFu)], (double)lam[yb + (e >> 16)]);
    }
  }
}

// lambda_next = g + alpha acc / out_degree; acc := 0; delta[k] = max change.
__global__ void k_pa_update(uint64_t V, const float* __restrict__ g,
                            const uint32_t* __restrict__ outdeg, double* __restrict__ acc,
                            const float* __restrict__ lam, float* __restrict__ lam_next,
                            uint32_t* __restrict__ delta, uint32_t k, float tol) {
  if (converged_before(delta, k, tol)) return;
  float m = 0.f;
  for (uint64_t v = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; v < V;
       v += (uint64_t)gridDim.x * blockDim.x) {
    const uint32_t d = outdeg[v];
    const float nl = g[v] + (d ? (float)((double)PPR_ALPHA * acc[v] / (double)d) : 0.f);
    m = fmaxf(m, fabsf(nl - lam[v]));
    lam_next[v] = nl;
    acc[v] = 0.0;
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) m = fmaxf(m, __shfl_xor(m, o, 64));
  if ((threadIdx.x & 63) == 0) atomicMax(&delta[k], __float_as_uint(m));
}

__global__ void k_pa_teleport(uint64_t V, const float* __restrict__ lam, float* __restrict__ ds) {
  for (uint64_t v = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; v < V;
       v += (uint64_t)gridDim.x * blockDim.x)
    ds[v] = (1.f - PPR_ALPHA) * lam[v];
}

// Live slots per chunk (the export positions are their exclusive scan).
__global__ void k_pa_live(const uint64_t* __restrict__ chunks,
                          const uint32_t* __restrict__ n_chunks_dev, StoreView S, uint32_t CH,
                          uint64_t* __restrict__ live) {
  const uint32_t lane = threadIdx.x & 63;
  const uint64_t n_chunks = *n_chunks_dev;
  for (uint64_t c = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6; c < n_chunks;
       c += ((uint64_t)gridDim.x * blockDim.x) >> 6) {
    const uint64_t ch = chunks[c];
    const uint32_t t = (uint32_t)(ch >> 32), s0 = (uint32_t)ch;
    const uint32_t n = min(CH, S.tile_cnt[t] - s0);
    const uint64_t off = S.tile_off[t] + s0;
    uint32_t cnt = 0;
    for (uint32_t i = lane; i < n; i += 64) cnt += S.pool[off + i] != TOMB ? 1u : 0u;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) cnt += __shfl_xor(cnt, o, 64);
    if (lane == 0) live[c] = cnt;
  }
}

// Every live edge's endpoints and dL/dw at its export position; mu_u is
// acc[u] / out_degree(u) (acc holds the final sweep's accumulation).
__global__ void k_pa_edges(const uint64_t* __restrict__ chunks,
                           const uint32_t* __restrict__ n_chunks_dev, StoreView S, uint32_t CH,
                           const uint64_t* __restrict__ base, const float* __restrict__ r,
                           const float* __restrict__ lam, const double* __restrict__ acc,
                           const uint32_t* __restrict__ outdeg, uint64_t cap,
                           uint32_t* __restrict__ src, uint32_t* __restrict__ dst,
                           float* __restrict__ dw) {
  const uint32_t lane = threadIdx.x & 63;
  const uint64_t n_chunks = *n_chunks_dev;
  for (uint64_t c = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6; c < n_chunks;
       c += ((uint64_t)gridDim.x * blockDim.x) >> 6) {
    const uint64_t ch = chunks[c];
    const uint32_t t = (uint32_t)(ch >> 32), s0 = (uint32_t)ch;
    const uint32_t n = min(CH, S.tile_cnt[t] - s0);
    const uint64_t off = S.tile_off[t] + s0;
    const uint32_t xb = (uint32_t)((t % S.T) << S.log2vpt);
    const uint32_t yb = (uint32_t)((t / S.T) << S.log2vpt);
    uint64_t p = base[c];
    for (uint32_t i0 = 0; i0 < n; i0 += 64) {
      const uint32_t i = i0 + lane;
      const uint32_t e = i < n ? S.pool[off + i] : TOMB;
      const bool live = e != TOMB;
      const uint64_t m = __ballot(live);
      if (live) {
        const uint64_t q = p + __popcll(m & ((1ull << lane) - 1ull));
        const uint32_t u = xb + (e & 0xFFFFu), v = yb + (e >> 16);
        if (q < cap) {
          const double d = (double)outdeg[u];  // >= 1: u has this live out-edge
          if (src) src[q] = u;
          if (dst) dst[q] = v;
          if (dw)
            dw[q] = (float)((double)PPR_ALPHA * (double)r[u] * ((double)lam[v] - acc[u] / d) / d);
        }
      }
      p += __popcll(m);
    }
  }
}

}  // namespace

int64_t ppr_program_adjoint(cytom_ctx* ctx, Program& P, const float* d_g, float tol,
                            uint32_t max_iterations, float* d_ds, uint32_t* d_src,
                            uint32_t* d_dst, float* d_dw, uint64_t cap, uint32_t* iterations) {
  if (P.kind != CYTOM_PPR)
    return set_err(ctx, CYTOM_EINVAL, "the adjoint is defined for personalized PageRank programs");
  if (!d_g) return set_err(ctx, CYTOM_EINVAL, "null loss gradient");
  if (!(tol >= 0.f)) return set_err(ctx, CYTOM_EINVAL, "tolerance must be >= 0");
  const bool want_edges = d_src || d_dst || d_dw;
  if (want_edges && cap < ctx->live_edges)
    return set_err(ctx, CYTOM_ERANGE, "edge outputs hold fewer entries than the live edges");
  int rc;
  if (ctx->jobs_dirty && (rc = build_jobs(ctx))) return rc;
  hipStream_t st = ctx->stream;
  const uint64_t V = ctx->V;
  const uint32_t maxit = max_iterations ? max_iterations : 1000;
  const uint64_t nc_max = std::max<uint64_t>(ctx->n_chunks, 1);
  const uint32_t* n_chunks_dev = static_cast<const uint32_t*>(ctx->g_cbase.p) + ctx->n_tiles;
  const uint64_t* chunks = static_cast<const uint64_t*>(ctx->g_chunks.p);
  const uint32_t CH = ctx->chunk_slots;
  double* acc = nullptr;
  float* lam = nullptr;
  uint32_t* delta = nullptr;
  uint64_t* live = nullptr;
  std::vector<uint32_t> h_delta(maxit);
  auto done = [&](int64_t r) {
    (void)hipStreamSynchronize(st);
    for (void* p : {(void*)acc, (void*)lam, (void*)delta, (void*)live})
      if (p) (void)hipFree(p);
    return r;
  };
  if (hipMalloc((void**)&acc, V * 8) != hipSuccess || hipMalloc((void**)&lam, V * 8) != hipSuccess ||
      hipMalloc((void**)&delta, (size_t)maxit * 4) != hipSuccess ||
      (want_edges && hipMalloc((void**)&live, 2 * (nc_max + 1) * 8) != hipSuccess))
    return done(set_err(ctx, CYTOM_ENOMEM, "adjoint scratch allocation failed"));
  float* lb[2] = {lam, lam + V};
  fill_u64(reinterpret_cast<uint64_t*>(acc), V, 0, st);
  fill_u32(delta, maxit, 0, st);
  if (hipMemcpyAsync(lb[0], d_g, V * 4, hipMemcpyDeviceToDevice, st) != hipSuccess)
    return done(set_err(ctx, CYTOM_EDEVICE, "adjoint copy failed"));
  const uint32_t gb = blocks_of(nc_max * 64, 256);
  const uint32_t vb = blocks_of(V, 256);
  // Sweeps in groups of 8 between host checks.
  uint32_t launched = 0, conv = 0;
  bool converged = false;
  while (!converged && launched < maxit) {
    const uint32_t n = std::min<uint32_t>(8, maxit - launched);
    for (uint32_t k = launched; k < launched + n; ++k) {
      k_pa_accumulate<<<gb, 256, 0, st>>>(chunks, n_chunks_dev, ctx->view(), CH, lb[k & 1], acc,
                                          delta, k, tol);
      k_pa_update<<<vb, 256, 0, st>>>(V, d_g, ctx->outdeg, acc, lb[k & 1], lb[(k + 1) & 1], delta,
                                      k, tol);
    }
    if (hipMemcpyAsync(h_delta.data() + launched, delta + launched, n * 4, hipMemcpyDeviceToHost,
                       st) != hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess)
      return done(set_err(ctx, CYTOM_EDEVICE, "adjoint sweep failed"));
    for (uint32_t k = launched; k < launched + n && !converged; ++k) {
      float d;
      std::memcpy(&d, &h_delta[k], 4);
      if (d <= tol) {
        converged = true;
        conv = k + 1;  // sweeps run, the last one within tol
      }
    }
    launched += n;
  }
  const uint32_t sweeps = converged ? conv : launched;
  const float* lam_final = lb[sweeps & 1];
  if (iterations) *iterations = sweeps;
  if (d_ds) k_pa_teleport<<<vb, 256, 0, st>>>(V, lam_final, d_ds);
  if (want_edges) {
    // mu: one more accumulation of the final lambda (acc is zero after the
    // last update that ran).
    k_pa_accumulate<<<gb, 256, 0, st>>>(chunks, n_chunks_dev, ctx->view(), CH, lam_final, acc,
                                        delta, 0, tol);
    k_pa_live<<<gb, 256, 0, st>>>(chunks, n_chunks_dev, ctx->view(), CH, live);
    if ((rc = ensure(ctx, ctx->g_scan_tmp, scan_tmp_bytes(nc_max + 1) + 256)))
      return done(rc);
    uint64_t* base = live + nc_max + 1;
    scan_exclusive_u64_devn(live, base, n_chunks_dev, nc_max, ctx->g_scan_tmp.p, st);
    k_pa_edges<<<gb, 256, 0, st>>>(chunks, n_chunks_dev, ctx->view(), CH, base, P.rank, lam_final,
                                   acc, ctx->outdeg, cap, d_src, d_dst, d_dw);
  }
  if (hipGetLastError() != hipSuccess || hipStreamSynchronize(st) != hipSuccess)
    return done(set_err(ctx, CYTOM_EDEVICE, "adjoint kernels failed"));
  return done((int64_t)ctx->live_edges);
}

}  // namespace cytom
