// Host-record ingest through pinned staging, and vertex results in the
// reference's on-disk vertex layouts.
//
// Record ingest follows the reference's binary inputs: file_edge_t
// {int32 src, int32 tgt} (util/datatypes.h:41-44, read by core/jobs.h:52-75)
// and the edge logs' edge_compact_t / edge_compact_weighted_t
// (util/datatypes.h:55-64, tile-manager/edges-writer.h:114-185).  Records are
// copied from the caller's (typically mmap'ed) pages into two pinned staging
// buffers in turn, so the CPU copy of one chunk overlaps the DMA of the
// previous one, then split into the ingest arrays on the device.
//
// Results follow core/results-writer.h:46-81 (raw VertexType[V] of the next
// array): u32 for BFS/CC, float for SSSP (FLT_MAX unreached, examples/sssp.cc:21),
// {delta_sum, rank, delta} floats for PageRank-delta
// (examples/pagerank-delta.cc:30-33).
#include <cfloat>
#include <cmath>
#include <cstring>

#include "internal.h"

namespace cytom {
namespace {

constexpr size_t STAGE_BYTES = size_t(8) << 20;  // per pinned staging buffer

inline uint32_t grid_for(uint64_t n, uint32_t threads) {
  return (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>((n + threads - 1) / threads, 65536));
}

// Weight of a weighted record: integral floats in [1, 2^24] convert exactly;
// anything else becomes 0, which the ingest validation rejects.
__device__ inline uint32_t record_weight(uint32_t bits) {
  const float f = __uint_as_float(bits);
  return (f >= 1.0f && f <= 16777216.0f && f == truncf(f)) ? (uint32_t)f : 0u;
}

// rec holds n records of `words` u32 words: src, tgt[, f32 weight].
__global__ void k_split_records(const uint32_t* __restrict__ rec, uint32_t words, uint64_t n,
                                uint32_t default_w, uint32_t* __restrict__ src,
                                uint32_t* __restrict__ dst, uint32_t* __restrict__ w) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x) {
    const uint32_t* r = rec + i * words;
    src[i] = r[0];
    dst[i] = r[1];
    if (w) w[i] = words == 3 ? record_weight(r[2]) : default_w;
  }
}

__global__ void k_results_u32(const uint32_t* __restrict__ cur, uint64_t V, int as_float,
                              uint32_t* __restrict__ out) {
  for (uint64_t v = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; v < V;
       v += (uint64_t)gridDim.x * blockDim.x) {
    const uint32_t x = cur[v];
    out[v] = as_float ? __float_as_uint(x == INF ? FLT_MAX : (float)x) : x;
  }
}

__global__ void k_results_prd(const long long* __restrict__ dsum, const float* __restrict__ rank,
                              const float* __restrict__ delta, uint64_t V, float inv_scale,
                              float* __restrict__ out) {
  for (uint64_t v = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; v < V;
       v += (uint64_t)gridDim.x * blockDim.x) {
    out[3 * v + 0] = (float)((double)dsum[v] * (double)inv_scale);
    out[3 * v + 1] = rank[v];
    out[3 * v + 2] = delta[v];
  }
}

// Both pinned buffers and both events, or none of them (a partial failure is
// released so the next call tries again).
int ensure_stage(cytom_ctx* ctx) {
  if (ctx->h_stage[0] && ctx->h_stage[1] && ctx->stage_ev[0] && ctx->stage_ev[1]) return 0;
  io_free(ctx);
  for (int b = 0; b < 2; ++b) {
    if (hipHostMalloc(&ctx->h_stage[b], STAGE_BYTES, hipHostMallocDefault) != hipSuccess) {
      ctx->h_stage[b] = nullptr;
      io_free(ctx);
      return set_err(ctx, CYTOM_ENOMEM, "pinned staging allocation failed");
    }
    const hipError_t e = hipEventCreateWithFlags(&ctx->stage_ev[b], hipEventDisableTiming);
    if (e != hipSuccess) {
      ctx->stage_ev[b] = nullptr;
      io_free(ctx);
      return hip_fail(ctx, e, "staging event");
    }
  }
  ctx->stage_busy[0] = ctx->stage_busy[1] = false;
  ctx->stage_next = 0;
  return 0;
}

// Copies `bytes` host bytes to device memory through the pinned buffers.
int stage_copy(cytom_ctx* ctx, void* d_dst, const void* h_src, size_t bytes) {
  hipStream_t st = ctx->stream;
  const char* s = static_cast<const char*>(h_src);
  char* d = static_cast<char*>(d_dst);
  for (size_t off = 0; off < bytes; off += STAGE_BYTES) {
    const int b = ctx->stage_next;
    ctx->stage_next ^= 1;
    const size_t n = std::min(STAGE_BYTES, bytes - off);
    if (ctx->stage_busy[b]) CY_HIP(ctx, hipEventSynchronize(ctx->stage_ev[b]));
    std::memcpy(ctx->h_stage[b], s + off, n);
    CY_HIP(ctx, hipMemcpyAsync(d + off, ctx->h_stage[b], n, hipMemcpyHostToDevice, st));
    CY_HIP(ctx, hipEventRecord(ctx->stage_ev[b], st));
    ctx->stage_busy[b] = true;
  }
  return 0;
}

}  // namespace

int ingest_records(cytom_ctx* ctx, const void* rec, uint32_t rec_bytes, uint64_t n_ins,
                   uint32_t default_w, const void* del_rec, uint64_t n_del,
                   cytom_batch_stats* stats) {
  if (rec_bytes != 8 && rec_bytes != 12)
    return set_err(ctx, CYTOM_EINVAL, "record size must be 8 (src, tgt) or 12 (src, tgt, weight)");
  if ((n_ins && !rec) || (n_del && !del_rec)) return set_err(ctx, CYTOM_EINVAL, "null records");
  if (ctx->weighted && rec_bytes == 8 && default_w == 0)
    return set_err(ctx, CYTOM_EINVAL, "weighted graph needs a nonzero default weight");
  if (n_ins > 0xFFFFFFFEull || n_del > 0xFFFFFFFEull)
    return set_err(ctx, CYTOM_ERANGE, "batch larger than 2^32-2 updates");
  int rc;
  const uint32_t words = rec_bytes / 4;
  const uint64_t ni = std::max<uint64_t>(n_ins, 1), nd = std::max<uint64_t>(n_del, 1);
  if ((rc = ensure_stage(ctx)) || (rc = ensure(ctx, ctx->g_rec, (ni + nd) * rec_bytes)) ||
      (rc = ensure(ctx, ctx->g_in_src, ni * 4)) || (rc = ensure(ctx, ctx->g_in_dst, ni * 4)) ||
      (rc = ensure(ctx, ctx->g_in_w, ni * 4)) || (rc = ensure(ctx, ctx->g_del_src, nd * 4)) ||
      (rc = ensure(ctx, ctx->g_del_dst, nd * 4)) || (rc = ensure(ctx, ctx->g_del_w, nd * 4)))
    return rc;
  hipStream_t st = ctx->stream;
  uint32_t* d_rec = static_cast<uint32_t*>(ctx->g_rec.p);
  uint32_t* d_del = d_rec + ni * words;
  uint32_t* ds = static_cast<uint32_t*>(ctx->g_in_src.p);
  uint32_t* dd = static_cast<uint32_t*>(ctx->g_in_dst.p);
  uint32_t* dw = ctx->weighted ? static_cast<uint32_t*>(ctx->g_in_w.p) : nullptr;
  uint32_t* xs = static_cast<uint32_t*>(ctx->g_del_src.p);
  uint32_t* xd = static_cast<uint32_t*>(ctx->g_del_dst.p);
  // Deletions match any weight unless the records carry one.
  uint32_t* xw = (ctx->weighted && words == 3) ? static_cast<uint32_t*>(ctx->g_del_w.p) : nullptr;
  if (n_ins) {
    if ((rc = stage_copy(ctx, d_rec, rec, n_ins * rec_bytes))) return rc;
    k_split_records<<<grid_for(n_ins, 256), 256, 0, st>>>(d_rec, words, n_ins, default_w, ds, dd,
                                                          dw);
  }
  if (n_del) {
    if ((rc = stage_copy(ctx, d_del, del_rec, n_del * rec_bytes))) return rc;
    k_split_records<<<grid_for(n_del, 256), 256, 0, st>>>(d_del, words, n_del, 1u, xs, xd, xw);
  }
  return ingest_batch(ctx, n_ins ? ds : nullptr, n_ins ? dd : nullptr, n_ins ? dw : nullptr,
                      n_ins, n_del ? xs : nullptr, n_del ? xd : nullptr, xw, n_del, stats);
}

int64_t read_results(cytom_ctx* ctx, const Program& P, void* out, size_t bytes) {
  const uint64_t V = ctx->V;
  const size_t rec = prd_like(P.kind) ? 12 : 4;
  if (!out) return (int64_t)(V * rec);
  if (bytes < V * rec) return set_err(ctx, CYTOM_EINVAL, "output buffer too small");
  int rc;
  if ((rc = ensure(ctx, ctx->g_rec, V * rec))) return rc;
  hipStream_t st = ctx->stream;
  if (prd_like(P.kind)) {
    const float inv = std::ldexp(1.0f, -ctx->prd_shift);
    k_results_prd<<<grid_for(V, 256), 256, 0, st>>>(P.dsum, P.rank, P.delta, V, inv,
                                                    static_cast<float*>(ctx->g_rec.p));
  } else {
    k_results_u32<<<grid_for(V, 256), 256, 0, st>>>(P.cur, V, P.kind == CYTOM_SSSP ? 1 : 0,
                                                    static_cast<uint32_t*>(ctx->g_rec.p));
  }
  CY_HIP(ctx, hipMemcpyAsync(out, ctx->g_rec.p, V * rec, hipMemcpyDeviceToHost, st));
  CY_HIP(ctx, hipStreamSynchronize(st));
  return (int64_t)(V * rec);
}

void io_free(cytom_ctx* ctx) {
  for (int b = 0; b < 2; ++b) {
    if (ctx->h_stage[b]) (void)hipHostFree(ctx->h_stage[b]);
    if (ctx->stage_ev[b]) (void)hipEventDestroy(ctx->stage_ev[b]);
    ctx->h_stage[b] = nullptr;
    ctx->stage_ev[b] = nullptr;
  }
}

}  // namespace cytom
