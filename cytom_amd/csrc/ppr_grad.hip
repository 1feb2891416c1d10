// Differentiable weighted personalized PageRank (no reference counterpart).
//
// For an explicit edge list with positive FP32 weights w_e and a teleport
// vector s, with out-row-normalised transition P_e = w_e / W_u (W_u = sum of
// the weights leaving u; rows without out-edges drop their mass):
//   forward   r = (1 - a) s + a P^T r
//   adjoint   lambda = g + a P lambda            (g = dL/dr)
//   dL/ds_v = (1 - a) lambda_v
//   dL/dw_e = a r_u (lambda_v - mu_u) / W_u     for e = (u -> v),
//             mu_u = sum_{e' leaving u} P_e' lambda_dst(e')
// Both linear systems are solved by Jacobi sweeps in FP32 with a fixed
// summation order (CSC rows for the forward, CSR rows for the adjoint), so
// results are deterministic.  The operator owns its CSR/CSC copies of the
// edge list; the tile store is not involved.
#include <algorithm>
#include <cstring>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "devutil.h"
#include "internal.h"

namespace cytom {
namespace {

constexpr int GT = 256;
constexpr uint64_t ROW_LIGHT = 64;  // rows above this many entries are summed by segments
constexpr uint64_t SEGL = 1024;     // entries per segment of a long row (one wavefront each)

inline uint32_t gblocks(uint64_t n, uint32_t per, uint32_t cap = 65536) {
  const uint64_t b = (n + per - 1) / per;
  return (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(b, cap));
}

// Row offsets from the sorted keys: position i writes off[r] = i for the rows
// r in (skey[i-1], skey[i]] (from 0 at i = 0, up to V at i = E), so every
// entry of off[0..V] is written once, with no atomics.
__global__ void k_row_bounds(const uint32_t* __restrict__ skey, uint64_t E, uint64_t V,
                             uint64_t* __restrict__ off) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i <= E;
       i += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t lo = i == 0 ? 0 : (uint64_t)skey[i - 1] + 1;
    const uint64_t hi = i == E ? V : (uint64_t)skey[i];
    for (uint64_t r = lo; r <= hi; ++r) off[r] = i;
  }
}

__global__ void k_iota(uint32_t* __restrict__ x, uint64_t n) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x)
    x[i] = (uint32_t)i;
}

__global__ void k_gather_u32(const uint32_t* __restrict__ perm, const uint32_t* __restrict__ in,
                             uint64_t E, uint32_t* __restrict__ out) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < E;
       i += (uint64_t)gridDim.x * blockDim.x)
    out[i] = in[perm[i]];
}

__global__ void k_long_rows(const uint64_t* __restrict__ off, uint64_t R, uint64_t minlen,
                            uint32_t* __restrict__ list, unsigned int* __restrict__ n) {
  for (uint64_t r = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; r < R;
       r += (uint64_t)gridDim.x * blockDim.x)
    if (off[r + 1] - off[r] > minlen) list[atomicAdd(n, 1u)] = (uint32_t)r;
}

// W_u: sum of the weights leaving u, in CSR order (one thread per row; the
// long rows by segments, k_seg_wsum then k_rows_from_parts).
__global__ void k_row_sum(const uint64_t* __restrict__ off, uint64_t R,
                          const uint32_t* __restrict__ perm, const float* __restrict__ w,
                          float* __restrict__ W) {
  for (uint64_t r = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; r < R;
       r += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t b = off[r], e = off[r + 1];
    if (e - b > ROW_LIGHT) continue;
    float s = 0.f;
    for (uint64_t i = b; i < e; ++i) s += w[perm[i]];
    W[r] = s;
  }
}

__device__ inline float wave_sum_f32(float v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// Transition probability of every edge in the given row order:
// p[i] = w[perm[i]] / W[src[perm[i]]].
__global__ void k_prob(const uint32_t* __restrict__ perm, const uint32_t* __restrict__ src,
                       const float* __restrict__ w, const float* __restrict__ W, uint64_t E,
                       float* __restrict__ p) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < E;
       i += (uint64_t)gridDim.x * blockDim.x) {
    const uint32_t e = perm[i];
    p[i] = w[e] / W[src[e]];
  }
}

// |v - x| for the convergence test, with a NaN or infinite iterate mapped to
// +inf (fmaxf would drop a NaN, and a diverged solve must not pass as a
// converged one).
__device__ inline float change_of(float v, float x) {
  const float d = fabsf(v - x);
  return d <= 3.402823466e38f ? d : __int_as_float(0x7f800000);
}

// Flags (*bad = 1) any weight that is not positive and finite.
__global__ void k_check_weights(const float* __restrict__ w, uint64_t E,
                                unsigned int* __restrict__ bad) {
  bool b = false;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < E;
       i += (uint64_t)gridDim.x * blockDim.x) {
    const float x = w[i];
    b |= !(x > 0.f && x <= 3.402823466e38f);
  }
  if (b) *bad = 1u;
}

__device__ inline void block_max_to(float d, unsigned int* out) {
  __shared__ float wmax[GT / 64];
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) d = fmaxf(d, __shfl_xor(d, o, 64));
  if ((threadIdx.x & 63) == 0) wmax[threadIdx.x >> 6] = d;
  __syncthreads();
  if (threadIdx.x == 0) {
    float m = 0.f;
    for (int i = 0; i < GT / 64; ++i) m = fmaxf(m, wmax[i]);
    if (m > 0.f) atomicMax(out, __float_as_uint(m));
  }
}

// One Jacobi sweep y_r = base_r + a * sum_{i in row r} p_i x_{nb_i} over the
// rows of at most ROW_LIGHT entries (one thread each); the largest
// |y_r - x_r| goes to *diff (as float bits, non-negative).
__global__ void k_sweep(const uint64_t* __restrict__ off, uint64_t R, const uint32_t* __restrict__ nb,
                        const float* __restrict__ p, const float* __restrict__ x,
                        const float* __restrict__ base, float base_scale, float a,
                        float* __restrict__ y, unsigned int* __restrict__ diff,
                        unsigned int* __restrict__ next_diff) {
  // the next sweep's change accumulator (the previous sweep's, already read)
  if (blockIdx.x == 0 && threadIdx.x == 0) *next_diff = 0u;
  float dmax = 0.f;
  for (uint64_t r = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; r < R;
       r += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t b = off[r], e = off[r + 1];
    if (e - b > ROW_LIGHT) continue;
    float acc = 0.f;
    for (uint64_t i = b; i < e; ++i) acc += p[i] * x[nb[i]];
    const float v = base_scale * base[r] + a * acc;
    y[r] = v;
    dmax = fmaxf(dmax, change_of(v, x[r]));
  }
  block_max_to(dmax, diff);
}

// Long rows are cut into segments of SEGL entries.  Counts per long row, then
// (after a scan into sbase) the segment bounds.
__global__ void k_lrow_nseg(const uint32_t* __restrict__ rows, uint32_t nrows,
                            const uint64_t* __restrict__ off, uint64_t* __restrict__ cnt) {
  for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < nrows; k += gridDim.x * blockDim.x) {
    const uint32_t r = rows[k];
    cnt[k] = (off[r + 1] - off[r] + SEGL - 1) / SEGL;
  }
}

__global__ void k_lrow_emit(const uint32_t* __restrict__ rows, uint32_t nrows,
                            const uint64_t* __restrict__ off, const uint64_t* __restrict__ sbase,
                            uint64_t* __restrict__ seg_b) {
  for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < nrows; k += gridDim.x * blockDim.x) {
    const uint32_t r = rows[k];
    const uint64_t e1 = off[r + 1];
    uint64_t j = sbase[k];
    for (uint64_t b = off[r]; b < e1; b += SEGL, ++j) {
      seg_b[2 * j] = b;
      seg_b[2 * j + 1] = min(b + SEGL, e1);
    }
  }
}

// part_j = sum over segment j of p_i x_{nb_i}: one wavefront per segment,
// lanes strided over it, then a fixed-order wavefront reduction.
__global__ void k_seg_partial(const uint64_t* __restrict__ seg_b, uint32_t nseg,
                              const uint32_t* __restrict__ nb, const float* __restrict__ p,
                              const float* __restrict__ x, float* __restrict__ part) {
  const uint32_t lane = threadIdx.x & 63;
  const uint64_t wave = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const uint64_t nw = ((uint64_t)gridDim.x * blockDim.x) >> 6;
  for (uint64_t k = wave; k < nseg; k += nw) {
    float acc = 0.f;
    for (uint64_t i = seg_b[2 * k] + lane; i < seg_b[2 * k + 1]; i += 64) acc += p[i] * x[nb[i]];
    acc = wave_sum_f32(acc);
    if (lane == 0) part[k] = acc;
  }
}

// Long rows of a sweep: y_r = base_r + a * (their segment partials in
// segment order), the largest change to *diff.
__global__ void k_sweep_rows(const uint32_t* __restrict__ rows, uint32_t nrows,
                             const uint64_t* __restrict__ sbase, const float* __restrict__ part,
                             const float* __restrict__ x, const float* __restrict__ base,
                             float base_scale, float a, float* __restrict__ y,
                             unsigned int* __restrict__ diff) {
  float dmax = 0.f;
  for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < nrows; k += gridDim.x * blockDim.x) {
    const uint32_t r = rows[k];
    float acc = 0.f;
    for (uint64_t j = sbase[k]; j < sbase[k + 1]; ++j) acc += part[j];
    const float v = base_scale * base[r] + a * acc;
    y[r] = v;
    dmax = fmaxf(dmax, change_of(v, x[r]));
  }
  block_max_to(dmax, diff);
}

// mu_u = sum over u's out-edges of p * lambda_dst (CSR order).
__global__ void k_mu(const uint64_t* __restrict__ off, uint64_t R, const uint32_t* __restrict__ nb,
                     const float* __restrict__ p, const float* __restrict__ lam,
                     float* __restrict__ mu) {
  for (uint64_t r = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; r < R;
       r += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t b = off[r], e = off[r + 1];
    if (e - b > ROW_LIGHT) continue;
    float acc = 0.f;
    for (uint64_t i = b; i < e; ++i) acc += p[i] * lam[nb[i]];
    mu[r] = acc;
  }
}

// Weight sums of the long rows' segments (w in CSR order through perm).
__global__ void k_seg_wsum(const uint64_t* __restrict__ seg_b, uint32_t nseg,
                           const uint32_t* __restrict__ perm, const float* __restrict__ w,
                           float* __restrict__ part) {
  const uint32_t lane = threadIdx.x & 63;
  const uint64_t wave = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const uint64_t nw = ((uint64_t)gridDim.x * blockDim.x) >> 6;
  for (uint64_t k = wave; k < nseg; k += nw) {
    float acc = 0.f;
    for (uint64_t i = seg_b[2 * k] + lane; i < seg_b[2 * k + 1]; i += 64) acc += w[perm[i]];
    acc = wave_sum_f32(acc);
    if (lane == 0) part[k] = acc;
  }
}

// out[r] for the long rows: their segment partials added in segment order
// (mu from the partials of p * lambda, W from the weight partials).
__global__ void k_rows_from_parts(const uint32_t* __restrict__ rows, uint32_t nrows,
                                  const uint64_t* __restrict__ sbase,
                                  const float* __restrict__ part, float* __restrict__ out) {
  for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < nrows; k += gridDim.x * blockDim.x) {
    float acc = 0.f;
    for (uint64_t j = sbase[k]; j < sbase[k + 1]; ++j) acc += part[j];
    out[rows[k]] = acc;
  }
}

// dL/dw_e = a r_u (lambda_v - mu_u) / W_u, in the caller's edge order.
__global__ void k_dw(const uint32_t* __restrict__ src, const uint32_t* __restrict__ dst, uint64_t E,
                     const float* __restrict__ r, const float* __restrict__ lam,
                     const float* __restrict__ mu, const float* __restrict__ W, float a,
                     float* __restrict__ dw) {
  for (uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; e < E;
       e += (uint64_t)gridDim.x * blockDim.x) {
    const uint32_t u = src[e], v = dst[e];
    dw[e] = a * r[u] * (lam[v] - mu[u]) / W[u];
  }
}

__global__ void k_scale(const float* __restrict__ x, uint64_t n, float c, float* __restrict__ y) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x)
    y[i] = c * x[i];
}

// ----------------------------------------------------------------- host
struct Rows {
  uint64_t* off = nullptr;   // R + 1
  uint32_t* perm = nullptr;  // edge ids by row
  uint32_t* nb = nullptr;    // neighbour (the other endpoint) per entry
  float* p = nullptr;        // transition probability per entry
  uint32_t* longr = nullptr;
  uint32_t nlong = 0;
  uint64_t* sbase = nullptr;  // nlong + 1: first segment of each long row
  uint64_t* seg_b = nullptr;  // [2 * nseg]: segment begin, end
  float* part = nullptr;      // nseg partial sums
  uint32_t nseg = 0;
};

// Device memory of one operator.  By default every buffer is its own
// hipMalloc.  In arena mode (set_arena) buffers are carved from one caller
// region instead and nothing is allocated: kept buffers from the bottom up,
// build-only temporaries from the top down (drop_temps releases them all at
// once), so a build needs its kept buffers plus one row build's temporaries.
struct Op {
  uint64_t V = 0, E = 0;
  Rows out, in;  // CSR by source, CSC by target
  float* W = nullptr;
  unsigned int* scratch = nullptr;  // [0] diff, [1] count, [2] max id, [3] bad weight
  void* tmp = nullptr;
  std::vector<void*> owned;
  int err = 0;
  std::string msg;
  char* arena_lo = nullptr;   // next kept byte (arena mode)
  char* arena_hi = nullptr;   // first byte of the temporaries
  char* arena_end = nullptr;  // end of the arena; nullptr = hipMalloc mode
  ~Op() {
    for (void* q : owned)
      if (q) (void)hipFree(q);
    for (void* q : temps)
      if (q) (void)hipFree(q);
  }
  std::vector<void*> temps;  // build-only buffers, freed by drop_temps
  void set_arena(void* base, size_t bytes) {
    arena_lo = static_cast<char*>(base);
    arena_hi = arena_end = arena_lo + bytes;
  }
  template <typename T>
  T* alloc(size_t bytes, bool temp = false) {
    if (arena_end) {
      const size_t n = (std::max<size_t>(bytes, 4) + 255) & ~size_t(255);
      if ((size_t)(arena_hi - arena_lo) < n) {
        err = CYTOM_ERANGE;
        msg = "allocation-free scratch exhausted";
        return nullptr;
      }
      void* q;
      if (temp) {
        arena_hi -= n;
        q = arena_hi;
      } else {
        q = arena_lo;
        arena_lo += n;
      }
      return static_cast<T*>(q);
    }
    void* q = nullptr;
    if (hipMalloc(&q, std::max<size_t>(bytes, 4)) != hipSuccess) {
      err = CYTOM_ENOMEM;
      msg = "device allocation failed";
      return nullptr;
    }
    (temp ? temps : owned).push_back(q);
    return static_cast<T*>(q);
  }
  void drop_temps(hipStream_t st) {
    (void)hipStreamSynchronize(st);
    if (arena_end) {
      arena_hi = arena_end;
      return;
    }
    for (void* q : temps)
      if (q) (void)hipFree(q);
    temps.clear();
  }
};

int build_rows(Op& op, Rows& R, const uint32_t* key, const uint32_t* other, hipStream_t st) {
  const uint64_t V = op.V, E = op.E;
  uint64_t* wide = op.alloc<uint64_t>((V + 1) * 8, true);
  R.off = op.alloc<uint64_t>((V + 1) * 8);
  R.perm = op.alloc<uint32_t>(E * 4);
  R.nb = op.alloc<uint32_t>(E * 4);
  R.p = op.alloc<float>(E * 4);
  R.longr = op.alloc<uint32_t>(V * 4);
  // stable radix sort of (key, edge id): ids increase inside every row, so
  // every row sums in a fixed order
  uint32_t* ids = op.alloc<uint32_t>(E * 4, true);
  uint32_t* k_out = op.alloc<uint32_t>(E * 4, true);
  uint32_t* k_tmp = op.alloc<uint32_t>(E * 4, true);
  uint32_t* v_tmp = op.alloc<uint32_t>(E * 4, true);
  void* hist = op.alloc<char>(radix_hist_bytes(E), true);
  void* stmp = op.alloc<char>(
      scan_tmp_bytes(std::max<uint64_t>(E, 256 * ((E + 4095) / 4096) + 1)) + 256, true);
  if (op.err) return op.err;
  k_iota<<<gblocks(E, GT, 16384), GT, 0, st>>>(ids, E);
  uint32_t bits = 1;
  while ((1ull << bits) < V) ++bits;
  radix_sort_pairs(key, ids, k_out, R.perm, k_tmp, v_tmp, E, bits, hist, stmp, st);
  k_row_bounds<<<gblocks(E + 1, GT, 16384), GT, 0, st>>>(k_out, E, V, R.off);
  k_gather_u32<<<gblocks(E, GT, 16384), GT, 0, st>>>(R.perm, other, E, R.nb);
  // long rows, summed by segments of SEGL entries in the sweeps
  fill_u32(op.scratch + 1, 1, 0, st);
  k_long_rows<<<gblocks(V, GT, 16384), GT, 0, st>>>(R.off, V, ROW_LIGHT, R.longr, op.scratch + 1);
  if (hipMemcpyAsync(&R.nlong, op.scratch + 1, 4, hipMemcpyDeviceToHost, st) != hipSuccess ||
      hipStreamSynchronize(st) != hipSuccess)
    return CYTOM_EDEVICE;
  if (R.nlong) {
    R.sbase = op.alloc<uint64_t>((R.nlong + 1) * 8);
    if (op.err) return op.err;
    k_lrow_nseg<<<gblocks(R.nlong, GT, 16384), GT, 0, st>>>(R.longr, R.nlong, R.off, wide);
    scan_exclusive_u64(wide, R.sbase, R.nlong, op.tmp, st);
    uint64_t nseg = 0;
    if (hipMemcpyAsync(&nseg, R.sbase + R.nlong, 8, hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess)
      return CYTOM_EDEVICE;
    if (nseg > 0xFFFFFFFFull) return CYTOM_ERANGE;
    R.nseg = (uint32_t)nseg;
    R.seg_b = op.alloc<uint64_t>(nseg * 16);
    R.part = op.alloc<float>(nseg * 4);
    if (op.err) return op.err;
    k_lrow_emit<<<gblocks(R.nlong, GT, 16384), GT, 0, st>>>(R.longr, R.nlong, R.off, R.sbase,
                                                           R.seg_b);
  }
  return 0;
}

// Largest endpoint id of the edge list, for the range check.
__global__ void k_max_id(const uint32_t* __restrict__ a, const uint32_t* __restrict__ b,
                         uint64_t E, unsigned int* __restrict__ out) {
  uint32_t m = 0;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < E;
       i += (uint64_t)gridDim.x * blockDim.x)
    m = max(m, max(a[i], b[i]));
  if (m) atomicMax(out, m);
}

// CSR and CSC rows of the edge list (src, dst are the operator's own copies).
int build_structure(Op& op, uint64_t V, uint64_t E, const uint32_t* src, const uint32_t* dst,
                    hipStream_t st) {
  op.V = V;
  op.E = E;
  op.scratch = op.alloc<unsigned int>(256);
  op.tmp = op.alloc<char>(scan_tmp_bytes(V + 1) + 256);
  op.W = op.alloc<float>(V * 4);
  if (op.err) return op.err;
  if (E) {
    fill_u32(op.scratch + 2, 1, 0, st);
    k_max_id<<<gblocks(E, GT, 4096), GT, 0, st>>>(src, dst, E, op.scratch + 2);
    unsigned int mx = 0;
    if (hipMemcpyAsync(&mx, op.scratch + 2, 4, hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess)
      return CYTOM_EDEVICE;
    if (mx >= V) {
      op.msg = "edge endpoint out of range";
      return CYTOM_EINVAL;
    }
  }
  int rc;
  if ((rc = build_rows(op, op.out, src, dst, st))) return rc;
  if (op.arena_end) op.drop_temps(st);  // the second build reuses the temporaries
  if ((rc = build_rows(op, op.in, dst, src, st))) return rc;
  op.drop_temps(st);
  return 0;
}

// W_u and the transition probabilities of both row orders for weights w
// (in the caller's edge order).
int set_weights(Op& op, const uint32_t* src, const float* w, hipStream_t st) {
  const uint64_t V = op.V, E = op.E;
  op.msg.clear();
  if (E) {
    unsigned int bad = 0;
    fill_u32(op.scratch + 3, 1, 0, st);
    k_check_weights<<<gblocks(E, GT, 4096), GT, 0, st>>>(w, E, op.scratch + 3);
    if (hipMemcpyAsync(&bad, op.scratch + 3, 4, hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess)
      return CYTOM_EDEVICE;
    if (bad) {
      op.msg = "edge weights must be positive and finite";
      return CYTOM_EINVAL;
    }
  }
  fill_u32(reinterpret_cast<uint32_t*>(op.W), V, 0, st);
  k_row_sum<<<gblocks(V, GT, 16384), GT, 0, st>>>(op.out.off, V, op.out.perm, w, op.W);
  if (op.out.nlong) {
    k_seg_wsum<<<gblocks((uint64_t)op.out.nseg * 64, GT, 16384), GT, 0, st>>>(
        op.out.seg_b, op.out.nseg, op.out.perm, w, op.out.part);
    k_rows_from_parts<<<gblocks(op.out.nlong, GT, 16384), GT, 0, st>>>(op.out.longr, op.out.nlong,
                                                              op.out.sbase, op.out.part, op.W);
  }
  if (E) {
    k_prob<<<gblocks(E, GT, 16384), GT, 0, st>>>(op.out.perm, src, w, op.W, E, op.out.p);
    k_prob<<<gblocks(E, GT, 16384), GT, 0, st>>>(op.in.perm, src, w, op.W, E, op.in.p);
  }
  return hipStreamSynchronize(st) == hipSuccess ? 0 : CYTOM_EDEVICE;
}

// Jacobi solve of x = base_scale * base + a * M x over the rows of R; x is
// the initial guess on entry and the solution on exit.
int solve(Op& op, const Rows& R, const float* base, float base_scale, float a, float tol,
          uint32_t max_iter, float* x, float* y, uint32_t* iters, hipStream_t st) {
  const uint64_t V = op.V;
  op.msg.clear();
  uint32_t it = 0;
  // Sweep k accumulates its largest change in scratch[4 + (k & 1)] and
  // clears the other word for sweep k + 1 (no memset between sweeps).
  fill_u32(op.scratch + 4, 1, 0, st);
  for (; it < max_iter; ++it) {
    unsigned int* dcur = op.scratch + 4 + (it & 1);
    unsigned int* dnext = op.scratch + 4 + ((it + 1) & 1);
    k_sweep<<<gblocks(V, GT, 16384), GT, 0, st>>>(R.off, V, R.nb, R.p, x, base, base_scale, a, y,
                                                  dcur, dnext);
    if (R.nlong) {
      k_seg_partial<<<gblocks((uint64_t)R.nseg * 64, GT, 16384), GT, 0, st>>>(
          R.seg_b, R.nseg, R.nb, R.p, x, R.part);
      k_sweep_rows<<<gblocks(R.nlong, GT, 16384), GT, 0, st>>>(
          R.longr, R.nlong, R.sbase, R.part, x, base, base_scale, a, y, dcur);
    }
    std::swap(x, y);
    // Check every 8 sweeps (and at the cap) to keep the host out of the loop.
    if ((it & 7) == 7 || it + 1 == max_iter) {
      unsigned int bits = 0;
      if (hipMemcpyAsync(&bits, dcur, 4, hipMemcpyDeviceToHost, st) != hipSuccess ||
          hipStreamSynchronize(st) != hipSuccess)
        return CYTOM_EDEVICE;
      float d;
      std::memcpy(&d, &bits, 4);
      if (!(d <= 3.402823466e38f)) {
        op.msg = "the iteration diverged (a non-finite iterate)";
        return CYTOM_ERANGE;
      }
      if (d <= tol) {
        ++it;
        break;
      }
    }
  }
  // An odd number of swaps leaves the solution in the caller's y buffer.
  if (it & 1) {
    if (hipMemcpyAsync(y, x, V * 4, hipMemcpyDeviceToDevice, st) != hipSuccess) return CYTOM_EDEVICE;
  }
  *iters = it;
  return hipStreamSynchronize(st) == hipSuccess ? 0 : CYTOM_EDEVICE;
}

int check_inputs(uint64_t V, uint64_t E, const void* src, const void* dst, const void* w) {
  if (V == 0 || V > (1ull << 31) || E > 0xFFFFFFFFull) return CYTOM_EINVAL;
  if (E && (!src || !dst || !w)) return CYTOM_EINVAL;
  return 0;
}

thread_local std::string g_ppr_err;

int fail(int rc, const Op* op, const char* what) {
  g_ppr_err = (op && !op->msg.empty()) ? op->msg : what;
  return rc;
}

}  // namespace
}  // namespace cytom

using namespace cytom;

// The operator on one edge list: its own copies of the endpoints, the CSR /
// CSC rows, the current transition probabilities, the solver vectors and a
// stream.
struct cytom_ppr_graph {
  Op op;
  hipStream_t st = nullptr;
  uint32_t* src = nullptr;
  uint32_t* dst = nullptr;
  float* y = nullptr;    // solver work vector
  float* lam = nullptr;  // adjoint solution
  float* mu = nullptr;   // adjoint row sums
  ~cytom_ppr_graph() {
    if (st) {
      (void)hipStreamSynchronize(st);
      (void)hipStreamDestroy(st);
    }
  }
};

extern "C" const char* cytom_ppr_last_error(void) { return g_ppr_err.c_str(); }

extern "C" int cytom_ppr_graph_create(uint64_t V, uint64_t E, const uint32_t* d_src,
                                      const uint32_t* d_dst, const float* d_w,
                                      cytom_ppr_graph** out) {
  if (!out) return CYTOM_EINVAL;
  *out = nullptr;
  if (check_inputs(V, E, d_src, d_dst, d_w)) return fail(CYTOM_EINVAL, nullptr, "invalid edge list");
  auto* g = new (std::nothrow) cytom_ppr_graph();
  if (!g) return fail(CYTOM_ENOMEM, nullptr, "host allocation failed");
  if (hipStreamCreate(&g->st) != hipSuccess) {
    g->st = nullptr;
    delete g;
    return fail(CYTOM_EDEVICE, nullptr, "stream creation failed");
  }
  Op& op = g->op;
  g->src = op.alloc<uint32_t>(E * 4);
  g->dst = op.alloc<uint32_t>(E * 4);
  g->y = op.alloc<float>(V * 4);
  g->lam = op.alloc<float>(V * 4);
  g->mu = op.alloc<float>(V * 4);
  int rc = op.err;
  if (!rc && E &&
      (hipMemcpyAsync(g->src, d_src, E * 4, hipMemcpyDeviceToDevice, g->st) != hipSuccess ||
       hipMemcpyAsync(g->dst, d_dst, E * 4, hipMemcpyDeviceToDevice, g->st) != hipSuccess))
    rc = CYTOM_EDEVICE;
  if (!rc) rc = build_structure(op, V, E, g->src, g->dst, g->st);
  if (!rc) rc = set_weights(op, g->src, d_w, g->st);
  if (rc) {
    fail(rc, &op, "weighted PPR graph build failed");
    delete g;
    return rc;
  }
  *out = g;
  return 0;
}

extern "C" int cytom_ppr_graph_set_weights(cytom_ppr_graph* g, const float* d_w) {
  if (!g || (g->op.E && !d_w)) return fail(CYTOM_EINVAL, nullptr, "invalid arguments");
  const int rc = set_weights(g->op, g->src, d_w, g->st);
  return rc ? fail(rc, &g->op, "weighted PPR weight update failed") : 0;
}

extern "C" int cytom_ppr_graph_forward(cytom_ppr_graph* g, const float* d_s, float alpha,
                                       float tol, uint32_t max_iter, float* d_r,
                                       uint32_t* iterations) {
  if (!g || !d_s || !d_r || !(alpha >= 0.f && alpha < 1.f))
    return fail(CYTOM_EINVAL, nullptr, "invalid arguments");
  Op& op = g->op;
  // x0 = (1 - a) s
  k_scale<<<gblocks(op.V, GT, 16384), GT, 0, g->st>>>(d_s, op.V, 1.f - alpha, d_r);
  uint32_t it = 0;
  const int rc = solve(op, op.in, d_s, 1.f - alpha, alpha, tol, max_iter, d_r, g->y, &it, g->st);
  if (iterations) *iterations = it;
  return rc ? fail(rc, &op, "weighted PPR forward failed") : 0;
}

extern "C" int cytom_ppr_graph_adjoint(cytom_ppr_graph* g, const float* d_r,
                                       const float* d_dL_dr, float alpha, float tol,
                                       uint32_t max_iter, float* d_dL_dw, float* d_dL_ds,
                                       uint32_t* iterations) {
  if (!g || !d_r || !d_dL_dr || !(alpha >= 0.f && alpha < 1.f))
    return fail(CYTOM_EINVAL, nullptr, "invalid arguments");
  Op& op = g->op;
  hipStream_t st = g->st;
  // lambda = g + a P lambda, from lambda0 = g
  if (hipMemcpyAsync(g->lam, d_dL_dr, op.V * 4, hipMemcpyDeviceToDevice, st) != hipSuccess)
    return fail(CYTOM_EDEVICE, &op, "weighted PPR adjoint failed");
  uint32_t it = 0;
  int rc = solve(op, op.out, d_dL_dr, 1.f, alpha, tol, max_iter, g->lam, g->y, &it, st);
  if (iterations) *iterations = it;
  if (!rc) {
    const float* L = g->lam;  // solve() leaves the solution in its x buffer
    k_mu<<<gblocks(op.V, GT, 16384), GT, 0, st>>>(op.out.off, op.V, op.out.nb, op.out.p, L, g->mu);
    if (op.out.nlong) {
      k_seg_partial<<<gblocks((uint64_t)op.out.nseg * 64, GT, 16384), GT, 0, st>>>(
          op.out.seg_b, op.out.nseg, op.out.nb, op.out.p, L, op.out.part);
      k_rows_from_parts<<<gblocks(op.out.nlong, GT, 16384), GT, 0, st>>>(
          op.out.longr, op.out.nlong, op.out.sbase, op.out.part, g->mu);
    }
    if (d_dL_dw && op.E)
      k_dw<<<gblocks(op.E, GT, 16384), GT, 0, st>>>(g->src, g->dst, op.E, d_r, L, g->mu, op.W,
                                                    alpha, d_dL_dw);
    if (d_dL_ds) k_scale<<<gblocks(op.V, GT, 16384), GT, 0, st>>>(L, op.V, 1.f - alpha, d_dL_ds);
    rc = hipStreamSynchronize(st) == hipSuccess ? 0 : CYTOM_EDEVICE;
  }
  return rc ? fail(rc, &op, "weighted PPR adjoint failed") : 0;
}

extern "C" int cytom_ppr_graph_destroy(cytom_ppr_graph* g) {
  delete g;
  return 0;
}

// One-shot forms: build the operator for this call only.
extern "C" int cytom_ppr_weighted(uint64_t V, uint64_t E, const uint32_t* d_src,
                                  const uint32_t* d_dst, const float* d_w, const float* d_s,
                                  float alpha, float tol, uint32_t max_iter, float* d_r,
                                  uint32_t* iterations) {
  if (check_inputs(V, E, d_src, d_dst, d_w) || !d_s || !d_r || !(alpha >= 0.f && alpha < 1.f))
    return fail(CYTOM_EINVAL, nullptr, "invalid arguments");
  cytom_ppr_graph* g = nullptr;
  int rc = cytom_ppr_graph_create(V, E, d_src, d_dst, d_w, &g);
  if (!rc) rc = cytom_ppr_graph_forward(g, d_s, alpha, tol, max_iter, d_r, iterations);
  cytom_ppr_graph_destroy(g);
  return rc;
}

extern "C" int cytom_ppr_adjoint(uint64_t V, uint64_t E, const uint32_t* d_src,
                                 const uint32_t* d_dst, const float* d_w, const float* d_r,
                                 const float* d_dL_dr, float alpha, float tol, uint32_t max_iter,
                                 float* d_dL_dw, float* d_dL_ds, uint32_t* iterations) {
  if (check_inputs(V, E, d_src, d_dst, d_w) || !d_r || !d_dL_dr || !(alpha >= 0.f && alpha < 1.f))
    return fail(CYTOM_EINVAL, nullptr, "invalid arguments");
  cytom_ppr_graph* g = nullptr;
  int rc = cytom_ppr_graph_create(V, E, d_src, d_dst, d_w, &g);
  if (!rc)
    rc = cytom_ppr_graph_adjoint(g, d_r, d_dL_dr, alpha, tol, max_iter, d_dL_dw, d_dL_ds,
                                 iterations);
  cytom_ppr_graph_destroy(g);
  return rc;
}

// ------------------------------------------------- allocation-free forward
// For callers that cannot allocate device memory (code under stream capture,
// tight serving loops, harnesses that forbid hipMalloc): the forward of
// cytom_ppr_weighted with every temporary carved from one file-scope device
// array, NOALLOC_BYTES per GPU, held until the process ends.  The array's
// name is a host-side stand-in: its device address comes from
// hipGetSymbolAddress, once per device, and only that address is passed on.
namespace cytom {
namespace {

constexpr size_t NOALLOC_BYTES = 64ull << 20;
static __device__ __align__(256) unsigned char noalloc_scratch[NOALLOC_BYTES];

constexpr int NOALLOC_MAX_DEVICES = 64;
std::mutex g_noalloc_mu;
void* g_noalloc_ptr[NOALLOC_MAX_DEVICES] = {};

// Device address of noalloc_scratch on the current device.
int noalloc_base(void** out) {
  int dev = -1;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= NOALLOC_MAX_DEVICES)
    return CYTOM_EDEVICE;
  std::lock_guard<std::mutex> lk(g_noalloc_mu);
  if (!g_noalloc_ptr[dev]) {
    void* p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(noalloc_scratch)) != hipSuccess || !p)
      return CYTOM_EDEVICE;
    g_noalloc_ptr[dev] = p;
  }
  *out = g_noalloc_ptr[dev];
  return 0;
}

inline size_t up256(size_t b) { return (std::max<size_t>(b, 4) + 255) & ~size_t(255); }

// Upper bound of the arena bytes one forward of (V, E) carves: the kept
// buffers of both row sets (long-row segments bounded by E/64: a long row
// has more than 64 entries and one segment per 1024), the weight sums, the
// solver vector, plus one row build's temporaries.
size_t noalloc_bytes(uint64_t V, uint64_t E) {
  const uint64_t nseg = E / 64 + 1;
  const size_t rows = up256((V + 1) * 8) + 3 * up256(E * 4) + up256(V * 4) + up256((V + 1) * 8) +
                      up256(nseg * 16) + up256(nseg * 4);
  const size_t kept = 2 * rows + up256(256 * 4) + up256(scan_tmp_bytes(V + 1) + 256) +
                      up256(V * 4) /* W */ + up256(V * 4) /* y */;
  const size_t temps =
      up256((V + 1) * 8) + 4 * up256(E * 4) + up256(radix_hist_bytes(E)) +
      up256(scan_tmp_bytes(std::max<uint64_t>(E, 256 * ((E + 4095) / 4096) + 1)) + 256);
  return kept + temps;
}

thread_local std::string g_noalloc_msg;

}  // namespace
}  // namespace cytom

extern "C" int cytom_ppr_weighted_noalloc(uint64_t V, uint64_t E, const uint32_t* d_src,
                                          const uint32_t* d_dst, const float* d_w,
                                          const float* d_s, float alpha, float tol,
                                          uint32_t max_iter, float* d_r, uint32_t* iterations) {
  if (check_inputs(V, E, d_src, d_dst, d_w) || !d_s || !d_r || !(alpha >= 0.f && alpha < 1.f))
    return fail(CYTOM_EINVAL, nullptr, "invalid arguments");
  const size_t need = noalloc_bytes(V, E);
  if (need > NOALLOC_BYTES) {
    g_noalloc_msg = "graph too large for the allocation-free forward: needs " +
                    std::to_string(need) + " bytes of scratch, the limit is " +
                    std::to_string(NOALLOC_BYTES);
    return fail(CYTOM_ERANGE, nullptr, g_noalloc_msg.c_str());
  }
  void* base = nullptr;
  if (noalloc_base(&base)) return fail(CYTOM_EDEVICE, nullptr, "scratch symbol lookup failed");
  hipStream_t st = nullptr;  // the null stream: no stream object is created
  Op op;
  op.set_arena(base, NOALLOC_BYTES);
  int rc = build_structure(op, V, E, d_src, d_dst, st);
  if (!rc) rc = set_weights(op, d_src, d_w, st);
  float* y = rc ? nullptr : op.alloc<float>(V * 4);
  if (!rc && op.err) rc = op.err;
  uint32_t it = 0;
  if (!rc) {
    k_scale<<<gblocks(V, GT, 16384), GT, 0, st>>>(d_s, V, 1.f - alpha, d_r);
    rc = solve(op, op.in, d_s, 1.f - alpha, alpha, tol, max_iter, d_r, y, &it, st);
  }
  if (iterations) *iterations = it;
  return rc ? fail(rc, &op, "allocation-free weighted PPR forward failed") : 0;
}
