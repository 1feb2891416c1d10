// Device primitives: exclusive scan and stable LSD radix sort (gfx950,
// 64-lane wavefronts).  Used by ingestion (sort of updates by tile key,
// segment heads) and by the executor work-list builder.
#include <algorithm>

#include "internal.h"
#include "devutil.h"

namespace cytom {
namespace {

// n_dev (nullable): the length, read from device memory, when it is known
// only on the device (n is then the launch bound).
__device__ inline uint64_t dev_len(uint64_t n, const uint32_t* n_dev) {
  return n_dev ? min(n, (uint64_t)*n_dev) : n;
}


__global__ void k_copy2_u32(const uint32_t* __restrict__ a, const uint32_t* __restrict__ b,
                            uint64_t n, const uint32_t* __restrict__ n_dev,
                            uint32_t* __restrict__ ao, uint32_t* __restrict__ bo) {
  n = dev_len(n, n_dev);
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x) {
    ao[i] = a[i];
    bo[i] = b[i];
  }
}

template <typename T>
__global__ void k_fill(T* __restrict__ p, uint64_t n, T v) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x)
    p[i] = v;
}

// Several fills in one launch: blockIdx.y picks the fill.
__global__ void k_fill_many(FillSet fs) {
  const FillOp& f = fs.op[blockIdx.y];
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < f.n;
       i += (uint64_t)gridDim.x * blockDim.x) {
    if (f.wide)
      static_cast<uint64_t*>(f.p)[i] = f.v;
    else
      static_cast<uint32_t*>(f.p)[i] = (uint32_t)f.v;
  }
}


constexpr int SCAN_THREADS = 256;
constexpr int SCAN_ITEMS = 8;
constexpr uint64_t SCAN_TILE = SCAN_THREADS * SCAN_ITEMS;  // 2048

inline size_t align256(size_t b) { return (b + 255) & ~size_t(255); }

// n_dev, when given, caps the length at run time (the launch is sized for n):
// elements from *n_dev on count as zero and are not written.
template <typename T>
__global__ void __launch_bounds__(SCAN_THREADS) k_scan_reduce(const T* __restrict__ in,
                                                              uint64_t n, T* __restrict__ bsum,
                                                              const uint32_t* __restrict__ n_dev) {
  if (n_dev) n = std::min<uint64_t>(n, *n_dev);
  const uint64_t base = (uint64_t)blockIdx.x * SCAN_TILE + (uint64_t)threadIdx.x * SCAN_ITEMS;
  T s = 0;
#pragma unroll
  for (int k = 0; k < SCAN_ITEMS; ++k) {
    uint64_t i = base + k;
    if (i < n) s += in[i];
  }
  T tot = block_sum<SCAN_THREADS, T>(s);
  if (threadIdx.x == 0) bsum[blockIdx.x] = tot;
}

template <typename T>
__global__ void __launch_bounds__(SCAN_THREADS) k_scan_down(const T* __restrict__ in,
                                                            T* __restrict__ out, uint64_t n,
                                                            const T* __restrict__ boff,
                                                            const uint32_t* __restrict__ n_dev) {
  if (n_dev) n = std::min<uint64_t>(n, *n_dev);
  const uint64_t base = (uint64_t)blockIdx.x * SCAN_TILE + (uint64_t)threadIdx.x * SCAN_ITEMS;
  T v[SCAN_ITEMS];
  T s = 0;
#pragma unroll
  for (int k = 0; k < SCAN_ITEMS; ++k) {
    uint64_t i = base + k;
    v[k] = i < n ? in[i] : T(0);
    s += v[k];
  }
  T tot;
  T p = block_scan_excl<SCAN_THREADS, T>(s, &tot);
  const T off = boff ? boff[blockIdx.x] : T(0);
  p += off;
#pragma unroll
  for (int k = 0; k < SCAN_ITEMS; ++k) {
    uint64_t i = base + k;
    if (i < n) out[i] = p;
    p += v[k];
  }
  // The block holding position n (or the last block) writes the total: the
  // elements from n on are zero, so its block total is the prefix at n.
  if (blockIdx.x == std::min<uint64_t>(n / SCAN_TILE, gridDim.x - 1) && threadIdx.x == 0)
    out[n] = off + tot;
}

// Down-sweep that finds its block's offset itself: the sum of the block
// sums before it (at most SCAN_FOLD_MAX of them, read by the whole block),
// so a scan of up to SCAN_FOLD_MAX tiles takes two launches instead of three.
constexpr uint64_t SCAN_FOLD_MAX = 8192;
template <typename T>
__global__ void __launch_bounds__(SCAN_THREADS) k_scan_down_fold(const T* __restrict__ in,
                                                                 T* __restrict__ out, uint64_t n,
                                                                 const T* __restrict__ bsum,
                                                                 const uint32_t* __restrict__ n_dev) {
  T part = 0;
  for (uint32_t b = threadIdx.x; b < blockIdx.x; b += SCAN_THREADS) part += bsum[b];
  const T off = block_sum<SCAN_THREADS, T>(part);
  if (n_dev) n = std::min<uint64_t>(n, *n_dev);
  const uint64_t base = (uint64_t)blockIdx.x * SCAN_TILE + (uint64_t)threadIdx.x * SCAN_ITEMS;
  T v[SCAN_ITEMS];
  T s = 0;
#pragma unroll
  for (int k = 0; k < SCAN_ITEMS; ++k) {
    uint64_t i = base + k;
    v[k] = i < n ? in[i] : T(0);
    s += v[k];
  }
  T tot;
  T p = block_scan_excl<SCAN_THREADS, T>(s, &tot) + off;
#pragma unroll
  for (int k = 0; k < SCAN_ITEMS; ++k) {
    uint64_t i = base + k;
    if (i < n) out[i] = p;
    p += v[k];
  }
  if (blockIdx.x == std::min<uint64_t>(n / SCAN_TILE, gridDim.x - 1) && threadIdx.x == 0)
    out[n] = off + tot;
}

template <typename T>
void scan_impl(const T* in, T* out, uint64_t n, void* tmp, hipStream_t st,
               const uint32_t* n_dev = nullptr) {
  if (n == 0) {
    k_fill<T><<<1, 64, 0, st>>>(out, 1, T(0));
    return;
  }
  if (n <= SCAN_TILE) {
    k_scan_down<T><<<1, SCAN_THREADS, 0, st>>>(in, out, n, nullptr, n_dev);
    return;
  }
  uint64_t nb = (n + SCAN_TILE - 1) / SCAN_TILE;
  T* bsum = static_cast<T*>(tmp);
  T* boff = reinterpret_cast<T*>(static_cast<char*>(tmp) + align256(nb * sizeof(T)));
  void* rest =
      static_cast<char*>(tmp) + align256(nb * sizeof(T)) + align256((nb + 1) * sizeof(T));
  k_scan_reduce<T><<<(uint32_t)nb, SCAN_THREADS, 0, st>>>(in, n, bsum, n_dev);
  if (nb <= SCAN_FOLD_MAX) {
    k_scan_down_fold<T><<<(uint32_t)nb, SCAN_THREADS, 0, st>>>(in, out, n, bsum, n_dev);
    return;
  }
  scan_impl<T>(bsum, boff, nb, rest, st);  // block sums past *n_dev are zero
  k_scan_down<T><<<(uint32_t)nb, SCAN_THREADS, 0, st>>>(in, out, n, boff, n_dev);
}

// ------------------------------------------------------------ radix sort
constexpr int RS_THREADS = 256;
constexpr int RS_ROUNDS = 16;
constexpr uint64_t RS_TILE = RS_THREADS * RS_ROUNDS;  // 4096 keys per block

// Blocks of a pass: with a device length, only the ones holding keys (the
// histogram is laid out for that count, so its scan and the scatter read
// 256 x that many entries; the launch is sized for the bound).
__device__ inline uint32_t pass_blocks(uint64_t n, uint32_t nb, const uint32_t* n_dev) {
  return n_dev ? (uint32_t)((dev_len(n, n_dev) + RS_TILE - 1) / RS_TILE) : nb;
}

__global__ void __launch_bounds__(RS_THREADS) k_radix_hist(const uint32_t* __restrict__ keys,
                                                           uint64_t n, uint32_t shift,
                                                           uint32_t* __restrict__ hist,
                                                           uint32_t nb,
                                                           const uint32_t* __restrict__ n_dev,
                                                           uint32_t* __restrict__ hist_len) {
  nb = pass_blocks(n, nb, n_dev);
  if (hist_len && blockIdx.x == 0 && threadIdx.x == 0) *hist_len = 256u * nb;
  if (blockIdx.x >= nb) return;  // uniform
  n = dev_len(n, n_dev);
  __shared__ uint32_t h[256];
  h[threadIdx.x] = 0;
  __syncthreads();
  const uint64_t base = (uint64_t)blockIdx.x * RS_TILE;
  for (int r = 0; r < RS_ROUNDS; ++r) {
    uint64_t i = base + (uint64_t)r * RS_THREADS + threadIdx.x;
    if (i < n) atomicAdd(&h[(keys[i] >> shift) & 255u], 1u);
  }
  __syncthreads();
  hist[(uint64_t)threadIdx.x * nb + blockIdx.x] = h[threadIdx.x];
}

// Stable scatter: ranks inside a round of 256 keys come from wavefront
// ballots (lanes with equal digits), ordered by wave then lane.
__global__ void __launch_bounds__(RS_THREADS) k_radix_scatter(
    const uint32_t* __restrict__ kin, const uint32_t* __restrict__ vin, uint64_t n,
    uint32_t shift, const uint32_t* __restrict__ hscan, uint32_t nb,
    uint32_t* __restrict__ kout, uint32_t* __restrict__ vout, const uint32_t* __restrict__ n_dev) {
  nb = pass_blocks(n, nb, n_dev);
  if (blockIdx.x >= nb) return;  // uniform
  n = dev_len(n, n_dev);
  __shared__ uint32_t run[256];
  __shared__ uint32_t wcnt[RS_THREADS / 64][256];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  run[tid] = hscan[(uint64_t)tid * nb + blockIdx.x];
  const uint64_t base = (uint64_t)blockIdx.x * RS_TILE;
  const uint64_t lt = (1ull << lane) - 1ull;
  for (int r = 0; r < RS_ROUNDS; ++r) {
    const uint64_t i = base + (uint64_t)r * RS_THREADS + tid;
    if (base + (uint64_t)r * RS_THREADS >= n) break;  // uniform across the block
    const bool valid = i < n;
    const uint32_t k = valid ? kin[i] : 0u;
    const uint32_t v = valid ? vin[i] : 0u;
    const uint32_t d = (k >> shift) & 255u;
#pragma unroll
    for (int w = 0; w < RS_THREADS / 64; ++w) wcnt[w][tid] = 0;
    __syncthreads();
    uint64_t peers = __ballot(valid);
#pragma unroll
    for (int b = 0; b < 8; ++b) {
      const bool bit = (d >> b) & 1u;
      const uint64_t m = __ballot(bit);
      peers &= bit ? m : ~m;
    }
    const uint32_t rank = (uint32_t)__popcll(peers & lt);
    if (valid && rank == 0) wcnt[wave][d] = (uint32_t)__popcll(peers);
    __syncthreads();
    if (valid) {
      uint32_t pos = run[d] + rank;
      for (int w = 0; w < wave; ++w) pos += wcnt[w][d];
      kout[pos] = k;
      vout[pos] = v;
    }
    __syncthreads();
    uint32_t add = 0;
#pragma unroll
    for (int w = 0; w < RS_THREADS / 64; ++w) add += wcnt[w][tid];
    run[tid] += add;
    __syncthreads();
  }
}

void radix_sort_impl(const uint32_t* k_in, const uint32_t* v_in, uint32_t* k_out,
                     uint32_t* v_out, uint32_t* k_tmp, uint32_t* v_tmp, uint64_t n,
                     const uint32_t* n_dev, uint32_t key_bits, void* hist, void* scan_tmp,
                     hipStream_t st) {
  if (n == 0) return;
  const uint32_t passes = (key_bits + 7) / 8;
  if (passes == 0) {
    const uint32_t g = (uint32_t)std::min<uint64_t>((n + 255) / 256, 16384);
    k_copy2_u32<<<g, 256, 0, st>>>(k_in, v_in, n, n_dev, k_out, v_out);
    return;
  }
  const uint32_t nb = (uint32_t)((n + RS_TILE - 1) / RS_TILE);
  uint32_t* h = static_cast<uint32_t*>(hist);
  uint32_t* hs = reinterpret_cast<uint32_t*>(static_cast<char*>(hist) + align256(256ull * nb * 4));
  // the histogram length of a device-length sort (256 x its blocks)
  uint32_t* hlen = n_dev ? reinterpret_cast<uint32_t*>(reinterpret_cast<char*>(hs) +
                                                       align256((256ull * nb + 1) * 4))
                         : nullptr;
  const uint32_t* ks = k_in;
  const uint32_t* vs = v_in;
  for (uint32_t p = 0; p < passes; ++p) {
    const bool last_to_out = ((passes - p) & 1u) == 1u;
    uint32_t* kd = last_to_out ? k_out : k_tmp;
    uint32_t* vd = last_to_out ? v_out : v_tmp;
    k_radix_hist<<<nb, RS_THREADS, 0, st>>>(ks, n, 8 * p, h, nb, n_dev, hlen);
    if (n_dev)
      scan_impl<uint32_t>(h, hs, 256ull * nb, scan_tmp, st, hlen);
    else
      scan_exclusive_u32(h, hs, 256ull * nb, scan_tmp, st);
    k_radix_scatter<<<nb, RS_THREADS, 0, st>>>(ks, vs, n, 8 * p, hs, nb, kd, vd, n_dev);
    ks = kd;
    vs = vd;
  }
}


}  // namespace

size_t scan_tmp_bytes(uint64_t n) {
  if (n <= SCAN_TILE) return 0;
  uint64_t nb = (n + SCAN_TILE - 1) / SCAN_TILE;
  return align256(nb * 8) + align256((nb + 1) * 8) + scan_tmp_bytes(nb);
}

void scan_exclusive_u32(const uint32_t* in, uint32_t* out, uint64_t n, void* tmp,
                        hipStream_t st) {
  scan_impl<uint32_t>(in, out, n, tmp, st);
}

void scan_exclusive_u64(const uint64_t* in, uint64_t* out, uint64_t n, void* tmp,
                        hipStream_t st) {
  scan_impl<uint64_t>(in, out, n, tmp, st);
}

void scan_exclusive_u64_devn(const uint64_t* in, uint64_t* out, const uint32_t* n_dev,
                             uint64_t n_max, void* tmp, hipStream_t st) {
  scan_impl<uint64_t>(in, out, std::max<uint64_t>(n_max, 1), tmp, st, n_dev);
}

size_t radix_hist_bytes(uint64_t n) {
  uint64_t nb = (n + RS_TILE - 1) / RS_TILE;
  if (nb == 0) nb = 1;
  // histogram + its scan (with the total slot) + the device-length word
  return align256(256 * nb * 4) + align256((256 * nb + 1) * 4) + 256;
}


void radix_sort_pairs(const uint32_t* k_in, const uint32_t* v_in, uint32_t* k_out,
                      uint32_t* v_out, uint32_t* k_tmp, uint32_t* v_tmp, uint64_t n,
                      uint32_t key_bits, void* hist, void* scan_tmp, hipStream_t st) {
  radix_sort_impl(k_in, v_in, k_out, v_out, k_tmp, v_tmp, n, nullptr, key_bits, hist, scan_tmp,
                  st);
}

void radix_sort_pairs_devn(const uint32_t* k_in, const uint32_t* v_in, uint32_t* k_out,
                           uint32_t* v_out, uint32_t* k_tmp, uint32_t* v_tmp,
                           const uint32_t* n_dev, uint64_t n_max, uint32_t key_bits, void* hist,
                           void* scan_tmp, hipStream_t st) {
  radix_sort_impl(k_in, v_in, k_out, v_out, k_tmp, v_tmp, std::max<uint64_t>(n_max, 1), n_dev,
                  key_bits, hist, scan_tmp, st);
}

void scan_exclusive_u32_devn(const uint32_t* in, uint32_t* out, const uint32_t* n_dev,
                             uint64_t n_max, void* tmp, hipStream_t st) {
  scan_impl<uint32_t>(in, out, std::max<uint64_t>(n_max, 1), tmp, st, n_dev);
}

void fill_u32(uint32_t* p, uint64_t n, uint32_t v, hipStream_t st) {
  if (!n) return;
  k_fill<uint32_t><<<(uint32_t)std::min<uint64_t>((n + 255) / 256, 16384), 256, 0, st>>>(p, n, v);
}

void fill_many(const FillOp* ops, int n_ops, hipStream_t st) {
  FillSet fs{};
  uint64_t most = 1;
  int k = 0;
  for (int i = 0; i < n_ops; ++i) {
    if (!ops[i].n) continue;
    if (k == MAX_FILLS) {  // a full set: launch it, start the next
      k_fill_many<<<dim3((uint32_t)std::min<uint64_t>((most + 255) / 256, 16384), k), 256, 0, st>>>(fs);
      k = 0;
      most = 1;
    }
    fs.op[k++] = ops[i];
    most = std::max(most, ops[i].n);
  }
  if (k)
    k_fill_many<<<dim3((uint32_t)std::min<uint64_t>((most + 255) / 256, 16384), k), 256, 0, st>>>(fs);
}

void fill_u64(uint64_t* p, uint64_t n, uint64_t v, hipStream_t st) {
  if (!n) return;
  k_fill<uint64_t><<<(uint32_t)std::min<uint64_t>((n + 255) / 256, 16384), 256, 0, st>>>(p, n, v);
}


}  // namespace cytom
