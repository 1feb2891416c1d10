// Counter-based RMAT update-stream generator on the device.
//
// Edge i of the stream is RMATContext::getEdge(i)
// (src/include/edge-inserter/rmat-context.h:39-73): per level a quadrant is
// picked by hashDouble(start + k*stride) against a, a+b, a+b+c
// (util.h:188-200), with start = hash(2*i*h), stride = hash((2*i+1)*h),
// h = hash(seed).  The thresholds are the float literals RMAT_A/B/C
// (datatypes.h:17-19) promoted to double, which the reference constructor
// never stores.  Weights are the stream's own: w(i) = 1 + hash(hw ^ i) % 255
// with hw = hash(seed ^ WEIGHT_SALT).
#include "internal.h"

namespace cytom {
namespace {

__global__ void __launch_bounds__(256) k_rmat(uint32_t scale, uint64_t h, uint64_t hw,
                                              uint64_t first, const uint64_t* __restrict__ index,
                                              uint64_t n, uint32_t* __restrict__ src,
                                              uint32_t* __restrict__ dst,
                                              uint32_t* __restrict__ w) {
  const double A = (double)0.57f;
  const double AB = A + (double)0.19f;
  const double ABC = AB + (double)0.19f;
  const double DEN = (double)UINT64_MAX;
  for (uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; j < n;
       j += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t i = index ? index[j] : first + j;
    const uint64_t start = hash64(2ull * i * h);
    const uint64_t stride = hash64((2ull * i + 1ull) * h);
    uint32_t s = 0, d = 0;
    for (uint32_t k = 1; k <= scale; ++k) {
      const double r = (double)hash64(start + (uint64_t)(scale - k) * stride) / DEN;
      const uint32_t half = 1u << (k - 1);
      if (r < A) {
      } else if (r < AB) {
        d += half;
      } else if (r < ABC) {
        s += half;
      } else {
        s += half;
        d += half;
      }
    }
    src[j] = s;
    dst[j] = d;
    if (w) w[j] = 1u + (uint32_t)(hash64(hw ^ i) % 255ull);
  }
}

}  // namespace

void launch_rmat(uint32_t scale, uint64_t seed, uint64_t first, const uint64_t* index,
                 uint64_t n, uint32_t* src, uint32_t* dst, uint32_t* w, hipStream_t st) {
  if (n == 0) return;
  uint64_t blocks = (n + 255) / 256;
  if (blocks > 65536) blocks = 65536;
  k_rmat<<<(uint32_t)blocks, 256, 0, st>>>(scale, hash64(seed), hash64(seed ^ WEIGHT_SALT),
                                           first, index, n, src, dst, w);
}

}  // namespace cytom
