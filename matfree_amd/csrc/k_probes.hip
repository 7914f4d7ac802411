// Rademacher probe generation on the device.
//
// Bit-identical to jax.random.rademacher(key, shape, float32) with JAX's
// partitionable threefry2x32 layout, which is what matfree's sampler_signs
// calls (stochtrace.py:932-937, 974-977; backend/prng.py:26-29):
//   word(m)   = y0 ^ y1, (y0, y1) = threefry2x32-20(key, (m >> 32, m & 0xffffffff))
//   sample(m) = +1 if bit 31 of word(m) is clear, else -1
// (uniform = bitcast((word >> 9) | 0x3f800000) - 1 < 0.5  <=>  bit 31 clear).
#include "slq_internal.h"

namespace slq {
namespace {

__device__ __forceinline__ uint32_t rotl32(uint32_t x, int r) {
  return (x << r) | (x >> (32 - r));
}

// threefry2x32 with 20 rounds: five groups of four mix rounds with key
// injection after every group (Salmon et al., SC'11).
__device__ __forceinline__ uint32_t threefry_word(uint32_t k0, uint32_t k1, uint32_t hi,
                                                  uint32_t lo) {
  const uint32_t k2 = k0 ^ k1 ^ 0x1BD11BDAu;
  uint32_t x0 = hi + k0;
  uint32_t x1 = lo + k1;
#define SLQ_MIX(r)      \
  x0 += x1;             \
  x1 = rotl32(x1, (r)); \
  x1 ^= x0;
  SLQ_MIX(13) SLQ_MIX(15) SLQ_MIX(26) SLQ_MIX(6)
  x0 += k1; x1 += k2 + 1u;
  SLQ_MIX(17) SLQ_MIX(29) SLQ_MIX(16) SLQ_MIX(24)
  x0 += k2; x1 += k0 + 2u;
  SLQ_MIX(13) SLQ_MIX(15) SLQ_MIX(26) SLQ_MIX(6)
  x0 += k0; x1 += k1 + 3u;
  SLQ_MIX(17) SLQ_MIX(29) SLQ_MIX(16) SLQ_MIX(24)
  x0 += k1; x1 += k2 + 4u;
  SLQ_MIX(13) SLQ_MIX(15) SLQ_MIX(26) SLQ_MIX(6)
  x0 += k2; x1 += k0 + 5u;
#undef SLQ_MIX
  return x0 ^ x1;
}

__device__ __forceinline__ float sign_of(uint32_t w) { return (w >> 31) ? -1.0f : 1.0f; }

__global__ void __launch_bounds__(256) k_rademacher(uint32_t k0, uint32_t k1, int64_t count,
                                                    float* __restrict__ out) {
  const int64_t groups = count / 4;
  const int64_t stride = int64_t(gridDim.x) * blockDim.x;
  float4* out4 = reinterpret_cast<float4*>(out);
  for (int64_t g = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; g < groups; g += stride) {
    const uint64_t m = uint64_t(g) * 4;
    float4 v;
    v.x = sign_of(threefry_word(k0, k1, uint32_t((m + 0) >> 32), uint32_t(m + 0)));
    v.y = sign_of(threefry_word(k0, k1, uint32_t((m + 1) >> 32), uint32_t(m + 1)));
    v.z = sign_of(threefry_word(k0, k1, uint32_t((m + 2) >> 32), uint32_t(m + 2)));
    v.w = sign_of(threefry_word(k0, k1, uint32_t((m + 3) >> 32), uint32_t(m + 3)));
    out4[g] = v;
  }
  // tail (count % 4 elements), handled by the first threads of block 0
  if (blockIdx.x == 0 && threadIdx.x < (count & 3)) {
    const uint64_t m = uint64_t(groups) * 4 + threadIdx.x;
    out[m] = sign_of(threefry_word(k0, k1, uint32_t(m >> 32), uint32_t(m)));
  }
}

__global__ void __launch_bounds__(256) k_rademacher_scalar(uint32_t k0, uint32_t k1,
                                                           int64_t count, float* __restrict__ out) {
  const int64_t stride = int64_t(gridDim.x) * blockDim.x;
  for (int64_t m = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; m < count; m += stride) {
    out[m] = sign_of(threefry_word(k0, k1, uint32_t(uint64_t(m) >> 32), uint32_t(m)));
  }
}

}  // namespace

hipError_t launch_rademacher(hipStream_t s, uint32_t k0, uint32_t k1, int64_t count, float* out) {
  if (count <= 0) return hipSuccess;
  const bool aligned = (reinterpret_cast<uintptr_t>(out) % 16) == 0;
  const int64_t work = aligned ? (count + 3) / 4 : count;
  int64_t blocks = (work + 255) / 256;
  if (blocks > 8192) blocks = 8192;
  if (blocks < 1) blocks = 1;
  if (aligned) {
    hipLaunchKernelGGL(k_rademacher, dim3(unsigned(blocks)), dim3(256), 0, s, k0, k1, count, out);
  } else {
    hipLaunchKernelGGL(k_rademacher_scalar, dim3(unsigned(blocks)), dim3(256), 0, s, k0, k1, count,
                       out);
  }
  return hipGetLastError();
}

}  // namespace slq
