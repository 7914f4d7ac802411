// Block matvec W = A V over all probes at once: the reference's matvec
// (decomp.py:460), vmapped over probes (stochtrace.py:49), becomes one dense
// contraction W[p][i] = sum_m A[i][m] V[p][m] on the matrix cores.
//
// FP32 path: v_mfma_f32_32x32x2f32 with D[i][p] (rows = output elements,
// columns = probes).  A workgroup of 4 waves owns 128 output elements; each
// wave owns 32 of them for 32 * NACC probes (NACC 32x32 accumulators).
//
//  * A operand: each lane builds its own fragment in registers -- row
//    i = i0 + 32 * wave + (lane & 31), contraction indices
//    m0 + 16 * (lane >> 5) + s, s = 0..15 of the current BK = 32 tile.  For a
//    stored matrix the 16 values are one contiguous 64-byte load (prefetched a
//    tile ahead); for the RBF kernel they are generated in place from the
//    lane's own point and the tile's 32 points, so every kernel entry is
//    evaluated exactly once per matvec and never touches LDS.
//  * B operand: the probe tile V[p0 .. p0 + 32 NACC)[m0 .. m0 + 32) is shared
//    by the four waves through a double-buffered LDS tile (one barrier per
//    tile; the next tile's global loads are in flight during the MFMAs).
//  * Blocked summation: every tile accumulates into fresh registers that are
//    then added to the running sums, so rounding grows with BK + n / BK rather
//    than n (as in a blocked CPU GEMM).
//
// Within an MFMA the two lane halves carry two different contraction indices;
// lane half h walks m in [16 h, 16 h + 16) of the tile, for both operands.
//
// RBF: K_im = exp2(c |x_i - x_m|^2), c = -log2(e) / (2 l^2); the nugget
// s2 * v is added in FP32 in the epilogue (never folded into a tile).
#include <cmath>

#include "slq_internal.h"

namespace slq {
namespace {

typedef float floatx16 __attribute__((ext_vector_type(16)));

constexpr int BI = 128;      // output elements per workgroup
constexpr int BK = 32;       // contraction tile
constexpr int LDK = BK + 4;  // padded LDS row (floats)
constexpr int kThreads = 256;

struct RBFParams {
  const float* X;
  int dim;
  float coef;  // -log2(e) / (2 l^2)
  float noise;
  float scale;  // sqrt(log2(e) / (2 l^2)) (FP32 kernels)
};

// Coordinate of the points past n in the FP32 kernels' staged tiles: far
// enough that exp2(-|s x_i - s x_m|^2) is exactly 0 for every point in use
// (the squared distance is at least ~1e36 times s^2, or +inf), so those
// entries need no mask.
constexpr float kFarPoint = 1e18f;

// 16 bytes at a 4-byte aligned address (rows of n % 4 != 0 floats start at
// every dword alignment): a memcpy, which gfx950 code (unaligned global
// access enabled by the HSA runtime) turns into one global_load_dwordx4 /
// global_store_dwordx4 whatever the alignment.
__device__ __forceinline__ float4 load4u(const float* p) {
  float4 v;
  __builtin_memcpy(&v, p, sizeof(v));
  return v;
}
__device__ __forceinline__ void store4u(float* p, const float4& v) {
  __builtin_memcpy(p, &v, sizeof(v));
}

// V tile: (32 NACC) probes x 32 m = 8 NACC float4, NACC per thread.
template <int NACC, bool VEC>
__device__ __forceinline__ void load_v_tile(const float* __restrict__ V, int64_t P, int64_t n,
                                            int64_t p0, int64_t m0, float4 (&reg)[NACC]) {
  const int t = threadIdx.x;
#pragma unroll
  for (int u = 0; u < NACC; ++u) {
    const int f = t + u * kThreads;
    const int row = f >> 3, c4 = f & 7;
    const int64_t p = p0 + row, m = m0 + 4 * c4;
    float4 v = make_float4(0.f, 0.f, 0.f, 0.f);
    if (p < P) {
      const float* src = V + p * n + m;
      if (VEC) {
        if (m < n) v = *reinterpret_cast<const float4*>(src);
      } else if (m + 4 <= n) {
        v = load4u(src);
      } else {
        if (m + 0 < n) v.x = src[0];
        if (m + 1 < n) v.y = src[1];
        if (m + 2 < n) v.z = src[2];
        if (m + 3 < n) v.w = src[3];
      }
    }
    reg[u] = v;
  }
}

template <int NACC>
__device__ __forceinline__ void store_v_tile(float* Vs, const float4 (&reg)[NACC]) {
  const int t = threadIdx.x;
#pragma unroll
  for (int u = 0; u < NACC; ++u) {
    const int f = t + u * kThreads;
    const int row = f >> 3, c4 = f & 7;
    *reinterpret_cast<float4*>(Vs + row * LDK + 4 * c4) = reg[u];
  }
}

// The lane's 16 stored-matrix values A[row][m0 + 16 h + 0..15].
template <bool VEC>
__device__ __forceinline__ void load_a_frag(const float* __restrict__ A, int64_t n, int64_t row,
                                            int64_t m0, int h, float (&a)[16]) {
  const int64_t m = m0 + 16 * h;
  if (row < n) {
    const float* src = A + row * n + m;
    if (m + 16 <= n) {
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const float4 v = VEC ? *reinterpret_cast<const float4*>(src + 4 * q) : load4u(src + 4 * q);
        a[4 * q + 0] = v.x;
        a[4 * q + 1] = v.y;
        a[4 * q + 2] = v.z;
        a[4 * q + 3] = v.w;
      }
    } else {
#pragma unroll
      for (int s = 0; s < 16; ++s) a[s] = (m + s < n) ? src[s] : 0.f;
    }
  } else {
#pragma unroll
    for (int s = 0; s < 16; ++s) a[s] = 0.f;
  }
}

// The lane's 16 generated RBF values exp2(-|x~_row - x~_m|^2) from scaled
// coordinates x~ = s x, s = sqrt(log2(e) / (2 l^2)) (points past n staged at
// kFarPoint, which gives exactly 0).
template <int DIM>
__device__ __forceinline__ void gen_rbf_frag(const float4* Xs, const float (&xr)[4], int h,
                                             float (&a)[16]) {
#pragma unroll
  for (int s = 0; s < 16; ++s) {
    const float4 xm = Xs[16 * h + s];
    float d = xr[0] - xm.x;
    float d2 = d * d;
    if (DIM > 1) { d = xr[1] - xm.y; d2 = fmaf(d, d, d2); }
    if (DIM > 2) { d = xr[2] - xm.z; d2 = fmaf(d, d, d2); }
    if (DIM > 3) { d = xr[3] - xm.w; d2 = fmaf(d, d, d2); }
    a[s] = __builtin_amdgcn_exp2f(-d2);
  }
}

// Epilogue: lane l of a 32x32 accumulator holds column (l & 31) and rows
// 8 * (r / 4) + 4 * (l / 32) + (r % 4), r = 0..15.
template <int NACC, bool VEC, bool NUGGET>
__device__ __forceinline__ void store_out(float* __restrict__ W, const float* __restrict__ V,
                                          int64_t P, int64_t n, int64_t i0w, int64_t p0,
                                          const floatx16 (&acc)[NACC], float noise) {
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int c = 0; c < NACC; ++c) {
    const int64_t p = p0 + 32 * c + (lane & 31);
    if (p >= P) continue;
#pragma unroll
    for (int g = 0; g < 4; ++g) {
      const int64_t i = i0w + 8 * g + 4 * (lane >> 5);
      float o[4] = {acc[c][4 * g + 0], acc[c][4 * g + 1], acc[c][4 * g + 2], acc[c][4 * g + 3]};
      if (VEC) {
        if (i < n) {
          if (NUGGET) {
            const float4 v = *reinterpret_cast<const float4*>(V + p * n + i);
            o[0] = fmaf(noise, v.x, o[0]);
            o[1] = fmaf(noise, v.y, o[1]);
            o[2] = fmaf(noise, v.z, o[2]);
            o[3] = fmaf(noise, v.w, o[3]);
          }
          *reinterpret_cast<float4*>(W + p * n + i) = make_float4(o[0], o[1], o[2], o[3]);
        }
      } else if (i + 4 <= n) {
        if (NUGGET) {
          const float4 v = load4u(V + p * n + i);
          o[0] = fmaf(noise, v.x, o[0]);
          o[1] = fmaf(noise, v.y, o[1]);
          o[2] = fmaf(noise, v.z, o[2]);
          o[3] = fmaf(noise, v.w, o[3]);
        }
        store4u(W + p * n + i, make_float4(o[0], o[1], o[2], o[3]));
      } else {
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          if (i + u < n) {
            float val = o[u];
            if (NUGGET) val = fmaf(noise, V[p * n + i + u], val);
            W[p * n + i + u] = val;
          }
        }
      }
    }
  }
}

template <int NACC, bool RBF, int DIM, bool VEC>
__global__ void __launch_bounds__(kThreads)
    k_matvec_f32(const float* __restrict__ A, RBFParams prm, const float* __restrict__ V,
                 int64_t P, int64_t n, float* __restrict__ W) {
  constexpr int BPW = 32 * NACC;  // probes per workgroup
  __shared__ __attribute__((aligned(16))) float Vs[2][BPW * LDK];
  __shared__ float4 Xs[2][BK];
  const int t = threadIdx.x;
  const int lane = t & 63, wave = t >> 6, r = lane & 31, h = lane >> 5;
  const int64_t i0w = int64_t(blockIdx.x) * BI + 32 * wave;
  const int64_t row = i0w + r;
  const int64_t p0 = int64_t(blockIdx.y) * BPW;
  // split-K: slice z of gridDim.z covers tiles [kt_begin, kt_end)
  const int64_t all_tiles = (n + BK - 1) / BK;
  const int64_t kt_begin = all_tiles * blockIdx.z / gridDim.z;
  const int64_t kt_end = all_tiles * (blockIdx.z + 1) / gridDim.z;
  const int64_t ntiles = kt_end - kt_begin;
  const int64_t mbase = kt_begin * BK;
  const bool split = gridDim.z > 1;
  float* Wout = split ? W + int64_t(blockIdx.z) * P * n : W;

  floatx16 acc[NACC];
#pragma unroll
  for (int c = 0; c < NACC; ++c)
#pragma unroll
    for (int e = 0; e < 16; ++e) acc[c][e] = 0.f;

  // points relative to point 0 (translation invariance; both sides form the
  // same FP32 difference), scaled
  float x0[4] = {0.f, 0.f, 0.f, 0.f};
  float xr[4] = {0.f, 0.f, 0.f, 0.f};
  if (RBF && n > 0) {
#pragma unroll
    for (int d = 0; d < DIM; ++d) x0[d] = prm.X[d];
  }
  if (RBF && row < n) {
#pragma unroll
    for (int d = 0; d < DIM; ++d) xr[d] = (prm.X[row * DIM + d] - x0[d]) * prm.scale;
  }
  auto load_x = [&](int64_t m0) {
    float4 x = make_float4(0.f, 0.f, 0.f, 0.f);
    const int64_t m = m0 + t;
    if (m < n) {
      const float* src = prm.X + m * DIM;
      x.x = (src[0] - x0[0]) * prm.scale;
      if (DIM > 1) x.y = (src[1] - x0[1]) * prm.scale;
      if (DIM > 2) x.z = (src[2] - x0[2]) * prm.scale;
      if (DIM > 3) x.w = (src[3] - x0[3]) * prm.scale;
    } else {
      x = make_float4(kFarPoint, kFarPoint, kFarPoint, kFarPoint);
    }
    return x;
  };

  float4 vreg[NACC];
  float areg[16];
  float4 xnext = make_float4(0.f, 0.f, 0.f, 0.f);
  load_v_tile<NACC, VEC>(V, P, n, p0, mbase, vreg);
  store_v_tile<NACC>(Vs[0], vreg);
  if (RBF) {
    if (t < BK) Xs[0][t] = load_x(mbase);
  } else {
    load_a_frag<VEC>(A, n, row, mbase, h, areg);
  }
  __syncthreads();

  for (int64_t kt = 0; kt < ntiles; ++kt) {
    const int buf = int(kt & 1);
    const int64_t m0 = mbase + kt * BK;
    const bool more = kt + 1 < ntiles;
    if (more) {
      load_v_tile<NACC, VEC>(V, P, n, p0, m0 + BK, vreg);
      if (RBF && t < BK) xnext = load_x(m0 + BK);
    }
    float a[16];
    if (RBF) {
      gen_rbf_frag<DIM>(Xs[buf], xr, h, a);
    } else {
#pragma unroll
      for (int s = 0; s < 16; ++s) a[s] = areg[s];
      if (more) load_a_frag<VEC>(A, n, row, m0 + BK, h, areg);
    }

    floatx16 tmp[NACC];
#pragma unroll
    for (int c = 0; c < NACC; ++c)
#pragma unroll
      for (int e = 0; e < 16; ++e) tmp[c][e] = 0.f;
    const float* vb = Vs[buf] + r * LDK + 16 * h;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      float4 b[NACC];
#pragma unroll
      for (int c = 0; c < NACC; ++c)
        b[c] = *reinterpret_cast<const float4*>(vb + 32 * c * LDK + 4 * q);
#pragma unroll
      for (int c = 0; c < NACC; ++c)
        tmp[c] = __builtin_amdgcn_mfma_f32_32x32x2f32(a[4 * q + 0], b[c].x, tmp[c], 0, 0, 0);
#pragma unroll
      for (int c = 0; c < NACC; ++c)
        tmp[c] = __builtin_amdgcn_mfma_f32_32x32x2f32(a[4 * q + 1], b[c].y, tmp[c], 0, 0, 0);
#pragma unroll
      for (int c = 0; c < NACC; ++c)
        tmp[c] = __builtin_amdgcn_mfma_f32_32x32x2f32(a[4 * q + 2], b[c].z, tmp[c], 0, 0, 0);
#pragma unroll
      for (int c = 0; c < NACC; ++c)
        tmp[c] = __builtin_amdgcn_mfma_f32_32x32x2f32(a[4 * q + 3], b[c].w, tmp[c], 0, 0, 0);
    }
#pragma unroll
    for (int c = 0; c < NACC; ++c)
#pragma unroll
      for (int e = 0; e < 16; ++e) acc[c][e] += tmp[c][e];

    if (more) {
      store_v_tile<NACC>(Vs[buf ^ 1], vreg);
      if (RBF && t < BK) Xs[buf ^ 1][t] = xnext;
    }
    __syncthreads();
  }
  if (split)
    store_out<NACC, VEC, false>(Wout, V, P, n, i0w, p0, acc, 0.f);
  else
    store_out<NACC, VEC, RBF>(Wout, V, P, n, i0w, p0, acc, prm.noise);
}

// ---------------------------------------------------------------------------
// BF16 paths: v_mfma_f32_32x32x16_bf16 (gfx950), FP32 accumulation.
// Lane (r, h) supplies A[i = r][k = 8 h .. 8 h + 7] of each 16-deep MFMA, so
// within a BK = 32 tile it owns 16 contraction slots (8 per MFMA).
//   NPROD = 1: hi*hi with x_hi = bf16(x);
//   NPROD = 2: lo*hi + hi*hi -- the matrix split x = hi + lo (lo = bf16(x -
//              hi)), the probes rounded once to BF16;
//   NPROD = 3: lo*hi + hi*lo + hi*hi -- both operands split, ~16 significant
//              bits recovered.
// BF16 products are exact in FP32, so the accumulation chain needs no
// blocking: its rounding (~sqrt(n / 16) ulp) stays below the representation
// error of the split operands.
//
// The probes are split into BF16 hi/lo planes once per matvec by
// k_split_probes (not once per workgroup), into a tile-major layout that is
// zero-padded to whole tiles and probe blocks, so the tile loads carry no
// bounds tests.  For the RBF kernel each 32-column tile is stored in the
// contraction order produced by the MFMA-generated fragments (column groups
// of four with bits 0 and 1 of the group index swapped).
// ---------------------------------------------------------------------------
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x4 __attribute__((ext_vector_type(4)));
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

constexpr int LDB = BK + 8;  // padded LDS row of bf16 (80 bytes)

__device__ __forceinline__ float bf16_to_f32(__bf16 b) { return float(b); }

typedef float f32x2 __attribute__((ext_vector_type(2)));
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));

// Split two floats into packed BF16 hi and lo words (one v_cvt_pk_bf16_f32
// each): hi = bf16(x), lo = bf16(x - hi).
__device__ __forceinline__ void split_pair(float x0, float x1, uint32_t& hi, uint32_t& lo) {
  const f32x2 x = {x0, x1};
  hi = __builtin_bit_cast(uint32_t, __builtin_convertvector(x, bf16x2));
  const f32x2 h = {__builtin_bit_cast(float, hi << 16),
                   __builtin_bit_cast(float, hi & 0xffff0000u)};
  lo = __builtin_bit_cast(uint32_t, __builtin_convertvector(x - h, bf16x2));
}

__device__ __forceinline__ int perm_group(int c4) {
  return (c4 & 4) | ((c4 & 1) << 1) | ((c4 & 2) >> 1);
}

// Split BF16 planes of V, zero-padded, tile-major: the 32 columns of tile kt
// of probe p sit at (kt * Ppad + p) * 32, so one workgroup's probe tile is a
// contiguous block of 64 * BPW bytes.  Vlo may be null (plane not read).
template <bool PERM>
__global__ void __launch_bounds__(256)
    k_split_probes(const float* __restrict__ V, int64_t P, int64_t n, int64_t Ppad, int64_t npad,
                   __bf16* __restrict__ Vhi, __bf16* __restrict__ Vlo) {
  const int64_t groups = Ppad * (npad / 4);
  const int64_t stride = int64_t(gridDim.x) * blockDim.x;
  for (int64_t g = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; g < groups; g += stride) {
    const int64_t p = g / (npad / 4);
    const int64_t m = 4 * (g % (npad / 4));
    float x[4] = {0.f, 0.f, 0.f, 0.f};
    if (p < P) {
#pragma unroll
      for (int e = 0; e < 4; ++e)
        if (m + e < n) x[e] = V[p * n + m + e];
    }
    uint32_t h0, l0, h1, l1;
    split_pair(x[0], x[1], h0, l0);
    split_pair(x[2], x[3], h1, l1);
    const int64_t tile = m / BK;
    const int c4 = int((m % BK) / 4);
    const int64_t dst = (tile * Ppad + p) * BK + 4 * (PERM ? perm_group(c4) : c4);
    *reinterpret_cast<uint2*>(Vhi + dst) = make_uint2(h0, h1);
    if (Vlo) *reinterpret_cast<uint2*>(Vlo + dst) = make_uint2(l0, l1);
  }
}

// The planes (layout and permuted order of k_split_probes<true>, Ppad
// columns) of columns c0 .. c0 + Ppad of a Krylov-layout set: column c at
// V + (c / Pin) ldc + (c % Pin) n, zero for c >= K.  One column per
// blockIdx.y.
__global__ void __launch_bounds__(256)
    k_split_columns(const float* __restrict__ V, int64_t ldc, int64_t Pin, int64_t n, int64_t K,
                    int64_t c0, int64_t Ppad, int64_t npad, __bf16* __restrict__ Vhi,
                    __bf16* __restrict__ Vlo) {
  const int64_t p = blockIdx.y, c = c0 + p;
  const float* src = c < K ? V + (c / Pin) * ldc + (c % Pin) * n : nullptr;
  const int64_t groups = npad / 4;
  for (int64_t g = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; g < groups;
       g += int64_t(gridDim.x) * blockDim.x) {
    const int64_t m = 4 * g;
    float x[4] = {0.f, 0.f, 0.f, 0.f};
    if (src) {
      if (m + 4 <= n) {
        const float4 v = load4u(src + m);
        x[0] = v.x;
        x[1] = v.y;
        x[2] = v.z;
        x[3] = v.w;
      } else {
#pragma unroll
        for (int e = 0; e < 4; ++e)
          if (m + e < n) x[e] = src[m + e];
      }
    }
    uint32_t h0, l0, h1, l1;
    split_pair(x[0], x[1], h0, l0);
    split_pair(x[2], x[3], h1, l1);
    const int64_t dst = ((m / BK) * Ppad + p) * BK + 4 * perm_group(int((m % BK) / 4));
    *reinterpret_cast<uint2*>(Vhi + dst) = make_uint2(h0, h1);
    *reinterpret_cast<uint2*>(Vlo + dst) = make_uint2(l0, l1);
  }
}

// One probe tile of one plane: BPW = 32 NACC rows x 32 bf16 = 4 BPW pieces of
// 16 bytes, spread over the 256 threads (NACC / 2 pieces each, or one piece
// for the first 128 threads when NACC = 1).
template <int NACC>
struct PlaneTile {
  static constexpr int kPieces = 4 * 32 * NACC;
  static constexpr int kPerThread = (kPieces + kThreads - 1) / kThreads;
};

// The contiguous BPW x 32 block of a plane tile starting at base.
template <int NACC>
__device__ __forceinline__ void load_plane_tile_at(const __bf16* __restrict__ base,
                                                   u32x4 (&reg)[PlaneTile<NACC>::kPerThread]) {
  const int t = threadIdx.x;
#pragma unroll
  for (int u = 0; u < PlaneTile<NACC>::kPerThread; ++u) {
    const int f = t + u * kThreads;
    if (PlaneTile<NACC>::kPieces % kThreads == 0 || f < PlaneTile<NACC>::kPieces)
      reg[u] = *reinterpret_cast<const u32x4*>(base + 8 * f);
  }
}

template <int NACC>
__device__ __forceinline__ void load_plane_tile(const __bf16* __restrict__ plane, int64_t Ppad,
                                                int64_t p0, int64_t kt,
                                                u32x4 (&reg)[PlaneTile<NACC>::kPerThread]) {
  load_plane_tile_at<NACC>(plane + (kt * Ppad + p0) * BK, reg);
}

template <int NACC>
__device__ __forceinline__ void store_plane_tile(__bf16* lds,
                                                 const u32x4 (&reg)[PlaneTile<NACC>::kPerThread]) {
  const int t = threadIdx.x;
#pragma unroll
  for (int u = 0; u < PlaneTile<NACC>::kPerThread; ++u) {
    const int f = t + u * kThreads;
    if (PlaneTile<NACC>::kPieces % kThreads == 0 || f < PlaneTile<NACC>::kPieces) {
      const int row = f >> 2, q = f & 3;
      *reinterpret_cast<u32x4*>(lds + row * LDB + 8 * q) = reg[u];
    }
  }
}

// Contraction index of fragment slot s (0..15) for lane half h (natural order).
__device__ __forceinline__ int frag_k(int s, int h) { return (s < 8) ? 8 * h + s : 16 + 8 * h + (s - 8); }

template <bool VEC>
__device__ __forceinline__ void load_a_frag16(const float* __restrict__ A, int64_t n, int64_t row,
                                              int64_t m0, int h, float (&a)[16]) {
  if (row < n) {
    const float* src = A + row * n + m0;
    if (m0 + BK <= n) {
#pragma unroll
      for (int half = 0; half < 2; ++half) {
        const float* q = src + 16 * half + 8 * h;
        const float4 v0 = VEC ? *reinterpret_cast<const float4*>(q) : load4u(q);
        const float4 v1 = VEC ? *reinterpret_cast<const float4*>(q + 4) : load4u(q + 4);
        a[8 * half + 0] = v0.x; a[8 * half + 1] = v0.y; a[8 * half + 2] = v0.z;
        a[8 * half + 3] = v0.w; a[8 * half + 4] = v1.x; a[8 * half + 5] = v1.y;
        a[8 * half + 6] = v1.z; a[8 * half + 7] = v1.w;
      }
    } else {
#pragma unroll
      for (int s = 0; s < 16; ++s) {
        const int64_t m = m0 + frag_k(s, h);
        a[s] = (m < n) ? src[frag_k(s, h)] : 0.f;
      }
    }
  } else {
#pragma unroll
    for (int s = 0; s < 16; ++s) a[s] = 0.f;
  }
}

template <int NPROD>
__device__ __forceinline__ void split_frag(const float (&a)[16], u32x4 (&hi)[2], u32x4 (&lo)[2]) {
#pragma unroll
  for (int kb = 0; kb < 2; ++kb)
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      uint32_t h, l;
      if (NPROD >= 2) {
        split_pair(a[8 * kb + 2 * e], a[8 * kb + 2 * e + 1], h, l);
        lo[kb][e] = l;
      } else {
        const f32x2 x = {a[8 * kb + 2 * e], a[8 * kb + 2 * e + 1]};
        h = __builtin_bit_cast(uint32_t, __builtin_convertvector(x, bf16x2));
      }
      hi[kb][e] = h;
    }
}

__device__ __forceinline__ floatx16 mfma_bf16(const u32x4& a, const u32x4& b, const floatx16& c) {
  return __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(bf16x8, a),
                                                 __builtin_bit_cast(bf16x8, b), c, 0, 0, 0);
}

// RBF exponents for the BF16 kernels.  The base-2 exponent of a 32 x 32 tile,
//   arg[m][i] = c |x_m|^2 + c |x_i|^2 + sum_d x_m,d (-2c x_i,d),
// is a short sum of products of FP32 numbers.  Each factor f is split into
// three BF16 parts f = f_h + f_m + f_l (round to nearest each), and every
// product is expanded to the parts whose size is at least 2^-16 of the
// product: for each coordinate  hH + hM + mH + hL + lH + mM  (6 slots), and
// the three parts of c |x_m|^2 and of c |x_i|^2 against an exact 1 -- 24 of
// the 32 contraction slots of two BF16 MFMAs (32x32x16), the other 8 zero.
// BF16 products are exact in FP32, so the exponent carries only the dropped
// parts (<= 2^-24 relative per product) and the FP32 accumulation, the same
// order as an FP32 evaluation (max abs error ~3e-6 for points in the unit
// cube and l = 0.5), and below the 2^-17 relative error of the split BF16
// representation of the entries.
//
// Slot s of the contraction (MFMA s / 16, lane half (s % 16) / 8):
//   6 d + 0..5 (d < 3): (x_m,d ; -2c x_i,d) as hH, hM, mH, hL, lH, mM
//   18..20: (c |x_m|^2 parts ; 1)     21..23: (1 ; c |x_i|^2 parts)
// The point side ("L") of each point is precomputed once per matvec as 64
// bytes (32 in the compact form below); the row side ("R") is built per lane
// at kernel start.  Both sides take the points relative to point 0 (the RBF
// kernel is translation invariant; each side computes the same FP32
// difference for a point), which keeps |x|^2, and with it the rounding of
// the expanded exponent's sums, small for data far from the origin.  The MFMA
// output gives lane (r, h) the exponents of its own row i = r at the 16
// points m = 8 (s / 4) + 4 h + (s % 4), s = 0..15 -- exactly the BF16
// A-fragment of that lane once the contraction index is ordered that way,
// which the permuted probe planes provide.  Points past n carry -2^100 in
// the slot of c |x_m|^2 (finite, so that d exp2(d) of the derivative kernel
// is 0 as well as exp2(d)), so their entries are exactly 0.
__device__ __forceinline__ void split3(float x, __bf16 (&p)[3]) {
  p[0] = __bf16(x);
  const float r1 = x - float(p[0]);
  p[1] = __bf16(r1);
  p[2] = __bf16(r1 - float(p[1]));
}

// The compact form (16 slots, one MFMA) for the split-BF16 matvec with the
// probes in one plane, whose entries carry 2^-17 relative error anyway:
// scaled coordinates y = s x (s = sqrt(-c)) are rounded to two BF16 parts,
// y~ = y_h + y_l, and the exponent is that of the points y~,
//   arg[m][i] = -|y~_m|^2 - |y~_i|^2 + sum_d y~_m,d (2 y~_i,d),
// with the cross products exact (hH, hL, lH, lL: slots 4 d + 0..3) and each
// squared norm (FP32, from y~) as two BF16 parts against an exact 1 (slots
// 12, 13 for the point, 14, 15 for the row).  Both sides round the same
// y~ for a point, so the tile is the RBF kernel of points moved by at most
// 2^-18 of their scaled coordinates, times D_m D_i with D = exp2(delta),
// |delta| <= 2^-18 |y~|^2, from the norms' rounding: a symmetric positive
// definite matrix whose log-determinant differs from the exact one by ~1e-9
// relative (measured in FP64 at n = 4096 and 8192, points in the unit cube,
// l = 0.5, noise 1e-2).  Points past n carry -2^100 in slot 12.
//
// The slots of one side (NS = 32: the full form above, NS = 16: the
// compact form).  side 0: point side of x (the L factors), side 1: row side
// (the R factors).
template <int NS>
__device__ __forceinline__ void exponent_slots(const float (&x)[3], float coef, float scale,
                                               int side, bool pad, __bf16 (&slot)[NS]) {
  const __bf16 one = __bf16(1.f), zero = __bf16(0.f);
  if constexpr (NS == 16) {
    float nrm = 0.f;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      const float y = scale * x[d];
      const __bf16 yh = __bf16(y);
      const __bf16 yl = __bf16(y - float(yh));
      const float yt = float(yh) + float(yl);
      nrm = fmaf(yt, yt, nrm);
      if (side == 0) {
        slot[4 * d + 0] = yh;
        slot[4 * d + 1] = yh;
        slot[4 * d + 2] = yl;
        slot[4 * d + 3] = yl;
      } else {
        const __bf16 h2 = __bf16(2.f * float(yh)), l2 = __bf16(2.f * float(yl));
        slot[4 * d + 0] = h2;
        slot[4 * d + 1] = l2;
        slot[4 * d + 2] = h2;
        slot[4 * d + 3] = l2;
      }
    }
    const __bf16 nh = __bf16(-nrm);
    const __bf16 nl = __bf16(-nrm - float(nh));
    slot[12] = side == 0 ? nh : one;
    slot[13] = side == 0 ? nl : one;
    slot[14] = side == 0 ? one : nh;
    slot[15] = side == 0 ? one : nl;
    if (pad) {
#pragma unroll
      for (int q = 0; q < 16; ++q) slot[q] = zero;
      slot[12] = __bf16(-0x1p100f);
      slot[14] = one;
      slot[15] = one;
    }
  } else {
#pragma unroll
    for (int s = 0; s < 32; ++s) slot[s] = zero;
    const float sq = fmaf(x[2] * x[2], coef, fmaf(x[1] * x[1], coef, x[0] * x[0] * coef));
    __bf16 q[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      split3(side == 0 ? x[d] : -2.f * coef * x[d], q);
      // hH, hM, mH, hL, lH, mM: point side takes (h, h, m, h, l, m), row side
      // (H, M, H, L, H, M)
      const int lp[6] = {0, 0, 1, 0, 2, 1}, rp[6] = {0, 1, 0, 2, 0, 1};
#pragma unroll
      for (int u = 0; u < 6; ++u) slot[6 * d + u] = q[side == 0 ? lp[u] : rp[u]];
    }
    split3(sq, q);
#pragma unroll
    for (int u = 0; u < 3; ++u) {
      slot[18 + u] = side == 0 ? q[u] : one;
      slot[21 + u] = side == 0 ? one : q[u];
    }
    if (pad) {  // a point past n: exponent -2^100 whatever the row
#pragma unroll
      for (int s = 0; s < 32; ++s) slot[s] = zero;
      slot[18] = __bf16(-0x1p100f);
      slot[21] = one;
      slot[22] = one;
      slot[23] = one;
    }
  }
}

// Exponent slots of the split-BF16 RBF kernels: the compact form where the
// entries are split into two BF16 parts and the probes kept in one (NPROD
// <= 2), the full form for bf16x3 and the derivative kernel.
constexpr int exponent_slot_count(int nprod, bool dk) { return (dk || nprod == 3) ? 32 : 16; }

// The exponents of a tile from NK = NS / 16 BF16 MFMAs.
template <int NK>
__device__ __forceinline__ floatx16 rbf_exponents_mfma(const u32x4 (&l)[NK],
                                                       const u32x4 (&r)[NK]) {
  floatx16 d;
#pragma unroll
  for (int e = 0; e < 16; ++e) d[e] = 0.f;
#pragma unroll
  for (int k = 0; k < NK; ++k) d = mfma_bf16(l[k], r[k], d);
  return d;
}

// Point side of the exponent slots for every m < npad (NS / 8 u32x4 each:
// slots 0-7, 8-15, ...).
template <int NS>
__global__ void __launch_bounds__(256)
    k_prep_points(const float* __restrict__ X, int dim, float coef, float scale, int64_t n,
                  int64_t npad, u32x4* __restrict__ rec) {
  const int64_t m = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (m >= npad) return;
  float x[3] = {0.f, 0.f, 0.f};
  if (m < n)
    for (int d = 0; d < dim; ++d) x[d] = X[m * dim + d] - X[d];  // centred on point 0
  __bf16 slot[NS];
  exponent_slots<NS>(x, coef, scale, 0, m >= n, slot);
#pragma unroll
  for (int q = 0; q < NS / 8; ++q) {
    bf16x8 v;
#pragma unroll
    for (int j = 0; j < 8; ++j) v[j] = slot[8 * q + j];
    rec[NS / 8 * m + q] = __builtin_bit_cast(u32x4, v);
  }
}

// The BF16 MFMAs of one tile: acc[c] += A-fragment (ahi, alo) x probe block c
// of the staged tile (vh / vl: this lane's row of the hi / lo plane).
template <int NACC, int NPROD>
__device__ __forceinline__ void tile_mfmas(const __bf16* vh, const __bf16* vl,
                                           const u32x4 (&ahi)[2], const u32x4 (&alo)[2],
                                           floatx16 (&acc)[NACC]) {
  constexpr bool LO = NPROD == 3;
#pragma unroll
  for (int kb = 0; kb < 2; ++kb) {
    u32x4 bh[NACC], bl[NACC];
#pragma unroll
    for (int c = 0; c < NACC; ++c) {
      bh[c] = *reinterpret_cast<const u32x4*>(vh + 32 * c * LDB + 16 * kb);
      if (LO) bl[c] = *reinterpret_cast<const u32x4*>(vl + 32 * c * LDB + 16 * kb);
    }
    if (NPROD >= 2) {
#pragma unroll
      for (int c = 0; c < NACC; ++c) acc[c] = mfma_bf16(alo[kb], bh[c], acc[c]);
    }
    if (LO) {
#pragma unroll
      for (int c = 0; c < NACC; ++c) acc[c] = mfma_bf16(ahi[kb], bl[c], acc[c]);
    }
#pragma unroll
    for (int c = 0; c < NACC; ++c) acc[c] = mfma_bf16(ahi[kb], bh[c], acc[c]);
  }
}

// Stored-matrix BF16 matvec: the fragment of tile kt + 1 is loaded a tile
// ahead and split after the MFMAs of tile kt are issued; the probe tiles are
// fetched two tiles ahead into alternating register sets (the loop is
// unrolled by two so the sets keep fixed registers).
template <int NACC, bool VEC, int NPROD>
__global__ void __launch_bounds__(kThreads)
    k_matvec_bf16_dense(const float* __restrict__ A, const __bf16* __restrict__ Vhi_g,
                        const __bf16* __restrict__ Vlo_g, int64_t Ppad,
                        const float* __restrict__ V, int64_t P, int64_t n,
                        float* __restrict__ W) {
  constexpr int BPW = 32 * NACC;
  constexpr bool LO = NPROD == 3;  // probes' low plane used
  constexpr int PT = PlaneTile<NACC>::kPerThread;
  __shared__ __attribute__((aligned(16))) __bf16 Vhi[2][BPW * LDB];
  __shared__ __attribute__((aligned(16))) __bf16 Vlo[LO ? 2 : 1][LO ? BPW * LDB : 8];
  const int t = threadIdx.x;
  const int lane = t & 63, wave = t >> 6, r = lane & 31, h = lane >> 5;
  const int64_t i0w = int64_t(blockIdx.x) * BI + 32 * wave;
  const int64_t row = i0w + r;
  const int64_t p0 = int64_t(blockIdx.y) * BPW;
  const int64_t all_tiles = (n + BK - 1) / BK;
  const int64_t kt_begin = all_tiles * blockIdx.z / gridDim.z;
  const int64_t kt_end = all_tiles * (blockIdx.z + 1) / gridDim.z;
  const int64_t ntiles = kt_end - kt_begin;
  const int64_t mbase = kt_begin * BK;
  const bool split = gridDim.z > 1;
  float* Wout = split ? W + int64_t(blockIdx.z) * P * n : W;

  floatx16 acc[NACC];
#pragma unroll
  for (int c = 0; c < NACC; ++c)
#pragma unroll
    for (int e = 0; e < 16; ++e) acc[c][e] = 0.f;

  // prologue: probe tile 0 staged, tile 1 in flight
  u32x4 hA[PT], hB[PT], lA[PT], lB[PT];
  float areg[16];
  const int64_t kt1 = ntiles > 1 ? 1 : 0;
  load_plane_tile<NACC>(Vhi_g, Ppad, p0, kt_begin, hA);
  if (LO) load_plane_tile<NACC>(Vlo_g, Ppad, p0, kt_begin, lA);
  load_plane_tile<NACC>(Vhi_g, Ppad, p0, kt_begin + kt1, hB);
  if (LO) load_plane_tile<NACC>(Vlo_g, Ppad, p0, kt_begin + kt1, lB);
  store_plane_tile<NACC>(Vhi[0], hA);
  if (LO) store_plane_tile<NACC>(&Vlo[0][0], lA);
  load_a_frag16<VEC>(A, n, row, mbase, h, areg);
  __syncthreads();
  u32x4 fAh[2], fAl[2], fBh[2], fBl[2];  // fragments of the current / next tile
  {
    float a[16];
#pragma unroll
    for (int s = 0; s < 16; ++s) a[s] = areg[s];
    if (ntiles > 1) load_a_frag16<VEC>(A, n, row, mbase + BK, h, areg);
    split_frag<NPROD>(a, fAh, fAl);
  }

  auto step = [&](int64_t kt, u32x4 (&ih)[PT], u32x4 (&il)[PT], const u32x4 (&sh)[PT],
                  const u32x4 (&sl)[PT], const u32x4 (&ahi)[2], const u32x4 (&alo)[2],
                  u32x4 (&nhi)[2], u32x4 (&nlo)[2]) {
    const int buf = int(kt & 1);
    const int64_t m0 = mbase + kt * BK;
    const int64_t kt2 = (kt + 2 < ntiles) ? kt + 2 : ntiles - 1;
    load_plane_tile<NACC>(Vhi_g, Ppad, p0, kt_begin + kt2, ih);
    if (LO) load_plane_tile<NACC>(Vlo_g, Ppad, p0, kt_begin + kt2, il);
    tile_mfmas<NACC, NPROD>(Vhi[buf] + r * LDB + 8 * h, &Vlo[LO ? buf : 0][0] + r * LDB + 8 * h,
                            ahi, alo, acc);
    {
      float a[16];
#pragma unroll
      for (int s = 0; s < 16; ++s) a[s] = areg[s];
      if (kt + 2 < ntiles) load_a_frag16<VEC>(A, n, row, m0 + 2 * BK, h, areg);
      split_frag<NPROD>(a, nhi, nlo);
    }
    store_plane_tile<NACC>(Vhi[buf ^ 1], sh);
    if (LO) store_plane_tile<NACC>(&Vlo[LO ? (buf ^ 1) : 0][0], sl);
    __syncthreads();
  };

  int64_t kt = 0;
  for (; kt + 1 < ntiles; kt += 2) {
    step(kt, hA, lA, hB, lB, fAh, fAl, fBh, fBl);
    step(kt + 1, hB, lB, hA, lA, fBh, fBl, fAh, fAl);
  }
  if (kt < ntiles) step(kt, hA, lA, hB, lB, fAh, fAl, fBh, fBl);
  store_out<NACC, VEC, false>(Wout, V, P, n, i0w, p0, acc, 0.f);
}

// ---------------------------------------------------------------------------
// Packed storage of a stored matrix (slq_storage in slq.h): 32 x 32 tiles of
// the zero-padded matrix in MFMA fragment order, so that the 64 lanes of a
// wave read a tile plane as contiguous 1 KB blocks (lane l = r + 32 h holds
// row r and the 16 contraction slots frag_k(s, h) of that row).
//   BF16 planes: 2 KB per tile, [half kb][lane][8 bf16] (slots 8 kb .. 8 kb + 7);
//   FP8 planes:  1 KB per tile, [lane][16 bytes] (byte s = slot s), OCP FP8
//                E4M3 with one power-of-two scale per tile and plane.
// ---------------------------------------------------------------------------
// (kTileBf16Bytes, kTileFp8Bytes, storage_planes and storage_fp8: slq_check.h)

// Exponent k of the power-of-two tile scale: the smallest k with
// m 2^-k <= 448 (the largest finite E4M3 value).  m = f 2^e, f in [0.5, 1):
// m 2^-(e - 9) = f 2^9 <= 448 iff f <= 0.875.
__device__ __forceinline__ int fp8_scale_exp(float m) {
  if (!(m > 0.f)) return 0;
  int e;
  const float f = frexpf(m, &e);
  const int k = f <= 0.875f ? e - 9 : e - 8;
  return k < -120 ? -120 : k;
}

__device__ __forceinline__ float wave_max_f32(float v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = fmaxf(v, __shfl_xor(v, off, 64));
  return v;
}

// 16 values -> 16 E4M3 bytes of v 2^-k (round to nearest even), byte s =
// value s; and back to floats (times 2^k).
__device__ __forceinline__ u32x4 fp8_quantize16(const float (&v)[16], int k) {
  u32x4 q;
#pragma unroll
  for (int w = 0; w < 4; ++w) {
    int word = __builtin_amdgcn_cvt_pk_fp8_f32(ldexpf(v[4 * w + 0], -k), ldexpf(v[4 * w + 1], -k),
                                               0, false);
    word = __builtin_amdgcn_cvt_pk_fp8_f32(ldexpf(v[4 * w + 2], -k), ldexpf(v[4 * w + 3], -k),
                                           word, true);
    q[w] = uint32_t(word);
  }
  return q;
}

__device__ __forceinline__ void fp8_dequantize16(const u32x4& q, float scale, float (&v)[16]) {
#pragma unroll
  for (int w = 0; w < 4; ++w) {
    const f32x2 lo = __builtin_amdgcn_cvt_pk_f32_fp8(int(q[w]), false);
    const f32x2 hi = __builtin_amdgcn_cvt_pk_f32_fp8(int(q[w]), true);
    v[4 * w + 0] = lo[0] * scale;
    v[4 * w + 1] = lo[1] * scale;
    v[4 * w + 2] = hi[0] * scale;
    v[4 * w + 3] = hi[1] * scale;
  }
}

// One wave per tile: reads the tile's 1024 FP32 entries (zero past n) and
// writes its planes (and scales).
template <int STORAGE>
__global__ void __launch_bounds__(256)
    k_pack_matrix(const float* __restrict__ A, int64_t n, int64_t nt, void* __restrict__ packed,
                  float* __restrict__ scales) {
  const int lane = threadIdx.x & 63;
  const int64_t tile = int64_t(blockIdx.x) * 4 + (threadIdx.x >> 6);
  if (tile >= nt * nt) return;  // the whole wave
  const int64_t rb = tile / nt, ct = tile - rb * nt;
  const int r = lane & 31, h = lane >> 5;
  const int64_t row = rb * BK + r;
  float a[16];
#pragma unroll
  for (int s = 0; s < 16; ++s) {
    const int64_t col = ct * BK + frag_k(s, h);
    a[s] = (row < n && col < n) ? A[row * n + col] : 0.f;
  }
  const int64_t plane_tiles = nt * nt;
  if constexpr (!storage_fp8(STORAGE)) {
    __bf16* hi = static_cast<__bf16*>(packed);
    __bf16* lo = hi + plane_tiles * (kTileBf16Bytes / 2);
#pragma unroll
    for (int kb = 0; kb < 2; ++kb) {
      u32x4 hv, lv;
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        uint32_t hw, lw;
        split_pair(a[8 * kb + 2 * e], a[8 * kb + 2 * e + 1], hw, lw);
        hv[e] = hw;
        lv[e] = lw;
      }
      const int64_t off = tile * (kTileBf16Bytes / 2) + kb * 512 + lane * 8;
      *reinterpret_cast<u32x4*>(hi + off) = hv;
      if (STORAGE == SLQ_STORE_BF16X2) *reinterpret_cast<u32x4*>(lo + off) = lv;
    }
  } else {
    uint8_t* q = static_cast<uint8_t*>(packed);
    float m = 0.f;
#pragma unroll
    for (int s = 0; s < 16; ++s) m = fmaxf(m, fabsf(a[s]));
    const int k = fp8_scale_exp(wave_max_f32(m));
    const u32x4 qh = fp8_quantize16(a, k);
    *reinterpret_cast<u32x4*>(q + tile * kTileFp8Bytes + lane * 16) = qh;
    if (lane == 0) scales[tile] = ldexpf(1.f, k);
    if (STORAGE == SLQ_STORE_FP8X2) {
      // the residual a - s q is exact (both within a factor 2 of each other,
      // or q = 0)
      float d[16];
      fp8_dequantize16(qh, ldexpf(1.f, k), d);
      float res[16], ml = 0.f;
#pragma unroll
      for (int s = 0; s < 16; ++s) {
        res[s] = a[s] - d[s];
        ml = fmaxf(ml, fabsf(res[s]));
      }
      const int kl = fp8_scale_exp(wave_max_f32(ml));
      *reinterpret_cast<u32x4*>(q + plane_tiles * kTileFp8Bytes + tile * kTileFp8Bytes +
                                lane * 16) = fp8_quantize16(res, kl);
      if (lane == 0) scales[plane_tiles + tile] = ldexpf(1.f, kl);
    }
  }
}

// The lane's BF16 fragment (two halves) of one tile plane, from its raw
// packed form: BF16 words as stored, or E4M3 bytes times the tile scale
// (exact in BF16: 4 significant bits).
template <bool FP8>
struct PackedRaw;
template <>
struct PackedRaw<false> {
  u32x4 w[2];
  __device__ __forceinline__ void load(const uint8_t* tile_base, int lane) {
    const __bf16* b = reinterpret_cast<const __bf16*>(tile_base);
    w[0] = *reinterpret_cast<const u32x4*>(b + lane * 8);
    w[1] = *reinterpret_cast<const u32x4*>(b + 512 + lane * 8);
  }
  __device__ __forceinline__ void zero() { w[0] = w[1] = u32x4{0u, 0u, 0u, 0u}; }
  __device__ __forceinline__ void to_frag(float, u32x4 (&f)[2]) const {
    f[0] = w[0];
    f[1] = w[1];
  }
};
template <>
struct PackedRaw<true> {
  u32x4 w;
  __device__ __forceinline__ void load(const uint8_t* tile_base, int lane) {
    w = *reinterpret_cast<const u32x4*>(tile_base + lane * 16);
  }
  __device__ __forceinline__ void zero() { w = u32x4{0u, 0u, 0u, 0u}; }
  __device__ __forceinline__ void to_frag(float scale, u32x4 (&f)[2]) const {
#pragma unroll
    for (int kb = 0; kb < 2; ++kb)
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const int src = int(w[2 * kb + (e >> 1)]);
        f32x2 x = (e & 1) ? __builtin_amdgcn_cvt_pk_f32_fp8(src, true)
                          : __builtin_amdgcn_cvt_pk_f32_fp8(src, false);
        x[0] *= scale;
        x[1] *= scale;
        f[kb][e] = __builtin_bit_cast(uint32_t, __builtin_convertvector(x, bf16x2));
      }
  }
};

// Stored-matrix matvec on packed storage: each wave streams its row block's
// tile planes (DEPTH tiles in flight, in registers), forms the BF16
// fragments (FP8: converted with the tile scale) and runs the BF16 MFMAs of
// tile_mfmas (NPL = 2: lo*hi + hi*hi as bf16x2) against the BF16 plane of the
// probes, staged through LDS as in k_matvec_bf16_dense.
template <int NACC, bool VEC, bool FP8, int NPL, int DEPTH>
__global__ void __launch_bounds__(kThreads)
    k_matvec_packed(const uint8_t* __restrict__ packed, const float* __restrict__ scales,
                    int64_t nt, const __bf16* __restrict__ Vhi_g, int64_t Ppad,
                    const float* __restrict__ V, int64_t P, int64_t n, float* __restrict__ W) {
  constexpr int BPW = 32 * NACC;
  constexpr int PT = PlaneTile<NACC>::kPerThread;
  constexpr int64_t TB = FP8 ? kTileFp8Bytes : kTileBf16Bytes;
  __shared__ __attribute__((aligned(16))) __bf16 Vhi[2][BPW * LDB];
  const int t = threadIdx.x;
  const int lane = t & 63, wave = t >> 6, r = lane & 31, h = lane >> 5;
  const int64_t i0w = int64_t(blockIdx.x) * BI + 32 * wave;
  const int64_t rb = i0w / BK;
  const bool rows_live = rb < nt;  // a wave past the last row block reads nothing
  const int64_t p0 = int64_t(blockIdx.y) * BPW;
  const int64_t kt_begin = nt * blockIdx.z / gridDim.z;
  const int64_t kt_end = nt * (blockIdx.z + 1) / gridDim.z;
  const int64_t ntiles = kt_end - kt_begin;
  const bool split = gridDim.z > 1;
  float* Wout = split ? W + int64_t(blockIdx.z) * P * n : W;
  const int64_t plane_bytes = nt * nt * TB;
  const uint8_t* base = packed + ((rows_live ? rb : 0) * nt + kt_begin) * TB;
  const float* sbase = scales ? scales + (rows_live ? rb : 0) * nt + kt_begin : nullptr;

  floatx16 acc[NACC];
#pragma unroll
  for (int c = 0; c < NACC; ++c)
#pragma unroll
    for (int e = 0; e < 16; ++e) acc[c][e] = 0.f;

  PackedRaw<FP8> raw[DEPTH][NPL];
  float rsc[DEPTH][NPL];
  auto load_tile = [&](int64_t kt, PackedRaw<FP8> (&rw)[NPL], float (&sc)[NPL]) {
#pragma unroll
    for (int pl = 0; pl < NPL; ++pl) {
      if (rows_live && kt < ntiles) {
        rw[pl].load(base + pl * plane_bytes + kt * TB, lane);
        sc[pl] = FP8 ? sbase[pl * nt * nt + kt] : 1.f;
      } else {
        rw[pl].zero();
        sc[pl] = 0.f;
      }
    }
  };
#pragma unroll
  for (int d = 0; d < DEPTH; ++d) load_tile(d, raw[d], rsc[d]);

  u32x4 hA[PT], hB[PT];
  load_plane_tile<NACC>(Vhi_g, Ppad, p0, kt_begin, hA);
  load_plane_tile<NACC>(Vhi_g, Ppad, p0, kt_begin + (ntiles > 1 ? 1 : 0), hB);
  store_plane_tile<NACC>(Vhi[0], hA);
  __syncthreads();

  // step kt uses ring slot d = kt % DEPTH and the probe tile in LDS buffer
  // kt & 1; (ih, sh): plane registers loaded for kt + 2 / staged for kt + 1
  auto step = [&](int64_t kt, PackedRaw<FP8> (&rw)[NPL], float (&sc)[NPL], u32x4 (&ih)[PT],
                  const u32x4 (&sh)[PT]) {
    const int buf = int(kt & 1);
    u32x4 ahi[2], alo[2];
    rw[0].to_frag(sc[0], ahi);
    if constexpr (NPL == 2) rw[1].to_frag(sc[1], alo);
    load_tile(kt + DEPTH, rw, sc);
    const int64_t kt2 = (kt + 2 < ntiles) ? kt + 2 : ntiles - 1;
    load_plane_tile<NACC>(Vhi_g, Ppad, p0, kt_begin + kt2, ih);
    tile_mfmas<NACC, NPL == 2 ? 2 : 1>(Vhi[buf] + r * LDB + 8 * h, Vhi[buf] + r * LDB + 8 * h,
                                        ahi, alo, acc);
    store_plane_tile<NACC>(Vhi[buf ^ 1], sh);
    __syncthreads();
  };

  // the ring slot and the plane register set of a step are fixed by kt mod
  // (DEPTH * 2 / gcd) -- unroll by 2 DEPTH so both stay static
  int64_t kt = 0;
  for (; kt + 2 * DEPTH <= ntiles; kt += 2 * DEPTH) {
#pragma unroll
    for (int u = 0; u < 2 * DEPTH; ++u) {
      if (u % 2 == 0)
        step(kt + u, raw[u % DEPTH], rsc[u % DEPTH], hA, hB);
      else
        step(kt + u, raw[u % DEPTH], rsc[u % DEPTH], hB, hA);
    }
  }
#pragma unroll
  for (int u = 0; u < 2 * DEPTH; ++u) {
    if (kt + u < ntiles) {
      if (u % 2 == 0)
        step(kt + u, raw[u % DEPTH], rsc[u % DEPTH], hA, hB);
      else
        step(kt + u, raw[u % DEPTH], rsc[u % DEPTH], hB, hA);
    }
  }
  store_out<NACC, VEC, false>(Wout, V, P, n, i0w, p0, acc, 0.f);
}

// Derivative mode of the RBF BF16 matvec: the parameter contraction
//   sum_c <L_c, D R_c>,  D_im = d_im exp2(d_im),  d_im = c |x_i - x_m|^2
// (d_im exp2(d_im) = c l^3 dK_im/dl), together with the trace sum_c <L_c, R_c>,
// over the columns c < K at base + (c / Pin) ldc + (c % Pin) n.  The planes of
// R (columns c0 .. c0 + Ppad of this launch) are the staged operand; the
// epilogue multiplies the product tile by L in FP32 instead of storing it.
// Each workgroup writes its two partial sums (FP64) to red[2 wg + {0, 1}],
// wg = (z * gridDim.y + y) * gridDim.x + x; the trace comes from the z = 0
// slices only, so every (i, c) enters it once.
struct DKArgs {
  const float* L;
  const float* R;
  int64_t ldc, Pin, K, c0;
  double* red;
};

template <int NACC>
__device__ __forceinline__ void dk_epilogue(const floatx16 (&acc)[NACC], const DKArgs& dk,
                                            int64_t i0w, int64_t p0, int64_t n, bool trace) {
  const int t = threadIdx.x, lane = t & 63;
  float part = 0.f, tr = 0.f;
#pragma unroll
  for (int c = 0; c < NACC; ++c) {
    const int64_t cg = dk.c0 + p0 + 32 * c + (lane & 31);
    if (cg >= dk.K) continue;
    const int64_t off = (cg / dk.Pin) * dk.ldc + (cg % dk.Pin) * n;
    const float* lc = dk.L + off;
    const float* rc = dk.R + off;
#pragma unroll
    for (int g = 0; g < 4; ++g) {
      const int64_t i = i0w + 8 * g + 4 * (lane >> 5);
      if (i + 4 <= n) {
        const float4 lv = load4u(lc + i);
        part = fmaf(acc[c][4 * g + 0], lv.x, part);
        part = fmaf(acc[c][4 * g + 1], lv.y, part);
        part = fmaf(acc[c][4 * g + 2], lv.z, part);
        part = fmaf(acc[c][4 * g + 3], lv.w, part);
        if (trace) {
          const float4 rv = load4u(rc + i);
          tr = fmaf(lv.x, rv.x, tr);
          tr = fmaf(lv.y, rv.y, tr);
          tr = fmaf(lv.z, rv.z, tr);
          tr = fmaf(lv.w, rv.w, tr);
        }
      } else {
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          if (i + u < n) {
            part = fmaf(acc[c][4 * g + u], lc[i + u], part);
            if (trace) tr = fmaf(lc[i + u], rc[i + u], tr);
          }
        }
      }
    }
  }
  __shared__ double red[2][kThreads];
  red[0][t] = double(part);
  red[1][t] = double(tr);
  __syncthreads();
  for (int w = kThreads / 2; w > 0; w >>= 1) {
    if (t < w) {
      red[0][t] += red[0][t + w];
      red[1][t] += red[1][t + w];
    }
    __syncthreads();
  }
  if (t == 0) {
    const int64_t wg =
        (int64_t(blockIdx.z) * gridDim.y + blockIdx.y) * int64_t(gridDim.x) + blockIdx.x;
    dk.red[2 * wg] = red[0][0];
    dk.red[2 * wg + 1] = red[1][0];
  }
}

// RBF BF16 matvec.  Per tile and wave: two BF16 MFMAs give the exponents of
// the NEXT tile, the 2 NACC NPROD BF16 MFMAs of this tile follow, and the
// exp2 / split of the next tile's fragment is interleaved between them.  The
// probe tiles and the point records are fetched two tiles ahead into
// alternating register sets; the loop body is one basic block (clamped tile
// indices, unconditional staging of the other LDS buffer), so that the
// instruction order fixed below holds.  DK: the derivative mode above (the
// entries d exp2(d), the epilogue dk_epilogue), run by k_rbf_dk_contract.
template <int NACC, bool VEC, int NPROD, bool DK>
__device__ __forceinline__ void rbf_bf16_body(const u32x4* __restrict__ rec, const RBFParams& prm,
                                              const __bf16* __restrict__ Vhi_g,
                                              const __bf16* __restrict__ Vlo_g, int64_t Ppad,
                                              const float* __restrict__ V, int64_t P, int64_t n,
                                              float* __restrict__ W, const DKArgs& dk) {
  constexpr int BPW = 32 * NACC;
  constexpr bool LO = NPROD == 3;  // probes' low plane used
  // exponent slots: the compact form for the forward matvec with the probes
  // in one plane, the full form for bf16x3 and the derivative mode
  constexpr int NS = exponent_slot_count(NPROD, DK);
  constexpr int NK = NS / 16, RW = NS / 8;  // exponent MFMAs, record words
  constexpr int PT = PlaneTile<NACC>::kPerThread;
  __shared__ __attribute__((aligned(16))) __bf16 Vhi[2][BPW * LDB];
  __shared__ __attribute__((aligned(16))) __bf16 Vlo[LO ? 2 : 1][LO ? BPW * LDB : 8];
  const int t = threadIdx.x;
  const int lane = t & 63, wave = t >> 6, r = lane & 31, h = lane >> 5;
  const int64_t i0w = int64_t(blockIdx.x) * BI + 32 * wave;
  const int64_t row = i0w + r;
  const int64_t p0 = int64_t(blockIdx.y) * BPW;
  const int64_t all_tiles = (n + BK - 1) / BK;
  const int64_t kt_begin = all_tiles * blockIdx.z / gridDim.z;
  const int64_t kt_end = all_tiles * (blockIdx.z + 1) / gridDim.z;
  const int64_t ntiles = kt_end - kt_begin;
  const bool split = gridDim.z > 1;
  float* Wout = (split && !DK) ? W + int64_t(blockIdx.z) * P * n : W;
  const int dim = prm.dim;

  floatx16 acc[NACC];
#pragma unroll
  for (int c = 0; c < NACC; ++c)
#pragma unroll
    for (int e = 0; e < 16; ++e) acc[c][e] = 0.f;

  // lane constants: the row side of the exponent slots of this lane's row,
  // as the B operands of the NK exponent MFMAs (slots 16 k + 8 h + 0..7)
  u32x4 rr[NK];
  {
    float x[3] = {0.f, 0.f, 0.f};
    if (row < n)
      for (int d = 0; d < dim; ++d) x[d] = prm.X[row * dim + d] - prm.X[d];
    __bf16 slot[NS];
    exponent_slots<NS>(x, prm.coef, prm.scale, 1, false, slot);
#pragma unroll
    for (int q = 0; q < NK; ++q) {
      bf16x8 v0, v1;
#pragma unroll
      for (int j = 0; j < 8; ++j) {
        v0[j] = slot[16 * q + j];
        v1[j] = slot[16 * q + 8 + j];
      }
      const u32x4 w0 = __builtin_bit_cast(u32x4, v0), w1 = __builtin_bit_cast(u32x4, v1);
#pragma unroll
      for (int j = 0; j < 4; ++j) rr[q][j] = h ? w1[j] : w0[j];
    }
  }
  // point side of this lane's tile point (m0 + r) for tile kt (relative to
  // kt_begin): the A operands of the exponent MFMAs (records of RW words)
  const u32x4* prow = rec + RW * (kt_begin * BK + r) + h;
  auto load_pts = [&](int64_t kt, u32x4 (&l)[NK]) {
#pragma unroll
    for (int k = 0; k < NK; ++k) l[k] = prow[RW * BK * kt + 2 * k];
  };

  // prologue: probe tile 0 staged, tile 1 in flight, fragment of tile 0
  auto clamp = [&](int64_t kt) { return kt < ntiles ? kt : ntiles - 1; };
  u32x4 hA[PT], hB[PT], lA[PT], lB[PT];
  load_plane_tile<NACC>(Vhi_g, Ppad, p0, kt_begin, hA);
  if (LO) load_plane_tile<NACC>(Vlo_g, Ppad, p0, kt_begin, lA);
  load_plane_tile<NACC>(Vhi_g, Ppad, p0, kt_begin + clamp(1), hB);
  if (LO) load_plane_tile<NACC>(Vlo_g, Ppad, p0, kt_begin + clamp(1), lB);
  u32x4 xA[NK], xB[NK];
  load_pts(clamp(1), xA);
  store_plane_tile<NACC>(Vhi[0], hA);
  if (LO) store_plane_tile<NACC>(&Vlo[0][0], lA);
  u32x4 fAh[2], fAl[2], fBh[2], fBl[2];  // fragments of the current / next tile
  {
    u32x4 x0[NK];
    load_pts(0, x0);
    const floatx16 d = rbf_exponents_mfma(x0, rr);
    float a[16];
#pragma unroll
    for (int e = 0; e < 16; ++e)
      a[e] = DK ? d[e] * __builtin_amdgcn_exp2f(d[e]) : __builtin_amdgcn_exp2f(d[e]);
    split_frag<NPROD>(a, fAh, fAl);
  }
  __syncthreads();

  // One tile: issue the loads of tile kt + 2 into (ih, il, xi) -- they held
  // tile kt, staged one iteration ago -- run the exponent MFMAs of tile kt + 1
  // from xs and the BF16 MFMAs of tile kt with fragment (ahi, alo), build the
  // next fragment into (nhi, nlo), and stage tile kt + 1 from (sh, sl) into
  // the other LDS buffer (on the last tile this copy is unused).  Loads past
  // the last tile are clamped.
  //
  // Instruction order is fixed with scheduling barriers: the loads, the
  // exponent MFMAs, the B-fragment reads, kLead BF16 MFMAs while the
  // exponents complete, then the remaining BF16 MFMAs each followed by a
  // share of the eight exp2 / split chunks of the next fragment (so the VALU
  // work issues while the matrix core runs), and the LDS stores last.
  constexpr int kM = 2 * NPROD * NACC;  // BF16 MFMAs per tile
  constexpr int kLead = kM > 3 ? 3 : kM - 1;
  // The loads of step kt read tile min(kt + 2, ntiles - 1): three pointers
  // advanced by one tile per step while a further tile exists (cheaper than
  // recomputing 64-bit tile offsets each step).
  const u32x4* pnext = prow + RW * BK * clamp(2);
  const int64_t pstep = Ppad * BK;
  const __bf16* hnext = Vhi_g + ((kt_begin + clamp(2)) * Ppad + p0) * BK;
  const __bf16* lnext = LO ? Vlo_g + ((kt_begin + clamp(2)) * Ppad + p0) * BK : nullptr;
  auto step = [&](int64_t kt, u32x4 (&ih)[PT], u32x4 (&il)[PT], const u32x4 (&sh)[PT],
                  const u32x4 (&sl)[PT], u32x4 (&xi)[NK], const u32x4 (&xs)[NK],
                  const u32x4 (&ahi)[2], const u32x4 (&alo)[2], u32x4 (&nhi)[2],
                  u32x4 (&nlo)[2]) {
    const int buf = int(kt & 1);
    // points (needed one step later) ahead of the planes (needed at the end
    // of the next step)
#pragma unroll
    for (int k = 0; k < NK; ++k) xi[k] = pnext[2 * k];
    load_plane_tile_at<NACC>(hnext, ih);
    if (LO) load_plane_tile_at<NACC>(lnext, il);
    if (kt + 3 < ntiles) {
      pnext += RW * BK;
      hnext += pstep;
      if (LO) lnext += pstep;
    }
    __builtin_amdgcn_sched_barrier(0);
    const floatx16 dnext = rbf_exponents_mfma(xs, rr);
    const __bf16* vh = Vhi[buf] + r * LDB + 8 * h;
    const __bf16* vl = &Vlo[LO ? buf : 0][0] + r * LDB + 8 * h;
    u32x4 bh[2][NACC], bl[2][NACC];
#pragma unroll
    for (int kb = 0; kb < 2; ++kb)
#pragma unroll
      for (int c = 0; c < NACC; ++c) {
        bh[kb][c] = *reinterpret_cast<const u32x4*>(vh + 32 * c * LDB + 16 * kb);
        if (LO) bl[kb][c] = *reinterpret_cast<const u32x4*>(vl + 32 * c * LDB + 16 * kb);
      }
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int i = 0; i < kM; ++i) {
      // MFMA i: kb outer, then the product (lo*hi, hi*lo, hi*hi), then c
      const int kb = i / (NPROD * NACC);
      const int prod = (i % (NPROD * NACC)) / NACC;
      const int c = i % NACC;
      const bool a_lo = NPROD >= 2 && prod == 0;
      const bool b_lo = LO && prod == 1;
      acc[c] = mfma_bf16(a_lo ? alo[kb] : ahi[kb], b_lo ? bl[kb][c] : bh[kb][c], acc[c]);
      if (i >= kLead) {
        const int j = i - kLead, slots = kM - kLead;
#pragma unroll
        for (int e = j * 8 / slots; e < (j + 1) * 8 / slots; ++e) {
          float x0 = __builtin_amdgcn_exp2f(dnext[2 * e]);
          float x1 = __builtin_amdgcn_exp2f(dnext[2 * e + 1]);
          if (DK) {
            x0 *= dnext[2 * e];
            x1 *= dnext[2 * e + 1];
          }
          uint32_t hw, lw;
          if (NPROD >= 2) {
            split_pair(x0, x1, hw, lw);
            nlo[e / 4][e % 4] = lw;
          } else {
            const f32x2 x = {x0, x1};
            hw = __builtin_bit_cast(uint32_t, __builtin_convertvector(x, bf16x2));
          }
          nhi[e / 4][e % 4] = hw;
        }
      }
      __builtin_amdgcn_sched_barrier(0);
    }
    store_plane_tile<NACC>(Vhi[buf ^ 1], sh);
    if (LO) store_plane_tile<NACC>(&Vlo[LO ? (buf ^ 1) : 0][0], sl);
    __syncthreads();
  };

  int64_t kt = 0;
  for (; kt + 1 < ntiles; kt += 2) {
    step(kt, hA, lA, hB, lB, xB, xA, fAh, fAl, fBh, fBl);
    step(kt + 1, hB, lB, hA, lA, xA, xB, fBh, fBl, fAh, fAl);
  }
  if (kt < ntiles) step(kt, hA, lA, hB, lB, xB, xA, fAh, fAl, fBh, fBl);
  if constexpr (DK)
    dk_epilogue<NACC>(acc, dk, i0w, p0, n, blockIdx.z == 0);
  else if (split)
    store_out<NACC, VEC, false>(Wout, V, P, n, i0w, p0, acc, 0.f);
  else
    store_out<NACC, VEC, true>(Wout, V, P, n, i0w, p0, acc, prm.noise);
}

template <int NACC, bool VEC, int NPROD>
__global__ void __launch_bounds__(kThreads)
    k_matvec_rbf_bf16(const u32x4* __restrict__ rec, RBFParams prm,
                      const __bf16* __restrict__ Vhi_g, const __bf16* __restrict__ Vlo_g,
                      int64_t Ppad, const float* __restrict__ V, int64_t P, int64_t n,
                      float* __restrict__ W) {
  rbf_bf16_body<NACC, VEC, NPROD, false>(rec, prm, Vhi_g, Vlo_g, Ppad, V, P, n, W, DKArgs{});
}

// The derivative-mode contraction over the Ppad (multiple of 128) columns of
// the planes (hi and lo) of one launch.
__global__ void __launch_bounds__(kThreads)
    k_rbf_dk_contract(const u32x4* __restrict__ rec, RBFParams prm,
                      const __bf16* __restrict__ Vhi_g, const __bf16* __restrict__ Vlo_g,
                      int64_t Ppad, int64_t n, DKArgs dk) {
  rbf_bf16_body<4, false, 3, true>(rec, prm, Vhi_g, Vlo_g, Ppad, nullptr, Ppad, n, nullptr, dk);
}

// RBF matvec for a few vectors (P <= kRowsMaxP), as CG's single right-hand
// side needs it: the MFMA kernels pad the probe dimension to 32, so here the
// kernel is evaluated on the vector units instead, in FP32 from the points:
// with the coordinates scaled by s = sqrt(log2(e) / (2 l^2)) when they are
// loaded, k = exp2(-sum_d (s x_i - s x_j)^2).  A workgroup owns kRowsRB rows
// (kRowsPerLane per lane) and a column slice; chunks of 256 points (scaled
// coordinates, with the vector entry in the fourth slot for one vector of
// dimension <= 3: one 16-byte LDS read per point) are staged in LDS and each
// wave takes a quarter of every chunk; the four waves' sums are added in LDS
// in a fixed order.  Column slices > 1 write split-K partial sums.
constexpr int kRowsMaxP = 4;
constexpr int kRowsPerLane = 2;
constexpr int kRowsRB = 64 * kRowsPerLane;
constexpr int kRowsChunk = 256;
constexpr int kRowsMaxSplit = 16;

template <int NP, int DIM>
__global__ void __launch_bounds__(256)
    k_matvec_rbf_rows(const float* __restrict__ X, float scale, float noise,
                      const float* __restrict__ V, int64_t n, float* __restrict__ W) {
  constexpr bool PACKV = DIM <= 3 && NP == 1;  // vector entry in the fourth slot
  __shared__ float4 xs[kRowsChunk];
  __shared__ float vs[PACKV ? 1 : kRowsChunk * NP];
  __shared__ float red[3][NP][kRowsRB];
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  const int64_t j0 = n * blockIdx.y / gridDim.y, j1 = n * (blockIdx.y + 1) / gridDim.y;
  // points relative to point 0, as in the other RBF kernels
  float x0[DIM];
#pragma unroll
  for (int d = 0; d < DIM; ++d) x0[d] = n > 0 ? X[d] : 0.f;
  float xi[kRowsPerLane][DIM];
#pragma unroll
  for (int u = 0; u < kRowsPerLane; ++u) {
    const int64_t row = int64_t(blockIdx.x) * kRowsRB + 64 * u + lane;
#pragma unroll
    for (int d = 0; d < DIM; ++d) xi[u][d] = row < n ? (X[row * DIM + d] - x0[d]) * scale : 0.f;
  }
  float acc[kRowsPerLane][NP];
#pragma unroll
  for (int u = 0; u < kRowsPerLane; ++u)
#pragma unroll
    for (int q = 0; q < NP; ++q) acc[u][q] = 0.f;
  for (int64_t c0 = j0; c0 < j1; c0 += kRowsChunk) {
    const int64_t cnt = j1 - c0 < kRowsChunk ? j1 - c0 : kRowsChunk;
    {
      const int64_t j = c0 + t;
      const bool ok = t < cnt;
      float c[4] = {0.f, 0.f, 0.f, 0.f};
      if (ok) {
#pragma unroll
        for (int d = 0; d < DIM; ++d) c[d] = (X[j * DIM + d] - x0[d]) * scale;
        if (PACKV) c[3] = V[j];
      }
      xs[t] = make_float4(c[0], c[1], c[2], c[3]);
      if (!PACKV)
#pragma unroll
        for (int q = 0; q < NP; ++q) vs[t * NP + q] = ok ? V[q * n + j] : 0.f;
    }
    __syncthreads();
    const int jb = wave * (kRowsChunk / 4);
    const int je = jb + kRowsChunk / 4 < cnt ? jb + kRowsChunk / 4 : int(cnt);
    for (int jj = jb; jj < je; ++jj) {
      const float4 xj = xs[jj];
      const float xc[4] = {xj.x, xj.y, xj.z, xj.w};
      float vj[NP];
      if (PACKV) {
        vj[0] = xj.w;
      } else {
#pragma unroll
        for (int q = 0; q < NP; ++q) vj[q] = vs[jj * NP + q];
      }
#pragma unroll
      for (int u = 0; u < kRowsPerLane; ++u) {
        const float d0 = xi[u][0] - xc[0];
        float d2 = d0 * d0;
#pragma unroll
        for (int d = 1; d < DIM; ++d) {
          const float dd = xi[u][d] - xc[d];
          d2 = fmaf(dd, dd, d2);
        }
        const float k = __builtin_amdgcn_exp2f(-d2);
#pragma unroll
        for (int q = 0; q < NP; ++q) acc[u][q] = fmaf(k, vj[q], acc[u][q]);
      }
    }
    __syncthreads();
  }
  if (wave > 0)
#pragma unroll
    for (int u = 0; u < kRowsPerLane; ++u)
#pragma unroll
      for (int q = 0; q < NP; ++q) red[wave - 1][q][64 * u + lane] = acc[u][q];
  __syncthreads();
  if (wave == 0) {
    const bool split = gridDim.y > 1;
#pragma unroll
    for (int u = 0; u < kRowsPerLane; ++u) {
      const int64_t row = int64_t(blockIdx.x) * kRowsRB + 64 * u + lane;
      if (row >= n) continue;
      const int li = 64 * u + lane;
#pragma unroll
      for (int q = 0; q < NP; ++q) {
        const float sum = ((acc[u][q] + red[0][q][li]) + red[1][q][li]) + red[2][q][li];
        if (split)
          W[int64_t(blockIdx.y) * NP * n + q * n + row] = sum;
        else
          W[q * n + row] = fmaf(noise, V[q * n + row], sum);
      }
    }
  }
}

// The few-vector RBF matvec with the kernel's exponents from two BF16 MFMAs
// per 32 x 32 tile (the split exponent slots of the block matvec: ~3e-6 at
// most in the base-2 exponent, the size of an FP32 evaluation's rounding)
// and everything else in FP32 on the vector units, for point dimensions <= 3
// (the records' slots).  A workgroup owns 128 rows (32 per wave) and a column
// slice, staged through LDS in chunks of kMxTiles tiles (the point records
// and the vectors' entries, the next chunk's loads in flight while the
// current one is used).  Lane (r, h) of a wave takes row i0w + r and, of each
// tile, the 16 points m0 + 8 (s / 4) + 4 h + (s % 4) whose exponents the
// MFMAs leave in it; it forms their exp2 and accumulates them times the
// vectors' entries, and the two lane halves of a row are added at the end.
// Per kernel entry: one exp2 and NP FMAs.  Split-K partial sums as in
// k_matvec_rbf_rows.
constexpr int kMxTiles = 8;
template <int NP>
__global__ void __launch_bounds__(kThreads)
    k_matvec_rbf_rows_mx(const u32x4* __restrict__ rec, const float* __restrict__ X, int dim,
                         float coef, float noise, const float* __restrict__ V, int64_t n,
                         float* __restrict__ W) {
  constexpr int CP = kMxTiles * BK;        // points per chunk
  constexpr int RW = CP * 4 / kThreads;    // record words (u32x4) per thread and chunk
  __shared__ u32x4 rs[CP * 4];
  __shared__ __attribute__((aligned(16))) float vs[NP][CP];
  const int t = threadIdx.x;
  const int lane = t & 63, wave = t >> 6, r = lane & 31, h = lane >> 5;
  const int64_t row = int64_t(blockIdx.x) * BI + 32 * wave + r;
  const int64_t all_tiles = (n + BK - 1) / BK;
  const int64_t kt_begin = all_tiles * blockIdx.y / gridDim.y;
  const int64_t kt_end = all_tiles * (blockIdx.y + 1) / gridDim.y;
  const int64_t rec_words = all_tiles * BK * 4;
  u32x4 rr[2];
  {
    float x[3] = {0.f, 0.f, 0.f};
    if (row < n)
      for (int d = 0; d < dim; ++d) x[d] = X[row * dim + d] - X[d];
    __bf16 slot[32];
    exponent_slots<32>(x, coef, 0.f, 1, false, slot);
#pragma unroll
    for (int q = 0; q < 2; ++q) {
      bf16x8 v;
#pragma unroll
      for (int j = 0; j < 8; ++j) v[j] = slot[16 * q + 8 * h + j];
      rr[q] = __builtin_bit_cast(u32x4, v);
    }
  }
  // chunk starting at tile c0: record words (clamped to the array) and the
  // vectors' entries (zero past n)
  u32x4 wr[RW];
  float wv[NP];
  auto load_chunk = [&](int64_t c0) {
#pragma unroll
    for (int u = 0; u < RW; ++u) {
      int64_t w = c0 * BK * 4 + int64_t(u) * kThreads + t;
      w = w < rec_words ? w : rec_words - 1;
      wr[u] = rec[w];
    }
    const int64_t m = c0 * BK + t;
#pragma unroll
    for (int q = 0; q < NP; ++q) wv[q] = (t < CP && m < n) ? V[q * n + m] : 0.f;
  };
  float acc[NP];
#pragma unroll
  for (int q = 0; q < NP; ++q) acc[q] = 0.f;
  if (kt_begin < kt_end) load_chunk(kt_begin);
  for (int64_t c0 = kt_begin; c0 < kt_end; c0 += kMxTiles) {
#pragma unroll
    for (int u = 0; u < RW; ++u) rs[u * kThreads + t] = wr[u];
    if (t < CP)
#pragma unroll
      for (int q = 0; q < NP; ++q) vs[q][t] = wv[q];
    __syncthreads();
    if (c0 + kMxTiles < kt_end) load_chunk(c0 + kMxTiles);
    const int tiles = int(kt_end - c0 < kMxTiles ? kt_end - c0 : kMxTiles);
    for (int j = 0; j < tiles; ++j) {
      const u32x4 l[2] = {rs[(j * BK + r) * 4 + h], rs[(j * BK + r) * 4 + h + 2]};
      float vq[NP][16];
#pragma unroll
      for (int q = 0; q < NP; ++q)
#pragma unroll
        for (int g = 0; g < 4; ++g) {
          const float4 v4 = *reinterpret_cast<const float4*>(&vs[q][j * BK + 8 * g + 4 * h]);
          vq[q][4 * g + 0] = v4.x;
          vq[q][4 * g + 1] = v4.y;
          vq[q][4 * g + 2] = v4.z;
          vq[q][4 * g + 3] = v4.w;
        }
      const floatx16 d = rbf_exponents_mfma(l, rr);
#pragma unroll
      for (int sidx = 0; sidx < 16; ++sidx) {
        const float k = __builtin_amdgcn_exp2f(d[sidx]);
#pragma unroll
        for (int q = 0; q < NP; ++q) acc[q] = fmaf(k, vq[q][sidx], acc[q]);
      }
    }
    __syncthreads();
  }
#pragma unroll
  for (int q = 0; q < NP; ++q) acc[q] += __shfl_xor(acc[q], 32);
  if (h == 0 && row < n) {
    const bool split = gridDim.y > 1;
#pragma unroll
    for (int q = 0; q < NP; ++q) {
      if (split)
        W[int64_t(blockIdx.y) * NP * n + q * n + row] = acc[q];
      else
        W[q * n + row] = fmaf(noise, V[q * n + row], acc[q]);
    }
  }
}

template <int DIM>
void launch_rows_dim(hipStream_t s, const float* X, float scale, float noise, const float* V,
                     int64_t P, int64_t n, float* W, dim3 grid) {
  switch (P) {
    case 1:
      hipLaunchKernelGGL((k_matvec_rbf_rows<1, DIM>), grid, dim3(256), 0, s, X, scale, noise, V,
                         n, W);
      break;
    case 2:
      hipLaunchKernelGGL((k_matvec_rbf_rows<2, DIM>), grid, dim3(256), 0, s, X, scale, noise, V,
                         n, W);
      break;
    case 3:
      hipLaunchKernelGGL((k_matvec_rbf_rows<3, DIM>), grid, dim3(256), 0, s, X, scale, noise, V,
                         n, W);
      break;
    default:
      hipLaunchKernelGGL((k_matvec_rbf_rows<4, DIM>), grid, dim3(256), 0, s, X, scale, noise, V,
                         n, W);
      break;
  }
}

// scale = sqrt(log2(e) / (2 l^2)); dim in 1..4
hipError_t launch_rows(hipStream_t s, const float* X, int dim, float scale, float noise,
                       const float* V, int64_t P, int64_t n, float* W, int S) {
  const dim3 grid(unsigned((n + kRowsRB - 1) / kRowsRB), unsigned(S));
  switch (dim) {
    case 1: launch_rows_dim<1>(s, X, scale, noise, V, P, n, W, grid); break;
    case 2: launch_rows_dim<2>(s, X, scale, noise, V, P, n, W, grid); break;
    case 3: launch_rows_dim<3>(s, X, scale, noise, V, P, n, W, grid); break;
    default: launch_rows_dim<4>(s, X, scale, noise, V, P, n, W, grid); break;
  }
  return hipGetLastError();
}

// Column slices of the row kernel: up to kRowsMaxSplit, while there are fewer
// than 1024 workgroups, every slice keeps two chunks and the partial sums
// fit the scratch.
int choose_rows_splitk(int64_t n, int64_t P, int64_t avail_floats) {
  const int64_t wgs = (n + kRowsRB - 1) / kRowsRB;
  int S = 1;
  while (S < kRowsMaxSplit && wgs * S < 1024 && n / (2 * S) >= 2 * kRowsChunk &&
         int64_t(2 * S) * P * n <= avail_floats)
    S *= 2;
  return S;
}

inline bool aligned16(const void* ptr) { return (reinterpret_cast<uintptr_t>(ptr) & 15) == 0; }

// W = sum_z partial[z] (+ s2 V), summed in the fixed order z = 0..S-1.
template <bool NUGGET>
__global__ void __launch_bounds__(256)
    k_splitk_reduce(const float* __restrict__ partial, int S, int64_t count,
                    const float* __restrict__ V, float noise, float* __restrict__ W) {
  const int64_t stride = int64_t(gridDim.x) * blockDim.x;
  for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < count; i += stride) {
    float v = partial[i];
    for (int z = 1; z < S; ++z) v += partial[int64_t(z) * count + i];
    if (NUGGET) v = fmaf(noise, V[i], v);
    W[i] = v;
  }
}

template <int NACC, bool RBF, int NPROD>
hipError_t launch_tpl_bf16_n(hipStream_t s, const float* A, const u32x4* rec,
                             const RBFParams& prm, const __bf16* Vhi, const __bf16* Vlo,
                             int64_t Ppad, const float* V, int64_t P, int64_t n, float* W, bool vec,
                             int S) {
  const dim3 grid(unsigned((n + BI - 1) / BI), unsigned((P + 32 * NACC - 1) / (32 * NACC)),
                  unsigned(S));
  const dim3 block(kThreads);
  if constexpr (RBF) {
    if (vec)
      hipLaunchKernelGGL((k_matvec_rbf_bf16<NACC, true, NPROD>), grid, block, 0, s, rec, prm, Vhi,
                         Vlo, Ppad, V, P, n, W);
    else
      hipLaunchKernelGGL((k_matvec_rbf_bf16<NACC, false, NPROD>), grid, block, 0, s, rec, prm,
                         Vhi, Vlo, Ppad, V, P, n, W);
  } else {
    if (vec)
      hipLaunchKernelGGL((k_matvec_bf16_dense<NACC, true, NPROD>), grid, block, 0, s, A, Vhi, Vlo,
                         Ppad, V, P, n, W);
    else
      hipLaunchKernelGGL((k_matvec_bf16_dense<NACC, false, NPROD>), grid, block, 0, s, A, Vhi,
                         Vlo, Ppad, V, P, n, W);
  }
  return hipGetLastError();
}

template <int NACC, bool RBF>
hipError_t launch_tpl_bf16(hipStream_t s, const float* A, const u32x4* rec, const RBFParams& prm,
                           const __bf16* Vhi, const __bf16* Vlo, int64_t Ppad, const float* V,
                           int64_t P, int64_t n, float* W, bool vec, int nprod, int S) {
  if (nprod == 3)
    return launch_tpl_bf16_n<NACC, RBF, 3>(s, A, rec, prm, Vhi, Vlo, Ppad, V, P, n, W, vec, S);
  if (nprod == 2)
    return launch_tpl_bf16_n<NACC, RBF, 2>(s, A, rec, prm, Vhi, Vlo, Ppad, V, P, n, W, vec, S);
  return launch_tpl_bf16_n<NACC, RBF, 1>(s, A, rec, prm, Vhi, Vlo, Ppad, V, P, n, W, vec, S);
}

template <int NACC, bool RBF, int DIM>
hipError_t launch_tpl(hipStream_t s, const float* A, const RBFParams& prm, const float* V,
                      int64_t P, int64_t n, float* W, bool vec, int S) {
  const dim3 grid(unsigned((n + BI - 1) / BI), unsigned((P + 32 * NACC - 1) / (32 * NACC)),
                  unsigned(S));
  if (vec)
    hipLaunchKernelGGL((k_matvec_f32<NACC, RBF, DIM, true>), grid, dim3(kThreads), 0, s, A, prm,
                       V, P, n, W);
  else
    hipLaunchKernelGGL((k_matvec_f32<NACC, RBF, DIM, false>), grid, dim3(kThreads), 0, s, A, prm,
                       V, P, n, W);
  return hipGetLastError();
}

template <bool RBF, int DIM>
hipError_t launch_f32(hipStream_t s, const float* A, const RBFParams& prm, const float* V,
                      int64_t P, int64_t n, float* W, bool vec, int S) {
  if (P <= 32) return launch_tpl<1, RBF, DIM>(s, A, prm, V, P, n, W, vec, S);
  if (P <= 64) return launch_tpl<2, RBF, DIM>(s, A, prm, V, P, n, W, vec, S);
  return launch_tpl<4, RBF, DIM>(s, A, prm, V, P, n, W, vec, S);
}

int nacc_for(int64_t P) { return P <= 32 ? 1 : (P <= 64 ? 2 : 4); }

// Ring depth of the packed matvec: 8 KB of tile planes in flight per wave
// (BF16 planes 2 KB per tile, FP8 planes 1 KB).
template <int NACC, bool FP8, int NPL>
hipError_t launch_packed_n(hipStream_t s, const uint8_t* packed, const float* scales, int64_t nt,
                           const __bf16* Vhi, int64_t Ppad, const float* V, int64_t P, int64_t n,
                           float* W, bool vec, int S) {
  constexpr int DEPTH = FP8 ? (NPL == 2 ? 4 : 8) : (NPL == 2 ? 2 : 4);
  const dim3 grid(unsigned((n + BI - 1) / BI), unsigned((P + 32 * NACC - 1) / (32 * NACC)),
                  unsigned(S));
  if (vec)
    hipLaunchKernelGGL((k_matvec_packed<NACC, true, FP8, NPL, DEPTH>), grid, dim3(kThreads), 0, s,
                       packed, scales, nt, Vhi, Ppad, V, P, n, W);
  else
    hipLaunchKernelGGL((k_matvec_packed<NACC, false, FP8, NPL, DEPTH>), grid, dim3(kThreads), 0,
                       s, packed, scales, nt, Vhi, Ppad, V, P, n, W);
  return hipGetLastError();
}

template <int NACC>
hipError_t launch_packed_acc(hipStream_t s, int storage, const uint8_t* packed,
                             const float* scales, int64_t nt, const __bf16* Vhi, int64_t Ppad,
                             const float* V, int64_t P, int64_t n, float* W, bool vec, int S) {
  switch (storage) {
    case SLQ_STORE_BF16X2:
      return launch_packed_n<NACC, false, 2>(s, packed, scales, nt, Vhi, Ppad, V, P, n, W, vec, S);
    case SLQ_STORE_BF16:
      return launch_packed_n<NACC, false, 1>(s, packed, scales, nt, Vhi, Ppad, V, P, n, W, vec, S);
    case SLQ_STORE_FP8X2:
      return launch_packed_n<NACC, true, 2>(s, packed, scales, nt, Vhi, Ppad, V, P, n, W, vec, S);
    default:
      return launch_packed_n<NACC, true, 1>(s, packed, scales, nt, Vhi, Ppad, V, P, n, W, vec, S);
  }
}

hipError_t launch_packed(hipStream_t s, int storage, const uint8_t* packed, const float* scales,
                         int64_t nt, const __bf16* Vhi, int64_t Ppad, const float* V, int64_t P,
                         int64_t n, float* W, bool vec, int S) {
  switch (nacc_for(P)) {
    case 1:
      return launch_packed_acc<1>(s, storage, packed, scales, nt, Vhi, Ppad, V, P, n, W, vec, S);
    case 2:
      return launch_packed_acc<2>(s, storage, packed, scales, nt, Vhi, Ppad, V, P, n, W, vec, S);
    default:
      return launch_packed_acc<4>(s, storage, packed, scales, nt, Vhi, Ppad, V, P, n, W, vec, S);
  }
}

template <bool RBF>
hipError_t launch_bf16(hipStream_t s, const float* A, const u32x4* rec, const RBFParams& prm,
                       const __bf16* Vhi, const __bf16* Vlo, int64_t Ppad, const float* V,
                       int64_t P, int64_t n, float* W, bool vec, int nprod, int S) {
  switch (nacc_for(P)) {
    case 1:
      return launch_tpl_bf16<1, RBF>(s, A, rec, prm, Vhi, Vlo, Ppad, V, P, n, W, vec, nprod, S);
    case 2:
      return launch_tpl_bf16<2, RBF>(s, A, rec, prm, Vhi, Vlo, Ppad, V, P, n, W, vec, nprod, S);
    default:
      return launch_tpl_bf16<4, RBF>(s, A, rec, prm, Vhi, Vlo, Ppad, V, P, n, W, vec, nprod, S);
  }
}

// Split-K factor: enough workgroups for two per CU (256 CUs), as long as
// every slice keeps at least 8 tiles and the partial sums fit the scratch.
// The FP32 kernels for up to 32 probes use under 128 registers and hold four
// waves per SIMD, so they get twice the workgroups (more waves to hide the
// latency of their short per-tile chains).
int choose_splitk(int64_t n, int64_t P, int64_t avail_floats, bool fp32) {
  const int64_t nacc = nacc_for(P);
  const int64_t wgs = ((n + BI - 1) / BI) * ((P + 32 * nacc - 1) / (32 * nacc));
  const int64_t tiles = (n + BK - 1) / BK;
  const int64_t target = (fp32 && nacc == 1) ? 1024 : 512;
  int S = 1;
  while (S < kMaxSplitK && wgs * S < target && tiles / (2 * S) >= 8 &&
         int64_t(2 * S) * P * n <= avail_floats)
    S *= 2;
  return S;
}

// Derivative-kernel contraction: columns in launches of G column blocks of
// 128 (NACC = 4; planes of 128 G x npad split-BF16 values, G <= 8) and S
// column slices (the slices add to the partial sums, so they need no partial
// planes; any S, each slice keeping at least 8 tiles).  Two workgroups fit a
// CU, so a launch runs in rounds of 512; (G, S) is chosen so that the
// workgroup count nbx G S fills its last round best (nbx = n / 128 is rarely a
// multiple of 256), among counts of at least one round and G S <= 32.
constexpr int64_t kDkMaxGroups = 8;
struct DkPlan {
  int64_t cols, groups, nbx;
  int S;
};
DkPlan dk_plan(int64_t n, int64_t K) {
  DkPlan p;
  p.groups = (K + 127) / 128;
  p.nbx = (n + BI - 1) / BI;
  const int64_t tiles = (n + BK - 1) / BK;
  const int64_t gmax = p.groups < kDkMaxGroups ? p.groups : kDkMaxGroups;
  int64_t bestG = gmax;
  int bestS = 1;
  double best = -1.0;
  for (int64_t G = gmax; G >= 1; --G) {
    for (int S = 1; S <= 16 && G * S <= 32; ++S) {
      if (S > 1 && tiles / S < 8) break;
      const double rounds = double(p.nbx * G * S) / 512.0;
      if (rounds < 1.0 && !(G == gmax && S == 1)) continue;
      const double fill = rounds / std::ceil(rounds);
      if (fill > best + 1e-3) {
        best = fill;
        bestG = G;
        bestS = S;
      }
    }
  }
  p.cols = bestG * 128;
  p.S = bestS;
  return p;
}

}  // namespace

bool packed_size(int64_t n, int storage, int64_t* bytes, int64_t* scale_count) {
  static_assert(kPackTile == BK, "packed tiles are the matvec's contraction tiles");
  return packed_layout(n, storage, bytes, scale_count);
}

hipError_t launch_pack_matrix(hipStream_t s, const float* A, int64_t n, int storage, void* packed,
                              float* scales) {
  const int64_t nt = (n + BK - 1) / BK;
  const dim3 grid(unsigned((nt * nt + 3) / 4));
  switch (storage) {
    case SLQ_STORE_BF16X2:
      hipLaunchKernelGGL(k_pack_matrix<SLQ_STORE_BF16X2>, grid, dim3(256), 0, s, A, n, nt, packed,
                         scales);
      break;
    case SLQ_STORE_BF16:
      hipLaunchKernelGGL(k_pack_matrix<SLQ_STORE_BF16>, grid, dim3(256), 0, s, A, n, nt, packed,
                         scales);
      break;
    case SLQ_STORE_FP8X2:
      hipLaunchKernelGGL(k_pack_matrix<SLQ_STORE_FP8X2>, grid, dim3(256), 0, s, A, n, nt, packed,
                         scales);
      break;
    case SLQ_STORE_FP8:
      hipLaunchKernelGGL(k_pack_matrix<SLQ_STORE_FP8>, grid, dim3(256), 0, s, A, n, nt, packed,
                         scales);
      break;
    default:
      return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

int64_t matvec_scratch_floats(int64_t n, int64_t P) {
  const int64_t bpw = 32 * nacc_for(P);
  const int64_t Ppad = (P + bpw - 1) / bpw * bpw;
  const int64_t npad = (n + BK - 1) / BK * BK;
  // the split-K slices a launch takes when scratch does not limit them (the
  // larger of the FP32 and the BF16 choice, and the few-vector row kernel's)
  const int64_t unlimited = int64_t(1) << 60;
  int S = choose_splitk(n, P, unlimited, true);
  const int S16 = choose_splitk(n, P, unlimited, false);
  if (S16 > S) S = S16;
  if (P <= kRowsMaxP) {
    const int Sr = choose_rows_splitk(n, P, unlimited);
    if (Sr > S) S = Sr;
  }
  // partial sums + two BF16 planes + the RBF point records (16 floats each)
  return int64_t(S) * P * n + Ppad * npad + 16 * npad;
}

PlaneOut matvec_planes(const slq_operator& op, int64_t P, float* scratch) {
  PlaneOut out;
  const bool rows = op.kind == SLQ_OP_RBF && P <= kRowsMaxP && op.dim >= 1 && op.dim <= 4;
  const bool packed = op.kind == SLQ_OP_DENSE && op.storage != SLQ_STORE_FP32;
  if (rows || (op.precision == SLQ_PREC_FP32 && !packed) || !scratch) return out;
  const int64_t npad = (op.n + BK - 1) / BK * BK;
  const int64_t bpw = 32 * nacc_for(P);
  out.Ppad = (P + bpw - 1) / bpw * bpw;
  out.hi = reinterpret_cast<__bf16*>(scratch);
  out.lo = (!packed && op.precision == SLQ_PREC_BF16X3) ? out.hi + out.Ppad * npad : nullptr;
  out.perm = op.kind == SLQ_OP_RBF ? 1 : 0;
  return out;
}

hipError_t launch_matvec(hipStream_t s, const slq_operator& op, const float* V, int64_t P,
                         float* W, float* scratch, int64_t scratch_floats, hipEvent_t ev0,
                         hipEvent_t ev1, int flags, MvSplitSum* defer) {
  if (defer) *defer = MvSplitSum{};
  const int64_t n = op.n;
  if (n <= 0 || P <= 0) return hipSuccess;
  if (op.kind == SLQ_OP_RBF && P <= kRowsMaxP && op.dim >= 1 && op.dim <= 4) {
    // a few vectors: the row kernels in FP32 (every precision gets FP32
    // here); for dimension <= 3 with the exponents from the matrix cores,
    // the point records at the start of the scratch, the partial sums after
    const float noise = float(op.noise);
    const int64_t npad = (n + BK - 1) / BK * BK;
    const bool mx = op.dim <= 3 && scratch && scratch_floats >= 16 * npad;
    float* partial = scratch ? (mx ? scratch + 16 * npad : scratch) : nullptr;
    const int S = partial ? choose_rows_splitk(n, P, scratch_floats - (mx ? 16 * npad : 0)) : 1;
    float* out = S > 1 ? partial : W;
    hipError_t err;
    if (mx) {
      const float coef = float(-1.4426950408889634 / (2.0 * op.lengthscale * op.lengthscale));
      u32x4* rec = reinterpret_cast<u32x4*>(scratch);
      if (!(flags & kMvPointsReady)) {
        hipLaunchKernelGGL(k_prep_points<32>, dim3(unsigned((npad + 255) / 256)), dim3(256), 0,
                           s, op.points, op.dim, coef, 0.f, n, npad, rec);
        err = hipGetLastError();
        if (err != hipSuccess) return err;
      }
      const dim3 grid(unsigned((n + BI - 1) / BI), unsigned(S));
      if (ev0) (void)hipEventRecord(ev0, s);
      switch (P) {
        case 1:
          hipLaunchKernelGGL(k_matvec_rbf_rows_mx<1>, grid, dim3(kThreads), 0, s, rec, op.points,
                             op.dim, coef, noise, V, n, out);
          break;
        case 2:
          hipLaunchKernelGGL(k_matvec_rbf_rows_mx<2>, grid, dim3(kThreads), 0, s, rec, op.points,
                             op.dim, coef, noise, V, n, out);
          break;
        case 3:
          hipLaunchKernelGGL(k_matvec_rbf_rows_mx<3>, grid, dim3(kThreads), 0, s, rec, op.points,
                             op.dim, coef, noise, V, n, out);
          break;
        default:
          hipLaunchKernelGGL(k_matvec_rbf_rows_mx<4>, grid, dim3(kThreads), 0, s, rec, op.points,
                             op.dim, coef, noise, V, n, out);
          break;
      }
      err = hipGetLastError();
      if (ev1) (void)hipEventRecord(ev1, s);
    } else {
      const float scale =
          float(sqrt(1.4426950408889634 / (2.0 * op.lengthscale * op.lengthscale)));
      if (ev0) (void)hipEventRecord(ev0, s);
      err = launch_rows(s, op.points, op.dim, scale, noise, V, P, n, out, S);
      if (ev1) (void)hipEventRecord(ev1, s);
    }
    if (err != hipSuccess || S == 1) return err;
    if (defer) {
      *defer = MvSplitSum{partial, S, V, noise, 1};
      return hipSuccess;
    }
    const int64_t count = P * n;
    int64_t blocks = (count + 255) / 256;
    if (blocks > 4096) blocks = 4096;
    hipLaunchKernelGGL(k_splitk_reduce<true>, dim3(unsigned(blocks)), dim3(256), 0, s, partial,
                       S, count, V, noise, W);
    return hipGetLastError();
  }
  const bool vec = (n % 4 == 0) && aligned16(V) && aligned16(W);
  const bool packed = op.kind == SLQ_OP_DENSE && op.storage != SLQ_STORE_FP32;
  const bool fp32 = op.precision == SLQ_PREC_FP32 && !packed;
  const int nprod = packed ? 1
                           : (op.precision == SLQ_PREC_BF16X3
                                  ? 3
                                  : (op.precision == SLQ_PREC_BF16X2 ? 2 : 1));
  // scratch layout: [BF16 planes, RBF point records (bf16 path only)]
  // [split-K partial sums]
  int64_t planes_floats = 0;
  __bf16 *Vhi = nullptr, *Vlo = nullptr;
  u32x4* rec = nullptr;
  const int64_t npad = (n + BK - 1) / BK * BK;
  const int64_t bpw = 32 * nacc_for(P);
  const int64_t Ppad = (P + bpw - 1) / bpw * bpw;
  if (!fp32) {
    if (!scratch) return hipErrorInvalidValue;
    planes_floats = Ppad * npad + 16 * npad;
    if (planes_floats > scratch_floats) return hipErrorInvalidValue;
    Vhi = reinterpret_cast<__bf16*>(scratch);
    Vlo = Vhi + Ppad * npad;
    rec = reinterpret_cast<u32x4*>(scratch + Ppad * npad);
    if (!(flags & kMvPlanesReady)) {
      int64_t blocks = (Ppad * npad / 4 + 255) / 256;
      if (blocks > 8192) blocks = 8192;
      __bf16* lo = nprod == 3 ? Vlo : nullptr;  // the low plane is read by bf16x3 only
      if (op.kind == SLQ_OP_RBF)
        hipLaunchKernelGGL(k_split_probes<true>, dim3(unsigned(blocks)), dim3(256), 0, s, V, P, n,
                           Ppad, npad, Vhi, lo);
      else
        hipLaunchKernelGGL(k_split_probes<false>, dim3(unsigned(blocks)), dim3(256), 0, s, V, P,
                           n, Ppad, npad, Vhi, lo);
      hipError_t e = hipGetLastError();
      if (e != hipSuccess) return e;
    }
  }
  float* partial = scratch ? scratch + planes_floats : nullptr;
  const int S = partial ? choose_splitk(n, P, scratch_floats - planes_floats, fp32) : 1;
  float* out = (S > 1) ? partial : W;
  RBFParams prm{};
  hipError_t err;
  const float rbf_coef = float(-1.4426950408889634 / (2.0 * op.lengthscale * op.lengthscale));
  const float rbf_scale = float(sqrt(1.4426950408889634 / (2.0 * op.lengthscale * op.lengthscale)));
  if (op.kind == SLQ_OP_RBF && !fp32 && !(flags & kMvPointsReady)) {
    const dim3 pgrid(unsigned((npad + 255) / 256));
    if (exponent_slot_count(nprod, false) == 16)
      hipLaunchKernelGGL(k_prep_points<16>, pgrid, dim3(256), 0, s, op.points, op.dim, rbf_coef,
                         rbf_scale, n, npad, rec);
    else
      hipLaunchKernelGGL(k_prep_points<32>, pgrid, dim3(256), 0, s, op.points, op.dim, rbf_coef,
                         rbf_scale, n, npad, rec);
    err = hipGetLastError();
    if (err != hipSuccess) return err;
  }
  if (ev0) (void)hipEventRecord(ev0, s);
  if (packed) {
    err = launch_packed(s, op.storage, static_cast<const uint8_t*>(op.packed), op.tile_scales,
                        npad / BK, Vhi, Ppad, V, P, n, out, vec, S);
  } else if (op.kind == SLQ_OP_DENSE) {
    const bool v2 = vec && aligned16(op.matrix);
    err = fp32 ? launch_f32<false, 1>(s, op.matrix, prm, V, P, n, out, v2, S)
               : launch_bf16<false>(s, op.matrix, nullptr, prm, Vhi, Vlo, Ppad, V, P, n, out, v2,
                                    nprod, S);
  } else {
    prm.X = op.points;
    prm.dim = op.dim;
    prm.coef = rbf_coef;
    prm.scale = rbf_scale;
    prm.noise = float(op.noise);
    if (!fp32) {
      if (op.dim > 3) return hipErrorInvalidValue;  // rejected by the API before launch
      err = launch_bf16<true>(s, nullptr, rec, prm, Vhi, Vlo, Ppad, V, P, n, out, vec, nprod, S);
    } else {
      switch (op.dim) {
        case 1: err = launch_f32<true, 1>(s, nullptr, prm, V, P, n, out, vec, S); break;
        case 2: err = launch_f32<true, 2>(s, nullptr, prm, V, P, n, out, vec, S); break;
        case 3: err = launch_f32<true, 3>(s, nullptr, prm, V, P, n, out, vec, S); break;
        default: err = launch_f32<true, 4>(s, nullptr, prm, V, P, n, out, vec, S); break;
      }
    }
  }
  if (ev1) (void)hipEventRecord(ev1, s);
  if (err != hipSuccess || S == 1) return err;
  if (defer) {
    *defer = MvSplitSum{partial, S, V, prm.noise, op.kind == SLQ_OP_RBF ? 1 : 0};
    return hipSuccess;
  }
  const int64_t count = P * n;
  int64_t blocks = (count + 255) / 256;
  if (blocks > 4096) blocks = 4096;
  if (op.kind == SLQ_OP_RBF)
    hipLaunchKernelGGL(k_splitk_reduce<true>, dim3(unsigned(blocks)), dim3(256), 0, s, partial, S,
                       count, V, prm.noise, W);
  else
    hipLaunchKernelGGL(k_splitk_reduce<false>, dim3(unsigned(blocks)), dim3(256), 0, s, partial,
                       S, count, V, 0.f, W);
  return hipGetLastError();
}

int64_t rbf_dk_scratch_floats(int64_t n, int64_t K) {
  const int64_t npad = (n + BK - 1) / BK * BK;
  return dk_plan(n, K).cols * npad + 16 * npad;
}

int64_t rbf_dk_partials(int64_t n, int64_t K) {
  const DkPlan p = dk_plan(n, K);
  return p.nbx * p.S * p.groups;
}

hipError_t launch_rbf_dk_contract(hipStream_t s, const float* X, int dim, double lengthscale,
                                  const float* L, const float* R, int64_t ldc, int64_t Pin,
                                  int64_t n, int64_t K, float* scratch, double* red) {
  if (n <= 0 || K <= 0) return hipSuccess;
  if (dim < 1 || dim > 3) return hipErrorInvalidValue;  // rejected by the API before launch
  const DkPlan pl = dk_plan(n, K);
  const int64_t npad = (n + BK - 1) / BK * BK;
  __bf16* Vhi = reinterpret_cast<__bf16*>(scratch);
  __bf16* Vlo = Vhi + pl.cols * npad;
  u32x4* rec = reinterpret_cast<u32x4*>(scratch + pl.cols * npad);
  RBFParams prm{};
  prm.X = X;
  prm.dim = dim;
  prm.coef = float(-1.4426950408889634 / (2.0 * lengthscale * lengthscale));
  static_assert(exponent_slot_count(3, true) == 32, "derivative mode: full exponent form");
  hipLaunchKernelGGL(k_prep_points<32>, dim3(unsigned((npad + 255) / 256)), dim3(256), 0, s, X,
                     dim, prm.coef, 0.f, n, npad, rec);
  hipError_t err = hipGetLastError();
  if (err != hipSuccess) return err;
  const int64_t gx = (npad / 4 + 255) / 256;
  int64_t g0 = 0;
  for (int64_t c0 = 0; c0 < K; c0 += pl.cols) {
    const int64_t left = (K - c0 + 127) / 128 * 128;
    const int64_t cols = left < pl.cols ? left : pl.cols;
    const int64_t ny = cols / 128;
    hipLaunchKernelGGL(k_split_columns, dim3(unsigned(gx), unsigned(cols)), dim3(256), 0, s, R,
                       ldc, Pin, n, K, c0, cols, npad, Vhi, Vlo);
    DKArgs dk{L, R, ldc, Pin, K, c0, red + 2 * pl.nbx * pl.S * g0};
    hipLaunchKernelGGL(k_rbf_dk_contract, dim3(unsigned(pl.nbx), unsigned(ny), unsigned(pl.S)),
                       dim3(kThreads), 0, s, rec, prm, Vhi, Vlo, cols, n, dk);
    err = hipGetLastError();
    if (err != hipSuccess) return err;
    g0 += ny;
  }
  return hipSuccess;
}

}  // namespace slq
