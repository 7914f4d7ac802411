// Block matvec W = A V over all probes at once: the reference's matvec
// (decomp.py:460), vmapped over probes (stochtrace.py:49), becomes one dense
// contraction W[p][i] = sum_m A[i][m] V[p][m] on the matrix cores.
//
// FP32 path: v_mfma_f32_32x32x2f32.  A workgroup (4 waves) owns a tile of
// BI = 128 output elements x BP = 64 probes and walks the contraction index m
// in steps of BK = 32 through LDS.  Each wave owns 64 x 32 outputs (two 32x32
// accumulators).  Within an MFMA the two lane halves carry two different m;
// lane half h walks m in [h * BK/2, (h + 1) * BK/2) of the tile so that each
// lane reads its operands as contiguous float4s from LDS.
//
// Operators:
//  * DENSE: the A tile is read from the stored row-major matrix;
//  * RBF:   the A tile is generated in place, K_im = exp2(c |x_i - x_m|^2)
//           with c = -log2(e) / (2 l^2); the nugget s2 * v is added in FP32
//           in the epilogue (never rounded into a low-precision tile).
#include "slq_internal.h"

namespace slq {
namespace {

typedef float floatx16 __attribute__((ext_vector_type(16)));

constexpr int BI = 128;
constexpr int BP = 64;
constexpr int BK = 32;
constexpr int LDK = BK + 4;  // padded LDS row (floats)
constexpr int kThreads = 256;

struct RBFParams {
  const float* X;
  int dim;
  float coef;  // -log2(e) / (2 l^2)
  float noise;
};

// Loads of the V tile: 64 probes x 32 m = 512 float4, two per thread.
template <bool VEC>
__device__ __forceinline__ void load_v_tile(const float* __restrict__ V, int64_t P, int64_t n,
                                            int64_t p0, int64_t m0, float4 (&reg)[2]) {
  const int t = threadIdx.x;
#pragma unroll
  for (int u = 0; u < 2; ++u) {
    const int f = t + u * kThreads;
    const int row = f >> 3, c4 = f & 7;
    const int64_t p = p0 + row, m = m0 + 4 * c4;
    float4 v = make_float4(0.f, 0.f, 0.f, 0.f);
    if (p < P) {
      const float* src = V + p * n + m;
      if (VEC) {
        if (m < n) v = *reinterpret_cast<const float4*>(src);
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

__device__ __forceinline__ void store_v_tile(float* Vs, const float4 (&reg)[2]) {
  const int t = threadIdx.x;
#pragma unroll
  for (int u = 0; u < 2; ++u) {
    const int f = t + u * kThreads;
    const int row = f >> 3, c4 = f & 7;
    *reinterpret_cast<float4*>(Vs + row * LDK + 4 * c4) = reg[u];
  }
}

// Loads of a stored A tile: 128 rows x 32 m = 1024 float4, four per thread.
template <bool VEC>
__device__ __forceinline__ void load_a_tile(const float* __restrict__ A, int64_t n, int64_t i0,
                                            int64_t m0, float4 (&reg)[4]) {
  const int t = threadIdx.x;
#pragma unroll
  for (int u = 0; u < 4; ++u) {
    const int f = t + u * kThreads;
    const int row = f >> 3, c4 = f & 7;
    const int64_t i = i0 + row, m = m0 + 4 * c4;
    float4 v = make_float4(0.f, 0.f, 0.f, 0.f);
    if (i < n) {
      const float* src = A + i * n + m;
      if (VEC) {
        if (m < n) v = *reinterpret_cast<const float4*>(src);
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

__device__ __forceinline__ void store_a_tile(float* As, const float4 (&reg)[4]) {
  const int t = threadIdx.x;
#pragma unroll
  for (int u = 0; u < 4; ++u) {
    const int f = t + u * kThreads;
    const int row = f >> 3, c4 = f & 7;
    *reinterpret_cast<float4*>(As + row * LDK + 4 * c4) = reg[u];
  }
}

// Generate an RBF tile: thread t owns row (t >> 1) and 16 columns starting at
// 16 * (t & 1).  xi holds the thread's own point (fixed for the kernel).
__device__ __forceinline__ void gen_rbf_tile(float* As, const float* Xs, const float (&xi)[4],
                                             int dim, float coef, int64_t n, int64_t m0) {
  const int t = threadIdx.x;
  const int row = t >> 1, cbase = 16 * (t & 1);
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    float vals[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int col = cbase + 4 * q + u;
      float d2 = 0.f;
      for (int d = 0; d < dim; ++d) {
        const float diff = xi[d] - Xs[col * 4 + d];
        d2 = fmaf(diff, diff, d2);
      }
      const float kv = __builtin_amdgcn_exp2f(coef * d2);
      vals[u] = (m0 + col < n) ? kv : 0.f;
    }
    *reinterpret_cast<float4*>(As + row * LDK + cbase + 4 * q) =
        make_float4(vals[0], vals[1], vals[2], vals[3]);
  }
}

// One BK-deep step of MFMAs from the LDS tiles.  The tile's partial sums
// start from zero and are added to the running total afterwards (blocked
// summation): the rounding error of an n-term contraction then grows with
// BK + n / BK instead of n, as in a blocked CPU GEMM.
__device__ __forceinline__ void mfma_tile(const float* As, const float* Vs, int wi, int wp,
                                          floatx16 (&acc)[2]) {
  const int lane = threadIdx.x & 63;
  const int r = lane & 31, h = lane >> 5;
  const float* a0p = As + (wi + r) * LDK + h * (BK / 2);
  const float* a1p = As + (wi + 32 + r) * LDK + h * (BK / 2);
  const float* bp = Vs + (wp + r) * LDK + h * (BK / 2);
  floatx16 t0, t1;
#pragma unroll
  for (int i = 0; i < 16; ++i) { t0[i] = 0.f; t1[i] = 0.f; }
#pragma unroll
  for (int q = 0; q < BK / 8; ++q) {
    const float4 a0 = *reinterpret_cast<const float4*>(a0p + 4 * q);
    const float4 a1 = *reinterpret_cast<const float4*>(a1p + 4 * q);
    const float4 b = *reinterpret_cast<const float4*>(bp + 4 * q);
    t0 = __builtin_amdgcn_mfma_f32_32x32x2f32(a0.x, b.x, t0, 0, 0, 0);
    t1 = __builtin_amdgcn_mfma_f32_32x32x2f32(a1.x, b.x, t1, 0, 0, 0);
    t0 = __builtin_amdgcn_mfma_f32_32x32x2f32(a0.y, b.y, t0, 0, 0, 0);
    t1 = __builtin_amdgcn_mfma_f32_32x32x2f32(a1.y, b.y, t1, 0, 0, 0);
    t0 = __builtin_amdgcn_mfma_f32_32x32x2f32(a0.z, b.z, t0, 0, 0, 0);
    t1 = __builtin_amdgcn_mfma_f32_32x32x2f32(a1.z, b.z, t1, 0, 0, 0);
    t0 = __builtin_amdgcn_mfma_f32_32x32x2f32(a0.w, b.w, t0, 0, 0, 0);
    t1 = __builtin_amdgcn_mfma_f32_32x32x2f32(a1.w, b.w, t1, 0, 0, 0);
  }
#pragma unroll
  for (int i = 0; i < 16; ++i) {
    acc[0][i] += t0[i];
    acc[1][i] += t1[i];
  }
}

// Epilogue: lane l of a 32x32 accumulator holds column (l & 31) and rows
// 8 * (r / 4) + 4 * (l / 32) + (r % 4), r = 0..15.
template <bool VEC, bool NUGGET>
__device__ __forceinline__ void store_out(float* __restrict__ W, const float* __restrict__ V,
                                          int64_t P, int64_t n, int64_t i0, int64_t p0, int wi,
                                          int wp, const floatx16 (&acc)[2], float noise) {
  const int lane = threadIdx.x & 63;
  const int64_t p = p0 + wp + (lane & 31);
  if (p >= P) return;
#pragma unroll
  for (int mi = 0; mi < 2; ++mi) {
#pragma unroll
    for (int g = 0; g < 4; ++g) {
      const int64_t i = i0 + wi + mi * 32 + 8 * g + 4 * (lane >> 5);
      float o[4] = {acc[mi][4 * g + 0], acc[mi][4 * g + 1], acc[mi][4 * g + 2],
                    acc[mi][4 * g + 3]};
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

template <bool VEC>
__global__ void __launch_bounds__(kThreads)
    k_matvec_dense_f32(const float* __restrict__ A, const float* __restrict__ V, int64_t P,
                       int64_t n, float* __restrict__ W) {
  __shared__ __attribute__((aligned(16))) float As[BI * LDK];
  __shared__ __attribute__((aligned(16))) float Vs[BP * LDK];
  const int64_t i0 = int64_t(blockIdx.x) * BI;
  const int64_t p0 = int64_t(blockIdx.y) * BP;
  const int wave = threadIdx.x >> 6;
  const int wi = (wave & 1) * 64, wp = (wave >> 1) * 32;
  floatx16 acc[2];
#pragma unroll
  for (int r = 0; r < 16; ++r) { acc[0][r] = 0.f; acc[1][r] = 0.f; }

  float4 areg[4], vreg[2];
  load_a_tile<VEC>(A, n, i0, 0, areg);
  load_v_tile<VEC>(V, P, n, p0, 0, vreg);
  for (int64_t m0 = 0; m0 < n; m0 += BK) {
    store_a_tile(As, areg);
    store_v_tile(Vs, vreg);
    __syncthreads();
    if (m0 + BK < n) {  // prefetch the next tile while the MFMAs run
      load_a_tile<VEC>(A, n, i0, m0 + BK, areg);
      load_v_tile<VEC>(V, P, n, p0, m0 + BK, vreg);
    }
    mfma_tile(As, Vs, wi, wp, acc);
    __syncthreads();
  }
  store_out<VEC, false>(W, V, P, n, i0, p0, wi, wp, acc, 0.f);
}

template <bool VEC>
__global__ void __launch_bounds__(kThreads)
    k_matvec_rbf_f32(RBFParams prm, const float* __restrict__ V, int64_t P, int64_t n,
                     float* __restrict__ W) {
  __shared__ __attribute__((aligned(16))) float As[BI * LDK];
  __shared__ __attribute__((aligned(16))) float Vs[BP * LDK];
  __shared__ float Xs[BK * 4];
  const int64_t i0 = int64_t(blockIdx.x) * BI;
  const int64_t p0 = int64_t(blockIdx.y) * BP;
  const int t = threadIdx.x;
  const int wave = t >> 6;
  const int wi = (wave & 1) * 64, wp = (wave >> 1) * 32;
  floatx16 acc[2];
#pragma unroll
  for (int r = 0; r < 16; ++r) { acc[0][r] = 0.f; acc[1][r] = 0.f; }

  float xi[4] = {0.f, 0.f, 0.f, 0.f};
  {
    const int64_t i = i0 + (t >> 1);
    if (i < n)
      for (int d = 0; d < prm.dim; ++d) xi[d] = prm.X[i * prm.dim + d];
  }
  float4 vreg[2];
  load_v_tile<VEC>(V, P, n, p0, 0, vreg);
  for (int64_t m0 = 0; m0 < n; m0 += BK) {
    if (t < BK * 4) {
      const int col = t >> 2, d = t & 3;
      const int64_t m = m0 + col;
      Xs[t] = (m < n && d < prm.dim) ? prm.X[m * prm.dim + d] : 0.f;
    }
    store_v_tile(Vs, vreg);
    __syncthreads();
    gen_rbf_tile(As, Xs, xi, prm.dim, prm.coef, n, m0);
    if (m0 + BK < n) load_v_tile<VEC>(V, P, n, p0, m0 + BK, vreg);
    __syncthreads();
    mfma_tile(As, Vs, wi, wp, acc);
    __syncthreads();
  }
  store_out<VEC, true>(W, V, P, n, i0, p0, wi, wp, acc, prm.noise);
}

inline bool aligned16(const void* ptr) { return (reinterpret_cast<uintptr_t>(ptr) & 15) == 0; }

}  // namespace

hipError_t launch_matvec(hipStream_t s, const slq_operator& op, const float* V, int64_t P,
                         float* W) {
  const int64_t n = op.n;
  if (n <= 0 || P <= 0) return hipSuccess;
  const dim3 grid(unsigned((n + BI - 1) / BI), unsigned((P + BP - 1) / BP));
  const dim3 block(kThreads);
  const bool vec = (n % 4 == 0) && aligned16(V) && aligned16(W);
  if (op.kind == SLQ_OP_DENSE) {
    const bool v2 = vec && aligned16(op.matrix);
    if (v2)
      hipLaunchKernelGGL(k_matvec_dense_f32<true>, grid, block, 0, s, op.matrix, V, P, n, W);
    else
      hipLaunchKernelGGL(k_matvec_dense_f32<false>, grid, block, 0, s, op.matrix, V, P, n, W);
  } else {
    RBFParams prm;
    prm.X = op.points;
    prm.dim = op.dim;
    prm.coef = float(-1.4426950408889634 / (2.0 * op.lengthscale * op.lengthscale));
    prm.noise = float(op.noise);
    if (vec)
      hipLaunchKernelGGL(k_matvec_rbf_f32<true>, grid, block, 0, s, prm, V, P, n, W);
    else
      hipLaunchKernelGGL(k_matvec_rbf_f32<false>, grid, block, 0, s, prm, V, P, n, W);
  }
  return hipGetLastError();
}

}  // namespace slq
