// Per-probe tridiagonal eigen-solve and Gauss quadrature in FP64.
//
// Replaces, for every probe, funm.dense_funm_sym_eigh(log) applied to the
// materialised tridiagonal matrix (funm.py:330-333, decomp.py:133-152) and the
// read-out |z|^2 e1^T f(T) e1 (funm.py:240-241), followed by the probe mean of
// stochtrace.py:50.  Only the first row of the eigenvector matrix enters
// e1^T U log(L) U^T e1 = sum_j U_0j^2 log(theta_j) (Golub-Welsch).
//
// Algorithm (one wave per probe, one eigenvalue per lane):
//  * theta_j by bisection on the Sturm count of T - x I (the number of
//    negative pivots of its LDL^T factorisation), to full FP64 resolution;
//  * the weight U_0j^2 from the residue of the resolvent entry
//    g(x) = [(T - x I)^{-1}]_00 = sum_j U_0j^2 / (theta_j - x).  Its
//    reciprocal r_0(x) is the bottom-up continued fraction
//      r_{m-1} = a_{m-1} - x,   r_i = a_i - x - b_i^2 / r_{i+1},
//    and U_0j^2 = -1 / r_0'(theta_j) with
//      r_i' = -1 + b_i^2 r_{i+1}' / r_{i+1}^2.
//    Every term of r_i' has the same sign (r_i' <= -1), so the recurrence
//    has no cancellation.
// Eigenvalues outside the leading unreduced block (below the first
// negligible off-diagonal, |b_m| <= eps (|a_m| + |a_{m+1}|)) have U_0j = 0 and
// do not contribute; the solver works on that block.  A non-positive Ritz
// value anywhere in T yields NaN like the reference's unguarded log
// (funm.py:201, 332), where 0 * log(theta) is NaN as well.
//
// The tridiagonal is extracted from the Arnoldi Hessenberg matrix exactly as
// decomp.py:133-135 does it, in FP32: alpha_i = H_ii and
// beta_i = 0.5 * (H_{i,i+1} + H_{i+1,i}); only then is it promoted to FP64.
#include <cfloat>

#include "slq_internal.h"

namespace slq {
namespace {

// Depths up to this keep (a, |b|, b^2) in LDS (at most 3 * 8 * 2048 = 48 KB).
constexpr int kLdsDepth = 2048;
constexpr int kMaxBisect = 256;

__device__ __forceinline__ double wave_sum_f64(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

__device__ __forceinline__ int wave_min_int(int v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = min(v, __shfl_xor(v, off, 64));
  return v;
}

__device__ __forceinline__ double wave_min_f64(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = fmin(v, __shfl_xor(v, off, 64));
  return v;
}

__device__ __forceinline__ double wave_max_f64(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
  return v;
}

// Number of eigenvalues of the leading m x m block that are < x.
__device__ __forceinline__ int sturm_count(const double* __restrict__ a,
                                           const double* __restrict__ b2, int m, double x,
                                           double pivmin) {
  double q = a[0] - x;
  if (fabs(q) <= pivmin) q = -pivmin;
  int cnt = q < 0.0;
  for (int i = 1; i < m; ++i) {
    q = (a[i] - x) - b2[i - 1] / q;
    if (fabs(q) <= pivmin) q = -pivmin;
    cnt += q < 0.0;
  }
  return cnt;
}

// U_0j^2 = -1 / r_0'(theta) for the leading m x m block.
__device__ __forceinline__ double first_component_sq(const double* __restrict__ a,
                                                     const double* __restrict__ b2, int m,
                                                     double theta, double pivmin) {
  double r = a[m - 1] - theta;
  double rp = -1.0;
  for (int i = m - 2; i >= 0; --i) {
    if (fabs(r) <= pivmin) r = -pivmin;
    const double t = 1.0 / r;
    const double bt = b2[i] * t;
    rp = -1.0 + bt * t * rp;
    r = (a[i] - theta) - bt;
  }
  return -1.0 / rp;
}

// One wave per probe.  (a, ab, b2) point at 3 arrays of k doubles (LDS or
// global workspace).
__device__ void quadrature_wave(const float* __restrict__ Hp, int k, double* a, double* ab,
                                double* b2, float* alpha, float* beta, double len,
                                double* out) {
  const int lane = threadIdx.x;
  // tridiagonal of T = (H + H^T)/2 in FP32 (decomp.py:133-135); H column-major
  for (int i = lane; i < k; i += 64) {
    const float av = Hp[int64_t(i) * k + i];
    a[i] = double(av);
    if (alpha) alpha[i] = av;
    if (i + 1 < k) {
      const float b = 0.5f * (Hp[int64_t(i + 1) * k + i] + Hp[int64_t(i) * k + i + 1]);
      ab[i] = fabs(double(b));
      b2[i] = double(b) * double(b);
      if (beta) beta[i] = b;
    } else {
      ab[i] = 0.0;
      b2[i] = 0.0;
    }
  }
  __syncthreads();

  // leading unreduced block [0, m), Gershgorin bounds, pivot threshold
  int split = k;
  double gl = DBL_MAX, gu = -DBL_MAX, bmax = 0.0;
  for (int i = lane; i < k; i += 64) {
    if (i + 1 < k && ab[i] <= DBL_EPSILON * (fabs(a[i]) + fabs(a[i + 1])))
      split = min(split, i + 1);
    const double rad = ab[i] + (i > 0 ? ab[i - 1] : 0.0);
    gl = fmin(gl, a[i] - rad);
    gu = fmax(gu, a[i] + rad);
    bmax = fmax(bmax, b2[i]);
  }
  split = wave_min_int(split);
  gl = wave_min_f64(gl);
  gu = wave_max_f64(gu);
  bmax = wave_max_f64(bmax);
  const double pivmin = DBL_MIN * fmax(1.0, bmax);
  const double wide = fmax(fabs(gl), fabs(gu));
  gl -= 2.0 * DBL_EPSILON * wide * k + 2.0 * pivmin;
  gu += 2.0 * DBL_EPSILON * wide * k + 2.0 * pivmin;

  // Any eigenvalue of T at or below zero makes the reference's log NaN.
  const bool nonpositive = sturm_count(a, b2, k, 0.0, pivmin) > 0;

  double acc = 0.0;
  const int m = split;
  for (int j = lane; j < m; j += 64) {
    double lo = gl, hi = gu;
    for (int it = 0; it < kMaxBisect; ++it) {
      const double mid = 0.5 * (lo + hi);
      if (mid <= lo || mid >= hi) break;
      if (hi - lo <= 2.0 * DBL_EPSILON * fmax(fabs(lo), fabs(hi))) break;
      if (sturm_count(a, b2, m, mid, pivmin) > j)
        hi = mid;
      else
        lo = mid;
    }
    const double theta = 0.5 * (lo + hi);
    const double w = first_component_sq(a, b2, m, theta, pivmin);
    acc += w * log(theta);
  }
  acc = wave_sum_f64(acc);
  if (lane == 0) *out = nonpositive ? __builtin_nan("") : len * len * acc;
}

__global__ void __launch_bounds__(64)
    k_tridiag_quadrature_lds(const float* __restrict__ H, const float* __restrict__ scale,
                             int64_t k, double* __restrict__ per_probe,
                             float* __restrict__ alpha, float* __restrict__ beta) {
  extern __shared__ double lds[];  // 3 k doubles
  const int64_t p = blockIdx.x;
  quadrature_wave(H + p * k * k, int(k), lds, lds + k, lds + 2 * k,
                  alpha ? alpha + p * k : nullptr, beta ? beta + p * (k - 1) : nullptr,
                  double(scale[p]), per_probe + p);
}

// Arrays in a global workspace (3 k P doubles) for depths beyond kLdsDepth.
__global__ void __launch_bounds__(64)
    k_tridiag_quadrature_global(const float* __restrict__ H, const float* __restrict__ scale,
                                int64_t k, double* __restrict__ work,
                                double* __restrict__ per_probe, float* __restrict__ alpha,
                                float* __restrict__ beta) {
  const int64_t p = blockIdx.x;
  double* w = work + p * 3 * k;
  quadrature_wave(H + p * k * k, int(k), w, w + k, w + 2 * k, alpha ? alpha + p * k : nullptr,
                  beta ? beta + p * (k - 1) : nullptr, double(scale[p]), per_probe + p);
}

// Probe mean in a fixed (sequential) order: deterministic for a given P.
__global__ void k_probe_mean(const double* __restrict__ per_probe, int64_t P,
                             double* __restrict__ mean) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    double s = 0.0;
    for (int64_t p = 0; p < P; ++p) s += per_probe[p];
    *mean = s / double(P);
  }
}

}  // namespace

hipError_t launch_tridiag_quadrature(hipStream_t s, const float* H, const float* scale, int64_t P,
                                     int64_t k, double* work, double* per_probe, double* mean,
                                     float* alpha, float* beta) {
  if (P <= 0) return hipSuccess;
  if (k > 0) {
    if (k <= kLdsDepth)
      hipLaunchKernelGGL(k_tridiag_quadrature_lds, dim3(unsigned(P)), dim3(64),
                         size_t(3 * k) * sizeof(double), s, H, scale, k, per_probe, alpha, beta);
    else
      hipLaunchKernelGGL(k_tridiag_quadrature_global, dim3(unsigned(P)), dim3(64), 0, s, H, scale,
                         k, work, per_probe, alpha, beta);
    hipError_t err = hipGetLastError();
    if (err != hipSuccess) return err;
  }
  if (mean) hipLaunchKernelGGL(k_probe_mean, dim3(1), dim3(64), 0, s, per_probe, P, mean);
  return hipGetLastError();
}

}  // namespace slq
