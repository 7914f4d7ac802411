// Per-probe tridiagonal eigen-solves in FP64: the Gauss quadrature of the
// log-determinant integrand, the eigen cotangent of the reverse pass and the
// small part of funm_lanczos_sym.
//
// Replaces, for every probe, funm.dense_funm_sym_eigh(log) applied to the
// materialised tridiagonal matrix (funm.py:330-333, decomp.py:133-152) and the
// read-out |z|^2 e1^T f(T) e1 (funm.py:240-241), followed by the probe mean of
// stochtrace.py:50.  Only the first row of the eigenvector matrix enters
// e1^T U log(L) U^T e1 = sum_j U_0j^2 log(theta_j) (Golub-Welsch).
//
// Algorithm (per probe):
//  * theta_j by bisection on the Sturm count of T - x I (the number of
//    negative pivots of its LDL^T factorisation), to full FP64 resolution --
//    for depths up to 64 as a multisection with four lanes per eigenvalue
//    (k_tridiag_quadrature_ms), beyond that one lane per eigenvalue;
//  * the weight U_0j^2 from the twisted factorisation of T - theta_j I
//    (Parlett and Dhillon, "Fernando's solution to Wilkinson's problem",
//    1997): the stationary pivots D+_i (top-down LDL^T) and D-_i (bottom-up
//    UDU^T), the twist r = argmin_i |gamma_i| with
//    gamma_i = D+_i + D-_i - (a_i - theta), and the eigenvector z with
//    z_r = 1, z_i = -(b_i / D+_i) z_{i+1} (i < r),
//    z_{i+1} = -(b_i / D-_{i+1}) z_i (i >= r); then U_0j^2 = z_0^2 / |z|^2,
//    accumulated as ratios so that no component is formed.  The plain residue
//    of the resolvent entry [(T - x I)^{-1}]_00 is the special case r = 0 and
//    is useless when U_0j is tiny, which is what the copies of converged
//    Ritz values in a Lanczos tridiagonal without reorthogonalisation have;
//    at r = argmin |gamma| the vector is accurate to O(eps |T| / gap).
//  * If the weights do not sum to one to kWeightSumTol (eigenvalues that
//    coincide to working precision, where the twisted vectors of a cluster
//    are not orthogonal), the wave discards them and takes U_0j^2 from
//    implicit QL with the first row of U accumulated (orthogonal by
//    construction).
// Eigenvalues outside the leading unreduced block (below the first
// negligible off-diagonal, |b_m| <= eps (|a_m| + |a_{m+1}|)) have U_0j = 0 and
// do not contribute; the solver works on that block.  A non-positive Ritz
// value anywhere in T yields NaN like the reference's unguarded log
// (funm.py:201, 332), where 0 * log(theta) is NaN as well.
//
// The tridiagonal is extracted from the Arnoldi Hessenberg matrix exactly as
// decomp.py:133-135 does it, in FP32: alpha_i = H_ii and
// beta_i = 0.5 * (H_{i,i+1} + H_{i+1,i}); only then is it promoted to FP64.
//
// Reverse pass (k_tridiag_vjp): the full eigendecomposition T = U diag(theta)
// U^T (implicit QL, eigenvectors accumulated) and the cotangent of
// [log T]_00 in the Daleckii-Krein form Tbar = U (L o u u^T) U^T with
// u = U[0, :] and L the divided differences of log at the eigenvalues --
// what the eigh VJP composed with fA[0, 0] evaluates to (funm.py:331-333).
// Through alpha_i = H_ii and beta_i = (H_{i,i+1} + H_{i+1,i}) / 2 this gives
// dH_ii = Tbar_ii and dH_{i,i+1} = dH_{i+1,i} = Tbar_{i,i+1}, scaled by
// |z|^2 / P.
//
// funm_lanczos_sym (k_tridiag_funm): f(T) e_1 = U f(theta) U^T e_1 from the
// same eigendecomposition (funm.py:145-147, dense_funm_sym_eigh).
#include <cfloat>

#include "slq_internal.h"

namespace slq {
namespace {

// Depths up to this keep (a, |b|, b^2) and the 64 lanes' bottom-up pivots in
// LDS (8 * (3 + 64) * 64 = 33.5 KB).
constexpr int kQuadLdsDepth = 64;
constexpr int kMaxBisect = 256;
constexpr int kMaxSweeps = 64;
// Twisted weights summing to one within this are accepted.
constexpr double kWeightSumTol = 1e-9;

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

// 1 / q for the Sturm recurrence: the hardware reciprocal refined by two
// Newton steps (q is finite and |q| >= pivmin there, so the special cases a
// full division handles do not arise).  The recurrence is a serial chain of
// these, so the shorter sequence is what bounds the bisection's speed.
__device__ __forceinline__ double sturm_recip(double q) {
  double r = __builtin_amdgcn_rcp(q);
  r = fma(r, fma(-q, r, 1.0), r);
  r = fma(r, fma(-q, r, 1.0), r);
  return r;
}

// Number of eigenvalues of the leading m x m block that are < x.
__device__ __forceinline__ int sturm_count(const double* __restrict__ a,
                                           const double* __restrict__ b2, int m, double x,
                                           double pivmin) {
  double q = a[0] - x;
  if (fabs(q) <= pivmin) q = -pivmin;
  int cnt = q < 0.0;
  for (int i = 1; i < m; ++i) {
    q = fma(-b2[i - 1], sturm_recip(q), a[i] - x);
    if (fabs(q) <= pivmin) q = -pivmin;
    cnt += q < 0.0;
  }
  return cnt;
}

// U_0j^2 of the eigenvalue theta of the leading m x m block from the twisted
// factorisation (see the file header).  dm: this lane's column of m doubles
// at stride 64 (the bottom-up pivots, kept for the twist search).  Running
// ratios: sp_i = sum_{l<=i} (z_l / z_i)^2, qq_i = (z_0 / z_i)^2 / sp_i (in
// [0, 1], so that it cannot overflow) and sm_r = sum_{l>=r} (z_l / z_r)^2;
// then z_0^2 / |z|^2 = qq_r / (1 + (sm_r - 1) / sp_r).
__device__ __forceinline__ double twisted_weight(const double* __restrict__ a,
                                                 const double* __restrict__ b2, int m,
                                                 double theta, double pivmin, double* dm) {
  double d = a[m - 1] - theta;
  if (fabs(d) <= pivmin) d = -pivmin;
  dm[(m - 1) * 64] = d;
  for (int i = m - 2; i >= 0; --i) {
    d = (a[i] - theta) - b2[i] / d;
    if (fabs(d) <= pivmin) d = -pivmin;
    dm[i * 64] = d;
  }
  double dp = a[0] - theta;
  if (fabs(dp) <= pivmin) dp = -pivmin;
  double sp = 1.0, qq = 1.0;
  double best = fabs(dp + dm[0] - (a[0] - theta));
  int r = 0;
  double sp_r = 1.0, qq_r = 1.0;
  for (int i = 1; i < m; ++i) {
    const double inv = 1.0 / dp;
    const double bt = b2[i - 1] * inv;  // b_{i-1}^2 / D+_{i-1}
    const double x = bt * inv * sp;     // (z_{i-1} / z_i)^2 sp_{i-1}
    qq *= x / (1.0 + x);
    sp = 1.0 + x;
    dp = (a[i] - theta) - bt;
    if (fabs(dp) <= pivmin) dp = -pivmin;
    const double g = fabs(dp + dm[i * 64] - (a[i] - theta));
    if (g < best) {
      best = g;
      r = i;
      sp_r = sp;
      qq_r = qq;
    }
  }
  double sm = 1.0;
  for (int i = m - 2; i >= r; --i) {
    const double inv = 1.0 / dm[(i + 1) * 64];
    sm = 1.0 + b2[i] * inv * inv * sm;
  }
  return qq_r / (1.0 + (sm - 1.0) / sp_r);
}

// ---------------------------------------------------------------------------
// Full eigendecomposition of T (from H as above) by implicit QL with the
// eigenvectors accumulated.  One wave per probe; (d, e) are k doubles, U is
// rows x k row-major with row stride ldu (U[a * ldu + j] = component a of
// eigenvector j; rows = k for the whole matrix, 1 for its first row only), in
// LDS for k <= kEigLdsDepth and in a global workspace beyond.  The QL
// recurrence is evaluated redundantly by every lane (uniform control flow);
// the rotations of U are spread over the lanes by rows.  Returns false if QL
// did not converge.  On return d holds the eigenvalues (unordered).
//
// The lanes of a rotation touch U[a * ldu + i] for a = lane: with an odd
// stride in doubles (eig_ld) these fall into distinct LDS bank pairs, where a
// stride of k (a multiple of 32 doubles for k = 32, 64) put all 64 lanes on
// one bank.  The wave routines of this file are inlined into their kernels:
// a called function sees generic pointers and reaches LDS arrays through
// flat instructions, with longer latency on these serial recurrences.
// ---------------------------------------------------------------------------
constexpr int kEigLdsDepth = 64;
__host__ __device__ constexpr int eig_ld(int k) { return k | 1; }

__device__ __forceinline__ bool tridiag_eigvecs_wave(const float* __restrict__ Hp, int k,
                                                     double* d, double* e, double* U, int rows,
                                                     int ldu) {
  const int lane = threadIdx.x;
  for (int i = lane; i < k; i += 64) {
    d[i] = double(Hp[int64_t(i) * k + i]);
    e[i] = (i + 1 < k) ? double(0.5f * (Hp[int64_t(i + 1) * k + i] + Hp[int64_t(i) * k + i + 1]))
                       : 0.0;
  }
  for (int a = lane; a < rows; a += 64)
    for (int j = 0; j < k; ++j) U[int64_t(a) * ldu + j] = (a == j) ? 1.0 : 0.0;
  __syncthreads();

  bool failed = false;
  for (int l = 0; l < k && !failed; ++l) {
    int sweeps = 0;
    while (true) {
      // m = first index >= l with a negligible e[m] (or k - 1): every lane
      // tests one index, a ballot picks the first
      int m = k - 1;
      for (int base = l; base < k - 1; base += 64) {
        const int mm = base + lane;
        bool stop = true;
        if (mm + 1 < k) stop = fabs(e[mm]) <= DBL_EPSILON * (fabs(d[mm]) + fabs(d[mm + 1]));
        const unsigned long long hit = __ballot(stop);
        if (hit) {
          m = base + __builtin_ctzll(hit);
          break;
        }
      }
      if (m == l) break;
      if (++sweeps > kMaxSweeps) {
        failed = true;
        break;
      }
      double g = (d[l + 1] - d[l]) / (2.0 * e[l]);
      double r = hypot(g, 1.0);
      g = d[m] - d[l] + e[l] / (g + copysign(r, g));
      double s = 1.0, c = 1.0, shift = 0.0;
      bool split = false;
      // e[i], d[i], d[i + 1] of the current step in registers; step i stores
      // only index i + 1, so the next step's inputs are read ahead
      double ei = e[m - 1], di = d[m - 1], di1 = d[m];
      for (int i = m - 1; i >= l; --i) {
        const double ei_next = i > l ? e[i - 1] : 0.0;
        const double di_next = i > l ? d[i - 1] : 0.0;
        const double f = s * ei;
        const double b = c * ei;
        // r = |(f, g)| and its reciprocal: one square root and one division
        // while f^2 + g^2 stays far from under- and overflow (the rotation
        // chain is the serial part of QL), hypot otherwise
        const double ss = fma(f, f, g * g);
        double ir;
        if (ss > 1e-280 && ss < 1e280) {
          r = sqrt(ss);
          ir = 1.0 / r;
        } else {
          r = hypot(f, g);
          ir = r == 0.0 ? 0.0 : 1.0 / r;
        }
        e[i + 1] = r;
        if (r == 0.0) {  // an off-diagonal underflowed: the matrix split
          d[i + 1] = di1 - shift;
          e[m] = 0.0;
          split = true;
          break;
        }
        s = f * ir;
        c = g * ir;
        g = di1 - shift;
        r = (di - g) * s + 2.0 * c * b;
        shift = s * r;
        d[i + 1] = g + shift;
        g = c * r - b;
        for (int a = lane; a < rows; a += 64) {
          double* ua = U + int64_t(a) * ldu;
          const double ui = ua[i], ui1 = ua[i + 1];
          ua[i + 1] = s * ui + c * ui1;
          ua[i] = c * ui - s * ui1;
        }
        ei = ei_next;
        di1 = di;
        di = di_next;
      }
      if (split) continue;
      d[l] -= shift;
      e[l] = g;
      e[m] = 0.0;
    }
  }
  __syncthreads();
  return !failed;
}

// One wave per probe.  (a, ab, b2) point at 3 arrays of k doubles and dm at
// 64 columns of k doubles, column l at dm + l with stride 64 (LDS or global
// workspace).
__device__ __forceinline__ void quadrature_wave(const float* __restrict__ Hp, int k, double* a,
                                                double* ab, double* b2, double* dm, float* alpha,
                                                float* beta, double len, double* out) {
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

  double acc = 0.0, wsum = 0.0;
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
    const double w = twisted_weight(a, b2, m, theta, pivmin, dm + lane);
    wsum += w;
    acc += w * log(theta);
  }
  acc = wave_sum_f64(acc);
  wsum = wave_sum_f64(wsum);
  if (!(fabs(wsum - 1.0) <= kWeightSumTol)) {
    // eigenvalues that coincide to working precision: first row of U by QL,
    // with (a, ab, b2) reused as (d, e, u)
    __syncthreads();
    const bool ok = tridiag_eigvecs_wave(Hp, k, a, ab, b2, 1, k);
    acc = 0.0;
    for (int j = lane; j < k; j += 64) acc += b2[j] * b2[j] * log(a[j]);
    acc = ok ? wave_sum_f64(acc) : __builtin_nan("");
  }
  if (lane == 0) *out = nonpositive ? __builtin_nan("") : len * len * acc;
}

// ---------------------------------------------------------------------------
// Four-wave forms for depths up to 64: four lanes per eigenvalue.
// ---------------------------------------------------------------------------
// The tridiagonal of T = (H + H^T)/2 in FP32 (decomp.py:133-135; H
// column-major) into LDS (a, |b|, b^2 and, if bs, the signed b), threads < k;
// alpha / beta outputs if given.
__device__ __forceinline__ void load_tridiag(const float* __restrict__ Hp, int k, double* a,
                                             double* ab, double* b2, double* bs, float* alpha,
                                             float* beta) {
  const int t = threadIdx.x;
  if (t < k) {
    const float av = Hp[int64_t(t) * k + t];
    a[t] = double(av);
    if (alpha) alpha[t] = av;
    if (t + 1 < k) {
      const float b = 0.5f * (Hp[int64_t(t + 1) * k + t] + Hp[int64_t(t) * k + t + 1]);
      ab[t] = fabs(double(b));
      b2[t] = double(b) * double(b);
      if (bs) bs[t] = double(b);
      if (beta) beta[t] = b;
    } else {
      ab[t] = 0.0;
      b2[t] = 0.0;
      if (bs) bs[t] = 0.0;
    }
  }
}

// Leading unreduced block [0, m), Gershgorin bounds widened as LAPACK's
// dstebz does, and the pivot threshold; every wave forms the same values.
struct TriBounds {
  int m;
  double gl, gu, pivmin;
};

__device__ __forceinline__ TriBounds tri_bounds(const double* a, const double* ab,
                                                const double* b2, int k) {
  const int lane = threadIdx.x & 63;
  int split = k;
  double gl = DBL_MAX, gu = -DBL_MAX, bmax = 0.0;
  if (lane < k) {
    const int i = lane;
    if (i + 1 < k && ab[i] <= DBL_EPSILON * (fabs(a[i]) + fabs(a[i + 1]))) split = i + 1;
    const double rad = ab[i] + (i > 0 ? ab[i - 1] : 0.0);
    gl = a[i] - rad;
    gu = a[i] + rad;
    bmax = b2[i];
  }
  split = wave_min_int(split);
  gl = wave_min_f64(gl);
  gu = wave_max_f64(gu);
  bmax = wave_max_f64(bmax);
  const double pivmin = DBL_MIN * fmax(1.0, bmax);
  const double wide = fmax(fabs(gl), fabs(gu));
  gl -= 2.0 * DBL_EPSILON * wide * k + 2.0 * pivmin;
  gu += 2.0 * DBL_EPSILON * wide * k + 2.0 * pivmin;
  return {split, gl, gu, pivmin};
}

// Eigenvalue e of the leading m x m block (ascending) by the four lanes of a
// group (sub = lane & 3): each evaluates the Sturm count at one of the four
// points that cut [lo, hi] into fifths, and the group keeps the fifth that
// holds the eigenvalue (log2 5 bits per round, the four counts in parallel),
// until the interval is as narrow as bisection leaves it or no interior
// point is left.  Every lane of the wave takes part (e >= m: idle).
__device__ __forceinline__ double multisect(const double* a, const double* b2, int m, int e,
                                            int sub, const TriBounds& tb) {
  const int lane = threadIdx.x & 63;
  double lo = tb.gl, hi = tb.gu;
  bool open = e < m;
  for (int it = 0; it < kMaxBisect; ++it) {
    const double w = (hi - lo) * 0.2;
    double x[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) x[q] = fma(double(q + 1), w, lo);
    if (open && (hi - lo <= 2.0 * DBL_EPSILON * fmax(fabs(lo), fabs(hi)) || !(lo < x[0]) ||
                 !(x[0] < x[1]) || !(x[1] < x[2]) || !(x[2] < x[3]) || !(x[3] < hi)))
      open = false;
    if (__ballot(open) == 0) break;
    const bool above = open && sturm_count(a, b2, m, x[sub], tb.pivmin) > e;
    const unsigned g = unsigned(__ballot(above) >> (lane & ~3)) & 0xFu;
    if (open) {
      if (g & 1u) {
        hi = x[0];
      } else if (g & 2u) {
        lo = x[0];
        hi = x[1];
      } else if (g & 4u) {
        lo = x[1];
        hi = x[2];
      } else if (g & 8u) {
        lo = x[2];
        hi = x[3];
      } else {
        lo = x[3];
      }
    }
  }
  return 0.5 * (lo + hi);
}

// Quadrature: the group's first lane forms the twisted weight.  Probes
// whose weights do not sum to one are flagged in retry[p] for
// k_tridiag_quadrature_retry (the one-wave path with its QL fallback);
// retry[p] is written for every probe.
__global__ void __launch_bounds__(256)
    k_tridiag_quadrature_ms(const float* __restrict__ H, const float* __restrict__ scale,
                            int64_t k, double* __restrict__ per_probe,
                            float* __restrict__ alpha, float* __restrict__ beta,
                            int* __restrict__ retry) {
  __shared__ double a[kQuadLdsDepth], ab[kQuadLdsDepth], b2[kQuadLdsDepth];
  __shared__ double dm[64 * kQuadLdsDepth];
  __shared__ double red[2][4];
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  const int64_t p = blockIdx.x;
  const int kk = int(k);
  load_tridiag(H + p * k * k, kk, a, ab, b2, nullptr, alpha ? alpha + p * k : nullptr,
               beta ? beta + p * (k - 1) : nullptr);
  __syncthreads();
  const TriBounds tb = tri_bounds(a, ab, b2, kk);
  const int m = tb.m;
  const int e = wave * 16 + (lane >> 2), sub = lane & 3;
  const double theta = multisect(a, b2, m, e, sub, tb);
  double acc = 0.0, wsum = 0.0;
  if (sub == 0 && e < m) {
    const double wt = twisted_weight(a, b2, m, theta, tb.pivmin, dm + e);
    wsum = wt;
    acc = wt * log(theta);
  }
  acc = wave_sum_f64(acc);
  wsum = wave_sum_f64(wsum);
  if (lane == 0) {
    red[0][wave] = acc;
    red[1][wave] = wsum;
  }
  __syncthreads();
  if (t == 0) {
    acc = (red[0][0] + red[0][1]) + (red[0][2] + red[0][3]);
    wsum = (red[1][0] + red[1][1]) + (red[1][2] + red[1][3]);
    const bool ok = fabs(wsum - 1.0) <= kWeightSumTol;
    retry[p] = ok ? 0 : 1;
    if (ok) {
      // any eigenvalue of T at or below zero makes the reference's log NaN
      const bool nonpositive = sturm_count(a, b2, kk, 0.0, tb.pivmin) > 0;
      const double len = double(scale[p]);
      per_probe[p] = nonpositive ? __builtin_nan("") : len * len * acc;
    }
  }
}

// The one-wave path for the probes k_tridiag_quadrature_ms flagged.
__global__ void __launch_bounds__(64)
    k_tridiag_quadrature_retry(const float* __restrict__ H, const float* __restrict__ scale,
                               int64_t k, const int* __restrict__ retry,
                               double* __restrict__ per_probe) {
  __shared__ double tri[3 * kQuadLdsDepth];
  __shared__ double dm[64 * kQuadLdsDepth];
  const int64_t p = blockIdx.x;
  if (!retry[p]) return;
  quadrature_wave(H + p * k * k, int(k), tri, tri + k, tri + 2 * k, dm, nullptr, nullptr,
                  double(scale[p]), per_probe + p);
}

// Arrays in a global workspace (67 k doubles per probe) beyond kQuadLdsDepth.
__global__ void __launch_bounds__(64)
    k_tridiag_quadrature_global(const float* __restrict__ H, const float* __restrict__ scale,
                                int64_t k, double* __restrict__ work,
                                double* __restrict__ per_probe, float* __restrict__ alpha,
                                float* __restrict__ beta) {
  const int64_t p = blockIdx.x;
  double* w = work + p * 67 * k;
  quadrature_wave(H + p * k * k, int(k), w, w + k, w + 2 * k, w + 3 * k,
                  alpha ? alpha + p * k : nullptr, beta ? beta + p * (k - 1) : nullptr,
                  double(scale[p]), per_probe + p);
}

// Cotangent of |z|^2 [log T]_00 / P w.r.t. H (see the file header), one wave
// per probe; U: k x eig_ld(k), lg: k and Wm: k x k doubles of scratch.
__device__ __forceinline__ void tridiag_vjp_wave(const float* __restrict__ Hp, int k, double* d,
                                                 double* e, double* lg, double* U, double* Wm,
                                                 double scale, float* __restrict__ dHp) {
  const int lane = threadIdx.x;
  const int ldu = eig_ld(k);
  const bool failed = !tridiag_eigvecs_wave(Hp, k, d, e, U, k, ldu);
  // log of the eigenvalues; a non-positive eigenvalue makes the reference's
  // log, and so its cotangent, NaN
  bool nonpositive = failed;
  for (int j = 0; j < k; ++j) nonpositive |= !(d[j] > 0.0);
  for (int j = lane; j < k; j += 64) lg[j] = log(d[j]);
  __syncthreads();

  // W_jl = u_j u_l L_jl (u = U[0, :], L the divided differences of log), all
  // k^2 entries once, spread over the lanes
  for (int idx = lane; idx < k * k; idx += 64) {
    const int j = idx / k, l = idx - j * k;
    const double thj = d[j], thl = d[l], del = thl - thj;
    double L;
    if (j == l || del == 0.0) {
      L = 1.0 / thj;
    } else if (fabs(del) < 1e-3 * fabs(thj)) {
      L = log1p(del / thj) / del;  // (log thl - log thj) / (thl - thj), no cancellation
    } else {
      L = (lg[l] - lg[j]) / del;
    }
    Wm[idx] = U[j] * U[l] * L;
  }
  __syncthreads();

  // Tbar = U W U^T; lane a forms the diagonal entry a and the superdiagonal
  // entry (a, a + 1):
  //   v_l = sum_j W_jl U_aj,  Tbar_aa = sum_l U_al v_l,
  //   Tbar_{a,a+1} = sum_l U_{a+1,l} v_l.
  for (int a = lane; a < k; a += 64) {
    double tdd = 0.0, tdo = 0.0;
    const double* ua = U + int64_t(a) * ldu;
    const double* ub = U + int64_t(a + 1 < k ? a + 1 : a) * ldu;
    for (int l = 0; l < k; ++l) {
      double v = 0.0;
      for (int j = 0; j < k; ++j) v += Wm[j * k + l] * ua[j];
      tdd += ua[l] * v;
      tdo += ub[l] * v;
    }
    const float nan = __builtin_nanf("");
    dHp[int64_t(a) * k + a] = nonpositive ? nan : float(scale * tdd);
    if (a + 1 < k) {
      const float o = nonpositive ? nan : float(scale * tdo);
      dHp[int64_t(a + 1) * k + a] = o;  // (a, a + 1)
      dHp[int64_t(a) * k + a + 1] = o;  // (a + 1, a)
    }
  }
}

// Unit eigenvector of the leading m x m block for the eigenvalue theta from
// the twisted factorisation (see the file header): bottom-up pivots D-_i into
// dm[i * 64], top-down pivots D+_i into z[i * ldz] (then overwritten), the
// twist r = argmin |gamma_i|, and z_r = 1, z_i = -(b_i / D+_i) z_{i+1}
// (i < r), z_{i+1} = -(b_i / D-_{i+1}) z_i (i >= r), normalised.  bs: the
// signed off-diagonal.
__device__ __forceinline__ void twisted_vector(const double* __restrict__ a,
                                               const double* __restrict__ bs,
                                               const double* __restrict__ b2, int m,
                                               double theta, double pivmin, double* dm, double* z,
                                               int ldz) {
  double d = a[m - 1] - theta;
  if (fabs(d) <= pivmin) d = -pivmin;
  dm[(m - 1) * 64] = d;
  for (int i = m - 2; i >= 0; --i) {
    d = (a[i] - theta) - b2[i] / d;
    if (fabs(d) <= pivmin) d = -pivmin;
    dm[i * 64] = d;
  }
  double dp = a[0] - theta;
  if (fabs(dp) <= pivmin) dp = -pivmin;
  z[0] = dp;
  double best = fabs(dp + dm[0] - (a[0] - theta));
  int r = 0;
  for (int i = 1; i < m; ++i) {
    dp = (a[i] - theta) - b2[i - 1] / dp;
    if (fabs(dp) <= pivmin) dp = -pivmin;
    z[i * ldz] = dp;
    const double g = fabs(dp + dm[i * 64] - (a[i] - theta));
    if (g < best) {
      best = g;
      r = i;
    }
  }
  double zi = 1.0, nrm = 1.0;
  for (int i = r - 1; i >= 0; --i) {
    zi = -(bs[i] / z[i * ldz]) * zi;
    z[i * ldz] = zi;
    nrm = fma(zi, zi, nrm);
  }
  z[r * ldz] = 1.0;
  zi = 1.0;
  for (int i = r; i + 1 < m; ++i) {
    zi = -(bs[i] / dm[(i + 1) * 64]) * zi;
    z[(i + 1) * ldz] = zi;
    nrm = fma(zi, zi, nrm);
  }
  const double inv = 1.0 / sqrt(nrm);
  for (int i = 0; i < m; ++i) z[i * ldz] *= inv;
}

// The eigen cotangent for depths up to 64, four waves per probe: the
// eigenvalues of the leading unreduced block by multisection, their
// eigenvectors by the twisted factorisation, accepted when the residuals
// |T u - theta u| / |T| and the entries of U^T U - I are all below
// kEigCheckTol (eigenvalues that coincide to working precision fail it);
// otherwise retry[p] = 1 and k_tridiag_vjp_retry (implicit QL, one wave)
// takes the probe.  Then Tbar = U (L o u u^T) U^T as in tridiag_vjp_wave; U
// has no component outside the leading block (U_0j = 0 there), so neither
// has Tbar.  retry[p] is written for every probe.
constexpr double kEigCheckTol = 1e-10;

// The eigenpairs of the leading m x m block into th[j], U[i * ldu + j]
// (i, j < m) with the checks described above; every thread of the block
// calls it and gets the same answer.  dm: 64 x 64 doubles of scratch.
__device__ __forceinline__ bool eigpairs_ms(const double* a, const double* bs, const double* b2,
                                            const TriBounds& tb, double* th, double* U, int ldu,
                                            double* dm, double* worst_w) {
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  const int m = tb.m;
  const int e = wave * 16 + (lane >> 2), sub = lane & 3;
  const double theta = multisect(a, b2, m, e, sub, tb);
  if (sub == 0 && e < m) {
    th[e] = theta;
    twisted_vector(a, bs, b2, m, theta, tb.pivmin, dm + e, U + e, ldu);
  }
  __syncthreads();
  // residuals (thread t: vector t) and orthogonality (pairs j <= l)
  double worst = 0.0;
  if (t < m) {
    const double tj = th[t];
    const double tnorm = fmax(fmax(fabs(tb.gl), fabs(tb.gu)), DBL_MIN);
    for (int i = 0; i < m; ++i) {
      double r = (a[i] - tj) * U[i * ldu + t];
      if (i > 0) r = fma(bs[i - 1], U[(i - 1) * ldu + t], r);
      if (i + 1 < m) r = fma(bs[i], U[(i + 1) * ldu + t], r);
      worst = fmax(worst, fabs(r) / tnorm);
    }
  }
  for (int idx = t; idx < m * m; idx += blockDim.x) {
    const int j = idx / m, l = idx - j * m;
    if (l < j) continue;
    double d = 0.0;
    for (int i = 0; i < m; ++i) d = fma(U[i * ldu + j], U[i * ldu + l], d);
    worst = fmax(worst, fabs(d - (j == l ? 1.0 : 0.0)));
  }
  // fmax drops a NaN operand: a NaN anywhere must fail the check
  if (worst != worst) worst = DBL_MAX;
  worst = wave_max_f64(worst);
  if (lane == 0) worst_w[wave] = worst;
  __syncthreads();
  const double wmax = fmax(fmax(worst_w[0], worst_w[1]), fmax(worst_w[2], worst_w[3]));
  return wmax <= kEigCheckTol;
}

__global__ void __launch_bounds__(256)
    k_tridiag_vjp_ms(const float* __restrict__ H, const float* __restrict__ len, int64_t P,
                     int64_t k, float* __restrict__ dH, int* __restrict__ retry) {
  constexpr int ldu = eig_ld(kEigLdsDepth);
  __shared__ double a[kEigLdsDepth], ab[kEigLdsDepth], b2[kEigLdsDepth], bs[kEigLdsDepth];
  __shared__ double th[kEigLdsDepth], lg[kEigLdsDepth];
  __shared__ double dm[64 * kEigLdsDepth];  // bottom-up pivots, then W
  __shared__ double U[kEigLdsDepth * ldu];  // U[i * ldu + j]: component i of vector j
  __shared__ double worst_w[4];
  const int t = threadIdx.x;
  const int64_t p = blockIdx.x;
  const int kk = int(k);
  load_tridiag(H + p * k * k, kk, a, ab, b2, bs, nullptr, nullptr);
  __syncthreads();
  const TriBounds tb = tri_bounds(a, ab, b2, kk);
  const int m = tb.m;
  const bool ok = eigpairs_ms(a, bs, b2, tb, th, U, ldu, dm, worst_w);
  if (t == 0) retry[p] = ok ? 0 : 1;
  if (!ok) return;  // the same decision in every thread of the block

  // any eigenvalue of T at or below zero makes the reference's log, and so
  // its cotangent, NaN
  const bool nonpositive = sturm_count(a, b2, kk, 0.0, tb.pivmin) > 0;
  if (t < m) lg[t] = log(th[t]);
  __syncthreads();
  // W_jl = u_j u_l L_jl (u = U[0, :], L the divided differences of log)
  for (int idx = t; idx < m * m; idx += 256) {
    const int j = idx / m, l = idx - j * m;
    const double thj = th[j], thl = th[l], del = thl - thj;
    double L;
    if (j == l || del == 0.0) {
      L = 1.0 / thj;
    } else if (fabs(del) < 1e-3 * fabs(thj)) {
      L = log1p(del / thj) / del;  // (log thl - log thj) / (thl - thj), no cancellation
    } else {
      L = (lg[l] - lg[j]) / del;
    }
    dm[idx] = U[j] * U[l] * L;
  }
  __syncthreads();
  // Tbar_aa = sum_l U_al v_l, Tbar_{a,a+1} = sum_l U_{a+1,l} v_l with
  // v_l = sum_j W_jl U_aj (zero outside the leading block)
  const double l2 = double(len[p]) * double(len[p]);
  const double scale = l2 / double(P);
  float* dHp = dH + p * k * k;
  if (t < kk) {
    double tdd = 0.0, tdo = 0.0;
    if (t < m) {
      const double* ua = U + t * ldu;
      const bool up = t + 1 < m;
      const double* ub = U + (up ? t + 1 : t) * ldu;
      for (int l = 0; l < m; ++l) {
        double v = 0.0;
        for (int j = 0; j < m; ++j) v = fma(dm[j * m + l], ua[j], v);
        tdd = fma(ua[l], v, tdd);
        if (up) tdo = fma(ub[l], v, tdo);
      }
    }
    const float nan = __builtin_nanf("");
    dHp[int64_t(t) * k + t] = nonpositive ? nan : float(scale * tdd);
    if (t + 1 < kk) {
      const float o = nonpositive ? nan : float(scale * tdo);
      dHp[int64_t(t + 1) * k + t] = o;  // (t, t + 1)
      dHp[int64_t(t) * k + t + 1] = o;  // (t + 1, t)
    }
  }
}

// The one-wave QL path for the probes k_tridiag_vjp_ms flagged.
__global__ void __launch_bounds__(64)
    k_tridiag_vjp_retry(const float* __restrict__ H, const float* __restrict__ len, int64_t P,
                        int64_t k, const int* __restrict__ retry, float* __restrict__ dH) {
  __shared__ double U[kEigLdsDepth * eig_ld(kEigLdsDepth)];
  __shared__ double Wm[kEigLdsDepth * kEigLdsDepth];
  __shared__ double d[kEigLdsDepth], e[kEigLdsDepth], lg[kEigLdsDepth];
  const int64_t p = blockIdx.x;
  if (!retry[p]) return;
  const double l2 = double(len[p]) * double(len[p]);
  tridiag_vjp_wave(H + p * k * k, int(k), d, e, lg, U, Wm, l2 / double(P), dH + p * k * k);
}

__global__ void __launch_bounds__(64)
    k_tridiag_vjp_global(const float* __restrict__ H, const float* __restrict__ len, int64_t P,
                         int64_t k, double* __restrict__ work, float* __restrict__ dH) {
  const int64_t p = blockIdx.x;
  const int64_t ku = k * eig_ld(int(k));  // U
  double* w = work + p * (ku + k * k + 3 * k);
  const double l2 = double(len[p]) * double(len[p]);
  tridiag_vjp_wave(H + p * k * k, int(k), w + ku, w + ku + k, w + ku + 2 * k, w, w + ku + 3 * k,
                   l2 / double(P), dH + p * k * k);
}


// f(theta) for the matrix functions of slq_funm_vectors (slq_matfun).
__device__ __forceinline__ double apply_matfun(int f, double x) {
  switch (f) {
    case SLQ_FUNM_LOG: return log(x);
    case SLQ_FUNM_EXP: return exp(x);
    case SLQ_FUNM_SIN: return sin(x);
    case SLQ_FUNM_COS: return cos(x);
    case SLQ_FUNM_SQRT: return sqrt(x);
    default: return 1.0 / x;  // SLQ_FUNM_INV
  }
}

// funm_lanczos_sym's small part (funm.py:145-147): y = f(T) e_1 =
// U f(theta) U^T e_1, written as the combination coefficients of
// |v| Q^T y for k_adjoint_combine (per probe 2 k + 4 floats: y, zeros, and
// the scale |v| = len).  NaN coefficients if QL fails.
__device__ __forceinline__ void tridiag_funm_wave(const float* __restrict__ Hp, int k, double* d,
                                                  double* e, double* fd, double* U, int f,
                                                  float scale, float* __restrict__ cp) {
  const int lane = threadIdx.x;
  const int ldu = eig_ld(k);
  const bool ok = tridiag_eigvecs_wave(Hp, k, d, e, U, k, ldu);
  for (int j = lane; j < k; j += 64) fd[j] = apply_matfun(f, d[j]);  // f(theta_j) once
  __syncthreads();
  for (int a = lane; a < k; a += 64) {
    double y = 0.0;
    const double* ua = U + int64_t(a) * ldu;
    for (int j = 0; j < k; ++j) y += ua[j] * fd[j] * U[j];
    cp[a] = ok ? float(y) : __builtin_nanf("");
    cp[k + a] = 0.f;
  }
  if (lane == 0) {
    cp[2 * k + 0] = 0.f;
    cp[2 * k + 1] = 0.f;
    cp[2 * k + 2] = 0.f;
    cp[2 * k + 3] = scale;
  }
}

__global__ void __launch_bounds__(64)
    k_tridiag_funm_global(const float* __restrict__ H, const float* __restrict__ len, int64_t k,
                          int f, double* __restrict__ work, float* __restrict__ coef) {
  const int64_t p = blockIdx.x;
  const int64_t ku = k * eig_ld(int(k));  // U
  double* w = work + p * (ku + 3 * k);
  tridiag_funm_wave(H + p * k * k, int(k), w + ku, w + ku + k, w + ku + 2 * k, w, f, len[p],
                    coef + p * (2 * k + 4));
}

// funm's small part for depths up to 64 (four waves per probe): the
// eigenpairs of eigpairs_ms when T has no negligible off-diagonal (the whole
// of T is the leading block) and they pass its checks; otherwise retry[p] = 1
// and k_tridiag_funm_retry (implicit QL, one wave) takes the probe.
__global__ void __launch_bounds__(256)
    k_tridiag_funm_ms(const float* __restrict__ H, const float* __restrict__ len, int64_t k,
                      int f, float* __restrict__ coef, int* __restrict__ retry) {
  constexpr int ldu = eig_ld(kEigLdsDepth);
  __shared__ double a[kEigLdsDepth], ab[kEigLdsDepth], b2[kEigLdsDepth], bs[kEigLdsDepth];
  __shared__ double th[kEigLdsDepth], fd[kEigLdsDepth];
  __shared__ double dm[64 * kEigLdsDepth];
  __shared__ double U[kEigLdsDepth * ldu];
  __shared__ double worst_w[4];
  const int t = threadIdx.x;
  const int64_t p = blockIdx.x;
  const int kk = int(k);
  load_tridiag(H + p * k * k, kk, a, ab, b2, bs, nullptr, nullptr);
  __syncthreads();
  const TriBounds tb = tri_bounds(a, ab, b2, kk);
  const bool ok = tb.m == kk && eigpairs_ms(a, bs, b2, tb, th, U, ldu, dm, worst_w);
  if (t == 0) retry[p] = ok ? 0 : 1;
  if (!ok) return;  // the same decision in every thread of the block
  if (t < kk) fd[t] = apply_matfun(f, th[t]);  // f(theta_j) once
  __syncthreads();
  float* cp = coef + p * (2 * k + 4);
  if (t < kk) {
    double y = 0.0;
    const double* ua = U + t * ldu;
    for (int j = 0; j < kk; ++j) y += ua[j] * fd[j] * U[j];
    cp[t] = float(y);
    cp[k + t] = 0.f;
  }
  if (t == 0) {
    cp[2 * k + 0] = 0.f;
    cp[2 * k + 1] = 0.f;
    cp[2 * k + 2] = 0.f;
    cp[2 * k + 3] = len[p];
  }
}

__global__ void __launch_bounds__(64)
    k_tridiag_funm_retry(const float* __restrict__ H, const float* __restrict__ len, int64_t k,
                         int f, const int* __restrict__ retry, float* __restrict__ coef) {
  __shared__ double U[kEigLdsDepth * eig_ld(kEigLdsDepth)];
  __shared__ double d[kEigLdsDepth], e[kEigLdsDepth], fd[kEigLdsDepth];
  const int64_t p = blockIdx.x;
  if (!retry[p]) return;
  tridiag_funm_wave(H + p * k * k, int(k), d, e, fd, U, f, len[p], coef + p * (2 * k + 4));
}

// mean and standard error of the mean (population std / sqrt(P)) in FP64,
// one block, fixed order (stochtrace.py:84-88).
__global__ void __launch_bounds__(256)
    k_mean_and_sem(const double* __restrict__ v, int64_t P, double* __restrict__ mean,
                   double* __restrict__ sem) {
  __shared__ double red[256];
  __shared__ double mu;
  double s = 0.0;
  for (int64_t i = threadIdx.x; i < P; i += 256) s += v[i];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) mu = red[0] / double(P);
  __syncthreads();
  double q = 0.0;
  for (int64_t i = threadIdx.x; i < P; i += 256) q += (v[i] - mu) * (v[i] - mu);
  red[threadIdx.x] = q;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    if (mean) *mean = mu;
    if (sem) *sem = sqrt(red[0] / double(P)) / sqrt(double(P));
  }
}

// Plain Lanczos coefficients from T (column-major P x k x k) and the
// remainder norms: alpha[p][i] = T_ii, beta[p][i] = T_{i+1,i} (i < k - 1) and
// beta[p][kk - 1] = blast[p], kk = max(k, 1).
__global__ void k_lanczos_coeffs(const float* __restrict__ T, const float* __restrict__ blast,
                                 int64_t P, int64_t k, float* __restrict__ alpha,
                                 float* __restrict__ beta) {
  const int64_t p = blockIdx.x;
  const int64_t kk = k > 0 ? k : 1;
  for (int64_t i = threadIdx.x; i < k; i += blockDim.x) {
    alpha[p * k + i] = T[p * k * k + i * k + i];
    if (i + 1 < k) beta[p * kk + i] = T[p * k * k + i * k + i + 1];
  }
  if (threadIdx.x == 0) beta[p * kk + kk - 1] = blast[p];
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
    if (k <= kQuadLdsDepth) {
      int* retry = reinterpret_cast<int*>(work);
      hipLaunchKernelGGL(k_tridiag_quadrature_ms, dim3(unsigned(P)), dim3(256), 0, s, H, scale, k,
                         per_probe, alpha, beta, retry);
      hipLaunchKernelGGL(k_tridiag_quadrature_retry, dim3(unsigned(P)), dim3(64), 0, s, H, scale,
                         k, retry, per_probe);
    } else {
      hipLaunchKernelGGL(k_tridiag_quadrature_global, dim3(unsigned(P)), dim3(64), 0, s, H, scale,
                         k, work, per_probe, alpha, beta);
    }
    hipError_t err = hipGetLastError();
    if (err != hipSuccess) return err;
  }
  if (mean) hipLaunchKernelGGL(k_probe_mean, dim3(1), dim3(64), 0, s, per_probe, P, mean);
  return hipGetLastError();
}

hipError_t launch_tridiag_vjp(hipStream_t s, const float* H, const float* len, int64_t P, int64_t k,
                              double* work, float* dH) {
  if (P <= 0 || k <= 0) return hipSuccess;
  if (k <= kEigLdsDepth) {
    int* retry = reinterpret_cast<int*>(work);
    hipLaunchKernelGGL(k_tridiag_vjp_ms, dim3(unsigned(P)), dim3(256), 0, s, H, len, P, k, dH,
                       retry);
    hipLaunchKernelGGL(k_tridiag_vjp_retry, dim3(unsigned(P)), dim3(64), 0, s, H, len, P, k,
                       retry, dH);
  } else {
    hipLaunchKernelGGL(k_tridiag_vjp_global, dim3(unsigned(P)), dim3(64), 0, s, H, len, P, k, work,
                       dH);
  }
  return hipGetLastError();
}

int64_t tridiag_quadrature_work_doubles(int64_t P, int64_t k) {
  // k <= kQuadLdsDepth: the retry flags (one int per probe, in doubles)
  return k <= kQuadLdsDepth ? P : P * 67 * k;
}

int64_t tridiag_vjp_work_doubles(int64_t P, int64_t k) {
  // k <= kEigLdsDepth: the retry flags (one int per probe, in doubles)
  return k <= kEigLdsDepth ? P : P * (k * eig_ld(int(k)) + k * k + 3 * k);
}

hipError_t launch_tridiag_funm(hipStream_t s, const float* H, const float* len, int64_t P,
                               int64_t k, int f, double* work, float* coef) {
  if (P <= 0 || k <= 0) return hipSuccess;
  if (k <= kEigLdsDepth) {
    int* retry = reinterpret_cast<int*>(work);
    hipLaunchKernelGGL(k_tridiag_funm_ms, dim3(unsigned(P)), dim3(256), 0, s, H, len, k, f, coef,
                       retry);
    hipLaunchKernelGGL(k_tridiag_funm_retry, dim3(unsigned(P)), dim3(64), 0, s, H, len, k, f,
                       retry, coef);
  } else {
    hipLaunchKernelGGL(k_tridiag_funm_global, dim3(unsigned(P)), dim3(64), 0, s, H, len, k, f,
                       work, coef);
  }
  return hipGetLastError();
}

hipError_t launch_lanczos_coeffs(hipStream_t s, const float* T, const float* blast, int64_t P,
                                 int64_t k, float* alpha, float* beta) {
  hipLaunchKernelGGL(k_lanczos_coeffs, dim3(unsigned(P)), dim3(64), 0, s, T, blast, P, k, alpha,
                     beta);
  return hipGetLastError();
}

hipError_t launch_mean_and_sem(hipStream_t s, const double* v, int64_t P, double* mean,
                               double* sem) {
  hipLaunchKernelGGL(k_mean_and_sem, dim3(1), dim3(256), 0, s, v, P, mean, sem);
  return hipGetLastError();
}

}  // namespace slq
