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
// Algorithm (one wave per probe, one eigenvalue per lane):
//  * theta_j by bisection on the Sturm count of T - x I (the number of
//    negative pivots of its LDL^T factorisation), to full FP64 resolution;
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

__global__ void __launch_bounds__(64)
    k_tridiag_quadrature_lds(const float* __restrict__ H, const float* __restrict__ scale,
                             int64_t k, double* __restrict__ per_probe,
                             float* __restrict__ alpha, float* __restrict__ beta) {
  __shared__ double tri[3 * kQuadLdsDepth];
  __shared__ double dm[64 * kQuadLdsDepth];
  const int64_t p = blockIdx.x;
  quadrature_wave(H + p * k * k, int(k), tri, tri + k, tri + 2 * k, dm,
                  alpha ? alpha + p * k : nullptr, beta ? beta + p * (k - 1) : nullptr,
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

__global__ void __launch_bounds__(64)
    k_tridiag_vjp_lds(const float* __restrict__ H, const float* __restrict__ len, int64_t P,
                      int64_t k, float* __restrict__ dH) {
  __shared__ double U[kEigLdsDepth * eig_ld(kEigLdsDepth)];
  __shared__ double Wm[kEigLdsDepth * kEigLdsDepth];
  __shared__ double d[kEigLdsDepth], e[kEigLdsDepth], lg[kEigLdsDepth];
  const int64_t p = blockIdx.x;
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
    k_tridiag_funm_lds(const float* __restrict__ H, const float* __restrict__ len, int64_t k, int f,
                       float* __restrict__ coef) {
  __shared__ double U[kEigLdsDepth * eig_ld(kEigLdsDepth)];
  __shared__ double d[kEigLdsDepth], e[kEigLdsDepth], fd[kEigLdsDepth];
  const int64_t p = blockIdx.x;
  tridiag_funm_wave(H + p * k * k, int(k), d, e, fd, U, f, len[p], coef + p * (2 * k + 4));
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
    if (k <= kQuadLdsDepth)
      hipLaunchKernelGGL(k_tridiag_quadrature_lds, dim3(unsigned(P)), dim3(64), 0, s, H, scale, k,
                         per_probe, alpha, beta);
    else
      hipLaunchKernelGGL(k_tridiag_quadrature_global, dim3(unsigned(P)), dim3(64), 0, s, H, scale,
                         k, work, per_probe, alpha, beta);
    hipError_t err = hipGetLastError();
    if (err != hipSuccess) return err;
  }
  if (mean) hipLaunchKernelGGL(k_probe_mean, dim3(1), dim3(64), 0, s, per_probe, P, mean);
  return hipGetLastError();
}

hipError_t launch_tridiag_vjp(hipStream_t s, const float* H, const float* len, int64_t P, int64_t k,
                              double* work, float* dH) {
  if (P <= 0 || k <= 0) return hipSuccess;
  if (k <= kEigLdsDepth)
    hipLaunchKernelGGL(k_tridiag_vjp_lds, dim3(unsigned(P)), dim3(64), 0, s, H, len, P, k, dH);
  else
    hipLaunchKernelGGL(k_tridiag_vjp_global, dim3(unsigned(P)), dim3(64), 0, s, H, len, P, k, work,
                       dH);
  return hipGetLastError();
}

int64_t tridiag_quadrature_work_doubles(int64_t P, int64_t k) {
  return k <= kQuadLdsDepth ? 0 : P * 67 * k;
}

int64_t tridiag_vjp_work_doubles(int64_t P, int64_t k) {
  return k <= kEigLdsDepth ? 0 : P * (k * eig_ld(int(k)) + k * k + 3 * k);
}

hipError_t launch_tridiag_funm(hipStream_t s, const float* H, const float* len, int64_t P,
                               int64_t k, int f, double* work, float* coef) {
  if (P <= 0 || k <= 0) return hipSuccess;
  if (k <= kEigLdsDepth)
    hipLaunchKernelGGL(k_tridiag_funm_lds, dim3(unsigned(P)), dim3(64), 0, s, H, len, k, f, coef);
  else
    hipLaunchKernelGGL(k_tridiag_funm_global, dim3(unsigned(P)), dim3(64), 0, s, H, len, k, f,
                       work, coef);
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
