// Per-probe tridiagonal eigen-solve and Gauss quadrature in FP64.
//
// Replaces, for every probe, funm.dense_funm_sym_eigh(log) applied to the
// materialised tridiagonal matrix (funm.py:330-333, decomp.py:133-152) and the
// read-out |z|^2 e1^T f(T) e1 (funm.py:240-241), followed by the probe mean of
// stochtrace.py:50.  Only the first row of the eigenvector matrix enters
// e1^T U log(L) U^T e1 = sum_j U_0j^2 log(lambda_j) (Golub-Welsch), so the
// solver tracks that row alone.
//
// Algorithm: implicit QL with Wilkinson-type shifts on the symmetric
// tridiagonal (d, e), applying every plane rotation to the tracked row.
// The bulge chase is sequential, so one lane owns one probe; the (d, e, z)
// arrays live in LDS (a global workspace for depths too large for it), laid
// out [index][lane] so that the lanes of a wave touch consecutive addresses.
//
// The tridiagonal is extracted from the Arnoldi Hessenberg matrix exactly as
// decomp.py:133-135 does it, in FP32: alpha_i = H_ii and
// beta_i = 0.5 * (H_{i,i+1} + H_{i+1,i}); only then is it promoted to FP64.
// A non-positive Ritz value yields NaN (or -inf for zero) like the
// reference's unguarded log (funm.py:201, 332).
#include <cfloat>

#include "slq_internal.h"

namespace slq {
namespace {

constexpr int kMaxSweeps = 64;
constexpr int kLdsBytes = 64 * 1024;

// The (d, e, z) arrays of lane `lane` among `lanes` probes: element i at
// base[i * lanes + lane], so the lanes of a wave touch consecutive addresses.
struct Arrays {
  double* d;
  double* e;
  double* z;
  int64_t stride;
};

__device__ void ql_quadrature(const float* __restrict__ Hp, int64_t k, const Arrays& a,
                              float* alpha, float* beta, double len, double* out) {
#define D(i) a.d[int64_t(i) * a.stride]
#define E(i) a.e[int64_t(i) * a.stride]
#define Z(i) a.z[int64_t(i) * a.stride]
  // tridiagonal of T = (H + H^T)/2 in FP32 (decomp.py:133-135); H column-major
  for (int64_t i = 0; i < k; ++i) {
    const float av = Hp[i * k + i];
    D(i) = double(av);
    Z(i) = (i == 0) ? 1.0 : 0.0;
    if (alpha) alpha[i] = av;
    if (i + 1 < k) {
      const float b = 0.5f * (Hp[(i + 1) * k + i] + Hp[i * k + i + 1]);
      E(i) = double(b);
      if (beta) beta[i] = b;
    } else {
      E(i) = 0.0;
    }
  }

  bool failed = false;
  for (int64_t l = 0; l < k && !failed; ++l) {
    int sweeps = 0;
    while (true) {
      int64_t m = l;
      for (; m + 1 < k; ++m) {
        const double dd = fabs(D(m)) + fabs(D(m + 1));
        if (fabs(E(m)) <= DBL_EPSILON * dd) break;
      }
      if (m == l) break;
      if (++sweeps > kMaxSweeps) {
        failed = true;
        break;
      }
      // shift from the leading 2x2 block of the unreduced part
      double g = (D(l + 1) - D(l)) / (2.0 * E(l));
      double r = hypot(g, 1.0);
      g = D(m) - D(l) + E(l) / (g + copysign(r, g));
      double s = 1.0, c = 1.0, shift = 0.0;
      bool split = false;
      for (int64_t i = m - 1; i >= l; --i) {
        const double f = s * E(i);
        const double b = c * E(i);
        r = hypot(f, g);
        E(i + 1) = r;
        if (r == 0.0) {  // an off-diagonal underflowed: the matrix split
          D(i + 1) -= shift;
          E(m) = 0.0;
          split = true;
          break;
        }
        s = f / r;
        c = g / r;
        g = D(i + 1) - shift;
        r = (D(i) - g) * s + 2.0 * c * b;
        shift = s * r;
        D(i + 1) = g + shift;
        g = c * r - b;
        const double zi = Z(i), zi1 = Z(i + 1);
        Z(i + 1) = s * zi + c * zi1;
        Z(i) = c * zi - s * zi1;
      }
      if (split) continue;
      D(l) -= shift;
      E(l) = g;
      E(m) = 0.0;
    }
  }

  double acc = 0.0;
  for (int64_t j = 0; j < k; ++j) acc += Z(j) * Z(j) * log(D(j));
  *out = failed ? __builtin_nan("") : len * len * acc;
#undef D
#undef E
#undef Z
}

// Arrays in LDS: `lanes` probes per block, 3 k doubles each.
__global__ void __launch_bounds__(64)
    k_tridiag_quadrature_lds(const float* __restrict__ H, const float* __restrict__ scale,
                             int64_t P, int64_t k, double* __restrict__ per_probe,
                             float* __restrict__ alpha, float* __restrict__ beta) {
  extern __shared__ double lds[];
  const int lanes = blockDim.x;
  const int lane = threadIdx.x;
  const int64_t p = int64_t(blockIdx.x) * lanes + lane;
  if (p >= P) return;
  Arrays a{lds + lane, lds + k * lanes + lane, lds + 2 * k * lanes + lane, lanes};
  ql_quadrature(H + p * k * k, k, a, alpha ? alpha + p * k : nullptr,
                beta ? beta + p * (k - 1) : nullptr, double(scale[p]), per_probe + p);
}

// Arrays in a global workspace (3 k P doubles) for depths too large for LDS.
__global__ void __launch_bounds__(64)
    k_tridiag_quadrature_global(const float* __restrict__ H, const float* __restrict__ scale,
                                int64_t P, int64_t k, double* __restrict__ work,
                                double* __restrict__ per_probe, float* __restrict__ alpha,
                                float* __restrict__ beta) {
  const int64_t p = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (p >= P) return;
  Arrays a{work + p, work + k * P + p, work + 2 * k * P + p, P};
  ql_quadrature(H + p * k * k, k, a, alpha ? alpha + p * k : nullptr,
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
    // probes per block: as many as fit 3 k doubles each into 64 KB of LDS
    int lanes = 64;
    while (lanes > 1 && int64_t(lanes) * 3 * k * 8 > kLdsBytes) lanes /= 2;
    if (int64_t(lanes) * 3 * k * 8 <= kLdsBytes) {
      const unsigned blocks = unsigned((P + lanes - 1) / lanes);
      const size_t shmem = size_t(lanes) * 3 * size_t(k) * sizeof(double);
      hipLaunchKernelGGL(k_tridiag_quadrature_lds, dim3(blocks), dim3(lanes), shmem, s, H, scale,
                         P, k, per_probe, alpha, beta);
    } else {
      const unsigned blocks = unsigned((P + 63) / 64);
      hipLaunchKernelGGL(k_tridiag_quadrature_global, dim3(blocks), dim3(64), 0, s, H, scale, P,
                         k, work, per_probe, alpha, beta);
    }
    hipError_t err = hipGetLastError();
    if (err != hipSuccess) return err;
  }
  if (mean) hipLaunchKernelGGL(k_probe_mean, dim3(1), dim3(64), 0, s, per_probe, P, mean);
  return hipGetLastError();
}

}  // namespace slq
