// The GP log-density caller of the SLQ path: partial Cholesky factorisations
// (low_rank.py:56-198), the Woodbury solve of the low-rank preconditioner
// (low_rank.py:7-53) and conjugate gradients over a batch of right-hand sides
// (backend/linalg.py:107-108, i.e. jax.scipy.sparse.linalg.cg, used by
// tutorials/8_gaussian_logpdf.py:38-52).
//
// Partial Cholesky, one launch sequence per rank step i (the step depends on
// the pivot of the previous one):
//  * k_chol_residual: per row r, |diag(A_perm)_r - sum_{j<i} L_rj^2| and a
//    block argmax (NaN ranks highest, ties go to the smaller row, as argmax);
//  * k_chol_pivot (one block): the global argmax, the swap of rows i and k of
//    L and of the permutation, row i of L staged for the next kernel, and
//    l_ii = sqrt(A_ii - <L_i, L_i>) with the success flag (l_ii^2 > 0);
//  * k_chol_column: column i, (A[perm[r], perm[i]] - <L_r, L_i>) / l_ii for
//    every row r (the reference writes the whole column, pivoted rows too).
// L is kept column-major (column j contiguous over the n rows) so that every
// per-row sum reads coalesced; k_chol_output writes the n x rank row-major
// factor with its rows back in the original order (L[argsort(P)]).
// Matrix elements come from the operator: a stored matrix A[r * n + c], or
// the RBF kernel exp(-|x_r - x_c|^2 / (2 l^2)) + noise [r == c] evaluated in
// FP32 from the points.
//
// Woodbury solve of (s I + L L^T) V = B: the capacitance C = I + L^T L / s is
// formed in FP64 from FP32 partial Gram tiles and factored by a left-looking
// Cholesky in one block (setup); each application is L^T b (chunk partials),
// a forward/backward substitution per right-hand side (one wave each, FP64),
// and b / s - L z / sqrt(s).
//
// CG: the per-column state machine of JAX's preconditioned CG under vmap
// (x0 = 0; rs = gamma without a preconditioner, r . r with one; a converged
// column keeps its state while the others iterate).  Per-column scalars live
// in small device arrays; every block of an iteration's kernels sums the
// partial dots it needs itself, and the scalars that change (gamma, active)
// are double-buffered by iteration parity, so an iteration is four launches:
// the matvec, d . A d (adding up the matvec's split-K partials), the x / r
// update, and beta with the new direction.  The host reads the number of
// active columns after batches of 1, 2, 4, 8, 8, ... iterations (an iteration
// with no active column changes nothing).
#include <cfloat>

#include "slq_internal.h"

namespace slq {
namespace {

constexpr int kT = 256;
constexpr int kCgPer = 8;  // elements per thread of the CG dot kernel
constexpr int kCgChunk = kT * kCgPer;
// The CG kernels of an iteration run kTq threads per chunk of kCgChunk
// elements (kQPer each): a column of a few thousand elements has only a few
// chunks, so more threads per chunk shorten these latency-bound launches.
constexpr int kTq = 1024;
constexpr int kQPer = kCgChunk / kTq;

__device__ __forceinline__ float wsum_f(float v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

__device__ __forceinline__ double wsum_d(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// Block sum (NT threads, red[NT / 64]); the result is valid in thread 0.
template <int NT = kT>
__device__ __forceinline__ float block_sum_f(float v, float* red) {
  v = wsum_f(v);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0) red[wave] = v;
  __syncthreads();
  float s = 0.f;
  if (threadIdx.x == 0)
    for (int w = 0; w < NT / 64; ++w) s += red[w];
  __syncthreads();
  return s;
}

struct Elem {
  const float* A;
  const float* X;
  int dim;
  int64_t n;
  float inv2l2;
  float noise;
  __device__ __forceinline__ float operator()(int64_t r, int64_t c) const {
    if (A) return A[r * n + c];
    float d2 = 0.f;
    for (int d = 0; d < dim; ++d) {
      const float t = X[r * dim + d] - X[c * dim + d];
      d2 = fmaf(t, t, d2);
    }
    float v = expf(-d2 * inv2l2);
    if (r == c) v += noise;
    return v;
  }
};

// a ranks above b: NaN first, then larger, then the smaller index
__device__ __forceinline__ bool ranks_above(float a, int64_t ia, float b, int64_t ib) {
  const bool an = a != a, bn = b != b;
  if (an || bn) return an && (!bn || ia < ib);
  return a > b || (a == b && ia < ib);
}

__device__ __forceinline__ void block_argmax(float& v, int64_t& idx, float* sv, int64_t* si) {
  const int t = threadIdx.x;
  sv[t] = v;
  si[t] = idx;
  __syncthreads();
  for (int w = kT / 2; w > 0; w >>= 1) {
    if (t < w && ranks_above(sv[t + w], si[t + w], sv[t], si[t])) {
      sv[t] = sv[t + w];
      si[t] = si[t + w];
    }
    __syncthreads();
  }
  v = sv[0];
  idx = si[0];
  __syncthreads();
}

__global__ void __launch_bounds__(kT)
    k_chol_residual(const float* __restrict__ Lc, int64_t n, int i,
                    const int64_t* __restrict__ perm, Elem E, float* __restrict__ bval,
                    int64_t* __restrict__ bidx) {
  __shared__ float sv[kT];
  __shared__ int64_t si[kT];
  const int64_t r = int64_t(blockIdx.x) * kT + threadIdx.x;
  float a = -1.f;
  int64_t ia = n;
  if (r < n) {
    const int64_t pr = perm[r];
    float acc = 0.f;
    for (int j = 0; j < i; ++j) {
      const float l = Lc[int64_t(j) * n + r];
      acc = fmaf(l, l, acc);
    }
    a = fabsf(E(pr, pr) - acc);
    ia = r;
  }
  block_argmax(a, ia, sv, si);
  if (threadIdx.x == 0) {
    bval[blockIdx.x] = a;
    bidx[blockIdx.x] = ia;
  }
}

__global__ void __launch_bounds__(kT)
    k_chol_pivot(float* __restrict__ Lc, int64_t n, int i, int64_t* __restrict__ perm, Elem E,
                 const float* __restrict__ bval, const int64_t* __restrict__ bidx, int nblk,
                 int pivot, float* __restrict__ rowi, float* __restrict__ lii,
                 int* __restrict__ success) {
  __shared__ float sv[kT];
  __shared__ int64_t si[kT];
  __shared__ float red[kT / 64];
  const int t = threadIdx.x;
  if (pivot) {
    float a = -1.f;
    int64_t ia = n;
    for (int b = t; b < nblk; b += kT)
      if (ranks_above(bval[b], bidx[b], a, ia)) {
        a = bval[b];
        ia = bidx[b];
      }
    block_argmax(a, ia, sv, si);
    const int64_t kp = ia;
    if (kp != i && kp < n) {
      for (int j = t; j < i; j += kT) {
        const float x = Lc[int64_t(j) * n + i];
        Lc[int64_t(j) * n + i] = Lc[int64_t(j) * n + kp];
        Lc[int64_t(j) * n + kp] = x;
      }
      if (t == 0) {
        const int64_t x = perm[i];
        perm[i] = perm[kp];
        perm[kp] = x;
      }
    }
    __syncthreads();
  }
  float acc = 0.f;
  for (int j = t; j < i; j += kT) {
    const float l = Lc[int64_t(j) * n + i];
    rowi[j] = l;
    acc = fmaf(l, l, acc);
  }
  const float dot = block_sum_f(acc, red);
  if (t == 0) {
    const int64_t pi = perm[i];
    const float l2 = E(pi, pi) - dot;
    *lii = sqrtf(l2);
    if (!(l2 > 0.f)) *success = 0;
  }
}

__global__ void __launch_bounds__(kT)
    k_chol_column(float* __restrict__ Lc, int64_t n, int i, const int64_t* __restrict__ perm,
                  Elem E, const float* __restrict__ rowi, const float* __restrict__ lii) {
  extern __shared__ float row[];  // i floats
  for (int j = threadIdx.x; j < i; j += kT) row[j] = rowi[j];
  __syncthreads();
  const int64_t r = int64_t(blockIdx.x) * kT + threadIdx.x;
  if (r >= n) return;
  float acc = 0.f;
  for (int j = 0; j < i; ++j) acc = fmaf(Lc[int64_t(j) * n + r], row[j], acc);
  Lc[int64_t(i) * n + r] = (E(perm[r], perm[i]) - acc) / *lii;
}

__global__ void __launch_bounds__(kT)
    k_chol_init(int64_t n, int64_t* __restrict__ perm, int* __restrict__ success) {
  const int64_t r = int64_t(blockIdx.x) * kT + threadIdx.x;
  if (r < n) perm[r] = r;
  if (r == 0) *success = 1;
}

__global__ void __launch_bounds__(kT)
    k_chol_output(const float* __restrict__ Lc, int64_t n, int rank,
                  const int64_t* __restrict__ perm, float* __restrict__ L,
                  int64_t* __restrict__ perm_out) {
  const int64_t e = int64_t(blockIdx.x) * kT + threadIdx.x;
  if (e >= n * rank) return;
  const int64_t j = e / n, r = e - j * n;
  L[perm[r] * rank + j] = Lc[e];
  if (perm_out && j == 0) perm_out[r] = perm[r];
}

// ---- Woodbury solve ---------------------------------------------------------
constexpr int kGT = 16;   // Gram tile edge (kGT x kGT threads)
constexpr int kGRows = 64;

__global__ void __launch_bounds__(kGT * kGT)
    k_gram_partial(const float* __restrict__ L, int64_t n, int r, int64_t rows_per_split,
                   float* __restrict__ Gp) {
  __shared__ float La[kGRows][kGT + 1], Lb[kGRows][kGT + 1];
  const int tx = threadIdx.x % kGT, ty = threadIdx.x / kGT;
  const int a0 = blockIdx.x * kGT, b0 = blockIdx.y * kGT;
  const int64_t m0 = int64_t(blockIdx.z) * rows_per_split;
  const int64_t m1 = m0 + rows_per_split < n ? m0 + rows_per_split : n;
  float acc = 0.f;
  for (int64_t mb = m0; mb < m1; mb += kGRows) {
    for (int e = threadIdx.x; e < kGRows * kGT; e += kGT * kGT) {
      const int mm = e / kGT, c = e % kGT;
      const int64_t m = mb + mm;
      La[mm][c] = (m < m1 && a0 + c < r) ? L[m * r + a0 + c] : 0.f;
      Lb[mm][c] = (m < m1 && b0 + c < r) ? L[m * r + b0 + c] : 0.f;
    }
    __syncthreads();
#pragma unroll 8
    for (int mm = 0; mm < kGRows; ++mm) acc = fmaf(La[mm][ty], Lb[mm][tx], acc);
    __syncthreads();
  }
  if (a0 + ty < r && b0 + tx < r)
    Gp[(int64_t(blockIdx.z) * r + a0 + ty) * r + b0 + tx] = acc;
}

__global__ void __launch_bounds__(kT)
    k_gram_reduce(const float* __restrict__ Gp, int nsplit, int r, double inv_s,
                  double* __restrict__ C) {
  const int64_t e = int64_t(blockIdx.x) * kT + threadIdx.x;
  const int64_t rr = int64_t(r) * r;
  if (e >= rr) return;
  double s = 0.0;
  for (int q = 0; q < nsplit; ++q) s += double(Gp[q * rr + e]);
  const int64_t a = e / r, b = e - a * r;
  C[e] = (a == b ? 1.0 : 0.0) + s * inv_s;
}

// In-place lower Cholesky of C (r x r, row-major, FP64), one block; ok = 0
// if a pivot is not positive.  rn[i] accumulates sum_{k<j} C_ik^2.
__global__ void __launch_bounds__(kT) k_chol_dense(double* __restrict__ C, int r,
                                                   int* __restrict__ ok) {
  extern __shared__ double rn[];  // r doubles
  __shared__ double djj;
  const int t = threadIdx.x;
  for (int i = t; i < r; i += kT) rn[i] = 0.0;
  if (t == 0) *ok = 1;
  __syncthreads();
  for (int j = 0; j < r; ++j) {
    if (t == 0) {
      const double d = C[int64_t(j) * r + j] - rn[j];
      if (!(d > 0.0)) *ok = 0;
      djj = sqrt(d);
      C[int64_t(j) * r + j] = djj;
    }
    __syncthreads();
    const double* cj = C + int64_t(j) * r;
    for (int i = j + 1 + t; i < r; i += kT) {
      double* ci = C + int64_t(i) * r;
      double s = ci[j];
      for (int k = 0; k < j; ++k) s -= ci[k] * cj[k];
      s /= djj;
      ci[j] = s;
      rn[i] += s * s;
    }
    __syncthreads();
  }
}

constexpr int kLtRows = 256;

// Wp[(c * P + p) * r + a] = sum over rows m of chunk c of L[m][a] V[p][m]
__global__ void __launch_bounds__(kT)
    k_lt_v(const float* __restrict__ L, int64_t n, int r, const float* __restrict__ V, int64_t P,
           float* __restrict__ Wp) {
  __shared__ float vs[kLtRows];
  const int64_t c = blockIdx.x, p = blockIdx.y;
  const int64_t m0 = c * kLtRows;
  for (int mm = threadIdx.x; mm < kLtRows; mm += kT)
    vs[mm] = (m0 + mm < n) ? V[p * n + m0 + mm] : 0.f;
  __syncthreads();
  const int64_t m1 = m0 + kLtRows < n ? m0 + kLtRows : n;
  for (int a = threadIdx.x; a < r; a += kT) {
    float acc = 0.f;
    for (int64_t m = m0; m < m1; ++m) acc = fmaf(L[m * r + a], vs[m - m0], acc);
    Wp[(c * P + p) * r + a] = acc;
  }
}

// z = C^{-1} (scale * sum_c Wp[c][p]) with C = G G^T (lower G in C), one
// wave per right-hand side.
__global__ void __launch_bounds__(64)
    k_cho_solve(const double* __restrict__ C, int r, const float* __restrict__ Wp, int64_t nchunk,
                int64_t P, double scale, double* __restrict__ Z) {
  extern __shared__ double y[];  // r doubles
  const int64_t p = blockIdx.x;
  const int lane = threadIdx.x;
  for (int a = lane; a < r; a += 64) {
    double s = 0.0;
    for (int64_t c = 0; c < nchunk; ++c) s += double(Wp[(c * P + p) * r + a]);
    y[a] = scale * s;
  }
  __syncthreads();
  for (int i = 0; i < r; ++i) {  // G y' = y
    const double* ci = C + int64_t(i) * r;
    double s = 0.0;
    for (int k = lane; k < i; k += 64) s += ci[k] * y[k];
    s = wsum_d(s);
    if (lane == 0) y[i] = (y[i] - s) / ci[i];
    __syncthreads();
  }
  for (int i = r - 1; i >= 0; --i) {  // G^T z = y'
    const double* ci = C + int64_t(i) * r;
    if (lane == 0) y[i] = y[i] / ci[i];
    __syncthreads();
    const double zi = y[i];
    for (int k = lane; k < i; k += 64) y[k] -= ci[k] * zi;
    __syncthreads();
  }
  for (int a = lane; a < r; a += 64) Z[p * r + a] = y[a];
}

// out[p][m] = V[p][m] / s - (sum_a L[m][a] z[p][a]) / sqrt(s), one wave per row
__global__ void __launch_bounds__(kT)
    k_woodbury_out(const float* __restrict__ L, int64_t n, int r, const float* __restrict__ V,
                   const double* __restrict__ Z, double inv_s, double inv_sqrt_s,
                   float* __restrict__ out) {
  extern __shared__ float zs[];  // r floats
  const int64_t p = blockIdx.y;
  for (int a = threadIdx.x; a < r; a += kT) zs[a] = float(Z[p * r + a]);
  __syncthreads();
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int64_t m = int64_t(blockIdx.x) * (kT / 64) + wave;
  if (m >= n) return;
  float acc = 0.f;
  for (int a = lane; a < r; a += 64) acc = fmaf(L[m * r + a], zs[a], acc);
  acc = wsum_f(acc);
  if (lane == 0) out[p * n + m] = float(double(V[p * n + m]) * inv_s - double(acc) * inv_sqrt_s);
}

// ---- conjugate gradients ------------------------------------------------------

// Per-column partial dots a . b over chunks of kCgChunk elements.
__global__ void __launch_bounds__(kT)
    k_cg_dot(const float* __restrict__ a, const float* __restrict__ b, int64_t n,
             float* __restrict__ part) {
  __shared__ float red[kT / 64];
  const int64_t p = blockIdx.y, c = blockIdx.x, nb = gridDim.x;
  const int64_t base = c * kCgChunk;
  float acc = 0.f;
#pragma unroll
  for (int e = 0; e < kCgPer; ++e) {
    const int64_t i = base + int64_t(e) * kT + threadIdx.x;
    if (i < n) acc = fmaf(a[p * n + i], b[p * n + i], acc);
  }
  const float s = block_sum_f(acc, red);
  if (threadIdx.x == 0) part[p * nb + c] = s;
}

// q = sum_z qsum.partial[z] (+ noise d), the deferred split-K sum of the
// matvec of d added in the order of launch_matvec's own reduction (so q is
// the same, bit for bit), written to q, and the per-column partial dots d . q
// over the same chunks as k_cg_dot; the partials' loads are issued 16 at a
// time.
__global__ void __launch_bounds__(kTq)
    k_cg_dot_q(const float* __restrict__ d, MvSplitSum qsum, int64_t n, float* __restrict__ q,
               float* __restrict__ part) {
  __shared__ float red[kTq / 64];
  const int64_t p = blockIdx.y, c = blockIdx.x, nb = gridDim.x;
  const int64_t count = int64_t(gridDim.y) * n;
  const int64_t base = c * kCgChunk;
  float acc = 0.f;
#pragma unroll
  for (int e = 0; e < kQPer; ++e) {
    const int64_t i = base + int64_t(e) * kTq + threadIdx.x;
    if (i < n) {
      const int64_t at = p * n + i;
      float v = qsum.partial[at];
      for (int z0 = 1; z0 < qsum.S; z0 += 16) {
        float pz[16];
#pragma unroll
        for (int z = 0; z < 16; ++z)
          pz[z] = z0 + z < qsum.S ? qsum.partial[int64_t(z0 + z) * count + at] : 0.f;
#pragma unroll
        for (int z = 0; z < 16; ++z)
          if (z0 + z < qsum.S) v += pz[z];
      }
      const float dv = d[at];
      if (qsum.nugget) v = fmaf(qsum.noise, dv, v);
      q[at] = v;
      acc = fmaf(dv, v, acc);
    }
  }
  const float s = block_sum_f<kTq>(acc, red);
  if (threadIdx.x == 0) part[p * nb + c] = s;
}

__device__ __forceinline__ float sum_parts(const float* part, int64_t p, int64_t nb) {
  float s = 0.f;
  for (int64_t c = 0; c < nb; ++c) s += part[p * nb + c];
  return s;
}

// generation 0: gamma = r . z, atol2 = max(tol^2 b.b, atol^2), active
__global__ void __launch_bounds__(kT)
    k_cg_start(const float* __restrict__ rz, const float* __restrict__ bb,
               const float* __restrict__ rr, int64_t P, int64_t nb, float tol, float atol,
               int64_t maxiter, CgScalars sc, int* __restrict__ nactive) {
  __shared__ float red[kT / 64];
  float cnt = 0.f;
  for (int64_t p = threadIdx.x; p < P; p += kT) {
    const float g = sum_parts(rz, p, nb);
    const float bs = sum_parts(bb, p, nb);
    const float a2 = fmaxf(tol * tol * bs, atol * atol);
    const float rs = rr ? sum_parts(rr, p, nb) : g;
    sc.gamma[p] = g;
    sc.atol2[p] = a2;
    sc.iters[p] = 0;
    const int act = (rs > a2) && (maxiter > 0);
    sc.active[p] = act;
    cnt += float(act);
  }
  const float tot = block_sum_f(cnt, red);
  if (threadIdx.x == 0) *nactive = int(tot + 0.5f);
}

// alpha = gamma / (p . A p) (every block sums the partials itself), then
// x += alpha p, r -= alpha A p for the columns active in generation gen and
// the partial r . r
__global__ void __launch_bounds__(kTq)
    k_cg_update(float* __restrict__ x, float* __restrict__ r, const float* __restrict__ d,
                const float* __restrict__ q, int64_t n, const float* __restrict__ pq,
                CgScalars sc, int gen, float* __restrict__ rr) {
  __shared__ float red[kTq / 64];
  const int64_t p = blockIdx.y, c = blockIdx.x, nb = gridDim.x, P = gridDim.y;
  const int64_t base = c * kCgChunk;
  const bool act = sc.active[gen * P + p] != 0;
  const float al = act ? sc.gamma[gen * P + p] / sum_parts(pq, p, nb) : 0.f;
  float acc = 0.f;
#pragma unroll
  for (int e = 0; e < kQPer; ++e) {
    const int64_t i = base + int64_t(e) * kTq + threadIdx.x;
    if (i < n) {
      float rv = r[p * n + i];
      if (act) {
        x[p * n + i] = fmaf(al, d[p * n + i], x[p * n + i]);
        rv = fmaf(-al, q[p * n + i], rv);
        r[p * n + i] = rv;
      }
      acc = fmaf(rv, rv, acc);
    }
  }
  const float s = block_sum_f<kTq>(acc, red);
  if (threadIdx.x == 0) rr[p * nb + c] = s;
}

// For the columns active in generation gen: gamma' = r . z, beta = gamma' /
// gamma and p = z + beta p (every block sums the partials itself); block 0 of
// each column writes generation gen ^ 1 of gamma and active (the stopping
// rule) and the iteration count.
__global__ void __launch_bounds__(kTq)
    k_cg_beta_direction(float* __restrict__ d, const float* __restrict__ z, int64_t n,
                        const float* __restrict__ rz, const float* __restrict__ rr, int precond,
                        int64_t maxiter, CgScalars sc, int gen) {
  const int64_t p = blockIdx.y, c = blockIdx.x, nb = gridDim.x, P = gridDim.y;
  const int64_t cur = gen * P + p, nxt = (gen ^ 1) * P + p;
  const int act = sc.active[cur];
  const float g = act ? sum_parts(precond ? rz : rr, p, nb) : 0.f;
  if (c == 0 && threadIdx.x == 0) {
    int next = 0;
    if (act) {
      const int64_t k = sc.iters[p] + 1;
      sc.iters[p] = k;
      const float rs = precond ? sum_parts(rr, p, nb) : g;
      next = (rs > sc.atol2[p]) && (k < maxiter);
    }
    sc.gamma[nxt] = act ? g : sc.gamma[cur];
    sc.active[nxt] = next;
  }
  if (!act) return;
  const float be = g / sc.gamma[cur];
  const int64_t base = c * kCgChunk;
#pragma unroll
  for (int e = 0; e < kQPer; ++e) {
    const int64_t i = base + int64_t(e) * kTq + threadIdx.x;
    if (i < n) d[p * n + i] = fmaf(be, d[p * n + i], z[p * n + i]);
  }
}

// *nactive = number of columns active in generation gen
__global__ void __launch_bounds__(kT)
    k_cg_count(CgScalars sc, int gen, int64_t P, int* __restrict__ nactive) {
  __shared__ float red[kT / 64];
  float cnt = 0.f;
  for (int64_t p = threadIdx.x; p < P; p += kT) cnt += float(sc.active[gen * P + p] != 0);
  const float tot = block_sum_f(cnt, red);
  if (threadIdx.x == 0) *nactive = int(tot + 0.5f);
}

inline unsigned blocks_for(int64_t count) { return unsigned((count + kT - 1) / kT); }

}  // namespace

int64_t cg_chunks(int64_t n) { return (n + kCgChunk - 1) / kCgChunk; }

hipError_t launch_cholesky_partial(hipStream_t s, const slq_operator& op, int rank, bool pivot,
                                   float* Lc, int64_t* perm, float* bval, int64_t* bidx,
                                   float* rowi, float* lii, int* success, float* L,
                                   int64_t* perm_out) {
  const int64_t n = op.n;
  Elem E{op.kind == SLQ_OP_DENSE ? op.matrix : nullptr,
         op.kind == SLQ_OP_RBF ? op.points : nullptr, op.dim, n,
         float(0.5 / (op.lengthscale * op.lengthscale)),
         op.kind == SLQ_OP_RBF ? float(op.noise) : 0.f};
  const unsigned nblk = blocks_for(n);
  hipLaunchKernelGGL(k_chol_init, dim3(nblk), dim3(kT), 0, s, n, perm, success);
  for (int i = 0; i < rank; ++i) {
    if (pivot)
      hipLaunchKernelGGL(k_chol_residual, dim3(nblk), dim3(kT), 0, s, Lc, n, i, perm, E, bval,
                         bidx);
    hipLaunchKernelGGL(k_chol_pivot, dim3(1), dim3(kT), 0, s, Lc, n, i, perm, E, bval, bidx,
                       int(nblk), pivot ? 1 : 0, rowi, lii, success);
    hipLaunchKernelGGL(k_chol_column, dim3(nblk), dim3(kT), size_t(i) * sizeof(float), s, Lc, n,
                       i, perm, E, rowi, lii);
  }
  hipLaunchKernelGGL(k_chol_output, dim3(blocks_for(n * rank)), dim3(kT), 0, s, Lc, n, rank, perm,
                     L, perm_out);
  return hipGetLastError();
}

int64_t gram_splits(int64_t n) {
  const int64_t per = 1024;
  const int64_t q = (n + per - 1) / per;
  return q < 1 ? 1 : q;
}

hipError_t launch_woodbury_setup(hipStream_t s, const float* L, int64_t n, int r, double shift,
                                 float* Gp, double* C, int* ok) {
  const int64_t nsplit = gram_splits(n);
  const int64_t rows = (n + nsplit - 1) / nsplit;
  const dim3 grid(unsigned((r + kGT - 1) / kGT), unsigned((r + kGT - 1) / kGT), unsigned(nsplit));
  hipLaunchKernelGGL(k_gram_partial, grid, dim3(kGT * kGT), 0, s, L, n, r, rows, Gp);
  hipLaunchKernelGGL(k_gram_reduce, dim3(blocks_for(int64_t(r) * r)), dim3(kT), 0, s, Gp,
                     int(nsplit), r, 1.0 / shift, C);
  hipLaunchKernelGGL(k_chol_dense, dim3(1), dim3(kT), size_t(r) * sizeof(double), s, C, r, ok);
  return hipGetLastError();
}

int64_t woodbury_chunks(int64_t n) { return (n + kLtRows - 1) / kLtRows; }

hipError_t launch_woodbury_apply(hipStream_t s, const float* L, int64_t n, int r, double shift,
                                 const double* C, const float* V, int64_t P, float* Wp, double* Z,
                                 float* out) {
  const int64_t nc = woodbury_chunks(n);
  hipLaunchKernelGGL(k_lt_v, dim3(unsigned(nc), unsigned(P)), dim3(kT), 0, s, L, n, r, V, P, Wp);
  // V^T (v / s) with V = L^T / sqrt(s) (low_rank.py:43-44, 50)
  const double scale = 1.0 / (shift * sqrt(shift));
  hipLaunchKernelGGL(k_cho_solve, dim3(unsigned(P)), dim3(64), size_t(r) * sizeof(double), s, C,
                     r, Wp, nc, P, scale, Z);
  const dim3 grid(unsigned((n + kT / 64 - 1) / (kT / 64)), unsigned(P));
  hipLaunchKernelGGL(k_woodbury_out, grid, dim3(kT), size_t(r) * sizeof(float), s, L, n, r, V, Z,
                     1.0 / shift, 1.0 / sqrt(shift), out);
  return hipGetLastError();
}

hipError_t launch_cg_dot(hipStream_t s, const float* a, const float* b, int64_t P, int64_t n,
                         float* part) {
  hipLaunchKernelGGL(k_cg_dot, dim3(unsigned(cg_chunks(n)), unsigned(P)), dim3(kT), 0, s, a, b, n,
                     part);
  return hipGetLastError();
}

hipError_t launch_cg_start(hipStream_t s, const float* rz, const float* bb, const float* rr,
                           int64_t P, int64_t n, float tol, float atol, int64_t maxiter,
                           const CgScalars& sc, int* nactive) {
  hipLaunchKernelGGL(k_cg_start, dim3(1), dim3(kT), 0, s, rz, bb, rr, P, cg_chunks(n), tol, atol,
                     maxiter, sc, nactive);
  return hipGetLastError();
}

hipError_t launch_cg_step(hipStream_t s, float* x, float* r, float* d, float* q,
                          const MvSplitSum& qsum, float* z, bool precond, int64_t P, int64_t n,
                          int64_t maxiter, const CgScalars& sc, int gen, float* pq, float* rr,
                          float* rz, const CgPrecond* M) {
  const int64_t nb = cg_chunks(n);
  const dim3 grid{unsigned(nb), unsigned(P)};
  if (qsum.partial)
    hipLaunchKernelGGL(k_cg_dot_q, grid, dim3(kTq), 0, s, d, qsum, n, q, pq);
  else
    hipLaunchKernelGGL(k_cg_dot, grid, dim3(kT), 0, s, d, q, n, pq);
  hipLaunchKernelGGL(k_cg_update, grid, dim3(kTq), 0, s, x, r, d, q, n, pq, sc, gen, rr);
  if (precond) {
    hipError_t e = launch_woodbury_apply(s, M->L, n, M->rank, M->shift, M->C, r, P, M->Wp, M->Z, z);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_cg_dot, grid, dim3(kT), 0, s, r, z, n, rz);
  }
  hipLaunchKernelGGL(k_cg_beta_direction, grid, dim3(kTq), 0, s, d, precond ? z : r, n, rz, rr,
                     precond ? 1 : 0, maxiter, sc, gen);
  return hipGetLastError();
}

hipError_t launch_cg_count(hipStream_t s, const CgScalars& sc, int gen, int64_t P, int* nactive) {
  hipLaunchKernelGGL(k_cg_count, dim3(1), dim3(kT), 0, s, sc, gen, P, nactive);
  return hipGetLastError();
}

}  // namespace slq
