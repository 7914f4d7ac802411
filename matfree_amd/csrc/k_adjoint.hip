// Reverse pass of the SLQ log-determinant: the Arnoldi adjoint with full
// reorthogonalisation (decomp.py:480-600) driven by the cotangent of
// |z|^2 [log T]_00 (funm.py:239-241, 331-333) and the probe mean
// (stochtrace.py:50).  The quadratic form discards Q, r and c
// (funm.py:237), so dQ = dr = dc = 0 and the adjoint's inputs reduce to dH.
//
// Kernels:
//  * the eigen cotangent dH comes from k_tridiag.hip (k_tridiag_vjp);
//  * k_adjoint_coef / k_adjoint_combine: one step of the backward
//    recursion (decomp.py:587-599) given the reductions of the streaming
//    kernels (k_lanczos.hip).
//  * k_outer_sum: the parameter cotangents from S = sum lambda q^T -- the
//    stored matrix's dA = S, or the RBF contractions <S, dK/dl>, <S, I>;
//    k_sum_fixed: their deterministic FP64 reduction.
#include "slq_internal.h"

namespace slq {
namespace {

// ---------------------------------------------------------------------------
// Coefficients of one backward step idx (decomp.py:592-599), one block per
// probe.  With s_j = q_j^T (A^T lambda) (sum of the partials over chunks):
//   Gamma[idx, j] = lower_mask[idx, j] (Pi_gamma[idx, j] - s_j), j <= idx,
//     lower_mask = tril(1) with 1/2 on the diagonal (decomp.py:490-495),
//     Pi_gamma = H dH^T (dQ = dc = 0, decomp.py:506), dH tridiagonal;
//   lambda_new = (gamma_idx r + sum_j (Gamma + Gamma^T)[idx, j] q_j
//                 - alpha_idx lambda + A^T lambda
//                 - sum_{j > idx} H[idx, j] Lambda_j) / beta_minus_idx
//     with gamma = dH[:, k-1] (dr = 0, decomp.py:498, 505),
//     beta_minus = (1, H[1,0], H[2,1], ...) (decomp.py:513) and beta_plus
//     the strict upper part of H beyond the diagonal (decomp.py:515).
// Output per probe (stride cs = 2 k + 4): gq[0..k), gl[k..2k), and
// (cr, cl, cv, scale) for r, lambda, A^T lambda and the final division.
// ---------------------------------------------------------------------------
//
// General cotangents (decomp.py:497-506): Pg (P x k x k, column-major) holds
// Pi_gamma = -dc c e_1 e_1^T + H dH^T - dQ^T Q and gam (P x k) the vector
// gamma = dH e_K - Q^T dr; when they are null, dQ = dr = dc = 0 and dH is
// tridiagonal (the SLQ composition) and both are formed here.
__global__ void __launch_bounds__(256)
    k_adjoint_coef(const float* __restrict__ H, const float* __restrict__ dH,
                   float* __restrict__ Gam, const float* __restrict__ part, int64_t nb,
                   int pstride, int64_t k, int64_t idx, float* __restrict__ coef,
                   const float* __restrict__ Pg, const float* __restrict__ gam) {
  extern __shared__ float grow[];  // Gamma[idx, 0..idx]
  const int64_t p = blockIdx.x;
  const float* Hp = H + p * k * k;    // column-major: H[row + col * k]
  const float* dHp = dH + p * k * k;  // column-major
  float* Gp = Gam + p * k * k;        // column-major
  float* cp = coef + p * (2 * k + 4);
  for (int64_t j = threadIdx.x; j <= idx; j += blockDim.x) {
    float s = 0.f;
    for (int64_t b = 0; b < nb; ++b) s += part[(p * nb + b) * pstride + j];
    float pg;
    if (Pg) {
      pg = Pg[p * k * k + idx + j * k];
    } else {
      pg = Hp[idx + j * k] * dHp[j + j * k];
      if (j > 0) pg += Hp[idx + (j - 1) * k] * dHp[j + (j - 1) * k];
      if (j + 1 < k) pg += Hp[idx + (j + 1) * k] * dHp[j + (j + 1) * k];
    }
    const float w = (j < idx) ? 1.f : 0.5f;
    const float g = w * (pg - s);
    grow[j] = g;
    Gp[idx + j * k] = g;
  }
  __syncthreads();
  for (int64_t j = threadIdx.x; j < k; j += blockDim.x) {
    float gq;
    if (j < idx)
      gq = grow[j];
    else if (j == idx)
      gq = 2.f * grow[j];
    else
      gq = Gp[j + idx * k];  // Gamma[j, idx], set at step j
    cp[j] = gq;
    cp[k + j] = (j > idx) ? -Hp[idx + j * k] : 0.f;
  }
  if (threadIdx.x == 0) {
    cp[2 * k + 0] = gam ? gam[p * k + idx] : dHp[idx + (k - 1) * k];  // gamma_idx
    cp[2 * k + 1] = -Hp[idx + idx * k];      // -alpha_idx
    cp[2 * k + 2] = 1.f;                     // + A^T lambda
    cp[2 * k + 3] = 1.f / (idx == 0 ? 1.f : Hp[idx + (idx - 1) * k]);
  }
}

// y = scale (cr r + cl x + cv v + d + sum_{j < ncq} gq_j Q_j + sum_{jl0 <= j < k} gl_j L_j)
// per probe, coefficients from coef (stride 2 k + 4); r, x, v, d may be null
// (their terms are skipped).  With save (non-null) x is also copied there.
template <int V>
__global__ void __launch_bounds__(kStreamThreads)
    k_adjoint_combine(const float* __restrict__ Q, const float* __restrict__ Lam, int64_t ldc,
                      const float* __restrict__ r, const float* __restrict__ x,
                      const float* __restrict__ v, float* __restrict__ y, float* __restrict__ save,
                      int64_t n, int64_t k, int ncq, int jl0, const float* __restrict__ coef,
                      const float* __restrict__ d) {
  extern __shared__ float cs[];
  const int64_t p = blockIdx.y;
  const int t = threadIdx.x;
  const float* cp = coef + p * (2 * k + 4);
  for (int64_t j = t; j < 2 * k + 4; j += blockDim.x) cs[j] = cp[j];
  __syncthreads();
  const float cr = cs[2 * k], cl = cs[2 * k + 1], cv = cs[2 * k + 2], sc = cs[2 * k + 3];
  const int64_t per_block = int64_t(kStreamThreads) * V;
  const int64_t i = int64_t(blockIdx.x) * per_block + int64_t(t) * V;
  if (i >= n) return;
  const int64_t off = p * n + i;
  float acc[V];
#pragma unroll
  for (int u = 0; u < V; ++u) acc[u] = 0.f;
  auto axpy = [&](float a, const float* src) {
    if constexpr (V == 4) {
      const float4 q = *reinterpret_cast<const float4*>(src);
      acc[0] = fmaf(a, q.x, acc[0]);
      acc[1] = fmaf(a, q.y, acc[1]);
      acc[2] = fmaf(a, q.z, acc[2]);
      acc[3] = fmaf(a, q.w, acc[3]);
    } else {
      acc[0] = fmaf(a, src[0], acc[0]);
    }
  };
  for (int j = 0; j < ncq; ++j) axpy(cs[j], Q + int64_t(j) * ldc + off);
  for (int64_t j = jl0; j < k; ++j) axpy(cs[k + j], Lam + j * ldc + off);
  if (r) axpy(cr, r + off);
  if (x) axpy(cl, x + off);
  if (v) axpy(cv, v + off);
  if (d) axpy(1.f, d + off);
  float out[V];
#pragma unroll
  for (int u = 0; u < V; ++u) out[u] = sc * acc[u];
  if constexpr (V == 4) {
    *reinterpret_cast<float4*>(y + off) = make_float4(out[0], out[1], out[2], out[3]);
    if (save) *reinterpret_cast<float4*>(save + off) = *reinterpret_cast<const float4*>(x + off);
  } else {
    y[off] = out[0];
    if (save) save[off] = x[off];
  }
}

// out[o] = f_o * sum_i partial[i * nout + o], i < count, o < nout <= 2: one
// block per output, a fixed strided split over the 256 threads and a fixed
// tree (deterministic).
__global__ void __launch_bounds__(256)
    k_sum_fixed(const double* __restrict__ partial, int64_t count, int nout, double f0, double f1,
                double* __restrict__ out) {
  __shared__ double red[256];
  const int o = blockIdx.x;
  double s = 0.0;
  for (int64_t i = threadIdx.x; i < count; i += 256) s += partial[i * nout + o];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[o] = (o == 0 ? f0 : f1) * red[0];
}

// Coefficients that make k_adjoint_combine compute lambda = Q gamma with
// gamma = dH[:, k - 1] (decomp.py:498-499 with dr = 0).
__global__ void k_adjoint_init_coef(const float* __restrict__ dH, int64_t k,
                                    float* __restrict__ coef) {
  const int64_t p = blockIdx.x;
  float* cp = coef + p * (2 * k + 4);
  const float* dHp = dH + p * k * k;
  for (int64_t j = threadIdx.x; j < 2 * k + 4; j += blockDim.x) {
    float v = 0.f;
    if (j < k) v = dHp[j + (k - 1) * k];
    if (j == 2 * k + 3) v = 1.f;
    cp[j] = v;
  }
}

// Row i of Pi_gamma = -dc c e_1 e_1^T + H dH^T - dQ^T Q (decomp.py:506) per
// probe: (H dH^T)[i, j] = sum_m H[i, m] dH[j, m] (both column-major), and
// (dQ^T Q)[i, j] = dQ_i . q_j from the chunk partials of launch_dots with
// x = dQ_i (null: dQ = 0); dc / c per probe (dc null: 0).  Pg column-major.
__global__ void __launch_bounds__(256)
    k_adjoint_pigamma_row(const float* __restrict__ H, const float* __restrict__ dH,
                          const float* __restrict__ part, int64_t nb, int pstride,
                          const float* __restrict__ c, const float* __restrict__ dc, int64_t k,
                          int64_t i, float* __restrict__ Pg) {
  const int64_t p = blockIdx.x;
  const float* Hp = H + p * k * k;
  const float* dHp = dH + p * k * k;
  for (int64_t j = threadIdx.x; j < k; j += blockDim.x) {
    float v = 0.f;
    for (int64_t m = 0; m < k; ++m) v = fmaf(Hp[i + m * k], dHp[j + m * k], v);
    if (part) {
      float g = 0.f;
      for (int64_t b = 0; b < nb; ++b) g += part[(p * nb + b) * pstride + j];
      v -= g;
    }
    if (dc && i == 0 && j == 0) v -= dc[p] * c[p];
    Pg[p * k * k + i + j * k] = v;
  }
}

// gamma = dH e_K - Q^T dr (decomp.py:498; part: chunk partials of Q^T dr,
// null for dr = 0) into gam (P x k), and the coefficients that make
// k_adjoint_combine compute lambda = dr + Q gamma (decomp.py:499; x = dr).
__global__ void __launch_bounds__(256)
    k_adjoint_gamma_init(const float* __restrict__ dH, const float* __restrict__ part, int64_t nb,
                         int pstride, int64_t k, float* __restrict__ gam,
                         float* __restrict__ coef) {
  const int64_t p = blockIdx.x;
  float* cp = coef + p * (2 * k + 4);
  const float* dHp = dH + p * k * k;
  for (int64_t j = threadIdx.x; j < 2 * k + 4; j += blockDim.x) {
    float v = 0.f;
    if (j < k) {
      v = dHp[j + (k - 1) * k];
      if (part) {
        float g = 0.f;
        for (int64_t b = 0; b < nb; ++b) g += part[(p * nb + b) * pstride + j];
        v -= g;
      }
      gam[p * k + j] = v;
    }
    if (j == 2 * k + 1) v = 1.f;  // x = dr
    if (j == 2 * k + 3) v = 1.f;
    cp[j] = v;
  }
}

// ---------------------------------------------------------------------------
// Stored-matrix cotangent dA = sum_c L_c R_c^T over the K = k P columns
// c = (idx, p) of two bases in the Krylov layout (column c at
// base + idx * ldc + p * n): the matvec VJPs lambda q^T of decomp.py:587-589
// summed over steps and probes.  FP32 MFMA (32x32x2): a workgroup owns a
// 128 x 128 output tile, each wave 32 rows x 128 columns (4 accumulators);
// both operands pass through LDS in stages of 8 columns (one float4 of each
// per thread), double-buffered.  In an MFMA the lane half h carries the
// contraction column 2 q + h of the stage, lane r the row / column r.
// ---------------------------------------------------------------------------
typedef float floatx16 __attribute__((ext_vector_type(16)));
constexpr int kOT = 128;         // output tile edge
constexpr int kOK = 8;           // columns per stage
constexpr int kOLd = kOT + 4;    // padded LDS row

// RBF contraction mode: instead of storing S = sum_c L_c R_c^T, reduce
//   sum_ij S_ij |x_i - x_j|^2 exp2(c |x_i - x_j|^2)   (l^3 times <S, dK/dl>)
// and sum_i S_ii (<S, I>) per workgroup into red[2 * block + {0, 1}] (FP64).
struct RBFContract {
  const float* X;
  int dim;
  float coef;
  double* red;
};

template <bool CONTRACT>
__global__ void __launch_bounds__(256)
    k_outer_sum(const float* __restrict__ L, const float* __restrict__ R, int64_t ldc, int64_t P,
                int64_t n, int64_t K, float* __restrict__ out, RBFContract rc) {
  __shared__ __attribute__((aligned(16))) float Ls[2][kOK * kOLd];
  __shared__ __attribute__((aligned(16))) float Rs[2][kOK * kOLd];
  const int t = threadIdx.x;
  const int lane = t & 63, wave = t >> 6, r = lane & 31, h = lane >> 5;
  const int64_t i0 = int64_t(blockIdx.y) * kOT, j0 = int64_t(blockIdx.x) * kOT;
  // staging: thread t moves float4 (t & 31) of stage column t >> 5
  const int sc = t >> 5, s4 = (t & 31) * 4;
  auto col_ptr = [&](const float* base, int64_t c) {
    const int64_t idx = c / P, p = c - idx * P;
    return base + idx * ldc + p * n;
  };
  auto load = [&](const float* base, int64_t c0, int64_t e0) {
    float4 v = make_float4(0.f, 0.f, 0.f, 0.f);
    const int64_t c = c0 + sc, e = e0 + s4;
    if (c < K) {
      const float* src = col_ptr(base, c) + e;
      if (e + 3 < n && (n & 3) == 0) {
        v = *reinterpret_cast<const float4*>(src);
      } else {
        if (e + 0 < n) v.x = src[0];
        if (e + 1 < n) v.y = src[1];
        if (e + 2 < n) v.z = src[2];
        if (e + 3 < n) v.w = src[3];
      }
    }
    return v;
  };
  floatx16 acc[4];
#pragma unroll
  for (int c = 0; c < 4; ++c)
#pragma unroll
    for (int e = 0; e < 16; ++e) acc[c][e] = 0.f;

  const int64_t nst = (K + kOK - 1) / kOK;
  float4 lv = load(L, 0, i0), rv = load(R, 0, j0);
  *reinterpret_cast<float4*>(&Ls[0][sc * kOLd + s4]) = lv;
  *reinterpret_cast<float4*>(&Rs[0][sc * kOLd + s4]) = rv;
  __syncthreads();
  for (int64_t st = 0; st < nst; ++st) {
    const int b = int(st & 1);
    const bool more = st + 1 < nst;
    if (more) {
      lv = load(L, (st + 1) * kOK, i0);
      rv = load(R, (st + 1) * kOK, j0);
    }
#pragma unroll
    for (int q = 0; q < kOK / 2; ++q) {
      const float a = Ls[b][(2 * q + h) * kOLd + 32 * wave + r];
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        const float bb = Rs[b][(2 * q + h) * kOLd + 32 * c + r];
        acc[c] = __builtin_amdgcn_mfma_f32_32x32x2f32(a, bb, acc[c], 0, 0, 0);
      }
    }
    if (more) {
      *reinterpret_cast<float4*>(&Ls[b ^ 1][sc * kOLd + s4]) = lv;
      *reinterpret_cast<float4*>(&Rs[b ^ 1][sc * kOLd + s4]) = rv;
    }
    __syncthreads();
  }
  // lane l of accumulator c holds column j0 + 32 c + (l & 31) at rows
  // i0 + 32 wave + 8 (s / 4) + 4 h + (s % 4)
  if constexpr (!CONTRACT) {
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      const int64_t j = j0 + 32 * c + r;
      if (j >= n) continue;
#pragma unroll
      for (int s = 0; s < 16; ++s) {
        const int64_t i = i0 + 32 * wave + 8 * (s / 4) + 4 * h + (s % 4);
        if (i < n) out[i * n + j] = acc[c][s];
      }
    }
  } else {
    double sd = 0.0, st2 = 0.0;
    float xi[16][3];
#pragma unroll
    for (int s = 0; s < 16; ++s) {
      const int64_t i = i0 + 32 * wave + 8 * (s / 4) + 4 * h + (s % 4);
#pragma unroll
      for (int d = 0; d < 3; ++d) xi[s][d] = (i < n && d < rc.dim) ? rc.X[i * rc.dim + d] : 0.f;
    }
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      const int64_t j = j0 + 32 * c + r;
      if (j >= n) continue;
      float xj[3];
#pragma unroll
      for (int d = 0; d < 3; ++d) xj[d] = d < rc.dim ? rc.X[j * rc.dim + d] : 0.f;
      float part = 0.f;
#pragma unroll
      for (int s = 0; s < 16; ++s) {
        const int64_t i = i0 + 32 * wave + 8 * (s / 4) + 4 * h + (s % 4);
        if (i >= n) continue;
        float d2 = 0.f;
#pragma unroll
        for (int d = 0; d < 3; ++d) {
          const float df = xi[s][d] - xj[d];
          d2 = fmaf(df, df, d2);
        }
        part = fmaf(acc[c][s], d2 * __builtin_amdgcn_exp2f(rc.coef * d2), part);
        if (i == j) st2 += double(acc[c][s]);
      }
      sd += double(part);
    }
    __shared__ double red[2][256];
    red[0][t] = sd;
    red[1][t] = st2;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
      if (t < w) {
        red[0][t] += red[0][t + w];
        red[1][t] += red[1][t + w];
      }
      __syncthreads();
    }
    if (t == 0) {
      const int64_t b = int64_t(blockIdx.y) * gridDim.x + blockIdx.x;
      rc.red[2 * b] = red[0][0];
      rc.red[2 * b + 1] = red[1][0];
    }
  }
}

}  // namespace

hipError_t launch_outer_sum(hipStream_t s, const float* L, const float* R, int64_t ldc, int64_t P,
                            int64_t n, int64_t K, float* out) {
  const unsigned tiles = unsigned((n + kOT - 1) / kOT);
  hipLaunchKernelGGL(k_outer_sum<false>, dim3(tiles, tiles), dim3(256), 0, s, L, R, ldc, P, n, K,
                     out, RBFContract{});
  return hipGetLastError();
}

int64_t outer_contract_blocks(int64_t n) {
  const int64_t tiles = (n + kOT - 1) / kOT;
  return tiles * tiles;
}

hipError_t launch_outer_rbf_contract(hipStream_t s, const float* L, const float* R, int64_t ldc,
                                     int64_t P, int64_t n, int64_t K, const float* X, int dim,
                                     double lengthscale, double* red) {
  const unsigned tiles = unsigned((n + kOT - 1) / kOT);
  RBFContract rc{X, dim, float(-1.4426950408889634 / (2.0 * lengthscale * lengthscale)), red};
  hipLaunchKernelGGL(k_outer_sum<true>, dim3(tiles, tiles), dim3(256), 0, s, L, R, ldc, P, n, K,
                     nullptr, rc);
  return hipGetLastError();
}

hipError_t launch_adjoint_coef(hipStream_t s, const float* H, const float* dH, float* Gam,
                               const float* part, int pstride, int64_t P, int64_t n, int64_t k,
                               int64_t idx, float* coef, const float* Pg, const float* gam) {
  const size_t shmem = size_t(idx + 1) * sizeof(float);
  hipLaunchKernelGGL(k_adjoint_coef, dim3(unsigned(P)), dim3(256), shmem, s, H, dH, Gam, part,
                     num_chunks(n), pstride, k, idx, coef, Pg, gam);
  return hipGetLastError();
}

hipError_t launch_adjoint_pigamma_row(hipStream_t s, const float* H, const float* dH,
                                      const float* part, int pstride, const float* c,
                                      const float* dc, int64_t P, int64_t n, int64_t k, int64_t i,
                                      float* Pg) {
  hipLaunchKernelGGL(k_adjoint_pigamma_row, dim3(unsigned(P)), dim3(256), 0, s, H, dH, part,
                     num_chunks(n), pstride, c, dc, k, i, Pg);
  return hipGetLastError();
}

hipError_t launch_adjoint_gamma_init(hipStream_t s, const float* dH, const float* part, int pstride,
                                     int64_t P, int64_t n, int64_t k, float* gam, float* coef) {
  hipLaunchKernelGGL(k_adjoint_gamma_init, dim3(unsigned(P)), dim3(256), 0, s, dH, part,
                     num_chunks(n), pstride, k, gam, coef);
  return hipGetLastError();
}

hipError_t launch_adjoint_combine(hipStream_t s, const float* Q, const float* Lam, int64_t ldc,
                                  const float* r, const float* x, const float* v, float* y,
                                  float* save, int64_t P, int64_t n, int64_t k, int ncq, int jl0,
                                  const float* coef, const float* d) {
  const auto al = [](const void* q) { return (reinterpret_cast<uintptr_t>(q) & 15) == 0; };
  const bool vec = n % 4 == 0 && ldc % 4 == 0 && al(Q) && al(y) && (!Lam || al(Lam)) &&
                   (!r || al(r)) && (!x || al(x)) && (!v || al(v)) && (!save || al(save)) &&
                   (!d || al(d));
  const int V = vec ? 4 : 1;
  const int64_t per_block = int64_t(kStreamThreads) * V;
  const dim3 grid(unsigned((n + per_block - 1) / per_block), unsigned(P));
  const size_t shmem = size_t(2 * k + 4) * sizeof(float);
  if (vec)
    hipLaunchKernelGGL(k_adjoint_combine<4>, grid, dim3(kStreamThreads), shmem, s, Q, Lam, ldc, r,
                       x, v, y, save, n, k, ncq, jl0, coef, d);
  else
    hipLaunchKernelGGL(k_adjoint_combine<1>, grid, dim3(kStreamThreads), shmem, s, Q, Lam, ldc, r,
                       x, v, y, save, n, k, ncq, jl0, coef, d);
  return hipGetLastError();
}

hipError_t launch_sum_fixed(hipStream_t s, const double* partial, int64_t count, int nout,
                            double f0, double f1, double* out) {
  hipLaunchKernelGGL(k_sum_fixed, dim3(unsigned(nout)), dim3(256), 0, s, partial, count, nout, f0,
                     f1, out);
  return hipGetLastError();
}

hipError_t launch_adjoint_init_coef(hipStream_t s, const float* dH, int64_t P, int64_t k,
                                    float* coef) {
  hipLaunchKernelGGL(k_adjoint_init_coef, dim3(unsigned(P)), dim3(256), 0, s, dH, k, coef);
  return hipGetLastError();
}

}  // namespace slq
