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
//    stored matrix's dA = S, or the RBF contractions <S, dK/dl>, <S, I> of
//    FP32 operators (for BF16 operators the block matvec's derivative mode
//    in k_matvec.hip computes them); k_sum_fixed: their deterministic FP64
//    reduction.
#include <type_traits>

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
    s = ordered_sum(part + p * nb * pstride + j, nb, pstride);
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
// V consecutive elements per thread (V = 4: 16-byte loads and stores at any
// dword alignment, which gfx950 code issues as single dwordx4 accesses, so
// rows of any length n stream as float4 groups, the last group of a row
// element by element; V = 1 for small P n, where one element per thread
// gives the launch four times the waves); the basis columns eight at a time,
// their loads in flight together.
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
  const int64_t i = (int64_t(blockIdx.x) * kStreamThreads + t) * V;
  if (i >= n) return;
  // W elements from offset off (W = 4: one 16-byte access per operand)
  auto combine = [&](auto width, int64_t off) {
    constexpr int W = decltype(width)::value;
    using VT = typename std::conditional<W == 4, float4, float>::type;
    auto ld = [](const float* q) {
      VT val;
      __builtin_memcpy(&val, q, sizeof(VT));
      return val;
    };
    float acc[W];
#pragma unroll
    for (int u = 0; u < W; ++u) acc[u] = 0.f;
    auto fma_in = [&](float a, const VT& q) {
      if constexpr (W == 4) {
        acc[0] = fmaf(a, q.x, acc[0]);
        acc[1] = fmaf(a, q.y, acc[1]);
        acc[2] = fmaf(a, q.z, acc[2]);
        acc[3] = fmaf(a, q.w, acc[3]);
      } else {
        acc[0] = fmaf(a, q, acc[0]);
      }
    };
    auto cols = [&](const float* base, int64_t j0, int64_t j1, const float* c) {
      int64_t j = j0;
      for (; j + 8 <= j1; j += 8) {
        VT q[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) q[u] = ld(base + (j + u) * ldc + off);
#pragma unroll
        for (int u = 0; u < 8; ++u) fma_in(c[j + u], q[u]);
      }
      for (; j < j1; ++j) fma_in(c[j], ld(base + j * ldc + off));
    };
    cols(Q, 0, ncq, cs);
    if (Lam) cols(Lam, jl0, k, cs + k);
    if (r) fma_in(cr, ld(r + off));
    if (x) fma_in(cl, ld(x + off));
    if (v) fma_in(cv, ld(v + off));
    if (d) fma_in(1.f, ld(d + off));
    VT out;
    if constexpr (W == 4)
      out = make_float4(sc * acc[0], sc * acc[1], sc * acc[2], sc * acc[3]);
    else
      out = sc * acc[0];
    __builtin_memcpy(y + off, &out, sizeof(VT));
    if (save) {
      const VT xs = ld(x + off);
      __builtin_memcpy(save + off, &xs, sizeof(VT));
    }
  };
  if (V == 1 || i + 4 <= n) {
    combine(std::integral_constant<int, V>{}, p * n + i);
  } else {  // the row's last, partial group
    for (int64_t e = i; e < n; ++e) combine(std::integral_constant<int, 1>{}, p * n + e);
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
      g = ordered_sum(part + p * nb * pstride + j, nb, pstride);
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
        g = ordered_sum(part + p * nb * pstride + j, nb, pstride);
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

// Order of the output tiles over the workgroups.  Workgroups go to the 8
// XCDs round-robin by their index, and each XCD caches in its own L2, so
// workgroup b (XCD b % 8, per-XCD slot b / 8) takes tile (r / S, r % S) of
// square q = (b / 8) / S^2 * 8 + b % 8 with r = (b / 8) % S^2: the S x S
// (S = 8) workgroups an XCD runs together read S row panels of L and S of R
// instead of S^2 of each.  Squares are laid over the T x T tiles row-major;
// workgroups whose tile lies outside (partial squares, the last ragged group
// of eight squares) write zeros.
struct TileMap {
  int64_t T, S, ST, blocks;
};

inline TileMap tile_map(int64_t n) {
  TileMap m;
  m.T = (n + kOT - 1) / kOT;
  m.S = m.T < 8 ? m.T : 8;
  m.ST = (m.T + m.S - 1) / m.S;
  m.blocks = (m.ST * m.ST + 7) / 8 * 8 * m.S * m.S;
  return m;
}

__device__ __forceinline__ bool tile_of_block(const TileMap& m, int64_t b, int64_t& ti,
                                              int64_t& tj) {
  const int64_t slot = b / 8, x = b % 8;
  const int64_t q = slot / (m.S * m.S), r = slot % (m.S * m.S);
  const int64_t g = q * 8 + x;
  if (g >= m.ST * m.ST) return false;
  ti = (g / m.ST) * m.S + r / m.S;
  tj = (g % m.ST) * m.S + r % m.S;
  return ti < m.T && tj < m.T;
}

// Column c = (idx, p) of a Krylov-layout basis (c = idx P + p) walked in
// steps of a fixed stride without a division per step.
struct ColWalk {
  int64_t idx, p;
  __device__ __forceinline__ ColWalk(int64_t c, int64_t P) : idx(c / P), p(c % P) {}
  __device__ __forceinline__ void advance(int64_t step, int64_t P) {
    p += step;
    while (p >= P) {
      p -= P;
      ++idx;
    }
  }
  __device__ __forceinline__ const float* ptr(const float* base, int64_t ldc, int64_t n) const {
    return base + idx * ldc + p * n;
  }
};

// RBF contraction mode: instead of storing S = sum_c L_c R_c^T, reduce
//   sum_ij S_ij |x_i - x_j|^2 exp2(c |x_i - x_j|^2)   (l^3 times <S, dK/dl>)
// and sum_i S_ii (<S, I>) per workgroup into red[2 * block + {0, 1}] (FP64).
struct RBFContract {
  const float* X;
  int dim;
  float coef;
  double* red;
};

// The points of rows r0 .. r0 + kOT - 1 as float4 (zeros past n and past dim)
// into LDS, for the contraction epilogues.
__device__ __forceinline__ void stage_points(float4* xs, const RBFContract& rc, int64_t r0,
                                             int64_t n) {
  for (int t = threadIdx.x; t < kOT; t += blockDim.x) {
    float4 v = make_float4(0.f, 0.f, 0.f, 0.f);
    const int64_t i = r0 + t;
    if (i < n) {
      const float* x = rc.X + i * rc.dim;
      v.x = x[0];
      if (rc.dim > 1) v.y = x[1];
      if (rc.dim > 2) v.z = x[2];
      if (rc.dim > 3) v.w = x[3];
    }
    xs[t] = v;
  }
}

// One 32 x 32 accumulator of S in the contraction: the lane holds column
// j0 + jl at rows i0 + ib + 8 (s / 4) + (s % 4), s = 0..15.  Adds
// sum_s S_ij d2 exp2(coef d2) to part and the diagonal entries to trace.
__device__ __forceinline__ void contract_acc(const floatx16& acc, const float4* xi,
                                             const float4* xj, int ib, int jl, int64_t i0,
                                             int64_t j0, int64_t n, float coef, float& part,
                                             double& trace) {
  const int64_t j = j0 + jl;
  if (j >= n) return;
  const float4 pj = xj[jl];
#pragma unroll
  for (int s = 0; s < 16; ++s) {
    const int il = ib + 8 * (s / 4) + (s % 4);
    const int64_t i = i0 + il;
    if (i >= n) continue;
    const float4 pi = xi[il];
    float d2 = 0.f, df;
    df = pi.x - pj.x;
    d2 = fmaf(df, df, d2);
    df = pi.y - pj.y;
    d2 = fmaf(df, df, d2);
    df = pi.z - pj.z;
    d2 = fmaf(df, df, d2);
    df = pi.w - pj.w;
    d2 = fmaf(df, df, d2);
    part = fmaf(acc[s], d2 * __builtin_amdgcn_exp2f(coef * d2), part);
    if (i == j) trace += double(acc[s]);
  }
}

template <bool CONTRACT>
__global__ void __launch_bounds__(256)
    k_outer_sum(const float* __restrict__ L, const float* __restrict__ R, int64_t ldc, int64_t P,
                int64_t n, int64_t K, float* __restrict__ out, RBFContract rc, TileMap tmap) {
  __shared__ __attribute__((aligned(16))) float Ls[2][kOK * kOLd];
  __shared__ __attribute__((aligned(16))) float Rs[2][kOK * kOLd];
  const int t = threadIdx.x;
  const int lane = t & 63, wave = t >> 6, r = lane & 31, h = lane >> 5;
  int64_t ti, tj;
  if (!tile_of_block(tmap, blockIdx.x, ti, tj)) {
    if (CONTRACT && t == 0) {
      rc.red[2 * int64_t(blockIdx.x)] = 0.0;
      rc.red[2 * int64_t(blockIdx.x) + 1] = 0.0;
    }
    return;
  }
  const int64_t i0 = ti * kOT, j0 = tj * kOT;
  // staging: thread t moves float4 (t & 31) of stage column t >> 5
  const int sc = t >> 5, s4 = (t & 31) * 4;
  ColWalk lw(sc, P), rw(sc, P);
  auto load = [&](const float* base, const ColWalk& cw, int64_t c0, int64_t e0) {
    float4 v = make_float4(0.f, 0.f, 0.f, 0.f);
    const int64_t c = c0 + sc, e = e0 + s4;
    if (c < K) {
      const float* src = cw.ptr(base, ldc, n) + e;
      if (e + 4 <= n) {
        __builtin_memcpy(&v, src, sizeof(v));  // dword alignment suffices on gfx950
      } else {
        if (e + 0 < n) v.x = src[0];
        if (e + 1 < n) v.y = src[1];
        if (e + 2 < n) v.z = src[2];
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
  float4 lv = load(L, lw, 0, i0), rv = load(R, rw, 0, j0);
  *reinterpret_cast<float4*>(&Ls[0][sc * kOLd + s4]) = lv;
  *reinterpret_cast<float4*>(&Rs[0][sc * kOLd + s4]) = rv;
  __syncthreads();
  for (int64_t st = 0; st < nst; ++st) {
    const int b = int(st & 1);
    const bool more = st + 1 < nst;
    if (more) {
      lw.advance(kOK, P);
      rw.advance(kOK, P);
      lv = load(L, lw, (st + 1) * kOK, i0);
      rv = load(R, rw, (st + 1) * kOK, j0);
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
    __shared__ float4 xi[kOT], xj[kOT];
    stage_points(xi, rc, i0, n);
    stage_points(xj, rc, j0, n);
    __syncthreads();
    double sd = 0.0, st2 = 0.0;
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      float part = 0.f;
      contract_acc(acc[c], xi, xj, 32 * wave + 4 * h, 32 * c + r, i0, j0, n, rc.coef, part, st2);
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
      const int64_t b = blockIdx.x;
      rc.red[2 * b] = red[0][0];
      rc.red[2 * b + 1] = red[1][0];
    }
  }
}

}  // namespace

hipError_t launch_outer_sum(hipStream_t s, const float* L, const float* R, int64_t ldc, int64_t P,
                            int64_t n, int64_t K, float* out) {
  const TileMap tm = tile_map(n);
  hipLaunchKernelGGL(k_outer_sum<false>, dim3(unsigned(tm.blocks)), dim3(256), 0, s, L, R, ldc, P,
                     n, K, out, RBFContract{}, tm);
  return hipGetLastError();
}

int64_t outer_contract_blocks(int64_t n) { return tile_map(n).blocks; }

hipError_t launch_outer_rbf_contract(hipStream_t s, const float* L, const float* R, int64_t ldc,
                                     int64_t P, int64_t n, int64_t K, const float* X, int dim,
                                     double lengthscale, double* red) {
  const TileMap tm = tile_map(n);
  RBFContract rc{X, dim, float(-1.4426950408889634 / (2.0 * lengthscale * lengthscale)), red};
  hipLaunchKernelGGL(k_outer_sum<true>, dim3(unsigned(tm.blocks)), dim3(256), 0, s, L, R, ldc, P,
                     n, K, nullptr, rc, tm);
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
  const bool vec = P * n >= (int64_t(1) << 20);
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
