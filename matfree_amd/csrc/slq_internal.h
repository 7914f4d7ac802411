// Internal declarations shared by the HIP translation units of libslq.
// Kernels live in k_*.hip; each exposes host launch functions declared here.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>

#include "../../include/slq.h"
#include "slq_check.h"

namespace slq {

// Threads per block of the streaming (HBM-bound) kernels.
constexpr int kStreamThreads = 256;
// Elements of one vector covered by one block of the streaming kernels
// (a multiple of 4 * kStreamThreads; 2048 = two float4 groups per thread).
#ifndef SLQ_STREAM_CHUNK
#define SLQ_STREAM_CHUNK 2048
#endif
constexpr int64_t kChunk = SLQ_STREAM_CHUNK;
// Extra floats between consecutive columns of the internal Krylov basis.
constexpr int64_t kBasisSkew = 1088;
// Basis columns held in registers by the fused project+dots kernel.
constexpr int kFusedMaxCols = 64;

inline int64_t num_chunks(int64_t n) { return (n + kChunk - 1) / kChunk; }

// sum_{b < count} v[b * stride], added in the order b = 0, 1, ... like a
// plain loop, but with the loads issued eight at a time: a serial loop costs
// one memory latency per partial sum, which dominates the prologue of a
// streaming kernel over many chunks.
__device__ __forceinline__ float ordered_sum(const float* __restrict__ v, int64_t count,
                                             int64_t stride) {
  float s = 0.f;
  for (int64_t b = 0; b < count; b += 8) {
    float r[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) r[u] = v[(b + u < count ? b + u : count - 1) * stride];
#pragma unroll
    for (int u = 0; u < 8; ++u)
      if (b + u < count) s += r[u];
  }
  return s;
}

// ---- probes (k_probes.hip) -------------------------------------------------
hipError_t launch_rademacher(hipStream_t s, uint32_t k0, uint32_t k1, int64_t count, float* out);

// BF16 planes of the block matvec's input (tile-major, zero-padded; see
// k_matvec.hip), written by the kernel that produces the input so that the
// matvec can skip its own split pass (kMvPlanesReady).  hi == nullptr: the
// matvec reads no planes for this operator and probe count.
struct PlaneOut {
  __bf16* hi = nullptr;
  __bf16* lo = nullptr;  // null when the matvec does not read the low plane
  int64_t Ppad = 0;
  int perm = 0;  // RBF fragment order of the columns within a tile
};

// A split-K sum left to the consumer: W = sum_{z < S} partial[z * P * n + .]
// (added in the order z = 0 .. S - 1) + noise * V when nugget (the reduction
// of launch_matvec, bit for bit).  partial == nullptr: W was written.
struct MvSplitSum {
  const float* partial = nullptr;
  int S = 1;
  const float* V = nullptr;
  float noise = 0.f;
  int nugget = 0;
};

// ---- streaming Gram-Schmidt kernels (k_lanczos.hip) ------------------------
// Basis column c of probe p starts at Q + c * ldc + p * n (ldc >= P * n).
// partial[p][b][c] = sum_{i in chunk b} Q[c][p][i] * x[p][i], c < ncols.
// Q may be null with ncols = 1 to mean "sum of squares of x" (partial[p][b][0]).
hipError_t launch_dots(hipStream_t s, const float* Q, int64_t ldc, const float* x, int64_t P,
                       int64_t n, int ncols, float* partial, int pstride);
// launch_dots of x = the split-K sum `sum` (a deferred launch_matvec result),
// which is also written to x (Q non-null), any ncols.
hipError_t launch_dots_sum(hipStream_t s, const float* Q, int64_t ldc, const MvSplitSum& sum,
                           float* x, int64_t P, int64_t n, int ncols, float* partial, int pstride);
// y = x - sum_c Q[c] h[c] with h[c] = sum_b hpart[p][b][c] - hsub[p][c];
// optional norm2part[p][b] = sum over chunk of y^2.  hout (if non-null)
// receives sum_b hpart[p][b][c] at hout + p * hout_stride (written by chunk 0
// only); hsub (if non-null) is read at hsub + p * hsub_stride.
hipError_t launch_project(hipStream_t s, const float* Q, int64_t ldc, const float* x, float* y,
                          int64_t P, int64_t n, int ncols, const float* hpart, int pstride,
                          float* norm2part, float* hout, int64_t hout_stride,
                          const float* hsub = nullptr, int64_t hsub_stride = 0);
// Fused: y = x - Q h1 (h1 reduced from h1part), then h2part = Q^T y per chunk.
hipError_t launch_project_dots(hipStream_t s, const float* Q, int64_t ldc, const float* x,
                               float* y, int64_t P, int64_t n, int ncols, const float* h1part,
                               float* h2part, int pstride, float* hout, int64_t hout_stride);
// out[p][i] = x[p][i] / sqrt(sum_b norm2part[p][b]); optionally writes the
// norm to norm_out[p] (float), to hslot[p * hslot_stride] and
// hslot2[p * hslot_stride], and its reciprocal to inv_out[p].
hipError_t launch_normalize(hipStream_t s, const float* x, float* out, int64_t P, int64_t n,
                            const float* norm2part, float* norm_out, float* hslot,
                            int64_t hslot_stride, float* inv_out, float* hslot2 = nullptr,
                            const PlaneOut* planes = nullptr);
// out[0 .. count) = 0 by a kernel (no memset node in a captured stream).
hipError_t launch_fill_zero(hipStream_t s, float* out, int64_t count);
// Plain Lanczos step: r = w - a q - b_prev qm with a = sum_b apart[p][b]
// (single-column dots of q and w), b_prev = bprev[p * bstride] (qm may be
// null), norm^2 partials of r, a written to aout[p * astride].
hipError_t launch_three_term(hipStream_t s, const float* w, const float* q, const float* qm,
                             int64_t P, int64_t n, const float* apart, const float* bprev,
                             int64_t bstride, float* r, float* norm2part, float* aout,
                             int64_t astride);
// out[p][i] = scale[p * sstride] * x[p][i]
hipError_t launch_scale_rows(hipStream_t s, const float* x, const float* scale, int64_t sstride,
                             int64_t P, int64_t n, float* out);

// Plain Lanczos adjoint (decomp.py:268-348), reverse step i.  Per-probe
// scalars: a_i = a[p * sstride], b_i = b[p * bstride], da_i = da[p * sstride],
// db_i = db1[p * sstride] + db2[p * sstride] (db1 null: 0); part: the step's
// partial dots (P x num_chunks(n) x 3: [lambda_plus . x_i, x_{i+1} . xi,
// x_i . xi], null: all zero).
// General cotangents add dx (the cotangent of x_i, subtracted in the xi
// update), separate strides for db1 / db2 and da (dbstride / dastride, 0:
// sstride; db2 may be null) and xplus_res: the xplus of the lambda kernel is the residual
// b x_{i+1} (the last step), so its coefficient is mu / b.
struct LanczosAdjStep {
  const float* a = nullptr;
  const float* b = nullptr;
  int64_t bstride = 0;
  const float* da = nullptr;
  const float* db1 = nullptr;
  const float* db2 = nullptr;
  int64_t sstride = 0;
  const float* part = nullptr;
  const float* dx = nullptr;
  int64_t dbstride = 0;
  int64_t dastride = 0;
  int xplus_res = 0;
};
// Start of the general Lanczos adjoint (decomp.py:290-297): xi = -dxk (dxk
// null: 0) and the partial dots of step k - 1 [0, x_k . xi, x_{k-1} . xi]
// with x_k = R / blast.
hipError_t launch_lanczos_adj_init(hipStream_t s, const float* dxk, const float* R,
                                   const float* blast, const float* xkm1, int64_t P, int64_t n,
                                   float* xi, float* part_out);
// Cotangent of the start vector (decomp.py:305): dv = c ((xi . x0) x0 - xi)
// with c = 1 / |v| per probe and the final xi.
hipError_t launch_lanczos_adj_initvec(hipStream_t s, const float* xi, const float* x0,
                                      const float* c, int64_t P, int64_t n, float* part,
                                      float* dv);
// lambda = -xi / b + mu x_{i+1} + nu x_i (xi null: 0; xplus null: the mu term
// is dropped, valid where mu = 0)
hipError_t launch_lanczos_adj_lambda(hipStream_t s, const float* xi, const float* x,
                                     const float* xplus, int64_t P, int64_t n,
                                     const LanczosAdjStep& st, float* lam);
// xi = -W + a lambda + b lambda_plus - b nu xplus (xplus_is_residual: xplus
// is b x_{i+1} already); with part_out, the partial dots of step i - 1
// [lambda . x_prev, x . xi, x_prev . xi]
hipError_t launch_lanczos_adj_xi(hipStream_t s, const float* W, const float* lam,
                                 const float* lam_plus, const float* xplus, bool xplus_is_residual,
                                 const float* x_prev, const float* x, int64_t P, int64_t n,
                                 const LanczosAdjStep& st, float* xi, float* part_out);

// ---- block matvec (k_matvec.hip) -------------------------------------------
// Largest split-K factor of the block matvec; its partial sums need up to
// kMaxSplitK * P * n floats of scratch (split-K is off when scratch is null;
// matvec_scratch_floats sizes it for the factor the launch will take).
constexpr int kMaxSplitK = 16;
// Scratch (in floats) the block matvec can use for P vectors of length n:
// the split-K partial sums and the BF16 hi/lo planes of the vectors.
int64_t matvec_scratch_floats(int64_t n, int64_t P);
// ev0 / ev1 (optional) are recorded on s right before and after the block
// matvec kernel (the launches around it excluded).
// flags: kMvPlanesReady -- the scratch planes already hold V (written by a
// PlaneOut producer after a call that split them, so the padding is zero);
// kMvPointsReady -- the RBF point records in the scratch are those of op
// (a previous call with the same operator and scratch).
constexpr int kMvPlanesReady = 1;
constexpr int kMvPointsReady = 2;
// With defer non-null a split-K result is not reduced into W; *defer
// describes it instead (W is then not written).
hipError_t launch_matvec(hipStream_t s, const slq_operator& op, const float* V, int64_t P,
                         float* W, float* scratch, int64_t scratch_floats,
                         hipEvent_t ev0 = nullptr, hipEvent_t ev1 = nullptr, int flags = 0,
                         MvSplitSum* defer = nullptr);
// The planes launch_matvec reads for (op, P) in scratch.
PlaneOut matvec_planes(const slq_operator& op, int64_t P, float* scratch);
// Packed storage of a stored matrix (slq_storage): sizes (false for FP32 or
// an unknown storage) and the packing kernel.
bool packed_size(int64_t n, int storage, int64_t* bytes, int64_t* scale_count);
hipError_t launch_pack_matrix(hipStream_t s, const float* A, int64_t n, int storage, void* packed,
                              float* scales);

// ---- tridiagonal eigen + quadrature (k_tridiag.hip) ------------------------
// For each probe p: T = (H + H^T)/2 restricted to its tridiagonal; returns
// per_probe[p] = scale[p]^2 * sum_j U_0j^2 log(lambda_j) in FP64, optional
// alpha/beta extraction, and the mean over probes.  work:
// tridiag_quadrature_work_doubles(P, k) doubles (may be 0).
hipError_t launch_tridiag_quadrature(hipStream_t s, const float* H, const float* scale,
                                     int64_t P, int64_t k, double* work, double* per_probe,
                                     double* mean, float* alpha, float* beta);
int64_t tridiag_quadrature_work_doubles(int64_t P, int64_t k);

// dH (P x k x k, column-major like H, written on the tridiagonal only; the
// caller zeroes it) = cotangent of mean_p |z_p|^2 [log T_p]_00 w.r.t. H, with
// len[p] = |z_p|.  work: tridiag_vjp_work_doubles(P, k) doubles (may be 0).
hipError_t launch_tridiag_vjp(hipStream_t s, const float* H, const float* len, int64_t P, int64_t k,
                              double* work, float* dH);
int64_t tridiag_vjp_work_doubles(int64_t P, int64_t k);
// Combination coefficients (per probe 2 k + 4 floats, k_adjoint_combine's
// layout) of |v| Q^T f(T) e_1 with |v| = len[p]; work as above.
hipError_t launch_tridiag_funm(hipStream_t s, const float* H, const float* len, int64_t P,
                               int64_t k, int f, double* work, float* coef);
// alpha (P x k) and beta (P x max(k, 1)) of plain Lanczos from its T (in an
// H-shaped buffer) and the remainder norms blast (P).
hipError_t launch_lanczos_coeffs(hipStream_t s, const float* T, const float* blast, int64_t P,
                                 int64_t k, float* alpha, float* beta);
// mean / standard error of P doubles (one block, fixed order).
hipError_t launch_mean_and_sem(hipStream_t s, const double* v, int64_t P, double* mean,
                               double* sem);

// ---- reverse pass (k_adjoint.hip) ------------------------------------------
// One backward step idx: Gamma row idx and the combination coefficients
// (coef, per probe 2 k + 4 floats) from the chunk partials of Q^T (A^T lambda).
// Pg / gam (optional): Pi_gamma and gamma of general cotangents (see
// launch_adjoint_pigamma_row / launch_adjoint_gamma_init); null: the SLQ
// composition's dQ = dr = dc = 0 with a tridiagonal dH.
hipError_t launch_adjoint_coef(hipStream_t s, const float* H, const float* dH, float* Gam,
                               const float* part, int pstride, int64_t P, int64_t n, int64_t k,
                               int64_t idx, float* coef, const float* Pg = nullptr,
                               const float* gam = nullptr);
// Row i of Pi_gamma = -dc c e_1 e_1^T + H dH^T - dQ^T Q per probe (P x k x k,
// column-major) from the chunk partials of Q^T dQ_i (part, null: dQ = 0);
// c, dc per probe (dc null: 0).
hipError_t launch_adjoint_pigamma_row(hipStream_t s, const float* H, const float* dH,
                                      const float* part, int pstride, const float* c,
                                      const float* dc, int64_t P, int64_t n, int64_t k, int64_t i,
                                      float* Pg);
// gamma = dH e_K - Q^T dr (part: chunk partials of Q^T dr, null: dr = 0) into
// gam (P x k) and the coefficients of lambda = dr + Q gamma (x = dr).
hipError_t launch_adjoint_gamma_init(hipStream_t s, const float* dH, const float* part, int pstride,
                                     int64_t P, int64_t n, int64_t k, float* gam, float* coef);
// Coefficients for lambda = Q dH[:, k - 1] (the start of the recursion).
hipError_t launch_adjoint_init_coef(hipStream_t s, const float* dH, int64_t P, int64_t k,
                                    float* coef);
// y = scale (cr r + cl x + cv v + d + sum_{j < ncq} gq_j Q_j + sum_{jl0 <= j < k} gl_j Lam_j),
// per probe (coef layout above); r, x, v, d may be null; save (if given) receives x.
hipError_t launch_adjoint_combine(hipStream_t s, const float* Q, const float* Lam, int64_t ldc,
                                  const float* r, const float* x, const float* v, float* y,
                                  float* save, int64_t P, int64_t n, int64_t k, int ncq, int jl0,
                                  const float* coef, const float* d = nullptr);
// out (n x n, row-major) = sum_{c < K} L_c R_c^T over columns c = (idx, p) at
// base + idx * ldc + p * n (K = k P), FP32 MFMA.
hipError_t launch_outer_sum(hipStream_t s, const float* L, const float* R, int64_t ldc, int64_t P,
                            int64_t n, int64_t K, float* out);
// With S = sum_{c < K} L_c R_c^T (as above) and the RBF points X (n x dim):
// red[2 b] = sum_ij S_ij d2_ij exp(-d2_ij / (2 l^2)), d2_ij = |x_i - x_j|^2, and
// red[2 b + 1] = sum_i S_ii, partial over the outer_contract_blocks(n) tiles b.
int64_t outer_contract_blocks(int64_t n);
// FP32 matrix cores (operators whose matvec runs in FP32).
hipError_t launch_outer_rbf_contract(hipStream_t s, const float* L, const float* R, int64_t ldc,
                                     int64_t P, int64_t n, int64_t K, const float* X, int dim,
                                     double lengthscale, double* red);
// The same two sums for an RBF operator with dim <= 3 in split BF16 from the
// block matvec's derivative mode (k_matvec.hip): red receives
// rbf_dk_partials(n, K) pairs (sum_ij S_ij d_ij exp2(d_ij), sum_i S_ii) with
// d_ij = c d2_ij and c = float(-log2(e) / (2 l^2)); scratch holds
// rbf_dk_scratch_floats(n, K) floats (16-byte aligned).
int64_t rbf_dk_partials(int64_t n, int64_t K);
int64_t rbf_dk_scratch_floats(int64_t n, int64_t K);
hipError_t launch_rbf_dk_contract(hipStream_t s, const float* X, int dim, double lengthscale,
                                  const float* L, const float* R, int64_t ldc, int64_t Pin,
                                  int64_t n, int64_t K, float* scratch, double* red);
// out[o] = f_o * sum_i partial[i * nout + o] (fixed order), nout <= 2.
hipError_t launch_sum_fixed(hipStream_t s, const double* partial, int64_t count, int nout,
                            double f0, double f1, double* out);

// ---- GP caller: partial Cholesky, Woodbury solve, CG (k_lowrank.hip) ------
// Partial Cholesky of the operator's matrix (elements A[r][c], or the RBF
// kernel plus noise on the diagonal) to rank `rank`, with or without pivoting.
// Scratch: Lc rank x n floats, perm n int64, bval/bidx ceil(n/256) each,
// rowi rank floats, lii 1 float.  Output: L (n x rank, row-major, original
// row order), perm_out (optional, n int64: position -> row), success (1 if
// every l_ii^2 > 0).
hipError_t launch_cholesky_partial(hipStream_t s, const slq_operator& op, int rank, bool pivot,
                                   float* Lc, int64_t* perm, float* bval, int64_t* bidx,
                                   float* rowi, float* lii, int* success, float* L,
                                   int64_t* perm_out);
// Woodbury factor of s I + L L^T: C = chol(I + L^T L / s) (r x r FP64, lower,
// row-major).  Gp: gram_splits(n) * r * r floats of scratch; ok = 0 if C is
// not positive definite.
int64_t gram_splits(int64_t n);
hipError_t launch_woodbury_setup(hipStream_t s, const float* L, int64_t n, int r, double shift,
                                 float* Gp, double* C, int* ok);
// out = (s I + L L^T)^{-1} V for the P rows of V.  Wp: woodbury_chunks(n) *
// P * r floats, Z: P * r doubles of scratch.
int64_t woodbury_chunks(int64_t n);
hipError_t launch_woodbury_apply(hipStream_t s, const float* L, int64_t n, int r, double shift,
                                 const double* C, const float* V, int64_t P, float* Wp, double* Z,
                                 float* out);
// CG over P columns: per-column scalars and the preconditioner's buffers.
// gamma and active hold two generations of P entries each: iteration t reads
// generation t % 2 and writes generation (t + 1) % 2, so that every block of
// a kernel can read the old values while one block writes the new ones.
struct CgScalars {
  float* gamma = nullptr;   // 2 P
  float* atol2 = nullptr;   // P
  int64_t* iters = nullptr; // P
  int* active = nullptr;    // 2 P
};
struct CgPrecond {
  const float* L = nullptr;
  int rank = 0;
  double shift = 1.0;
  const double* C = nullptr;
  float* Wp = nullptr;
  double* Z = nullptr;
};
// partial dots of the CG kernels: P x cg_chunks(n) floats
int64_t cg_chunks(int64_t n);
hipError_t launch_cg_dot(hipStream_t s, const float* a, const float* b, int64_t P, int64_t n,
                         float* part);
// generation 0: gamma = sum rz, atol2 = max(tol^2 sum bb, atol^2), rs = sum
// rr (or gamma when rr is null), active = rs > atol2 && maxiter > 0;
// *nactive = count
hipError_t launch_cg_start(hipStream_t s, const float* rz, const float* bb, const float* rr,
                           int64_t P, int64_t n, float tol, float atol, int64_t maxiter,
                           const CgScalars& sc, int* nactive);
// Iteration t (gen = t % 2) after the matvec of d: q = A d (the deferred
// split-K sum `qsum` is added up and written to q when its partial is
// non-null), alpha, x and r, z = M r (precond), gamma, beta, the stopping
// rule and d = z + beta d for the columns active in generation gen.
hipError_t launch_cg_step(hipStream_t s, float* x, float* r, float* d, float* q,
                          const MvSplitSum& qsum, float* z, bool precond, int64_t P, int64_t n,
                          int64_t maxiter, const CgScalars& sc, int gen, float* pq, float* rr,
                          float* rz, const CgPrecond* M);
// *nactive = number of columns active in generation gen
hipError_t launch_cg_count(hipStream_t s, const CgScalars& sc, int gen, int64_t P, int* nactive);

}  // namespace slq
