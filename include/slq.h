/*
 * slq.h -- C ABI of the MI355X (gfx950) stochastic Lanczos quadrature library.
 *
 * The library replaces the device side of matfree's SLQ log-determinant path
 * (Rademacher probes -> decomp.tridiag_sym(reortho="full") ->
 * funm.monte_carlo_funm_sym_logdet -> stochtrace.estimator_monte_carlo).
 * In matfree that path is pure JAX; the functions below are what a JAX FFI
 * custom call (or any other host) binds in its place.
 *
 * Conventions
 *  - Every array argument is a DEVICE pointer on the context's device,
 *    caller-owned, row-major, float32 unless stated otherwise.
 *  - Every call is asynchronous and ordered on the context's stream.
 *  - The return value is an slq_status; on failure slq_last_error() gives a
 *    message.  Argument errors are detected on the host before any launch.
 *  - Scratch memory (partial sums, work vectors, the Krylov basis of
 *    slq_logdet) is owned and pooled by the context.  A context is used by
 *    one host thread at a time.
 *  - Layouts: a batch of P vectors of length n is P x n (vector p at
 *    p * n).  A Krylov basis is k x P x n (vector j of probe p at
 *    (j * P + p) * n).  A Hessenberg matrix batch is P x k x k stored column
 *    by column (H[p][col][row] at (p * k + col) * k + row), the order in
 *    which the Arnoldi iteration writes it.
 */
#ifndef SLQ_H_
#define SLQ_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum slq_status {
  SLQ_OK = 0,
  SLQ_EINVAL = -1,       /* bad argument (null pointer, bad enum, bad size)      */
  SLQ_EINVAL_K = -2,     /* num_matvecs outside [0, n]: matfree's "exceeds" error */
  SLQ_ENOMEM = -3,       /* device allocation failed                              */
  SLQ_EHIP = -4,         /* HIP runtime error                                     */
  SLQ_EUNSUPPORTED = -5, /* valid request this build does not implement           */
  SLQ_EADJOINT_K0 = -6   /* adjoint requested for num_matvecs = 0                 */
} slq_status;

typedef enum slq_operator_kind {
  SLQ_OP_DENSE = 0, /* stored symmetric matrix A (n x n): v -> A v             */
  SLQ_OP_RBF = 1    /* K_ij = exp(-|x_i - x_j|^2 / (2 l^2)) + s2 delta_ij      */
} slq_operator_kind;

typedef enum slq_precision {
  SLQ_PREC_FP32 = 0,   /* FP32 MFMA, FP32 operands                                */
  SLQ_PREC_BF16 = 1,   /* one BF16 MFMA product, FP32 accumulation                 */
  SLQ_PREC_BF16X3 = 2, /* split hi+lo BF16: hi*hi + hi*lo + lo*hi, FP32 accumulation */
  SLQ_PREC_BF16X2 = 3  /* matrix split hi+lo, vectors BF16: hi*hi + lo*hi           */
} slq_precision;

/* Storage of a DENSE operator's matrix.  SLQ_STORE_FP32 is the plain n x n
 * FP32 matrix (`matrix`, products in the slq_precision `precision`).  The
 * other forms are packed once by slq_pack_matrix into `packed` (+
 * `tile_scales`); their block matvec runs BF16 MFMAs on the stored planes
 * with the vectors rounded to BF16 and FP32 accumulation (`precision` is
 * ignored), reading 4 (BF16X2), 2 (BF16, FP8X2) or 1 (FP8) bytes per matrix
 * element from HBM instead of 4 with an FP32 matrix.
 *   BF16X2: A = hi + lo, hi = bf16(A), lo = bf16(A - hi) (~16 significant bits);
 *   BF16:   hi only (8 bits);
 *   FP8X2:  A = s_hi q_hi + s_lo q_lo, q = OCP FP8 E4M3 (round to nearest
 *           even), one power-of-two scale per 32 x 32 tile and plane
 *           (max |A_tile| / s <= 448), q_lo from the exact residual A - s_hi q_hi;
 *   FP8:    s q only (4 significant bits).
 * Packed layout: the matrix zero-padded to npad = 32 ceil(n / 32) in both
 * dimensions, cut into 32 x 32 tiles, tile (rb, ct) (rows 32 rb.., columns
 * 32 ct..) at index rb * (npad / 32) + ct; planes one after the other (hi,
 * then lo), each npad^2 elements; within a tile the elements are in the
 * order the MFMA fragments read them (lane l = r + 32 h, r < 32, h < 2, holds
 * row 32 rb + r, columns 32 ct + {8 h .. 8 h + 7} and 32 ct + {16 + 8 h ..
 * 23 + 8 h}): BF16 planes as [half (first / second column group)][lane][8],
 * FP8 planes as [lane][16] (the first column group in bytes 0-7).
 * tile_scales: per FP8 plane (npad / 32)^2 floats, tile index as above. */
typedef enum slq_storage {
  SLQ_STORE_FP32 = 0,
  SLQ_STORE_BF16X2 = 1,
  SLQ_STORE_BF16 = 2,
  SLQ_STORE_FP8X2 = 3,
  SLQ_STORE_FP8 = 4
} slq_storage;

/* Operator descriptor: the reference's matvec(v, *params) plugin
 * (decomp.py:460, user code such as tests/test_decomp/test_tridiag_sym.py:16-18
 * and tutorials/8_gaussian_logpdf.py:21-31), restricted to the two operator
 * families this library runs on the GPU. */
typedef struct slq_operator {
  int32_t kind;        /* slq_operator_kind                                   */
  int32_t precision;   /* slq_precision of the block matvec                    */
  int64_t n;           /* dimension                                            */
  const float* matrix; /* DENSE, SLQ_STORE_FP32: n x n symmetric, row-major    */
  const float* points; /* RBF: n x dim points                                  */
  int32_t dim;         /* RBF: point dimension, 1..4                           */
  int32_t storage;     /* DENSE: slq_storage (0 for RBF)                       */
  double lengthscale;  /* RBF: l > 0                                          */
  double noise;        /* RBF: s2 (added on the diagonal in FP32)             */
  const void* packed;  /* DENSE, packed storage: the planes (see slq_storage)  */
  const float* tile_scales; /* DENSE, FP8 storage: the per-tile scales        */
} slq_operator;

/* Sizes of the packed form of an n x n matrix in `storage` (not FP32):
 * *packed_bytes for `packed`, *scale_count floats for `tile_scales` (0 for
 * the BF16 forms).  Host only. */
int slq_packed_size(int64_t n, int storage, int64_t* packed_bytes, int64_t* scale_count);

/* Context (device, stream, pooled scratch). stream is a hipStream_t or NULL. */
typedef struct slq_context slq_context;
int slq_context_create(int device, void* stream, slq_context** out);
int slq_context_destroy(slq_context* ctx);
int slq_context_set_stream(slq_context* ctx, void* stream);
const char* slq_last_error(const slq_context* ctx);
/* Bytes of device scratch the context currently holds. */
int64_t slq_context_scratch_bytes(const slq_context* ctx);
/* Synchronise the context's stream and free all of its pooled scratch (the
 * next call allocates again).  No counterpart in the reference (JAX owns its
 * buffers): for a host that admits work by device-memory footprint and must
 * give an idle context's memory back. */
int slq_context_release(slq_context* ctx);

/* Pack the n x n FP32 matrix A (device, row-major) into `storage` (see
 * slq_storage; buffers of the sizes slq_packed_size gives, device).  No
 * counterpart in the reference (its operator is a user function): the
 * low-precision stored-matrix variant of the operator plugin of
 * decomp.py:460. */
int slq_pack_matrix(slq_context* ctx, const float* A, int64_t n, int storage, void* packed,
                    float* tile_scales);

/* Rademacher probes, bit-identical to jax.random.rademacher(key, (num, n),
 * float32) with the partitionable threefry2x32 layout.
 * Replaces stochtrace.sampler_signs -> _sampler_from_jax_random ->
 * _uniform_signs -> prng.rademacher (stochtrace.py:932-937, 957-964, 974-977;
 * backend/prng.py:26-29).  key = {key[0], key[1]} as in jax.random.PRNGKey
 * (backend/prng.py:6-7).  out: num x n. */
int slq_sampler_signs(slq_context* ctx, uint32_t key0, uint32_t key1, int64_t num, int64_t n,
                      float* out);

/* Block matvec W = A V over P vectors (rows of V): the reference's matvec
 * (decomp.py:460) batched by vmap (stochtrace.py:49).  V, W: P x n. */
int slq_matvec(slq_context* ctx, const slq_operator* op, const float* V, int64_t P, float* W);

/* Arnoldi with full reorthogonalisation (classical Gram-Schmidt twice) for P
 * start vectors: decomp.hessenberg(k, reortho="full") forward pass,
 * decomp.py:426-477.
 *   V0: P x n start vectors (normalised internally, as decomp.py:435, 456)
 *   Q:  k x P x n Krylov basis (output)
 *   H:  P x k x k Hessenberg matrices, column-major per probe (output)
 *   R:  P x n unnormalised residuals (output)
 *   c:  P values 1 / |v0| (output; decomp.py:450)
 * Fails with SLQ_EINVAL_K unless 0 <= k <= n (decomp.py:427-429). */
int slq_hessenberg(slq_context* ctx, const slq_operator* op, const float* V0, int64_t P, int64_t k,
                   float* Q, float* H, float* R, float* c);

/* slq_hessenberg with the forward's reorthogonalisation selectable:
 * reortho = 0 runs one classical Gram-Schmidt pass per step (the forward of
 * decomp.hessenberg(..., reortho_vjp="none"), decomp.py:394-396, 466-468),
 * reortho != 0 two (as slq_hessenberg). */
int slq_arnoldi(slq_context* ctx, const slq_operator* op, const float* V0, int64_t P, int64_t k,
                int reortho, float* Q, float* H, float* R, float* c);

/* Reverse mode of the Arnoldi decomposition for general output cotangents:
 * the custom VJP of decomp.hessenberg (decomp.py:398-422) and
 * _hessenberg_adjoint (decomp.py:480-600).  The forward pass on V0 (as
 * slq_arnoldi with `reortho`) is recomputed, then the adjoint runs with its
 * reorthogonalisation on (reortho_adjoint != 0, the reference's
 * reortho="full") or off.  Cotangents (device, any may be NULL = zero):
 *   dQ: k x P x n (the layout of slq_hessenberg's Q), dH: P x k x k
 *   (column-major like H), dr: P x n, dc: P.
 * Outputs: dv (P x n, cotangent of V0 = c lambda, decomp.py:602; may be
 * NULL), and the parameter cotangents as in slq_logdet_vjp (dparams for
 * RBF, dmatrix for stored matrices).  Fails with SLQ_EADJOINT_K0 for k = 0. */
int slq_arnoldi_vjp(slq_context* ctx, const slq_operator* op, const float* V0, int64_t P,
                    int64_t k, int reortho, int reortho_adjoint, const float* dQ, const float* dH,
                    const float* dr, const float* dc, float* dv, double* dparams, float* dmatrix);

/* Reverse mode of plain Lanczos (tridiag_sym(k, reortho="none")) for general
 * cotangents: the custom VJP of decomp.py:181-197 and _tridiag_adjoint
 * (decomp.py:268-348).  The forward (slq_lanczos) is recomputed.
 * Cotangents (device, any may be NULL = zero): dQ (k x P x n, the vectors
 * x_0 .. x_{k-1}), dalpha (P x k), dbeta (P x k: the k - 1 off-diagonals,
 * then the remainder's norm b_k), dxk (P x n, the remainder's direction x_k,
 * residual = b_k x_k).  Outputs: dv (P x n, cotangent of V0, decomp.py:305;
 * may be NULL) and the parameter cotangents as in slq_logdet_vjp. */
int slq_lanczos_vjp(slq_context* ctx, const slq_operator* op, const float* V0, int64_t P,
                    int64_t k, const float* dQ, const float* dalpha, const float* dbeta,
                    const float* dxk, float* dv, double* dparams, float* dmatrix);

/* SLQ log-determinant integrand and estimator for P probes z_p:
 *   per_probe[p] = |z_p|^2 [log T_p]_00 with T_p = (H_p + H_p^T)/2 from
 *   slq_hessenberg on z_p / |z_p|, evaluated in FP64 (eigenvalues and first
 *   eigenvector components of the tridiagonal T_p);
 *   mean = (1/P) sum_p per_probe[p].
 * Replaces funm.monte_carlo_funm_sym_logdet -> monte_carlo_funm_sym.quadform
 * (funm.py:186-243), dense_funm_sym_eigh(log) (funm.py:322-335) and
 * stochtrace.estimator_monte_carlo (stochtrace.py:47-50).
 *   Z: P x n probes; per_probe: P doubles; mean: 1 double (either may be
 *   NULL); alpha: P x k and beta: P x (k-1) tridiagonal coefficients (may be
 *   NULL).  The Krylov basis lives in context scratch (k x P x n floats). */
int slq_logdet(slq_context* ctx, const slq_operator* op, const float* Z, int64_t P, int64_t k,
               double* per_probe, double* mean, float* alpha, float* beta);

/* slq_logdet without any device allocation.  No counterpart in the
 * reference: for callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, hosts that forbid it).  Same arguments,
 * checks and results as slq_logdet (the same kernels in the same order), but
 * every scratch buffer is carved from a static device array of
 * SLQ_NOALLOC_SCRATCH_BYTES bytes per GPU (held by each GPU the library runs
 * on until the process ends) instead of the context's pool; the call makes no
 * hipMalloc/hipFree.  A call needing more scratch than that (roughly
 * 4 (k + 6) P n bytes) fails with SLQ_EINVAL before anything is queued.  The
 * array is located when the first context on a device is created.  All calls
 * on one GPU share its array: two calls whose work overlaps on the device
 * (different streams, or different contexts) race on it; serialise them. */
#define SLQ_NOALLOC_SCRATCH_BYTES (256ull << 20)
int slq_logdet_noalloc(slq_context* ctx, const slq_operator* op, const float* Z, int64_t P,
                       int64_t k, double* per_probe, double* mean, float* alpha, float* beta);

/* Value and parameter cotangent of the SLQ log-determinant estimate
 *   mean = (1/P) sum_p |z_p|^2 [log T_p]_00
 * for P probes (reverse mode through the estimator): the custom VJP of
 * decomp.hessenberg (decomp.py:398-422), _hessenberg_adjoint with
 * reortho="full" (decomp.py:480-600) driven by dH alone (the quadratic form
 * discards Q, r and c, funm.py:237), the eigh VJP at [log T]_00
 * (funm.py:331-333) and the probe mean (stochtrace.py:47-50).
 *   RBF operators: dparams[0] = d mean / d lengthscale,
 *                  dparams[1] = d mean / d noise (two device doubles);
 *   stored matrices: dmatrix = d mean / d A (n x n device floats, row-major,
 *                  the matvec VJP lambda q^T of decomp.py:587-589 summed).
 * per_probe and mean as in slq_logdet (either may be NULL).  The forward
 * pass is recomputed and its basis kept in context scratch; the transposed
 * product A^T lambda uses A itself (the operator of tridiag_sym is
 * symmetric).  Fails with SLQ_EADJOINT_K0 for k = 0 (decomp.py:483-486). */
int slq_logdet_vjp(slq_context* ctx, const slq_operator* op, const float* Z, int64_t P, int64_t k,
                   double* per_probe, double* mean, double* dparams, float* dmatrix);

/* Saved forward state: the residuals of the custom VJP (decomp.py:398-401:
 * basis, Hessenberg or tridiagonal matrix, residual) kept on the device
 * between the forward and the reverse pass, so that the reverse pass need
 * not recompute the forward (jax.custom_vjp stores them the same way).
 *
 * slq_logdet_save runs slq_logdet (reortho_full != 0) or slq_logdet_lanczos
 * (reortho_full = 0, its k vectors kept) and moves the buffers the reverse
 * pass reads out of the context's pool into *out; per_probe and mean as in
 * slq_logdet.  slq_logdet_vjp_saved then computes exactly what
 * slq_logdet_vjp / slq_logdet_lanczos_vjp computes for the same operator
 * and probes (without the forward pass), on the context's stream, any
 * number of times; op must be the operator of the save (same n, kind and
 * parameters).  slq_saved_destroy synchronises ctx's stream (the device
 * when ctx is NULL) and frees the state.  Errors as the corresponding
 * slq_logdet_vjp call. */
typedef struct slq_saved slq_saved;
int slq_logdet_save(slq_context* ctx, const slq_operator* op, const float* Z, int64_t P, int64_t k,
                    int reortho_full, double* per_probe, double* mean, slq_saved** out);
int slq_logdet_vjp_saved(slq_context* ctx, const slq_operator* op, const slq_saved* saved,
                         double* dparams, float* dmatrix);
int64_t slq_saved_bytes(const slq_saved* saved);
int slq_saved_destroy(slq_context* ctx, slq_saved* saved);

/* Device scratch (bytes) that one call of an entry point needs for (op, P,
 * k), the caller's inputs and outputs excluded: what a fresh context grows
 * to (up to 256-byte alignment per buffer).  No counterpart in the
 * reference; for admission by device-memory footprint. */
typedef enum slq_entry {
  SLQ_ENTRY_LOGDET = 0,             /* slq_logdet             */
  SLQ_ENTRY_LOGDET_LANCZOS = 1,     /* slq_logdet_lanczos     */
  SLQ_ENTRY_LOGDET_VJP = 2,         /* slq_logdet_vjp         */
  SLQ_ENTRY_LOGDET_LANCZOS_VJP = 3  /* slq_logdet_lanczos_vjp */
} slq_entry;
int slq_workspace_bytes(slq_context* ctx, int entry, const slq_operator* op, int64_t P, int64_t k,
                        int64_t* bytes);
/* The bytes by which ctx's pooled scratch grows in one call of entry for
 * (op, P, k), given what the pool holds now (a pool keeps each buffer at the
 * largest size any call needed): after the call slq_context_scratch_bytes(ctx)
 * is at most its value before plus this (exactly, up to 256-byte alignment
 * per grown buffer).  No counterpart in the reference; for admission by
 * footprint when a request reuses a context that already holds scratch. */
int slq_workspace_growth(slq_context* ctx, int entry, const slq_operator* op, int64_t P,
                         int64_t k, int64_t* bytes);

/* Plain Lanczos (tridiag_sym(k, reortho="none"), decomp.py:155-266) for P
 * start vectors: v0 = V0 / |V0|; per step a = q . A q, r = A q - a q - b q_prev,
 * b = |r|, q_next = r / b -- no reorthogonalisation, one matvec per step.
 *   Q:     k x P x n Lanczos vectors q_0 .. q_{k-1} (may be NULL: then only
 *          three vectors per probe are kept);
 *   alpha: P x k diagonal of T;
 *   beta:  P x max(k, 1): the k - 1 off-diagonals of T, then the norm b of the
 *          remainder (decomp.py:227-229);
 *   R:     P x n residual b q_k (decomp.py:167);
 *   c:     P values 1 / |V0| (decomp.py:171-178; unlike the "full" variant).
 * k = 0 runs the initial step only (decomp.py:215-218).  Fails with
 * SLQ_EINVAL_K unless 0 <= k <= n. */
int slq_lanczos(slq_context* ctx, const slq_operator* op, const float* V0, int64_t P, int64_t k,
                float* Q, float* alpha, float* beta, float* R, float* c);

/* slq_logdet with tridiag_sym(k, reortho="none") (plain Lanczos, no basis
 * kept: O(n) memory per probe); same outputs. */
int slq_logdet_lanczos(slq_context* ctx, const slq_operator* op, const float* Z, int64_t P,
                       int64_t k, double* per_probe, double* mean, float* alpha, float* beta);

/* slq_logdet_vjp with tridiag_sym(k, reortho="none"): the custom VJP of the
 * plain Lanczos decomposition (decomp.py:181-197), _tridiag_adjoint and
 * _tridiag_adjoint_step (decomp.py:268-348) driven by the cotangents of the
 * diagonals alone (zero for the vectors, the residual and the remainder's
 * norm: funm.py:237 keeps only T), the eigh VJP at [log T]_00
 * (funm.py:331-333) and the probe mean (stochtrace.py:47-50).  dparams and
 * dmatrix as in slq_logdet_vjp (dmatrix: sum of the matvec VJPs x_i lambda_i^T
 * of decomp.py:337-338).  The forward pass is recomputed with its k vectors
 * kept in context scratch.  Fails with SLQ_EINVAL for k = 0 (the integrand
 * has no tridiagonal) and SLQ_EINVAL_K unless k <= n. */
int slq_logdet_lanczos_vjp(slq_context* ctx, const slq_operator* op, const float* Z, int64_t P,
                           int64_t k, double* per_probe, double* mean, double* dparams,
                           float* dmatrix);

/* Matrix functions of slq_funm_vectors (dense_funm_sym_eigh(matfun),
 * funm.py:322-335, for these matfun). */
typedef enum slq_matfun {
  SLQ_FUNM_LOG = 0,
  SLQ_FUNM_EXP = 1,
  SLQ_FUNM_SIN = 2,
  SLQ_FUNM_COS = 3,
  SLQ_FUNM_SQRT = 4,
  SLQ_FUNM_INV = 5 /* 1 / x */
} slq_matfun;

/* Matrix-function-vector products f(A) v_p ~ |v_p| Q_p^T f(T_p) e_1 for P
 * vectors: funm.funm_lanczos_sym(dense_funm_sym_eigh(f), tridiag_sym(k))
 * (funm.py:114-147) with the forward pass of slq_hessenberg (reortho_full
 * != 0) or of slq_lanczos (reortho_full = 0) and f applied to the FP64
 * eigendecomposition of each T_p.
 *   V: P x n vectors; out: P x n results (may alias neither V nor scratch).
 * Fails with SLQ_EINVAL_K unless 0 < k <= n. */
int slq_funm_vectors(slq_context* ctx, const slq_operator* op, const float* V, int64_t P,
                     int64_t k, int reortho_full, int matfun, float* out);

/* mean and standard error of the mean (population standard deviation over
 * sqrt(P)) of P device doubles, in FP64 and a fixed order:
 * estimator_monte_carlo_mean_and_sem's reductions (stochtrace.py:84-88).
 * Either output may be NULL. */
int slq_mean_and_sem(slq_context* ctx, const double* values, int64_t P, double* mean,
                     double* sem);

/* ---- the GP log-density caller (tutorials/8_gaussian_logpdf.py:38-52) ------ */

/* Partial Cholesky factorisation of the operator's matrix, the element
 * function mat_el(i, j, *params) of low_rank.cholesky_partial
 * (low_rank.py:56-99, pivot = 0) and low_rank.cholesky_partial_pivot
 * (low_rank.py:102-198, pivot = 1): A[i][j] for a stored matrix, the RBF
 * kernel exp(-|x_i - x_j|^2 / (2 l^2)) plus noise on the diagonal.
 *   L:       n x rank device floats, row-major (the reference's (nrows, rank)
 *            result, rows in the original order);
 *   perm:    n device int64 (may be NULL): the pivot order, position -> row;
 *   success: one device int, 1 if every l_ii^2 was positive (the pivoted
 *            variant's info["success"]).
 * Fails with SLQ_EINVAL ("Rank exceeds nrows" / "Rank must be positive",
 * low_rank.py:112-117) unless 1 <= rank <= n, SLQ_EUNSUPPORTED above 4096. */
int slq_cholesky_partial(slq_context* ctx, const slq_operator* op, int64_t rank, int pivot,
                         float* L, int64_t* perm, int* success);

/* The preconditioner's solve (low_rank.py:33-51):
 * out = (s I + L L^T)^{-1} V for the P rows of V (P x n), by the Woodbury
 * identity with the capacitance I + L^T L / s factored in FP64 (factored on
 * every call, as the reference does). */
int slq_lowrank_solve(slq_context* ctx, const float* L, int64_t n, int64_t rank, double shift,
                      const float* V, int64_t P, float* out);

/* Conjugate gradients for A X = B over the P rows of B (backend/linalg.py:
 * 107-108, jax.scipy.sparse.linalg.cg with x0 = 0 under vmap): stop a row
 * when rs <= max(tol^2 |b|^2, atol^2) (rs = r . z, or r . r with a
 * preconditioner) or after maxiter steps; converged rows keep their state.
 * With L non-NULL the preconditioner is (s I + L L^T)^{-1} (rank columns of
 * L, n x rank row-major).  iterations (P device int64, may be NULL) receives
 * the steps each row took.  The host reads one int (the number of rows still
 * active) after batches of 1, 2, 4, 8, 8, ... iterations; iterations run
 * after the last row stopped change nothing. */
int slq_cg(slq_context* ctx, const slq_operator* op, const float* B, int64_t P, double tol,
           double atol, int64_t maxiter, const float* L, int64_t rank, double shift, float* X,
           int64_t* iterations);

/* Cotangents of sum_p x_p^T A(theta) x_p w.r.t. the operator parameters (the
 * matvec VJP with cotangent x at x; what the derivative of the quadratic form
 * y^T A^{-1} y needs, as -(these) at x = A^{-1} y): RBF: dparams[0] =
 * sum x^T dK/dl x, dparams[1] = sum x^T x; stored matrix: dmatrix = sum x x^T. */
int slq_quadform_vjp(slq_context* ctx, const slq_operator* op, const float* X, int64_t P,
                     double* dparams, float* dmatrix);

/* ---- measurement hooks (no counterpart in the reference) -------------------
 * With profiling on, every launch of the selected kernel classes made through
 * this context is bracketed by a pair of HIP events on the context's stream
 * (each event record costs device time too: select only the classes that are
 * read).  slq_profile_enable(ctx, on) selects all classes (on != 0) or none;
 * slq_profile_select(ctx, mask) selects the classes kc with bit (1 << kc) of
 * mask set.  slq_profile_read synchronises the stream, returns the summed
 * device time (ms) and launch count of one class since the last read, and
 * resets it. */
typedef enum slq_kernel_class {
  SLQ_KC_MATVEC = 0,       /* block matvec                              */
  SLQ_KC_DOTS = 1,         /* Gram-Schmidt coefficients Q^T w            */
  SLQ_KC_PROJECT_DOTS = 2, /* fused w - Q h and Q^T (w - Q h)            */
  SLQ_KC_PROJECT = 3,      /* w - Q h (+ norm partials)                  */
  SLQ_KC_NORMALIZE = 4,    /* v / |v| and basis store                    */
  SLQ_KC_QUADRATURE = 5,   /* FP64 tridiagonal eigen + quadrature        */
  SLQ_KC_ADJOINT = 6,      /* whole reverse pass of slq_logdet(_lanczos)_vjp */
  SLQ_KC_COUNT = 7
} slq_kernel_class;
int slq_profile_enable(slq_context* ctx, int on);
int slq_profile_select(slq_context* ctx, int mask);
int slq_profile_read(slq_context* ctx, int kernel_class, double* total_ms, int64_t* launches);

#ifdef __cplusplus
}
#endif

#endif /* SLQ_H_ */
