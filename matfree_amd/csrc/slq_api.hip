// C ABI of libslq: argument checking, scratch pooling and the launch sequence
// of the SLQ log-determinant path.  Kernels live in k_*.hip.
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "slq_internal.h"

struct slq_context {
  int device = 0;
  hipStream_t stream = nullptr;
  std::string error;
  struct Buf {
    void* ptr = nullptr;
    size_t bytes = 0;
  };
  // pooled scratch, grown on demand (never shrunk until destroy)
  Buf partial1, partial2, norm2, work_w, work_w2, basis, hmat, resid, vec0, scal1, scal2, eig,
      per_probe, mean, splitk;
  // reverse pass
  Buf adj_dh, adj_gam, adj_coef, adj_lam, adj_v1, adj_v2, adj_v3, adj_work, adj_sum, adj_dk;
  // plain Lanczos: three rotating vectors and the last off-diagonal
  Buf ring, blast;
  // general-cotangent reverse passes: Pi_gamma, gamma, the internal dH
  Buf adj_pg, adj_gvec;
  // partial Cholesky, Woodbury solve and CG
  Buf lr_lc, lr_perm, lr_bval, lr_bidx, lr_rowi, lr_small, lr_gp, lr_c, lr_wp, lr_z, cg_r, cg_z,
      cg_d, cg_q, cg_parts, cg_scal;
  std::vector<Buf*> all_bufs() {
    return {&partial1, &partial2, &norm2,   &work_w,  &work_w2,  &basis,   &hmat,
            &resid,    &vec0,     &scal1,   &scal2,   &eig,      &per_probe, &mean,
            &splitk,   &adj_dh,   &adj_gam, &adj_coef, &adj_lam, &adj_v1,  &adj_v2,
            &adj_v3,   &adj_work, &adj_sum, &adj_dk,  &ring,    &blast,    &adj_pg,  &adj_gvec,
            &lr_lc,    &lr_perm,
            &lr_bval,  &lr_bidx,  &lr_rowi, &lr_small, &lr_gp,   &lr_c,    &lr_wp,
            &lr_z,     &cg_r,     &cg_z,    &cg_d,    &cg_q,     &cg_parts, &cg_scal};
  }
  // pinned host word for the CG loop's count of active columns
  int* host_count = nullptr;              // pinned, 16 ints (CG's polled counts)
  hipEvent_t cg_done[2] = {nullptr, nullptr};  // CG: count copied to host_count[i]
  // Arena mode: ensure() carves buffers out of [arena, arena + arena_size)
  // instead of the pool (nothing is allocated).  With dry_run set the arena
  // is only counted (its base is a placeholder) and nothing is launched:
  // arena_off then says how many bytes the call needs.
  bool arena_mode = false;
  bool dry_run = false;
  char* arena = nullptr;
  size_t arena_size = 0, arena_off = 0;
  // arena mode: bytes of the buffers that were replaced by larger ones (a
  // dry run seeded with a pool's sizes: the pool frees them when it grows)
  size_t arena_replaced = 0;
  // measurement hooks: event pairs per kernel class, reused across reads
  unsigned profile_mask = 0;  // bit kc: kernel class kc is bracketed by events
  std::vector<std::pair<hipEvent_t, hipEvent_t>> events[SLQ_KC_COUNT];
  size_t used[SLQ_KC_COUNT] = {};
};

// Device state of a forward solve kept for a later reverse pass (the
// residuals of the custom VJP): buffers moved out of a context's pool.
struct slq_saved {
  int device = 0;
  int full = 1;  // reortho="full" (Arnoldi) or "none" (plain Lanczos)
  int32_t kind = 0;
  int64_t n = 0, P = 0, k = 0, ldc = 0;
  slq_context::Buf basis, hmat, resid, scal1, scal2, blast;
  std::vector<slq_context::Buf*> all_bufs() {
    return {&basis, &hmat, &resid, &scal1, &scal2, &blast};
  }
};

// Static scratch of slq_logdet_noalloc: a file-scope device array, one copy
// per GPU the library runs on, held until the process ends.  On the host the
// name `scratch` only stands for the symbol (it is not a device address and is
// never passed to a kernel or a copy): the device address of each GPU's copy
// is looked up once per device with hipGetSymbolAddress, with that device
// current, when the first context on the device is created, and that address
// is what the entry point hands to its kernels.
static __device__ __align__(256) char scratch[SLQ_NOALLOC_SCRATCH_BYTES];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mutex;
struct StaticScratch {
  bool looked_up = false;
  hipError_t status = hipSuccess;
  void* ptr = nullptr;  // device address of this GPU's copy of `scratch`
};
StaticScratch g_scratch[kMaxDevices];

// Called with `device` current: the lookup happens once per device.
void setup_static_scratch(int device) {
  if (device < 0 || device >= kMaxDevices) return;
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  StaticScratch& st = g_scratch[device];
  if (st.looked_up) return;
  void* p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess) {
    (void)hipGetLastError();
    p = nullptr;
  }
  st.looked_up = true;
  st.status = e;
  st.ptr = p;
}
}  // namespace

namespace slq {
namespace {

int fail(slq_context* ctx, int code, const std::string& msg) {
  if (ctx) ctx->error = msg;
  return code;
}

int hip_fail(slq_context* ctx, hipError_t err, const char* where) {
  std::string msg = std::string(where) + ": " + hipGetErrorString(err);
  return fail(ctx, SLQ_EHIP, msg);
}

// A HIP call or launch, skipped in a dry run (which only sizes scratch).
#define SLQ_HIP(ctx, expr)                                       \
  do {                                                           \
    if (!(ctx)->dry_run) {                                       \
      hipError_t _e = (expr);                                    \
      if (_e != hipSuccess) return hip_fail((ctx), _e, #expr);   \
    }                                                            \
  } while (0)

// Grow a pooled buffer to at least `bytes`.  The stream is synchronised
// before an old allocation is released so in-flight work never sees it freed.
int ensure(slq_context* ctx, slq_context::Buf& b, size_t bytes, void** out) {
  if (bytes == 0) bytes = 16;
  if (ctx->arena_mode) {
    if (b.bytes < bytes) {
      ctx->arena_replaced += b.bytes;
      const size_t off = (ctx->arena_off + 255) / 256 * 256;
      if (off + bytes > ctx->arena_size)
        return fail(ctx, SLQ_ENOMEM,
                    "the call needs more than the " + std::to_string(ctx->arena_size) +
                        " bytes of scratch available to it");
      b.ptr = ctx->arena + off;
      b.bytes = bytes;
      ctx->arena_off = off + bytes;
    }
    *out = b.ptr;
    return SLQ_OK;
  }
  if (b.bytes < bytes) {
    if (b.ptr) {
      SLQ_HIP(ctx, hipStreamSynchronize(ctx->stream));
      SLQ_HIP(ctx, hipFree(b.ptr));
      b.ptr = nullptr;
      b.bytes = 0;
    }
    void* p = nullptr;
    hipError_t e = hipMalloc(&p, bytes);
    if (e != hipSuccess) {
      (void)hipGetLastError();
      return fail(ctx, SLQ_ENOMEM,
                  "device allocation of " + std::to_string(bytes) + " bytes failed: " +
                      hipGetErrorString(e));
    }
    b.ptr = p;
    b.bytes = bytes;
  }
  *out = b.ptr;
  return SLQ_OK;
}

template <typename T>
int ensure_t(slq_context* ctx, slq_context::Buf& b, size_t count, T** out) {
  void* p = nullptr;
  int rc = ensure(ctx, b, count * sizeof(T), &p);
  *out = static_cast<T*>(p);
  return rc;
}

#define SLQ_TRY(expr)              \
  do {                             \
    int _rc = (expr);              \
    if (_rc != SLQ_OK) return _rc; \
  } while (0)

std::string exceeds_msg(int64_t k, int64_t n) {
  // Same wording as matfree's decomp._error_num_matvecs (decomp.py:753-756).
  return "Parameter 'num_matvecs'=" + std::to_string(k) +
         " exceeds the acceptable range. Expected: 0 <= num_matvecs <= " + std::to_string(n) + ".";
}

int check_operator(slq_context* ctx, const slq_operator* op) {
  std::string msg;
  const int rc = operator_status(op, &msg);
  return rc == SLQ_OK ? SLQ_OK : fail(ctx, rc, msg);
}

// Sizes whose products the library forms (limits in slq_check.h).
int check_footprint(slq_context* ctx, int64_t P, int64_t n, int64_t k) {
  std::string msg;
  const int rc = footprint_status(P, n, k, &msg);
  return rc == SLQ_OK ? SLQ_OK : fail(ctx, rc, msg);
}

// The next (start, end) event pair of kernel class kc, or nulls when the
// class is not profiled (or an event cannot be created).
std::pair<hipEvent_t, hipEvent_t> take_events(slq_context* ctx, int kc) {
  if (!((ctx->profile_mask >> kc) & 1u)) return {nullptr, nullptr};
  auto& vec = ctx->events[kc];
  size_t& u = ctx->used[kc];
  if (u == vec.size()) {
    hipEvent_t a = nullptr, b = nullptr;
    if (hipEventCreate(&a) != hipSuccess || hipEventCreate(&b) != hipSuccess) {
      (void)hipGetLastError();
      return {nullptr, nullptr};
    }
    vec.emplace_back(a, b);
  }
  return vec[u++];
}

// Brackets the launches of its scope with the next event pair of class kc.
struct Timed {
  slq_context* ctx;
  hipEvent_t end = nullptr;
  Timed(slq_context* c, int kc) : ctx(c) {
    const auto ev = take_events(ctx, kc);
    if (ev.first) (void)hipEventRecord(ev.first, ctx->stream);
    end = ev.second;
  }
  ~Timed() {
    if (end) (void)hipEventRecord(end, ctx->stream);
  }
};

int set_device(slq_context* ctx) {
  SLQ_HIP(ctx, hipSetDevice(ctx->device));
  return SLQ_OK;
}

// Column stride of the internal Krylov basis: P * n rounded up to 64 floats
// plus a 1088-float skew.
inline int64_t basis_stride(int64_t P, int64_t n) {
  return (P * n + 63) / 64 * 64 + kBasisSkew;
}

// Forward Arnoldi (decomp.py:426-477) for P start vectors; see slq.h.
// reortho = false: one classical Gram-Schmidt pass per step (the forward of
// hessenberg(..., reortho_vjp="none"), decomp.py:466-468).
int hessenberg_core(slq_context* ctx, const slq_operator& op, const float* V0, int64_t P,
                    int64_t k, float* Q, int64_t ldc, float* H, float* R, float* c,
                    bool reortho = true) {
  const int64_t n = op.n;
  hipStream_t s = ctx->stream;
  const int64_t nb = num_chunks(n);
  const int pstride = int(k > 0 ? k : 1);
  float *part1, *part2, *norm2, *w, *w2, *mvs;
  const int64_t mv_floats = matvec_scratch_floats(n, P);
  SLQ_TRY(ensure_t(ctx, ctx->splitk, size_t(mv_floats), &mvs));
  SLQ_TRY(ensure_t(ctx, ctx->partial1, size_t(P * nb * pstride), &part1));
  SLQ_TRY(ensure_t(ctx, ctx->partial2, size_t(P * nb * pstride), &part2));
  SLQ_TRY(ensure_t(ctx, ctx->norm2, size_t(P * nb), &norm2));
  SLQ_TRY(ensure_t(ctx, ctx->work_w, size_t(P * n), &w));
  SLQ_TRY(ensure_t(ctx, ctx->work_w2, size_t(P * n), &w2));

  if (k > 0) SLQ_HIP(ctx, launch_fill_zero(s, H, P * k * k));

  // initlength = sqrt(<v, v>) (decomp.py:435); c = 1 / initlength (:450)
  SLQ_HIP(ctx, launch_dots(s, nullptr, 0, V0, P, n, 1, norm2, 1));
  if (k == 0) {
    SLQ_HIP(ctx, launch_normalize(s, V0, w, P, n, norm2, nullptr, nullptr, 0, c));
    SLQ_HIP(ctx, hipMemcpyAsync(R, V0, size_t(P * n) * sizeof(float), hipMemcpyDeviceToDevice, s));
    return SLQ_OK;
  }
  // step 0: v /= length; Q[:, 0] = v (decomp.py:456-457)
  SLQ_HIP(ctx, launch_normalize(s, V0, Q, P, n, norm2, nullptr, nullptr, 0, c));
  // after the first matvec (which splits its input and prepares the point
  // records) the normalisation writes the next input's BF16 planes itself
  const PlaneOut planes = matvec_planes(op, P, mvs);
  MvSplitSum pend;
  for (int64_t j = 0; j < k; ++j) {
    const float* qj = Q + j * ldc;
    const int ncols = int(j + 1);
    float* hcol = H + j * k;  // column j of every probe, stride k * k
    // w = A v (decomp.py:460); the events bracket the block-matvec kernel
    // itself (not the probe split or the split-K reduction around it).  A
    // split-K result is summed into w by the first dots pass (pend).
    {
      const auto ev = take_events(ctx, SLQ_KC_MATVEC);
      const int flags = j == 0 ? 0 : (kMvPointsReady | (planes.hi ? kMvPlanesReady : 0));
      SLQ_HIP(ctx, launch_matvec(s, op, qj, P, w, mvs, mv_floats, ev.first, ev.second, flags,
                                 &pend));
    }
    // h = Q^T w (decomp.py:463)
    {
      Timed tm(ctx, SLQ_KC_DOTS);
      if (pend.partial)
        SLQ_HIP(ctx, launch_dots_sum(s, Q, ldc, pend, w, P, n, ncols, part1, pstride));
      else
        SLQ_HIP(ctx, launch_dots(s, Q, ldc, w, P, n, ncols, part1, pstride));
    }
    const bool last = (j + 1 == k);
    float* dst = last ? R : w;
    if (!reortho) {
      // w = w - Q h and <w, w> (decomp.py:464, 471), h into H[:, j]
      {
        Timed tm(ctx, SLQ_KC_PROJECT);
        SLQ_HIP(ctx, launch_project(s, Q, ldc, w, dst, P, n, ncols, part1, pstride, norm2, hcol,
                                    k * k));
      }
      if (!last) {
        Timed tm(ctx, SLQ_KC_NORMALIZE);
        SLQ_HIP(ctx, launch_normalize(s, w, Q + (j + 1) * ldc, P, n, norm2, nullptr, hcol + j + 1,
                                      k * k, nullptr, nullptr, planes.hi ? &planes : nullptr));
      }
      continue;
    }
    // w = w - Q h (decomp.py:464), then the second pass's Q^T w (decomp.py:468)
    if (ncols <= kFusedMaxCols) {
      Timed tm(ctx, SLQ_KC_PROJECT_DOTS);
      SLQ_HIP(ctx, launch_project_dots(s, Q, ldc, w, w2, P, n, ncols, part1, part2, pstride, hcol,
                                        k * k));
    } else {
      SLQ_HIP(ctx, launch_project(s, Q, ldc, w, w2, P, n, ncols, part1, pstride, nullptr, hcol, k * k));
      SLQ_HIP(ctx, launch_dots(s, Q, ldc, w2, P, n, ncols, part2, pstride));
    }
    // w = w - Q (Q^T w) and <w, w> (decomp.py:468, 471)
    {
      Timed tm(ctx, SLQ_KC_PROJECT);
      SLQ_HIP(ctx, launch_project(s, Q, ldc, w2, dst, P, n, ncols, part2, pstride, norm2, nullptr, 0));
    }
    if (!last) {
      // length = sqrt(<w, w>); h[j+1] = length (decomp.py:471-475); next
      // step's v /= length; Q[:, j+1] = v (decomp.py:456-457)
      Timed tm(ctx, SLQ_KC_NORMALIZE);
      SLQ_HIP(ctx, launch_normalize(s, w, Q + (j + 1) * ldc, P, n, norm2, nullptr, hcol + j + 1,
                                    k * k, nullptr, nullptr, planes.hi ? &planes : nullptr));
    }
  }
  return SLQ_OK;
}

// Plain Lanczos (tridiag_sym(reortho="none"), decomp.py:198-266) for P start
// vectors.  T_p goes into H (P x k x k, column-major: alpha on the diagonal,
// beta on both off-diagonals), the last off-diagonal (the remainder's norm)
// into blast[p], R = blast * q_k, c = 1 / |v0|.  With Q non-null the k
// vectors q_0 .. q_{k-1} are kept at Q + j * ldc; otherwise three rotating
// vectors are used (no basis).  k = 0 runs the initial step only.
int lanczos_core(slq_context* ctx, const slq_operator& op, const float* V0, int64_t P, int64_t k,
                 float* Q, int64_t ldc, float* H, float* blast, float* R, float* c) {
  const int64_t n = op.n;
  hipStream_t s = ctx->stream;
  const int64_t nb = num_chunks(n);
  float *apart, *norm2, *w, *r, *ring, *mvs;
  const int64_t mv_floats = matvec_scratch_floats(n, P);
  SLQ_TRY(ensure_t(ctx, ctx->splitk, size_t(mv_floats), &mvs));
  SLQ_TRY(ensure_t(ctx, ctx->partial1, size_t(P * nb), &apart));
  SLQ_TRY(ensure_t(ctx, ctx->norm2, size_t(P * nb), &norm2));
  SLQ_TRY(ensure_t(ctx, ctx->work_w, size_t(P * n), &w));
  SLQ_TRY(ensure_t(ctx, ctx->work_w2, size_t(P * n), &r));
  SLQ_TRY(ensure_t(ctx, ctx->ring, size_t(3 * P * n), &ring));
  if (k > 0) SLQ_HIP(ctx, launch_fill_zero(s, H, P * k * k));
  // vector j: basis column j when kept (j < k), else ring slot j % 3
  auto vec = [&](int64_t j) -> float* {
    if (Q && j < k) return Q + j * ldc;
    return ring + (j % 3) * P * n;
  };
  // v0 = vec / |vec| (decomp.py:208-209); c = 1 / |vec| (:178)
  SLQ_HIP(ctx, launch_dots(s, nullptr, 0, V0, P, n, 1, norm2, 1));
  SLQ_HIP(ctx, launch_normalize(s, V0, vec(0), P, n, norm2, nullptr, nullptr, 0, c));
  const int64_t steps = k > 0 ? k : 1;
  const PlaneOut planes = matvec_planes(op, P, mvs);
  MvSplitSum pend;
  for (int64_t j = 0; j < steps; ++j) {
    const float* q = vec(j);
    const float* qm = j > 0 ? vec(j - 1) : nullptr;
    {
      const auto ev = take_events(ctx, SLQ_KC_MATVEC);
      const int flags = j == 0 ? 0 : (kMvPointsReady | (planes.hi ? kMvPlanesReady : 0));
      SLQ_HIP(ctx, launch_matvec(s, op, q, P, w, mvs, mv_floats, ev.first, ev.second, flags,
                                 &pend));
    }
    // a = q . A q; r = A q - a q - b q_prev (decomp.py:243-245, 254-260); a
    // split-K matvec result is summed into w by the dot kernel
    {
      Timed tm(ctx, SLQ_KC_DOTS);
      if (pend.partial)
        SLQ_HIP(ctx, launch_dots_sum(s, q, 0, pend, w, P, n, 1, apart, 1));
      else
        SLQ_HIP(ctx, launch_dots(s, q, 0, w, P, n, 1, apart, 1));
    }
    {
      Timed tm(ctx, SLQ_KC_PROJECT);
      SLQ_HIP(ctx, launch_three_term(s, w, q, qm, P, n, apart,
                                     j > 0 ? H + (j - 1) * k + j : nullptr, k * k, r, norm2,
                                     j < k ? H + j * k + j : nullptr, k * k));
    }
    // b = |r|; x = r / b (:246-247, 261-262); b is T's off-diagonal except
    // after the last step, where it is the remainder's norm
    {
      Timed tm(ctx, SLQ_KC_NORMALIZE);
      const bool inner = j + 1 < k;
      SLQ_HIP(ctx, launch_normalize(s, r, vec(j + 1), P, n, norm2, inner ? nullptr : blast,
                                    inner ? H + j * k + j + 1 : nullptr, k * k, nullptr,
                                    inner ? H + (j + 1) * k + j : nullptr,
                                    planes.hi ? &planes : nullptr));
    }
  }
  // residual = b * x (decomp.py:167)
  SLQ_HIP(ctx, launch_scale_rows(s, vec(steps), blast, 1, P, n, R));
  return SLQ_OK;
}

// Device state of a forward solve that the reverse pass reads.
struct ForwardState {
  float *Q = nullptr, *H = nullptr, *R = nullptr, *c = nullptr, *ell = nullptr;
  float* blast = nullptr;  // plain Lanczos: the remainder's norm per probe
  int64_t ldc = 0;
};

// The SLQ forward pass (funm.py:226-241 for all probes at once), basis kept.
int logdet_forward(slq_context* ctx, const slq_operator& op, const float* Z, int64_t P, int64_t k,
                   double* per_probe, double* mean, float* alpha, float* beta, ForwardState& st) {
  const int64_t n = op.n;
  hipStream_t s = ctx->stream;
  float *v0, *norm2;
  double *eig, *pp;
  SLQ_TRY(ensure_t(ctx, ctx->vec0, size_t(P * n), &v0));
  // basis columns padded so that consecutive columns do not sit a power of
  // two apart (all columns of a step are streamed together)
  st.ldc = basis_stride(P, n);
  SLQ_TRY(ensure_t(ctx, ctx->basis, size_t(k * st.ldc), &st.Q));
  SLQ_TRY(ensure_t(ctx, ctx->hmat, size_t(P * k * k), &st.H));
  SLQ_TRY(ensure_t(ctx, ctx->resid, size_t(P * n), &st.R));
  SLQ_TRY(ensure_t(ctx, ctx->scal1, size_t(P), &st.c));
  SLQ_TRY(ensure_t(ctx, ctx->scal2, size_t(P), &st.ell));
  SLQ_TRY(ensure_t(ctx, ctx->norm2, size_t(P * num_chunks(n)), &norm2));
  SLQ_TRY(ensure_t(ctx, ctx->eig, size_t(tridiag_quadrature_work_doubles(P, k)), &eig));
  if (per_probe) {
    pp = per_probe;
  } else {
    SLQ_TRY(ensure_t(ctx, ctx->per_probe, size_t(P), &pp));
  }
  // length = |z|; v0 = z / length (funm.py:228-229)
  SLQ_HIP(ctx, launch_dots(s, nullptr, 0, Z, P, n, 1, norm2, 1));
  SLQ_HIP(ctx, launch_normalize(s, Z, v0, P, n, norm2, st.ell, nullptr, 0, nullptr));
  // tridiag_sym(matvec, v0) (funm.py:237 -> decomp.py:130-143)
  SLQ_TRY(hessenberg_core(ctx, op, v0, P, k, st.Q, st.ldc, st.H, st.R, st.c));
  // |z|^2 [log T]_00 (funm.py:239-241) and the probe mean (stochtrace.py:50)
  {
    Timed tm(ctx, SLQ_KC_QUADRATURE);
    SLQ_HIP(ctx, launch_tridiag_quadrature(s, st.H, st.ell, P, k, eig, pp, mean, alpha, beta));
  }
  return SLQ_OK;
}

// The SLQ forward pass on plain Lanczos (tridiag_sym(reortho="none"),
// funm.py:226-241 for all probes); with keep_basis the vectors
// q_0 .. q_{k-1} stay in st.Q (basis layout) for the reverse pass, otherwise
// only three rotating vectors are used.
int logdet_lanczos_forward(slq_context* ctx, const slq_operator& op, const float* Z, int64_t P,
                           int64_t k, double* per_probe, double* mean, float* alpha, float* beta,
                           bool keep_basis, ForwardState& st) {
  const int64_t n = op.n;
  hipStream_t s = ctx->stream;
  float *v0, *norm2;
  double *eig, *pp;
  SLQ_TRY(ensure_t(ctx, ctx->vec0, size_t(P * n), &v0));
  if (keep_basis) {
    st.ldc = basis_stride(P, n);
    SLQ_TRY(ensure_t(ctx, ctx->basis, size_t(k * st.ldc), &st.Q));
  }
  SLQ_TRY(ensure_t(ctx, ctx->hmat, size_t(P * k * k), &st.H));
  SLQ_TRY(ensure_t(ctx, ctx->resid, size_t(P * n), &st.R));
  SLQ_TRY(ensure_t(ctx, ctx->scal1, size_t(P), &st.c));
  SLQ_TRY(ensure_t(ctx, ctx->scal2, size_t(P), &st.ell));
  SLQ_TRY(ensure_t(ctx, ctx->blast, size_t(P), &st.blast));
  SLQ_TRY(ensure_t(ctx, ctx->norm2, size_t(P * num_chunks(n)), &norm2));
  SLQ_TRY(ensure_t(ctx, ctx->eig, size_t(tridiag_quadrature_work_doubles(P, k)), &eig));
  if (per_probe) {
    pp = per_probe;
  } else {
    SLQ_TRY(ensure_t(ctx, ctx->per_probe, size_t(P), &pp));
  }
  // length = |z|; v0 = z / length (funm.py:228-229)
  SLQ_HIP(ctx, launch_dots(s, nullptr, 0, Z, P, n, 1, norm2, 1));
  SLQ_HIP(ctx, launch_normalize(s, Z, v0, P, n, norm2, st.ell, nullptr, 0, nullptr));
  SLQ_TRY(lanczos_core(ctx, op, v0, P, k, keep_basis ? st.Q : nullptr, st.ldc, st.H, st.blast,
                       st.R, st.c));
  {
    Timed tm(ctx, SLQ_KC_QUADRATURE);
    SLQ_HIP(ctx, launch_tridiag_quadrature(s, st.H, st.ell, P, k, eig, pp, mean, alpha, beta));
  }
  return SLQ_OK;
}

// RBF parameter cotangents from S = sum_{c < K} L_c R_c^T (columns c at
// base + (c / P) ldc + (c % P) n): dparams = (<S, dK/dl>, <S, I>), the VJPs of
// the matvec in (lengthscale, noise).  FP32 operators: the FP32 matrix-core
// contraction over the tiles of S; BF16 operators: the derivative kernel
// d exp2(d) = c l^3 dK/dl (c = -log2(e) / (2 l^2)) applied to the columns of R
// as a block matvec in split BF16 (hi*hi + hi*lo + lo*hi), dotted with L.
int rbf_param_contract(slq_context* ctx, const slq_operator& op, const float* L, const float* R,
                       int64_t ldc, int64_t P, int64_t K, double* dparams) {
  const int64_t n = op.n;
  hipStream_t s = ctx->stream;
  const double ls = op.lengthscale;
  double* dsum;
  if (op.precision == SLQ_PREC_FP32) {
    const int64_t cnt = outer_contract_blocks(n);
    SLQ_TRY(ensure_t(ctx, ctx->adj_sum, size_t(2 * cnt), &dsum));
    SLQ_HIP(ctx, launch_outer_rbf_contract(s, L, R, ldc, P, n, K, op.points, op.dim, ls, dsum));
    SLQ_HIP(ctx, launch_sum_fixed(s, dsum, cnt, 2, 1.0 / (ls * ls * ls), 1.0, dparams));
  } else {
    const int64_t cnt = rbf_dk_partials(n, K);
    float* dk;
    SLQ_TRY(ensure_t(ctx, ctx->adj_sum, size_t(2 * cnt), &dsum));
    SLQ_TRY(ensure_t(ctx, ctx->adj_dk, size_t(rbf_dk_scratch_floats(n, K)), &dk));
    SLQ_HIP(ctx, launch_rbf_dk_contract(s, op.points, op.dim, ls, L, R, ldc, P, n, K, dk, dsum));
    const double c = double(float(-1.4426950408889634 / (2.0 * ls * ls)));
    SLQ_HIP(ctx, launch_sum_fixed(s, dsum, cnt, 2, 1.0 / (ls * ls * ls * c), 1.0, dparams));
  }
  return SLQ_OK;
}

// The reverse pass: cotangent of the probe mean w.r.t. H (eigh VJP), the
// Arnoldi adjoint recursion (decomp.py:480-600, reortho="full"), and the
// operator parameters' cotangents (the matvec VJPs of decomp.py:587-589).
int logdet_reverse(slq_context* ctx, const slq_operator& op, int64_t P, int64_t k,
                   const ForwardState& st, double* dparams, float* dmatrix) {
  const int64_t n = op.n, ldc = st.ldc;
  hipStream_t s = ctx->stream;
  const int pstride = int(k);
  const int64_t nb = num_chunks(n);
  float *dH, *Gam, *coef, *Lam, *lam, *lam2, *vm, *part1, *part2, *mvs;
  double* vwork = nullptr;
  const int64_t mv_floats = matvec_scratch_floats(n, P);
  SLQ_TRY(ensure_t(ctx, ctx->splitk, size_t(mv_floats), &mvs));
  SLQ_TRY(ensure_t(ctx, ctx->partial1, size_t(P * nb * pstride), &part1));
  SLQ_TRY(ensure_t(ctx, ctx->partial2, size_t(P * nb * pstride), &part2));
  SLQ_TRY(ensure_t(ctx, ctx->adj_dh, size_t(P * k * k), &dH));
  SLQ_TRY(ensure_t(ctx, ctx->adj_gam, size_t(P * k * k), &Gam));
  SLQ_TRY(ensure_t(ctx, ctx->adj_coef, size_t(P * (2 * k + 4)), &coef));
  SLQ_TRY(ensure_t(ctx, ctx->adj_lam, size_t(k * ldc), &Lam));
  SLQ_TRY(ensure_t(ctx, ctx->adj_v1, size_t(P * n), &lam));
  SLQ_TRY(ensure_t(ctx, ctx->adj_v2, size_t(P * n), &lam2));
  SLQ_TRY(ensure_t(ctx, ctx->adj_v3, size_t(P * n), &vm));
  const int64_t vw = tridiag_vjp_work_doubles(P, k);
  if (vw > 0) SLQ_TRY(ensure_t(ctx, ctx->adj_work, size_t(vw), &vwork));

  // dH from the eigh VJP of |z|^2 [log T]_00 / P (funm.py:239-241, 331-333)
  SLQ_HIP(ctx, launch_fill_zero(s, dH, P * k * k));
  SLQ_HIP(ctx, launch_fill_zero(s, Gam, P * k * k));
  SLQ_HIP(ctx, launch_tridiag_vjp(s, st.H, st.ell, P, k, vwork, dH));

  // lambda = dr + Q gamma, gamma = dH e_K - Q^T dr with dr = 0 (decomp.py:498-499)
  SLQ_HIP(ctx, launch_adjoint_init_coef(s, dH, P, k, coef));
  SLQ_HIP(ctx, launch_adjoint_combine(s, st.Q, nullptr, ldc, nullptr, nullptr, nullptr, lam,
                                      nullptr, P, n, k, int(k), int(k), coef));
  // reverse scan (decomp.py:538, 575-599)
  for (int64_t idx = k - 1; idx >= 0; --idx) {
    // lambda -= Q_m (Q_m^T lambda - dH[:, idx]_m), columns j <= idx + 1 (decomp.py:576-584)
    const int ncols = int(idx + 2 < k ? idx + 2 : k);
    SLQ_HIP(ctx, launch_dots(s, st.Q, ldc, lam, P, n, ncols, part1, pstride));
    SLQ_HIP(ctx, launch_project(s, st.Q, ldc, lam, lam2, P, n, ncols, part1, pstride, nullptr,
                                nullptr, 0, dH + idx * k, k * k));
    // A^T lambda (decomp.py:587-588; A symmetric)
    SLQ_HIP(ctx, launch_matvec(s, op, lam2, P, vm, mvs, mv_floats));
    // Q^T (A^T lambda) for Gamma[idx, :] (decomp.py:592-593)
    SLQ_HIP(ctx, launch_dots(s, st.Q, ldc, vm, P, n, int(idx + 1), part2, pstride));
    SLQ_HIP(ctx, launch_adjoint_coef(s, st.H, dH, Gam, part2, pstride, P, n, k, idx, coef));
    // Lambda[:, idx] = lambda and the next lambda (decomp.py:596-599)
    SLQ_HIP(ctx, launch_adjoint_combine(s, st.Q, Lam, ldc, st.R, lam2, vm, lam, Lam + idx * ldc,
                                        P, n, k, int(k), int(idx + 1), coef));
  }

  // parameter cotangents: sum over steps and probes of the matvec VJPs
  // lambda^T (dA) q_idx with lambda = Lambda[:, idx] (decomp.py:587-589)
  if (op.kind == SLQ_OP_RBF) {
    // with S = sum_{idx, p} lambda q^T: d/d noise = <S, I> and
    // d/d lengthscale = <S, dK/dl> = <S, |x_i - x_j|^2 K> / l^3, in one pass
    // over the output tiles of S (never stored)
    SLQ_TRY(rbf_param_contract(ctx, op, Lam, st.Q, ldc, P, k * P, dparams));
  } else {
    // d/dA = sum_idx sum_p lambda q^T (the VJP of A -> A q)
    SLQ_HIP(ctx, launch_outer_sum(s, Lam, st.Q, ldc, P, n, k * P, dmatrix));
  }
  return SLQ_OK;
}

// Parameter cotangents of a reverse pass: S = sum over steps and probes of
// L_c R_c^T (the matvec VJPs), contracted for the RBF parameters or stored
// as the matrix cotangent.
int param_cotangents(slq_context* ctx, const slq_operator& op, const float* L, const float* R,
                     int64_t ldc, int64_t P, int64_t k, double* dparams, float* dmatrix) {
  const int64_t n = op.n;
  hipStream_t s = ctx->stream;
  if (op.kind == SLQ_OP_RBF) {
    SLQ_TRY(rbf_param_contract(ctx, op, L, R, ldc, P, k * P, dparams));
  } else {
    SLQ_HIP(ctx, launch_outer_sum(s, L, R, ldc, P, n, k * P, dmatrix));
  }
  return SLQ_OK;
}

// The Arnoldi adjoint for general cotangents (decomp.py:480-600): dQ (k x P
// x n, column j of probe p at dQ + j P n + p n), dH (P x k x k column-major),
// dr (P x n), dc (P); any may be null (zero).  adj_reortho: the
// reorthogonalisation of the adjoint (decomp.py:575-584, reortho="full").
// Writes dv = c lambda (P x n, decomp.py:602) when dv is non-null and the
// parameter cotangents.
int arnoldi_reverse(slq_context* ctx, const slq_operator& op, int64_t P, int64_t k,
                    const ForwardState& st, bool adj_reortho, const float* dQ, const float* dH_in,
                    const float* dr, const float* dc, float* dv, double* dparams,
                    float* dmatrix) {
  const int64_t n = op.n, ldc = st.ldc, ldq = P * n;
  hipStream_t s = ctx->stream;
  const int pstride = int(k);
  const int64_t nb = num_chunks(n);
  float *dH, *Gam, *coef, *Lam, *lam, *lam2, *vm, *part1, *part2, *mvs, *Pg, *gam;
  const int64_t mv_floats = matvec_scratch_floats(n, P);
  SLQ_TRY(ensure_t(ctx, ctx->splitk, size_t(mv_floats), &mvs));
  SLQ_TRY(ensure_t(ctx, ctx->partial1, size_t(P * nb * pstride), &part1));
  SLQ_TRY(ensure_t(ctx, ctx->partial2, size_t(P * nb * pstride), &part2));
  SLQ_TRY(ensure_t(ctx, ctx->adj_dh, size_t(P * k * k), &dH));
  SLQ_TRY(ensure_t(ctx, ctx->adj_gam, size_t(P * k * k), &Gam));
  SLQ_TRY(ensure_t(ctx, ctx->adj_pg, size_t(P * k * k), &Pg));
  SLQ_TRY(ensure_t(ctx, ctx->adj_gvec, size_t(P * k), &gam));
  SLQ_TRY(ensure_t(ctx, ctx->adj_coef, size_t(P * (2 * k + 4)), &coef));
  SLQ_TRY(ensure_t(ctx, ctx->adj_lam, size_t(k * ldc), &Lam));
  SLQ_TRY(ensure_t(ctx, ctx->adj_v1, size_t(P * n), &lam));
  SLQ_TRY(ensure_t(ctx, ctx->adj_v2, size_t(P * n), &lam2));
  SLQ_TRY(ensure_t(ctx, ctx->adj_v3, size_t(P * n), &vm));

  if (dH_in)
    SLQ_HIP(ctx, hipMemcpyAsync(dH, dH_in, size_t(P * k * k) * sizeof(float),
                                hipMemcpyDeviceToDevice, s));
  else
    SLQ_HIP(ctx, launch_fill_zero(s, dH, P * k * k));
  SLQ_HIP(ctx, launch_fill_zero(s, Gam, P * k * k));
  // Pi_gamma = -dc c e_1 e_1^T + H dH^T - dQ^T Q, row by row (decomp.py:506)
  for (int64_t i = 0; i < k; ++i) {
    if (dQ) SLQ_HIP(ctx, launch_dots(s, st.Q, ldc, dQ + i * ldq, P, n, int(k), part1, pstride));
    SLQ_HIP(ctx, launch_adjoint_pigamma_row(s, st.H, dH, dQ ? part1 : nullptr, pstride, st.c, dc,
                                            P, n, k, i, Pg));
  }
  // gamma = dH e_K - Q^T dr; lambda = dr + Q gamma (decomp.py:498-499)
  if (dr) SLQ_HIP(ctx, launch_dots(s, st.Q, ldc, dr, P, n, int(k), part1, pstride));
  SLQ_HIP(ctx, launch_adjoint_gamma_init(s, dH, dr ? part1 : nullptr, pstride, P, n, k, gam,
                                         coef));
  SLQ_HIP(ctx, launch_adjoint_combine(s, st.Q, nullptr, ldc, nullptr, dr, nullptr, lam, nullptr,
                                      P, n, k, int(k), int(k), coef));
  // reverse scan (decomp.py:538, 575-599)
  for (int64_t idx = k - 1; idx >= 0; --idx) {
    const float* lin = lam;
    if (adj_reortho) {
      // lambda -= Q_m (Q_m^T lambda - dH[:, idx]_m), columns j <= idx + 1
      const int ncols = int(idx + 2 < k ? idx + 2 : k);
      SLQ_HIP(ctx, launch_dots(s, st.Q, ldc, lam, P, n, ncols, part1, pstride));
      SLQ_HIP(ctx, launch_project(s, st.Q, ldc, lam, lam2, P, n, ncols, part1, pstride, nullptr,
                                  nullptr, 0, dH + idx * k, k * k));
      lin = lam2;
    }
    // A^T lambda (decomp.py:587-588; A symmetric)
    SLQ_HIP(ctx, launch_matvec(s, op, lin, P, vm, mvs, mv_floats));
    // Q^T (A^T lambda) for Gamma[idx, :] (decomp.py:592-593)
    SLQ_HIP(ctx, launch_dots(s, st.Q, ldc, vm, P, n, int(idx + 1), part2, pstride));
    SLQ_HIP(ctx, launch_adjoint_coef(s, st.H, dH, Gam, part2, pstride, P, n, k, idx, coef, Pg,
                                     gam));
    // Lambda[:, idx] = lambda and the next lambda, with Pi_xi[idx] = dQ_idx +
    // gamma_idx r (decomp.py:505, 596-599)
    SLQ_HIP(ctx, launch_adjoint_combine(s, st.Q, Lam, ldc, st.R, lin, vm, adj_reortho ? lam : lam2,
                                        Lam + idx * ldc, P, n, k, int(k), int(idx + 1), coef,
                                        dQ ? dQ + idx * ldq : nullptr));
    if (!adj_reortho) std::swap(lam, lam2);
  }
  // dv = lambda c (decomp.py:602)
  if (dv) SLQ_HIP(ctx, launch_scale_rows(s, lam, st.c, 1, P, n, dv));
  return param_cotangents(ctx, op, Lam, st.Q, ldc, P, k, dparams, dmatrix);
}

// The plain-Lanczos adjoint for general cotangents (decomp.py:181-197,
// 268-348): dQ (k x P x n: the vectors x_0 .. x_{k-1}), dalpha (P x k),
// dbeta (P x k: the k - 1 off-diagonals, then the remainder's norm), dxk (P x
// n: the remainder's direction x_k); any may be null (zero).  Writes dv (P x
// n, decomp.py:305) when non-null and the parameter cotangents.
int lanczos_reverse(slq_context* ctx, const slq_operator& op, int64_t P, int64_t k,
                    const ForwardState& st, const float* dQ, const float* dalpha,
                    const float* dbeta, const float* dxk, float* dv, double* dparams,
                    float* dmatrix) {
  const int64_t n = op.n, ldc = st.ldc, kk = k * k, ldq = P * n;
  hipStream_t s = ctx->stream;
  const int64_t nb = num_chunks(n);
  float *db, *Lam, *xi, *W, *part1, *part2, *mvs;
  const int64_t mv_floats = matvec_scratch_floats(n, P);
  SLQ_TRY(ensure_t(ctx, ctx->splitk, size_t(mv_floats), &mvs));
  SLQ_TRY(ensure_t(ctx, ctx->partial1, size_t(P * nb * 3), &part1));
  SLQ_TRY(ensure_t(ctx, ctx->partial2, size_t(P * nb * 3), &part2));
  SLQ_TRY(ensure_t(ctx, ctx->adj_gvec, size_t(2 * P * k), &db));
  SLQ_TRY(ensure_t(ctx, ctx->adj_lam, size_t(k * ldc), &Lam));
  SLQ_TRY(ensure_t(ctx, ctx->adj_v1, size_t(P * n), &xi));
  SLQ_TRY(ensure_t(ctx, ctx->adj_v2, size_t(P * n), &W));
  // da and db as (P x k) arrays (zero when not given)
  float* da = db + P * k;
  if (dalpha)
    SLQ_HIP(ctx, hipMemcpyAsync(da, dalpha, size_t(P * k) * sizeof(float), hipMemcpyDeviceToDevice,
                                s));
  else
    SLQ_HIP(ctx, launch_fill_zero(s, da, P * k));
  if (dbeta)
    SLQ_HIP(ctx, hipMemcpyAsync(db, dbeta, size_t(P * k) * sizeof(float), hipMemcpyDeviceToDevice,
                                s));
  else
    SLQ_HIP(ctx, launch_fill_zero(s, db, P * k));

  // xi = -dx_k, lambda_plus = 0 (decomp.py:290-297) and the dots of step k - 1
  float* part[2] = {part1, part2};
  SLQ_HIP(ctx, launch_lanczos_adj_init(s, dxk, st.R, st.blast, st.Q + (k - 1) * ldc, P, n, xi,
                                       part[(k - 1) & 1]));
  for (int64_t i = k - 1; i >= 0; --i) {
    const bool last = (i + 1 == k);
    LanczosAdjStep sc;
    sc.a = st.H + i * k + i;
    sc.sstride = kk;
    sc.da = da + i;
    sc.dastride = k;
    sc.db1 = db + i;
    sc.dbstride = k;
    if (last) {
      sc.b = st.blast;  // the remainder's norm
      sc.bstride = 1;
      sc.xplus_res = 1;
    } else {
      sc.b = st.H + i * k + i + 1;  // T_{i+1,i}
      sc.bstride = kk;
    }
    sc.part = part[i & 1];
    sc.dx = dQ ? dQ + i * ldq : nullptr;
    const float* x = st.Q + i * ldc;
    const float* xplus = last ? st.R : st.Q + (i + 1) * ldc;
    float* lam = Lam + i * ldc;
    SLQ_HIP(ctx, launch_lanczos_adj_lambda(s, xi, x, xplus, P, n, sc, lam));
    // A lambda (decomp.py:337; the operator is symmetric)
    SLQ_HIP(ctx, launch_matvec(s, op, lam, P, W, mvs, mv_floats));
    // xi for step i - 1 (decomp.py:341) and that step's dots
    SLQ_HIP(ctx, launch_lanczos_adj_xi(s, W, lam, last ? nullptr : Lam + (i + 1) * ldc, xplus,
                                       last, i > 0 ? st.Q + (i - 1) * ldc : x, x, P, n, sc, xi,
                                       i > 0 ? part[(i - 1) & 1] : nullptr));
  }
  // dv = ((xi . x_0) x_0 - xi) / |v| (decomp.py:305)
  if (dv) SLQ_HIP(ctx, launch_lanczos_adj_initvec(s, xi, st.Q, st.c, P, n, part1, dv));
  return param_cotangents(ctx, op, st.Q, Lam, ldc, P, k, dparams, dmatrix);
}

// Largest rank of the partial Cholesky / Woodbury kernels (LDS staging).// Largest rank of the partial Cholesky / Woodbury kernels (LDS staging).
constexpr int64_t kMaxRank = 4096;

int check_rank(slq_context* ctx, int64_t rank, int64_t n) {
  if (n <= 0 || n > kMaxDim)
    return fail(ctx, SLQ_EINVAL, "dimension n must be in 1.." + std::to_string(kMaxDim));
  if (rank > n)
    return fail(ctx, SLQ_EINVAL,
                "Rank exceeds nrows: " + std::to_string(rank) + " >= " + std::to_string(n) + ".");
  if (rank < 1)
    return fail(ctx, SLQ_EINVAL, "Rank must be positive, but " + std::to_string(rank) + " < 1.");
  if (rank > kMaxRank)
    return fail(ctx, SLQ_EUNSUPPORTED,
                "rank " + std::to_string(rank) + " exceeds the supported maximum " +
                    std::to_string(kMaxRank));
  return SLQ_OK;
}

// C = chol(I + L^T L / s) in context scratch (Woodbury factor of s I + L L^T).
int woodbury_factor(slq_context* ctx, const float* L, int64_t n, int64_t r, double shift,
                    double** C) {
  float* gp;
  int* ok;
  SLQ_TRY(ensure_t(ctx, ctx->lr_gp, size_t(gram_splits(n) * r * r), &gp));
  SLQ_TRY(ensure_t(ctx, ctx->lr_c, size_t(r * r), C));
  SLQ_TRY(ensure_t(ctx, ctx->lr_small, size_t(16), &ok));
  SLQ_HIP(ctx, launch_woodbury_setup(ctx->stream, L, n, int(r), shift, gp, *C, ok));
  return SLQ_OK;
}

// The plain-Lanczos reverse pass: dT from the eigh VJP (as above; with
// T = diag(a) + offdiag(b, +-1) the chain gives da_i = dT_ii and
// db_i = dT_{i,i+1} + dT_{i+1,i}), the adjoint recursion of
// decomp.py:268-348 with zero cotangents for the vectors and the remainder
// (quadform discards them, funm.py:237), and the parameter cotangents
// sum_i d/dp <x_i, A(p) lambda_i> (the matvec VJPs of decomp.py:337-338).
// The recursion's last xi feeds only the start vector's cotangent
// (decomp.py:305), which the estimator does not need, so the final matvec is
// skipped.
int logdet_lanczos_reverse(slq_context* ctx, const slq_operator& op, int64_t P, int64_t k,
                           const ForwardState& st, double* dparams, float* dmatrix) {
  const int64_t n = op.n, ldc = st.ldc, kk = k * k;
  hipStream_t s = ctx->stream;
  const int64_t nb = num_chunks(n);
  float *dH, *Lam, *xi, *W, *part1, *part2, *mvs;
  double* vwork = nullptr;
  const int64_t mv_floats = matvec_scratch_floats(n, P);
  SLQ_TRY(ensure_t(ctx, ctx->splitk, size_t(mv_floats), &mvs));
  SLQ_TRY(ensure_t(ctx, ctx->partial1, size_t(P * nb * 3), &part1));
  SLQ_TRY(ensure_t(ctx, ctx->partial2, size_t(P * nb * 3), &part2));
  SLQ_TRY(ensure_t(ctx, ctx->adj_dh, size_t(P * kk), &dH));
  SLQ_TRY(ensure_t(ctx, ctx->adj_lam, size_t(k * ldc), &Lam));
  SLQ_TRY(ensure_t(ctx, ctx->adj_v1, size_t(P * n), &xi));
  SLQ_TRY(ensure_t(ctx, ctx->adj_v2, size_t(P * n), &W));
  const int64_t vw = tridiag_vjp_work_doubles(P, k);
  if (vw > 0) SLQ_TRY(ensure_t(ctx, ctx->adj_work, size_t(vw), &vwork));

  // dT from the eigh VJP of |z|^2 [log T]_00 / P (funm.py:239-241, 331-333)
  SLQ_HIP(ctx, launch_fill_zero(s, dH, P * kk));
  SLQ_HIP(ctx, launch_tridiag_vjp(s, st.H, st.ell, P, k, vwork, dH));

  // reverse scan i = k-1 .. 0 (decomp.py:290-300) from xi = 0, lambda_plus = 0
  float* part[2] = {part1, part2};
  for (int64_t i = k - 1; i >= 0; --i) {
    const bool last = (i + 1 == k);
    LanczosAdjStep sc;
    sc.a = st.H + i * k + i;
    sc.da = dH + i * k + i;
    sc.sstride = kk;
    if (last) {
      // b_{k-1} is the remainder's norm; its cotangent is zero, and so are
      // xi and lambda_plus: mu = 0 and lambda = da x_{k-1}
      sc.b = st.blast;
      sc.bstride = 1;
    } else {
      sc.b = st.H + i * k + i + 1;  // T_{i+1,i}
      sc.bstride = kk;
      sc.db1 = dH + i * k + i + 1;
      sc.db2 = dH + (i + 1) * k + i;
      sc.part = part[i & 1];
    }
    const float* x = st.Q + i * ldc;
    float* lam = Lam + i * ldc;
    SLQ_HIP(ctx, launch_lanczos_adj_lambda(s, last ? nullptr : xi, x,
                                           last ? nullptr : st.Q + (i + 1) * ldc, P, n, sc, lam));
    if (i == 0) break;
    // A lambda (decomp.py:337; the operator is symmetric)
    SLQ_HIP(ctx, launch_matvec(s, op, lam, P, W, mvs, mv_floats));
    // xi for step i - 1 (decomp.py:341) and that step's dots
    SLQ_HIP(ctx, launch_lanczos_adj_xi(s, W, lam, last ? nullptr : Lam + (i + 1) * ldc,
                                       last ? st.R : st.Q + (i + 1) * ldc, last,
                                       st.Q + (i - 1) * ldc, x, P, n, sc, xi, part[(i - 1) & 1]));
  }

  // parameter cotangents: S = sum_{i, p} x_i lambda_i^T
  if (op.kind == SLQ_OP_RBF) {
    SLQ_TRY(rbf_param_contract(ctx, op, st.Q, Lam, ldc, P, k * P, dparams));
  } else {
    // d/dA of <x, A lambda> is x lambda^T
    SLQ_HIP(ctx, launch_outer_sum(s, st.Q, Lam, ldc, P, n, k * P, dmatrix));
  }
  return SLQ_OK;
}

}  // namespace
}  // namespace slq

using namespace slq;

extern "C" {

int slq_context_create(int device, void* stream, slq_context** out) {
  if (!out) return SLQ_EINVAL;
  *out = nullptr;
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || device < 0 || device >= count) {
    (void)hipGetLastError();
    return SLQ_EINVAL;
  }
  if (hipSetDevice(device) != hipSuccess) return SLQ_EHIP;
  setup_static_scratch(device);  // a failed lookup only disables slq_logdet_noalloc
  slq_context* ctx = new slq_context();
  ctx->device = device;
  ctx->stream = static_cast<hipStream_t>(stream);
  *out = ctx;
  return SLQ_OK;
}

int slq_context_destroy(slq_context* ctx) {
  if (!ctx) return SLQ_OK;
  (void)hipSetDevice(ctx->device);
  (void)hipStreamSynchronize(ctx->stream);
  for (auto& vec : ctx->events)
    for (auto& ev : vec) {
      (void)hipEventDestroy(ev.first);
      (void)hipEventDestroy(ev.second);
    }
  for (auto* b : ctx->all_bufs())
    if (b->ptr) (void)hipFree(b->ptr);
  if (ctx->host_count) (void)hipHostFree(ctx->host_count);
  for (auto& ev : ctx->cg_done)
    if (ev) (void)hipEventDestroy(ev);
  delete ctx;
  return SLQ_OK;
}

int slq_context_set_stream(slq_context* ctx, void* stream) {
  if (!ctx) return SLQ_EINVAL;
  ctx->stream = static_cast<hipStream_t>(stream);
  return SLQ_OK;
}

const char* slq_last_error(const slq_context* ctx) { return ctx ? ctx->error.c_str() : ""; }

int64_t slq_context_scratch_bytes(const slq_context* ctx) {
  if (!ctx) return 0;
  int64_t total = 0;
  for (auto* b : const_cast<slq_context*>(ctx)->all_bufs()) total += int64_t(b->bytes);
  return total;
}

int slq_packed_size(int64_t n, int storage, int64_t* packed_bytes, int64_t* scale_count) {
  if (!packed_bytes || !scale_count) return SLQ_EINVAL;
  *packed_bytes = 0;
  *scale_count = 0;
  return packed_size(n, storage, packed_bytes, scale_count) ? SLQ_OK : SLQ_EINVAL;
}

int slq_pack_matrix(slq_context* ctx, const float* A, int64_t n, int storage, void* packed,
                    float* tile_scales) {
  if (!ctx) return SLQ_EINVAL;
  int64_t bytes = 0, scales = 0;
  if (!packed_size(n, storage, &bytes, &scales))
    return fail(ctx, SLQ_EINVAL, "slq_pack_matrix: n must be positive and storage one of the "
                                 "packed forms (not SLQ_STORE_FP32)");
  if (!A || !packed || (scales > 0 && !tile_scales))
    return fail(ctx, SLQ_EINVAL, "input or output pointer is NULL");
  SLQ_TRY(set_device(ctx));
  SLQ_HIP(ctx, launch_pack_matrix(ctx->stream, A, n, storage, packed, tile_scales));
  return SLQ_OK;
}

int slq_sampler_signs(slq_context* ctx, uint32_t key0, uint32_t key1, int64_t num, int64_t n,
                      float* out) {
  if (!ctx) return SLQ_EINVAL;
  if (num < 0 || n < 0) return fail(ctx, SLQ_EINVAL, "num and n must be non-negative");
  if (num > kMaxCount || n > kMaxDim)
    return fail(ctx, SLQ_EINVAL, "num or n exceeds the library's size limits");
  if (num * n > 0 && !out) return fail(ctx, SLQ_EINVAL, "output pointer is NULL");
  SLQ_TRY(set_device(ctx));
  SLQ_HIP(ctx, launch_rademacher(ctx->stream, key0, key1, num * n, out));
  return SLQ_OK;
}

int slq_matvec(slq_context* ctx, const slq_operator* op, const float* V, int64_t P, float* W) {
  if (!ctx) return SLQ_EINVAL;
  SLQ_TRY(check_operator(ctx, op));
  if (P <= 0) return fail(ctx, SLQ_EINVAL, "number of vectors P must be positive");
  SLQ_TRY(check_footprint(ctx, P, op->n, 0));
  if (!V || !W) return fail(ctx, SLQ_EINVAL, "input or output pointer is NULL");
  SLQ_TRY(set_device(ctx));
  float* mvs;
  const int64_t mv_floats = matvec_scratch_floats(op->n, P);
  SLQ_TRY(ensure_t(ctx, ctx->splitk, size_t(mv_floats), &mvs));
  SLQ_HIP(ctx, launch_matvec(ctx->stream, *op, V, P, W, mvs, mv_floats));
  return SLQ_OK;
}

int slq_hessenberg(slq_context* ctx, const slq_operator* op, const float* V0, int64_t P,
                   int64_t k, float* Q, float* H, float* R, float* c) {
  if (!ctx) return SLQ_EINVAL;
  SLQ_TRY(check_operator(ctx, op));
  if (k < 0 || k > op->n) return fail(ctx, SLQ_EINVAL_K, exceeds_msg(k, op->n));
  if (P <= 0) return fail(ctx, SLQ_EINVAL, "number of vectors P must be positive");
  SLQ_TRY(check_footprint(ctx, P, op->n, k));
  if (!V0 || !R || !c || (k > 0 && (!Q || !H)))
    return fail(ctx, SLQ_EINVAL, "input or output pointer is NULL");
  SLQ_TRY(set_device(ctx));
  return hessenberg_core(ctx, *op, V0, P, k, Q, P * op->n, H, R, c);
}

int slq_arnoldi(slq_context* ctx, const slq_operator* op, const float* V0, int64_t P, int64_t k,
                int reortho, float* Q, float* H, float* R, float* c) {
  if (!ctx) return SLQ_EINVAL;
  SLQ_TRY(check_operator(ctx, op));
  if (k < 0 || k > op->n) return fail(ctx, SLQ_EINVAL_K, exceeds_msg(k, op->n));
  if (P <= 0) return fail(ctx, SLQ_EINVAL, "number of vectors P must be positive");
  SLQ_TRY(check_footprint(ctx, P, op->n, k));
  if (!V0 || !R || !c || (k > 0 && (!Q || !H)))
    return fail(ctx, SLQ_EINVAL, "input or output pointer is NULL");
  SLQ_TRY(set_device(ctx));
  return hessenberg_core(ctx, *op, V0, P, k, Q, P * op->n, H, R, c, reortho != 0);
}

int slq_arnoldi_vjp(slq_context* ctx, const slq_operator* op, const float* V0, int64_t P,
                    int64_t k, int reortho, int reortho_adjoint, const float* dQ, const float* dH,
                    const float* dr, const float* dc, float* dv, double* dparams, float* dmatrix) {
  if (!ctx) return SLQ_EINVAL;
  SLQ_TRY(check_operator(ctx, op));
  const int64_t n = op->n;
  if (k < 0 || k > n) return fail(ctx, SLQ_EINVAL_K, exceeds_msg(k, n));
  if (k == 0)
    return fail(ctx, SLQ_EADJOINT_K0,
                "Custom Hessenberg-adjoints are not implemented for num_matvecs = 0.");
  if (P <= 0) return fail(ctx, SLQ_EINVAL, "number of vectors P must be positive");
  SLQ_TRY(check_footprint(ctx, P, op->n, k));
  if (!V0) return fail(ctx, SLQ_EINVAL, "start-vector pointer is NULL");
  if (op->kind == SLQ_OP_RBF && !dparams)
    return fail(ctx, SLQ_EINVAL, "dparams (2 doubles) is NULL");
  if (op->kind == SLQ_OP_DENSE && !dmatrix)
    return fail(ctx, SLQ_EINVAL, "dmatrix (n x n floats) is NULL");
  SLQ_TRY(set_device(ctx));
  ForwardState st;
  st.ldc = basis_stride(P, n);
  SLQ_TRY(ensure_t(ctx, ctx->basis, size_t(k * st.ldc), &st.Q));
  SLQ_TRY(ensure_t(ctx, ctx->hmat, size_t(P * k * k), &st.H));
  SLQ_TRY(ensure_t(ctx, ctx->resid, size_t(P * n), &st.R));
  SLQ_TRY(ensure_t(ctx, ctx->scal1, size_t(P), &st.c));
  SLQ_TRY(hessenberg_core(ctx, *op, V0, P, k, st.Q, st.ldc, st.H, st.R, st.c, reortho != 0));
  Timed tm(ctx, SLQ_KC_ADJOINT);
  return arnoldi_reverse(ctx, *op, P, k, st, reortho_adjoint != 0, dQ, dH, dr, dc, dv, dparams,
                         dmatrix);
}

int slq_lanczos_vjp(slq_context* ctx, const slq_operator* op, const float* V0, int64_t P,
                    int64_t k, const float* dQ, const float* dalpha, const float* dbeta,
                    const float* dxk, float* dv, double* dparams, float* dmatrix) {
  if (!ctx) return SLQ_EINVAL;
  SLQ_TRY(check_operator(ctx, op));
  const int64_t n = op->n;
  if (k < 0 || k > n) return fail(ctx, SLQ_EINVAL_K, exceeds_msg(k, n));
  if (k == 0)
    return fail(ctx, SLQ_EADJOINT_K0, "the Lanczos adjoint needs num_matvecs >= 1");
  if (P <= 0) return fail(ctx, SLQ_EINVAL, "number of vectors P must be positive");
  SLQ_TRY(check_footprint(ctx, P, op->n, k));
  if (!V0) return fail(ctx, SLQ_EINVAL, "start-vector pointer is NULL");
  if (op->kind == SLQ_OP_RBF && !dparams)
    return fail(ctx, SLQ_EINVAL, "dparams (2 doubles) is NULL");
  if (op->kind == SLQ_OP_DENSE && !dmatrix)
    return fail(ctx, SLQ_EINVAL, "dmatrix (n x n floats) is NULL");
  SLQ_TRY(set_device(ctx));
  ForwardState st;
  st.ldc = basis_stride(P, n);
  SLQ_TRY(ensure_t(ctx, ctx->basis, size_t(k * st.ldc), &st.Q));
  SLQ_TRY(ensure_t(ctx, ctx->hmat, size_t(P * k * k), &st.H));
  SLQ_TRY(ensure_t(ctx, ctx->resid, size_t(P * n), &st.R));
  SLQ_TRY(ensure_t(ctx, ctx->scal1, size_t(P), &st.c));
  SLQ_TRY(ensure_t(ctx, ctx->blast, size_t(P), &st.blast));
  SLQ_TRY(lanczos_core(ctx, *op, V0, P, k, st.Q, st.ldc, st.H, st.blast, st.R, st.c));
  Timed tm(ctx, SLQ_KC_ADJOINT);
  return lanczos_reverse(ctx, *op, P, k, st, dQ, dalpha, dbeta, dxk, dv, dparams, dmatrix);
}

int slq_logdet(slq_context* ctx, const slq_operator* op, const float* Z, int64_t P, int64_t k,
               double* per_probe, double* mean, float* alpha, float* beta) {
  if (!ctx) return SLQ_EINVAL;
  SLQ_TRY(check_operator(ctx, op));
  const int64_t n = op->n;
  if (k < 0 || k > n) return fail(ctx, SLQ_EINVAL_K, exceeds_msg(k, n));
  if (k == 0)
    return fail(ctx, SLQ_EINVAL, "the log-determinant integrand needs num_matvecs >= 1");
  if (P <= 0) return fail(ctx, SLQ_EINVAL, "number of probes P must be positive");
  SLQ_TRY(check_footprint(ctx, P, op->n, k));
  if (!Z) return fail(ctx, SLQ_EINVAL, "probe pointer is NULL");
  SLQ_TRY(set_device(ctx));
  ForwardState st;
  return logdet_forward(ctx, *op, Z, P, k, per_probe, mean, alpha, beta, st);
}

int slq_logdet_noalloc(slq_context* ctx, const slq_operator* op, const float* Z, int64_t P,
                       int64_t k, double* per_probe, double* mean, float* alpha, float* beta) {
  if (!ctx) return SLQ_EINVAL;
  SLQ_TRY(check_operator(ctx, op));
  const int64_t n = op->n;
  if (k < 0 || k > n) return fail(ctx, SLQ_EINVAL_K, exceeds_msg(k, n));
  if (k == 0)
    return fail(ctx, SLQ_EINVAL, "the log-determinant integrand needs num_matvecs >= 1");
  if (P <= 0) return fail(ctx, SLQ_EINVAL, "number of probes P must be positive");
  SLQ_TRY(check_footprint(ctx, P, op->n, k));
  if (!Z) return fail(ctx, SLQ_EINVAL, "probe pointer is NULL");
  char* base = nullptr;
  {
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (ctx->device < 0 || ctx->device >= kMaxDevices || !g_scratch[ctx->device].looked_up)
      return fail(ctx, SLQ_EINVAL, "slq_logdet_noalloc: no static scratch on device " +
                                       std::to_string(ctx->device));
    const StaticScratch& st = g_scratch[ctx->device];
    if (st.status != hipSuccess || !st.ptr)
      return fail(ctx, SLQ_EHIP,
                  std::string("slq_logdet_noalloc: the static scratch could not be located: ") +
                      hipGetErrorString(st.status));
    base = static_cast<char*>(st.ptr);
  }
  // The call's buffers are carved from the static scratch in the order the
  // call takes them; a dry run (nothing launched) sizes them first, so a call
  // that does not fit is rejected before any work is queued.
  slq_context sizing;
  sizing.device = ctx->device;
  sizing.stream = ctx->stream;
  sizing.arena_mode = true;
  sizing.dry_run = true;
  sizing.arena = base;
  sizing.arena_size = SIZE_MAX / 2;
  ForwardState st;
  int rc = logdet_forward(&sizing, *op, Z, P, k, per_probe, mean, alpha, beta, st);
  if (rc != SLQ_OK) return fail(ctx, rc, sizing.error);
  if (sizing.arena_off > size_t(SLQ_NOALLOC_SCRATCH_BYTES))
    return fail(ctx, SLQ_EINVAL,
                "slq_logdet_noalloc: the call needs " + std::to_string(sizing.arena_off) +
                    " bytes of scratch, more than the " +
                    std::to_string(size_t(SLQ_NOALLOC_SCRATCH_BYTES)) +
                    " bytes of the static scratch (use slq_logdet)");
  slq_context run;
  run.device = ctx->device;
  run.stream = ctx->stream;
  run.arena_mode = true;
  run.arena = base;
  run.arena_size = size_t(SLQ_NOALLOC_SCRATCH_BYTES);
  rc = set_device(&run);
  if (rc == SLQ_OK) rc = logdet_forward(&run, *op, Z, P, k, per_probe, mean, alpha, beta, st);
  if (rc != SLQ_OK) return fail(ctx, rc, run.error);
  return SLQ_OK;
}

int slq_lanczos(slq_context* ctx, const slq_operator* op, const float* V0, int64_t P, int64_t k,
                float* Q, float* alpha, float* beta, float* R, float* c) {
  if (!ctx) return SLQ_EINVAL;
  SLQ_TRY(check_operator(ctx, op));
  const int64_t n = op->n;
  if (k < 0 || k > n) return fail(ctx, SLQ_EINVAL_K, exceeds_msg(k, n));
  if (P <= 0) return fail(ctx, SLQ_EINVAL, "number of vectors P must be positive");
  SLQ_TRY(check_footprint(ctx, P, op->n, k));
  if (!V0 || !R || !c || !beta || (k > 0 && !alpha))
    return fail(ctx, SLQ_EINVAL, "input or output pointer is NULL");
  SLQ_TRY(set_device(ctx));
  hipStream_t s = ctx->stream;
  float* H = nullptr;
  if (k > 0) SLQ_TRY(ensure_t(ctx, ctx->hmat, size_t(P * k * k), &H));
  float* bl;
  SLQ_TRY(ensure_t(ctx, ctx->blast, size_t(P), &bl));
  SLQ_TRY(lanczos_core(ctx, *op, V0, P, k, Q, P * n, H, bl, R, c));
  // (alpha, beta[0 .. k-2]) from T, beta[k-1] = the remainder's norm
  SLQ_HIP(ctx, launch_lanczos_coeffs(s, H, bl, P, k, alpha, beta));
  return SLQ_OK;
}

int slq_logdet_lanczos(slq_context* ctx, const slq_operator* op, const float* Z, int64_t P,
                       int64_t k, double* per_probe, double* mean, float* alpha, float* beta) {
  if (!ctx) return SLQ_EINVAL;
  SLQ_TRY(check_operator(ctx, op));
  const int64_t n = op->n;
  if (k < 0 || k > n) return fail(ctx, SLQ_EINVAL_K, exceeds_msg(k, n));
  if (k == 0)
    return fail(ctx, SLQ_EINVAL, "the log-determinant integrand needs num_matvecs >= 1");
  if (P <= 0) return fail(ctx, SLQ_EINVAL, "number of probes P must be positive");
  SLQ_TRY(check_footprint(ctx, P, op->n, k));
  if (!Z) return fail(ctx, SLQ_EINVAL, "probe pointer is NULL");
  SLQ_TRY(set_device(ctx));
  ForwardState st;
  return logdet_lanczos_forward(ctx, *op, Z, P, k, per_probe, mean, alpha, beta, false, st);
}

int slq_funm_vectors(slq_context* ctx, const slq_operator* op, const float* V, int64_t P,
                     int64_t k, int reortho_full, int matfun, float* out) {
  if (!ctx) return SLQ_EINVAL;
  SLQ_TRY(check_operator(ctx, op));
  const int64_t n = op->n;
  if (k < 0 || k > n) return fail(ctx, SLQ_EINVAL_K, exceeds_msg(k, n));
  if (k == 0) return fail(ctx, SLQ_EINVAL, "funm_lanczos_sym needs num_matvecs >= 1");
  if (P <= 0) return fail(ctx, SLQ_EINVAL, "number of vectors P must be positive");
  SLQ_TRY(check_footprint(ctx, P, op->n, k));
  if (!V || !out) return fail(ctx, SLQ_EINVAL, "input or output pointer is NULL");
  if (matfun < SLQ_FUNM_LOG || matfun > SLQ_FUNM_INV)
    return fail(ctx, SLQ_EUNSUPPORTED, "matrix function " + std::to_string(matfun) +
                                           " is not available");
  SLQ_TRY(set_device(ctx));
  hipStream_t s = ctx->stream;
  float *v0, *Q, *H, *R, *c, *ell, *norm2, *coef;
  double* work = nullptr;
  const int64_t ldc = basis_stride(P, n);
  SLQ_TRY(ensure_t(ctx, ctx->vec0, size_t(P * n), &v0));
  SLQ_TRY(ensure_t(ctx, ctx->basis, size_t(k * ldc), &Q));
  SLQ_TRY(ensure_t(ctx, ctx->hmat, size_t(P * k * k), &H));
  SLQ_TRY(ensure_t(ctx, ctx->resid, size_t(P * n), &R));
  SLQ_TRY(ensure_t(ctx, ctx->scal1, size_t(P), &c));
  SLQ_TRY(ensure_t(ctx, ctx->scal2, size_t(P), &ell));
  SLQ_TRY(ensure_t(ctx, ctx->norm2, size_t(P * num_chunks(n)), &norm2));
  SLQ_TRY(ensure_t(ctx, ctx->adj_coef, size_t(P * (2 * k + 4)), &coef));
  const int64_t vw = tridiag_vjp_work_doubles(P, k);
  if (vw > 0) SLQ_TRY(ensure_t(ctx, ctx->adj_work, size_t(vw), &work));
  // length = |vec|; vec / length (funm.py:138-139); tridiag_sym (:144)
  SLQ_HIP(ctx, launch_dots(s, nullptr, 0, V, P, n, 1, norm2, 1));
  SLQ_HIP(ctx, launch_normalize(s, V, v0, P, n, norm2, ell, nullptr, 0, nullptr));
  if (reortho_full) {
    SLQ_TRY(hessenberg_core(ctx, *op, v0, P, k, Q, ldc, H, R, c));
  } else {
    float* bl;
    SLQ_TRY(ensure_t(ctx, ctx->blast, size_t(P), &bl));
    SLQ_TRY(lanczos_core(ctx, *op, v0, P, k, Q, ldc, H, bl, R, c));
  }
  // length * Q^T (f(T) e_1) (:145-147)
  SLQ_HIP(ctx, launch_tridiag_funm(s, H, ell, P, k, matfun, work, coef));
  SLQ_HIP(ctx, launch_adjoint_combine(s, Q, nullptr, ldc, nullptr, nullptr, nullptr, out,
                                      nullptr, P, n, k, int(k), int(k), coef));
  return SLQ_OK;
}

int slq_mean_and_sem(slq_context* ctx, const double* values, int64_t P, double* mean,
                     double* sem) {
  if (!ctx) return SLQ_EINVAL;
  if (P <= 0) return fail(ctx, SLQ_EINVAL, "number of values P must be positive");
  SLQ_TRY(check_footprint(ctx, P, 1, 0));
  if (!values) return fail(ctx, SLQ_EINVAL, "values pointer is NULL");
  SLQ_TRY(set_device(ctx));
  SLQ_HIP(ctx, launch_mean_and_sem(ctx->stream, values, P, mean, sem));
  return SLQ_OK;
}

int slq_logdet_vjp(slq_context* ctx, const slq_operator* op, const float* Z, int64_t P, int64_t k,
                   double* per_probe, double* mean, double* dparams, float* dmatrix) {
  if (!ctx) return SLQ_EINVAL;
  SLQ_TRY(check_operator(ctx, op));
  const int64_t n = op->n;
  if (k == 0)
    return fail(ctx, SLQ_EADJOINT_K0,
                "Custom Hessenberg-adjoints are not implemented for num_matvecs = 0.");
  if (k < 0 || k > n) return fail(ctx, SLQ_EINVAL_K, exceeds_msg(k, n));
  if (P <= 0) return fail(ctx, SLQ_EINVAL, "number of probes P must be positive");
  SLQ_TRY(check_footprint(ctx, P, op->n, k));
  if (!Z) return fail(ctx, SLQ_EINVAL, "probe pointer is NULL");
  if (op->kind == SLQ_OP_RBF && !dparams)
    return fail(ctx, SLQ_EINVAL, "dparams (2 doubles) is NULL");
  if (op->kind == SLQ_OP_DENSE && !dmatrix)
    return fail(ctx, SLQ_EINVAL, "dmatrix (n x n floats) is NULL");
  SLQ_TRY(set_device(ctx));
  ForwardState st;
  SLQ_TRY(logdet_forward(ctx, *op, Z, P, k, per_probe, mean, nullptr, nullptr, st));
  Timed tm(ctx, SLQ_KC_ADJOINT);
  return logdet_reverse(ctx, *op, P, k, st, dparams, dmatrix);
}

int slq_cholesky_partial(slq_context* ctx, const slq_operator* op, int64_t rank, int pivot,
                         float* L, int64_t* perm, int* success) {
  if (!ctx) return SLQ_EINVAL;
  SLQ_TRY(check_operator(ctx, op));
  if (op->kind == SLQ_OP_DENSE && op->storage != SLQ_STORE_FP32)
    return fail(ctx, SLQ_EUNSUPPORTED,
                "the partial Cholesky factorisation reads matrix entries: it needs FP32 storage");
  const int64_t n = op->n;
  SLQ_TRY(check_rank(ctx, rank, n));
  if (!L || !success) return fail(ctx, SLQ_EINVAL, "output pointer is NULL");
  SLQ_TRY(set_device(ctx));
  const int64_t nblk = (n + 255) / 256;
  float *lc, *bval, *rowi, *lii;
  int64_t *pm, *bidx;
  SLQ_TRY(ensure_t(ctx, ctx->lr_lc, size_t(rank * n), &lc));
  SLQ_TRY(ensure_t(ctx, ctx->lr_perm, size_t(n), &pm));
  SLQ_TRY(ensure_t(ctx, ctx->lr_bval, size_t(nblk), &bval));
  SLQ_TRY(ensure_t(ctx, ctx->lr_bidx, size_t(nblk), &bidx));
  SLQ_TRY(ensure_t(ctx, ctx->lr_rowi, size_t(rank + 1), &rowi));
  lii = rowi + rank;
  SLQ_HIP(ctx, launch_cholesky_partial(ctx->stream, *op, int(rank), pivot != 0, lc, pm, bval,
                                       bidx, rowi, lii, success, L, perm));
  return SLQ_OK;
}

int slq_lowrank_solve(slq_context* ctx, const float* L, int64_t n, int64_t rank, double shift,
                      const float* V, int64_t P, float* out) {
  if (!ctx) return SLQ_EINVAL;
  SLQ_TRY(check_rank(ctx, rank, n));
  if (!(shift > 0.0)) return fail(ctx, SLQ_EINVAL, "the shift s must be positive");
  if (P <= 0) return fail(ctx, SLQ_EINVAL, "number of vectors P must be positive");
  SLQ_TRY(check_footprint(ctx, P, n, rank));
  if (!L || !V || !out) return fail(ctx, SLQ_EINVAL, "input or output pointer is NULL");
  SLQ_TRY(set_device(ctx));
  double *C, *z;
  float* wp;
  SLQ_TRY(woodbury_factor(ctx, L, n, rank, shift, &C));
  SLQ_TRY(ensure_t(ctx, ctx->lr_wp, size_t(woodbury_chunks(n) * P * rank), &wp));
  SLQ_TRY(ensure_t(ctx, ctx->lr_z, size_t(P * rank), &z));
  SLQ_HIP(ctx, launch_woodbury_apply(ctx->stream, L, n, int(rank), shift, C, V, P, wp, z, out));
  return SLQ_OK;
}

int slq_cg(slq_context* ctx, const slq_operator* op, const float* B, int64_t P, double tol,
           double atol, int64_t maxiter, const float* L, int64_t rank, double shift, float* X,
           int64_t* iterations) {
  if (!ctx) return SLQ_EINVAL;
  SLQ_TRY(check_operator(ctx, op));
  const int64_t n = op->n;
  if (P <= 0) return fail(ctx, SLQ_EINVAL, "number of right-hand sides P must be positive");
  SLQ_TRY(check_footprint(ctx, P, op->n, 0));
  if (!B || !X) return fail(ctx, SLQ_EINVAL, "input or output pointer is NULL");
  if (!(tol >= 0.0) || !(atol >= 0.0) || maxiter < 0)
    return fail(ctx, SLQ_EINVAL, "tol, atol and maxiter must be non-negative");
  const bool precond = L != nullptr;
  if (precond) {
    SLQ_TRY(check_rank(ctx, rank, n));
    if (!(shift > 0.0)) return fail(ctx, SLQ_EINVAL, "the shift s must be positive");
  }
  SLQ_TRY(set_device(ctx));
  hipStream_t s = ctx->stream;
  if (!ctx->host_count) {
    void* h = nullptr;
    if (hipHostMalloc(&h, 64, hipHostMallocDefault) != hipSuccess) {
      (void)hipGetLastError();
      return fail(ctx, SLQ_ENOMEM, "pinned host allocation failed");
    }
    ctx->host_count = static_cast<int*>(h);
  }
  for (auto& ev : ctx->cg_done)
    if (!ev && hipEventCreateWithFlags(&ev, hipEventDisableTiming) != hipSuccess) {
      ev = nullptr;
      (void)hipGetLastError();
      return fail(ctx, SLQ_EHIP, "event creation failed");
    }
  const int64_t nb = cg_chunks(n);
  float *r, *z = nullptr, *d, *q, *parts, *scal, *mvs;
  SLQ_TRY(ensure_t(ctx, ctx->cg_r, size_t(P * n), &r));
  SLQ_TRY(ensure_t(ctx, ctx->cg_d, size_t(P * n), &d));
  SLQ_TRY(ensure_t(ctx, ctx->cg_q, size_t(P * n), &q));
  if (precond) SLQ_TRY(ensure_t(ctx, ctx->cg_z, size_t(P * n), &z));
  SLQ_TRY(ensure_t(ctx, ctx->cg_parts, size_t(4 * P * nb), &parts));
  // scalars: gamma (2 generations), atol2 (floats), iters (int64), active
  // (2 generations), the count (ints); 16 words per column plus the count
  SLQ_TRY(ensure_t(ctx, ctx->cg_scal, size_t(16 * P + 64), &scal));
  const int64_t mv_floats = matvec_scratch_floats(n, P);
  SLQ_TRY(ensure_t(ctx, ctx->splitk, size_t(mv_floats), &mvs));
  float *pq = parts, *rr = parts + P * nb, *rz = parts + 2 * P * nb, *bb = parts + 3 * P * nb;
  CgScalars sc;
  sc.gamma = scal;
  sc.atol2 = scal + 2 * P;
  sc.iters = reinterpret_cast<int64_t*>(scal + 4 * P);  // 16 P bytes in: 8-byte aligned
  sc.active = reinterpret_cast<int*>(scal + 8 * P + 2);
  int* nactive = reinterpret_cast<int*>(scal + 10 * P + 2);
  CgPrecond M;
  if (precond) {
    double* C;
    SLQ_TRY(woodbury_factor(ctx, L, n, rank, shift, &C));
    M.L = L;
    M.rank = int(rank);
    M.shift = shift;
    M.C = C;
    SLQ_TRY(ensure_t(ctx, ctx->lr_wp, size_t(woodbury_chunks(n) * P * rank), &M.Wp));
    SLQ_TRY(ensure_t(ctx, ctx->lr_z, size_t(P * rank), &M.Z));
  }
  // x0 = 0; r0 = b - A x0 = b; p0 = z0 = M r0; gamma0 = r0 . z0
  SLQ_HIP(ctx, launch_fill_zero(s, X, P * n));
  SLQ_HIP(ctx, hipMemcpyAsync(r, B, size_t(P * n) * sizeof(float), hipMemcpyDeviceToDevice, s));
  if (precond) {
    SLQ_HIP(ctx, launch_woodbury_apply(s, L, n, int(rank), shift, M.C, r, P, M.Wp, M.Z, z));
    SLQ_HIP(ctx, launch_cg_dot(s, r, z, P, n, rz));
    SLQ_HIP(ctx, launch_cg_dot(s, r, r, P, n, rr));
  } else {
    SLQ_HIP(ctx, launch_cg_dot(s, r, r, P, n, rz));
  }
  SLQ_HIP(ctx, hipMemcpyAsync(d, precond ? z : r, size_t(P * n) * sizeof(float),
                              hipMemcpyDeviceToDevice, s));
  SLQ_HIP(ctx, launch_cg_dot(s, B, B, P, n, bb));
  SLQ_HIP(ctx, launch_cg_start(s, rz, bb, precond ? rr : nullptr, P, n, float(tol), float(atol),
                               maxiter, sc, nactive));
  // Iterations run in batches of 1, 2, 4, 8, 8, ...; after batch i the count
  // of active columns is copied to pinned host memory (slot (i + 1) % 2).
  // While the batches grow the host waits for each batch's own count; once
  // they are 8 long it keeps one batch in flight: it enqueues batch i, then
  // waits for the count of batch i - 1, so that the GPU does not idle while
  // the host polls.  It stops at a zero count; iterations that find no
  // active column (the batch still in flight) change nothing.  Iteration t
  // reads scalar generation t % 2 and writes the other.
  int* hc = ctx->host_count;
  int64_t t = 0;
  for (int64_t i = 0, batch = 1;; ++i, batch = batch < 8 ? 2 * batch : 8) {
    for (int64_t b = 0; b < batch; ++b, ++t) {
      MvSplitSum qsum;
      {
        const auto ev = take_events(ctx, SLQ_KC_MATVEC);
        // the RBF point records of the first iteration's matvec are reused
        SLQ_HIP(ctx, launch_matvec(s, *op, d, P, q, mvs, mv_floats, ev.first, ev.second,
                                   t == 0 ? 0 : kMvPointsReady, &qsum));
      }
      SLQ_HIP(ctx, launch_cg_step(s, X, r, d, q, qsum, z, precond, P, n, maxiter, sc,
                                  int(t & 1), pq, rr, rz, precond ? &M : nullptr));
    }
    SLQ_HIP(ctx, launch_cg_count(s, sc, int(t & 1), P, nactive));
    const int slot = int((i + 1) & 1);
    SLQ_HIP(ctx, hipMemcpyAsync(hc + slot, nactive, sizeof(int), hipMemcpyDeviceToHost, s));
    SLQ_HIP(ctx, hipEventRecord(ctx->cg_done[slot], s));
    const int wait = batch < 8 ? slot : int(i & 1);
    SLQ_HIP(ctx, hipEventSynchronize(ctx->cg_done[wait]));
    if (hc[wait] == 0) break;
  }
  if (iterations)
    SLQ_HIP(ctx, hipMemcpyAsync(iterations, sc.iters, size_t(P) * sizeof(int64_t),
                                hipMemcpyDeviceToDevice, s));
  return SLQ_OK;
}

int slq_quadform_vjp(slq_context* ctx, const slq_operator* op, const float* X, int64_t P,
                     double* dparams, float* dmatrix) {
  if (!ctx) return SLQ_EINVAL;
  SLQ_TRY(check_operator(ctx, op));
  const int64_t n = op->n;
  if (P <= 0) return fail(ctx, SLQ_EINVAL, "number of vectors P must be positive");
  SLQ_TRY(check_footprint(ctx, P, op->n, 0));
  if (!X) return fail(ctx, SLQ_EINVAL, "input pointer is NULL");
  if (op->kind == SLQ_OP_RBF && !dparams)
    return fail(ctx, SLQ_EINVAL, "dparams (2 doubles) is NULL");
  if (op->kind == SLQ_OP_DENSE && !dmatrix)
    return fail(ctx, SLQ_EINVAL, "dmatrix (n x n floats) is NULL");
  SLQ_TRY(set_device(ctx));
  hipStream_t s = ctx->stream;
  if (op->kind == SLQ_OP_RBF) {
    SLQ_TRY(rbf_param_contract(ctx, *op, X, X, P * n, P, P, dparams));
  } else {
    SLQ_HIP(ctx, launch_outer_sum(s, X, X, P * n, P, n, P, dmatrix));
  }
  return SLQ_OK;
}

int slq_logdet_lanczos_vjp(slq_context* ctx, const slq_operator* op, const float* Z, int64_t P,
                           int64_t k, double* per_probe, double* mean, double* dparams,
                           float* dmatrix) {
  if (!ctx) return SLQ_EINVAL;
  SLQ_TRY(check_operator(ctx, op));
  const int64_t n = op->n;
  if (k < 0 || k > n) return fail(ctx, SLQ_EINVAL_K, exceeds_msg(k, n));
  if (k == 0)
    return fail(ctx, SLQ_EINVAL, "the log-determinant integrand needs num_matvecs >= 1");
  if (P <= 0) return fail(ctx, SLQ_EINVAL, "number of probes P must be positive");
  SLQ_TRY(check_footprint(ctx, P, op->n, k));
  if (!Z) return fail(ctx, SLQ_EINVAL, "probe pointer is NULL");
  if (op->kind == SLQ_OP_RBF && !dparams)
    return fail(ctx, SLQ_EINVAL, "dparams (2 doubles) is NULL");
  if (op->kind == SLQ_OP_DENSE && !dmatrix)
    return fail(ctx, SLQ_EINVAL, "dmatrix (n x n floats) is NULL");
  SLQ_TRY(set_device(ctx));
  ForwardState st;
  SLQ_TRY(logdet_lanczos_forward(ctx, *op, Z, P, k, per_probe, mean, nullptr, nullptr, true, st));
  Timed tm(ctx, SLQ_KC_ADJOINT);
  return logdet_lanczos_reverse(ctx, *op, P, k, st, dparams, dmatrix);
}

int slq_context_release(slq_context* ctx) {
  if (!ctx) return SLQ_EINVAL;
  SLQ_TRY(set_device(ctx));
  SLQ_HIP(ctx, hipStreamSynchronize(ctx->stream));
  for (auto* b : ctx->all_bufs()) {
    if (b->ptr) SLQ_HIP(ctx, hipFree(b->ptr));
    b->ptr = nullptr;
    b->bytes = 0;
  }
  return SLQ_OK;
}

// The scratch of one call of entry, by running the call as a dry run: from an
// empty pool (fresh), or from a pool with ctx's buffer sizes (grow: the bytes
// by which ctx's pool grows).
static int workspace_query(slq_context* ctx, int entry, const slq_operator* op, int64_t P,
                           int64_t k, bool grow, int64_t* bytes) {
  if (!ctx) return SLQ_EINVAL;
  if (!bytes) return fail(ctx, SLQ_EINVAL, "output pointer is NULL");
  *bytes = 0;
  SLQ_TRY(check_operator(ctx, op));
  const int64_t n = op->n;
  if (k < 0 || k > n) return fail(ctx, SLQ_EINVAL_K, exceeds_msg(k, n));
  if (k == 0) return fail(ctx, SLQ_EINVAL, "the log-determinant integrand needs num_matvecs >= 1");
  if (P <= 0) return fail(ctx, SLQ_EINVAL, "number of probes P must be positive");
  SLQ_TRY(check_footprint(ctx, P, op->n, k));
  // the call itself in a dry run: every buffer it would take is counted and
  // nothing is launched, so the placeholders below are never dereferenced
  slq_context v;
  v.device = ctx->device;
  v.stream = ctx->stream;
  v.arena_mode = true;
  v.dry_run = true;
  v.arena = reinterpret_cast<char*>(uintptr_t(4096));  // placeholder base (never dereferenced)
  v.arena_size = SIZE_MAX / 2;
  if (grow) {
    const auto from = ctx->all_bufs();
    const auto to = v.all_bufs();
    for (size_t i = 0; i < from.size(); ++i) {
      to[i]->bytes = from[i]->bytes;
      to[i]->ptr = from[i]->bytes ? v.arena : nullptr;  // placeholder
    }
  }
  const float* Z = reinterpret_cast<const float*>(op);
  double dd = 0.0;
  float df = 0.f;
  ForwardState st;
  int rc;
  switch (entry) {
    case SLQ_ENTRY_LOGDET:
      rc = logdet_forward(&v, *op, Z, P, k, &dd, &dd, nullptr, nullptr, st);
      break;
    case SLQ_ENTRY_LOGDET_LANCZOS:
      rc = logdet_lanczos_forward(&v, *op, Z, P, k, &dd, &dd, nullptr, nullptr, false, st);
      break;
    case SLQ_ENTRY_LOGDET_VJP:
      rc = logdet_forward(&v, *op, Z, P, k, &dd, &dd, nullptr, nullptr, st);
      if (rc == SLQ_OK) rc = logdet_reverse(&v, *op, P, k, st, &dd, &df);
      break;
    case SLQ_ENTRY_LOGDET_LANCZOS_VJP:
      rc = logdet_lanczos_forward(&v, *op, Z, P, k, &dd, &dd, nullptr, nullptr, true, st);
      if (rc == SLQ_OK) rc = logdet_lanczos_reverse(&v, *op, P, k, st, &dd, &df);
      break;
    default:
      return fail(ctx, SLQ_EINVAL, "unknown entry " + std::to_string(entry));
  }
  if (rc != SLQ_OK) return fail(ctx, rc, v.error);
  *bytes = int64_t(v.arena_off - (grow ? v.arena_replaced : 0));
  return SLQ_OK;
}

int slq_workspace_bytes(slq_context* ctx, int entry, const slq_operator* op, int64_t P, int64_t k,
                        int64_t* bytes) {
  return workspace_query(ctx, entry, op, P, k, false, bytes);
}

int slq_workspace_growth(slq_context* ctx, int entry, const slq_operator* op, int64_t P,
                         int64_t k, int64_t* bytes) {
  return workspace_query(ctx, entry, op, P, k, true, bytes);
}

int slq_logdet_save(slq_context* ctx, const slq_operator* op, const float* Z, int64_t P, int64_t k,
                    int reortho_full, double* per_probe, double* mean, slq_saved** out) {
  if (!ctx) return SLQ_EINVAL;
  if (!out) return fail(ctx, SLQ_EINVAL, "output pointer is NULL");
  *out = nullptr;
  SLQ_TRY(check_operator(ctx, op));
  const int64_t n = op->n;
  if (k == 0 && reortho_full)
    return fail(ctx, SLQ_EADJOINT_K0,
                "Custom Hessenberg-adjoints are not implemented for num_matvecs = 0.");
  if (k < 0 || k > n) return fail(ctx, SLQ_EINVAL_K, exceeds_msg(k, n));
  if (k == 0) return fail(ctx, SLQ_EINVAL, "the log-determinant integrand needs num_matvecs >= 1");
  if (P <= 0) return fail(ctx, SLQ_EINVAL, "number of probes P must be positive");
  SLQ_TRY(check_footprint(ctx, P, op->n, k));
  if (!Z) return fail(ctx, SLQ_EINVAL, "probe pointer is NULL");
  SLQ_TRY(set_device(ctx));
  ForwardState st;
  if (reortho_full)
    SLQ_TRY(logdet_forward(ctx, *op, Z, P, k, per_probe, mean, nullptr, nullptr, st));
  else
    SLQ_TRY(logdet_lanczos_forward(ctx, *op, Z, P, k, per_probe, mean, nullptr, nullptr, true, st));
  // the forward's buffers leave the pool: later calls through ctx take new ones
  slq_saved* sv = new slq_saved();
  sv->device = ctx->device;
  sv->full = reortho_full ? 1 : 0;
  sv->kind = op->kind;
  sv->n = n;
  sv->P = P;
  sv->k = k;
  sv->ldc = st.ldc;
  slq_context::Buf* from[] = {&ctx->basis, &ctx->hmat, &ctx->resid, &ctx->scal1, &ctx->scal2,
                              &ctx->blast};
  auto to = sv->all_bufs();
  for (size_t i = 0; i < to.size(); ++i) {
    *to[i] = *from[i];
    *from[i] = slq_context::Buf{};
  }
  *out = sv;
  return SLQ_OK;
}

int slq_logdet_vjp_saved(slq_context* ctx, const slq_operator* op, const slq_saved* saved,
                         double* dparams, float* dmatrix) {
  if (!ctx) return SLQ_EINVAL;
  if (!saved) return fail(ctx, SLQ_EINVAL, "saved state is NULL");
  SLQ_TRY(check_operator(ctx, op));
  if (op->n != saved->n || op->kind != saved->kind || ctx->device != saved->device)
    return fail(ctx, SLQ_EINVAL, "the operator or device differs from the saved forward pass");
  if (op->kind == SLQ_OP_RBF && !dparams)
    return fail(ctx, SLQ_EINVAL, "dparams (2 doubles) is NULL");
  if (op->kind == SLQ_OP_DENSE && !dmatrix)
    return fail(ctx, SLQ_EINVAL, "dmatrix (n x n floats) is NULL");
  SLQ_TRY(set_device(ctx));
  ForwardState st;
  st.Q = static_cast<float*>(saved->basis.ptr);
  st.H = static_cast<float*>(saved->hmat.ptr);
  st.R = static_cast<float*>(saved->resid.ptr);
  st.c = static_cast<float*>(saved->scal1.ptr);
  st.ell = static_cast<float*>(saved->scal2.ptr);
  st.blast = static_cast<float*>(saved->blast.ptr);
  st.ldc = saved->ldc;
  Timed tm(ctx, SLQ_KC_ADJOINT);
  if (saved->full) return logdet_reverse(ctx, *op, saved->P, saved->k, st, dparams, dmatrix);
  return logdet_lanczos_reverse(ctx, *op, saved->P, saved->k, st, dparams, dmatrix);
}

int64_t slq_saved_bytes(const slq_saved* saved) {
  if (!saved) return 0;
  int64_t total = 0;
  for (auto* b : const_cast<slq_saved*>(saved)->all_bufs()) total += int64_t(b->bytes);
  return total;
}

int slq_saved_destroy(slq_context* ctx, slq_saved* saved) {
  if (!saved) return SLQ_OK;
  (void)hipSetDevice(saved->device);
  if (ctx)
    (void)hipStreamSynchronize(ctx->stream);
  else
    (void)hipDeviceSynchronize();
  for (auto* b : saved->all_bufs())
    if (b->ptr) (void)hipFree(b->ptr);
  delete saved;
  return SLQ_OK;
}

}  // extern "C"

extern "C" {

int slq_profile_enable(slq_context* ctx, int on) {
  if (!ctx) return SLQ_EINVAL;
  ctx->profile_mask = on != 0 ? (1u << SLQ_KC_COUNT) - 1u : 0u;
  return SLQ_OK;
}

int slq_profile_select(slq_context* ctx, int mask) {
  if (!ctx) return SLQ_EINVAL;
  if (mask < 0 || mask >= (1 << SLQ_KC_COUNT))
    return fail(ctx, SLQ_EINVAL, "profile mask has bits beyond the kernel classes");
  ctx->profile_mask = unsigned(mask);
  return SLQ_OK;
}

int slq_profile_read(slq_context* ctx, int kernel_class, double* total_ms, int64_t* launches) {
  if (!ctx) return SLQ_EINVAL;
  if (kernel_class < 0 || kernel_class >= SLQ_KC_COUNT || !total_ms || !launches)
    return fail(ctx, SLQ_EINVAL, "bad kernel class or NULL output");
  SLQ_TRY(set_device(ctx));
  SLQ_HIP(ctx, hipStreamSynchronize(ctx->stream));
  double sum = 0.0;
  auto& vec = ctx->events[kernel_class];
  const size_t used = ctx->used[kernel_class];
  for (size_t i = 0; i < used; ++i) {
    float ms = 0.f;
    SLQ_HIP(ctx, hipEventElapsedTime(&ms, vec[i].first, vec[i].second));
    sum += ms;
  }
  ctx->used[kernel_class] = 0;
  *total_ms = sum;
  *launches = int64_t(used);
  return SLQ_OK;
}

}  // extern "C"
