// Streaming kernels of the Arnoldi step with full reorthogonalisation
// (classical Gram-Schmidt applied twice), decomp.py:454-477:
//
//   v /= length; Q[:, j] = v; w = A v                    (normalize + matvec)
//   h = Q^T w; w = w - Q h                               (pass 1: dots, pass 2: project)
//   w = w - Q (Q^T w)                                     (pass 2: dots, pass 3: project)
//   length = sqrt(<w, w>); h[j + 1] = length; H[:, j] = h
//
// Only the j + 1 live basis columns are read (the reference multiplies by the
// full-width zero-padded Q; the extra products are exact zeros).  All kernels
// are HBM-bound.  A block owns one chunk of kChunk elements of one probe's
// vector; cross-block sums go through per-chunk partials that the consumer
// reduces in a fixed order (deterministic results).
//
// Column reductions use a wave-level reduce-scatter: with NC partial sums
// per lane, log2(NC) exchange steps halve the live values each time
// (NC/2 + NC/4 + ... + 1 shuffles in total instead of 6 per column), after
// which lane l holds the wave total of column l >> (6 - log2 NC).
//
// Basis layout: column c of probe p at Q + c * ldc + p * n (ldc >= P * n).
//
// Plain Lanczos (reortho="none", decomp.py:198-266) uses the single-column
// dots, k_three_term and k_normalize: no basis is read.
#include <initializer_list>
#include <type_traits>

#include "slq_internal.h"

namespace slq {
namespace {

constexpr int kWaves = kStreamThreads / 64;

__device__ __forceinline__ float wave_sum(float v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// Reduce-scatter of NC per-lane values (NC a power of two, <= 64).
template <int L, int OFF>
struct ReduceScatter {
  __device__ static __forceinline__ void run(float* acc, int lane) {
    if constexpr (OFF == 0) {
      return;
    } else if constexpr (L > 1) {
      const bool upper = (lane & OFF) != 0;
#pragma unroll
      for (int i = 0; i < L / 2; ++i) {
        const float lo = acc[i], hi = acc[i + L / 2];
        const float send = upper ? lo : hi;
        const float keep = upper ? hi : lo;
        acc[i] = keep + __shfl_xor(send, OFF, 64);
      }
      ReduceScatter<L / 2, OFF / 2>::run(acc, lane);
    } else {
      acc[0] += __shfl_xor(acc[0], OFF, 64);
      ReduceScatter<1, OFF / 2>::run(acc, lane);
    }
  }
};

constexpr int log2i(int v) { return v <= 1 ? 0 : 1 + log2i(v / 2); }

// Block-wide column sums of acc[0..NC) (one set per thread); the sums of the
// columns c < ncols are written to out[c] (out is global memory).
template <int NC>
__device__ __forceinline__ void block_column_sums(float (&acc)[NC], int ncols, float* red,
                                                  float* out) {
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  ReduceScatter<NC, 32>::run(acc, lane);
  constexpr int shift = 6 - log2i(NC);
  if ((lane & ((1 << shift) - 1)) == 0) red[wave * NC + (lane >> shift)] = acc[0];
  __syncthreads();
  if (t < ncols) {
    float s = 0.f;
#pragma unroll
    for (int w = 0; w < kWaves; ++w) s += red[w * NC + t];
    out[t] = s;
  }
  __syncthreads();
}

template <int V>
struct VecIO;

template <>
struct VecIO<4> {
  __device__ static __forceinline__ void load(const float* p, float (&r)[4]) {
    const float4 t = *reinterpret_cast<const float4*>(p);
    r[0] = t.x; r[1] = t.y; r[2] = t.z; r[3] = t.w;
  }
  __device__ static __forceinline__ void store(float* p, const float (&r)[4]) {
    *reinterpret_cast<float4*>(p) = make_float4(r[0], r[1], r[2], r[3]);
  }
};

template <>
struct VecIO<2> {
  __device__ static __forceinline__ void load(const float* p, float (&r)[2]) {
    const float2 t = *reinterpret_cast<const float2*>(p);
    r[0] = t.x; r[1] = t.y;
  }
  __device__ static __forceinline__ void store(float* p, const float (&r)[2]) {
    *reinterpret_cast<float2*>(p) = make_float2(r[0], r[1]);
  }
};

template <>
struct VecIO<1> {
  __device__ static __forceinline__ void load(const float* p, float (&r)[1]) { r[0] = *p; }
  __device__ static __forceinline__ void store(float* p, const float (&r)[1]) { *p = r[0]; }
};

// Elements handled by one thread: NE groups of V contiguous floats; group e
// starts at chunk_base + (e * kStreamThreads + t) * V.
template <int V>
struct Layout {
  static constexpr int NE = int(kChunk / (kStreamThreads * V));
};

// Block-wide sum of one value per thread; every thread gets the result.
__device__ __forceinline__ float block_sum(float v, float* red) {
  v = wave_sum(v);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0) red[wave] = v;
  __syncthreads();
  float s = 0.f;
#pragma unroll
  for (int w = 0; w < kWaves; ++w) s += red[w];
  __syncthreads();
  return s;
}

// ---------------------------------------------------------------------------
// dots: partial[p][b][c] = sum over chunk b of Q[c][p][:] . x[p][:]
// Columns in batches of NC (all held as per-lane partial sums).
// ---------------------------------------------------------------------------
// x of probe p, elements i .. i + V - 1: loaded, or (sm.partial) formed
// from a deferred split-K sum exactly as the matvec's reduction does and
// stored to x.
// (The partial planes z >= 1 start at z P n floats, which need not keep the
// 16-byte alignment of plane 0 when P n % 4 != 0: they are read by memcpy,
// one global_load_dwordx4 at any dword alignment on gfx950.)
template <int V>
__device__ __forceinline__ void load_x(float* x, const MvSplitSum& sm, int64_t P, int64_t n,
                                       int64_t p, int64_t i, float (&xv)[V]) {
  const int64_t count = P * n, off = p * n + i;
  __builtin_memcpy(xv, sm.partial + off, sizeof(xv));
  for (int z = 1; z < sm.S; ++z) {
    float pz[V];
    __builtin_memcpy(pz, sm.partial + int64_t(z) * count + off, sizeof(pz));
#pragma unroll
    for (int v = 0; v < V; ++v) xv[v] += pz[v];
  }
  if (sm.nugget) {
    float vv[V];
    __builtin_memcpy(vv, sm.V + off, sizeof(vv));
#pragma unroll
    for (int v = 0; v < V; ++v) xv[v] = fmaf(sm.noise, vv[v], xv[v]);
  }
  VecIO<V>::store(x + p * n + i, xv);
}

// Basis columns whose loads are issued together (one batch in flight per
// lane before the first product): a loop that loads, waits and multiplies one
// column at a time leaves each wave one load in flight, and the kernels then
// reach HBM bandwidth only at full occupancy.  Column indices past the last
// live column are clamped to it (the loads stay in bounds, the products are
// not used), and element groups past n read the row's last group with x = 0.
constexpr int kColBatch = 8;

template <int V, int NC, bool SUM>
__global__ void __launch_bounds__(kStreamThreads)
    k_dots(const float* __restrict__ Q, int64_t ldc, const float* __restrict__ x, float* xw,
           int64_t n, int ncols, float* __restrict__ partial, int pstride, MvSplitSum sm) {
  constexpr int NE = Layout<V>::NE;
  constexpr int CB = NC < kColBatch ? NC : kColBatch;
  __shared__ float red[kWaves * NC];
  const int64_t p = blockIdx.y;
  const int64_t b = blockIdx.x;
  const int64_t nb = gridDim.x;
  const int t = threadIdx.x;
  const int64_t base = b * kChunk;

  float xv[NE][V];
  int64_t off[NE];
#pragma unroll
  for (int e = 0; e < NE; ++e) {
    const int64_t i = base + (int64_t(e) * kStreamThreads + t) * V;
    off[e] = i < n ? i : n - V;  // V divides n
    if (i < n) {
      if constexpr (SUM)
        load_x<V>(xw, sm, int64_t(gridDim.y), n, p, i, xv[e]);
      else
        VecIO<V>::load(x + p * n + i, xv[e]);
    } else {
#pragma unroll
      for (int v = 0; v < V; ++v) xv[e][v] = 0.f;
    }
  }

  const float* qp = Q + p * n;
  for (int c0 = 0; c0 < ncols; c0 += NC) {
    float acc[NC];
#pragma unroll
    for (int c = 0; c < NC; ++c) acc[c] = 0.f;
#pragma unroll
    for (int g = 0; g < NC; g += CB) {
      if (c0 + g < ncols) {
        float qv[CB][NE][V];
#pragma unroll
        for (int u = 0; u < CB; ++u) {
          const int c = c0 + g + u < ncols ? c0 + g + u : ncols - 1;
          const float* q = qp + int64_t(c) * ldc;
#pragma unroll
          for (int e = 0; e < NE; ++e) VecIO<V>::load(q + off[e], qv[u][e]);
        }
#pragma unroll
        for (int u = 0; u < CB; ++u)
#pragma unroll
          for (int e = 0; e < NE; ++e)
#pragma unroll
            for (int v = 0; v < V; ++v) acc[g + u] = fmaf(qv[u][e][v], xv[e][v], acc[g + u]);
      }
    }
    const int live = ncols - c0 < NC ? ncols - c0 : NC;
    block_column_sums<NC>(acc, live, red, partial + (p * nb + b) * pstride + c0);
  }
}

// Rows whose length n is not a multiple of 4 start at every alignment (probe
// rows are contiguous).  The *_shifted kernels read row p as float4 groups
// from its first 16-byte boundary: group e of thread t in chunk b starts at
// sh_p + b kChunk + (e kStreamThreads + t) 4, sh_p = (-p n) mod 4, and the
// groups that end by n are the full ones, [sh_p, gend_p).  The up to three
// head elements [0, sh_p) and up to three tail elements [gend_p, n) are
// single floats, one per thread of block 0 (threads 0-2 the head, threads
// 4-6 the tail).  The basis stride ldc is a multiple of 4, so every column of
// the row shares its alignment.  Launched for n >= 8 (at least one full
// group).
__device__ __forceinline__ int row_shift(int64_t p, int64_t n) { return int((4 - (p * n) % 4) % 4); }

struct ShiftedRow {
  int64_t gend;  // end of the full groups
  int64_t ix;    // the thread's single element, or -1
  __device__ __forceinline__ ShiftedRow(int64_t p, int64_t n, int64_t b, int t) {
    const int sh = row_shift(p, n);
    gend = sh + (n - sh) / 4 * 4;
    ix = -1;
    if (b == 0) {
      if (t < sh) ix = t;
      else if (t >= 4 && t - 4 < n - gend) ix = gend + (t - 4);
    }
  }
  // start of group e of thread t in chunk b, and its in-bounds load offset
  __device__ __forceinline__ static int64_t group(int64_t p, int64_t n, int64_t b, int e, int t) {
    return row_shift(p, n) + b * kChunk + (int64_t(e) * kStreamThreads + t) * 4;
  }
};

// SUM: x is the deferred split-K sum `sm` (formed as the matvec's reduction
// does it and written to xw), as in k_dots.
template <int NC, bool SUM>
__global__ void __launch_bounds__(kStreamThreads)
    k_dots_shifted(const float* __restrict__ Q, int64_t ldc, const float* __restrict__ x,
                   float* xw, int64_t n, int ncols, float* __restrict__ partial, int pstride,
                   MvSplitSum sm) {
  constexpr int NE = Layout<4>::NE;
  constexpr int CB = NC < kColBatch ? NC : kColBatch;
  __shared__ float red[kWaves * NC];
  const int64_t p = blockIdx.y, b = blockIdx.x, nb = gridDim.x;
  const int t = threadIdx.x;
  const ShiftedRow row(p, n, b, t);
  const float* xr = x + p * n;
  float xv[NE][4];
  int64_t off[NE];
#pragma unroll
  for (int e = 0; e < NE; ++e) {
    const int64_t i = ShiftedRow::group(p, n, b, e, t);
    const bool full = i < row.gend;
    off[e] = full ? i : row.gend - 4;
    if (full) {
      if constexpr (SUM)
        load_x<4>(xw, sm, int64_t(gridDim.y), n, p, i, xv[e]);
      else
        VecIO<4>::load(xr + i, xv[e]);
    } else {
#pragma unroll
      for (int v = 0; v < 4; ++v) xv[e][v] = 0.f;
    }
  }
  const int64_t ixo = row.ix >= 0 ? row.ix : 0;
  float xh = 0.f;
  if (row.ix >= 0) {
    if constexpr (SUM) {
      float one[1];
      load_x<1>(xw, sm, int64_t(gridDim.y), n, p, ixo, one);
      xh = one[0];
    } else {
      xh = xr[ixo];
    }
  }

  const float* qp = Q + p * n;
  for (int c0 = 0; c0 < ncols; c0 += NC) {
    float acc[NC];
#pragma unroll
    for (int c = 0; c < NC; ++c) acc[c] = 0.f;
#pragma unroll
    for (int g = 0; g < NC; g += CB) {
      if (c0 + g < ncols) {
        float qv[CB][NE][4], qh[CB];
#pragma unroll
        for (int u = 0; u < CB; ++u) {
          const int c = c0 + g + u < ncols ? c0 + g + u : ncols - 1;
          const float* q = qp + int64_t(c) * ldc;
#pragma unroll
          for (int e = 0; e < NE; ++e) VecIO<4>::load(q + off[e], qv[u][e]);
          qh[u] = q[ixo];
        }
#pragma unroll
        for (int u = 0; u < CB; ++u) {
#pragma unroll
          for (int e = 0; e < NE; ++e)
#pragma unroll
            for (int v = 0; v < 4; ++v) acc[g + u] = fmaf(qv[u][e][v], xv[e][v], acc[g + u]);
          acc[g + u] = fmaf(qh[u], xh, acc[g + u]);
        }
      }
    }
    const int live = ncols - c0 < NC ? ncols - c0 : NC;
    block_column_sums<NC>(acc, live, red, partial + (p * nb + b) * pstride + c0);
  }
}

// sum of squares: partial[p][b] = sum over chunk b of x[p][:]^2
template <int V>
__global__ void __launch_bounds__(kStreamThreads)
    k_sumsq(const float* __restrict__ x, int64_t n, float* __restrict__ partial) {
  constexpr int NE = Layout<V>::NE;
  __shared__ float red[kWaves];
  const int64_t p = blockIdx.y, b = blockIdx.x, nb = gridDim.x;
  const int t = threadIdx.x;
  const int64_t base = b * kChunk;
  float acc = 0.f;
#pragma unroll
  for (int e = 0; e < NE; ++e) {
    const int64_t i = base + (int64_t(e) * kStreamThreads + t) * V;
    if (i < n) {
      float xv[V];
      VecIO<V>::load(x + p * n + i, xv);
#pragma unroll
      for (int v = 0; v < V; ++v) acc = fmaf(xv[v], xv[v], acc);
    }
  }
  const float s = block_sum(acc, red);
  if (t == 0) partial[p * nb + b] = s;
}

// Reduce coefficient partials of probe p into LDS (fixed order over chunks);
// hout (if given) receives the sums, and hsub (if given) is subtracted from
// them afterwards: hs[c] = sum_b hpart[p][b][c] - hsub[p * hsub_stride + c].
__device__ __forceinline__ void load_coeffs(const float* __restrict__ hpart, int64_t p,
                                            int64_t nb, int ncols, int pstride, float* hs,
                                            float* hout, int64_t hout_stride, bool write_out,
                                            const float* __restrict__ hsub = nullptr,
                                            int64_t hsub_stride = 0) {
  for (int c = threadIdx.x; c < ncols; c += blockDim.x) {
    float s = 0.f;
    s = ordered_sum(hpart + p * nb * pstride + c, nb, pstride);
    if (write_out && hout) hout[p * hout_stride + c] = s;
    if (hsub) s -= hsub[p * hsub_stride + c];
    hs[c] = s;
  }
  __syncthreads();
}

// ---------------------------------------------------------------------------
// project: y = x - Q h (h from partials); optional norm^2 partials of y.
// Columns in batches of kColBatch loaded together, the last ncols % kColBatch
// in batches of 4, 2 and 1 (no column is read twice).
// ---------------------------------------------------------------------------
template <int V>
__global__ void __launch_bounds__(kStreamThreads)
    k_project(const float* __restrict__ Q, int64_t ldc, const float* __restrict__ x,
              float* __restrict__ y, int64_t n, int ncols, const float* __restrict__ hpart,
              int pstride, float* __restrict__ norm2part, float* __restrict__ hout,
              int64_t hout_stride, const float* __restrict__ hsub, int64_t hsub_stride) {
  constexpr int NE = Layout<V>::NE;
  extern __shared__ float hs[];
  __shared__ float red[kWaves];
  const int64_t p = blockIdx.y, b = blockIdx.x, nb = gridDim.x;
  const int t = threadIdx.x;
  const int64_t base = b * kChunk;
  load_coeffs(hpart, p, nb, ncols, pstride, hs, hout, hout_stride, b == 0, hsub, hsub_stride);

  float s[NE][V];
  int64_t off[NE];
#pragma unroll
  for (int e = 0; e < NE; ++e) {
    const int64_t i = base + (int64_t(e) * kStreamThreads + t) * V;
    off[e] = i < n ? i : n - V;  // V divides n
#pragma unroll
    for (int v = 0; v < V; ++v) s[e][v] = 0.f;
  }

  const float* qp = Q + p * n;
  // columns [c0, c0 + B), loaded together, added in column order
  auto batch = [&](auto bsz, int c0) {
    constexpr int B = decltype(bsz)::value;
    float qv[B][NE][V];
#pragma unroll
    for (int u = 0; u < B; ++u) {
      const float* q = qp + int64_t(c0 + u) * ldc;
#pragma unroll
      for (int e = 0; e < NE; ++e) VecIO<V>::load(q + off[e], qv[u][e]);
    }
#pragma unroll
    for (int u = 0; u < B; ++u) {
      const float hc = hs[c0 + u];
#pragma unroll
      for (int e = 0; e < NE; ++e)
#pragma unroll
        for (int v = 0; v < V; ++v) s[e][v] = fmaf(qv[u][e][v], hc, s[e][v]);
    }
  };
  int c0 = 0;
  for (; c0 + kColBatch <= ncols; c0 += kColBatch) batch(std::integral_constant<int, kColBatch>{}, c0);
  if (c0 + 4 <= ncols) { batch(std::integral_constant<int, 4>{}, c0); c0 += 4; }
  if (c0 + 2 <= ncols) { batch(std::integral_constant<int, 2>{}, c0); c0 += 2; }
  if (c0 < ncols) batch(std::integral_constant<int, 1>{}, c0);

  float acc = 0.f;
#pragma unroll
  for (int e = 0; e < NE; ++e) {
    const int64_t i = base + (int64_t(e) * kStreamThreads + t) * V;
    if (i < n) {
      float xv[V];
      VecIO<V>::load(x + p * n + i, xv);
#pragma unroll
      for (int v = 0; v < V; ++v) {
        xv[v] = xv[v] - s[e][v];
        acc = fmaf(xv[v], xv[v], acc);
      }
      VecIO<V>::store(y + p * n + i, xv);
    }
  }
  if (norm2part) {
    const float tot = block_sum(acc, red);
    if (t == 0) norm2part[p * nb + b] = tot;
  }
}

// k_project for rows that start at every alignment (see k_dots_shifted).
__global__ void __launch_bounds__(kStreamThreads)
    k_project_shifted(const float* __restrict__ Q, int64_t ldc, const float* __restrict__ x,
                      float* __restrict__ y, int64_t n, int ncols,
                      const float* __restrict__ hpart, int pstride,
                      float* __restrict__ norm2part, float* __restrict__ hout,
                      int64_t hout_stride, const float* __restrict__ hsub, int64_t hsub_stride) {
  constexpr int NE = Layout<4>::NE;
  extern __shared__ float hs[];
  __shared__ float red[kWaves];
  const int64_t p = blockIdx.y, b = blockIdx.x, nb = gridDim.x;
  const int t = threadIdx.x;
  const ShiftedRow row(p, n, b, t);
  load_coeffs(hpart, p, nb, ncols, pstride, hs, hout, hout_stride, b == 0, hsub, hsub_stride);

  float s[NE][4];
  int64_t off[NE];
#pragma unroll
  for (int e = 0; e < NE; ++e) {
    const int64_t i = ShiftedRow::group(p, n, b, e, t);
    off[e] = i < row.gend ? i : row.gend - 4;
#pragma unroll
    for (int v = 0; v < 4; ++v) s[e][v] = 0.f;
  }
  const int64_t ixo = row.ix >= 0 ? row.ix : 0;
  float sx = 0.f;

  const float* qp = Q + p * n;
  auto batch = [&](auto bsz, int c0) {
    constexpr int B = decltype(bsz)::value;
    float qv[B][NE][4], qh[B];
#pragma unroll
    for (int u = 0; u < B; ++u) {
      const float* q = qp + int64_t(c0 + u) * ldc;
#pragma unroll
      for (int e = 0; e < NE; ++e) VecIO<4>::load(q + off[e], qv[u][e]);
      qh[u] = q[ixo];
    }
#pragma unroll
    for (int u = 0; u < B; ++u) {
      const float hc = hs[c0 + u];
#pragma unroll
      for (int e = 0; e < NE; ++e)
#pragma unroll
        for (int v = 0; v < 4; ++v) s[e][v] = fmaf(qv[u][e][v], hc, s[e][v]);
      sx = fmaf(qh[u], hc, sx);
    }
  };
  int c0 = 0;
  for (; c0 + kColBatch <= ncols; c0 += kColBatch) batch(std::integral_constant<int, kColBatch>{}, c0);
  if (c0 + 4 <= ncols) { batch(std::integral_constant<int, 4>{}, c0); c0 += 4; }
  if (c0 + 2 <= ncols) { batch(std::integral_constant<int, 2>{}, c0); c0 += 2; }
  if (c0 < ncols) batch(std::integral_constant<int, 1>{}, c0);

  float acc = 0.f;
  const float* xr = x + p * n;
  float* yr = y + p * n;
  if (row.ix >= 0) {
    const float v = xr[row.ix] - sx;
    acc = fmaf(v, v, acc);
    yr[row.ix] = v;
  }
#pragma unroll
  for (int e = 0; e < NE; ++e) {
    const int64_t i = ShiftedRow::group(p, n, b, e, t);
    if (i < row.gend) {
      float xv[4];
      VecIO<4>::load(xr + i, xv);
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        xv[v] = xv[v] - s[e][v];
        acc = fmaf(xv[v], xv[v], acc);
      }
      VecIO<4>::store(yr + i, xv);
    }
  }
  if (norm2part) {
    const float tot = block_sum(acc, red);
    if (t == 0) norm2part[p * nb + b] = tot;
  }
}

// ---------------------------------------------------------------------------
// fused project + dots (ncols <= NC): y = x - Q h1, h2part = Q^T y.
// E consecutive elements per lane per iteration with their NC basis values
// in registers, so Q is read once for both operations; the element loop is
// not unrolled (one iteration's values live at a time), and h1 is read from
// LDS.
// ---------------------------------------------------------------------------
template <int E>
struct VecE;
template <>
struct VecE<1> {
  using T = float;
  __device__ static __forceinline__ float get(const T& v, int) { return v; }
  __device__ static __forceinline__ void set(T& v, int, float f) { v = f; }
};
template <>
struct VecE<2> {
  using T = float2;
  __device__ static __forceinline__ float get(const T& v, int e) { return e ? v.y : v.x; }
  __device__ static __forceinline__ void set(T& v, int e, float f) { (e ? v.y : v.x) = f; }
};
template <>
struct VecE<4> {
  using T = float4;
  __device__ static __forceinline__ float get(const T& v, int e) {
    return e == 0 ? v.x : e == 1 ? v.y : e == 2 ? v.z : v.w;
  }
  __device__ static __forceinline__ void set(T& v, int e, float f) {
    (e == 0 ? v.x : e == 1 ? v.y : e == 2 ? v.z : v.w) = f;
  }
};

template <int NC, int E>
__global__ void __launch_bounds__(kStreamThreads)
    k_project_dots(const float* __restrict__ Q, int64_t ldc, const float* __restrict__ x,
                   float* __restrict__ y, int64_t n, int ncols,
                   const float* __restrict__ h1part, float* __restrict__ h2part, int pstride,
                   float* __restrict__ hout, int64_t hout_stride) {
  using VT = typename VecE<E>::T;
  __shared__ __attribute__((aligned(16))) float hs[NC];
  __shared__ float red[kWaves * NC];
  const int64_t p = blockIdx.y, b = blockIdx.x, nb = gridDim.x;
  const int t = threadIdx.x;
  const int64_t base = b * kChunk;
  if (t < NC) hs[t] = 0.f;
  __syncthreads();
  load_coeffs(h1part, p, nb, ncols, pstride, hs, hout, hout_stride, b == 0);

  float acc[NC];
#pragma unroll
  for (int c = 0; c < NC; ++c) acc[c] = 0.f;

  const float* qp = Q + p * n;
#pragma unroll 1
  for (int64_t it = 0; it < kChunk / (kStreamThreads * E); ++it) {
    const int64_t i = base + (it * kStreamThreads + t) * E;
    if (i >= n) break;
    // Column offsets are formed by a chain that the compiler may not
    // re-associate: otherwise it keeps NC loop-invariant 64-bit column
    // offsets live across the loop (2 NC registers).  The chain runs on the
    // integer offset, not on the pointer, so that the loads keep the global
    // address space (a pointer through the asm becomes a generic pointer and
    // its loads flat loads, which also wait on the LDS counter).
    VT q[NC];
    int64_t coff = i;
#pragma unroll
    for (int c = 0; c < NC; ++c) {
      if (c < ncols) {
        q[c] = *reinterpret_cast<const VT*>(qp + coff);
      } else {
#pragma unroll
        for (int e = 0; e < E; ++e) VecE<E>::set(q[c], e, 0.f);
      }
      coff += ldc;
      asm volatile("" : "+v"(coff));
    }
    VT yv = *reinterpret_cast<const VT*>(x + p * n + i);
#pragma unroll
    for (int e = 0; e < E; ++e) {
      float s = 0.f;
#pragma unroll
      for (int c = 0; c < NC; ++c) s = fmaf(VecE<E>::get(q[c], e), hs[c], s);
      VecE<E>::set(yv, e, VecE<E>::get(yv, e) - s);
    }
    *reinterpret_cast<VT*>(y + p * n + i) = yv;
#pragma unroll
    for (int c = 0; c < NC; ++c)
#pragma unroll
      for (int e = 0; e < E; ++e)
        acc[c] = fmaf(VecE<E>::get(q[c], e), VecE<E>::get(yv, e), acc[c]);
  }
  block_column_sums<NC>(acc, ncols, red, h2part + (p * nb + b) * pstride);
}

// ---------------------------------------------------------------------------
// three-term step of plain Lanczos (decomp.py:254-260 / 243-245):
//   a = sum_b apart[p][b] (= q . A q), r = A q - a q - b_prev q_prev,
// plus the norm^2 partials of r; a is written to aout[p * astride] (chunk 0).
// qm (and b_prev) may be null for the first step.
// ---------------------------------------------------------------------------
template <int V>
__global__ void __launch_bounds__(kStreamThreads)
    k_three_term(const float* __restrict__ w, const float* __restrict__ q,
                 const float* __restrict__ qm, int64_t n, const float* __restrict__ apart,
                 const float* __restrict__ bprev, int64_t bstride, float* __restrict__ r,
                 float* __restrict__ norm2part, float* __restrict__ aout, int64_t astride) {
  constexpr int NE = Layout<V>::NE;
  __shared__ float red[kWaves];
  const int64_t p = blockIdx.y, b = blockIdx.x, nb = gridDim.x;
  const int t = threadIdx.x;
  const int64_t base = b * kChunk;
  float a = 0.f;
  a = ordered_sum(apart + p * nb, nb, 1);
  const float bm = qm ? bprev[p * bstride] : 0.f;
  float acc = 0.f;
#pragma unroll
  for (int e = 0; e < NE; ++e) {
    const int64_t i = base + (int64_t(e) * kStreamThreads + t) * V;
    if (i < n) {
      float wv[V], qv[V], mv[V];
      VecIO<V>::load(w + p * n + i, wv);
      VecIO<V>::load(q + p * n + i, qv);
      if (qm) {
        VecIO<V>::load(qm + p * n + i, mv);
      } else {
#pragma unroll
        for (int u = 0; u < V; ++u) mv[u] = 0.f;
      }
#pragma unroll
      for (int u = 0; u < V; ++u) {
        wv[u] = (wv[u] - a * qv[u]) - bm * mv[u];
        acc = fmaf(wv[u], wv[u], acc);
      }
      VecIO<V>::store(r + p * n + i, wv);
    }
  }
  const float tot = block_sum(acc, red);
  if (t == 0) {
    norm2part[p * nb + b] = tot;
    if (b == 0 && aout) aout[p * astride] = a;
  }
}

// BF16 hi / lo planes of V elements i .. i + V - 1 of probe p, in the layout
// of the block matvec's split pass (k_matvec.hip: tile-major, 32 columns per
// tile and probe; RBF tiles with column groups of four ordered as the MFMA
// fragments expect): hi = bf16(x), lo = bf16(x - hi), round to nearest.
__device__ __forceinline__ int plane_group(int c4, int perm) {
  return perm ? ((c4 & 4) | ((c4 & 1) << 1) | ((c4 & 2) >> 1)) : c4;
}

template <int V>
__device__ __forceinline__ void store_planes(const PlaneOut& pl, int64_t p, int64_t i,
                                             const float (&v)[V]) {
  if constexpr (V == 4) {  // i % 4 == 0: one group of four, one 8-byte store per plane
    const int64_t dst = ((i / 32) * pl.Ppad + p) * 32 + 4 * plane_group(int((i % 32) / 4), pl.perm);
    __bf16 h[4], l[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      h[u] = __bf16(v[u]);
      l[u] = __bf16(v[u] - float(h[u]));
    }
    *reinterpret_cast<uint2*>(pl.hi + dst) = __builtin_bit_cast(uint2, h);
    if (pl.lo) *reinterpret_cast<uint2*>(pl.lo + dst) = __builtin_bit_cast(uint2, l);
  } else {
#pragma unroll
    for (int u = 0; u < V; ++u) {
      const int64_t m = i + u;
      const int64_t dst =
          ((m / 32) * pl.Ppad + p) * 32 + 4 * plane_group(int((m % 32) / 4), pl.perm) + (m % 4);
      const __bf16 h = __bf16(v[u]);
      pl.hi[dst] = h;
      if (pl.lo) pl.lo[dst] = __bf16(v[u] - float(h));
    }
  }
}

// ---------------------------------------------------------------------------
// normalize: out = x / sqrt(sum_b norm2part), plus bookkeeping of the norm
// ---------------------------------------------------------------------------
constexpr int kNormVec = 2;  // float4 groups per thread

template <int V>
__global__ void __launch_bounds__(kStreamThreads)
    k_normalize(const float* __restrict__ x, float* __restrict__ out, int64_t n, int64_t nbn,
                const float* __restrict__ norm2part, float* __restrict__ norm_out,
                float* __restrict__ hslot, int64_t hslot_stride, float* __restrict__ hslot2,
                float* __restrict__ inv_out, PlaneOut pl) {
  __shared__ float nrm;
  const int64_t p = blockIdx.y;
  const int t = threadIdx.x;
  if (t == 0) {
    float s = 0.f;
    s = ordered_sum(norm2part + p * nbn, nbn, 1);
    const float len = sqrtf(s);
    nrm = len;
    if (blockIdx.x == 0) {
      if (norm_out) norm_out[p] = len;
      if (hslot) hslot[p * hslot_stride] = len;
      if (hslot2) hslot2[p * hslot_stride] = len;
      if (inv_out) inv_out[p] = 1.0f / len;
    }
  }
  __syncthreads();
  const float len = nrm;
  const int64_t per_block = int64_t(kStreamThreads) * V * kNormVec;
  const int64_t base = int64_t(blockIdx.x) * per_block;
#pragma unroll
  for (int e = 0; e < kNormVec; ++e) {
    const int64_t i = base + (int64_t(e) * kStreamThreads + t) * V;
    if (i < n) {
      float v[V];
      VecIO<V>::load(x + p * n + i, v);
#pragma unroll
      for (int u = 0; u < V; ++u) v[u] = v[u] / len;
      VecIO<V>::store(out + p * n + i, v);
      if (pl.hi) store_planes<V>(pl, p, i, v);
    }
  }
}

inline bool aligned16(const void* ptr) { return (reinterpret_cast<uintptr_t>(ptr) & 15) == 0; }

// Vector width of the streaming kernels for rows of n floats at these
// pointers: 4 (float4) when n % 4 == 0 and all are 16-byte aligned, 2 (float2)
// when n is even and all are 8-byte aligned (odd-multiple-of-2 n: every row
// then starts on an 8-byte boundary), else 1.
// The shifted float4 kernels (rows of any n >= 8): a basis stride that is a
// multiple of 4 and 16-byte aligned bases.
inline bool shifted_ok(int64_t n, int64_t ld, std::initializer_list<const void*> ptrs) {
  uintptr_t bits = 0;
  for (const void* q : ptrs) bits |= reinterpret_cast<uintptr_t>(q);
  return n >= 8 && ld % 4 == 0 && (bits & 15) == 0;
}

inline int stream_width(int64_t n, int64_t ld, std::initializer_list<const void*> ptrs) {
  uintptr_t bits = 0;
  for (const void* q : ptrs) bits |= reinterpret_cast<uintptr_t>(q);
  if (n % 4 == 0 && ld % 4 == 0 && (bits & 15) == 0) return 4;
  if (n % 2 == 0 && ld % 2 == 0 && (bits & 7) == 0) return 2;
  return 1;
}

template <int V>
hipError_t launch_dots_v(hipStream_t s, const float* Q, int64_t ldc, const float* x, int64_t P,
                         int64_t n, int ncols, float* partial, int pstride) {
  const dim3 grid(unsigned(num_chunks(n)), unsigned(P));
  const dim3 block(kStreamThreads);
  const MvSplitSum sm;
  if (ncols <= 8)
    hipLaunchKernelGGL((k_dots<V, 8, false>), grid, block, 0, s, Q, ldc, x, nullptr, n, ncols,
                       partial, pstride, sm);
  else if (ncols <= 16)
    hipLaunchKernelGGL((k_dots<V, 16, false>), grid, block, 0, s, Q, ldc, x, nullptr, n, ncols,
                       partial, pstride, sm);
  else if (ncols <= 32)
    hipLaunchKernelGGL((k_dots<V, 32, false>), grid, block, 0, s, Q, ldc, x, nullptr, n, ncols,
                       partial, pstride, sm);
  else
    hipLaunchKernelGGL((k_dots<V, 64, false>), grid, block, 0, s, Q, ldc, x, nullptr, n, ncols,
                       partial, pstride, sm);
  return hipGetLastError();
}

}  // namespace

hipError_t launch_dots(hipStream_t s, const float* Q, int64_t ldc, const float* x, int64_t P,
                       int64_t n, int ncols, float* partial, int pstride) {
  const int V = Q ? stream_width(n, ldc, {x, Q}) : stream_width(n, n, {x});
  if (Q == nullptr) {
    const dim3 grid(unsigned(num_chunks(n)), unsigned(P));
    if (V == 4)
      hipLaunchKernelGGL(k_sumsq<4>, grid, dim3(kStreamThreads), 0, s, x, n, partial);
    else if (V == 2)
      hipLaunchKernelGGL(k_sumsq<2>, grid, dim3(kStreamThreads), 0, s, x, n, partial);
    else
      hipLaunchKernelGGL(k_sumsq<1>, grid, dim3(kStreamThreads), 0, s, x, n, partial);
    return hipGetLastError();
  }
  if (V == 4) return launch_dots_v<4>(s, Q, ldc, x, P, n, ncols, partial, pstride);
  if (shifted_ok(n, ldc, {x, Q})) {  // measured faster than the float2 kernels for even n
    const dim3 grid(unsigned(num_chunks(n)), unsigned(P));
    const dim3 block(kStreamThreads);
    if (ncols <= 8)
      hipLaunchKernelGGL((k_dots_shifted<8, false>), grid, block, 0, s, Q, ldc, x, nullptr, n,
                         ncols, partial, pstride, MvSplitSum{});
    else if (ncols <= 16)
      hipLaunchKernelGGL((k_dots_shifted<16, false>), grid, block, 0, s, Q, ldc, x, nullptr, n,
                         ncols, partial, pstride, MvSplitSum{});
    else if (ncols <= 32)
      hipLaunchKernelGGL((k_dots_shifted<32, false>), grid, block, 0, s, Q, ldc, x, nullptr, n,
                         ncols, partial, pstride, MvSplitSum{});
    else
      hipLaunchKernelGGL((k_dots_shifted<64, false>), grid, block, 0, s, Q, ldc, x, nullptr, n,
                         ncols, partial, pstride, MvSplitSum{});
    return hipGetLastError();
  }
  if (V == 2) return launch_dots_v<2>(s, Q, ldc, x, P, n, ncols, partial, pstride);
  return launch_dots_v<1>(s, Q, ldc, x, P, n, ncols, partial, pstride);
}

hipError_t launch_dots_sum(hipStream_t s, const float* Q, int64_t ldc, const MvSplitSum& sum,
                           float* x, int64_t P, int64_t n, int ncols, float* partial, int pstride) {
  const dim3 grid(unsigned(num_chunks(n)), unsigned(P));
  const dim3 block(kStreamThreads);
  const std::initializer_list<const void*> ptrs = {x, Q, sum.partial, sum.nugget ? sum.V : x};
  const int V = stream_width(n, ldc, ptrs);
#define SLQ_DS(VV, NC)                                                                          \
  hipLaunchKernelGGL((k_dots<VV, NC, true>), grid, block, 0, s, Q, ldc, nullptr, x, n, ncols, \
                     partial, pstride, sum)
#define SLQ_DSH(NC)                                                                          \
  hipLaunchKernelGGL((k_dots_shifted<NC, true>), grid, block, 0, s, Q, ldc, nullptr, x, n, \
                     ncols, partial, pstride, sum)
#define SLQ_DS_NC(M)         \
  if (ncols <= 8) M(8);      \
  else if (ncols <= 16) M(16); \
  else if (ncols <= 32) M(32); \
  else M(64)
  if (V == 4) {
#define SLQ_DS4(NC) SLQ_DS(4, NC)
    SLQ_DS_NC(SLQ_DS4);
#undef SLQ_DS4
  } else if (shifted_ok(n, ldc, ptrs)) {
    SLQ_DS_NC(SLQ_DSH);
  } else if (V == 2) {
#define SLQ_DS2(NC) SLQ_DS(2, NC)
    SLQ_DS_NC(SLQ_DS2);
#undef SLQ_DS2
  } else {
#define SLQ_DS1(NC) SLQ_DS(1, NC)
    SLQ_DS_NC(SLQ_DS1);
#undef SLQ_DS1
  }
#undef SLQ_DS_NC
#undef SLQ_DSH
#undef SLQ_DS
  return hipGetLastError();
}

hipError_t launch_project(hipStream_t s, const float* Q, int64_t ldc, const float* x, float* y,
                          int64_t P, int64_t n, int ncols, const float* hpart, int pstride,
                          float* norm2part, float* hout, int64_t hout_stride, const float* hsub,
                          int64_t hsub_stride) {
  const dim3 grid(unsigned(num_chunks(n)), unsigned(P));
  const size_t shmem = size_t(ncols > 0 ? ncols : 1) * sizeof(float);
  const int V = stream_width(n, ldc, {x, y, Q});
  if (V == 4)
    hipLaunchKernelGGL(k_project<4>, grid, dim3(kStreamThreads), shmem, s, Q, ldc, x, y, n, ncols,
                       hpart, pstride, norm2part, hout, hout_stride, hsub, hsub_stride);
  else if (shifted_ok(n, ldc, {x, y, Q}))  // measured faster than the float2 kernel for even n
    hipLaunchKernelGGL(k_project_shifted, grid, dim3(kStreamThreads), shmem, s, Q, ldc, x, y, n,
                       ncols, hpart, pstride, norm2part, hout, hout_stride, hsub, hsub_stride);
  else if (V == 2)
    hipLaunchKernelGGL(k_project<2>, grid, dim3(kStreamThreads), shmem, s, Q, ldc, x, y, n, ncols,
                       hpart, pstride, norm2part, hout, hout_stride, hsub, hsub_stride);
  else
    hipLaunchKernelGGL(k_project<1>, grid, dim3(kStreamThreads), shmem, s, Q, ldc, x, y, n, ncols,
                       hpart, pstride, norm2part, hout, hout_stride, hsub, hsub_stride);
  return hipGetLastError();
}

hipError_t launch_project_dots(hipStream_t s, const float* Q, int64_t ldc, const float* x,
                               float* y, int64_t P, int64_t n, int ncols, const float* h1part,
                               float* h2part, int pstride, float* hout, int64_t hout_stride) {
  const dim3 grid(unsigned(num_chunks(n)), unsigned(P));
  const dim3 block(kStreamThreads);
  const bool v2 = (n % 2 == 0) && ldc % 2 == 0 && ((reinterpret_cast<uintptr_t>(Q) |
                   reinterpret_cast<uintptr_t>(x) | reinterpret_cast<uintptr_t>(y)) & 7) == 0;
#define SLQ_PD(NC, E)                                                                        \
  hipLaunchKernelGGL((k_project_dots<NC, E>), grid, block, 0, s, Q, ldc, x, y, n, ncols, h1part, \
                     h2part, pstride, hout, hout_stride)
  if (ncols <= 8) {
    if (v2) SLQ_PD(8, 2); else SLQ_PD(8, 1);
  } else if (ncols <= 16) {
    if (v2) SLQ_PD(16, 2); else SLQ_PD(16, 1);
  } else if (ncols <= 32) {
    if (v2) SLQ_PD(32, 2); else SLQ_PD(32, 1);
  } else if (ncols <= kFusedMaxCols) {
    if (v2) SLQ_PD(kFusedMaxCols, 2); else SLQ_PD(kFusedMaxCols, 1);
  } else {
    return hipErrorInvalidValue;
  }
#undef SLQ_PD
  return hipGetLastError();
}

hipError_t launch_three_term(hipStream_t s, const float* w, const float* q, const float* qm,
                             int64_t P, int64_t n, const float* apart, const float* bprev,
                             int64_t bstride, float* r, float* norm2part, float* aout,
                             int64_t astride) {
  const dim3 grid(unsigned(num_chunks(n)), unsigned(P));
  const bool vec = (n % 4 == 0) && aligned16(w) && aligned16(q) && aligned16(r) &&
                   (!qm || aligned16(qm));
  if (vec)
    hipLaunchKernelGGL(k_three_term<4>, grid, dim3(kStreamThreads), 0, s, w, q, qm, n, apart,
                       bprev, bstride, r, norm2part, aout, astride);
  else
    hipLaunchKernelGGL(k_three_term<1>, grid, dim3(kStreamThreads), 0, s, w, q, qm, n, apart,
                       bprev, bstride, r, norm2part, aout, astride);
  return hipGetLastError();
}

// out[p][i] = scale[p] * x[p][i]
template <int V>
__global__ void __launch_bounds__(kStreamThreads)
    k_scale_rows(const float* __restrict__ x, const float* __restrict__ scale, int64_t sstride,
                 int64_t n, float* __restrict__ out) {
  const int64_t p = blockIdx.y;
  const int64_t i = (int64_t(blockIdx.x) * kStreamThreads + threadIdx.x) * V;
  if (i >= n) return;
  const float sc = scale[p * sstride];
  float v[V];
  VecIO<V>::load(x + p * n + i, v);
#pragma unroll
  for (int u = 0; u < V; ++u) v[u] = sc * v[u];
  VecIO<V>::store(out + p * n + i, v);
}

hipError_t launch_scale_rows(hipStream_t s, const float* x, const float* scale, int64_t sstride,
                             int64_t P, int64_t n, float* out) {
  const bool vec = (n % 4 == 0) && aligned16(x) && aligned16(out);
  const int V = vec ? 4 : 1;
  const int64_t per_block = int64_t(kStreamThreads) * V;
  const dim3 grid(unsigned((n + per_block - 1) / per_block), unsigned(P));
  if (vec)
    hipLaunchKernelGGL(k_scale_rows<4>, grid, dim3(kStreamThreads), 0, s, x, scale, sstride, n,
                       out);
  else
    hipLaunchKernelGGL(k_scale_rows<1>, grid, dim3(kStreamThreads), 0, s, x, scale, sstride, n,
                       out);
  return hipGetLastError();
}

// out[0 .. count) = 0 (a kernel rather than hipMemsetAsync, so that the
// clear is an ordinary kernel node when a caller captures the stream)
__global__ void __launch_bounds__(kStreamThreads)
    k_fill_zero(float* __restrict__ out, int64_t count) {
  const int64_t stride = int64_t(gridDim.x) * kStreamThreads;
  for (int64_t i = int64_t(blockIdx.x) * kStreamThreads + threadIdx.x; i < count; i += stride)
    out[i] = 0.f;
}

hipError_t launch_fill_zero(hipStream_t s, float* out, int64_t count) {
  if (count <= 0) return hipSuccess;
  int64_t blocks = (count + kStreamThreads - 1) / kStreamThreads;
  if (blocks > 2048) blocks = 2048;
  hipLaunchKernelGGL(k_fill_zero, dim3(unsigned(blocks)), dim3(kStreamThreads), 0, s, out, count);
  return hipGetLastError();
}

hipError_t launch_normalize(hipStream_t s, const float* x, float* out, int64_t P, int64_t n,
                            const float* norm2part, float* norm_out, float* hslot,
                            int64_t hslot_stride, float* inv_out, float* hslot2,
                            const PlaneOut* planes) {
  const PlaneOut pl = planes ? *planes : PlaneOut{};
  const int V = stream_width(n, n, {x, out});
  const int64_t per_block = int64_t(kStreamThreads) * V * kNormVec;
  const dim3 grid(unsigned((n + per_block - 1) / per_block), unsigned(P));
  const int64_t nbn = num_chunks(n);
  if (V == 4)
    hipLaunchKernelGGL(k_normalize<4>, grid, dim3(kStreamThreads), 0, s, x, out, n, nbn, norm2part,
                       norm_out, hslot, hslot_stride, hslot2, inv_out, pl);
  else if (V == 2)
    hipLaunchKernelGGL(k_normalize<2>, grid, dim3(kStreamThreads), 0, s, x, out, n, nbn, norm2part,
                       norm_out, hslot, hslot_stride, hslot2, inv_out, pl);
  else
    hipLaunchKernelGGL(k_normalize<1>, grid, dim3(kStreamThreads), 0, s, x, out, n, nbn, norm2part,
                       norm_out, hslot, hslot_stride, hslot2, inv_out, pl);
  return hipGetLastError();
}


// ---------------------------------------------------------------------------
// Plain Lanczos adjoint (decomp.py:268-348): reverse step i in two passes.
// k_ladj_lambda: lambda = -xi / b + mu x_{i+1} + nu x_i with
//   mu = db - lambda_plus . x_i + x_{i+1} . xi / b,  nu = da + x_i . xi / b
// from the step's partial dots [lambda_plus . x_i, x_{i+1} . xi, x_i . xi].
// k_ladj_xi: xi' = -A lambda + a lambda + b lambda_plus - b nu x_{i+1}
//   (b x_{i+1} is the residual R at the last step), fused with the partial
//   dots of step i - 1: [lambda . x_{i-1}, x_i . xi', x_{i-1} . xi'].
// ---------------------------------------------------------------------------
namespace {

__device__ __forceinline__ void ladj_coeffs(const LanczosAdjStep& st, int64_t p, int64_t nb,
                                            float& b, float& mu, float& nu) {
  b = st.b[p * st.bstride];
  float d0 = 0.f, d1 = 0.f, d2 = 0.f;
  if (st.part) {
    const float* pp = st.part + p * nb * 3;
    d0 = ordered_sum(pp + 0, nb, 3);
    d1 = ordered_sum(pp + 1, nb, 3);
    d2 = ordered_sum(pp + 2, nb, 3);
  }
  const int64_t dbs = st.dbstride ? st.dbstride : st.sstride;
  const float db =
      st.db1 ? st.db1[p * dbs] + (st.db2 ? st.db2[p * dbs] : 0.f) : 0.f;
  mu = db - d0 + d1 / b;
  nu = st.da[p * (st.dastride ? st.dastride : st.sstride)] + d2 / b;
}

template <int V>
__global__ void __launch_bounds__(kStreamThreads)
    k_ladj_lambda(const float* __restrict__ xi, const float* __restrict__ x,
                  const float* __restrict__ xplus, int64_t n, LanczosAdjStep st,
                  float* __restrict__ lam) {
  constexpr int NE = Layout<V>::NE;
  const int64_t p = blockIdx.y, nb = gridDim.x;
  const int t = threadIdx.x;
  const int64_t base = int64_t(blockIdx.x) * kChunk;
  float b, mu, nu;
  ladj_coeffs(st, p, nb, b, mu, nu);
  const float cmu = st.xplus_res ? mu / b : mu;
#pragma unroll
  for (int e = 0; e < NE; ++e) {
    const int64_t i = base + (int64_t(e) * kStreamThreads + t) * V;
    if (i < n) {
      float xv[V], out[V];
      VecIO<V>::load(x + p * n + i, xv);
#pragma unroll
      for (int u = 0; u < V; ++u) out[u] = nu * xv[u];
      if (xi) {
        float zv[V];
        VecIO<V>::load(xi + p * n + i, zv);
#pragma unroll
        for (int u = 0; u < V; ++u) out[u] -= zv[u] / b;
      }
      if (xplus) {
        float yv[V];
        VecIO<V>::load(xplus + p * n + i, yv);
#pragma unroll
        for (int u = 0; u < V; ++u) out[u] = fmaf(cmu, yv[u], out[u]);
      }
      VecIO<V>::store(lam + p * n + i, out);
    }
  }
}

template <int V>
__global__ void __launch_bounds__(kStreamThreads)
    k_ladj_xi(const float* __restrict__ W, const float* __restrict__ lam,
              const float* __restrict__ lam_plus, const float* __restrict__ xplus,
              int xplus_is_residual, const float* __restrict__ x_prev,
              const float* __restrict__ x, int64_t n, LanczosAdjStep st,
              float* __restrict__ xi_out, float* __restrict__ part_out) {
  constexpr int NE = Layout<V>::NE;
  __shared__ float red[kWaves * 4];
  const int64_t p = blockIdx.y, nb = gridDim.x;
  const int t = threadIdx.x;
  const int64_t base = int64_t(blockIdx.x) * kChunk;
  float b, mu, nu;
  ladj_coeffs(st, p, nb, b, mu, nu);
  const float a = st.a[p * st.sstride];
  const float cx = xplus_is_residual ? nu : b * nu;
  float acc[4] = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int e = 0; e < NE; ++e) {
    const int64_t i = base + (int64_t(e) * kStreamThreads + t) * V;
    if (i < n) {
      float wv[V], lv[V], yv[V], out[V];
      VecIO<V>::load(W + p * n + i, wv);
      VecIO<V>::load(lam + p * n + i, lv);
      VecIO<V>::load(xplus + p * n + i, yv);
#pragma unroll
      for (int u = 0; u < V; ++u) out[u] = fmaf(a, lv[u], -wv[u]);
      if (lam_plus) {
        float mv[V];
        VecIO<V>::load(lam_plus + p * n + i, mv);
#pragma unroll
        for (int u = 0; u < V; ++u) out[u] = fmaf(b, mv[u], out[u]);
      }
#pragma unroll
      for (int u = 0; u < V; ++u) out[u] = fmaf(-cx, yv[u], out[u]);
      if (st.dx) {
        float dv[V];
        VecIO<V>::load(st.dx + p * n + i, dv);
#pragma unroll
        for (int u = 0; u < V; ++u) out[u] -= dv[u];
      }
      VecIO<V>::store(xi_out + p * n + i, out);
      if (part_out) {
        float xv[V], pv[V];
        VecIO<V>::load(x + p * n + i, xv);
        VecIO<V>::load(x_prev + p * n + i, pv);
#pragma unroll
        for (int u = 0; u < V; ++u) {
          acc[0] = fmaf(lv[u], pv[u], acc[0]);
          acc[1] = fmaf(xv[u], out[u], acc[1]);
          acc[2] = fmaf(pv[u], out[u], acc[2]);
        }
      }
    }
  }
  if (part_out) block_column_sums<4>(acc, 3, red, part_out + (p * nb + blockIdx.x) * 3);
}

// xi = -dxk and the partial dots [0, (R / blast) . xi, xkm1 . xi] (scalar).
__global__ void __launch_bounds__(kStreamThreads)
    k_ladj_init(const float* __restrict__ dxk, const float* __restrict__ R,
                const float* __restrict__ blast, const float* __restrict__ xkm1, int64_t n,
                float* __restrict__ xi, float* __restrict__ part_out) {
  __shared__ float red[kWaves * 4];
  const int64_t p = blockIdx.y, nb = gridDim.x;
  const int t = threadIdx.x;
  const int64_t base = int64_t(blockIdx.x) * kChunk;
  const float inv_b = 1.f / blast[p];
  float acc[4] = {0.f, 0.f, 0.f, 0.f};
  for (int64_t i = base + t; i < base + kChunk && i < n; i += kStreamThreads) {
    const int64_t off = p * n + i;
    const float z = dxk ? -dxk[off] : 0.f;
    xi[off] = z;
    acc[1] = fmaf(R[off] * inv_b, z, acc[1]);
    acc[2] = fmaf(xkm1[off], z, acc[2]);
  }
  block_column_sums<4>(acc, 3, red, part_out + (p * nb + blockIdx.x) * 3);
}

// partial dots of xi . x0 (one column), then dv = c ((xi . x0) x0 - xi)
__global__ void __launch_bounds__(kStreamThreads)
    k_ladj_initvec_dot(const float* __restrict__ xi, const float* __restrict__ x0, int64_t n,
                       float* __restrict__ part) {
  __shared__ float red[kWaves];
  const int64_t p = blockIdx.y, nb = gridDim.x;
  const int t = threadIdx.x;
  const int64_t base = int64_t(blockIdx.x) * kChunk;
  float acc = 0.f;
  for (int64_t i = base + t; i < base + kChunk && i < n; i += kStreamThreads)
    acc = fmaf(xi[p * n + i], x0[p * n + i], acc);
  const float s = block_sum(acc, red);
  if (t == 0) part[p * nb + blockIdx.x] = s;
}

__global__ void __launch_bounds__(kStreamThreads)
    k_ladj_initvec(const float* __restrict__ xi, const float* __restrict__ x0,
                   const float* __restrict__ c, const float* __restrict__ part, int64_t nb,
                   int64_t n, float* __restrict__ dv) {
  const int64_t p = blockIdx.y;
  float d = 0.f;
  d = ordered_sum(part + p * nb, nb, 1);
  const float cp = c[p];
  const int64_t i = int64_t(blockIdx.x) * kStreamThreads + threadIdx.x;
  if (i < n) dv[p * n + i] = cp * fmaf(d, x0[p * n + i], -xi[p * n + i]);
}

}  // namespace

hipError_t launch_lanczos_adj_init(hipStream_t s, const float* dxk, const float* R,
                                   const float* blast, const float* xkm1, int64_t P, int64_t n,
                                   float* xi, float* part_out) {
  const dim3 grid(unsigned(num_chunks(n)), unsigned(P));
  hipLaunchKernelGGL(k_ladj_init, grid, dim3(kStreamThreads), 0, s, dxk, R, blast, xkm1, n, xi,
                     part_out);
  return hipGetLastError();
}

hipError_t launch_lanczos_adj_initvec(hipStream_t s, const float* xi, const float* x0,
                                      const float* c, int64_t P, int64_t n, float* part,
                                      float* dv) {
  const int64_t nb = num_chunks(n);
  hipLaunchKernelGGL(k_ladj_initvec_dot, dim3(unsigned(nb), unsigned(P)), dim3(kStreamThreads), 0,
                     s, xi, x0, n, part);
  const dim3 grid(unsigned((n + kStreamThreads - 1) / kStreamThreads), unsigned(P));
  hipLaunchKernelGGL(k_ladj_initvec, grid, dim3(kStreamThreads), 0, s, xi, x0, c, part, nb, n, dv);
  return hipGetLastError();
}

hipError_t launch_lanczos_adj_lambda(hipStream_t s, const float* xi, const float* x,
                                     const float* xplus, int64_t P, int64_t n,
                                     const LanczosAdjStep& st, float* lam) {
  const dim3 grid(unsigned(num_chunks(n)), unsigned(P));
  const bool vec = (n % 4 == 0) && aligned16(x) && aligned16(lam) && (!xi || aligned16(xi)) &&
                   (!xplus || aligned16(xplus));
  if (vec)
    hipLaunchKernelGGL(k_ladj_lambda<4>, grid, dim3(kStreamThreads), 0, s, xi, x, xplus, n, st,
                       lam);
  else
    hipLaunchKernelGGL(k_ladj_lambda<1>, grid, dim3(kStreamThreads), 0, s, xi, x, xplus, n, st,
                       lam);
  return hipGetLastError();
}

hipError_t launch_lanczos_adj_xi(hipStream_t s, const float* W, const float* lam,
                                 const float* lam_plus, const float* xplus, bool xplus_is_residual,
                                 const float* x_prev, const float* x, int64_t P, int64_t n,
                                 const LanczosAdjStep& st, float* xi, float* part_out) {
  const dim3 grid(unsigned(num_chunks(n)), unsigned(P));
  const bool vec = (n % 4 == 0) && aligned16(W) && aligned16(lam) && aligned16(xplus) &&
                   aligned16(xi) && (!lam_plus || aligned16(lam_plus)) &&
                   (!part_out || (aligned16(x) && aligned16(x_prev))) &&
                   (!st.dx || aligned16(st.dx));
  const int res = xplus_is_residual ? 1 : 0;
  if (vec)
    hipLaunchKernelGGL(k_ladj_xi<4>, grid, dim3(kStreamThreads), 0, s, W, lam, lam_plus, xplus,
                       res, x_prev, x, n, st, xi, part_out);
  else
    hipLaunchKernelGGL(k_ladj_xi<1>, grid, dim3(kStreamThreads), 0, s, W, lam, lam_plus, xplus,
                       res, x_prev, x, n, st, xi, part_out);
  return hipGetLastError();
}

}  // namespace slq
