// Grouped kernels of the serving loop: many small eigenproblems advance
// together, one launch per operation for the whole batch.
//
// Each launch reads a device table (header + items). The grids are fixed, the
// work units (row tiles, row chunks, items) are handed out by a loop over the
// table, so a graph captured once replays for any batch. A work unit's
// arithmetic depends only on its item and its index inside the item, and
// partial sums are combined in chunk order: a problem gets bitwise the same
// results whatever else is in the batch.
#include "fe_common.hpp"

#include <algorithm>

namespace fe {

constexpr int GRP_GRID = 2048;   // persistent grid of the tiled kernels (8 per CU)
constexpr int GRP_SUB = 32;      // rows staged per LDS pass in gram / resid

template <typename Item>
__device__ __forceinline__ const Item* grp_items(const void* table) {
  return reinterpret_cast<const Item*>(reinterpret_cast<const char*>(table) + sizeof(fe_grp_header));
}

// item holding work unit t: the last item whose tile0 <= t (items sorted)
template <typename Item>
__device__ __forceinline__ int grp_find(const Item* it, int nitems, int64_t t) {
  int lo = 0, hi = nitems - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (it[mid].tile0 <= t) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

// ------------------------------------------------------------------ SpMM

template <typename T, int LPR>
__device__ void grp_spmm_tile(const fe_grp_spmm& it, int64_t lt) {
  constexpr int V = 16 / (int)sizeof(T);
  constexpr int UNR = 8;
  const int lane = threadIdx.x & 63;
  const int sub = lane % LPR;
  const int gbase = lane - sub;
  const int64_t row = lt * (256 / LPR) + threadIdx.x / LPR;
  const int c = sub * V;
  const int k = it.k;
  const bool rowok = row < it.n;
  const int beg = rowok ? it.rowptr[row] : 0;
  const int end = rowok ? it.rowptr[row + 1] : 0;
  const T* X = static_cast<const T*>(it.X);
  const T* vals = static_cast<const T*>(it.vals);
  T acc[V];
#pragma unroll
  for (int e = 0; e < V; ++e) acc[e] = T(0);
  for (int p = beg; p < end; p += LPR) {
    const int q = p + sub;
    const int mycol = q < end ? it.colind[q] : 0;
    const T myval = q < end ? vals[q] : T(0);
    const int cnt = end - p < LPR ? end - p : LPR;
    for (int t = 0; t < cnt; t += UNR) {
      T xv[UNR][V];
      T vv[UNR];
#pragma unroll
      for (int u = 0; u < UNR; ++u) {
        const int src = gbase + ((t + u) & (LPR - 1));
        const int cc = __shfl(mycol, src, 64);
        vv[u] = __shfl(myval, src, 64);
        if (t + u < cnt && c < k) {
          const T* xp = X + (int64_t)cc * it.ldx + c;
          if (c + V <= k) {
            const float4 f = *reinterpret_cast<const float4*>(xp);
            __builtin_memcpy(xv[u], &f, 16);
          } else {
#pragma unroll
            for (int e = 0; e < V; ++e) xv[u][e] = c + e < k ? xp[e] : T(0);
          }
        }
      }
#pragma unroll
      for (int u = 0; u < UNR; ++u)
        if (t + u < cnt && c < k)
#pragma unroll
          for (int e = 0; e < V; ++e) acc[e] += vv[u] * xv[u][e];
    }
  }
  if (!rowok || c >= k) return;
  const T alpha = static_cast<T>(it.alpha), beta = static_cast<T>(it.beta), gamma = static_cast<T>(it.gamma);
  const T* Z = static_cast<const T*>(it.Z);
  T* Y = static_cast<T*>(it.Y);
#pragma unroll
  for (int e = 0; e < V; ++e) {
    if (c + e >= k) break;
    T o = alpha * acc[e];
    if (beta != T(0)) o += beta * X[row * it.ldx + c + e];
    if (Z && gamma != T(0)) o += gamma * Z[row * it.ldz + c + e];
    Y[row * it.ldy + c + e] = o;
  }
}

__global__ __launch_bounds__(256) void k_grp_spmm(const void* table) {
  const fe_grp_header* h = static_cast<const fe_grp_header*>(table);
  const fe_grp_spmm* items = grp_items<fe_grp_spmm>(table);
  const int nitems = h->nitems;
  const int64_t ntiles = h->ntiles;
  for (int64_t t = blockIdx.x; t < ntiles; t += gridDim.x) {
    const int i = grp_find(items, nitems, t);
    const fe_grp_spmm& it = items[i];
    const int64_t lt = t - it.tile0;
    if (it.prec == FE_F64) {
      switch (it.lpr) {
        case 1: grp_spmm_tile<double, 1>(it, lt); break;
        case 2: grp_spmm_tile<double, 2>(it, lt); break;
        case 4: grp_spmm_tile<double, 4>(it, lt); break;
        case 8: grp_spmm_tile<double, 8>(it, lt); break;
        case 16: grp_spmm_tile<double, 16>(it, lt); break;
        default: grp_spmm_tile<double, 32>(it, lt); break;
      }
    } else {
      switch (it.lpr) {
        case 1: grp_spmm_tile<float, 1>(it, lt); break;
        case 2: grp_spmm_tile<float, 2>(it, lt); break;
        case 4: grp_spmm_tile<float, 4>(it, lt); break;
        case 8: grp_spmm_tile<float, 8>(it, lt); break;
        default: grp_spmm_tile<float, 16>(it, lt); break;
      }
    }
  }
}

// ------------------------------------------------------------------ Gram

template <typename T>
__device__ void grp_gram_chunk(const fe_grp_gram& it, int64_t lc, double* xs, double* ys) {
  const int m = it.m, p = it.p, tid = threadIdx.x;
  const int64_t r0 = lc * FE_GRP_CHUNK;
  const int64_t r1 = r0 + FE_GRP_CHUNK < it.n ? r0 + FE_GRP_CHUNK : it.n;
  const T* X = static_cast<const T*>(it.X);
  const T* Y = static_cast<const T*>(it.Y);
  // thread -> 4 x 4 block of the 64 x 64 tile: rows 4 (tid % 16) .., columns
  // 4 (tid / 16) ..; per staged row 8 LDS values feed 16 FMAs
  const int a0 = 4 * (tid % 16), b0 = 4 * (tid / 16);
  double acc[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[i][j] = 0.0;
  for (int64_t s0 = r0; s0 < r1; s0 += GRP_SUB) {
    const int rows = static_cast<int>(r1 - s0 < GRP_SUB ? r1 - s0 : GRP_SUB);
    __syncthreads();
    for (int e = tid; e < GRP_SUB * FE_GRP_MAXCOLS; e += 256) {
      const int r = e / FE_GRP_MAXCOLS, j = e % FE_GRP_MAXCOLS;
      xs[e] = (r < rows && j < m) ? static_cast<double>(X[(s0 + r) * it.ldx + j]) : 0.0;
      ys[e] = (r < rows && j < p) ? static_cast<double>(Y[(s0 + r) * it.ldy + j]) : 0.0;
    }
    __syncthreads();
    if (a0 < m && b0 < p) {
      for (int r = 0; r < rows; ++r) {
        const double2 x01 = *reinterpret_cast<const double2*>(xs + r * FE_GRP_MAXCOLS + a0);
        const double2 x23 = *reinterpret_cast<const double2*>(xs + r * FE_GRP_MAXCOLS + a0 + 2);
        const double2 y01 = *reinterpret_cast<const double2*>(ys + r * FE_GRP_MAXCOLS + b0);
        const double2 y23 = *reinterpret_cast<const double2*>(ys + r * FE_GRP_MAXCOLS + b0 + 2);
        const double xv[4] = {x01.x, x01.y, x23.x, x23.y}, yv[4] = {y01.x, y01.y, y23.x, y23.y};
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
          for (int j = 0; j < 4; ++j) acc[i][j] += xv[i] * yv[j];
      }
    }
  }
  double* part = it.part + lc * (int64_t)m * p;
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j)
      if (a0 + i < m && b0 + j < p) part[(a0 + i) + (int64_t)(b0 + j) * m] = acc[i][j];
}

__global__ __launch_bounds__(256) void k_grp_gram_part(const void* table) {
  __shared__ __align__(16) double xs[GRP_SUB * FE_GRP_MAXCOLS];
  __shared__ __align__(16) double ys[GRP_SUB * FE_GRP_MAXCOLS];
  const fe_grp_header* h = static_cast<const fe_grp_header*>(table);
  const fe_grp_gram* items = grp_items<fe_grp_gram>(table);
  for (int64_t t = blockIdx.x; t < h->ntiles; t += gridDim.x) {
    const fe_grp_gram& it = items[grp_find(items, h->nitems, t)];
    if (it.prec == FE_F64) grp_gram_chunk<double>(it, t - it.tile0, xs, ys);
    else grp_gram_chunk<float>(it, t - it.tile0, xs, ys);
  }
}

// grid (items, 16): workgroup (i, y) sums elements 256 y .. 256 y + 255 of
// item i's m x p partials, each in chunk order
__global__ __launch_bounds__(256) void k_grp_gram_reduce(const void* table) {
  const fe_grp_header* h = static_cast<const fe_grp_header*>(table);
  if (static_cast<int>(blockIdx.x) >= h->nitems) return;
  const fe_grp_gram& it = grp_items<fe_grp_gram>(table)[blockIdx.x];
  const int64_t nch = (it.n + FE_GRP_CHUNK - 1) / FE_GRP_CHUNK;
  const int mp = it.m * it.p;
  const int e = static_cast<int>(blockIdx.y) * 256 + static_cast<int>(threadIdx.x);
  if (e >= mp) return;
  double s = 0.0;
  for (int64_t c = 0; c < nch; ++c) s += it.part[c * mp + e];
  const int a = e % it.m, b = e / it.m;
  it.G[a + b * it.ldg] = s;
}

// ------------------------------------------------------------------ residual norms

template <typename T>
__device__ void grp_resid_chunk(const fe_grp_resid& it, int64_t lc, double* red) {
  const int p = it.p, tid = threadIdx.x;
  const int j = tid % FE_GRP_MAXCOLS, g = tid / FE_GRP_MAXCOLS;  // 4 row groups
  const int64_t r0 = lc * FE_GRP_CHUNK;
  const int64_t r1 = r0 + FE_GRP_CHUNK < it.n ? r0 + FE_GRP_CHUNK : it.n;
  const T* Y = static_cast<const T*>(it.Y);
  const T* W = static_cast<const T*>(it.W);
  double sr = 0.0, sy = 0.0;
  if (j < p) {
    const double th = it.theta[j];
    for (int64_t r = r0 + g; r < r1; r += 4) {
      const double y = static_cast<double>(Y[r * it.ldy + j]);
      const double d = static_cast<double>(W[r * it.ldw + j]) - th * y;
      sr += d * d;
      sy += y * y;
    }
  }
  __syncthreads();
  red[tid] = sr;
  red[256 + tid] = sy;
  __syncthreads();
  if (g == 0 && j < p) {
    double* part = it.part + lc * 2 * (int64_t)p;
    part[j] = ((red[j] + red[64 + j]) + red[128 + j]) + red[192 + j];
    part[p + j] = ((red[256 + j] + red[320 + j]) + red[384 + j]) + red[448 + j];
  }
}

__global__ __launch_bounds__(256) void k_grp_resid_part(const void* table) {
  __shared__ double red[512];
  const fe_grp_header* h = static_cast<const fe_grp_header*>(table);
  const fe_grp_resid* items = grp_items<fe_grp_resid>(table);
  for (int64_t t = blockIdx.x; t < h->ntiles; t += gridDim.x) {
    const fe_grp_resid& it = items[grp_find(items, h->nitems, t)];
    if (it.prec == FE_F64) grp_resid_chunk<double>(it, t - it.tile0, red);
    else grp_resid_chunk<float>(it, t - it.tile0, red);
  }
}

__global__ __launch_bounds__(64) void k_grp_resid_reduce(const void* table) {
  const fe_grp_header* h = static_cast<const fe_grp_header*>(table);
  if (static_cast<int>(blockIdx.x) >= h->nitems) return;
  const fe_grp_resid& it = grp_items<fe_grp_resid>(table)[blockIdx.x];
  const int j = threadIdx.x;
  if (j >= it.p) return;
  const int64_t nch = (it.n + FE_GRP_CHUNK - 1) / FE_GRP_CHUNK;
  double sr = 0.0, sy = 0.0;
  for (int64_t c = 0; c < nch; ++c) {
    sr += it.part[c * 2 * it.p + j];
    sy += it.part[c * 2 * it.p + it.p + j];
  }
  it.rnorm[j] = sqrt(sr);
  it.ynorm[j] = sqrt(sy);
}

// ------------------------------------------------------------------ GEMM

// 64-row tile: X rows staged in LDS, C in LDS (as T), each thread one row and
// 16 output columns, sums over l in ascending order.
// C (m x p, FP64 column-major) into LDS as T, [l][j]
template <typename T>
__device__ void grp_gemm_stage_c(const fe_grp_gemm& it, unsigned char* smem) {
  constexpr int MC = FE_GRP_MAXCOLS;
  T* cs = reinterpret_cast<T*>(smem) + 64 * (MC + 1);
  for (int e = threadIdx.x; e < MC * MC; e += 256) {
    const int l = e / MC, j = e % MC;
    cs[e] = (l < it.m && j < it.p) ? static_cast<T>(it.C[l + j * it.ldc]) : T(0);
  }
}

// Y[64 rows of tile lt] = X C with C already in LDS; thread -> 4 x 4 block
// (rows 4 (tid % 16) .., columns 4 (tid / 16) ..), sums over l in order.
template <typename T>
__device__ void grp_gemm_tile(const fe_grp_gemm& it, int64_t lt, unsigned char* smem) {
  constexpr int MC = FE_GRP_MAXCOLS;
  T* xs = reinterpret_cast<T*>(smem);           // 64 x (MC + 1)
  const T* cs = xs + 64 * (MC + 1);             // MC x MC, [l][j]
  const int m = it.m, p = it.p, tid = threadIdx.x;
  const int64_t r0 = lt * 64;
  const int rows = static_cast<int>(it.n - r0 < 64 ? it.n - r0 : 64);
  const T* X = static_cast<const T*>(it.X);
  for (int e = tid; e < 64 * MC; e += 256) {
    const int r = e / MC, l = e % MC;
    xs[r * (MC + 1) + l] = (r < rows && l < m) ? X[(r0 + r) * it.ldx + l] : T(0);
  }
  __syncthreads();
  const int rb = 4 * (tid % 16), c0 = 4 * (tid / 16);
  if (rb < rows && c0 < p) {
    T acc[4][4];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int j = 0; j < 4; ++j) acc[i][j] = T(0);
    for (int l = 0; l < m; ++l) {
      T xv[4], cv[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) xv[i] = xs[(rb + i) * (MC + 1) + l];
#pragma unroll
      for (int j = 0; j < 4; ++j) cv[j] = cs[l * MC + c0 + j];
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] += xv[i] * cv[j];
    }
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      if (rb + i >= rows) break;
      T* Y = static_cast<T*>(it.Y) + (r0 + rb + i) * it.ldy;
#pragma unroll
      for (int j = 0; j < 4; ++j)
        if (c0 + j < p) Y[c0 + j] = acc[i][j];
    }
  }
  __syncthreads();
}

// each workgroup takes 4 consecutive tiles per pass and stages C again only
// when the item changes
__global__ __launch_bounds__(256) void k_grp_gemm(const void* table) {
  extern __shared__ __align__(16) unsigned char smem[];
  const fe_grp_header* h = static_cast<const fe_grp_header*>(table);
  const fe_grp_gemm* items = grp_items<fe_grp_gemm>(table);
  const int64_t groups = (h->ntiles + 3) / 4;
  for (int64_t g = blockIdx.x; g < groups; g += gridDim.x) {
    int cur = -1;
    for (int64_t t = 4 * g; t < 4 * g + 4 && t < h->ntiles; ++t) {
      const int i = grp_find(items, h->nitems, t);
      const fe_grp_gemm& it = items[i];
      if (i != cur) {
        __syncthreads();
        if (it.prec == FE_F64) grp_gemm_stage_c<double>(it, smem);
        else grp_gemm_stage_c<float>(it, smem);
        cur = i;
      }
      if (it.prec == FE_F64) grp_gemm_tile<double>(it, t - it.tile0, smem);
      else grp_gemm_tile<float>(it, t - it.tile0, smem);
    }
  }
}

// ------------------------------------------------------------------ small dense

constexpr int SM_LD = FE_GRP_MAXCOLS + 1;

// equilibrated Cholesky with shifted retries, then C = D R^{-1}
__global__ __launch_bounds__(256) void k_grp_cholinv(const void* table) {
  __shared__ double a[FE_GRP_MAXCOLS * SM_LD];
  __shared__ double d[FE_GRP_MAXCOLS];
  __shared__ int fail;
  const fe_grp_header* h = static_cast<const fe_grp_header*>(table);
  if (static_cast<int>(blockIdx.x) >= h->nitems) return;
  const fe_grp_small& it = grp_items<fe_grp_small>(table)[blockIdx.x];
  const int m = it.m, tid = threadIdx.x;
  for (int i = tid; i < m; i += 256) {
    const double g = it.A[i + i * it.lda];
    d[i] = g > 0.0 ? 1.0 / sqrt(g) : 1.0;
  }
  __syncthreads();
  double shift = 0.0;
  int tries = 0;
  for (;; ++tries) {
    for (int e = tid; e < m * m; e += 256) {
      const int i = e % m, j = e / m;
      a[i * SM_LD + j] = d[i] * it.A[i + j * it.lda] * d[j] + (i == j ? shift : 0.0);
    }
    if (tid == 0) fail = 0;
    __syncthreads();
    // upper R with D A D + s I = R^T R, stored in the upper triangle of a
    for (int k = 0; k < m; ++k) {
      if (tid == 0) {
        const double piv = a[k * SM_LD + k];
        if (!(piv > 1e-300)) fail = 1;
        a[k * SM_LD + k] = piv > 1e-300 ? sqrt(piv) : 1.0;
      }
      __syncthreads();
      const double rkk = a[k * SM_LD + k];
      for (int j = k + 1 + tid; j < m; j += 256) a[k * SM_LD + j] /= rkk;
      __syncthreads();
      for (int e = tid; e < (m - k - 1) * (m - k - 1); e += 256) {
        const int i = k + 1 + e / (m - k - 1), j = k + 1 + e % (m - k - 1);
        if (j >= i) a[i * SM_LD + j] -= a[k * SM_LD + i] * a[k * SM_LD + j];
      }
      __syncthreads();
    }
    if (!fail || tries >= 12) break;
    shift = shift == 0.0 ? 1e-14 * m : shift * 100.0;
    __syncthreads();
  }
  // C[:, c] = D R^{-1} e_c by back substitution, one column per thread
  for (int c = tid; c < m; c += 256) {
    for (int i = m - 1; i >= 0; --i) {
      double s = (i == c) ? 1.0 : 0.0;
      if (i > c) s = 0.0;
      else
        for (int k = i + 1; k <= c; ++k) s -= a[i * SM_LD + k] * it.C[k + c * it.ldc] / d[k];
      it.C[i + c * it.ldc] = d[i] * s / a[i * SM_LD + i];
    }
  }
  if (tid == 0) it.info[0] = fail ? -1 : tries;
}

// cyclic Jacobi with round-robin (tournament) pair ordering; disjoint pairs of
// a round rotate in parallel
__global__ __launch_bounds__(256) void k_grp_eig(const void* table) {
  __shared__ double a[FE_GRP_MAXCOLS * SM_LD];
  __shared__ double v[FE_GRP_MAXCOLS * SM_LD];
  __shared__ double cs[FE_GRP_MAXCOLS / 2], sn[FE_GRP_MAXCOLS / 2];
  __shared__ int pp[FE_GRP_MAXCOLS / 2], qq[FE_GRP_MAXCOLS / 2];
  __shared__ int ord[FE_GRP_MAXCOLS];
  __shared__ double red[256];
  const fe_grp_header* h = static_cast<const fe_grp_header*>(table);
  if (static_cast<int>(blockIdx.x) >= h->nitems) return;
  const fe_grp_small& it = grp_items<fe_grp_small>(table)[blockIdx.x];
  const int m = it.m, tid = threadIdx.x;
  const int mm = (m + 1) & ~1;  // players, one bye when m is odd
  for (int e = tid; e < m * m; e += 256) {
    const int i = e % m, j = e / m;
    a[i * SM_LD + j] = 0.5 * (it.A[i + j * it.lda] + it.A[j + i * it.lda]);
    v[i * SM_LD + j] = i == j ? 1.0 : 0.0;
  }
  __syncthreads();
  int sweeps = 0;
  for (; sweeps < 40; ++sweeps) {
    // off-diagonal and total Frobenius mass
    double off = 0.0, tot = 0.0;
    for (int e = tid; e < m * m; e += 256) {
      const int i = e % m, j = e / m;
      const double x = a[i * SM_LD + j];
      tot += x * x;
      if (i != j) off += x * x;
    }
    red[tid] = off;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
      if (tid < s) red[tid] += red[tid + s];
      __syncthreads();
    }
    off = red[0];
    __syncthreads();
    red[tid] = tot;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
      if (tid < s) red[tid] += red[tid + s];
      __syncthreads();
    }
    tot = red[0];
    __syncthreads();
    if (off <= 1e-30 * tot || off == 0.0) break;
    for (int round = 0; round < mm - 1; ++round) {
      if (tid < mm / 2) {
        // player 0 fixed, the others rotate: position x -> player
        auto player = [&](int x) { return x == 0 ? 0 : 1 + (x - 1 + round) % (mm - 1); };
        int p = player(tid), q = player(mm - 1 - tid);
        if (p > q) {
          const int t = p;
          p = q;
          q = t;
        }
        double c = 1.0, s = 0.0;
        if (q < m) {
          const double apq = a[p * SM_LD + q];
          if (apq != 0.0) {
            const double tau = (a[q * SM_LD + q] - a[p * SM_LD + p]) / (2.0 * apq);
            const double t = (tau >= 0.0 ? 1.0 : -1.0) / (fabs(tau) + sqrt(1.0 + tau * tau));
            c = 1.0 / sqrt(1.0 + t * t);
            s = t * c;
          }
        }
        cs[tid] = c;
        sn[tid] = s;
        pp[tid] = p;
        qq[tid] = q < m ? q : p;
      }
      __syncthreads();
      // columns: A <- A J, V <- V J
      for (int e = tid; e < (mm / 2) * m; e += 256) {
        const int k = e / m, i = e % m;
        const int p = pp[k], q = qq[k];
        if (p == q) continue;
        const double c = cs[k], s = sn[k];
        const double aip = a[i * SM_LD + p], aiq = a[i * SM_LD + q];
        a[i * SM_LD + p] = c * aip - s * aiq;
        a[i * SM_LD + q] = s * aip + c * aiq;
        const double vip = v[i * SM_LD + p], viq = v[i * SM_LD + q];
        v[i * SM_LD + p] = c * vip - s * viq;
        v[i * SM_LD + q] = s * vip + c * viq;
      }
      __syncthreads();
      // rows: A <- J^T A
      for (int e = tid; e < (mm / 2) * m; e += 256) {
        const int k = e / m, j = e % m;
        const int p = pp[k], q = qq[k];
        if (p == q) continue;
        const double c = cs[k], s = sn[k];
        const double apj = a[p * SM_LD + j], aqj = a[q * SM_LD + j];
        a[p * SM_LD + j] = c * apj - s * aqj;
        a[q * SM_LD + j] = s * apj + c * aqj;
      }
      __syncthreads();
    }
  }
  // ascending order (ties by index)
  for (int i = tid; i < m; i += 256) {
    const double wi = a[i * SM_LD + i];
    int r = 0;
    for (int j = 0; j < m; ++j) {
      const double wj = a[j * SM_LD + j];
      r += (wj < wi) || (wj == wi && j < i);
    }
    ord[r] = i;
  }
  __syncthreads();
  for (int e = tid; e < m * m; e += 256) {
    const int i = e % m, r = e / m;
    it.C[i + r * it.ldc] = v[i * SM_LD + ord[r]];
  }
  for (int r = tid; r < m; r += 256) it.w[r] = a[ord[r] * SM_LD + ord[r]];
  if (tid == 0) it.info[0] = sweeps;
}

// ------------------------------------------------------------------ random fill

__device__ __forceinline__ uint64_t splitmix64(uint64_t x) {
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  return x ^ (x >> 31);
}

template <typename T>
__global__ void k_fill_random(int64_t n, int k, T* X, int64_t ldx, uint64_t seed) {
  const int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (e >= n * k) return;
  const int64_t i = e / k;
  const int j = static_cast<int>(e % k);
  const uint64_t z = splitmix64(seed ^ splitmix64(static_cast<uint64_t>(i) * 0x100000001B3ull + j));
  const double u = static_cast<double>(z >> 11) * (1.0 / 9007199254740992.0);  // [0, 1)
  X[i * ldx + j] = static_cast<T>(2.0 * u - 1.0);
}

}  // namespace fe

using namespace fe;

namespace {

int lanes_for(int k, int prec) {
  const int v = prec == FE_F64 ? 2 : 4;
  const int runs = (k + v - 1) / v;
  int l = 1;
  while (l < runs) l <<= 1;
  return l;
}

bool aligned16(const void* p, int64_t ld, int prec) {
  return p && (reinterpret_cast<uintptr_t>(p) % 16) == 0 && ld % (prec == FE_F64 ? 2 : 4) == 0;
}

int64_t chunks_of(int n) { return (static_cast<int64_t>(n) + FE_GRP_CHUNK - 1) / FE_GRP_CHUNK; }

}  // namespace

extern "C" {

int fe_fill_random(fe_ctx ctx, int prec, int64_t n, int k, void* X, int64_t ldx, uint64_t seed) {
  FE_ARG(ctx && X && n >= 0 && k >= 0 && ldx >= k, "fe_fill_random: bad argument");
  if (n * k == 0) return 0;
  const unsigned blocks = blocks_for(n * k, 256);
  if (prec == FE_F64)
    hipLaunchKernelGGL(k_fill_random<double>, dim3(blocks), dim3(256), 0, ctx->stream, n, k, (double*)X, ldx, seed);
  else
    hipLaunchKernelGGL(k_fill_random<float>, dim3(blocks), dim3(256), 0, ctx->stream, n, k, (float*)X, ldx, seed);
  FE_CHECK_LAUNCH();
  return 0;
}

int fe_grp_spmm_plan(fe_grp_header* hdr, fe_grp_spmm* items, int nitems) {
  FE_ARG(hdr && (items || nitems == 0) && nitems >= 0, "fe_grp_spmm_plan: null argument");
  int64_t t = 0;
  for (int i = 0; i < nitems; ++i) {
    fe_grp_spmm& it = items[i];
    FE_ARG(it.prec == FE_F64 || it.prec == FE_F32, "fe_grp_spmm_plan: bad precision");
    FE_ARG(it.n >= 0 && it.k >= 1 && it.k <= FE_GRP_MAXCOLS, "fe_grp_spmm_plan: k must be 1..64");
    FE_ARG(it.rowptr && it.X && it.Y && it.X != it.Y, "fe_grp_spmm_plan: null or aliased block");
    FE_ARG(aligned16(it.X, it.ldx, it.prec) && aligned16(it.Y, it.ldy, it.prec) &&
               (!it.Z || aligned16(it.Z, it.ldz, it.prec)),
           "fe_grp_spmm_plan: blocks need 16-byte aligned rows");
    it.lpr = lanes_for(it.k, it.prec);
    it.tile0 = t;
    t += (static_cast<int64_t>(it.n) + 256 / it.lpr - 1) / (256 / it.lpr);
  }
  hdr->nitems = nitems;
  hdr->kind = FE_GRP_KIND_SPMM;
  hdr->ntiles = t;
  return 0;
}

int fe_grp_gram_plan(fe_grp_header* hdr, fe_grp_gram* items, int nitems) {
  FE_ARG(hdr && (items || nitems == 0) && nitems >= 0, "fe_grp_gram_plan: null argument");
  int64_t t = 0;
  for (int i = 0; i < nitems; ++i) {
    fe_grp_gram& it = items[i];
    FE_ARG(it.prec == FE_F64 || it.prec == FE_F32, "fe_grp_gram_plan: bad precision");
    FE_ARG(it.m >= 1 && it.m <= FE_GRP_MAXCOLS && it.p >= 1 && it.p <= FE_GRP_MAXCOLS && it.n >= 1,
           "fe_grp_gram_plan: m, p must be 1..64 and n >= 1");
    FE_ARG(it.X && it.Y && it.part && it.G && it.ldg >= it.m, "fe_grp_gram_plan: null argument");
    it.tile0 = t;
    t += chunks_of(it.n);
  }
  hdr->nitems = nitems;
  hdr->kind = FE_GRP_KIND_GRAM;
  hdr->ntiles = t;
  return 0;
}

int fe_grp_resid_plan(fe_grp_header* hdr, fe_grp_resid* items, int nitems) {
  FE_ARG(hdr && (items || nitems == 0) && nitems >= 0, "fe_grp_resid_plan: null argument");
  int64_t t = 0;
  for (int i = 0; i < nitems; ++i) {
    fe_grp_resid& it = items[i];
    FE_ARG(it.prec == FE_F64 || it.prec == FE_F32, "fe_grp_resid_plan: bad precision");
    FE_ARG(it.p >= 1 && it.p <= FE_GRP_MAXCOLS && it.n >= 1, "fe_grp_resid_plan: p must be 1..64, n >= 1");
    FE_ARG(it.Y && it.W && it.theta && it.part && it.rnorm && it.ynorm, "fe_grp_resid_plan: null argument");
    it.tile0 = t;
    t += chunks_of(it.n);
  }
  hdr->nitems = nitems;
  hdr->kind = FE_GRP_KIND_RESID;
  hdr->ntiles = t;
  return 0;
}

int fe_grp_gemm_plan(fe_grp_header* hdr, fe_grp_gemm* items, int nitems) {
  FE_ARG(hdr && (items || nitems == 0) && nitems >= 0, "fe_grp_gemm_plan: null argument");
  int64_t t = 0;
  for (int i = 0; i < nitems; ++i) {
    fe_grp_gemm& it = items[i];
    FE_ARG(it.prec == FE_F64 || it.prec == FE_F32, "fe_grp_gemm_plan: bad precision");
    FE_ARG(it.m >= 1 && it.m <= FE_GRP_MAXCOLS && it.p >= 1 && it.p <= FE_GRP_MAXCOLS && it.n >= 0,
           "fe_grp_gemm_plan: m, p must be 1..64");
    FE_ARG(it.X && it.C && it.Y && it.X != it.Y && it.ldc >= it.m, "fe_grp_gemm_plan: null or aliased argument");
    it.tile0 = t;
    t += (static_cast<int64_t>(it.n) + 63) / 64;
  }
  hdr->nitems = nitems;
  hdr->kind = FE_GRP_KIND_GEMM;
  hdr->ntiles = t;
  return 0;
}

int fe_grp_small_plan(fe_grp_header* hdr, fe_grp_small* items, int nitems) {
  FE_ARG(hdr && (items || nitems == 0) && nitems >= 0, "fe_grp_small_plan: null argument");
  for (int i = 0; i < nitems; ++i) {
    const fe_grp_small& it = items[i];
    FE_ARG(it.m >= 1 && it.m <= FE_GRP_MAXCOLS, "fe_grp_small_plan: m must be 1..64");
    FE_ARG(it.A && it.C && it.info && it.lda >= it.m && it.ldc >= it.m, "fe_grp_small_plan: null argument");
  }
  hdr->nitems = nitems;
  hdr->kind = FE_GRP_KIND_SMALL;
  hdr->ntiles = nitems;
  return 0;
}

int fe_grp_spmm_run(fe_ctx ctx, const void* table) {
  FE_ARG(ctx && table, "fe_grp_spmm_run: null argument");
  hipLaunchKernelGGL(k_grp_spmm, dim3(GRP_GRID), dim3(256), 0, ctx->stream, table);
  FE_CHECK_LAUNCH();
  return 0;
}

int fe_grp_gram_run(fe_ctx ctx, const void* table, int max_items) {
  FE_ARG(ctx && table && max_items >= 1, "fe_grp_gram_run: bad argument");
  hipLaunchKernelGGL(k_grp_gram_part, dim3(GRP_GRID), dim3(256), 0, ctx->stream, table);
  hipLaunchKernelGGL(k_grp_gram_reduce, dim3(max_items, FE_GRP_MAXCOLS * FE_GRP_MAXCOLS / 256), dim3(256), 0,
                     ctx->stream, table);
  FE_CHECK_LAUNCH();
  return 0;
}

int fe_grp_resid_run(fe_ctx ctx, const void* table, int max_items) {
  FE_ARG(ctx && table && max_items >= 1, "fe_grp_resid_run: bad argument");
  hipLaunchKernelGGL(k_grp_resid_part, dim3(GRP_GRID), dim3(256), 0, ctx->stream, table);
  hipLaunchKernelGGL(k_grp_resid_reduce, dim3(max_items), dim3(64), 0, ctx->stream, table);
  FE_CHECK_LAUNCH();
  return 0;
}

int fe_grp_gemm_run(fe_ctx ctx, const void* table) {
  FE_ARG(ctx && table, "fe_grp_gemm_run: null argument");
  const size_t lds = sizeof(double) * (64 * (FE_GRP_MAXCOLS + 1) + FE_GRP_MAXCOLS * FE_GRP_MAXCOLS);
  hipLaunchKernelGGL(k_grp_gemm, dim3(GRP_GRID), dim3(256), lds, ctx->stream, table);
  FE_CHECK_LAUNCH();
  return 0;
}

int fe_grp_cholinv_run(fe_ctx ctx, const void* table, int max_items) {
  FE_ARG(ctx && table && max_items >= 1, "fe_grp_cholinv_run: bad argument");
  hipLaunchKernelGGL(k_grp_cholinv, dim3(max_items), dim3(256), 0, ctx->stream, table);
  FE_CHECK_LAUNCH();
  return 0;
}

int fe_grp_eig_run(fe_ctx ctx, const void* table, int max_items) {
  FE_ARG(ctx && table && max_items >= 1, "fe_grp_eig_run: bad argument");
  hipLaunchKernelGGL(k_grp_eig, dim3(max_items), dim3(256), 0, ctx->stream, table);
  FE_CHECK_LAUNCH();
  return 0;
}

}  // extern "C"

// ------------------------------------------------------------------ runtime helpers

struct fe_graph_s {
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
};

extern "C" {

int fe_csr_arrays(fe_csr A, const int** rowptr, const int** colind, const double** vals64, const float** vals32) {
  FE_ARG(A, "fe_csr_arrays: null matrix");
  if (rowptr) *rowptr = A->rowptr;
  if (colind) *colind = A->colind;
  if (vals64) *vals64 = A->vals64;
  if (vals32) *vals32 = A->vals32;
  return 0;
}

int fe_memcpy_h2d_async(fe_ctx ctx, void* dst, const void* src, size_t bytes) {
  FE_ARG(ctx && (bytes == 0 || (dst && src)), "fe_memcpy_h2d_async: null argument");
  if (bytes) FE_HIP_TRY(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->stream));
  return 0;
}

int fe_memcpy_d2h_async(fe_ctx ctx, void* dst, const void* src, size_t bytes) {
  FE_ARG(ctx && (bytes == 0 || (dst && src)), "fe_memcpy_d2h_async: null argument");
  if (bytes) FE_HIP_TRY(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->stream));
  return 0;
}

int fe_host_alloc(size_t bytes, void** ptr) {
  FE_ARG(ptr, "fe_host_alloc: null argument");
  FE_HIP_TRY(hipHostMalloc(ptr, bytes ? bytes : 1, hipHostMallocDefault));
  return 0;
}

int fe_host_free(void* ptr) {
  if (ptr) FE_HIP_TRY(hipHostFree(ptr));
  return 0;
}

int fe_graph_begin(fe_ctx ctx) {
  FE_ARG(ctx, "fe_graph_begin: null context");
  FE_HIP_TRY(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
  return 0;
}

int fe_graph_end(fe_ctx ctx, fe_graph* out) {
  FE_ARG(ctx && out, "fe_graph_end: null argument");
  auto* g = new fe_graph_s;
  hipError_t e = hipStreamEndCapture(ctx->stream, &g->graph);
  if (e == hipSuccess) e = hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0);
  if (e != hipSuccess) {
    if (g->graph) (void)hipGraphDestroy(g->graph);
    delete g;
    return fail(-3, std::string("fe_graph_end: ") + hipGetErrorString(e));
  }
  *out = g;
  return 0;
}

int fe_graph_launch(fe_ctx ctx, fe_graph g) {
  FE_ARG(ctx && g, "fe_graph_launch: null argument");
  FE_HIP_TRY(hipGraphLaunch(g->exec, ctx->stream));
  return 0;
}

int fe_graph_destroy(fe_graph g) {
  if (!g) return 0;
  if (g->exec) (void)hipGraphExecDestroy(g->exec);
  if (g->graph) (void)hipGraphDestroy(g->graph);
  delete g;
  return 0;
}

}  // extern "C"
