// Fused block SpMM for gfx950: Y = alpha * A X + beta * X + gamma * Z.
//
// Blocks are row-major, so the k values that one nonzero a_ij multiplies are
// the contiguous row j of X. A group of LPR lanes (power of two, <= 64) owns
// one row of A and VEC consecutive columns per lane (16-byte vector loads);
// a 64-lane wavefront therefore holds 64/LPR rows. The group first loads up to
// LPR (column, value) pairs of its row with one coalesced load, then walks
// them, broadcasting each pair with a lane shuffle (a scalar readlane when the
// whole wavefront is one row) and gathering the 16*LPR-byte X row segment.
// A is read once per column tile; X rows are re-read from L2 by the rows that
// share them; Y is written once. Workgroups are mapped XCD-contiguously.
#include "fe_common.hpp"

namespace fe {

typedef float nt_f4 __attribute__((ext_vector_type(4)));  // 16 bytes for non-temporal access

template <typename T, int VEC>
struct vec_t;
template <>
struct vec_t<float, 4> {
  using type = float4;
};
template <>
struct vec_t<float, 2> {
  using type = float2;
};
template <>
struct vec_t<float, 1> {
  using type = float;
};
template <>
struct vec_t<double, 2> {
  using type = double2;
};
template <>
struct vec_t<double, 1> {
  using type = double;
};

template <typename T, int VEC>
__device__ __forceinline__ void load_vec(const T* p, T (&v)[VEC]) {
  if constexpr (VEC == 1) {
    v[0] = *p;
  } else {
    using V = typename vec_t<T, VEC>::type;
    const V x = *reinterpret_cast<const V*>(p);
    if constexpr (VEC == 2) {
      v[0] = x.x;
      v[1] = x.y;
    } else {
      v[0] = x.x;
      v[1] = x.y;
      v[2] = x.z;
      v[3] = x.w;
    }
  }
}

template <typename T, int VEC>
__device__ __forceinline__ void store_vec(T* p, const T (&v)[VEC]) {
  if constexpr (VEC == 1) {
    *p = v[0];
  } else {
    using V = typename vec_t<T, VEC>::type;
    V x;
    if constexpr (VEC == 2) {
      x.x = v[0];
      x.y = v[1];
    } else {
      x.x = v[0];
      x.y = v[1];
      x.z = v[2];
      x.w = v[3];
    }
    *reinterpret_cast<V*>(p) = x;
  }
}

template <int LPR>
__device__ __forceinline__ int bcast_int(int v, int src_lane) {
  if constexpr (LPR == 64) return __builtin_amdgcn_readlane(v, src_lane);
  else return __shfl(v, src_lane, 64);
}

template <int LPR, typename T>
__device__ __forceinline__ T bcast_val(T v, int src_lane) {
  if constexpr (LPR == 64) {
    if constexpr (sizeof(T) == 8) {
      const long long b = __double_as_longlong(v);
      const int lo = __builtin_amdgcn_readlane(static_cast<int>(b & 0xffffffffll), src_lane);
      const int hi = __builtin_amdgcn_readlane(static_cast<int>(b >> 32), src_lane);
      return __longlong_as_double((static_cast<long long>(hi) << 32) | static_cast<unsigned int>(lo));
    } else {
      return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), src_lane));
    }
  } else {
    return __shfl(v, src_lane, 64);
  }
}

template <typename T, int RS>
__device__ __forceinline__ void ld_runs(const T* p, int stride, bool full, int c, int k, T (&v)[RS][16 / sizeof(T)]) {
  constexpr int V = 16 / sizeof(T);
#pragma unroll
  for (int r = 0; r < RS; ++r) {
    const int cr = c + r * stride;
    if (full || cr + V <= k) {
      const float4 f = *reinterpret_cast<const float4*>(p + r * stride);
      __builtin_memcpy(&v[r][0], &f, 16);
    } else {
#pragma unroll
      for (int e = 0; e < V; ++e) v[r][e] = (cr + e < k) ? p[r * stride + e] : T(0);
    }
  }
}

// gathers in flight per lane: rows of 16+ lanes with one run, and two runs
constexpr int CSR_UNR_WIDE = 16;
constexpr int CSR_UNR_RS2 = 8;

// CSR kernel, one row per group of LPR lanes, RS 16-byte runs per lane spaced
// LPR * 16 bytes apart (each load instruction of the wavefront reads one
// contiguous 16 * LPR-byte stretch of the X row). Row tiles map to XCDs in
// contiguous ranges. Scalar variant (V16 = 1 element runs) for unaligned blocks.
//
// Latency: the epilogue's own-row X and Z loads are issued before the gathers,
// and the gathers of a row go out in batches of UNR loads; the last batch is
// a full batch with the slots past the row's end masked off, so a 27-entry
// row costs two dependent gather round trips at UNR = 16.
template <typename T, int LPR, int RS, bool VECTOR>
__global__ __launch_bounds__(256) void k_spmm_csr(
    int n, const int* __restrict__ rowptr, const int* __restrict__ colind, const T* __restrict__ vals,
    int k, const T* __restrict__ X, int64_t ldx, T* Y, int64_t ldy, T alpha, T beta,
    const T* Z, int64_t ldz, T gamma, int64_t per_xcd, int64_t ntiles) {
  constexpr int V = VECTOR ? 16 / (int)sizeof(T) : 1;
  constexpr int RPW = 64 / LPR;
  constexpr int RPB = 4 * RPW;
  constexpr int STRIDE = LPR * V;        // columns between a lane's runs
  constexpr int COLS = STRIDE * RS;      // columns per column tile
  constexpr int UNR = RS == 1 ? (LPR >= 16 ? CSR_UNR_WIDE : 8) : CSR_UNR_RS2;
  const int64_t tile = xcd_tile(blockIdx.x, per_xcd);
  if (tile >= ntiles) return;
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int64_t row = tile * RPB + wave * RPW + lane / LPR;
  const int sub = lane % LPR;
  const int gbase = lane - sub;
  const int c = blockIdx.y * COLS + sub * V;
  // the lane's runs are all complete (no column tail) -> unpredicated loads
  const bool full = c + (RS - 1) * STRIDE + V <= k;
  const bool rowok = row < n;
  const int beg = rowok ? rowptr[row] : 0;
  const int end = rowok ? rowptr[row + 1] : 0;
  // epilogue operands (own X row, Z row) of complete vector runs, loaded early
  T xe[RS][V == 1 ? 1 : V], ze[RS][V == 1 ? 1 : V];
  const bool early = VECTOR && rowok && full;
  if constexpr (VECTOR) {
    if (VECTOR && early) {
      if (beta != T(0)) ld_runs<T, RS>(X + row * ldx + c, STRIDE, true, c, k, xe);
      if (Z && gamma != T(0)) ld_runs<T, RS>(Z + row * ldz + c, STRIDE, true, c, k, ze);
    }
  }
  T acc[RS][V == 1 ? 1 : V];
#pragma unroll
  for (int r = 0; r < RS; ++r)
#pragma unroll
    for (int e = 0; e < V; ++e) acc[r][e] = T(0);
  for (int p = beg; p < end; p += LPR) {
    const int q = p + sub;
    const int mycol = q < end ? colind[q] : 0;
    const T myval = q < end ? vals[q] : T(0);
    const int cnt = end - p < LPR ? end - p : LPR;
    for (int t = 0; t < cnt; t += UNR) {
      int cc[UNR];
      T vv[UNR];
#pragma unroll
      for (int u = 0; u < UNR; ++u) {
        cc[u] = bcast_int<LPR>(mycol, gbase + ((t + u) & (LPR - 1)));
        vv[u] = bcast_val<LPR>(myval, gbase + ((t + u) & (LPR - 1)));
      }
      T xv[UNR][RS][V == 1 ? 1 : V];
#pragma unroll
      for (int u = 0; u < UNR; ++u) {
        if (t + u < cnt) {
          const T* xp = X + (int64_t)cc[u] * ldx + c;
          if constexpr (VECTOR) {
            ld_runs<T, RS>(xp, STRIDE, full, c, k, xv[u]);
          } else {
#pragma unroll
            for (int r = 0; r < RS; ++r) xv[u][r][0] = (c + r * STRIDE < k) ? xp[r * STRIDE] : T(0);
          }
        }
      }
#pragma unroll
      for (int u = 0; u < UNR; ++u)
        if (t + u < cnt)
#pragma unroll
          for (int r = 0; r < RS; ++r)
#pragma unroll
            for (int e = 0; e < V; ++e) acc[r][e] += vv[u] * xv[u][r][e];
    }
  }
  if (!rowok) return;
#pragma unroll
  for (int r = 0; r < RS; ++r) {
    const int cr = c + r * STRIDE;
    if (cr >= k) continue;
    T out[V == 1 ? 1 : V];
#pragma unroll
    for (int e = 0; e < V; ++e) out[e] = alpha * acc[r][e];
    if (VECTOR && early) {
      if (beta != T(0))
#pragma unroll
        for (int e = 0; e < V; ++e) out[e] += beta * xe[r][e];
      if (Z && gamma != T(0))
#pragma unroll
        for (int e = 0; e < V; ++e) out[e] += gamma * ze[r][e];
      nt_f4 o;  // Y streams out once: non-temporal, keeps L2 for the gathered X rows
      __builtin_memcpy(&o, out, 16);
      __builtin_nontemporal_store(o, reinterpret_cast<nt_f4*>(Y + row * ldy + cr));
    } else if (VECTOR && cr + V <= k) {
      if (beta != T(0)) {
        const float4 f = *reinterpret_cast<const float4*>(X + row * ldx + cr);
        T xv[V];
        __builtin_memcpy(xv, &f, 16);
#pragma unroll
        for (int e = 0; e < V; ++e) out[e] += beta * xv[e];
      }
      if (Z && gamma != T(0)) {
        const float4 f = *reinterpret_cast<const float4*>(Z + row * ldz + cr);
        T zv[V];
        __builtin_memcpy(zv, &f, 16);
#pragma unroll
        for (int e = 0; e < V; ++e) out[e] += gamma * zv[e];
      }
      nt_f4 o;  // Y streams out once: non-temporal, keeps L2 for the gathered X rows
      __builtin_memcpy(&o, out, 16);
      __builtin_nontemporal_store(o, reinterpret_cast<nt_f4*>(Y + row * ldy + cr));
    } else {
#pragma unroll
      for (int e = 0; e < V; ++e) {
        if (cr + e >= k) continue;
        T o = out[e];
        if (beta != T(0)) o += beta * X[row * ldx + cr + e];
        if (Z && gamma != T(0)) o += gamma * Z[row * ldz + cr + e];
        Y[row * ldy + cr + e] = o;
      }
    }
  }
}

// SELL-C-sigma variant: one wavefront per slice row group. Slices hold C rows;
// within a slice the entries are stored column by column ([w][C]), so lane
// groups of consecutive sorted rows read consecutive column/value entries.
template <typename T, int VEC, int LPR>
__global__ __launch_bounds__(256) void k_spmm_sell(
    int n, int C, const int64_t* __restrict__ sptr, const int* __restrict__ slen,
    const int* __restrict__ perm, const int* __restrict__ scol, const T* __restrict__ sval, int k,
    const T* __restrict__ X, int64_t ldx, T* Y, int64_t ldy, T alpha, T beta, const T* Z, int64_t ldz,
    T gamma, int64_t per_xcd, int64_t ntiles) {
  constexpr int RPW = 64 / LPR;
  constexpr int RPB = 4 * RPW;
  constexpr int COLS = LPR * VEC;
  const int64_t tile = xcd_tile(blockIdx.x, per_xcd);
  if (tile >= ntiles) return;
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int64_t srow = tile * RPB + wave * RPW + lane / LPR;  // sorted row position
  const int sub = lane % LPR;
  const int c = blockIdx.y * COLS + sub * VEC;
  if (srow >= n || c >= k) return;
  const int64_t slice = srow / C;
  const int within = static_cast<int>(srow - slice * C);
  const int64_t base = sptr[slice];
  const int w = slen[slice];
  T acc[VEC];
#pragma unroll
  for (int v = 0; v < VEC; ++v) acc[v] = T(0);
  int t = 0;
  for (; t + 2 <= w; t += 2) {
    const int64_t e0 = base + (int64_t)t * C + within;
    const int64_t e1 = e0 + C;
    const int c0i = scol[e0], c1i = scol[e1];
    const T v0 = sval[e0], v1 = sval[e1];
    T x0[VEC], x1[VEC];
    load_vec<T, VEC>(X + (int64_t)c0i * ldx + c, x0);
    load_vec<T, VEC>(X + (int64_t)c1i * ldx + c, x1);
#pragma unroll
    for (int v = 0; v < VEC; ++v) acc[v] += v0 * x0[v] + v1 * x1[v];
  }
  if (t < w) {
    const int64_t e0 = base + (int64_t)t * C + within;
    const int c0i = scol[e0];
    const T v0 = sval[e0];
    T x0[VEC];
    load_vec<T, VEC>(X + (int64_t)c0i * ldx + c, x0);
#pragma unroll
    for (int v = 0; v < VEC; ++v) acc[v] += v0 * x0[v];
  }
  const int64_t row = perm[srow];
  T out[VEC];
#pragma unroll
  for (int v = 0; v < VEC; ++v) out[v] = alpha * acc[v];
  if (beta != T(0)) {
    T xv[VEC];
    load_vec<T, VEC>(X + row * ldx + c, xv);
#pragma unroll
    for (int v = 0; v < VEC; ++v) out[v] += beta * xv[v];
  }
  if (Z && gamma != T(0)) {
    T zv[VEC];
    load_vec<T, VEC>(Z + row * ldz + c, zv);
#pragma unroll
    for (int v = 0; v < VEC; ++v) out[v] += gamma * zv[v];
  }
  store_vec<T, VEC>(Y + row * ldy + c, out);
}

template <typename T, int LPR, int RS, bool VECTOR>
static int launch_csr(fe_ctx ctx, fe_csr A, int k, const T* X, int64_t ldx, T* Y, int64_t ldy, T alpha,
                      T beta, const T* Z, int64_t ldz, T gamma, const T* vals) {
  constexpr int V = VECTOR ? 16 / (int)sizeof(T) : 1;
  constexpr int RPB = 4 * (64 / LPR);
  constexpr int COLS = LPR * V * RS;
  const int64_t ntiles = (A->n + RPB - 1) / RPB;
  const int64_t nb8 = (ntiles + 7) / 8 * 8;
  const unsigned ycols = static_cast<unsigned>((k + COLS - 1) / COLS);
  hipLaunchKernelGGL((k_spmm_csr<T, LPR, RS, VECTOR>), dim3((unsigned)nb8, ycols), dim3(256), 0, ctx->stream, A->n,
                     A->rowptr, A->colind, vals, k, X, ldx, Y, ldy, alpha, beta, Z, ldz, gamma, nb8 / 8, ntiles);
  FE_CHECK_LAUNCH();
  return 0;
}

template <typename T, int VEC, int LPR>
static int launch_sell(fe_ctx ctx, fe_csr A, int k, const T* X, int64_t ldx, T* Y, int64_t ldy, T alpha,
                       T beta, const T* Z, int64_t ldz, T gamma, const T* vals) {
  constexpr int RPB = 4 * (64 / LPR);
  constexpr int COLS = LPR * VEC;
  const int64_t ntiles = (A->n + RPB - 1) / RPB;
  const int64_t nb8 = (ntiles + 7) / 8 * 8;
  const unsigned ycols = static_cast<unsigned>((k + COLS - 1) / COLS);
  hipLaunchKernelGGL((k_spmm_sell<T, VEC, LPR>), dim3((unsigned)nb8, ycols), dim3(256), 0, ctx->stream, A->n,
                     A->sell_C, A->sell_ptr, A->sell_len, A->sell_perm, A->sell_col, vals, k, X, ldx, Y, ldy,
                     alpha, beta, Z, ldz, gamma, nb8 / 8, ntiles);
  FE_CHECK_LAUNCH();
  return 0;
}

// Sparse matrix-vector product (k = 1): a group of G lanes per row walks the
// row's nonzeros with coalesced loads of values and indices, gathers x, and
// reduces across the group with lane exchanges (CSR-vector layout).
template <typename T, int G>
__global__ __launch_bounds__(256) void k_spmv_csr(int n, const int* __restrict__ rowptr,
                                                  const int* __restrict__ colind, const T* __restrict__ vals,
                                                  const T* __restrict__ x, int64_t ldx, T* y, int64_t ldy, T alpha,
                                                  T beta, const T* z, int64_t ldz, T gamma, int64_t per_xcd,
                                                  int64_t ntiles) {
  const int64_t tile = xcd_tile(blockIdx.x, per_xcd);
  if (tile >= ntiles) return;
  const int sub = threadIdx.x % G;
  const int64_t row = tile * (256 / G) + threadIdx.x / G;
  const bool rowok = row < n;
  const int beg = rowok ? rowptr[row] : 0;
  const int end = rowok ? rowptr[row + 1] : 0;
  T acc = T(0);
  int p = beg + sub;
  // four of the lane's entries per round: their index, value and x loads are
  // in flight together; the sum keeps the entry order
  for (; p + 3 * G < end; p += 4 * G) {
    int c[4];
    T v[4], xv[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) c[u] = colind[p + u * G];
#pragma unroll
    for (int u = 0; u < 4; ++u) v[u] = vals[p + u * G];
#pragma unroll
    for (int u = 0; u < 4; ++u) xv[u] = x[(int64_t)c[u] * ldx];
#pragma unroll
    for (int u = 0; u < 4; ++u) acc += v[u] * xv[u];
  }
  for (; p < end; p += G) acc += vals[p] * x[(int64_t)colind[p] * ldx];
#pragma unroll
  for (int o = G / 2; o > 0; o >>= 1) acc += __shfl_xor(acc, o, 64);
  if (!rowok || sub != 0) return;
  T out = alpha * acc;
  if (beta != T(0)) out += beta * x[row * ldx];
  if (z && gamma != T(0)) out += gamma * z[row * ldz];
  y[row * ldy] = out;
}

template <typename T>
static int launch_spmv(fe_ctx ctx, fe_csr A, const T* X, int64_t ldx, T* Y, int64_t ldy, T alpha, T beta,
                       const T* Z, int64_t ldz, T gamma) {
  constexpr int G = 8;
  const int64_t ntiles = (A->n + 256 / G - 1) / (256 / G);
  const int64_t nb8 = (ntiles + 7) / 8 * 8;
  const T* v = sizeof(T) == 8 ? (const T*)A->vals64 : (const T*)A->vals32;
  hipLaunchKernelGGL((k_spmv_csr<T, G>), dim3((unsigned)nb8), dim3(256), 0, ctx->stream, A->n, A->rowptr,
                     A->colind, v, X, ldx, Y, ldy, alpha, beta, Z, ldz, gamma, nb8 / 8, ntiles);
  FE_CHECK_LAUNCH();
  return 0;
}

// row-block gather kernel (fe_spmm_bg.hip); needs 16-byte aligned rows
int spmm_blocked(fe_ctx ctx, fe_csr A, int prec, int k, const void* X, int64_t ldx, void* Y, int64_t ldy,
                 double alpha, double beta, const void* Z, int64_t ldz, double gamma);
int spmm_tiled(fe_ctx ctx, fe_csr A, int prec, int k, const void* X, int64_t ldx, void* Y, int64_t ldy,
               double alpha, double beta, const void* Z, int64_t ldz, double gamma);

// Choose lanes-per-row so that one row group spans the k columns with
// 16-byte vectors when the layout allows it.
template <typename T>
static int dispatch(fe_ctx ctx, fe_csr A, int k, const T* X, int64_t ldx, T* Y, int64_t ldy, double a,
                    double b, const T* Z, int64_t ldz, double g) {
  constexpr int V16 = 16 / sizeof(T);  // elements per 16-byte vector
  auto aligned = [&](const void* p, int64_t ld) {
    return ((reinterpret_cast<uintptr_t>(p) % 16) == 0) && (ld % V16 == 0);
  };
  const bool al = aligned(X, ldx) && aligned(Y, ldy) && (!Z || aligned(Z, ldz));
  const bool vec = (k % V16 == 0) && al;
  const T alpha = static_cast<T>(a), beta = static_cast<T>(b), gamma = static_cast<T>(g);
  if (k == 1) return launch_spmv<T>(ctx, A, X, ldx, Y, ldy, alpha, beta, Z, ldz, gamma);
  // the row-block kernel handles a partial last run element-wise
  if (A->use_tl && al) return spmm_tiled(ctx, A, prec_of<T>::value, k, X, ldx, Y, ldy, a, b, Z, ldz, g);
  if (A->use_bg && al) return spmm_blocked(ctx, A, prec_of<T>::value, k, X, ldx, Y, ldy, a, b, Z, ldz, g);
  const T* v = A->use_sell ? nullptr : (sizeof(T) == 8 ? (const T*)A->vals64 : (const T*)A->vals32);
  if (A->use_sell) {
    const T* sv = sizeof(T) == 8 ? (const T*)A->sell_v64 : (const T*)A->sell_v32;
    const int groups = vec ? k / V16 : k;
    if (vec) {
      if (groups <= 4) return launch_sell<T, V16, 4>(ctx, A, k, X, ldx, Y, ldy, alpha, beta, Z, ldz, gamma, sv);
      if (groups <= 8) return launch_sell<T, V16, 8>(ctx, A, k, X, ldx, Y, ldy, alpha, beta, Z, ldz, gamma, sv);
      if (groups <= 16) return launch_sell<T, V16, 16>(ctx, A, k, X, ldx, Y, ldy, alpha, beta, Z, ldz, gamma, sv);
      if (groups <= 32) return launch_sell<T, V16, 32>(ctx, A, k, X, ldx, Y, ldy, alpha, beta, Z, ldz, gamma, sv);
      return launch_sell<T, V16, 64>(ctx, A, k, X, ldx, Y, ldy, alpha, beta, Z, ldz, gamma, sv);
    }
    if (groups <= 8) return launch_sell<T, 1, 8>(ctx, A, k, X, ldx, Y, ldy, alpha, beta, Z, ldz, gamma, sv);
    if (groups <= 32) return launch_sell<T, 1, 32>(ctx, A, k, X, ldx, Y, ldy, alpha, beta, Z, ldz, gamma, sv);
    return launch_sell<T, 1, 64>(ctx, A, k, X, ldx, Y, ldy, alpha, beta, Z, ldz, gamma, sv);
  }
  // CSR: 16-byte runs per lane when the block layout allows (a partial last
  // run is handled element-wise); up to 2 runs per lane for the widest blocks.
#define FE_CSR(LPR, RS, VEC) \
  return launch_csr<T, LPR, RS, VEC>(ctx, A, k, X, ldx, Y, ldy, alpha, beta, Z, ldz, gamma, v)
  if (al) {
    const int runs = (k + V16 - 1) / V16;
    if (runs <= 1) FE_CSR(1, 1, true);
    if (runs <= 2) FE_CSR(2, 1, true);
    if (runs <= 4) FE_CSR(4, 1, true);
    if (runs <= 8) FE_CSR(8, 1, true);
    if (runs <= 16) FE_CSR(16, 1, true);
    if (runs <= 32) FE_CSR(32, 1, true);
    if (runs <= 64) FE_CSR(64, 1, true);
    FE_CSR(64, 2, true);
  }
  if (k <= 1) FE_CSR(1, 1, false);
  if (k <= 2) FE_CSR(2, 1, false);
  if (k <= 4) FE_CSR(4, 1, false);
  if (k <= 8) FE_CSR(8, 1, false);
  if (k <= 16) FE_CSR(16, 1, false);
  if (k <= 32) FE_CSR(32, 1, false);
  FE_CSR(64, 1, false);
#undef FE_CSR
}

// ------------------------------------------------------------ SELL-C-sigma build

// Host-side construction (a one-time format conversion): rows are sorted by
// length, descending, inside windows of sigma rows; slices of C sorted rows
// are padded to their longest row. Stored column-major within a slice.
static int build_sell_host(fe_ctx ctx, fe_csr A, int C, int sigma) {
  const int n = A->n;
  std::vector<int> rp(n + 1), ci(A->nnz);
  std::vector<double> v(A->nnz);
  FE_HIP_TRY(hipMemcpyAsync(rp.data(), A->rowptr, sizeof(int) * (n + 1), hipMemcpyDeviceToHost, ctx->stream));
  if (A->nnz) {
    FE_HIP_TRY(hipMemcpyAsync(ci.data(), A->colind, sizeof(int) * A->nnz, hipMemcpyDeviceToHost, ctx->stream));
    FE_HIP_TRY(hipMemcpyAsync(v.data(), A->vals64, sizeof(double) * A->nnz, hipMemcpyDeviceToHost, ctx->stream));
  }
  FE_HIP_TRY(hipStreamSynchronize(ctx->stream));
  std::vector<int> perm(n);
  for (int i = 0; i < n; ++i) perm[i] = i;
  for (int w0 = 0; w0 < n; w0 += sigma) {
    const int w1 = w0 + sigma < n ? w0 + sigma : n;
    std::stable_sort(perm.begin() + w0, perm.begin() + w1,
                     [&](int a, int b) { return (rp[a + 1] - rp[a]) > (rp[b + 1] - rp[b]); });
  }
  const int64_t slices = (n + C - 1) / C;
  std::vector<int64_t> sptr(slices + 1, 0);
  std::vector<int> slen(slices, 0);
  for (int64_t s = 0; s < slices; ++s) {
    int w = 0;
    for (int r = 0; r < C; ++r) {
      const int64_t sr = s * C + r;
      if (sr < n) w = std::max(w, rp[perm[sr] + 1] - rp[perm[sr]]);
    }
    slen[s] = w;
    sptr[s + 1] = sptr[s] + (int64_t)w * C;
  }
  const int64_t total = sptr[slices];
  std::vector<int> scol(total > 0 ? total : 1, 0);
  std::vector<double> sv(total > 0 ? total : 1, 0.0);
  for (int64_t s = 0; s < slices; ++s)
    for (int r = 0; r < C; ++r) {
      const int64_t sr = s * C + r;
      if (sr >= n) continue;
      const int row = perm[sr];
      const int len = rp[row + 1] - rp[row];
      for (int t = 0; t < slen[s]; ++t) {
        const int64_t e = sptr[s] + (int64_t)t * C + r;
        if (t < len) {
          scol[e] = ci[rp[row] + t];
          sv[e] = v[rp[row] + t];
        } else {
          scol[e] = row;  // padding reads the row's own X entry with value 0
          sv[e] = 0.0;
        }
      }
    }
  std::vector<float> sv32(sv.size());
  for (size_t t = 0; t < sv.size(); ++t) sv32[t] = static_cast<float>(sv[t]);
  FE_HIP_TRY(hipMalloc(&A->sell_ptr, sizeof(int64_t) * (slices + 1)));
  FE_HIP_TRY(hipMalloc(&A->sell_len, sizeof(int) * slices));
  FE_HIP_TRY(hipMalloc(&A->sell_perm, sizeof(int) * n));
  FE_HIP_TRY(hipMalloc(&A->sell_col, sizeof(int) * scol.size()));
  FE_HIP_TRY(hipMalloc(&A->sell_v64, sizeof(double) * sv.size()));
  FE_HIP_TRY(hipMalloc(&A->sell_v32, sizeof(float) * sv.size()));
  FE_HIP_TRY(hipMemcpyAsync(A->sell_ptr, sptr.data(), sizeof(int64_t) * (slices + 1), hipMemcpyHostToDevice, ctx->stream));
  FE_HIP_TRY(hipMemcpyAsync(A->sell_len, slen.data(), sizeof(int) * slices, hipMemcpyHostToDevice, ctx->stream));
  FE_HIP_TRY(hipMemcpyAsync(A->sell_perm, perm.data(), sizeof(int) * n, hipMemcpyHostToDevice, ctx->stream));
  FE_HIP_TRY(hipMemcpyAsync(A->sell_col, scol.data(), sizeof(int) * scol.size(), hipMemcpyHostToDevice, ctx->stream));
  FE_HIP_TRY(hipMemcpyAsync(A->sell_v64, sv.data(), sizeof(double) * sv.size(), hipMemcpyHostToDevice, ctx->stream));
  FE_HIP_TRY(hipMemcpyAsync(A->sell_v32, sv32.data(), sizeof(float) * sv.size(), hipMemcpyHostToDevice, ctx->stream));
  FE_HIP_TRY(hipStreamSynchronize(ctx->stream));
  A->sell_C = C;
  A->sell_sigma = sigma;
  A->sell_slices = slices;
  A->sell_padded = total;
  return 0;
}

}  // namespace fe

using namespace fe;

extern "C" int fe_spmm(fe_ctx ctx, fe_csr A, int prec, int k, const void* X, int64_t ldx, void* Y,
                       int64_t ldy, double alpha, double beta, const void* Z, int64_t ldz, double gamma) {
  FE_ARG(ctx && A && X && Y, "fe_spmm: null argument");
  FE_ARG(k >= 0 && ldx >= k && ldy >= k && (!Z || ldz >= k), "fe_spmm: bad leading dimension");
  FE_ARG(X != Y, "fe_spmm: Y must not alias X");
  FE_ARG(prec == FE_F64 || prec == FE_F32, "fe_spmm: bad precision");
  if (k == 0 || A->n == 0) return 0;
  if (prec == FE_F64)
    return dispatch<double>(ctx, A, k, (const double*)X, ldx, (double*)Y, ldy, alpha, beta, (const double*)Z,
                            ldz, gamma);
  return dispatch<float>(ctx, A, k, (const float*)X, ldx, (float*)Y, ldy, alpha, beta, (const float*)Z, ldz,
                         gamma);
}

extern "C" int fe_csr_build_sell(fe_ctx ctx, fe_csr A, int C, int sigma) {
  FE_ARG(ctx && A, "fe_csr_build_sell: null argument");
  FE_ARG(C > 0 && C <= 1024 && sigma >= 1, "fe_csr_build_sell: bad C or sigma");
  FE_ARG(!A->sell_col, "fe_csr_build_sell: already built");
  return build_sell_host(ctx, A, C, sigma);
}
