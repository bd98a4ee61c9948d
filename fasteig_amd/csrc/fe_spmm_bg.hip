// Row-blocked SpMM: Y = alpha A X + beta X + gamma Z, one gather per column
// of a row block instead of one per nonzero.
//
// The rows are grouped in blocks of RB consecutive rows. A block stores the
// sorted union of its rows' column indices and, per union entry, the RB
// values of that column in the block's rows (zero where a row has no such
// entry): a BCSR layout with RB x 1 blocks. The product gathers each union X
// row once and feeds it to RB accumulators. For matrices with locality the
// union is much smaller than the nonzero count (27-point stencil, RB = 4: 54
// union columns for 108 nonzeros), which halves the gathered X bytes, the
// dominant traffic of the block product through the vector memory path;
// the explicit zeros cost FMAs and a little more A traffic. A row's sums run
// over its entries in ascending column order, as the CSR kernel does on a
// matrix with sorted rows (added explicit zeros leave the sums unchanged).
#include "fe_common.hpp"

#include <algorithm>

namespace fe {

typedef float nt_f4 __attribute__((ext_vector_type(4)));  // 16 bytes for non-temporal access

// Same lane layout as the CSR kernel: a group of LPR lanes owns one row block
// and RS 16-byte runs per lane, spaced LPR * 16 bytes apart. The group reads
// the union entries' columns and values directly (one address per group:
// scalar loads when the group is the whole wavefront).
template <typename T, int RB, int LPR, int RS>
__global__ __launch_bounds__(256) void k_spmm_rb(int n, const int64_t* __restrict__ bptr,
                                                 const int* __restrict__ ucol, const T* __restrict__ bval, int k,
                                                 const T* __restrict__ X, int64_t ldx, T* Y, int64_t ldy, T alpha,
                                                 T beta, const T* Z, int64_t ldz, T gamma, int64_t per_xcd,
                                                 int64_t ntiles, int64_t nblk, const int* __restrict__ sched) {
  constexpr int V = 16 / (int)sizeof(T);
  constexpr int GPW = 64 / LPR;  // row blocks per wavefront
  constexpr int GPB = 4 * GPW;   // row blocks per workgroup
  constexpr int STRIDE = LPR * V;
  constexpr int COLS = STRIDE * RS;
  constexpr int UNR = (RS == 1 && RB == 4) ? 8 : 4;
  const int64_t tile = xcd_tile(blockIdx.x, per_xcd);
  if (tile >= ntiles) return;
  const int lane = threadIdx.x & 63;
  // wave index made wavefront-uniform so that whole-wave groups use scalar loads
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int grp = LPR == 64 ? 0 : lane / LPR;
  // slot in the processing order; the row block is the slot itself or the
  // schedule's entry for it
  const int64_t slot = tile * GPB + wave * GPW + grp;
  const bool blkok = slot < nblk;
  const int64_t blk = (sched && blkok) ? static_cast<int64_t>(sched[slot]) : slot;
  const int sub = LPR == 64 ? lane : lane % LPR;
  const int c = blockIdx.y * COLS + sub * V;
  const bool full = c + (RS - 1) * STRIDE + V <= k;
  const int64_t e0 = blkok ? bptr[blk] : 0;
  const int64_t e1 = blkok ? bptr[blk + 1] : 0;
  T acc[RB][RS][V];
#pragma unroll
  for (int r = 0; r < RB; ++r)
#pragma unroll
    for (int s = 0; s < RS; ++s)
#pragma unroll
      for (int e = 0; e < V; ++e) acc[r][s][e] = T(0);
  for (int64_t e = e0; e < e1; e += UNR) {
    T xv[UNR][RS][V];
#pragma unroll
    for (int u = 0; u < UNR; ++u) {
      if (e + u < e1) {
        const T* xp = X + (int64_t)ucol[e + u] * ldx + c;
#pragma unroll
        for (int s = 0; s < RS; ++s) {
          const int cs = c + s * STRIDE;
          if (full || cs + V <= k) {
            const float4 f = *reinterpret_cast<const float4*>(xp + s * STRIDE);
            __builtin_memcpy(&xv[u][s][0], &f, 16);
          } else {
#pragma unroll
            for (int q = 0; q < V; ++q) xv[u][s][q] = (cs + q < k) ? xp[s * STRIDE + q] : T(0);
          }
        }
      }
    }
#pragma unroll
    for (int u = 0; u < UNR; ++u) {
      if (e + u < e1) {
        const T* w = bval + (e + u) * RB;
#pragma unroll
        for (int r = 0; r < RB; ++r) {
          const T wr = w[r];
#pragma unroll
          for (int s = 0; s < RS; ++s)
#pragma unroll
            for (int q = 0; q < V; ++q) acc[r][s][q] += wr * xv[u][s][q];
        }
      }
    }
  }
  if (!blkok) return;
#pragma unroll
  for (int r = 0; r < RB; ++r) {
    const int64_t row = blk * RB + r;
    if (row >= n) break;
#pragma unroll
    for (int s = 0; s < RS; ++s) {
      const int cs = c + s * STRIDE;
      if (cs >= k) continue;
      T out[V];
#pragma unroll
      for (int q = 0; q < V; ++q) out[q] = alpha * acc[r][s][q];
      if (cs + V <= k) {
        if (beta != T(0)) {
          const float4 f = *reinterpret_cast<const float4*>(X + row * ldx + cs);
          T xv[V];
          __builtin_memcpy(xv, &f, 16);
#pragma unroll
          for (int q = 0; q < V; ++q) out[q] += beta * xv[q];
        }
        if (Z && gamma != T(0)) {
          // Z and Y stream through once: non-temporal, so they do not evict
          // the gathered X rows from L2
          const nt_f4 f = __builtin_nontemporal_load(reinterpret_cast<const nt_f4*>(Z + row * ldz + cs));
          T zv[V];
          __builtin_memcpy(zv, &f, 16);
#pragma unroll
          for (int q = 0; q < V; ++q) out[q] += gamma * zv[q];
        }
        nt_f4 o;
        __builtin_memcpy(&o, out, 16);
        __builtin_nontemporal_store(o, reinterpret_cast<nt_f4*>(Y + row * ldy + cs));
      } else {
#pragma unroll
        for (int q = 0; q < V; ++q) {
          if (cs + q >= k) continue;
          T o = out[q];
          if (beta != T(0)) o += beta * X[row * ldx + cs + q];
          if (Z && gamma != T(0)) o += gamma * Z[row * ldz + cs + q];
          Y[row * ldy + cs + q] = o;
        }
      }
    }
  }
}

__device__ __forceinline__ float lane_value(float v, int u) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), u));
}
__device__ __forceinline__ double lane_value(double v, int u) {
  const int64_t b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane(static_cast<int>(b), u);
  const int hi = __builtin_amdgcn_readlane(static_cast<int>(b >> 32), u);
  return __longlong_as_double((static_cast<int64_t>(hi) << 32) | static_cast<uint32_t>(lo));
}

// One wavefront per row block (the LPR = 64 layout of k_spmm_rb, same sums in
// the same order). The gather loop does not wait on index loads: every 64
// union entries, lane j reads entry j's column and RB values with one vector
// load each, and the loop takes them from lane u with readlane (uniform, so
// the gather address is scalar). The gathers of the next UNR entries are
// issued before the FMAs of the current UNR, so a wavefront keeps two groups
// of row loads in flight.
template <typename T, int RB, int RS, int UNR>
__global__ __launch_bounds__(256) void k_spmm_rbw(int n, const int64_t* __restrict__ bptr,
                                                  const int* __restrict__ ucol, const T* __restrict__ bval, int k,
                                                  const T* __restrict__ X, int64_t ldx, T* Y, int64_t ldy, T alpha,
                                                  T beta, const T* Z, int64_t ldz, T gamma, int64_t per_xcd,
                                                  int64_t ntiles, int64_t nblk, const int* __restrict__ sched) {
  constexpr int V = 16 / (int)sizeof(T);
  constexpr int STRIDE = 64 * V;
  constexpr int COLS = STRIDE * RS;
  constexpr int WV = RB * (int)sizeof(T) / 16;  // 16-byte loads per entry's values
  const int64_t tile = xcd_tile(blockIdx.x, per_xcd);
  if (tile >= ntiles) return;
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int64_t slot = tile * 4 + wave;
  if (slot >= nblk) return;  // wavefront-uniform; the kernel has no barrier
  const int64_t blk = sched ? static_cast<int64_t>(sched[slot]) : slot;
  const int c = blockIdx.y * COLS + lane * V;
  const bool full = c + (RS - 1) * STRIDE + V <= k;
  const int64_t e0 = bptr[blk], e1 = bptr[blk + 1];
  T acc[RB][RS][V];
#pragma unroll
  for (int r = 0; r < RB; ++r)
#pragma unroll
    for (int s = 0; s < RS; ++s)
#pragma unroll
      for (int e = 0; e < V; ++e) acc[r][s][e] = T(0);
  T xa[UNR][RS][V], xb[UNR][RS][V];
  for (int64_t ch = e0; ch < e1; ch += 64) {
    const int cnt = static_cast<int>(e1 - ch < 64 ? e1 - ch : 64);
    int myc = 0;
    T myw[RB];
    if (lane < cnt) {
      myc = ucol[ch + lane];
#pragma unroll
      for (int j = 0; j < WV; ++j) {
        const float4 f = reinterpret_cast<const float4*>(bval + (ch + lane) * RB)[j];
        __builtin_memcpy(&myw[j * (16 / (int)sizeof(T))], &f, 16);
      }
    } else {
#pragma unroll
      for (int r = 0; r < RB; ++r) myw[r] = T(0);
    }
    auto load = [&](T (&x)[UNR][RS][V], int g) {
#pragma unroll
      for (int u = 0; u < UNR; ++u) {
        if (g + u < cnt) {
          const T* xp = X + static_cast<int64_t>(__builtin_amdgcn_readlane(myc, g + u)) * ldx + c;
#pragma unroll
          for (int s = 0; s < RS; ++s) {
            const int cs = c + s * STRIDE;
            if (full || cs + V <= k) {
              const float4 f = *reinterpret_cast<const float4*>(xp + s * STRIDE);
              __builtin_memcpy(&x[u][s][0], &f, 16);
            } else {
#pragma unroll
              for (int q = 0; q < V; ++q) x[u][s][q] = (cs + q < k) ? xp[s * STRIDE + q] : T(0);
            }
          }
        }
      }
    };
    auto fma = [&](const T (&x)[UNR][RS][V], int g) {
#pragma unroll
      for (int u = 0; u < UNR; ++u) {
        if (g + u < cnt) {
#pragma unroll
          for (int r = 0; r < RB; ++r) {
            // the value is wavefront-uniform: a zero (a row of the block
            // without this column) skips its FMAs with a scalar branch
            const T wr = lane_value(myw[r], g + u);
            if (wr != T(0)) {
#pragma unroll
              for (int s = 0; s < RS; ++s)
#pragma unroll
                for (int q = 0; q < V; ++q) acc[r][s][q] += wr * x[u][s][q];
            }
          }
        }
      }
    };
    load(xa, 0);
    for (int g = 0; g < cnt; g += 2 * UNR) {
      if (g + UNR < cnt) load(xb, g + UNR);
      fma(xa, g);
      if (g + 2 * UNR < cnt) load(xa, g + 2 * UNR);
      if (g + UNR < cnt) fma(xb, g + UNR);
    }
  }
#pragma unroll
  for (int r = 0; r < RB; ++r) {
    const int64_t row = blk * RB + r;
    if (row >= n) break;
#pragma unroll
    for (int s = 0; s < RS; ++s) {
      const int cs = c + s * STRIDE;
      if (cs >= k) continue;
      T out[V];
#pragma unroll
      for (int q = 0; q < V; ++q) out[q] = alpha * acc[r][s][q];
      if (cs + V <= k) {
        if (beta != T(0)) {
          const float4 f = *reinterpret_cast<const float4*>(X + row * ldx + cs);
          T xv[V];
          __builtin_memcpy(xv, &f, 16);
#pragma unroll
          for (int q = 0; q < V; ++q) out[q] += beta * xv[q];
        }
        if (Z && gamma != T(0)) {
          const nt_f4 f = __builtin_nontemporal_load(reinterpret_cast<const nt_f4*>(Z + row * ldz + cs));
          T zv[V];
          __builtin_memcpy(zv, &f, 16);
#pragma unroll
          for (int q = 0; q < V; ++q) out[q] += gamma * zv[q];
        }
        nt_f4 o;
        __builtin_memcpy(&o, out, 16);
        __builtin_nontemporal_store(o, reinterpret_cast<nt_f4*>(Y + row * ldy + cs));
      } else {
#pragma unroll
        for (int q = 0; q < V; ++q) {
          if (cs + q >= k) continue;
          T o = out[q];
          if (beta != T(0)) o += beta * X[row * ldx + cs + q];
          if (Z && gamma != T(0)) o += gamma * Z[row * ldz + cs + q];
          Y[row * ldy + cs + q] = o;
        }
      }
    }
  }
}

template <typename T, int RB, int LPR, int RS>
static int launch_rb(fe_ctx ctx, fe_csr A, int k, const T* X, int64_t ldx, T* Y, int64_t ldy, T alpha, T beta,
                     const T* Z, int64_t ldz, T gamma) {
  constexpr int V = 16 / (int)sizeof(T);
  constexpr int GPB = 4 * (64 / LPR);
  constexpr int COLS = LPR * V * RS;
  const int64_t ntiles = (A->bg_nblk + GPB - 1) / GPB;
  const int64_t nb8 = (ntiles + 7) / 8 * 8;
  const unsigned ycols = static_cast<unsigned>((k + COLS - 1) / COLS);
  const T* bv = sizeof(T) == 8 ? (const T*)A->bg_v64 : (const T*)A->bg_v32;
  if constexpr (LPR == 64) {
    constexpr int UNR = 4;  // two groups of 4 x RS 16-byte loads in flight
    hipLaunchKernelGGL((k_spmm_rbw<T, RB, RS, UNR>), dim3((unsigned)nb8, ycols), dim3(256), 0, ctx->stream, A->n,
                       A->bg_bptr, A->bg_ucol, bv, k, X, ldx, Y, ldy, alpha, beta, Z, ldz, gamma, nb8 / 8, ntiles,
                       A->bg_nblk, A->bg_sched);
  } else {
    hipLaunchKernelGGL((k_spmm_rb<T, RB, LPR, RS>), dim3((unsigned)nb8, ycols), dim3(256), 0, ctx->stream, A->n,
                       A->bg_bptr, A->bg_ucol, bv, k, X, ldx, Y, ldy, alpha, beta, Z, ldz, gamma, nb8 / 8, ntiles,
                       A->bg_nblk, A->bg_sched);
  }
  FE_CHECK_LAUNCH();
  return 0;
}

template <typename T, int RB>
static int dispatch_rb(fe_ctx ctx, fe_csr A, int k, const T* X, int64_t ldx, T* Y, int64_t ldy, T alpha, T beta,
                       const T* Z, int64_t ldz, T gamma) {
  constexpr int V = 16 / (int)sizeof(T);
  const int runs = (k + V - 1) / V;
#define FE_RB(LPR, RS) return launch_rb<T, RB, LPR, RS>(ctx, A, k, X, ldx, Y, ldy, alpha, beta, Z, ldz, gamma)
  if (runs <= 4) FE_RB(4, 1);
  if (runs <= 8) FE_RB(8, 1);
  if (runs <= 16) FE_RB(16, 1);
  if (runs <= 32) FE_RB(32, 1);
  if (runs <= 64) FE_RB(64, 1);
  FE_RB(64, 2);
#undef FE_RB
}

// Called by fe_spmm for 16-byte aligned blocks when the row-block copy is active.
int spmm_blocked(fe_ctx ctx, fe_csr A, int prec, int k, const void* X, int64_t ldx, void* Y, int64_t ldy,
                 double alpha, double beta, const void* Z, int64_t ldz, double gamma) {
  if (prec == FE_F64) {
    auto f = A->bg_rb == 8 ? dispatch_rb<double, 8> : dispatch_rb<double, 4>;
    return f(ctx, A, k, (const double*)X, ldx, (double*)Y, ldy, alpha, beta, (const double*)Z, ldz, gamma);
  }
  auto f = A->bg_rb == 8 ? dispatch_rb<float, 8> : dispatch_rb<float, 4>;
  return f(ctx, A, k, (const float*)X, ldx, (float*)Y, ldy, (float)alpha, (float)beta, (const float*)Z, ldz,
           (float)gamma);
}

}  // namespace fe

using namespace fe;

extern "C" int fe_csr_build_blocked(fe_ctx ctx, fe_csr A, int rb, double* reuse) {
  FE_ARG(ctx && A, "fe_csr_build_blocked: null argument");
  FE_ARG(rb == 4 || rb == 8, "fe_csr_build_blocked: the row block size is 4 or 8");
  if (A->bg_ucol) {
    FE_ARG(A->bg_rb == rb, "fe_csr_build_blocked: already built with another row block size");
    if (reuse) *reuse = A->bg_reuse;
    return 0;
  }
  const int n = A->n;
  std::vector<int> rp(n + 1), ci(A->nnz);
  std::vector<double> v(A->nnz);
  FE_HIP_TRY(hipMemcpyAsync(rp.data(), A->rowptr, sizeof(int) * (n + 1), hipMemcpyDeviceToHost, ctx->stream));
  if (A->nnz) {
    FE_HIP_TRY(hipMemcpyAsync(ci.data(), A->colind, sizeof(int) * A->nnz, hipMemcpyDeviceToHost, ctx->stream));
    FE_HIP_TRY(hipMemcpyAsync(v.data(), A->vals64, sizeof(double) * A->nnz, hipMemcpyDeviceToHost, ctx->stream));
  }
  FE_HIP_TRY(hipStreamSynchronize(ctx->stream));
  const int64_t nblk = (n + rb - 1) / rb;
  // pass 1: union sizes
  std::vector<int64_t> bptr(nblk + 1, 0);
  std::vector<int> u;
  for (int64_t b = 0; b < nblk; ++b) {
    const int r0 = static_cast<int>(b * rb), r1 = std::min(n, r0 + rb);
    u.assign(ci.begin() + rp[r0], ci.begin() + rp[r1]);
    std::sort(u.begin(), u.end());
    bptr[b + 1] = bptr[b] + (std::unique(u.begin(), u.end()) - u.begin());
  }
  const int64_t nu = bptr[nblk];
  const double r = nu ? static_cast<double>(A->nnz) / static_cast<double>(nu) : 0.0;
  if (reuse) *reuse = r;
  if (r < 1.5) return 1;  // too little column sharing: the CSR gather is as good
  // pass 2: union columns and the rb values per union entry
  std::vector<int> ucol(static_cast<size_t>(std::max<int64_t>(nu, 1)));
  std::vector<double> bv(static_cast<size_t>(std::max<int64_t>(nu, 1)) * rb, 0.0);
  for (int64_t b = 0; b < nblk; ++b) {
    const int r0 = static_cast<int>(b * rb), r1 = std::min(n, r0 + rb);
    u.assign(ci.begin() + rp[r0], ci.begin() + rp[r1]);
    std::sort(u.begin(), u.end());
    u.erase(std::unique(u.begin(), u.end()), u.end());
    std::copy(u.begin(), u.end(), ucol.begin() + bptr[b]);
    for (int row = r0; row < r1; ++row)
      for (int p = rp[row]; p < rp[row + 1]; ++p) {
        const int64_t pos = bptr[b] + (std::lower_bound(u.begin(), u.end(), ci[p]) - u.begin());
        bv[pos * rb + (row - r0)] += v[p];  // duplicates in a CSR row add up
      }
  }
  std::vector<float> bv32(bv.size());
  for (size_t i = 0; i < bv.size(); ++i) bv32[i] = static_cast<float>(bv[i]);
  FE_HIP_TRY(hipMalloc(&A->bg_bptr, sizeof(int64_t) * (nblk + 1)));
  FE_HIP_TRY(hipMalloc(&A->bg_ucol, sizeof(int) * ucol.size()));
  FE_HIP_TRY(hipMalloc(&A->bg_v64, sizeof(double) * bv.size()));
  FE_HIP_TRY(hipMalloc(&A->bg_v32, sizeof(float) * bv32.size()));
  FE_HIP_TRY(hipMemcpyAsync(A->bg_bptr, bptr.data(), sizeof(int64_t) * (nblk + 1), hipMemcpyHostToDevice, ctx->stream));
  FE_HIP_TRY(hipMemcpyAsync(A->bg_ucol, ucol.data(), sizeof(int) * ucol.size(), hipMemcpyHostToDevice, ctx->stream));
  FE_HIP_TRY(hipMemcpyAsync(A->bg_v64, bv.data(), sizeof(double) * bv.size(), hipMemcpyHostToDevice, ctx->stream));
  FE_HIP_TRY(hipMemcpyAsync(A->bg_v32, bv32.data(), sizeof(float) * bv32.size(), hipMemcpyHostToDevice, ctx->stream));
  FE_HIP_TRY(hipStreamSynchronize(ctx->stream));
  A->bg_rb = rb;
  A->bg_nblk = nblk;
  A->bg_nu = nu;
  A->bg_reuse = r;
  return 0;
}

namespace fe {
// The distinct positive column offsets j - i of a sample of rows form three
// clusters (within a line, to the neighbouring lines, to the neighbouring
// planes) whose centres are 1, the line length and the plane length.
bool detect_grid(const std::vector<int>& rp, const std::vector<int>& ci, int n, int64_t& line, int64_t& plane) {
  std::vector<int64_t> off;
  const int step = std::max(1, n / 4096);
  for (int i = 0; i < n; i += step)
    for (int p = rp[i]; p < rp[i + 1]; ++p)
      if (ci[p] > i) off.push_back(static_cast<int64_t>(ci[p]) - i);
  std::sort(off.begin(), off.end());
  off.erase(std::unique(off.begin(), off.end()), off.end());
  if (off.empty() || off.size() > 64) return false;
  std::vector<std::pair<int64_t, int64_t>> cl;  // [first, last] of each cluster
  for (int64_t d : off) {
    if (cl.empty() || d > 4 * cl.back().second) cl.push_back({d, d});
    else cl.back().second = d;
  }
  if (cl.size() != 3 || cl[0].second > 2) return false;
  line = (cl[1].first + cl[1].second) / 2;
  plane = (cl[2].first + cl[2].second) / 2;
  return line >= 4 && plane >= 4 * line && plane % line == 0 && n >= 3 * plane;
}
}  // namespace fe

extern "C" int fe_csr_block_schedule(fe_ctx ctx, fe_csr A, int64_t plane, int64_t strip) {
  FE_ARG(ctx && A, "fe_csr_block_schedule: null argument");
  FE_ARG(A->bg_ucol, "fe_csr_block_schedule: build the row-block copy first");
  if (A->bg_sched) {
    FE_HIP_TRY(hipStreamSynchronize(ctx->stream));
    FE_HIP_TRY(hipFree(A->bg_sched));
    A->bg_sched = nullptr;
    A->bg_plane = A->bg_strip = 0;
  }
  if (plane == 0) return 0;
  const int n = A->n, rb = A->bg_rb;
  int64_t line = 0;
  if (plane < 0) {
    std::vector<int> rp(n + 1), ci(A->nnz);
    FE_HIP_TRY(hipMemcpyAsync(rp.data(), A->rowptr, sizeof(int) * (n + 1), hipMemcpyDeviceToHost, ctx->stream));
    if (A->nnz)
      FE_HIP_TRY(hipMemcpyAsync(ci.data(), A->colind, sizeof(int) * A->nnz, hipMemcpyDeviceToHost, ctx->stream));
    FE_HIP_TRY(hipStreamSynchronize(ctx->stream));
    if (!detect_grid(rp, ci, n, line, plane)) return 1;
    // one strip per XCD: each XCD sweeps its eighth of every plane
    if (strip <= 0) strip = line * ((plane / line + 7) / 8);
    if (plane % rb || strip % rb) return 1;
  }
  FE_ARG(strip > 0 && plane % rb == 0 && strip % rb == 0 && strip <= plane,
         "fe_csr_block_schedule: plane and strip must be multiples of the row block size, strip <= plane");
  const int64_t nblk = A->bg_nblk, planes = (n + plane - 1) / plane;
  std::vector<int> order;
  order.reserve(static_cast<size_t>(nblk));
  for (int64_t s0 = 0; s0 < plane; s0 += strip)
    for (int64_t z = 0; z < planes; ++z)
      for (int64_t r = z * plane + s0; r < z * plane + std::min(s0 + strip, plane) && r < n; r += rb)
        order.push_back(static_cast<int>(r / rb));
  FE_ARG(static_cast<int64_t>(order.size()) == nblk, "fe_csr_block_schedule: schedule does not cover the row blocks");
  FE_HIP_TRY(hipMalloc(&A->bg_sched, sizeof(int) * order.size()));
  FE_HIP_TRY(hipMemcpyAsync(A->bg_sched, order.data(), sizeof(int) * order.size(), hipMemcpyHostToDevice, ctx->stream));
  FE_HIP_TRY(hipStreamSynchronize(ctx->stream));
  A->bg_plane = plane;
  A->bg_strip = strip;
  return 0;
}
