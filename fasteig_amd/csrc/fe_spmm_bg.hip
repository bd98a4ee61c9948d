// Row-block gather SpMM: Y = alpha A X + beta X + gamma Z with X staged in LDS.
//
// For matrices with locality (stencils, FEM meshes) the rows of a block of RB
// consecutive rows touch a union of X rows much smaller than their nonzero
// count (27-point stencil, RB = 32: ~306 union rows for ~860 nonzeros). A
// workgroup takes one block and one column tile of KT columns (128 bytes per
// X row), copies the union's X row segments into LDS with coalesced 16-byte
// loads, then runs the nonzeros of its 32 rows against LDS: each X row
// segment crosses L2 once per block instead of once per nonzero. The
// nonzeros keep their CSR order, so the sums are the same as the CSR kernel's.
// Column tiles of one block are consecutive on one XCD, so the block's
// nonzeros (values, 16-bit union positions) come from that XCD's L2 after the
// first tile.
#include "fe_common.hpp"

#include <algorithm>

namespace fe {

constexpr int BG_RB = 32;          // rows per block
constexpr int BG_MAXU = 448;       // union rows that fit the LDS tile
constexpr int BG_ROWB = 128 + 16;  // LDS bytes per staged row (128 B + pad)

template <typename T>
__global__ __launch_bounds__(256) void k_spmm_bg(int n, const int* __restrict__ rowptr, const uint16_t* __restrict__ lcol,
                                                 const T* __restrict__ vals, const int64_t* __restrict__ bptr,
                                                 const int* __restrict__ ucol, int k, const T* __restrict__ X,
                                                 int64_t ldx, T* Y, int64_t ldy, T alpha, T beta, const T* Z,
                                                 int64_t ldz, T gamma, int ntc, int64_t per_xcd, int64_t ntiles) {
  constexpr int V = 16 / (int)sizeof(T);    // elements per 16-byte run
  constexpr int KT = 128 / (int)sizeof(T);  // columns per tile
  constexpr int RUNS = KT / V;              // 8 runs per staged row
  constexpr int LPR = 256 / BG_RB;          // 8 lanes per row
  extern __shared__ __align__(16) unsigned char smem[];
  const int64_t t = xcd_tile(blockIdx.x, per_xcd);
  if (t >= ntiles) return;
  const int64_t blk = t / ntc;
  const int ct = static_cast<int>(t - blk * ntc);
  const int c0 = ct * KT;
  const int tid = threadIdx.x;
  const int64_t u0 = bptr[blk];
  const int nu = static_cast<int>(bptr[blk + 1] - u0);
  // stage the union rows' column tile
  for (int idx = tid; idx < nu * RUNS; idx += 256) {
    const int u = idx / RUNS, r = idx % RUNS;
    const int c = c0 + r * V;
    const T* src = X + (int64_t)ucol[u0 + u] * ldx + c;
    T* dst = reinterpret_cast<T*>(smem + u * BG_ROWB) + r * V;
    if (c + V <= k) {
      *reinterpret_cast<float4*>(dst) = *reinterpret_cast<const float4*>(src);
    } else {
#pragma unroll
      for (int e = 0; e < V; ++e) dst[e] = (c + e < k) ? src[e] : T(0);
    }
  }
  __syncthreads();
  const int lane = tid & 63;
  const int sub = lane % LPR, gbase = lane - sub;
  const int64_t row = blk * BG_RB + tid / LPR;
  const int c = c0 + sub * V;
  const bool rowok = row < n;
  const int beg = rowok ? rowptr[row] : 0;
  const int end = rowok ? rowptr[row + 1] : 0;
  T acc[V];
#pragma unroll
  for (int e = 0; e < V; ++e) acc[e] = T(0);
  for (int p = beg; p < end; p += LPR) {
    const int q = p + sub;
    const int myl = q < end ? (int)lcol[q] : 0;
    const T myv = q < end ? vals[q] : T(0);
    const int cnt = end - p < LPR ? end - p : LPR;
    for (int s = 0; s < cnt; ++s) {
      const int j = __shfl(myl, gbase + s, 64);
      const T v = __shfl(myv, gbase + s, 64);
      const float4 f = *reinterpret_cast<const float4*>(smem + j * BG_ROWB + sub * 16);
      T x[V];
      __builtin_memcpy(x, &f, 16);
#pragma unroll
      for (int e = 0; e < V; ++e) acc[e] += v * x[e];
    }
  }
  if (!rowok || c >= k) return;
  T out[V];
#pragma unroll
  for (int e = 0; e < V; ++e) out[e] = alpha * acc[e];
  if (c + V <= k) {
    if (beta != T(0)) {
      const float4 f = *reinterpret_cast<const float4*>(X + row * ldx + c);
      T xv[V];
      __builtin_memcpy(xv, &f, 16);
#pragma unroll
      for (int e = 0; e < V; ++e) out[e] += beta * xv[e];
    }
    if (Z && gamma != T(0)) {
      const float4 f = *reinterpret_cast<const float4*>(Z + row * ldz + c);
      T zv[V];
      __builtin_memcpy(zv, &f, 16);
#pragma unroll
      for (int e = 0; e < V; ++e) out[e] += gamma * zv[e];
    }
    float4 o;
    __builtin_memcpy(&o, out, 16);
    *reinterpret_cast<float4*>(Y + row * ldy + c) = o;
  } else {
#pragma unroll
    for (int e = 0; e < V; ++e) {
      if (c + e >= k) continue;
      T o = out[e];
      if (beta != T(0)) o += beta * X[row * ldx + c + e];
      if (Z && gamma != T(0)) o += gamma * Z[row * ldz + c + e];
      Y[row * ldy + c + e] = o;
    }
  }
}

template <typename T>
int spmm_blocked_t(fe_ctx ctx, fe_csr A, int k, const T* X, int64_t ldx, T* Y, int64_t ldy, T alpha, T beta,
                   const T* Z, int64_t ldz, T gamma) {
  constexpr int KT = 128 / (int)sizeof(T);
  const int ntc = (k + KT - 1) / KT;
  const int64_t ntiles = A->bg_nblk * ntc;
  const int64_t nb8 = (ntiles + 7) / 8 * 8;
  const size_t lds = static_cast<size_t>(A->bg_maxu) * BG_ROWB;
  const T* v = sizeof(T) == 8 ? (const T*)A->vals64 : (const T*)A->vals32;
  hipLaunchKernelGGL(k_spmm_bg<T>, dim3((unsigned)nb8), dim3(256), lds, ctx->stream, A->n, A->rowptr, A->bg_lcol, v,
                     A->bg_bptr, A->bg_ucol, k, X, ldx, Y, ldy, alpha, beta, Z, ldz, gamma, ntc, nb8 / 8, ntiles);
  FE_CHECK_LAUNCH();
  return 0;
}

int spmm_blocked(fe_ctx ctx, fe_csr A, int prec, int k, const void* X, int64_t ldx, void* Y, int64_t ldy,
                 double alpha, double beta, const void* Z, int64_t ldz, double gamma) {
  if (prec == FE_F64)
    return spmm_blocked_t<double>(ctx, A, k, (const double*)X, ldx, (double*)Y, ldy, alpha, beta, (const double*)Z,
                                  ldz, gamma);
  return spmm_blocked_t<float>(ctx, A, k, (const float*)X, ldx, (float*)Y, ldy, (float)alpha, (float)beta,
                               (const float*)Z, ldz, (float)gamma);
}

}  // namespace fe

using namespace fe;

extern "C" int fe_csr_build_blocked(fe_ctx ctx, fe_csr A, int rb, double* reuse) {
  FE_ARG(ctx && A, "fe_csr_build_blocked: null argument");
  FE_ARG(rb == BG_RB, "fe_csr_build_blocked: the row block size is 32");
  if (A->bg_lcol) {
    if (reuse) *reuse = A->bg_reuse;
    return 0;
  }
  const int n = A->n;
  std::vector<int> rp(n + 1), ci(A->nnz);
  FE_HIP_TRY(hipMemcpyAsync(rp.data(), A->rowptr, sizeof(int) * (n + 1), hipMemcpyDeviceToHost, ctx->stream));
  if (A->nnz)
    FE_HIP_TRY(hipMemcpyAsync(ci.data(), A->colind, sizeof(int) * A->nnz, hipMemcpyDeviceToHost, ctx->stream));
  FE_HIP_TRY(hipStreamSynchronize(ctx->stream));
  const int64_t nblk = (n + rb - 1) / rb;
  std::vector<int64_t> bptr(nblk + 1, 0);
  std::vector<int> ucol;
  std::vector<uint16_t> lcol(A->nnz > 0 ? A->nnz : 1);
  ucol.reserve(static_cast<size_t>(A->nnz / 2 + 16));
  std::vector<int> u;
  int maxu = 0;
  for (int64_t b = 0; b < nblk; ++b) {
    const int r0 = static_cast<int>(b * rb), r1 = std::min(n, r0 + rb);
    u.assign(ci.begin() + rp[r0], ci.begin() + rp[r1]);
    std::sort(u.begin(), u.end());
    u.erase(std::unique(u.begin(), u.end()), u.end());
    maxu = std::max(maxu, static_cast<int>(u.size()));
    if (maxu > BG_MAXU) {
      if (reuse) *reuse = 0.0;
      return 1;  // union too large for the LDS tile
    }
    for (int p = rp[r0]; p < rp[r1]; ++p)
      lcol[p] = static_cast<uint16_t>(std::lower_bound(u.begin(), u.end(), ci[p]) - u.begin());
    ucol.insert(ucol.end(), u.begin(), u.end());
    bptr[b + 1] = static_cast<int64_t>(ucol.size());
  }
  const double r = ucol.empty() ? 0.0 : static_cast<double>(A->nnz) / static_cast<double>(ucol.size());
  if (reuse) *reuse = r;
  if (r < 2.0) return 1;  // too little reuse: the direct gather is as good
  FE_HIP_TRY(hipMalloc(&A->bg_bptr, sizeof(int64_t) * (nblk + 1)));
  FE_HIP_TRY(hipMalloc(&A->bg_ucol, sizeof(int) * std::max<size_t>(ucol.size(), 1)));
  FE_HIP_TRY(hipMalloc(&A->bg_lcol, sizeof(uint16_t) * lcol.size()));
  FE_HIP_TRY(hipMemcpyAsync(A->bg_bptr, bptr.data(), sizeof(int64_t) * (nblk + 1), hipMemcpyHostToDevice, ctx->stream));
  if (!ucol.empty())
    FE_HIP_TRY(hipMemcpyAsync(A->bg_ucol, ucol.data(), sizeof(int) * ucol.size(), hipMemcpyHostToDevice, ctx->stream));
  FE_HIP_TRY(hipMemcpyAsync(A->bg_lcol, lcol.data(), sizeof(uint16_t) * lcol.size(), hipMemcpyHostToDevice, ctx->stream));
  FE_HIP_TRY(hipStreamSynchronize(ctx->stream));
  A->bg_rb = rb;
  A->bg_nblk = nblk;
  A->bg_maxu = std::max(maxu, 1);
  A->bg_reuse = r;
  return 0;
}
