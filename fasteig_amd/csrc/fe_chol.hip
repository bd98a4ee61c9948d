// Blocked FP64 Cholesky factorization and triangular inverse of the k x k Gram
// matrix on the device (column-major), block size NB = 32:
//
//   potrf  for each block column jb:
//            k_chol_diag    factor the diagonal block in LDS (unblocked,
//                           right-looking, one wavefront, one barrier per
//                           column), invert it (a column per thread by
//                           forward substitution), keep the inverse
//            k_chol_panel   L_ib = A_ib L_jj^{-T} for the blocks below
//            k_chol_update  A_ij -= L_ib L_jb^T for the trailing lower blocks
//   rinv   (L^T)^{-1} = (L^{-1})^T:
//            k_trtri_diag   invert every diagonal block (all in parallel)
//            k_trtri_off    block diagonal d = i - j: X_ij = -L_ii^{-1} sum_p L_ip X_pj
//
// Each launch is a handful of 256-thread workgroups; the k = 512 Gram of the
// benchmark takes 48 + 17 small launches instead of one latency-bound kernel.
#include "fe_common.hpp"

#include <cmath>

namespace fe {

constexpr int NB = 32;

// Inverse of the lower-triangular block D (LDS, D[r][c], unit diagonal past
// nb) into Di: thread c < NB solves L x = e_c by forward substitution and
// writes column c of Di; D is only read and each thread reads back only its
// own column, so no barrier is needed inside.
__device__ __forceinline__ void invert_lower_cols(const double (*D)[NB + 1], double (*Di)[NB + 1]) {
  const int c = threadIdx.x;
  if (c >= NB) return;
  for (int i = 0; i < NB; ++i) {
    if (i < c) {
      Di[i][c] = 0.0;
      continue;
    }
    double s = i == c ? 1.0 : 0.0;
    for (int p = c; p < i; ++p) s -= D[i][p] * Di[p][c];
    Di[i][c] = s / D[i][i];
  }
}

// Factor the diagonal block j0 in LDS (right-looking, one barrier per column:
// the trailing update reads the unscaled column j, whose scaled copy goes to
// Lc), write it back, and keep its inverse in dinv. One wavefront.
__global__ __launch_bounds__(64) void k_chol_diag(int k, double* G, int64_t ldg, int j0, double* dinv, int* info) {
  __shared__ double D[NB][NB + 1];
  __shared__ double Lc[NB][NB + 1];
  __shared__ double Di[NB][NB + 1];
  const int tid = threadIdx.x;
  if (info[0] != 0) return;  // an earlier block failed: leave G as is
  const int nb = min(NB, k - j0);
  for (int t = tid; t < NB * NB; t += blockDim.x) {
    const int r = t % NB, c = t / NB;
    D[r][c] = (r < nb && c < nb) ? G[(int64_t)(j0 + c) * ldg + j0 + r] : (r == c ? 1.0 : 0.0);
    Lc[r][c] = r == c ? 1.0 : 0.0;
  }
  __syncthreads();
  for (int j = 0; j < nb; ++j) {
    const double d = D[j][j];
    if (!(d > 0.0) || !isfinite(d)) {  // the same value in every thread
      if (tid == 0) info[0] = j0 + j + 1;
      return;
    }
    const double sd = sqrt(d), inv = 1.0 / d;
    for (int r = j + tid; r < NB; r += blockDim.x) Lc[r][j] = r == j ? sd : D[r][j] / sd;
    for (int t = tid; t < NB * NB; t += blockDim.x) {
      const int r = t % NB, c = t / NB;
      if (c > j && r >= c && r < nb) D[r][c] -= D[r][j] * D[c][j] * inv;
    }
    __syncthreads();
  }
  for (int t = tid; t < NB * NB; t += blockDim.x) {
    const int r = t % NB, c = t / NB;
    if (r < nb && c < nb && r >= c) G[(int64_t)(j0 + c) * ldg + j0 + r] = Lc[r][c];
  }
  invert_lower_cols(Lc, Di);
  __syncthreads();
  for (int t = tid; t < NB * NB; t += blockDim.x) {
    const int r = t % NB, c = t / NB;
    dinv[c * NB + r] = (r >= c) ? Di[r][c] : 0.0;
  }
}

// L(rows of block i, cols of block jb) = A(...) * Dinv^T, for blocks below jb.
__global__ __launch_bounds__(256) void k_chol_panel(int k, double* G, int64_t ldg, int j0, const double* dinv,
                                                   const int* info) {
  __shared__ double Ab[NB][NB + 1];
  __shared__ double Di[NB][NB + 1];
  if (info[0] != 0) return;
  const int tid = threadIdx.x;
  const int nb = min(NB, k - j0);
  const int r0 = j0 + NB * (1 + blockIdx.x);
  const int rows = min(NB, k - r0);
  for (int t = tid; t < NB * NB; t += blockDim.x) {
    const int r = t % NB, c = t / NB;
    Ab[r][c] = (r < rows && c < nb) ? G[(int64_t)(j0 + c) * ldg + r0 + r] : 0.0;
    Di[r][c] = dinv[c * NB + r];
  }
  __syncthreads();
  for (int t = tid; t < NB * NB; t += blockDim.x) {
    const int r = t % NB, c = t / NB;
    if (r >= rows || c >= nb) continue;
    double s = 0.0;
    for (int p = 0; p <= c; ++p) s += Ab[r][p] * Di[c][p];
    G[(int64_t)(j0 + c) * ldg + r0 + r] = s;
  }
}

// Trailing update of the lower blocks (ti >= tj) right of block column jb.
__global__ __launch_bounds__(256) void k_chol_update(int k, double* G, int64_t ldg, int j0, const int* info) {
  __shared__ double Li[NB][NB + 1];
  __shared__ double Lj[NB][NB + 1];
  if (info[0] != 0) return;
  const int tid = threadIdx.x;
  const int nb = min(NB, k - j0);
  // lower-triangular tile enumeration: x -> (ti, tj), tj <= ti
  const int x = blockIdx.x;
  int ti = static_cast<int>((sqrt(8.0 * x + 1.0) - 1.0) * 0.5);
  while ((ti + 1) * (ti + 2) / 2 <= x) ++ti;
  while (ti * (ti + 1) / 2 > x) --ti;
  const int tj = x - ti * (ti + 1) / 2;
  const int R = j0 + NB + ti * NB, C = j0 + NB + tj * NB;
  for (int t = tid; t < NB * NB; t += blockDim.x) {
    const int r = t % NB, p = t / NB;
    Li[r][p] = (R + r < k && p < nb) ? G[(int64_t)(j0 + p) * ldg + R + r] : 0.0;
    Lj[r][p] = (C + r < k && p < nb) ? G[(int64_t)(j0 + p) * ldg + C + r] : 0.0;
  }
  __syncthreads();
  for (int t = tid; t < NB * NB; t += blockDim.x) {
    const int r = t % NB, c = t / NB;
    if (R + r >= k || C + c >= k || (ti == tj && r < c)) continue;
    double s = 0.0;
    for (int p = 0; p < nb; ++p) s += Li[r][p] * Lj[c][p];
    G[(int64_t)(C + c) * ldg + R + r] -= s;
  }
}

// Invert every diagonal block of L; write its transpose into Rinv's diagonal
// block (Rinv = L^{-T} is upper triangular) and keep the inverse in dinv.
__global__ __launch_bounds__(64) void k_trtri_diag(int k, const double* __restrict__ L, int64_t ldl, double* dinv,
                                                   double* Rinv, int64_t ldr) {
  __shared__ double D[NB][NB + 1];
  __shared__ double Di[NB][NB + 1];
  const int tid = threadIdx.x;
  const int j0 = blockIdx.x * NB;
  const int nb = min(NB, k - j0);
  for (int t = tid; t < NB * NB; t += blockDim.x) {
    const int r = t % NB, c = t / NB;
    D[r][c] = (r < nb && c < nb) ? (r >= c ? L[(int64_t)(j0 + c) * ldl + j0 + r] : 0.0) : (r == c ? 1.0 : 0.0);
  }
  __syncthreads();
  invert_lower_cols(D, Di);
  __syncthreads();
  double* dv = dinv + (int64_t)blockIdx.x * NB * NB;
  for (int t = tid; t < NB * NB; t += blockDim.x) {
    const int r = t % NB, c = t / NB;
    const double v = (r >= c) ? Di[r][c] : 0.0;
    dv[c * NB + r] = v;
    // Rinv(j0 + c, j0 + r) = Linv(j0 + r, j0 + c)
    if (r < nb && c < nb) Rinv[(int64_t)(j0 + r) * ldr + j0 + c] = v;
  }
}

// Block diagonal d of L^{-1}: X_ij = -Dinv_i * sum_{p=j}^{i-1} L_ip X_pj, i = j + d,
// with X_pj read back from Rinv (X_pj = Rinv_jp^T). Written as Rinv_ji = X_ij^T.
__global__ __launch_bounds__(256) void k_trtri_off(int k, const double* __restrict__ L, int64_t ldl,
                                                   const double* __restrict__ dinv, double* Rinv, int64_t ldr,
                                                   int d) {
  __shared__ double Lb[NB][NB + 1];
  __shared__ double Xb[NB][NB + 1];
  __shared__ double Acc[NB][NB + 1];
  const int tid = threadIdx.x;
  const int jb = blockIdx.x, ib = jb + d;
  const int i0 = ib * NB, j0 = jb * NB;
  const int ni = min(NB, k - i0), nj = min(NB, k - j0);
  for (int t = tid; t < NB * NB; t += blockDim.x) Acc[t % NB][t / NB] = 0.0;
  for (int pb = jb; pb < ib; ++pb) {
    const int p0 = pb * NB;
    const int np = min(NB, k - p0);
    __syncthreads();
    for (int t = tid; t < NB * NB; t += blockDim.x) {
      const int r = t % NB, c = t / NB;
      Lb[r][c] = (r < ni && c < np) ? L[(int64_t)(p0 + c) * ldl + i0 + r] : 0.0;           // L_ip[r][c]
      Xb[r][c] = (r < np && c < nj) ? Rinv[(int64_t)(p0 + r) * ldr + j0 + c] : 0.0;        // X_pj[r][c]
    }
    __syncthreads();
    for (int t = tid; t < NB * NB; t += blockDim.x) {
      const int r = t % NB, c = t / NB;
      double s = 0.0;
      for (int q = 0; q < NB; ++q) s += Lb[r][q] * Xb[q][c];
      Acc[r][c] += s;
    }
  }
  __syncthreads();
  const double* dv = dinv + (int64_t)ib * NB * NB;  // Dinv_i, column-major
  for (int t = tid; t < NB * NB; t += blockDim.x) {
    const int r = t % NB, c = t / NB;
    if (r >= ni || c >= nj) continue;
    double s = 0.0;
    for (int q = 0; q <= r; ++q) s += dv[q * NB + r] * Acc[q][c];
    Rinv[(int64_t)(i0 + r) * ldr + j0 + c] = -s;
  }
}

}  // namespace fe

using namespace fe;

extern "C" int fe_potrf(fe_ctx ctx, int k, double* G, int64_t ldg) {
  FE_ARG(ctx && G && k >= 0 && ldg >= k, "fe_potrf: bad argument");
  if (k == 0) return 0;
  const int nblk = (k + NB - 1) / NB;
  void* ws = nullptr;
  int rc = scratch(ctx, 1, sizeof(double) * NB * NB, &ws);
  if (rc) return rc;
  rc = fe_zero2d(ctx, ctx->info_dev, 1, 1, 1, 4);  // a fill kernel, not a memset node
  if (rc) return rc;
  for (int jb = 0; jb < nblk; ++jb) {
    const int j0 = jb * NB;
    hipLaunchKernelGGL(k_chol_diag, dim3(1), dim3(64), 0, ctx->stream, k, G, ldg, j0, (double*)ws, ctx->info_dev);
    const int below = nblk - jb - 1;
    if (below > 0) {
      hipLaunchKernelGGL(k_chol_panel, dim3(below), dim3(256), 0, ctx->stream, k, G, ldg, j0, (const double*)ws,
                         ctx->info_dev);
      hipLaunchKernelGGL(k_chol_update, dim3(below * (below + 1) / 2), dim3(256), 0, ctx->stream, k, G, ldg, j0,
                         ctx->info_dev);
    }
  }
  FE_CHECK_LAUNCH();
  int info = 0;
  FE_HIP_TRY(hipMemcpyAsync(&info, ctx->info_dev, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
  FE_HIP_TRY(hipStreamSynchronize(ctx->stream));
  return info;
}

extern "C" int fe_chol_rinv(fe_ctx ctx, int k, const double* L, int64_t ldl, double* Rinv, int64_t ldr) {
  FE_ARG(ctx && L && Rinv && k >= 0 && ldl >= k && ldr >= k, "fe_chol_rinv: bad argument");
  if (k == 0) return 0;
  const int nblk = (k + NB - 1) / NB;
  void* ws = nullptr;
  int rc = scratch(ctx, 1, sizeof(double) * NB * NB * nblk, &ws);
  if (rc) return rc;
  rc = fe_zero2d(ctx, Rinv, k, k, ldr, 8);
  if (rc) return rc;
  hipLaunchKernelGGL(k_trtri_diag, dim3(nblk), dim3(64), 0, ctx->stream, k, L, ldl, (double*)ws, Rinv, ldr);
  for (int d = 1; d < nblk; ++d)
    hipLaunchKernelGGL(k_trtri_off, dim3(nblk - d), dim3(256), 0, ctx->stream, k, L, ldl, (const double*)ws, Rinv,
                       ldr, d);
  FE_CHECK_LAUNCH();
  return 0;
}
