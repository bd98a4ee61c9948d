// Blocked FP64 Cholesky factorization and triangular inverse of the k x k Gram
// matrix on the device (column-major), block size NB = 32:
//
//   potrf  for each block column jb:
//            k_chol_diag    factor the diagonal block and invert it in one
//                           right-looking pass over the registers of one
//                           wavefront, keep the inverse
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

// The 32 x 32 diagonal blocks live in registers of one wavefront: lane l
// holds the 4 x 4 sub-block (bi, bj) = (l >> 3, l & 7), rows 4 bi .. 4 bi + 3
// and columns 4 bj .. 4 bj + 3. Step j publishes column j of L and row j of
// L^{-1} (32 values each) through LDS, with zeros where the update must not
// apply, so every lane runs the same unmasked FMAs; the diagonal entry comes
// from its owner lane by readlane. Entries above the diagonal are never read
// back or stored, so what the unmasked updates leave there does not matter,
// except in L^{-1}, whose upper part stays exactly zero (every row j it is
// updated with has zeros right of column j). The LDS version of this kernel,
// LDS round trips per entry and a correctly rounded divide per step, took
// 60-84 us per block.
struct Sub {
  int bi, bj;
  __device__ Sub() : bi(threadIdx.x >> 3), bj(threadIdx.x & 7) {}
};

__device__ __forceinline__ double readlane_f64(double v, int lane) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane(static_cast<int>(b), lane);
  const int hi = __builtin_amdgcn_readlane(static_cast<int>(b >> 32), lane);
  return __longlong_as_double((static_cast<long long>(hi) << 32) | static_cast<unsigned>(lo));
}

// Lower triangle of the nb x nb block at (j0, j0) of the column-major matrix
// M; rows and columns past nb are those of the identity. Every lane issues
// its 16 loads (invalid ones from the block's first entry) before using any.
__device__ __forceinline__ void load_lower(const Sub& s, const double* M, int64_t ldm, int j0, int nb,
                                           double (&a)[4][4]) {
  const double* base = M + (int64_t)j0 * ldm + j0;
#pragma unroll
  for (int c = 0; c < 4; ++c)
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int R = 4 * s.bi + r, C = 4 * s.bj + c;
      const bool ok = R >= C && R < nb && C < nb;
      a[r][c] = base[ok ? (int64_t)C * ldm + R : 0];
    }
#pragma unroll
  for (int c = 0; c < 4; ++c)
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int R = 4 * s.bi + r, C = 4 * s.bj + c;
      if (!(R >= C && R < nb && C < nb)) a[r][c] = R == C ? 1.0 : 0.0;
    }
}

// FACTOR: Cholesky of the block in a (lower triangle in, L out) and x = L^{-1},
// right-looking, in one pass. Otherwise a holds L already and only x = L^{-1}
// is formed. Returns the first column whose pivot is not positive and finite,
// or -1 (the same in every lane).
template <bool FACTOR>
__device__ int chol_inv_regs(const Sub& s, double (&a)[4][4], double (&x)[4][4], double* col, double* row) {
#pragma unroll
  for (int r = 0; r < 4; ++r)
#pragma unroll
    for (int c = 0; c < 4; ++c) x[r][c] = (s.bi == s.bj && r == c) ? 1.0 : 0.0;
  for (int jb = 0; jb < NB / 4; ++jb) {
#pragma unroll
    for (int jj = 0; jj < 4; ++jj) {
      const int j = 4 * jb + jj;
      const double d = readlane_f64(a[jj][jj], 9 * jb);  // lane 9 jb holds sub-block (jb, jb)
      double inv;  // 1 / L[j][j]
      if (FACTOR) {
        if (!(d > 0.0) || !isfinite(d)) return j;
        inv = rsqrt_nr(d);
      } else {
        inv = rcp_nr(d);
      }
      if (s.bj == jb) {
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int R = 4 * s.bi + r;
          if (FACTOR) {
            if (R > j) a[r][jj] *= inv;
            else if (R == j) a[r][jj] = d * inv;
          }
          col[R] = R > j ? a[r][jj] : 0.0;  // all 32 rows: no stale entries
        }
      }
      if (s.bi == jb) {
#pragma unroll
        for (int c = 0; c < 4; ++c) {
          x[jj][c] *= inv;  // row j of L^{-1} is final
          row[4 * s.bj + c] = x[jj][c];
        }
      }
      __syncthreads();
      double lr[4], lc[4], xr[4];
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        lr[e] = col[4 * s.bi + e];
        xr[e] = row[4 * s.bj + e];
        if (FACTOR) lc[e] = col[4 * s.bj + e];
      }
#pragma unroll
      for (int r = 0; r < 4; ++r)
#pragma unroll
        for (int c = 0; c < 4; ++c) {
          if (FACTOR) a[r][c] -= lr[r] * lc[c];
          x[r][c] -= lr[r] * xr[c];
        }
      __syncthreads();
    }
  }
  return -1;
}

// x as the column-major NB x NB block dv (zeros above the diagonal).
__device__ __forceinline__ void store_block(const Sub& s, const double (&x)[4][4], double* dv) {
#pragma unroll
  for (int c = 0; c < 4; ++c)
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int R = 4 * s.bi + r, C = 4 * s.bj + c;
      dv[C * NB + R] = R >= C ? x[r][c] : 0.0;
    }
}

// Factor the diagonal block j0, write L back into G's lower triangle, and
// keep L^{-1} in dinv (column-major NB x NB). One wavefront.
__global__ __launch_bounds__(64) void k_chol_diag(int k, double* G, int64_t ldg, int j0, double* dinv, int* info) {
  __shared__ double col[NB], row[NB];
  if (info[0] != 0) return;  // an earlier block failed: leave G as is
  const Sub s;
  const int nb = min(NB, k - j0);
  double a[4][4], x[4][4];
  load_lower(s, G, ldg, j0, nb, a);
  const int bad = chol_inv_regs<true>(s, a, x, col, row);
  if (bad >= 0) {
    if (threadIdx.x == 0) info[0] = j0 + bad + 1;
    return;
  }
  if (s.bi >= s.bj) {
#pragma unroll
    for (int c = 0; c < 4; ++c)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int R = 4 * s.bi + r, C = 4 * s.bj + c;
        if (R < nb && C < nb && R >= C) G[(int64_t)(j0 + C) * ldg + j0 + R] = a[r][c];
      }
  }
  store_block(s, x, dinv);
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
// One wavefront per block.
__global__ __launch_bounds__(64) void k_trtri_diag(int k, const double* __restrict__ L, int64_t ldl, double* dinv,
                                                   double* Rinv, int64_t ldr) {
  __shared__ double col[NB], row[NB];
  const Sub s;
  const int j0 = blockIdx.x * NB;
  const int nb = min(NB, k - j0);
  double a[4][4], x[4][4];
  load_lower(s, L, ldl, j0, nb, a);
  chol_inv_regs<false>(s, a, x, col, row);
  store_block(s, x, dinv + (int64_t)blockIdx.x * NB * NB);
#pragma unroll
  for (int c = 0; c < 4; ++c)
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int R = 4 * s.bi + r, C = 4 * s.bj + c;
      // Rinv(j0 + C, j0 + R) = Linv(j0 + R, j0 + C)
      if (R < nb && C < nb && R >= C) Rinv[(int64_t)(j0 + R) * ldr + j0 + C] = x[r][c];
    }
}

// Block diagonal d of L^{-1}: X_ij = -Dinv_i * sum_{p=j}^{i-1} L_ip X_pj, i = j + d,
// with X_pj read back from Rinv (X_pj = Rinv_jp^T). Written as Rinv_ji = X_ij^T.
// Thread t accumulates row t & 31, columns 4 (t >> 5) .. + 3 of the sum in
// registers; the tiles of the next p are loaded while the current ones are
// multiplied.
__global__ __launch_bounds__(256) void k_trtri_off(int k, const double* __restrict__ L, int64_t ldl,
                                                   const double* __restrict__ dinv, double* Rinv, int64_t ldr,
                                                   int d) {
  __shared__ double Lb[NB][NB + 1];
  __shared__ double Xb[NB][NB + 4];
  const int tid = threadIdx.x;
  const int jb = blockIdx.x, ib = jb + d;
  const int i0 = ib * NB, j0 = jb * NB;
  const int ni = min(NB, k - i0), nj = min(NB, k - j0);
  const int r = tid & 31, c0 = 4 * (tid >> 5);
  double lv[4], xv[4];
  // staging element h: L_ip[tid & 31][(tid >> 5) + 8 h] (columns of L are
  // contiguous), X_pj[(tid >> 5) + 8 h][tid & 31] (rows of Rinv are contiguous)
  auto fetch = [&](int pb) {
    const int p0 = pb * NB, np = min(NB, k - p0);
#pragma unroll
    for (int h = 0; h < 4; ++h) {
      const int q = (tid >> 5) + 8 * h;
      lv[h] = (r < ni && q < np) ? L[(int64_t)(p0 + q) * ldl + i0 + r] : 0.0;
      xv[h] = (q < np && r < nj) ? Rinv[(int64_t)(p0 + q) * ldr + j0 + r] : 0.0;
    }
  };
  double acc[4] = {0.0, 0.0, 0.0, 0.0};
  fetch(jb);
  for (int pb = jb; pb < ib; ++pb) {
    __syncthreads();
#pragma unroll
    for (int h = 0; h < 4; ++h) {
      const int q = (tid >> 5) + 8 * h;
      Lb[r][q] = lv[h];
      Xb[q][r] = xv[h];
    }
    __syncthreads();
    if (pb + 1 < ib) fetch(pb + 1);
#pragma unroll 8
    for (int q = 0; q < NB; ++q) {
      const double l = Lb[r][q];
#pragma unroll
      for (int c = 0; c < 4; ++c) acc[c] += l * Xb[q][c0 + c];
    }
  }
  // Rinv_ji^T = -Dinv_i * Acc: Acc through LDS, Dinv_i (column-major) into Lb
  const double* dv = dinv + (int64_t)ib * NB * NB;
  __syncthreads();
#pragma unroll
  for (int c = 0; c < 4; ++c) Xb[r][c0 + c] = acc[c];
#pragma unroll
  for (int h = 0; h < 4; ++h) {
    const int q = (tid >> 5) + 8 * h;
    Lb[r][q] = dv[q * NB + r];
  }
  __syncthreads();
  double out[4] = {0.0, 0.0, 0.0, 0.0};
  for (int q = 0; q <= r; ++q) {
    const double l = Lb[r][q];
#pragma unroll
    for (int c = 0; c < 4; ++c) out[c] += l * Xb[q][c0 + c];
  }
  if (r < ni) {
#pragma unroll
    for (int c = 0; c < 4; ++c)
      if (c0 + c < nj) Rinv[(int64_t)(i0 + r) * ldr + j0 + c0 + c] = -out[c];
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
