// MFMA kernels for the tall-skinny dense products of the eigensolver:
//   Gram   G = X^T diag(w) Y   (split-K over row chunks, fixed-order reduction)
//   GEMM   Y = alpha X C + beta Y (C small, FP64 on the device)
//
// Both use 16x16x4 MFMA tiles (FP64: v_mfma_f64_16x16x4_f64, FP32:
// v_mfma_f32_16x16x4_f32), one operand element per lane:
//   A operand: lane l holds A[l % 16][l / 16]      (16 x 4)
//   B operand: lane l holds B[l / 16][l % 16]      (4 x 16)
//   D result : lane l, item i holds D[4 (l / 16) + i][l % 16] (FP32)
//              and D[l / 16 + 4 i][l % 16] (FP64: rows interleaved)
// The compensated BF16 products (FP32 blocks) split each operand into 2 or 3
// bfloat16 parts and run v_mfma_f32_16x16x32_bf16 (8 bf16 per lane: lane l
// holds A[l % 16][8 (l / 16) + 0..7] and B[8 (l / 16) + 0..7][l % 16]).
#include "fe_common.hpp"

#include <cstdlib>

#include <algorithm>
#include <type_traits>

// Fixed device scratch of fe_gram_static (split-K partial tiles): every GPU the
// library runs on holds its own copy until the process ends. Host code never
// uses this name as an address: the per-device address comes from
// hipGetSymbolAddress (gram_static_init) and travels in fe_ctx_s::gram_static.
static __device__ __align__(256) double scratch[FE_GRAM_STATIC_BYTES / sizeof(double)];

namespace fe {

typedef double d4_t __attribute__((ext_vector_type(4)));
typedef float f4_t __attribute__((ext_vector_type(4)));
typedef __bf16 bf8_t __attribute__((ext_vector_type(8)));

template <typename T>
struct mfma;
template <>
struct mfma<double> {
  using acc = d4_t;
  __device__ static __forceinline__ acc run(double a, double b, acc c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  // output row of accumulator item e in lane `lane`
  __device__ static __forceinline__ int row(int lane, int e) { return (lane >> 4) + 4 * e; }
};
template <>
struct mfma<float> {
  using acc = f4_t;
  __device__ static __forceinline__ acc run(float a, float b, acc c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  __device__ static __forceinline__ int row(int lane, int e) { return 4 * (lane >> 4) + e; }
};

constexpr int TILE = 64;   // output tile edge per workgroup
constexpr int KCG = 32;    // rows per LDS stage in the Gram kernel
constexpr int KCM = 16;    // inner-dimension chunk in the GEMM kernel
constexpr int PAD = 1;     // LDS row padding (elements)

// Tile t of the output: row-major over (ta, tb) tiles, or, for a symmetric
// Gram (sym != 0), over the upper triangle ta <= tb only.
__device__ __forceinline__ void gram_tile(int t, int tiles_b, int sym, int& ta, int& tb) {
  if (!sym) {
    ta = t / tiles_b;
    tb = t % tiles_b;
    return;
  }
  int a = 0;
  while (t >= tiles_b - a) {
    t -= tiles_b - a;
    ++a;
  }
  ta = a;
  tb = a + t;
}

// ---------------------------------------------------------------- Gram (native)
// Work unit u = chunk * ntiles + tile; units map to XCDs in contiguous ranges,
// so all tiles of a row chunk run on one XCD and share its X / Y rows in L2.
// A workgroup computes a 128 x 128 tile (4 wavefronts of 64 x 64 = 4 x 4 MFMA
// blocks) over the chunk's rows, staging GKS rows of X and Y per pass with
// 16-byte loads; the next pass's rows are loaded into registers while the
// current one runs on MFMA.
constexpr int WT = 128;   // workgroup tile edge (Gram, GEMM)
constexpr int GKS = 16;   // rows per LDS stage (Gram)
constexpr int GPAD = 16;  // LDS row padding (elements): row stride = 16 banks mod 64

template <typename T>
__device__ __forceinline__ void load8(const T* p, bool vec, int valid, T (&v)[8]) {
  if (vec && valid >= 8) {
    if constexpr (sizeof(T) == 4) {
      const float4 a = *reinterpret_cast<const float4*>(p), b = *reinterpret_cast<const float4*>(p + 4);
      __builtin_memcpy(&v[0], &a, 16);
      __builtin_memcpy(&v[4], &b, 16);
    } else {
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const double2 a = *reinterpret_cast<const double2*>(p + 2 * q);
        v[2 * q] = a.x;
        v[2 * q + 1] = a.y;
      }
    }
  } else {
#pragma unroll
    for (int e = 0; e < 8; ++e) v[e] = e < valid ? p[e] : T(0);
  }
}

template <typename T>
__global__ __launch_bounds__(256, 2) void k_gram(int64_t n, int m, int p, const T* __restrict__ X, int64_t ldx,
                                              const T* __restrict__ Y, int64_t ldy,
                                              const double* __restrict__ w, int tiles_b, int ntiles,
                                              int sym, int64_t rows_per, int64_t per_xcd, int64_t nunits,
                                              int vec, double* __restrict__ part) {
  // rows per stage: 32, two per staging thread (for FP64 too: 16 rows per
  // stage kept the FP64 MFMA pipe 68% busy at n 884,736, m 512; 32 rows, 75%)
  constexpr int GK = 2 * GKS, NR = GK / 16;
  __shared__ T Xs[GK][WT + GPAD];
  __shared__ T Ys[GK][WT + GPAD];
  const int64_t u = xcd_tile(blockIdx.x, per_xcd);
  if (u >= nunits) return;
  const int64_t chunk = u / ntiles;
  const int tile = static_cast<int>(u - chunk * ntiles);
  int ta, tb;
  gram_tile(tile, tiles_b, sym, ta, tb);
  const int a0 = ta * WT, b0 = tb * WT;
  const int64_t r0 = chunk * rows_per;
  const int64_t r1 = r0 + rows_per < n ? r0 + rows_per : n;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave >> 1, wn = wave & 1;
  // valid 16-wide blocks of the wavefront's 64 x 64 tile (wave-uniform)
  const int vi = min(4, max(0, (m - a0 - wm * 64 + 15) >> 4)), vj = min(4, max(0, (p - b0 - wn * 64 + 15) >> 4));
  using acc_t = typename mfma<T>::acc;
  acc_t acc[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[i][j] = acc_t{};
  // staging: thread -> rows tid / 16 + 16 h (h < NR), columns (tid % 16) * 8 .. + 7 of the tile
  const int sr = tid >> 4, sc = (tid & 15) * 8;
  const int va = m - (a0 + sc), vb = p - (b0 + sc);
  T xr[NR][8], yr[NR][8];
  auto fetch = [&](int64_t rb) {
#pragma unroll
    for (int h = 0; h < NR; ++h) {
      const int64_t r = rb + sr + 16 * h;
      if (r < r1) {
        load8(X + r * ldx + a0 + sc, vec, va, xr[h]);
        load8(Y + r * ldy + b0 + sc, vec, vb, yr[h]);
        if (w) {
          const double wr = w[r];
#pragma unroll
          for (int e = 0; e < 8; ++e) yr[h][e] = static_cast<T>(static_cast<double>(yr[h][e]) * wr);
        }
      } else {
#pragma unroll
        for (int e = 0; e < 8; ++e) xr[h][e] = yr[h][e] = T(0);
      }
    }
  };
  fetch(r0);
  for (int64_t rb = r0; rb < r1; rb += GK) {
    __syncthreads();
#pragma unroll
    for (int h = 0; h < NR; ++h)
#pragma unroll
      for (int e = 0; e < 8; ++e) {
        Xs[sr + 16 * h][sc + e] = xr[h][e];
        Ys[sr + 16 * h][sc + e] = yr[h][e];
      }
    __syncthreads();
    if (rb + GK < r1) fetch(rb + GK);
#pragma unroll
    for (int kk = 0; kk < GK; kk += 4) {
      const int kr = kk + (lane >> 4);
      T af[4], bf[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) af[i] = Xs[kr][wm * 64 + i * 16 + (lane & 15)];
#pragma unroll
      for (int j = 0; j < 4; ++j) bf[j] = Ys[kr][wn * 64 + j * 16 + (lane & 15)];
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j)
          if (i < vi && j < vj) acc[i][j] = mfma<T>::run(af[i], bf[j], acc[i][j]);
    }
  }
  // partial tile, column-major 128 x 128
  double* P = part + u * (WT * WT);
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j)
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const int row = wm * 64 + i * 16 + mfma<T>::row(lane, e);
        const int col = wn * 64 + j * 16 + (lane & 15);
        P[col * WT + row] = static_cast<double>(acc[i][j][e]);
      }
}

// Gram with 64 x 64 tiles for thin or small operands: wavefront w owns the
// 16-row strip 16w..16w+15 of the tile (four 16 x 16 MFMA blocks, blocks
// beyond the valid columns skipped), so the accumulators take 32 VGPRs and
// four workgroups fit on a CU to hide the load latency of the row stream.
constexpr int G64 = 64;   // tile edge
constexpr int G64R = 32;  // rows per LDS stage

template <typename T>
__global__ __launch_bounds__(256, 4) void k_gram64(int64_t n, int m, int p, const T* __restrict__ X, int64_t ldx,
                                                const T* __restrict__ Y, int64_t ldy,
                                                const double* __restrict__ w, int tiles_b, int ntiles, int sym,
                                                int64_t rows_per, int64_t per_xcd, int64_t nunits, int vec,
                                                double* __restrict__ part) {
  __shared__ T Xs[G64R][G64 + GPAD];
  __shared__ T Ys[G64R][G64 + GPAD];
  const int64_t u = xcd_tile(blockIdx.x, per_xcd);
  if (u >= nunits) return;
  const int64_t chunk = u / ntiles;
  const int tile = static_cast<int>(u - chunk * ntiles);
  int ta, tb;
  gram_tile(tile, tiles_b, sym, ta, tb);
  const int a0 = ta * G64, b0 = tb * G64;
  const int64_t r0 = chunk * rows_per;
  const int64_t r1 = r0 + rows_per < n ? r0 + rows_per : n;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  // valid 16-wide blocks (wave-uniform): this wave's strip, and the columns
  const bool strip = m - a0 - 16 * wave > 0;
  const int vj = strip ? min(4, max(0, (p - b0 + 15) >> 4)) : 0;
  using acc_t = typename mfma<T>::acc;
  acc_t acc[4];
#pragma unroll
  for (int j = 0; j < 4; ++j) acc[j] = acc_t{};
  // staging: thread -> row tid / 8, columns (tid % 8) * 8 .. + 7 of the tile
  const int sr = tid >> 3, sc = (tid & 7) * 8;
  const int va = m - (a0 + sc), vb = p - (b0 + sc);
  T xr[8], yr[8];
  auto fetch = [&](int64_t rb) {
    const int64_t r = rb + sr;
    if (r < r1) {
      load8(X + r * ldx + a0 + sc, vec, va, xr);
      load8(Y + r * ldy + b0 + sc, vec, vb, yr);
      if (w) {
        const double wr = w[r];
#pragma unroll
        for (int e = 0; e < 8; ++e) yr[e] = static_cast<T>(static_cast<double>(yr[e]) * wr);
      }
    } else {
#pragma unroll
      for (int e = 0; e < 8; ++e) xr[e] = yr[e] = T(0);
    }
  };
  fetch(r0);
  for (int64_t rb = r0; rb < r1; rb += G64R) {
    __syncthreads();
#pragma unroll
    for (int e = 0; e < 8; ++e) {
      Xs[sr][sc + e] = xr[e];
      Ys[sr][sc + e] = yr[e];
    }
    __syncthreads();
    if (rb + G64R < r1) fetch(rb + G64R);
#pragma unroll
    for (int kk = 0; kk < G64R; kk += 4) {
      const int kr = kk + (lane >> 4);
      const T af = Xs[kr][16 * wave + (lane & 15)];
      T bf[4];
#pragma unroll
      for (int j = 0; j < 4; ++j) bf[j] = Ys[kr][j * 16 + (lane & 15)];
#pragma unroll
      for (int j = 0; j < 4; ++j)
        if (j < vj) acc[j] = mfma<T>::run(af, bf[j], acc[j]);
    }
  }
  // partial tile, column-major 64 x 64
  double* P = part + u * (G64 * G64);
#pragma unroll
  for (int j = 0; j < 4; ++j)
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const int row = 16 * wave + mfma<T>::row(lane, e);
      const int col = j * 16 + (lane & 15);
      P[col * G64 + row] = static_cast<double>(acc[j][e]);
    }
}

// ---------------------------------------------------------------- fused Ritz residual
// Epilogue of the rotation AY = AQ C of a Rayleigh-Ritz extraction: per
// element v = AY[row, col] (stored if `store`), R[row, col] = v - theta[col]
// Y[row, col] with Y = Q C read from memory (R stored if not null), and the
// column sums of R^2 and Y^2 over the workgroup's rows into part_r / part_y
// [row tile][p]. A lane sums its rows of each of its four columns, the four
// lanes of a column are added by shuffles, the row wavefronts in order
// through LDS: a fixed order, so the sums are reproducible.
template <typename T>
struct ResArgs {
  const T* Y;
  int64_t ldy;
  const double* theta;
  T* R;
  int64_t ldr;
  double* part_r;
  double* part_y;
  int store;
};

// acc[i][j] of the wavefront's 64 x 64 block at rows rw0 (+ 16 i + row(lane, e))
// and columns cw0 (+ 16 j + lane % 16); red: 2 NWM TC doubles of LDS
template <typename T, typename ROW, int NWM, int TC, typename ACC>
__device__ __forceinline__ void resid_epilogue(const ACC (&acc)[4][4], int64_t n, int p, int64_t rw0, int cw0,
                                               T alpha, T* AY, int64_t lday, const ResArgs<T>& res, double* red,
                                               int wm, int wn, int64_t rt, int c0) {
  const int tid = threadIdx.x, lane = tid & 63;
  // opaque copies of the tile origin: the element addresses are then formed
  // here, after the product loop, instead of being hoisted above it (where
  // they held registers through the loop and spilled)
  int64_t rb = rw0;
  int cb = cw0;
  asm volatile("" : "+v"(rb), "+v"(cb));
  double sr[4], sy[4];
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    sr[j] = sy[j] = 0.0;
    const int col = cb + j * 16 + (lane & 15);
    if (col >= p) continue;
    const double th = res.theta[col];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const int64_t row = rb + i * 16 + ROW::row(lane, e);
        if (row < n) {
          const T v = alpha * static_cast<T>(acc[i][j][e]);
          if (res.store) AY[row * lday + col] = v;
          const double yv = static_cast<double>(res.Y[row * res.ldy + col]);
          const T rv = static_cast<T>(static_cast<double>(v) - th * yv);
          if (res.R) res.R[row * res.ldr + col] = rv;
          sr[j] += static_cast<double>(rv) * static_cast<double>(rv);
          sy[j] += yv * yv;
        }
      }
  }
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    sr[j] += __shfl_xor(sr[j], 16, 64);
    sy[j] += __shfl_xor(sy[j], 16, 64);
    sr[j] += __shfl_xor(sr[j], 32, 64);
    sy[j] += __shfl_xor(sy[j], 32, 64);
  }
  __syncthreads();  // the LDS of the product stages is free
  if (lane < 16)
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      red[(2 * wm) * TC + wn * 64 + j * 16 + lane] = sr[j];
      red[(2 * wm + 1) * TC + wn * 64 + j * 16 + lane] = sy[j];
    }
  __syncthreads();
  if (tid < TC && c0 + tid < p) {
    double a = 0.0, b = 0.0;
#pragma unroll
    for (int w = 0; w < NWM; ++w) {
      a += red[(2 * w) * TC + tid];
      b += red[(2 * w + 1) * TC + tid];
    }
    res.part_r[rt * p + c0 + tid] = a;
    res.part_y[rt * p + c0 + tid] = b;
  }
}

// ---------------------------------------------------------------- compensated BF16 products
// An FP32 value x is split into NS bfloat16 parts, x_0 = bf16(x),
// x_1 = bf16(x - x_0), x_2 = bf16(x - x_0 - x_1) (each difference is exact in
// FP32), and a product x y is summed on BF16 MFMA (v_mfma_f32_16x16x32_bf16,
// FP32 accumulation) over the part pairs s + t < NS:
//   NS = 2 (BF16x3): x_1 y_0 + x_0 y_1 + x_0 y_0, about 16 significant bits;
//   NS = 3 (BF16x6): the six pairs s + t <= 2, within a few units of the FP32
//                    rounding of the product,
// smallest terms first. Every element is split once, when its stage is
// written to LDS, where the parts sit k-contiguous so that a lane reads its
// 8 consecutive k values of a part with one 16-byte load.
typedef float f2v_t __attribute__((ext_vector_type(2)));
typedef __bf16 b2v_t __attribute__((ext_vector_type(2)));
typedef uint32_t u4v_t __attribute__((ext_vector_type(4)));  // 16 bytes of packed bf16

// two FP32 values -> NS packed bfloat16 pairs (first value in the low half)
template <int NS>
__device__ __forceinline__ void split_pair(float a, float b, uint32_t (&o)[NS]) {
#pragma unroll
  for (int s = 0; s < NS; ++s) {
    const uint32_t bits = __builtin_bit_cast(uint32_t, __builtin_convertvector(f2v_t{a, b}, b2v_t));
    o[s] = bits;
    if (s + 1 < NS) {
      a -= __uint_as_float(bits << 16);
      b -= __uint_as_float(bits & 0xffff0000u);
    }
  }
}

// acc += sum over part pairs s + t < NS of a[s] b[t], smallest first
template <int NS>
__device__ __forceinline__ f4_t mfma_split(const bf8_t (&a)[NS], const bf8_t (&b)[NS], f4_t acc) {
  if constexpr (NS == 3) {
    acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[2], b[0], acc, 0, 0, 0);
    acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[1], b[1], acc, 0, 0, 0);
    acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[0], b[2], acc, 0, 0, 0);
  }
  acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[1], b[0], acc, 0, 0, 0);
  acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[0], b[1], acc, 0, 0, 0);
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[0], b[0], acc, 0, 0, 0);
}

constexpr int BK = 32;       // k values per stage (one 16x16x32 MFMA)
constexpr int BKP = BK + 8;  // LDS row stride in bf16: 80 bytes, 16-byte reads of 16 lanes hit distinct banks

// Gram G = X^T diag(w) Y on GT x GT = 256 x 256 tiles, split-K over row
// chunks as k_gram, 512 threads (8 wavefronts of 128 x 64). Each staged row is
// used for a 256-wide slice of the output: with 128 x 128 tiles the tile
// loads alone moved 14.5 GB per 884,736 x 512 Gram and held it at the native
// FP32 product's time. Stage: 32 rows; thread t (lane l of wavefront w) loads
// rows 4 (l & 7) .. + 3 and columns 4 ((l >> 3) + 8 w) .. + 3 of X and Y (four
// 16-byte loads each) and writes each column's 4 rows as 8-byte runs of the
// parts, laid out [column][row] (eight row groups per wavefront keep the
// transposed writes at two lanes per bank). For a diagonal tile of a
// symmetric Gram the Y stage is the X stage.
constexpr int GT = 256;
template <int NS>
__global__ __launch_bounds__(512, 1) void k_gram_bf16(int64_t n, int m, int p, const float* __restrict__ X,
                                                   int64_t ldx, const float* __restrict__ Y, int64_t ldy,
                                                   const double* __restrict__ w, int tiles_b, int ntiles, int sym,
                                                   int64_t rows_per, int64_t per_xcd, int64_t nunits, int vec,
                                                   double* __restrict__ part) {
  __shared__ __attribute__((aligned(16))) __bf16 Xl[NS][GT][BKP];
  __shared__ __attribute__((aligned(16))) __bf16 Yl[NS][GT][BKP];
  const int64_t u = xcd_tile(blockIdx.x, per_xcd);
  if (u >= nunits) return;
  const int64_t chunk = u / ntiles;
  const int tile = static_cast<int>(u - chunk * ntiles);
  int ta, tb;
  gram_tile(tile, tiles_b, sym, ta, tb);
  const bool same = sym == 1 && ta == tb;  // identical operands (workgroup-uniform)
  const int a0 = ta * GT, b0 = tb * GT;
  const int64_t r0 = chunk * rows_per;
  const int64_t r1 = r0 + rows_per < n ? r0 + rows_per : n;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave >> 2, wn = wave & 3;  // rows wm * 128 .. + 127, columns wn * 64 .. + 63
  const int vi = min(8, max(0, (m - a0 - wm * 128 + 15) >> 4)), vj = min(4, max(0, (p - b0 - wn * 64 + 15) >> 4));
  f4_t acc[8][4];
#pragma unroll
  for (int i = 0; i < 8; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[i][j] = f4_t{};
  const int rb = 4 * (lane & 7), cb = 4 * ((lane >> 3) + 8 * wave);
  const int va = m - (a0 + cb), vb = p - (b0 + cb);
  float xr[4][4], yr[4][4];
  auto load4 = [&](const float* q, int valid, float (&v)[4]) {
    if (vec && valid >= 4) {
      const float4 t = *reinterpret_cast<const float4*>(q);
      v[0] = t.x, v[1] = t.y, v[2] = t.z, v[3] = t.w;
    } else {
#pragma unroll
      for (int e = 0; e < 4; ++e) v[e] = e < valid ? q[e] : 0.0f;
    }
  };
  auto fetch = [&](int64_t r) {
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int64_t row = r + rb + i;
      if (row < r1) {
        load4(X + row * ldx + a0 + cb, va, xr[i]);
        if (!same) {
          load4(Y + row * ldy + b0 + cb, vb, yr[i]);
          if (w) {
            const double wr = w[row];
#pragma unroll
            for (int e = 0; e < 4; ++e) yr[i][e] = static_cast<float>(static_cast<double>(yr[i][e]) * wr);
          }
        }
      } else {
#pragma unroll
        for (int e = 0; e < 4; ++e) xr[i][e] = yr[i][e] = 0.0f;
      }
    }
  };
  typedef uint32_t u2v_t __attribute__((ext_vector_type(2)));
  auto put = [&](const float (&v)[4][4], __bf16 (*L)[GT][BKP]) {
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      uint32_t lo[NS], hi[NS];
      split_pair<NS>(v[0][c], v[1][c], lo);
      split_pair<NS>(v[2][c], v[3][c], hi);
#pragma unroll
      for (int s = 0; s < NS; ++s) *reinterpret_cast<u2v_t*>(&L[s][cb + c][rb]) = u2v_t{lo[s], hi[s]};
    }
  };
  fetch(r0);
  for (int64_t r = r0; r < r1; r += BK) {
    __syncthreads();
    put(xr, Xl);
    if (!same) put(yr, Yl);
    __syncthreads();
    if (r + BK < r1) fetch(r + BK);
    const int kb = 8 * (lane >> 4);
    bf8_t bfr[4][NS];
#pragma unroll
    for (int j = 0; j < 4; ++j)
#pragma unroll
      for (int s = 0; s < NS; ++s)
        bfr[j][s] = *reinterpret_cast<const bf8_t*>(same ? &Xl[s][wn * 64 + j * 16 + (lane & 15)][kb]
                                                         : &Yl[s][wn * 64 + j * 16 + (lane & 15)][kb]);
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      bf8_t ai[NS];
#pragma unroll
      for (int s = 0; s < NS; ++s) ai[s] = *reinterpret_cast<const bf8_t*>(&Xl[s][wm * 128 + i * 16 + (lane & 15)][kb]);
#pragma unroll
      for (int j = 0; j < 4; ++j)
        if (i < vi && j < vj) acc[i][j] = mfma_split<NS>(ai, bfr[j], acc[i][j]);
    }
  }
  double* P = part + u * (GT * GT);
#pragma unroll
  for (int i = 0; i < 8; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j)
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const int row = wm * 128 + i * 16 + 4 * (lane >> 4) + e;
        const int col = wn * 64 + j * 16 + (lane & 15);
        P[col * GT + row] = static_cast<double>(acc[i][j][e]);
      }
}

// Compensated Gram, fast path: the tiles' 256 columns all inside m and p,
// 16-byte aligned blocks, no weights. Same tiles, stages, parts and sums as
// k_gram_bf16 (bitwise the same partials), but two stages of loads in flight
// instead of one, and no condition on any load of the stage loop: rows past
// the chunk load its last row and are zeroed by a select, a diagonal tile of
// a symmetric Gram loads its Y stage (the X rows again, cache hits) and only
// skips writing it. With one stage in flight and a full barrier (which also
// waits for every load) a stage took about 5.9 us at the C3 shape against
// 0.4 us of MFMA work: one memory round trip per stage.
__device__ __forceinline__ void lds_sync() {
  asm volatile("" ::: "memory");
  __builtin_amdgcn_s_waitcnt(0xc07f);  // gfx9 encoding: vmcnt 63, expcnt 7, lgkmcnt 0
  __builtin_amdgcn_s_barrier();
  asm volatile("" ::: "memory");
}

template <int NS>
__global__ __launch_bounds__(512, 1) void k_gram_bf16f(int64_t n, const float* __restrict__ X, int64_t ldx,
                                                    const float* __restrict__ Y, int64_t ldy, int tiles_b,
                                                    int ntiles, int sym, int64_t rows_per, int64_t per_xcd,
                                                    int64_t nunits, double* __restrict__ part) {
  __shared__ __attribute__((aligned(16))) __bf16 Xl[NS][GT][BKP];
  __shared__ __attribute__((aligned(16))) __bf16 Yl[NS][GT][BKP];
  const int64_t u = xcd_tile(blockIdx.x, per_xcd);
  if (u >= nunits) return;
  const int64_t chunk = u / ntiles;
  const int tile = static_cast<int>(u - chunk * ntiles);
  int ta, tb;
  gram_tile(tile, tiles_b, sym, ta, tb);
  const bool same = sym == 1 && ta == tb;
  const int a0 = ta * GT, b0 = tb * GT;
  const int64_t r0 = chunk * rows_per;
  const int64_t r1 = r0 + rows_per < n ? r0 + rows_per : n;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave >> 2, wn = wave & 3;
  f4_t acc[8][4];
#pragma unroll
  for (int i = 0; i < 8; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[i][j] = f4_t{};
  const int rb = 4 * (lane & 7), cb = 4 * ((lane >> 3) + 8 * wave);
  const float* Xc = X + a0 + cb;
  const float* Yc = Y + b0 + cb;
  struct Stage {
    float4 x[4], y[4];
  };
  auto fetch = [&](int64_t r, Stage& st) {
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int64_t row = r + rb + i;
      const int64_t rc = row < r1 ? row : r1 - 1;
      st.x[i] = *reinterpret_cast<const float4*>(Xc + rc * ldx);
      st.y[i] = *reinterpret_cast<const float4*>(Yc + rc * ldy);
    }
  };
  typedef uint32_t u2v_t __attribute__((ext_vector_type(2)));
  auto put = [&](int64_t r, const float4 (&v)[4], __bf16 (*L)[GT][BKP]) {
    float e[4][4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const bool ok = r + rb + i < r1;  // rows past the chunk: zeros (per component: no select of objects)
      e[i][0] = ok ? v[i].x : 0.0f;
      e[i][1] = ok ? v[i].y : 0.0f;
      e[i][2] = ok ? v[i].z : 0.0f;
      e[i][3] = ok ? v[i].w : 0.0f;
    }
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      uint32_t lo[NS], hi[NS];
      split_pair<NS>(e[0][c], e[1][c], lo);
      split_pair<NS>(e[2][c], e[3][c], hi);
#pragma unroll
      for (int s = 0; s < NS; ++s) *reinterpret_cast<u2v_t*>(&L[s][cb + c][rb]) = u2v_t{lo[s], hi[s]};
    }
  };
  auto products = [&]() {
    const int kb = 8 * (lane >> 4);
    bf8_t bfr[4][NS];
#pragma unroll
    for (int j = 0; j < 4; ++j)
#pragma unroll
      for (int s = 0; s < NS; ++s)
        bfr[j][s] = *reinterpret_cast<const bf8_t*>(same ? &Xl[s][wn * 64 + j * 16 + (lane & 15)][kb]
                                                         : &Yl[s][wn * 64 + j * 16 + (lane & 15)][kb]);
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      bf8_t ai[NS];
#pragma unroll
      for (int s = 0; s < NS; ++s) ai[s] = *reinterpret_cast<const bf8_t*>(&Xl[s][wm * 128 + i * 16 + (lane & 15)][kb]);
#pragma unroll
      for (int j = 0; j < 4; ++j) acc[i][j] = mfma_split<NS>(ai, bfr[j], acc[i][j]);
    }
  };
  Stage st[2];
  fetch(r0, st[0]);
  fetch(r0 + BK, st[1]);
  auto stage = [&](auto Pc, int64_t r) {
    constexpr int P = decltype(Pc)::value;
    lds_sync();  // every wave is done reading the previous stage
    put(r, st[P].x, Xl);
    if (!same) put(r, st[P].y, Yl);
    lds_sync();
    fetch(r + 2 * BK, st[P]);  // two stages ahead (past the chunk: its last row again)
    products();
  };
  // stages in pairs: a chunk with an odd count runs one stage of zeros
  for (int64_t r = r0; r < r1; r += 2 * BK) {
    stage(std::integral_constant<int, 0>{}, r);
    stage(std::integral_constant<int, 1>{}, r + BK);
  }
  double* Pp = part + u * (GT * GT);
#pragma unroll
  for (int i = 0; i < 8; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j)
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const int row = wm * 128 + i * 16 + 4 * (lane >> 4) + e;
        const int col = wn * 64 + j * 16 + (lane & 15);
        Pp[col * GT + row] = static_cast<double>(acc[i][j][e]);
      }
}

// C (FP64 column-major m x p, ldc) -> NS bfloat16 parts of float(C), part s at
// S + s p mpad, column j at + j mpad, k contiguous, zero for k >= m (and, with
// upper, below the diagonal k > j).
template <int NS>
__global__ void k_split_coeffs(int m, int p, const double* __restrict__ C, int64_t ldc, int upper, int mpad,
                               __bf16* __restrict__ S) {
  const int64_t total = (int64_t)p * (mpad / 2);
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total; t += (int64_t)gridDim.x * blockDim.x) {
    const int j = static_cast<int>(t / (mpad / 2)), k = 2 * static_cast<int>(t % (mpad / 2));
    const bool ok0 = k < m && (!upper || k <= j), ok1 = k + 1 < m && (!upper || k + 1 <= j);
    const float c0 = ok0 ? static_cast<float>(C[(int64_t)j * ldc + k]) : 0.0f;
    const float c1 = ok1 ? static_cast<float>(C[(int64_t)j * ldc + k + 1]) : 0.0f;
    uint32_t o[NS];
    split_pair<NS>(c0, c1, o);
#pragma unroll
    for (int s = 0; s < NS; ++s)
      *reinterpret_cast<uint32_t*>(S + (int64_t)s * p * mpad + (int64_t)j * mpad + k) = o[s];
  }
}

// Y = alpha X C + beta Y on 128 x 128 output tiles from the split parts of X
// (staged: thread t loads 16 k values of row t >> 1 with four 16-byte loads)
// and of C (pre-split by k_split_coeffs: two 16-byte loads per part). Tile
// mapping, triangular pairing and epilogue as k_gemm_nn.
template <int NS, bool RES>
__global__ __launch_bounds__(256, 2) void k_gemm_bf16(int64_t n, int m, int p, const float* __restrict__ X,
                                                   int64_t ldx, const __bf16* __restrict__ S, int mpad, float* Y,
                                                   int64_t ldy, float alpha, float beta, int tiles_c,
                                                   int64_t per_xcd, int64_t ntiles, int vec, int upper,
                                                   ResArgs<float> res) {
  __shared__ __attribute__((aligned(16))) __bf16 Xl[NS][WT][BKP];
  __shared__ __attribute__((aligned(16))) __bf16 Cl[NS][WT][BKP];
  static_assert(sizeof(Xl) >= sizeof(double) * 4 * WT, "the residual sums reuse the X stage");
  const int64_t t = xcd_tile(blockIdx.x, per_xcd);
  if (t >= ntiles) return;
  const int per_slab = upper ? (tiles_c + 1) / 2 : tiles_c;
  const int64_t rt = t / per_slab;
  const int q = static_cast<int>(t - rt * per_slab);
  const int npass = upper && tiles_c - 1 - q != q ? 2 : 1;
  const int64_t r0 = rt * WT;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave >> 1, wn = wave & 1;
  const int64_t rows_left = n - r0 - wm * 64;
  const int vi = rows_left >= 64 ? 4 : static_cast<int>(max<int64_t>(0, (rows_left + 15) >> 4));
  const int sr = tid >> 1, sk = 16 * (tid & 1);  // staging: row (X) / column (C) sr, k sk .. sk + 15
  const int64_t xrow = r0 + sr;
  const size_t pstride = (size_t)p * mpad;
  for (int pass = 0; pass < npass; ++pass) {
  const int ct = pass == 0 ? q : tiles_c - 1 - q;
  const int c0 = ct * WT;
  const int vj = min(4, max(0, (p - c0 - wn * 64 + 15) >> 4));
  const int ccol = c0 + sr;
  const int kend = upper ? min(m, c0 + WT) : m;
  f4_t acc[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[i][j] = f4_t{};
  // rows past n and columns past p are read from the last row / column (their
  // outputs are not stored); k past m reads zeros
  const float* xp = X + min<int64_t>(xrow, n - 1) * ldx;
  const __bf16* cp = S + (size_t)min(ccol, p - 1) * mpad;
  float xr[16];
  u4v_t cr[NS][2];
  auto fetch = [&](int k0) {
    const int kk = k0 + sk;
#pragma unroll
    for (int h = 0; h < 4; ++h) {
      const int kh = kk + 4 * h;
      if (vec && kh + 4 <= m) {
        const float4 v = *reinterpret_cast<const float4*>(xp + kh);
        xr[4 * h] = v.x, xr[4 * h + 1] = v.y, xr[4 * h + 2] = v.z, xr[4 * h + 3] = v.w;
      } else {
#pragma unroll
        for (int e = 0; e < 4; ++e) xr[4 * h + e] = kh + e < m ? xp[kh + e] : 0.0f;
      }
    }
#pragma unroll
    for (int s = 0; s < NS; ++s) {
      const u4v_t* src = reinterpret_cast<const u4v_t*>(cp + s * pstride + kk);
      cr[s][0] = src[0];
      cr[s][1] = src[1];
    }
  };
  fetch(0);
  for (int k0 = 0; k0 < kend; k0 += BK) {
    __syncthreads();
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      uint32_t o[4][NS];
#pragma unroll
      for (int e = 0; e < 4; ++e) split_pair<NS>(xr[8 * h + 2 * e], xr[8 * h + 2 * e + 1], o[e]);
#pragma unroll
      for (int s = 0; s < NS; ++s)
        *reinterpret_cast<u4v_t*>(&Xl[s][sr][sk + 8 * h]) = u4v_t{o[0][s], o[1][s], o[2][s], o[3][s]};
    }
#pragma unroll
    for (int s = 0; s < NS; ++s) {
      *reinterpret_cast<u4v_t*>(&Cl[s][sr][sk]) = cr[s][0];
      *reinterpret_cast<u4v_t*>(&Cl[s][sr][sk + 8]) = cr[s][1];
    }
    __syncthreads();
    if (k0 + BK < kend) fetch(k0 + BK);
    const int kb = 8 * (lane >> 4);
    bf8_t af[4][NS];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int s = 0; s < NS; ++s) af[i][s] = *reinterpret_cast<const bf8_t*>(&Xl[s][wm * 64 + i * 16 + (lane & 15)][kb]);
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      bf8_t bj[NS];  // one column block at a time (all four spilled at NS = 3)
#pragma unroll
      for (int s = 0; s < NS; ++s) bj[s] = *reinterpret_cast<const bf8_t*>(&Cl[s][wn * 64 + j * 16 + (lane & 15)][kb]);
#pragma unroll
      for (int i = 0; i < 4; ++i)
        if (i < vi && j < vj) acc[i][j] = mfma_split<NS>(af[i], bj, acc[i][j]);
    }
  }
  if constexpr (RES) {
    resid_epilogue<float, mfma<float>, 2, WT>(acc, n, p, r0 + wm * 64, c0 + wn * 64, alpha, Y, ldy, res,
                                               reinterpret_cast<double*>(&Xl[0][0][0]), wm, wn, rt, c0);
  } else {
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j)
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const int64_t row = r0 + wm * 64 + i * 16 + 4 * (lane >> 4) + e;
        const int col = c0 + wn * 64 + j * 16 + (lane & 15);
        if (row < n && col < p) {
          float& y = Y[row * ldy + col];
          const float v = alpha * acc[i][j][e];
          y = beta == 0.0f ? v : v + beta * y;
        }
      }
  }
  }  // pass (the next pass's first barrier orders its LDS writes after these reads)
}

// Fixed-order sum over chunks; mirrors the lower triangle for symmetric Grams.
// Grid: (tile slots, tsz * tsz / 64 element slices), 256 threads: 64 elements
// x 4 chunk residues mod 4, each summed in chunk order and then the four
// residues in order.
__global__ __launch_bounds__(256) void k_gram_reduce(int m, int p, int tiles_b, int ntiles, int chunks,
                                                     const double* __restrict__ part, double* G, int64_t ldg, int sym,
                                                     int tsz) {
  __shared__ double red[4][64];
  const int t = blockIdx.x;  // computed tile slot
  int ta, tb;
  gram_tile(t, tiles_b, sym, ta, tb);
  const int el = threadIdx.x & 63, q = threadIdx.x >> 6;
  const int e = blockIdx.y * 64 + el;
  const int col = e / tsz, row = e % tsz;
  const int a = ta * tsz + row, b = tb * tsz + col;
  const bool ok = e < tsz * tsz && a < m && b < p;
  double s = 0.0;
  if (ok)
    for (int c = q; c < chunks; c += 4) s += part[((int64_t)c * ntiles + t) * (tsz * tsz) + e];
  red[q][el] = s;
  __syncthreads();
  if (q != 0 || !ok) return;
  s = red[0][el] + red[1][el] + red[2][el] + red[3][el];
  if (sym) {
    if (ta == tb) {
      if (a <= b) {
        G[(int64_t)b * ldg + a] = s;
        G[(int64_t)a * ldg + b] = s;
      }
    } else {
      G[(int64_t)b * ldg + a] = s;
      G[(int64_t)a * ldg + b] = s;
    }
  } else {
    G[(int64_t)b * ldg + a] = s;
  }
}

// For symmetric diagonal tiles the MFMA result is symmetric up to rounding of
// the same products in a different order; using the upper entry for both
// halves makes G exactly symmetric.

// ---------------------------------------------------------------- GEMM NN
// Y[r0:r0+128, c0:c0+128] = alpha X[r0:r0+128, :] C[:, c0:c0+128] + beta Y.
// Tiles map to XCDs in contiguous ranges with the column tiles of a row slab
// consecutive, so the slab is read from HBM once and from L2 for the other
// column tiles. X is staged [row][k] and C (column-major FP64, rounded to T)
// [k][col], MKS values of k per stage, the next stage prefetched into
// registers during the MFMA work.
constexpr int MKS = 16;

// upper != 0: C is upper triangular (m == p; nothing below its diagonal is
// read) and a column tile's inner loop stops at its last column. Column tile
// ct then costs ct + 1 units, so one workgroup computes the pair ct = q and
// ct = tiles_c - 1 - q of its row slab (tiles_c + 1 units for every q): with
// one tile per workgroup the CUs that kept receiving the last column tile set
// the finish time and the triangular product took as long as the full one.
template <typename T, int TR, int TC, bool RES>
__global__ __launch_bounds__(256, 2) void k_gemm_nn(int64_t n, int m, int p, const T* __restrict__ X, int64_t ldx,
                                                 const double* __restrict__ C, int64_t ldc, T* Y, int64_t ldy,
                                                 T alpha, T beta, int tiles_c, int64_t per_xcd, int64_t ntiles,
                                                 int vec, int cvec, int upper, ResArgs<T> res) {
  static_assert((TR / 64) * (TC / 64) == 4, "four 64 x 64 wavefront tiles");
  constexpr int MK = MKS;  // k values per stage
  constexpr int TPR = 256 / TR, XK = MK / TPR;  // X staging: threads per row, k values per thread
  constexpr int TPC = 256 / TC, CK = MK / TPC;  // C staging: threads per column, k values per thread
  __shared__ T Xs[TR][MK + 1];
  __shared__ T Cs[MK][TC + GPAD];
  __shared__ double red[RES ? 2 * (TR / 64) * TC : 1];  // residual column sums per row wavefront
  const int64_t t = xcd_tile(blockIdx.x, per_xcd);
  if (t >= ntiles) return;
  const int per_slab = upper ? (tiles_c + 1) / 2 : tiles_c;  // workgroups per row slab
  const int64_t rt = t / per_slab;
  const int q = static_cast<int>(t - rt * per_slab);
  const int npass = upper && tiles_c - 1 - q != q ? 2 : 1;
  const int64_t r0 = rt * TR;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave / (TC / 64), wn = wave % (TC / 64);
  // valid 16-wide row blocks of the wavefront's 64 x 64 tile (wave-uniform)
  const int64_t rows_left = n - r0 - wm * 64;
  const int vi = rows_left >= 64 ? 4 : static_cast<int>(max<int64_t>(0, (rows_left + 15) >> 4));
  // staging: X row tid / TPR, k (tid % TPR) * XK .. + XK - 1; C column tid / TPC, k (tid % TPC) * CK ..
  const int sr = tid / TPR, sk = (tid % TPR) * XK;
  const int scc = tid / TPC, sck = (tid % TPC) * CK;
  const int64_t xrow = r0 + sr;
  using acc_t = typename mfma<T>::acc;
  for (int pass = 0; pass < npass; ++pass) {
  const int ct = pass == 0 ? q : tiles_c - 1 - q;
  const int c0 = ct * TC;
  const int vj = min(4, max(0, (p - c0 - wn * 64 + 15) >> 4));
  acc_t acc[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[i][j] = acc_t{};
  const int ccol = c0 + scc;
  T xr[XK], cr[CK];
  auto fetch = [&](int k0) {
#pragma unroll
    for (int h = 0; h < XK / 8; ++h) {
      T* dst = xr + 8 * h;
      if (xrow < n) load8(X + xrow * ldx + k0 + sk + 8 * h, vec, m - (k0 + sk + 8 * h), *reinterpret_cast<T(*)[8]>(dst));
      else
#pragma unroll
        for (int e = 0; e < 8; ++e) dst[e] = T(0);
    }
    // with 16-byte aligned coefficient columns (cvec), whole runs are read
    // as 16-byte vectors: with 8-byte reads the FP32 rotation ran at 58% of
    // the MFMA peak and the FP64 one at 75%
    if (cvec && ccol < p && k0 + sck + CK <= m && (!upper || k0 + sck + CK - 1 <= ccol)) {
#pragma unroll
      for (int e = 0; e < CK; e += 2) {
        const double2 d = *reinterpret_cast<const double2*>(C + (int64_t)ccol * ldc + k0 + sck + e);
        cr[e] = static_cast<T>(d.x);
        cr[e + 1] = static_cast<T>(d.y);
      }
      return;
    }
#pragma unroll
    for (int e = 0; e < CK; ++e) {
      const int k = k0 + sck + e;
      cr[e] = (ccol < p && k < m && (!upper || k <= ccol)) ? static_cast<T>(C[(int64_t)ccol * ldc + k]) : T(0);
    }
  };
  const int kend = upper ? min(m, c0 + TC) : m;  // rows of C this column tile reads
  fetch(0);
  for (int k0 = 0; k0 < kend; k0 += MK) {
    __syncthreads();
#pragma unroll
    for (int e = 0; e < XK; ++e) Xs[sr][sk + e] = xr[e];
#pragma unroll
    for (int e = 0; e < CK; ++e) Cs[sck + e][scc] = cr[e];
    __syncthreads();
    if (k0 + MK < kend) fetch(k0 + MK);
#pragma unroll
    for (int kk = 0; kk < MK; kk += 4) {
      const int kr = kk + (lane >> 4);
      T af[4], bf[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) af[i] = Xs[wm * 64 + i * 16 + (lane & 15)][kr];
#pragma unroll
      for (int j = 0; j < 4; ++j) bf[j] = Cs[kr][wn * 64 + j * 16 + (lane & 15)];
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j)
          if (i < vi && j < vj) acc[i][j] = mfma<T>::run(af[i], bf[j], acc[i][j]);
    }
  }
  if constexpr (RES) {
    resid_epilogue<T, mfma<T>, TR / 64, TC>(acc, n, p, r0 + wm * 64, c0 + wn * 64, alpha, Y, ldy, res, red, wm, wn,
                                            rt, c0);
  } else {
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j)
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const int64_t row = r0 + wm * 64 + i * 16 + mfma<T>::row(lane, e);
        const int col = c0 + wn * 64 + j * 16 + (lane & 15);
        if (row < n && col < p) {
          T& y = Y[row * ldy + col];
          const T v = alpha * static_cast<T>(acc[i][j][e]);
          y = beta == T(0) ? v : v + beta * y;
        }
      }
  }
  }  // pass (the next pass's first barrier orders its LDS writes after these reads)
}

// ---------------------------------------------------------------- one-column products
// Gram against one column, g = X^T diag(w) y (X n x m, m <= GV_MAX): a
// thread sums its rows' products for all m columns in registers, the 256
// threads of a workgroup add theirs in a fixed tree order, one partial per
// workgroup chunk of rows; k_gram_vec_reduce adds the chunk partials in order.
// The MFMA Gram's 64 x 64 tiles would carry one useful column in 64.
constexpr int GV_MAX = 32;

template <typename T, int MM>
__global__ __launch_bounds__(256) void k_gram_vec(int64_t n, int m, const T* __restrict__ X, int64_t ldx,
                                                  const T* __restrict__ y, int64_t ldy, const double* __restrict__ w,
                                                  int64_t rows_per, double* __restrict__ part) {
  __shared__ double red[4][MM];
  const int64_t r0 = blockIdx.x * rows_per, r1 = min(n, r0 + rows_per);
  double acc[MM];
#pragma unroll
  for (int j = 0; j < MM; ++j) acc[j] = 0.0;
  for (int64_t i = r0 + threadIdx.x; i < r1; i += 256) {
    double yv = static_cast<double>(y[i * ldy]);
    if (w) yv *= w[i];
    const T* xr = X + i * ldx;
#pragma unroll
    for (int j = 0; j < MM; ++j)
      if (j < m) acc[j] += static_cast<double>(xr[j]) * yv;
  }
  // each column's sum over the wavefront by shuffles, then the four
  // wavefronts in order
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int j = 0; j < MM; ++j) {
    const double v = wave_sum(acc[j]);
    if (lane == 0) red[wave][j] = v;
  }
  __syncthreads();
  if (static_cast<int>(threadIdx.x) < m)
    part[(int64_t)blockIdx.x * m + threadIdx.x] =
        red[0][threadIdx.x] + red[1][threadIdx.x] + red[2][threadIdx.x] + red[3][threadIdx.x];
}

// G[j] = sum over the chunks of part[c * m + j], j < m <= GV_MAX: 8 chunk
// residues of 32 columns, each residue in chunk order, then the residues in
// order
__global__ __launch_bounds__(256) void k_gram_vec_reduce(int chunks, int m, const double* __restrict__ part,
                                                         double* G) {
  __shared__ double red[8][GV_MAX];
  const int j = threadIdx.x & (GV_MAX - 1), q = threadIdx.x / GV_MAX;
  double s = 0.0;
  if (j < m)
    for (int c = q; c < chunks; c += 8) s += part[(int64_t)c * m + j];
  red[q][j] = s;
  __syncthreads();
  if (q == 0 && j < m) {
    double t = 0.0;
#pragma unroll
    for (int r = 0; r < 8; ++r) t += red[r][j];
    G[j] = t;
  }
}

// y = alpha X c + beta y for one column c (m <= 32 FP64 values, rounded to T
// as in k_gemm_nn): a workgroup stages 128 rows through LDS with coalesced
// loads (consecutive threads read consecutive elements of the row-major
// block), then a thread sums its row
constexpr int GEMV_ROWS = 128, GEMV_MAX = 32;
template <typename T>
__global__ __launch_bounds__(GEMV_ROWS) void k_gemv(int64_t n, int m, const T* __restrict__ X, int64_t ldx,
                                                    const double* __restrict__ c, T* y, int64_t ldy, T alpha,
                                                    T beta) {
  __shared__ T xs[GEMV_ROWS * (GEMV_MAX + 1)];
  __shared__ T cs[GEMV_MAX];
  const int64_t r0 = blockIdx.x * (int64_t)GEMV_ROWS;
  const int rows = static_cast<int>(min<int64_t>(GEMV_ROWS, n - r0));
  if (static_cast<int>(threadIdx.x) < m) cs[threadIdx.x] = static_cast<T>(c[threadIdx.x]);
  for (int e = threadIdx.x; e < rows * m; e += GEMV_ROWS) {
    const int r = e / m, j = e - r * m;
    xs[r * (GEMV_MAX + 1) + j] = X[(r0 + r) * ldx + j];
  }
  __syncthreads();
  if (static_cast<int>(threadIdx.x) >= rows) return;
  T acc = T(0);
  for (int j = 0; j < m; ++j) acc += xs[threadIdx.x * (GEMV_MAX + 1) + j] * cs[j];
  T& out = y[(r0 + threadIdx.x) * ldy];
  out = beta == T(0) ? alpha * acc : alpha * acc + beta * out;
}

}  // namespace fe

using namespace fe;

namespace fe {

// Device addresses of the fixed scratch, one per device (filled once each).
static std::mutex g_static_mu;
static std::map<int, void*> g_static_ptr;

int gram_static_init(fe_ctx ctx) {
  std::lock_guard<std::mutex> lock(g_static_mu);
  auto it = g_static_ptr.find(ctx->device);
  if (it == g_static_ptr.end()) {
    void* ptr = nullptr;  // ctx->device is the current device (fe_ctx_create)
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(::scratch));
    if (e != hipSuccess || !ptr)
      return fail(-3, std::string("hipGetSymbolAddress(fe_gram_static scratch): ") + hipGetErrorString(e));
    it = g_static_ptr.emplace(ctx->device, ptr).first;
  }
  ctx->gram_static = it->second;
  return 0;
}

// The Gram launch of fe_gram and fe_gram_static. fixed == nullptr: the split-K
// partials go to the context's growable scratch. Otherwise they go to `fixed`
// (fixed_bytes long): the row chunks are cut so that their partials fit, and a
// product whose single chunk does not fit is rejected; nothing is allocated.
static int gram_run(fe_ctx ctx, int mode, int prec, int64_t n, int m, int p, const void* X, int64_t ldx,
                    const void* Y, int64_t ldy, const double* w, double* G, int64_t ldg, void* fixed,
                    size_t fixed_bytes, const char* who) {
  const std::string me(who);
  FE_ARG(ctx && X && Y && G, me + ": null argument");
  FE_ARG(m >= 0 && p >= 0 && n >= 0 && ldx >= m && ldy >= p && ldg >= m, me + ": bad dimensions");
  FE_ARG(prec == FE_F64 || prec == FE_F32, me + ": bad precision");
  {
    const int base = mode & ~FE_GRAM_SYMMETRIC;
    FE_ARG(base == FE_GRAM_NATIVE || ((base == FE_GRAM_BF16X3 || base == FE_GRAM_BF16X6) && prec == FE_F32),
           me + ": bad mode");
  }
  // FE_GRAM_SYMMETRIC: the caller asserts a symmetric product (m == p), so
  // only the tiles on and above the diagonal are computed and mirrored
  const bool sym_result = (mode & FE_GRAM_SYMMETRIC) != 0;
  mode &= ~FE_GRAM_SYMMETRIC;
  FE_ARG(!sym_result || m == p, me + ": a symmetric product needs m == p");
  if (m == 0 || p == 0) return 0;
  // symi 1: identical operands (the diagonal tiles stage one operand); 2: a
  // symmetric product of distinct operands
  const bool same = (X == Y) && (m == p) && (ldx == ldy) && !w;
  const bool sym = same || sym_result;
  const int symi = same ? 1 : sym ? 2 : 0;
  if (n == 0) return fe_zero2d(ctx, G, p, m, ldg, 8);
  const int cus = ctx->cus > 0 ? ctx->cus : 256;
  // partial tiles of one chunk, and the chunk count that fits the fixed buffer
  auto cap_chunks = [&](int64_t& chunks, int64_t& rows_per, int ntiles, int tsz, int stage) -> int {
    if (!fixed) return 0;
    const size_t per_chunk = sizeof(double) * (size_t)ntiles * tsz * tsz;
    const int64_t fit = static_cast<int64_t>(fixed_bytes / per_chunk);
    if (fit < 1)
      return fail(-2, me + ": the partial tiles of one row chunk (" + std::to_string(per_chunk) +
                          " bytes) exceed the fixed scratch of " + std::to_string(fixed_bytes) + " bytes");
    if (chunks > fit) {
      rows_per = (n + fit - 1) / fit;
      rows_per = (rows_per + stage - 1) / stage * stage;
      chunks = (n + rows_per - 1) / rows_per;
    }
    return 0;
  };
  if (mode == FE_GRAM_NATIVE && p == 1 && m <= GV_MAX && !fixed) {
    // one column (the reorthogonalization of a single vector against a
    // narrow basis): row chunks of about 2,048 rows, at most 1,024 of them
    const int64_t chunks = std::min<int64_t>(1024, std::max<int64_t>(1, (n + 2047) / 2048));
    const int64_t rows_per = (n + chunks - 1) / chunks;
    void* ws = nullptr;
    if (int rc = fe::scratch(ctx, 0, sizeof(double) * (size_t)chunks * m, &ws)) return rc;
    auto launch = [&](auto mm) {
      constexpr int MM = decltype(mm)::value;
      if (prec == FE_F64)
        hipLaunchKernelGGL((k_gram_vec<double, MM>), dim3((unsigned)chunks), dim3(256), 0, ctx->stream, n, m,
                           (const double*)X, ldx, (const double*)Y, ldy, w, rows_per, (double*)ws);
      else
        hipLaunchKernelGGL((k_gram_vec<float, MM>), dim3((unsigned)chunks), dim3(256), 0, ctx->stream, n, m,
                           (const float*)X, ldx, (const float*)Y, ldy, w, rows_per, (double*)ws);
    };
    if (m <= 8) launch(std::integral_constant<int, 8>{});
    else if (m <= 16) launch(std::integral_constant<int, 16>{});
    else launch(std::integral_constant<int, GV_MAX>{});
    FE_CHECK_LAUNCH();
    hipLaunchKernelGGL(k_gram_vec_reduce, dim3(1), dim3(256), 0, ctx->stream, (int)chunks, m, (const double*)ws, G);
    FE_CHECK_LAUNCH();
    return 0;
  }
  if (mode == FE_GRAM_NATIVE) {
    // 64 x 64 tiles (rows split over the waves) for thin or small products,
    // 128 x 128 tiles otherwise
    const bool small = std::min(m, p) <= 64 || std::max(m, p) <= 160;
    const int TT = small ? G64 : WT, RS = small ? G64R : 2 * GKS;
    const int tiles_a = (m + TT - 1) / TT, tiles_b = (p + TT - 1) / TT;
    const int ntiles = sym ? tiles_a * (tiles_a + 1) / 2 : tiles_a * tiles_b;
    // at most 4 work units per CU (two rounds of two resident workgroups:
    // rounding up would leave a third round of a few units); chunks of whole
    // stages
    int64_t chunks = std::max<int64_t>(1, 4 * cus / ntiles);
    int64_t rows_per = (n + chunks - 1) / chunks;
    rows_per = (rows_per + RS - 1) / RS * RS;
    chunks = (n + rows_per - 1) / rows_per;
    if (int rc = cap_chunks(chunks, rows_per, ntiles, TT, RS)) return rc;
    void* ws = fixed;
    if (!ws)
      if (int rc = fe::scratch(ctx, 0, sizeof(double) * (size_t)chunks * ntiles * TT * TT, &ws)) return rc;
    const int64_t units = chunks * ntiles, nb8 = (units + 7) / 8 * 8;
    const int V = prec == FE_F64 ? 2 : 4;
    const int vec = (reinterpret_cast<uintptr_t>(X) % 16 == 0) && (reinterpret_cast<uintptr_t>(Y) % 16 == 0) &&
                    ldx % V == 0 && ldy % V == 0;
    const dim3 grid((unsigned)nb8), block(256);
    if (prec == FE_F64) {
      if (small)
        hipLaunchKernelGGL(k_gram64<double>, grid, block, 0, ctx->stream, n, m, p, (const double*)X, ldx,
                           (const double*)Y, ldy, w, tiles_b, ntiles, symi, rows_per, nb8 / 8, units, vec,
                           (double*)ws);
      else
        hipLaunchKernelGGL(k_gram<double>, grid, block, 0, ctx->stream, n, m, p, (const double*)X, ldx,
                           (const double*)Y, ldy, w, tiles_b, ntiles, symi, rows_per, nb8 / 8, units, vec,
                           (double*)ws);
    } else {
      if (small)
        hipLaunchKernelGGL(k_gram64<float>, grid, block, 0, ctx->stream, n, m, p, (const float*)X, ldx,
                           (const float*)Y, ldy, w, tiles_b, ntiles, symi, rows_per, nb8 / 8, units, vec,
                           (double*)ws);
      else
        hipLaunchKernelGGL(k_gram<float>, grid, block, 0, ctx->stream, n, m, p, (const float*)X, ldx,
                           (const float*)Y, ldy, w, tiles_b, ntiles, symi, rows_per, nb8 / 8, units, vec,
                           (double*)ws);
    }
    FE_CHECK_LAUNCH();
    hipLaunchKernelGGL(k_gram_reduce, dim3((unsigned)ntiles, TT * TT / 64), dim3(256), 0, ctx->stream, m, p,
                       tiles_b, ntiles, (int)chunks, (const double*)ws, G, ldg, symi, TT);
    FE_CHECK_LAUNCH();
    return 0;
  }
  // compensated BF16 products: 256 x 256 tiles, one 512-thread workgroup per
  // CU (two rounds), chunks of whole 32-row stages
  const int tiles_a = (m + GT - 1) / GT, tiles_b = (p + GT - 1) / GT;
  const int ntiles = sym ? tiles_a * (tiles_a + 1) / 2 : tiles_a * tiles_b;
  int64_t chunks = std::max<int64_t>(1, 2 * cus / ntiles);
  int64_t rows_per = (n + chunks - 1) / chunks;
  rows_per = (rows_per + BK - 1) / BK * BK;
  chunks = (n + rows_per - 1) / rows_per;
  if (int rc = cap_chunks(chunks, rows_per, ntiles, GT, BK)) return rc;
  void* ws = fixed;
  if (!ws)
    if (int rc = fe::scratch(ctx, 0, sizeof(double) * (size_t)chunks * ntiles * GT * GT, &ws)) return rc;
  const int64_t units = chunks * ntiles, nb8 = (units + 7) / 8 * 8;
  const int vec = (reinterpret_cast<uintptr_t>(X) % 16 == 0) && (reinterpret_cast<uintptr_t>(Y) % 16 == 0) &&
                  ldx % 4 == 0 && ldy % 4 == 0;
  // fast path: whole tiles, aligned blocks, no weights (FE_GRAM_FAST=0 turns
  // it off); BF16x3 only: the BF16x6 parts with two stages in flight need more
  // than the 256 registers of two waves per SIMD and spill, also with the B
  // fragments taken in halves
  const char* fe_env = std::getenv("FE_GRAM_FAST");
  const bool fast = !(fe_env && std::atoi(fe_env) == 0) && vec && !w && m % GT == 0 && p % GT == 0 && n > 0;
  if (fast && mode == FE_GRAM_BF16X3)
    hipLaunchKernelGGL(k_gram_bf16f<2>, dim3((unsigned)nb8), dim3(512), 0, ctx->stream, n, (const float*)X, ldx,
                       (const float*)Y, ldy, tiles_b, ntiles, symi, rows_per, nb8 / 8, units, (double*)ws);
  else if (mode == FE_GRAM_BF16X3)
    hipLaunchKernelGGL(k_gram_bf16<2>, dim3((unsigned)nb8), dim3(512), 0, ctx->stream, n, m, p, (const float*)X, ldx,
                       (const float*)Y, ldy, w, tiles_b, ntiles, symi, rows_per, nb8 / 8, units, vec, (double*)ws);
  else
    hipLaunchKernelGGL(k_gram_bf16<3>, dim3((unsigned)nb8), dim3(512), 0, ctx->stream, n, m, p, (const float*)X, ldx,
                       (const float*)Y, ldy, w, tiles_b, ntiles, symi, rows_per, nb8 / 8, units, vec, (double*)ws);
  FE_CHECK_LAUNCH();
  hipLaunchKernelGGL(k_gram_reduce, dim3((unsigned)ntiles, GT * GT / 64), dim3(256), 0, ctx->stream, m, p, tiles_b,
                     ntiles, (int)chunks, (const double*)ws, G, ldg, symi, GT);
  FE_CHECK_LAUNCH();
  return 0;
}

}  // namespace fe

extern "C" int fe_gram(fe_ctx ctx, int mode, int prec, int64_t n, int m, int p, const void* X, int64_t ldx,
                       const void* Y, int64_t ldy, const double* w, double* G, int64_t ldg) {
  return fe::gram_run(ctx, mode, prec, n, m, p, X, ldx, Y, ldy, w, G, ldg, nullptr, 0, "fe_gram");
}

extern "C" int fe_gram_static(fe_ctx ctx, int mode, int prec, int64_t n, int m, int p, const void* X,
                              int64_t ldx, const void* Y, int64_t ldy, const double* w, double* G,
                              int64_t ldg) {
  FE_ARG(ctx, "fe_gram_static: null context");
  FE_ARG(ctx->gram_static, "fe_gram_static: the context has no fixed scratch address");
  return fe::gram_run(ctx, mode, prec, n, m, p, X, ldx, Y, ldy, w, G, ldg, ctx->gram_static,
                      FE_GRAM_STATIC_BYTES, "fe_gram_static");
}

namespace fe {

// Residual arguments of a fused rotation (fe_gemm_nn_resid), untyped.
struct ResHost {
  const void* Y;
  int64_t ldy;
  const double* theta;
  void* R;
  int64_t ldr;
  double* part_r;
  double* part_y;
  int store;
};

template <typename T>
static ResArgs<T> typed(const ResHost* r) {
  if (!r) return ResArgs<T>{nullptr, 0, nullptr, nullptr, 0, nullptr, nullptr, 1};
  return ResArgs<T>{static_cast<const T*>(r->Y), r->ldy, r->theta, static_cast<T*>(r->R), r->ldr, r->part_r,
                    r->part_y, r->store};
}

// rh != nullptr: the fused residual epilogue (beta 0, not triangular)
static int gemm_run(fe_ctx ctx, int mode, int prec, int64_t n, int m, int p, const void* X, int64_t ldx,
                    const double* C, int64_t ldc, void* Y, int64_t ldy, double alpha, double beta, int upper,
                    const ResHost* rh = nullptr) {
  if (mode == FE_GRAM_BF16X3 || mode == FE_GRAM_BF16X6) {
    // C split once into its bfloat16 parts (scratch slot 4), then the product
    const int NS = mode == FE_GRAM_BF16X3 ? 2 : 3;
    const int mpad = (m + BK - 1) / BK * BK;
    void* sp = nullptr;
    if (int rc = fe::scratch(ctx, 4, sizeof(__bf16) * (size_t)NS * p * mpad, &sp)) return rc;
    const unsigned sg = std::min<int64_t>(1024, ((int64_t)p * mpad / 2 + 255) / 256);
    if (NS == 2)
      hipLaunchKernelGGL(k_split_coeffs<2>, dim3(sg), dim3(256), 0, ctx->stream, m, p, C, ldc, upper, mpad, (__bf16*)sp);
    else
      hipLaunchKernelGGL(k_split_coeffs<3>, dim3(sg), dim3(256), 0, ctx->stream, m, p, C, ldc, upper, mpad, (__bf16*)sp);
    FE_CHECK_LAUNCH();
    const int tiles_c = (p + WT - 1) / WT;
    const int64_t tiles_r = (n + WT - 1) / WT;
    const int64_t ntiles = tiles_r * (upper ? (tiles_c + 1) / 2 : tiles_c);
    const int64_t nb8 = (ntiles + 7) / 8 * 8;
    const int vec = (reinterpret_cast<uintptr_t>(X) % 16 == 0) && ldx % 4 == 0;
    const ResArgs<float> ra = typed<float>(rh);
    auto kern = NS == 2 ? (rh ? k_gemm_bf16<2, true> : k_gemm_bf16<2, false>)
                        : (rh ? k_gemm_bf16<3, true> : k_gemm_bf16<3, false>);
    hipLaunchKernelGGL(kern, dim3((unsigned)nb8), dim3(256), 0, ctx->stream, n, m, p, (const float*)X, ldx,
                       (const __bf16*)sp, mpad, (float*)Y, ldy, (float)alpha, (float)beta, tiles_c, nb8 / 8, ntiles,
                       vec, upper, ra);
    FE_CHECK_LAUNCH();
    return 0;
  }
  if (p == 1 && m <= GEMV_MAX && !rh && !upper) {  // one column against a narrow block
    const unsigned g = static_cast<unsigned>((n + GEMV_ROWS - 1) / GEMV_ROWS);
    if (prec == FE_F64)
      hipLaunchKernelGGL(k_gemv<double>, dim3(g), dim3(GEMV_ROWS), 0, ctx->stream, n, m, (const double*)X, ldx, C,
                         (double*)Y, ldy, alpha, beta);
    else
      hipLaunchKernelGGL(k_gemv<float>, dim3(g), dim3(GEMV_ROWS), 0, ctx->stream, n, m, (const float*)X, ldx, C,
                         (float*)Y, ldy, (float)alpha, (float)beta);
    FE_CHECK_LAUNCH();
    return 0;
  }
  // 256 x 64 tiles when the output has at most 64 columns, else 128 x 128
  const bool thin = p <= 64;
  const int TR = thin ? 256 : WT, TC = thin ? 64 : WT;
  const int tiles_c = (p + TC - 1) / TC;
  const int64_t tiles_r = (n + TR - 1) / TR;
  // workgroups: one per output tile, or per pair of column tiles when C is triangular
  const int64_t ntiles = tiles_r * (upper ? (tiles_c + 1) / 2 : tiles_c);
  const int64_t nb8 = (ntiles + 7) / 8 * 8;
  const int V = prec == FE_F64 ? 2 : 4;
  const int vec = (reinterpret_cast<uintptr_t>(X) % 16 == 0) && ldx % V == 0;
  const dim3 grid((unsigned)nb8), block(256);
  // coefficient columns read as 16-byte vectors when C and ldc allow it
  const int cvec = (reinterpret_cast<uintptr_t>(C) % 16 == 0) && ldc % 2 == 0;
  if (prec == FE_F64) {
    const ResArgs<double> ra = typed<double>(rh);
    auto kern = thin ? (rh ? k_gemm_nn<double, 256, 64, true> : k_gemm_nn<double, 256, 64, false>)
                     : (rh ? k_gemm_nn<double, 128, 128, true> : k_gemm_nn<double, 128, 128, false>);
    hipLaunchKernelGGL(kern, grid, block, 0, ctx->stream, n, m, p, (const double*)X, ldx, C, ldc, (double*)Y, ldy,
                       alpha, beta, tiles_c, nb8 / 8, ntiles, vec, cvec, upper, ra);
  } else {
    const ResArgs<float> ra = typed<float>(rh);
    auto kern = thin ? (rh ? k_gemm_nn<float, 256, 64, true> : k_gemm_nn<float, 256, 64, false>)
                     : (rh ? k_gemm_nn<float, 128, 128, true> : k_gemm_nn<float, 128, 128, false>);
    hipLaunchKernelGGL(kern, grid, block, 0, ctx->stream, n, m, p, (const float*)X, ldx, C, ldc, (float*)Y, ldy,
                       (float)alpha, (float)beta, tiles_c, nb8 / 8, ntiles, vec, cvec, upper, ra);
  }
  FE_CHECK_LAUNCH();
  return 0;
}

// Row tiles of the product (the residual partials per column)
static int64_t gemm_row_tiles(int mode, int64_t n, int p) {
  const int TR = (mode == FE_GRAM_BF16X3 || mode == FE_GRAM_BF16X6) ? WT : (p <= 64 ? 256 : WT);
  return (n + TR - 1) / TR;
}

}  // namespace fe

extern "C" int fe_gemm_nn_mode(fe_ctx ctx, int mode, int prec, int64_t n, int m, int p, const void* X,
                               int64_t ldx, const double* C, int64_t ldc, void* Y, int64_t ldy, double alpha,
                               double beta) {
  FE_ARG(ctx && X && C && Y, "fe_gemm_nn: null argument");
  FE_ARG(X != Y, "fe_gemm_nn: Y must not alias X");
  FE_ARG(n >= 0 && m >= 0 && p >= 0 && ldx >= m && ldy >= p && ldc >= m, "fe_gemm_nn: bad dimensions");
  FE_ARG(mode == FE_GRAM_NATIVE || ((mode == FE_GRAM_BF16X3 || mode == FE_GRAM_BF16X6) && prec == FE_F32),
         "fe_gemm_nn: bad mode");
  if (n == 0 || p == 0) return 0;
  if (m == 0) {
    if (beta == 0.0) return fe_zero2d(ctx, Y, n, p, ldy, (int)prec_bytes(prec));
    return fe_axpby_cols(ctx, prec, n, p, Y, ldy, Y, ldy, 0.0, beta, nullptr);
  }
  return fe::gemm_run(ctx, mode, prec, n, m, p, X, ldx, C, ldc, Y, ldy, alpha, beta, 0);
}

extern "C" int fe_gemm_nn(fe_ctx ctx, int prec, int64_t n, int m, int p, const void* X, int64_t ldx,
                          const double* C, int64_t ldc, void* Y, int64_t ldy, double alpha, double beta) {
  return fe_gemm_nn_mode(ctx, FE_GRAM_NATIVE, prec, n, m, p, X, ldx, C, ldc, Y, ldy, alpha, beta);
}

extern "C" int fe_trmm_upper_mode(fe_ctx ctx, int mode, int prec, int64_t n, int m, const void* X, int64_t ldx,
                                  const double* C, int64_t ldc, void* Y, int64_t ldy) {
  FE_ARG(ctx && X && C && Y, "fe_trmm_upper: null argument");
  FE_ARG(X != Y, "fe_trmm_upper: Y must not alias X");
  FE_ARG(n >= 0 && m >= 0 && ldx >= m && ldy >= m && ldc >= m, "fe_trmm_upper: bad dimensions");
  FE_ARG(mode == FE_GRAM_NATIVE || ((mode == FE_GRAM_BF16X3 || mode == FE_GRAM_BF16X6) && prec == FE_F32),
         "fe_trmm_upper: bad mode");
  if (n == 0 || m == 0) return 0;
  return fe::gemm_run(ctx, mode, prec, n, m, m, X, ldx, C, ldc, Y, ldy, 1.0, 0.0, 1);
}

extern "C" int fe_trmm_upper(fe_ctx ctx, int prec, int64_t n, int m, const void* X, int64_t ldx, const double* C,
                             int64_t ldc, void* Y, int64_t ldy) {
  return fe_trmm_upper_mode(ctx, FE_GRAM_NATIVE, prec, n, m, X, ldx, C, ldc, Y, ldy);
}

extern "C" int fe_gemm_nn_resid(fe_ctx ctx, int mode, int prec, int64_t n, int m, int p, const void* X, int64_t ldx,
                                const double* C, int64_t ldc, void* AY, int64_t lday, const void* Y, int64_t ldy,
                                const double* theta, void* R, int64_t ldr, double* rnorm, double* ynorm) {
  FE_ARG(ctx && X && C && Y && theta && rnorm && ynorm, "fe_gemm_nn_resid: null argument");
  FE_ARG(X != AY && X != Y && X != R && (!AY || AY != Y) && (!R || (R != Y && R != AY)),
         "fe_gemm_nn_resid: outputs must not alias inputs");
  FE_ARG(n >= 0 && m > 0 && p >= 0 && ldx >= m && ldy >= p && ldc >= m && (!AY || lday >= p) && (!R || ldr >= p),
         "fe_gemm_nn_resid: bad dimensions");
  FE_ARG(prec == FE_F64 || prec == FE_F32, "fe_gemm_nn_resid: bad precision");
  FE_ARG(mode == FE_GRAM_NATIVE || ((mode == FE_GRAM_BF16X3 || mode == FE_GRAM_BF16X6) && prec == FE_F32),
         "fe_gemm_nn_resid: bad mode");
  if (p == 0) return 0;
  if (n == 0) {
    if (int rc = fe_zero2d(ctx, rnorm, 1, p, p, 8)) return rc;
    return fe_zero2d(ctx, ynorm, 1, p, p, 8);
  }
  const int64_t tiles_r = fe::gemm_row_tiles(mode, n, p);
  void* ws = nullptr;
  if (int rc = fe::scratch(ctx, 3, sizeof(double) * 2 * (size_t)tiles_r * p, &ws)) return rc;
  double* part_r = static_cast<double*>(ws);
  double* part_y = part_r + tiles_r * p;
  const fe::ResHost rh{Y, ldy, theta, R, ldr, part_r, part_y, AY ? 1 : 0};
  if (int rc = fe::gemm_run(ctx, mode, prec, n, m, p, X, ldx, C, ldc, AY, lday, 1.0, 0.0, 0, &rh)) return rc;
  if (int rc = fe::reduce_cols_sqrt(ctx, static_cast<int>(tiles_r), p, part_r, rnorm)) return rc;
  return fe::reduce_cols_sqrt(ctx, static_cast<int>(tiles_r), p, part_y, ynorm);
}
