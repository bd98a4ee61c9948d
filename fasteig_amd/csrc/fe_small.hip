// Small dense FP64 kernels for the Rayleigh-Ritz and Cholesky-QR steps.
// All matrices are column-major on the device.
//
//   k_syev_tri    one workgroup (m <= 16): Householder tridiagonalization
//                 (Golub & Van Loan, Alg. 8.3.1), Sturm-count bisection for
//                 every eigenvalue (GvL 8.4.1), inverse iteration for the
//                 eigenvectors of the tridiagonal (Wilkinson; shifts and
//                 clusters as in LAPACK dstein), clusters of close eigenvalues
//                 orthogonalized by a workgroup between the steps
//   for m > 16:
//   k_tri_step    the multi-workgroup tridiagonalization (one launch per column)
//   k_syev_msect  multisection of the tridiagonal's eigenvalues
//   k_syev_inviter  inverse iteration, one thread per eigenvalue
//   k_syev_clusters the clusters of close eigenvalues, a workgroup each
//   k_syev_back   multi-workgroup back-transformation V = H_0 ... H_{m-3} Z with
//                 the eigenvector slice of each workgroup held in LDS
#include "fe_common.hpp"

#include <cfloat>

namespace fe {

constexpr int SYEV_MAX = 2048;  // largest projected matrix handled on the device

__device__ __forceinline__ double block_sum(double v, double* red) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nw = (blockDim.x + 63) >> 6;
  v = wave_sum(v);
  __syncthreads();
  if (lane == 0) red[wave] = v;
  __syncthreads();
  double s = 0.0;
  for (int w = 0; w < nw; ++w) s += red[w];  // fixed order: every thread gets the same value
  return s;
}

__device__ __forceinline__ int sturm_count(int m, const double* d, const double* e2, double x, double pivmin) {
  int cnt = 0;
  double q = d[0] - x;
  if (fabs(q) < pivmin) q = -pivmin;
  if (q < 0.0) ++cnt;
  for (int i = 1; i < m; ++i) {
    q = d[i] - x - e2[i - 1] / q;
    if (fabs(q) < pivmin) q = -pivmin;
    if (q < 0.0) ++cnt;
  }
  return cnt;
}

__device__ __forceinline__ double hash_unit(uint32_t a, uint32_t b) {
  uint32_t h = a * 0x9E3779B1u ^ (b + 0x7F4A7C15u) * 0x85EBCA77u;
  h ^= h >> 15;
  h *= 0x2C1B3C6Du;
  h ^= h >> 12;
  h *= 0x297A2D39u;
  h ^= h >> 15;
  return (static_cast<double>(h) / 4294967295.0) * 2.0 - 1.0;
}

__device__ __forceinline__ double block_max(double v, double* red) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nw = (blockDim.x + 63) >> 6;
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
  __syncthreads();
  if (lane == 0) red[wave] = v;
  __syncthreads();
  double s = 0.0;
  for (int w = 0; w < nw; ++w) s = fmax(s, red[w]);
  return s;
}

// Solve (T - x_j I) y = in_scale * z in place with the LU factors of slot j
// (interleaved [i * m + j], u1 holding the reciprocals of the pivots, whose
// magnitudes are at least tiny); return max |y_i| and, in *ss, sum y_i^2. The
// recurrences carry their last values in registers, and the loads are
// software-pipelined: the factors and z of the next TB steps are in flight
// while the current TB steps are computed (a step waiting on its own loads
// made each of the 2 m steps a global memory round trip).
template <int TB>
__device__ __forceinline__ double tri_lu_solve_raw(int m, int j, const double* __restrict__ u1,
                                                   const double* __restrict__ u2, const double* __restrict__ u3,
                                                   const double* __restrict__ ml, const double* __restrict__ pv,
                                                   double tiny, double* __restrict__ Z, double in_scale,
                                                   double* ss = nullptr) {
  const int64_t M = m;
  // forward: row interchanges and L; step i reads z_{i+1}, pv_i, ml_i
  double fz[2][TB], fp[2][TB], fl[2][TB];
  auto fload = [&](int i0, double (&zb)[TB], double (&pb)[TB], double (&lb)[TB]) {
#pragma unroll
    for (int u = 0; u < TB; ++u) {
      const int i = i0 + u;
      if (i + 1 < m) {
        zb[u] = Z[(i + 1) * M + j];
        pb[u] = pv[i * M + j];
        lb[u] = ml[i * M + j];
      }
    }
  };
  double yc = Z[j] * in_scale;
  auto fsteps = [&](int i0, const double (&zb)[TB], const double (&pb)[TB], const double (&lb)[TB]) {
#pragma unroll
    for (int u = 0; u < TB; ++u) {
      const int i = i0 + u;
      if (i + 1 < m) {
        double yn = zb[u] * in_scale;
        if (pb[u] != 0.0) {
          const double tmp = yc;
          yc = yn;
          yn = tmp;
        }
        yn -= lb[u] * yc;
        Z[i * M + j] = yc;
        yc = yn;
      }
    }
  };
  fload(0, fz[0], fp[0], fl[0]);
  for (int i0 = 0; i0 + 1 < m; i0 += 2 * TB) {
    fload(i0 + TB, fz[1], fp[1], fl[1]);
    fsteps(i0, fz[0], fp[0], fl[0]);
    fload(i0 + 2 * TB, fz[0], fp[0], fl[0]);
    fsteps(i0 + TB, fz[1], fp[1], fl[1]);
  }
  // backward: U y = z; step i reads z_i, u1_i, u2_i, u3_i
  double bz[2][TB], b1[2][TB], b2[2][TB], b3[2][TB];
  auto bload = [&](int i0, double (&zb)[TB], double (&p1)[TB], double (&p2)[TB], double (&p3)[TB]) {
#pragma unroll
    for (int u = 0; u < TB; ++u) {
      const int i = i0 - u;
      if (i >= 0) {
        zb[u] = i == m - 1 ? 0.0 : Z[i * M + j];
        p1[u] = u1[i * M + j];
        p2[u] = i + 1 < m ? u2[i * M + j] : 0.0;
        p3[u] = i + 2 < m ? u3[i * M + j] : 0.0;
      }
    }
  };
  double mx = 0.0, sq = 0.0, z1 = 0.0, z2 = 0.0;  // y_{i+1}, y_{i+2}
  auto bsteps = [&](int i0, const double (&zb)[TB], const double (&p1)[TB], const double (&p2)[TB],
                    const double (&p3)[TB]) {
#pragma unroll
    for (int u = 0; u < TB; ++u) {
      const int i = i0 - u;
      if (i >= 0) {
        // y_{i+2} term first, off the chain through y_{i+1}; p1 holds 1 / u_ii
        const double s = fma(-p2[u], z1, (i == m - 1 ? yc : zb[u]) - p3[u] * z2);
        const double yi = s * p1[u];
        Z[i * M + j] = yi;
        mx = fmax(mx, fabs(yi));
        sq += yi * yi;
        z2 = z1;
        z1 = yi;
      }
    }
  };
  bload(m - 1, bz[0], b1[0], b2[0], b3[0]);
  for (int i0 = m - 1; i0 >= 0; i0 -= 2 * TB) {
    bload(i0 - TB, bz[1], b1[1], b2[1], b3[1]);
    bsteps(i0, bz[0], b1[0], b2[0], b3[0]);
    bload(i0 - 2 * TB, bz[0], b1[0], b2[0], b3[0]);
    bsteps(i0 - TB, bz[1], b1[1], b2[1], b3[1]);
  }
  if (ss) *ss = sq;
  return mx;
}

// The same solve with z scaled to max-norm one afterwards.
template <int TB>
__device__ __forceinline__ void tri_lu_solve(int m, int j, const double* u1, const double* u2, const double* u3,
                                             const double* ml, const double* pv, double tiny, double* Z) {
  const double mx = tri_lu_solve_raw<TB>(m, j, u1, u2, u3, ml, pv, tiny, Z, 1.0);
  if (mx > 0.0)
    for (int i = 0; i < m; ++i) Z[(int64_t)i * m + j] /= mx;
}

// Columns t..last of Z in order, each made orthonormal to the cluster's earlier
// columns by two classical Gram-Schmidt passes; the whole workgroup works on
// one column at a time (the dot products one wavefront each). The cluster is
// copied to C with its columns contiguous (C[q * m + i]) for the sweep, so that
// both the dot products and the updates read whole cache lines.
__device__ __forceinline__ void cluster_orth(int m, int t, int last, double* Z, double* C, double* zs, double* dots, double* red,
                             int* info) {
  const int tid = threadIdx.x, nt = blockDim.x, lane = tid & 63, wave = tid >> 6, nw = (nt + 63) >> 6;
  const int c = last - t + 1;
  for (int64_t idx = tid; idx < (int64_t)c * m; idx += nt) {
    const int i = static_cast<int>(idx / c), q = t + static_cast<int>(idx % c);
    C[(int64_t)q * m + i] = Z[(int64_t)i * m + q];
  }
  __syncthreads();
  for (int j = t; j <= last; ++j) {
    double* cj = C + (int64_t)j * m;
    for (int i = tid; i < m; i += nt) zs[i] = cj[i];
    __syncthreads();
    for (int pass = 0; pass < 2 && j > t; ++pass) {
      for (int q = t + wave; q < j; q += nw) {
        const double* cq = C + (int64_t)q * m;
        double sv = 0.0;
        for (int i = lane; i < m; i += 64) sv += cq[i] * zs[i];
        sv = wave_sum(sv);
        if (lane == 0) dots[q - t] = sv;
      }
      __syncthreads();
      for (int i = tid; i < m; i += nt) {
        double acc = zs[i];
        for (int q = t; q < j; ++q) acc -= dots[q - t] * C[(int64_t)q * m + i];
        zs[i] = acc;
      }
      __syncthreads();
    }
    double nr = 0.0;
    for (int i = tid; i < m; i += nt) nr += zs[i] * zs[i];
    nr = sqrt(block_sum(nr, red));
    if (!(nr > 0.0)) {
      if (tid == 0) info[0] = 2;
      nr = 1.0;
    }
    for (int i = tid; i < m; i += nt) cj[i] = zs[i] / nr;
    __syncthreads();
  }
  for (int64_t idx = tid; idx < (int64_t)c * m; idx += nt) {
    const int i = static_cast<int>(idx / c), q = t + static_cast<int>(idx % c);
    Z[(int64_t)i * m + q] = C[(int64_t)q * m + i];
  }
  __syncthreads();
}

// Eigenvectors of clustered eigenvalues (consecutive gaps <= ortol). Every
// member has had three inverse-iteration steps on its own (perturbed) shift,
// independently of the other members, with its LU factors kept in slot j of the
// work arrays. One workgroup per cluster (clusters dealt round-robin over the
// launch's workgroups) then orthogonalizes the members in order, runs one more
// inverse-iteration step on every member at once (one thread each), which
// removes what the orthogonalization left outside the cluster's invariant
// subspace, and orthogonalizes again. The members' order and shifts are those
// of LAPACK dstein; dstein reorthogonalizes after every step, one member at a
// time, which leaves the O(m) triangular solves on a single thread.
template <int TB>
__device__ __forceinline__ void syev_clusters(int m, const double* lam, double onenrm, double pivmin, double* work, double* zs,
                              double* dots, double* red, int* info) {
  const double ortol = 1.0e-5 * onenrm;
  const double tiny = fmax(DBL_EPSILON * onenrm, pivmin);
  const int64_t MM = (int64_t)m * m;
  const double* u1 = work;
  const double* u2 = work + MM;
  const double* u3 = work + 2 * MM;
  const double* ml = work + 3 * MM;
  const double* pv = work + 4 * MM;
  double* Z = work + 5 * MM;
  double* C = work + 6 * MM;
  const int tid = threadIdx.x, nt = blockDim.x;
  int ordinal = 0;
  for (int t = 0; t < m;) {
    int last = t;
    while (last + 1 < m && lam[last + 1] - lam[last] <= ortol) ++last;
    if (last > t && (ordinal++ % static_cast<int>(gridDim.x)) == static_cast<int>(blockIdx.x)) {
      cluster_orth(m, t, last, Z, C, zs, dots, red, info);
      for (int j = t + tid; j <= last; j += nt) tri_lu_solve<TB>(m, j, u1, u2, u3, ml, pv, tiny, Z);
      __syncthreads();
      cluster_orth(m, t, last, Z, C, zs, dots, red, info);
    }
    t = last + 1;
  }
}

// Gershgorin interval [gl, gu] of the tridiagonal (d, e) widened by a few
// units of roundoff, its 1-norm estimate and the pivot threshold of the
// Sturm counts (LAPACK dstebz). One thread.
__device__ void tri_bounds(int m, const double* d, const double* e, double* gl_out, double* gu_out,
                           double* onenrm_out, double* pivmin_out) {
  double gl = DBL_MAX, gu = -DBL_MAX, emax2 = 0.0;
  for (int i = 0; i < m; ++i) {
    const double el = i > 0 ? fabs(e[i - 1]) : 0.0;
    const double er = i + 1 < m ? fabs(e[i]) : 0.0;
    gl = fmin(gl, d[i] - el - er);
    gu = fmax(gu, d[i] + el + er);
    if (i + 1 < m) emax2 = fmax(emax2, e[i] * e[i]);
  }
  const double onenrm = fmax(fabs(gl), fabs(gu));
  const double on = onenrm > 0.0 ? onenrm : 1.0;
  const double pivmin = DBL_MIN * fmax(1.0, emax2);
  const double pad = 2.0 * DBL_EPSILON * on * m + 2.0 * pivmin;  // widen the interval slightly
  *onenrm_out = on;
  *pivmin_out = pivmin;
  *gl_out = gl - pad;
  *gu_out = gu + pad;
}

// Eigenvector t of the tridiagonal (d, e) by inverse iteration at the
// eigenvalue lam[t] (shifts of cluster members as LAPACK dstein), into
// column t of Z (interleaved [i * m + t]) with its LU factors in slot t.
// One thread; the caller's launch bounds must leave it the registers of the
// pipelined solves (at 128 VGPRs the loop ran out of scratch memory).
template <int TB>
__device__ __forceinline__ void inverse_iteration(int m, int t, const double* lam, const double* d, const double* ps, double onenrm,
                                  double pivmin, double* work, int* info) {
  const double ortol = 1.0e-5 * onenrm;
  const double pertol = 10.0 * DBL_EPSILON * onenrm;
  const double tiny = fmax(DBL_EPSILON * onenrm, pivmin);
  const int64_t MM = (int64_t)m * m;
  double* u1 = work;
  double* u2 = work + MM;
  double* u3 = work + 2 * MM;
  double* ml = work + 3 * MM;
  double* pv_ = work + 4 * MM;
  double* Z = work + 5 * MM;
      // shift: inside a cluster each member sits at least pertol above the
      // previous member's shift, so that no two factorizations coincide
      int s0 = t;
      while (s0 > 0 && lam[s0] - lam[s0 - 1] <= ortol) --s0;
      double xj = lam[s0];
      for (int q = s0 + 1; q <= t; ++q) xj = fmax(lam[q], xj + pertol);
      // LU with partial pivoting of T - xj I into slot t (interleaved), as
      // LAPACK dlagtf, with pivots of magnitude below tiny replaced by +-tiny as
      // its solver dlagts (job -1) does; u1 keeps the reciprocals of the pivots.
      // The pivot row's diagonal and superdiagonal entries are carried in
      // registers and the step has no branches (the lanes of a wavefront take
      // different pivots) and one reciprocal.
      double u1c = d[0] - xj, u2c = m > 1 ? ps[0] : 0.0;
      for (int i = 0; i + 1 < m; ++i) {
        const double bi = ps[i];
        const double a1 = d[i + 1] - xj;
        const double c1 = i + 2 < m ? ps[i + 1] : 0.0;
        const bool keep = fabs(u1c) >= fabs(bi);  // no row interchange
        double piv = keep ? u1c : bi;
        if (fabs(piv) < tiny) piv = copysign(tiny, piv);
        const double rp = rcp_nr(piv);
        const double l = (keep ? bi : u1c) * rp;
        u1[(int64_t)i * m + t] = rp;
        u2[(int64_t)i * m + t] = keep ? u2c : a1;
        u3[(int64_t)i * m + t] = keep ? 0.0 : c1;
        ml[(int64_t)i * m + t] = l;
        pv_[(int64_t)i * m + t] = keep ? 0.0 : 1.0;
        const double u1n = keep ? a1 - l * u2c : u2c - l * a1;
        u2c = keep ? c1 : -l * c1;
        u1c = u1n;
      }
      u1[(int64_t)(m - 1) * m + t] = rcp_nr(fabs(u1c) < tiny ? copysign(tiny, u1c) : u1c);
      // seeded start vector, three inverse-iteration steps; each step's
      // max-norm scaling is applied as the next one reads z (the iteration is
      // invariant to scaling) and the 2-norm one at the end
      for (int i = 0; i < m; ++i) Z[(int64_t)i * m + t] = hash_unit(t + 1, i + 1);
      double scale = 1.0, ss = 0.0;
      for (int iter = 0; iter < 3; ++iter) {
        double mx = tri_lu_solve_raw<TB>(m, t, u1, u2, u3, ml, pv_, tiny, Z, scale, &ss);
        if (!(mx > 0.0)) {
          info[0] = 2;
          mx = 1.0;
        }
        scale = 1.0 / mx;
      }
      // the last solve's sum of squares gives the 2-norm (its values are at
      // most 1 / tiny times the max-normalized input, far from overflow)
      double nrm = sqrt(ss);
      if (!(nrm > 0.0) || !isfinite(nrm)) {
        info[0] = 2;
        nrm = 1.0;
      }
      const double f = 1.0 / nrm;
  #pragma unroll 8
      for (int i = 0; i < m; ++i) Z[(int64_t)i * m + t] *= f;
}

// work layout (doubles): u1, u2, u3, ml, piv, Z  -- each m*m, interleaved [i * m + t];
// then m*m for the cluster copies (column q at C[q * m])
__global__ __launch_bounds__(1024) void k_syev_tri(int m, double* A, int64_t lda, double* W, double* tau,
                                                   double* work, int* info) {
  // m <= 16: everything in one workgroup
  __shared__ double vs[SYEV_MAX], ps[SYEV_MAX], wsv[SYEV_MAX];
  __shared__ double d[SYEV_MAX], e2[SYEV_MAX], lam[SYEV_MAX];
  __shared__ double zs[SYEV_MAX], dots[SYEV_MAX];
  __shared__ double red[16];
  __shared__ double s_tau, s_scale, s_gl, s_gu, s_pivmin, s_onenrm;
  __shared__ int s_skip;
  const int tid = threadIdx.x, nt = blockDim.x;
  const int lane = tid & 63, wave = tid >> 6, nw = nt >> 6;
  if (tid == 0) info[0] = 0;
  {
    // symmetrize from the lower triangle
    for (int64_t t = tid; t < (int64_t)m * m; t += nt) {
      const int j = static_cast<int>(t / m), i = static_cast<int>(t % m);
      if (i < j) A[(int64_t)j * lda + i] = A[(int64_t)i * lda + j];
    }
    __syncthreads();

    // ---- 1. Householder tridiagonalization
    for (int k = 0; k + 2 < m; ++k) {
      const int len = m - k - 1;
      double* x = A + (int64_t)k * lda + (k + 1);
      double part = 0.0;
      for (int i = 1 + tid; i < len; i += nt) part += x[i] * x[i];
      const double xn2 = block_sum(part, red);
      if (tid == 0) {
        const double alpha = x[0];
        if (xn2 == 0.0) {
          s_skip = 1;
          tau[k] = 0.0;
          e2[k] = alpha;  // store e temporarily (squared later)
        } else {
          s_skip = 0;
          const double beta = -copysign(sqrt(alpha * alpha + xn2), alpha);
          s_tau = (beta - alpha) / beta;
          s_scale = 1.0 / (alpha - beta);
          tau[k] = s_tau;
          e2[k] = beta;
        }
      }
      __syncthreads();
      if (s_skip) continue;
      const double t_k = s_tau, sc = s_scale;
      for (int i = tid; i < len; i += nt) {
        const double v = i == 0 ? 1.0 : x[i] * sc;
        vs[i] = v;
      }
      __syncthreads();
      for (int i = tid; i < len; i += nt) x[i] = vs[i];  // keep v for the back-transformation
      // p = tau * A22 v : one wavefront per column of A22
      for (int j = wave; j < len; j += nw) {
        const double* col = A + (int64_t)(k + 1 + j) * lda + (k + 1);
        double s = 0.0;
        for (int i = lane; i < len; i += 64) s += col[i] * vs[i];
        s = wave_sum(s);
        if (lane == 0) ps[j] = t_k * s;
      }
      __syncthreads();
      double pv = 0.0;
      for (int i = tid; i < len; i += nt) pv += ps[i] * vs[i];
      const double K = 0.5 * t_k * block_sum(pv, red);
      for (int i = tid; i < len; i += nt) wsv[i] = ps[i] - K * vs[i];
      __syncthreads();
      for (int j = wave; j < len; j += nw) {
        double* col = A + (int64_t)(k + 1 + j) * lda + (k + 1);
        const double vj = vs[j], wj = wsv[j];
        for (int i = lane; i < len; i += 64) col[i] -= vs[i] * wj + wsv[i] * vj;
      }
      __syncthreads();
    }
    // diagonal and off-diagonal of T
    for (int i = tid; i < m; i += nt) d[i] = A[(int64_t)i * lda + i];
    if (tid == 0 && m >= 2) e2[m - 2] = A[(int64_t)(m - 2) * lda + (m - 1)];
    __syncthreads();
  }
  // e2 currently holds e; Gershgorin bounds and squares
  if (tid == 0) tri_bounds(m, d, e2, &s_gl, &s_gu, &s_onenrm, &s_pivmin);
  __syncthreads();
  // keep e (unsquared) in ps for the inverse iteration, squares in e2
  for (int i = tid; i + 1 < m; i += nt) {
    ps[i] = e2[i];
    e2[i] = e2[i] * e2[i];
  }
  __syncthreads();

  // ---- 2. bisection: thread t finds the t-th smallest eigenvalue
  const double pivmin = s_pivmin, onenrm = s_onenrm;
  for (int t = tid; t < m; t += nt) {
    double lo = s_gl, hi = s_gu;
    for (int it = 0; it < 200; ++it) {
      const double mid = 0.5 * (lo + hi);
      if (hi - lo <= 2.0 * DBL_EPSILON * fmax(fabs(lo), fabs(hi)) + pivmin || mid == lo || mid == hi) break;
      if (sturm_count(m, d, e2, mid, pivmin) > t) hi = mid;
      else lo = mid;
    }
    lam[t] = 0.5 * (lo + hi);
    W[t] = lam[t];
  }
  __syncthreads();
  // ---- 3. inverse iteration, then the clusters of close eigenvalues
  // (m <= 16: the solves are not pipelined, which would spill at 1024 threads)
  for (int t = tid; t < m; t += nt) inverse_iteration<1>(m, t, lam, d, ps, onenrm, pivmin, work, info);
  __syncthreads();
  syev_clusters<1>(m, lam, onenrm, pivmin, work, zs, dots, red, info);
}

// Phase 2 of the multi-workgroup path as a kernel of its own: one thread per
// eigenvalue, 64 per workgroup, so the solves keep their values in registers.
__global__ __launch_bounds__(64) void k_syev_inviter(int m, const double* __restrict__ dg,
                                                     const double* __restrict__ eg, const double* __restrict__ W,
                                                     double* work, int* info) {
  __shared__ double d[SYEV_MAX], e[SYEV_MAX], lam[SYEV_MAX];
  __shared__ double s_gl, s_gu, s_onenrm, s_pivmin;
  const int tid = threadIdx.x;
  for (int i = tid; i < m; i += blockDim.x) {
    d[i] = dg[i];
    e[i] = i + 1 < m ? eg[i] : 0.0;
    lam[i] = W[i];
  }
  __syncthreads();
  if (tid == 0) tri_bounds(m, d, e, &s_gl, &s_gu, &s_onenrm, &s_pivmin);
  __syncthreads();
  const int t = blockIdx.x * blockDim.x + tid;
  if (t < m) inverse_iteration<8>(m, t, lam, d, e, s_onenrm, s_pivmin, work, info);
}

// Phase 3: the clusters of close eigenvalues, one workgroup of 256 threads
// per cluster (dealt round-robin).
__global__ __launch_bounds__(256) void k_syev_clusters(int m, const double* __restrict__ dg,
                                                       const double* __restrict__ eg,
                                                       const double* __restrict__ W, double* work, int* info) {
  __shared__ double d[SYEV_MAX], e[SYEV_MAX], lam[SYEV_MAX], zs[SYEV_MAX], dots[SYEV_MAX];
  __shared__ double red[16];
  __shared__ double s_gl, s_gu, s_onenrm, s_pivmin;
  const int tid = threadIdx.x;
  for (int i = tid; i < m; i += blockDim.x) {
    d[i] = dg[i];
    e[i] = i + 1 < m ? eg[i] : 0.0;
    lam[i] = W[i];
  }
  __syncthreads();
  if (tid == 0) tri_bounds(m, d, e, &s_gl, &s_gu, &s_onenrm, &s_pivmin);
  __syncthreads();
  syev_clusters<8>(m, lam, s_onenrm, s_pivmin, work, zs, dots, red, info);
}

// ---- multi-workgroup Householder tridiagonalization (one launch per column)
// Launch k applies the deferred rank-2 update of step k-1 to the workgroup's
// columns of the trailing matrix, recomputes the updated column k (every
// workgroup redundantly: O(m) work), forms the reflector of column k and
// accumulates p_k = tau_k A22 v_k for its columns in the same pass over the
// matrix. Reflector k is kept in Vh[k * m + i], i > k (Vh[k * m + k + 1] = 1).
__device__ __forceinline__ double wg_sum256(double v, double* red) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  v = wave_sum(v);
  __syncthreads();
  if (lane == 0) red[wave] = v;
  __syncthreads();
  return red[0] + red[1] + red[2] + red[3];
}

// Eigenvalues of the tridiagonal (dg, eg: diagonal and off-diagonal) by
// multisection: one wavefront per eigenvalue t, whose 64 lanes take Sturm
// counts at 64 interior points of the bracket; the first point with more than
// t eigenvalues below it and the one before it form the next bracket, 65x
// narrower (9 rounds from the Gershgorin interval to the bisection's
// stopping width instead of about 55). Same bounds, pivmin and stopping rule
// as the bisection of k_syev_tri.
__global__ __launch_bounds__(256) void k_syev_msect(int m, const double* __restrict__ dg,
                                                    const double* __restrict__ eg, double* W, int* info) {
  __shared__ double d[SYEV_MAX], e2[SYEV_MAX];
  __shared__ double s_gl, s_gu, s_pivmin;
  const int tid = threadIdx.x, lane = tid & 63;
  if (tid == 0 && blockIdx.x == 0) info[0] = 0;
  for (int i = tid; i < m; i += blockDim.x) {
    d[i] = dg[i];
    if (i + 1 < m) e2[i] = eg[i];
  }
  __syncthreads();
  if (tid == 0) {
    double gl = DBL_MAX, gu = -DBL_MAX, emax2 = 0.0;
    for (int i = 0; i < m; ++i) {
      const double el = i > 0 ? fabs(e2[i - 1]) : 0.0;
      const double er = i + 1 < m ? fabs(e2[i]) : 0.0;
      gl = fmin(gl, d[i] - el - er);
      gu = fmax(gu, d[i] + el + er);
      if (i + 1 < m) emax2 = fmax(emax2, e2[i] * e2[i]);
    }
    const double onenrm = fmax(fabs(gl), fabs(gu));
    s_pivmin = DBL_MIN * fmax(1.0, emax2);
    const double pad = 2.0 * DBL_EPSILON * (onenrm > 0.0 ? onenrm : 1.0) * m + 2.0 * s_pivmin;
    s_gl = gl - pad;
    s_gu = gu + pad;
  }
  __syncthreads();
  for (int i = tid; i + 1 < m; i += blockDim.x) e2[i] = e2[i] * e2[i];
  __syncthreads();
  const int t = blockIdx.x * (blockDim.x >> 6) + (tid >> 6);
  if (t >= m) return;  // whole wavefront; no barrier follows
  const double pivmin = s_pivmin;
  double lo = s_gl, hi = s_gu;
  for (int it = 0; it < 64; ++it) {
    if (hi - lo <= 2.0 * DBL_EPSILON * fmax(fabs(lo), fabs(hi)) + pivmin) break;
    const double h = (hi - lo) / 65.0;
    const double x = lo + h * (lane + 1);
    const uint64_t above = __ballot(sturm_count(m, d, e2, x, pivmin) > t);
    const int l = above ? __ffsll(static_cast<unsigned long long>(above)) - 1 : 64;
    const double nlo = l == 0 ? lo : lo + h * l;
    const double nhi = l == 64 ? hi : lo + h * (l + 1);
    if (!(nlo > lo || nhi < hi)) break;  // no narrower bracket representable
    lo = nlo;
    hi = nhi;
  }
  if (lane == 0) W[t] = 0.5 * (lo + hi);
}

// All of a step's global reads are issued first, so that one memory round
// trip serves them: the reflector and p of step k-1 and column k (NQ values
// per thread, rows k + tid + 256 q) and the first column the wavefront
// updates (PE values per lane, rows k + 1 + lane + 64 q).
template <int NQ, int PE>
__global__ __launch_bounds__(256) void k_tri_step(int m, int k, double* A, int64_t lda, double* Vh, double* tau,
                                                  double* dd, double* ee, const double* __restrict__ pprev,
                                                  double* pcur) {
  __shared__ double vp[SYEV_MAX], wp[SYEV_MAX], col[SYEV_MAX], vk[SYEV_MAX];
  __shared__ double red[4];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const bool upd = k > 0;
  double vq[NQ], pq[NQ], aq[NQ];
#pragma unroll
  for (int q = 0; q < NQ; ++q) {
    const int i = k + tid + 256 * q;
    vq[q] = pq[q] = aq[q] = 0.0;
    if (i < m) {
      aq[q] = A[(int64_t)k * lda + i];
      if (upd) {
        vq[q] = (i == k) ? 1.0 : Vh[(int64_t)(k - 1) * m + i];
        pq[q] = pprev[i];
      }
    }
  }
  // wave g updates the columns j = g (mod nwaves) right of k, the same ones
  // in every launch; its first two are read with the step's other loads
  const int nwaves = gridDim.x * 4, g = blockIdx.x * 4 + wave;
  const int jfirst = g + nwaves * ((k + 1 - g + nwaves - 1) / nwaves > 0 ? (k + 1 - g + nwaves - 1) / nwaves : 0);
  const int jsecond = jfirst + nwaves;
  double cq[PE], cq2[PE];
#pragma unroll
  for (int q = 0; q < PE; ++q) {
    const int i = k + 1 + lane + 64 * q;
    cq[q] = (jfirst < m && i < m) ? A[(int64_t)jfirst * lda + i] : 0.0;
    cq2[q] = (jsecond < m && i < m) ? A[(int64_t)jsecond * lda + i] : 0.0;
  }
  const double tprev = upd ? tau[k - 1] : 0.0;
  if (upd) {
    // v_{k-1} lives on rows k.. ; w_{k-1} = p_{k-1} - (tau/2)(p^T v) v
    double part = 0.0;
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
      const int i = k + tid + 256 * q;
      if (i < m) {
        vp[i] = vq[q];
        part += pq[q] * vq[q];
      }
    }
    const double K = 0.5 * tprev * wg_sum256(part, red);
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
      const int i = k + tid + 256 * q;
      if (i < m) wp[i] = pq[q] - K * vq[q];
    }
    __syncthreads();
  }
  // updated column k (rows k..m-1)
#pragma unroll
  for (int q = 0; q < NQ; ++q) {
    const int i = k + tid + 256 * q;
    if (i < m) {
      double x = aq[q];
      if (upd) x -= vp[i] * wp[k] + wp[i] * vp[k];
      col[i] = x;
    }
  }
  __syncthreads();
  double part = 0.0;
  for (int i = k + 2 + tid; i < m; i += 256) part += col[i] * col[i];
  const double xn2 = wg_sum256(part, red);
  const double alpha = col[k + 1];
  double tk = 0.0, sc = 0.0, beta = alpha;
  if (xn2 != 0.0) {
    beta = -copysign(sqrt(alpha * alpha + xn2), alpha);
    tk = (beta - alpha) / beta;
    sc = 1.0 / (alpha - beta);
  }
  for (int i = k + 1 + tid; i < m; i += 256) vk[i] = (i == k + 1) ? 1.0 : (tk != 0.0 ? col[i] * sc : 0.0);
  __syncthreads();
  if (blockIdx.x == 0) {
    if (tid == 0) {
      tau[k] = tk;
      dd[k] = col[k];
      ee[k] = beta;
    }
    for (int i = k + 1 + tid; i < m; i += 256) Vh[(int64_t)k * m + i] = vk[i];
  }
  // columns j > k: apply the update of step k-1 (rows > k) and form p_k[j]
  for (int j = jfirst; j < m; j += nwaves) {
    double* cj = A + (int64_t)j * lda;
    double s = 0.0;
    if (j == jfirst || j == jsecond) {
#pragma unroll
      for (int q = 0; q < PE; ++q) {
        const int i = k + 1 + lane + 64 * q;
        if (i < m) {
          double a = j == jfirst ? cq[q] : cq2[q];
          if (upd) {
            a -= vp[i] * wp[j] + wp[i] * vp[j];
            cj[i] = a;
          }
          s += a * vk[i];
        }
      }
    } else {
      for (int i = k + 1 + lane; i < m; i += 64) {
        double a = cj[i];
        if (upd) {
          a -= vp[i] * wp[j] + wp[i] * vp[j];
          cj[i] = a;
        }
        s += a * vk[i];
      }
    }
    s = wave_sum(s);
    if (lane == 0) pcur[j] = tk * s;
  }
}

// Last 2 x 2 block: apply the final deferred update and read d, e.
__global__ void k_tri_finish(int m, double* A, int64_t lda, const double* __restrict__ Vh,
                             const double* __restrict__ tau, double* dd, double* ee, const double* __restrict__ plast) {
  if (threadIdx.x != 0) return;
  const int k = m - 2;  // rows/cols k, k+1 remain
  double a00 = A[(int64_t)k * lda + k], a10 = A[(int64_t)k * lda + k + 1], a11 = A[(int64_t)(k + 1) * lda + k + 1];
  if (m >= 3) {
    const double t = tau[k - 1];
    const double v0 = 1.0, v1 = Vh[(int64_t)(k - 1) * m + k + 1];
    const double K = 0.5 * t * (plast[k] * v0 + plast[k + 1] * v1);
    const double w0 = plast[k] - K * v0, w1 = plast[k + 1] - K * v1;
    a00 -= 2.0 * v0 * w0;
    a10 -= v1 * w0 + w1 * v0;
    a11 -= 2.0 * v1 * w1;
  }
  dd[k] = a00;
  dd[k + 1] = a11;
  ee[k] = a10;
}

// V(:, cols) = H_0 ... H_{m-3} Z(:, cols); Householder vectors in A's lower part.
// Each workgroup owns CB columns of Z, held in LDS as [row][CB]. The
// reflectors are staged into LDS R at a time (one memory round trip per R
// steps instead of one per step), so a step is LDS work and one barrier.
__global__ __launch_bounds__(256) void k_syev_back(int m, const double* __restrict__ A, int64_t lda,
                                                   const double* __restrict__ tau, const double* __restrict__ Z,
                                                   double* V, int64_t ldv, int CB, int R) {
  extern __shared__ double lds[];
  double* zs = lds;          // m * CB
  double* vs = lds + m * CB;  // R * m: reflector kh - r at vs[r * m + i], zero for rows i <= kh - r
  __shared__ double ts[32];
  const int tid = threadIdx.x;
  const int c0 = blockIdx.x * CB;
  const int ncol = min(CB, m - c0);
  for (int t = tid; t < m * CB; t += blockDim.x) {
    const int i = t / CB, c = t % CB;
    zs[t] = c < ncol ? Z[(int64_t)i * m + c0 + c] : 0.0;
  }
  const int TPC = 256 / CB;  // threads per column
  const int c = tid / TPC, sub = tid % TPC;
  const int RC = R > 0 ? R : 1;  // R == 0: no staging (the slice fills the LDS), v read from A
  for (int kh = m - 3; kh >= 0; kh -= RC) {
    const int nr = min(RC, kh + 1);  // reflectors kh, kh - 1, .., kh - nr + 1
    __syncthreads();                 // the previous chunk is no longer read
    for (int t = tid; R > 0 && t < nr * m; t += blockDim.x) {
      const int r = t / m, i = t % m, k = kh - r;
      vs[t] = i > k ? A[(int64_t)k * lda + i] : 0.0;
    }
    if (tid < nr) ts[tid] = tau[kh - tid];
    __syncthreads();
    for (int r = 0; r < nr; ++r) {
      const int k = kh - r;
      const double tk = ts[r];
      if (tk == 0.0) continue;  // the same in every thread
      const double* v = R > 0 ? vs + r * m : A + (int64_t)k * lda;
      double s = 0.0;
      for (int i = k + 1 + sub; i < m; i += TPC) s += v[i] * zs[i * CB + c];
      for (int o = TPC / 2; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
      s *= tk;
      for (int i = k + 1 + sub; i < m; i += TPC) zs[i * CB + c] -= s * v[i];
      __syncthreads();  // the next step's sums read rows other threads updated
    }
  }
  __syncthreads();
  for (int t = tid; t < m * CB; t += blockDim.x) {
    const int cc = t / m, i = t % m;
    if (cc < ncol) V[(int64_t)(c0 + cc) * ldv + i] = zs[i * CB + cc];
  }
}

}  // namespace fe

using namespace fe;

extern "C" int fe_syev(fe_ctx ctx, int m, double* A, int64_t lda, double* W, double* V, int64_t ldv) {
  FE_ARG(ctx && A && W && V && m >= 0 && lda >= m && ldv >= m, "fe_syev: bad argument");
  FE_ARG(m <= SYEV_MAX, "fe_syev: matrix larger than 2048 is not supported");
  if (m == 0) return 0;
  const int64_t MM = (int64_t)m * m;
  void* ws = nullptr;
  // tau | d | e | p (2 m) | Vh (m*m) | work (7 m*m)
  int rc = scratch(ctx, 1, sizeof(double) * (8 * MM + 5 * (int64_t)m + 64), &ws);
  if (rc) return rc;
  double* tau = (double*)ws;
  double* dd = tau + m;
  double* ee = dd + m;
  double* pb = ee + m;
  double* Vh = pb + 2 * m;
  double* work = Vh + MM;
  // multi-workgroup steps, multisection and per-eigenvalue inverse iteration
  // from m = 17 up (m = 96: 0.98 ms against 1.46 ms in one workgroup)
  const bool multi = m > 16;
  if (multi) {
    rc = fe_sym_small(ctx, m, A, lda);
    if (rc) return rc;
    const int G = 64;
    auto steps = [&]() {
      for (int k = 0; k + 2 < m; ++k) {
        const double* pprev = pb + ((k + 1) & 1) * m;
        double* pcur = pb + (k & 1) * m;
        // rows past k held in registers: 2 x 256 / 8 x 64 up to m = 512, else 8 x 256 / 32 x 64
        if (m <= 512)
          hipLaunchKernelGGL((k_tri_step<2, 8>), dim3(G), dim3(256), 0, ctx->stream, m, k, A, lda, Vh, tau, dd, ee,
                             pprev, pcur);
        else
          hipLaunchKernelGGL((k_tri_step<8, 32>), dim3(G), dim3(256), 0, ctx->stream, m, k, A, lda, Vh, tau, dd, ee,
                             pprev, pcur);
      }
      hipLaunchKernelGGL(k_tri_finish, dim3(1), dim3(64), 0, ctx->stream, m, A, lda, Vh, tau, dd, ee,
                         pb + ((m - 3) & 1) * m);
    };
    // The m - 2 dependent column steps as one captured graph, replayed when
    // the size and the addresses match (the solvers call this with the same
    // workspace every cycle): one submission instead of m - 2 launches. A
    // caller that is itself capturing the stream gets the launches recorded
    // into its own graph instead.
    if (m < 64) {
      steps();
      FE_CHECK_LAUNCH();
    } else {
      const int64_t key[8] = {1, m, (int64_t)(uintptr_t)A, lda, (int64_t)(uintptr_t)ws, 0, 0, 0};
      if (int rc2 = run_captured(ctx, key, steps)) return rc2;
    }
  }
  if (multi) {
    const unsigned g = blocks_for(m, 64);
    hipLaunchKernelGGL(k_syev_msect, dim3(blocks_for(m, 4)), dim3(256), 0, ctx->stream, m, dd, ee, W, ctx->info_dev);
    hipLaunchKernelGGL(k_syev_inviter, dim3(g), dim3(64), 0, ctx->stream, m, dd, ee, W, work, ctx->info_dev);
    hipLaunchKernelGGL(k_syev_clusters, dim3(16), dim3(256), 0, ctx->stream, m, dd, ee, W, work, ctx->info_dev);
  } else {
    hipLaunchKernelGGL(k_syev_tri, dim3(1), dim3(1024), 0, ctx->stream, m, A, lda, W, tau, work, ctx->info_dev);
  }
  FE_CHECK_LAUNCH();
  // columns per workgroup: about m / 128 (so that m = 512 spreads over 128
  // workgroups: the reflectors are applied one after another, so more and
  // narrower workgroups shorten each step), at least 4 (a column's 64 threads
  // reduce within one wavefront), LDS slice m*CB doubles within 64 KiB
  int CB = 32;
  while (CB > 4 && (CB * 128 > m || (size_t)m * CB * sizeof(double) > 65536)) CB >>= 1;
  // reflectors staged R at a time, the slice and R reflectors within 64 KiB
  // of LDS (at most 32; none when the slice alone fills it)
  const int R = std::max(0, std::min(32, (8192 - m * CB) / m));
  const size_t lds = sizeof(double) * ((size_t)m * CB + (size_t)R * m);
  hipLaunchKernelGGL(k_syev_back, dim3(blocks_for(m, CB)), dim3(256), lds, ctx->stream, m,
                     multi ? (const double*)Vh : (const double*)A, multi ? (int64_t)m : lda, tau, work + 5 * MM, V,
                     ldv, CB, R);
  FE_CHECK_LAUNCH();
  int info = 0;
  FE_HIP_TRY(hipMemcpyAsync(&info, ctx->info_dev, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
  FE_HIP_TRY(hipStreamSynchronize(ctx->stream));
  return info;
}
