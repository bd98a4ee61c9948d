// CPU restatement of the fasteig block-eigensolver path (FP64, C++17 + OpenMP).
//
// TEST INFRASTRUCTURE ONLY. This file is the checker for the HIP library in
// fasteig_amd/: only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline
// leg may load it. The product path never links or calls it.
//
// The reference (BerkeleyLab/fasteig) cannot be built here: it needs Eigen 3.4.0,
// which is not vendored and is fetched from the network by its CMake. So every
// Eigen routine the path relies on is restated below from the textbook method it
// implements (named at each function), and every solver follows the reference
// source line by line (cited as file:line relative to the reference root).
//
// Storage: dense blocks are column-major (like Eigen's MatrixXd), n x k with
// leading dimension n. Sparse matrices are full (both triangles) CSR with int32
// indices. Exported entry points have C linkage and plain pointers.
//
// Parity pinning (what checks each function; see tests/test_oracle.py):
//   orc_syev            numpy.linalg.eigh on random symmetric + Wilkinson matrices
//   orc_householder_qr  Q^T Q = I, Q R = X, R upper (invariants)
//   orc_cholqr          same invariants, -1 on a singular Gram
//   orc_jacobi_svd      numpy.linalg.svd singular values; U^T U = I
//   orc_spmm            dense brute force (numpy) on small matrices
//   orc_gram / orc_right_multiply
//                       numpy X^T Y and X C
//   orc_lanczos_bounds  bounds enclose the spectrum of small matrices (eigh)
//   orc_cheb_scalars    closed form T_d(c*lambda+e) vs numpy Chebyshev
//   orc_cheb_filter_scaled
//                       T_d((lambda-c)/e) / T_d((a0-c)/e) (numpy Chebyshev) on a
//                       diagonal matrix, V p(Lambda) V^T X (eigh) on a 27-point one
//   orc_chebyshev(ritz_bounds=1)
//                       2D Laplacian closed form (SR, LR), eigh on a small 27-point
//                       Hamiltonian (values and subspace)
//   orc_chebyshev / orc_lobpcg / orc_davidson
//                       2D Laplacian closed form, SiO spectrum from the
//                       reference's tests/reference/SiO.out.121mpi (golden
//                       fixture), dense eigh on small random problems.
//   orc_lanczos / orc_krylov_schur / orc_block_lanczos
//                       2D Laplacian closed form (SR, LR, BE ends), SiO golden
//                       spectrum, dense eigh on small random problems, a
//                       matrix with a repeated eigenvalue (invariant-subspace
//                       restart), Lanczos at m = n against eigh exactly.
//   orc_residual_norms  numpy on the reference's definition (unnormalized, zero
//                       and exact-eigenpair columns; negative and zero lambda)
//   orc_orthonormalize_block
//                       Q^T Q = I, B^T Q = 0 and the rank count on a block with
//                       dependent columns (invariants)
//   orc_block_krylov_schur
//                       2D Laplacian closed form (SR, LR), SiO golden spectrum,
//                       full-dimension exactness against eigh, a start block in
//                       an invariant subspace (seeded replenishment)
//   orc_random_block    reference's seeded mt19937/uniform(-1,1) fill: pinned by
//                       the C++ standard's required 10000th mt19937 output
//                       (4123659995) and by a golden vector.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <numeric>
#include <random>
#include <stdexcept>
#include <string>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace orc {

using Vec = std::vector<double>;

struct Mat {
  int r = 0, c = 0;
  Vec a;
  Mat() = default;
  Mat(int rows, int cols) : r(rows), c(cols), a(static_cast<size_t>(rows) * cols, 0.0) {}
  double& operator()(int i, int j) { return a[static_cast<size_t>(j) * r + i]; }
  double operator()(int i, int j) const { return a[static_cast<size_t>(j) * r + i]; }
  double* col(int j) { return a.data() + static_cast<size_t>(j) * r; }
  const double* col(int j) const { return a.data() + static_cast<size_t>(j) * r; }
};

struct Csr {
  int n = 0;
  const int* rowptr = nullptr;
  const int* colind = nullptr;
  const double* vals = nullptr;
};

// ---------------------------------------------------------------- dense kernels

// The tall-skinny products below are cache-blocked: rows are taken in chunks,
// packed into panels of 4 (left operand) and 8 (right operand) columns, and
// multiplied in 4 x 8 register blocks (AVX2 vectors of 4 doubles). Same sums
// as the textbook loops, in an order that keeps the operands in cache, so
// that the checker finishes full-size problems.
typedef double v4d __attribute__((vector_size(32)));

namespace {
inline v4d bcast(double x) { return v4d{x, x, x, x}; }
inline v4d loadu(const double* p) {
  v4d v;
  std::memcpy(&v, p, sizeof(v));
  return v;
}
}  // namespace

// C = A^T B over the first ka columns of A and kb of B (A,B n-row blocks).
// Rows are cut into a fixed number of slabs (independent of the thread count,
// so results do not depend on it); each slab's partial product is summed over
// its rows in order, and the slab partials are added in slab order.
Mat AtB(const Mat& A, int ka, const Mat& B, int kb) {
  Mat C(ka, kb);
  const long n = A.r;
  if (n == 0 || ka == 0 || kb == 0) return C;
  constexpr long RB = 256;  // rows per packed chunk
  const int pa = (ka + 3) / 4, pb = (kb + 7) / 8;  // panels
  const long slabs = std::max(1L, std::min(64L, n / 2048));
  const long per = (n + slabs - 1) / slabs;
  std::vector<double> part(static_cast<size_t>(slabs) * pa * 4 * pb * 8, 0.0);
#pragma omp parallel
  {
    std::vector<double> ap(static_cast<size_t>(pa) * RB * 4), bp(static_cast<size_t>(pb) * RB * 8);
#pragma omp for schedule(dynamic, 1)
    for (long sl = 0; sl < slabs; ++sl) {
      double* P = part.data() + static_cast<size_t>(sl) * pa * 4 * pb * 8;  // [pa*4][pb*8] row-major
      const long r0 = sl * per, r1 = std::min(n, r0 + per);
      for (long c0 = r0; c0 < r1; c0 += RB) {
        const long rb = std::min(RB, r1 - c0);
        // panels: ap[(panel * RB + t) * 4 + u] = A(c0 + t, 4 panel + u), zero padded
        for (int i = 0; i < pa * 4; ++i) {
          double* dst = &ap[(static_cast<size_t>(i / 4) * RB) * 4 + (i % 4)];
          if (i < ka) {
            const double* x = A.col(i) + c0;
            for (long t = 0; t < rb; ++t) dst[t * 4] = x[t];
          } else {
            for (long t = 0; t < rb; ++t) dst[t * 4] = 0.0;
          }
        }
        for (int j = 0; j < pb * 8; ++j) {
          double* dst = &bp[(static_cast<size_t>(j / 8) * RB) * 8 + (j % 8)];
          if (j < kb) {
            const double* y = B.col(j) + c0;
            for (long t = 0; t < rb; ++t) dst[t * 8] = y[t];
          } else {
            for (long t = 0; t < rb; ++t) dst[t * 8] = 0.0;
          }
        }
        for (int ip = 0; ip < pa; ++ip) {
          const double* ar = &ap[static_cast<size_t>(ip) * RB * 4];
          for (int jp = 0; jp < pb; ++jp) {
            const double* br = &bp[static_cast<size_t>(jp) * RB * 8];
            v4d acc[4][2] = {};
            for (long t = 0; t < rb; ++t) {
              const v4d b0 = loadu(br + t * 8), b1 = loadu(br + t * 8 + 4);
              for (int u = 0; u < 4; ++u) {
                const v4d a = bcast(ar[t * 4 + u]);
                acc[u][0] += a * b0;
                acc[u][1] += a * b1;
              }
            }
            for (int u = 0; u < 4; ++u) {
              double* prow = P + static_cast<size_t>(ip * 4 + u) * pb * 8 + jp * 8;
              for (int v = 0; v < 4; ++v) {
                prow[v] += acc[u][0][v];
                prow[4 + v] += acc[u][1][v];
              }
            }
          }
        }
      }
    }
  }
  const size_t ld = static_cast<size_t>(pb) * 8;
#pragma omp parallel for schedule(static)
  for (int i = 0; i < ka; ++i)
    for (int j = 0; j < kb; ++j) {
      double s = 0.0;
      for (long sl = 0; sl < slabs; ++sl) s += part[static_cast<size_t>(sl) * pa * 4 * ld + i * ld + j];
      C(i, j) = s;
    }
  return C;
}

// Y = A[:, :k] * C   (A n x >=k, C k x p); each entry sums over i in order.
Mat AC(const Mat& A, int k, const Mat& C) {
  Mat Y(A.r, C.c);
  const long n = A.r;
  const int p = C.c;
  if (n == 0 || p == 0) return Y;
  const int pp = (p + 7) / 8;
  std::vector<double> ct(static_cast<size_t>(pp) * std::max(k, 1) * 8, 0.0);  // [panel][i][8]
  for (int i = 0; i < k; ++i)
    for (int j = 0; j < p; ++j) ct[(static_cast<size_t>(j / 8) * k + i) * 8 + (j % 8)] = C(i, j);
  constexpr long RB = 64;  // rows per chunk (16 row quads)
  const long chunks = (n + RB - 1) / RB;
#pragma omp parallel
  {
    std::vector<double> ap(static_cast<size_t>(RB / 4) * std::max(k, 1) * 4);  // [quad][i][4]
#pragma omp for schedule(static)
    for (long ch = 0; ch < chunks; ++ch) {
      const long c0 = ch * RB, rb = std::min(RB, n - c0);
      const long quads = (rb + 3) / 4;
      for (int i = 0; i < k; ++i) {
        const double* x = A.col(i) + c0;
        for (long t = 0; t < quads * 4; ++t)
          ap[(static_cast<size_t>(t / 4) * k + i) * 4 + (t % 4)] = t < rb ? x[t] : 0.0;
      }
      for (long q = 0; q < quads; ++q) {
        const double* ar = &ap[static_cast<size_t>(q) * k * 4];
        for (int jp = 0; jp < pp; ++jp) {
          const double* cr = &ct[static_cast<size_t>(jp) * k * 8];
          v4d acc[4][2] = {};
          for (int i = 0; i < k; ++i) {
            const v4d c0v = loadu(cr + i * 8), c1v = loadu(cr + i * 8 + 4);
            for (int u = 0; u < 4; ++u) {
              const v4d a = bcast(ar[i * 4 + u]);
              acc[u][0] += a * c0v;
              acc[u][1] += a * c1v;
            }
          }
          for (int u = 0; u < 4; ++u) {
            const long t = q * 4 + u;
            if (t >= rb) break;
            for (int v = 0; v < 8; ++v) {
              const int j = jp * 8 + v;
              if (j < p) Y(static_cast<int>(c0 + t), j) = v < 4 ? acc[u][0][v] : acc[u][1][v - 4];
            }
          }
        }
      }
    }
  }
  return Y;
}

Mat small_mul(const Mat& A, const Mat& B) {  // small dense A*B
  Mat C(A.r, B.c);
  for (int j = 0; j < B.c; ++j)
    for (int k = 0; k < A.c; ++k) {
      const double b = B(k, j);
      for (int i = 0; i < A.r; ++i) C(i, j) += A(i, k) * b;
    }
  return C;
}

Mat transpose(const Mat& A) {
  Mat T(A.c, A.r);
  for (int j = 0; j < A.c; ++j)
    for (int i = 0; i < A.r; ++i) T(j, i) = A(i, j);
  return T;
}

Vec column_norms(const Mat& X, int k) {
  Vec out(k);
#pragma omp parallel for schedule(static)
  for (int j = 0; j < k; ++j) {
    const double* x = X.col(j);
    double s = 0.0;
    for (int t = 0; t < X.r; ++t) s += x[t] * x[t];
    out[j] = std::sqrt(s);
  }
  return out;
}

Mat left_cols(const Mat& X, int k) {
  Mat Y(X.r, k);
  std::copy(X.a.begin(), X.a.begin() + static_cast<size_t>(X.r) * k, Y.a.begin());
  return Y;
}

void copy_cols(Mat& dst, int dst_first, const Mat& src, int src_first, int count) {
  for (int j = 0; j < count; ++j)
    std::copy(src.col(src_first + j), src.col(src_first + j) + src.r, dst.col(dst_first + j));
}

// Cholesky G = L L^T (lower). Mirrors Eigen::LLT's success test: every pivot
// must be strictly positive (reference src/core/Orthogonalization.cpp:51-52).
int cholesky_lower(const Mat& G, Mat& L) {
  const int k = G.r;
  L = Mat(k, k);
  for (int j = 0; j < k; ++j) {
    double d = G(j, j);
    for (int p = 0; p < j; ++p) d -= L(j, p) * L(j, p);
    if (!(d > 0.0) || !std::isfinite(d)) return -1;
    const double ljj = std::sqrt(d);
    L(j, j) = ljj;
    for (int i = j + 1; i < k; ++i) {
      double s = G(i, j);
      for (int p = 0; p < j; ++p) s -= L(i, p) * L(j, p);
      L(i, j) = s / ljj;
    }
  }
  return 0;
}

// Inverse of an upper-triangular R (back substitution on the identity).
Mat upper_inverse(const Mat& R) {
  const int k = R.r;
  Mat Ri(k, k);
  for (int j = 0; j < k; ++j) {
    Ri(j, j) = 1.0 / R(j, j);
    for (int i = j - 1; i >= 0; --i) {
      double s = 0.0;
      for (int p = i + 1; p <= j; ++p) s += R(i, p) * Ri(p, j);
      Ri(i, j) = -s / R(i, i);
    }
  }
  return Ri;
}

// Cholesky QR: G = X^T X = R^T R, Q = X R^{-1}.
// Reference: src/core/Orthogonalization.cpp:49-57 (cholesky_qr_from_gram), :290-297.
int cholesky_qr(Mat& X, Mat& R) {
  Mat G = AtB(X, X.c, X, X.c);
  Mat L;
  if (cholesky_lower(G, L) != 0) return -1;
  R = transpose(L);
  Mat Q = AC(X, X.c, upper_inverse(R));
  X = std::move(Q);
  return 0;
}

// Householder thin QR (Golub & Van Loan, Algorithm 5.2.1). Reflector sign follows
// LAPACK/Eigen (beta = -sign(alpha)*||x||), so R's diagonal may be negative.
// Reference: src/core/Orthogonalization.cpp:351-365 (Eigen::HouseholderQR).
int householder_qr(Mat& X, Mat& R) {
  const int n = X.r, k = X.c;
  Mat H = X;                // reflectors stored below the diagonal
  Vec tau(k, 0.0);
  for (int j = 0; j < k; ++j) {
    double* h = H.col(j);
    double xnorm2 = 0.0;
    for (int i = j + 1; i < n; ++i) xnorm2 += h[i] * h[i];
    const double alpha = h[j];
    if (xnorm2 == 0.0) {
      tau[j] = 0.0;
      continue;
    }
    const double beta = -std::copysign(std::sqrt(alpha * alpha + xnorm2), alpha);
    tau[j] = (beta - alpha) / beta;
    const double scale = 1.0 / (alpha - beta);
    for (int i = j + 1; i < n; ++i) h[i] *= scale;
    h[j] = beta;
    // apply I - tau v v^T (v = [1; h[j+1:]]) to the trailing columns
#pragma omp parallel for schedule(static) if (static_cast<long>(n) * (k - j) > 16384)
    for (int c = j + 1; c < k; ++c) {
      double* y = H.col(c);
      double s = y[j];
      for (int i = j + 1; i < n; ++i) s += h[i] * y[i];
      s *= tau[j];
      y[j] -= s;
      for (int i = j + 1; i < n; ++i) y[i] -= s * h[i];
    }
  }
  R = Mat(k, k);
  for (int j = 0; j < k; ++j)
    for (int i = 0; i <= j; ++i) R(i, j) = H(i, j);
  // Q = H_0 ... H_{k-1} [I; 0], formed backwards.
  Mat Q(n, k);
  for (int j = 0; j < k; ++j) Q(j, j) = 1.0;
  for (int j = k - 1; j >= 0; --j) {
    if (tau[j] == 0.0) continue;
    const double* h = H.col(j);
#pragma omp parallel for schedule(static) if (static_cast<long>(n) * (k - j) > 16384)
    for (int c = j; c < k; ++c) {
      double* y = Q.col(c);
      double s = y[j];
      for (int i = j + 1; i < n; ++i) s += h[i] * y[i];
      s *= tau[j];
      y[j] -= s;
      for (int i = j + 1; i < n; ++i) y[i] -= s * h[i];
    }
  }
  X = std::move(Q);
  return 0;
}

// Symmetric eigensolver: Householder tridiagonalization with accumulated Q
// (Golub & Van Loan, Algorithm 8.3.1) followed by the implicit symmetric QR
// algorithm with Wilkinson shifts (GvL Algorithms 8.3.2-8.3.3), rotations
// accumulated into Q. Eigenvalues ascending; V columns are the eigenvectors.
// Stands in for Eigen::SelfAdjointEigenSolver (src/core/RitzExtraction.cpp:104).
int syev(const Mat& Ain, Vec& w, Mat& V) {
  const int m = Ain.r;
  Mat A = Ain;
  Vec d(m, 0.0), e(std::max(m - 1, 1), 0.0);
  Vec tau(std::max(m, 1), 0.0);
  // Reduction: for column k, reflect A[k+1:, k] onto e1.
  for (int k = 0; k + 2 < m; ++k) {
    const int len = m - k - 1;
    double* x = A.col(k) + (k + 1);
    double xn2 = 0.0;
    for (int i = 1; i < len; ++i) xn2 += x[i] * x[i];
    const double alpha = x[0];
    if (xn2 == 0.0) {
      tau[k] = 0.0;
      e[k] = alpha;
      continue;
    }
    const double beta = -std::copysign(std::sqrt(alpha * alpha + xn2), alpha);
    const double t = (beta - alpha) / beta;
    const double sc = 1.0 / (alpha - beta);
    Vec v(len);
    v[0] = 1.0;
    for (int i = 1; i < len; ++i) v[i] = x[i] * sc;
    tau[k] = t;
    e[k] = beta;
    // p = t * A22 v ; w = p - (t/2)(p^T v) v ; A22 -= v w^T + w v^T
    Vec p(len, 0.0);
    for (int j = 0; j < len; ++j) {
      const double* a = A.col(k + 1 + j) + (k + 1);
      double s = 0.0;
      for (int i = 0; i < len; ++i) s += a[i] * v[i];
      p[j] = t * s;
    }
    double pv = 0.0;
    for (int i = 0; i < len; ++i) pv += p[i] * v[i];
    const double kfac = 0.5 * t * pv;
    for (int i = 0; i < len; ++i) p[i] -= kfac * v[i];
    for (int j = 0; j < len; ++j) {
      double* a = A.col(k + 1 + j) + (k + 1);
      for (int i = 0; i < len; ++i) a[i] -= v[i] * p[j] + p[i] * v[j];
    }
    // keep v for accumulation
    x[0] = 1.0;
    for (int i = 1; i < len; ++i) x[i] = v[i];
  }
  for (int i = 0; i < m; ++i) d[i] = A(i, i);
  if (m >= 2) e[m - 2] = A(m - 1, m - 2);
  if (m >= 3) {
    // e[k] for k < m-2 set in the loop; recompute the last coupling only.
  }
  // Accumulate Q = H_0 H_1 ... H_{m-3}.
  V = Mat(m, m);
  for (int i = 0; i < m; ++i) V(i, i) = 1.0;
  for (int k = m - 3; k >= 0; --k) {
    if (tau[k] == 0.0) continue;
    const int len = m - k - 1;
    const double* v = A.col(k) + (k + 1);
    for (int c = k + 1; c < m; ++c) {
      double* y = V.col(c) + (k + 1);
      double s = 0.0;
      for (int i = 0; i < len; ++i) s += v[i] * y[i];
      s *= tau[k];
      for (int i = 0; i < len; ++i) y[i] -= s * v[i];
    }
  }
  // Implicit symmetric QR on (d, e); e[i] couples i and i+1.
  const double eps = std::numeric_limits<double>::epsilon();
  int hi = m - 1;
  long guard = 0;
  while (hi > 0) {
    for (int i = 0; i < hi; ++i)
      if (std::fabs(e[i]) <= eps * (std::fabs(d[i]) + std::fabs(d[i + 1]))) e[i] = 0.0;
    if (e[hi - 1] == 0.0) {
      --hi;
      continue;
    }
    int lo = hi - 1;
    while (lo > 0 && e[lo - 1] != 0.0) --lo;
    if (++guard > 60L * m + 100) return -1;
    const double a = d[hi - 1], b = e[hi - 1], c = d[hi];
    const double delta = 0.5 * (a - c);
    const double den = delta + std::copysign(std::hypot(delta, b), delta == 0.0 ? 1.0 : delta);
    const double mu = c - b * b / den;
    double x = d[lo] - mu, z = e[lo];
    for (int k = lo; k < hi; ++k) {
      const double r = std::hypot(x, z);
      double cs = 1.0, sn = 0.0;
      if (r != 0.0) {
        cs = x / r;
        sn = z / r;
      }
      if (k > lo) e[k - 1] = r;
      const double dk = d[k], dk1 = d[k + 1], ek = e[k];
      d[k] = cs * cs * dk + 2.0 * cs * sn * ek + sn * sn * dk1;
      d[k + 1] = sn * sn * dk - 2.0 * cs * sn * ek + cs * cs * dk1;
      e[k] = cs * sn * (dk1 - dk) + (cs * cs - sn * sn) * ek;
      if (k + 1 < hi) {
        z = sn * e[k + 1];
        e[k + 1] = cs * e[k + 1];
        x = e[k];
      }
      double* q0 = V.col(k);
      double* q1 = V.col(k + 1);
      for (int i = 0; i < m; ++i) {
        const double u = q0[i], t = q1[i];
        q0[i] = cs * u + sn * t;
        q1[i] = -sn * u + cs * t;
      }
    }
  }
  // Sort ascending (stable), permute V.
  std::vector<int> idx(m);
  std::iota(idx.begin(), idx.end(), 0);
  std::stable_sort(idx.begin(), idx.end(), [&](int i, int j) { return d[i] < d[j]; });
  w.assign(m, 0.0);
  Mat Vs(m, m);
  for (int j = 0; j < m; ++j) {
    w[j] = d[idx[j]];
    std::copy(V.col(idx[j]), V.col(idx[j]) + m, Vs.col(j));
  }
  V = std::move(Vs);
  return 0;
}

// One-sided (Hestenes) Jacobi SVD of a small square matrix R = U S W^T.
// Singular values descending; U's columns for zero singular values are zero.
// Stands in for Eigen::JacobiSVD (src/core/Orthogonalization.cpp:756).
void jacobi_svd(const Mat& R, Vec& s, Mat& U) {
  const int k = R.c, r = R.r;
  Mat A = R;
  const double eps = std::numeric_limits<double>::epsilon();
  for (int sweep = 0; sweep < 80; ++sweep) {
    bool rotated = false;
    for (int p = 0; p < k - 1; ++p)
      for (int q = p + 1; q < k; ++q) {
        double al = 0, be = 0, ga = 0;
        for (int i = 0; i < r; ++i) {
          al += A(i, p) * A(i, p);
          be += A(i, q) * A(i, q);
          ga += A(i, p) * A(i, q);
        }
        if (std::fabs(ga) <= eps * std::sqrt(al * be) || ga == 0.0) continue;
        rotated = true;
        const double zeta = (be - al) / (2.0 * ga);
        const double t = std::copysign(1.0, zeta) / (std::fabs(zeta) + std::sqrt(1.0 + zeta * zeta));
        const double c = 1.0 / std::sqrt(1.0 + t * t), sn = c * t;
        for (int i = 0; i < r; ++i) {
          const double x = A(i, p), y = A(i, q);
          A(i, p) = c * x - sn * y;
          A(i, q) = sn * x + c * y;
        }
      }
    if (!rotated) break;
  }
  Vec sv(k);
  for (int j = 0; j < k; ++j) {
    double t = 0;
    for (int i = 0; i < r; ++i) t += A(i, j) * A(i, j);
    sv[j] = std::sqrt(t);
  }
  std::vector<int> idx(k);
  std::iota(idx.begin(), idx.end(), 0);
  std::stable_sort(idx.begin(), idx.end(), [&](int a, int b) { return sv[a] > sv[b]; });
  s.assign(k, 0.0);
  U = Mat(r, k);
  for (int j = 0; j < k; ++j) {
    s[j] = sv[idx[j]];
    if (s[j] > 0)
      for (int i = 0; i < r; ++i) U(i, j) = A(i, idx[j]) / s[j];
  }
}

// ---------------------------------------------------------------- sparse

// Y[:, :k] = alpha * A * X[:, xf:xf+k] + beta * Y  (full CSR).
// Reference: LinearOperator::apply_block semantics, include/core/LinearOperator.hpp:32-37;
// CPU product src/fasteig_problem.cpp:649-655.
void spmm(const Csr& A, const Mat& X, int xf, int k, Mat& Y, double alpha, double beta) {
  const int n = A.n;
  // columns in groups of JB, packed row-major so a nonzero reads JB
  // consecutive values; each entry still sums its row's nonzeros in CSR order
  constexpr int JB = 16;
  std::vector<double> xr(static_cast<size_t>(X.r) * JB);
  for (int j0 = 0; j0 < k; j0 += JB) {
    const int jb = std::min(JB, k - j0);
#pragma omp parallel for schedule(static)
    for (int t = 0; t < X.r; ++t)
      for (int j = 0; j < JB; ++j) xr[static_cast<size_t>(t) * JB + j] = j < jb ? X(t, xf + j0 + j) : 0.0;
#pragma omp parallel for schedule(dynamic, 256)
    for (int i = 0; i < n; ++i) {
      double s[JB] = {};
      for (int p = A.rowptr[i]; p < A.rowptr[i + 1]; ++p) {
        const double v = A.vals[p];
        const double* x = &xr[static_cast<size_t>(A.colind[p]) * JB];
        for (int j = 0; j < JB; ++j) s[j] += v * x[j];
      }
      for (int j = 0; j < jb; ++j) {
        double& y = Y(i, j0 + j);
        y = alpha * s[j] + (beta == 0.0 ? 0.0 : beta * y);
      }
    }
  }
}

// ---------------------------------------------------------------- random

// Reference: src/core/Executor.cpp:44-62 (mt19937(seed), uniform(-1,1), column-major).
Mat random_block(int n, int k, unsigned seed) {
  std::mt19937 rng(seed);
  std::uniform_real_distribution<double> dist(-1.0, 1.0);
  Mat X(n, k);
  for (int j = 0; j < k; ++j)
    for (int i = 0; i < n; ++i) X(i, j) = dist(rng);
  return X;
}

// ---------------------------------------------------------------- orthogonalization

// Two-pass classical block Gram-Schmidt: H = V^T S; R = S - V H; S2 = V^T R;
// R -= V S2; H += S2. Reference: src/core/Orthogonalization.cpp:482-527.
void block_gram_schmidt(const Mat& V, int active, const Mat& S, Mat& R, Mat& H) {
  R = S;
  H = Mat(active, S.c);
  if (active == 0) return;
  H = AtB(V, active, S, S.c);
  Mat VH = AC(V, active, H);
  for (size_t t = 0; t < R.a.size(); ++t) R.a[t] -= VH.a[t];
  Mat S2 = AtB(V, active, R, R.c);
  Mat VS = AC(V, active, S2);
  for (size_t t = 0; t < R.a.size(); ++t) R.a[t] -= VS.a[t];
  for (size_t t = 0; t < H.a.size(); ++t) H.a[t] += S2.a[t];
}

// Rank-revealing block orthonormalization: two passes of [GS against basis,
// Householder QR, Jacobi SVD of R, keep sigma > tol*max||col||, Q U].
// Returns the number of kept columns (0 = nothing survived).
// Reference: src/core/Orthogonalization.cpp:717-775.
int orthonormalize_block(const Mat& basis, int active, const Mat& source, Mat& out,
                         double rank_tol = 1e-12) {
  if (active == basis.r) return 0;
  Mat current = source;
  for (int pass = 0; pass < 2; ++pass) {
    Vec norms = column_norms(current, current.c);
    double scale = 0.0;
    for (double v : norms) {
      if (!std::isfinite(v)) throw std::runtime_error("orthonormalize_block: non-finite column norms.");
      scale = std::max(scale, v);
    }
    if (scale == 0.0) return 0;
    Mat orth, proj, R;
    block_gram_schmidt(basis, active, current, orth, proj);
    householder_qr(orth, R);
    Vec sv;
    Mat U;
    jacobi_svd(R, sv, U);
    int cap = std::min(current.c, source.r - active);
    int rank = 0;
    while (rank < cap && sv[rank] > rank_tol * scale) ++rank;
    if (rank == 0) return 0;
    Mat Ur(U.r, rank);
    copy_cols(Ur, 0, U, 0, rank);
    current = AC(orth, orth.c, Ur);
  }
  out = std::move(current);
  return out.c;
}

// ---------------------------------------------------------------- Ritz extraction

struct Ritz {
  Vec values, relres;
  Mat coeffs;
};

// stable index sort by `which` (include/fasteig_solver.hpp:59-82)
std::vector<int> sort_indices(const Vec& v, const std::string& which) {
  std::vector<int> idx(v.size());
  std::iota(idx.begin(), idx.end(), 0);
  if (which == "sr")
    std::stable_sort(idx.begin(), idx.end(), [&](int a, int b) { return v[a] < v[b]; });
  else if (which == "lr")
    std::stable_sort(idx.begin(), idx.end(), [&](int a, int b) { return v[a] > v[b]; });
  else if (which == "sm")
    std::stable_sort(idx.begin(), idx.end(), [&](int a, int b) { return std::fabs(v[a]) < std::fabs(v[b]); });
  else if (which == "lm")
    std::stable_sort(idx.begin(), idx.end(), [&](int a, int b) { return std::fabs(v[a]) > std::fabs(v[b]); });
  else if (which == "be") {
    const int m = static_cast<int>(v.size());
    int pos = 0;
    for (int i = 0; i < m / 2; ++i) {
      idx[pos++] = m - 1 - i;
      idx[pos++] = i;
    }
    if (m % 2 == 1) idx[pos] = m / 2;
  }
  return idx;
}

// Reference: src/core/RitzExtraction.cpp:65-141.
Ritz extract_ritz(const Mat& projected, int m, const Mat& Q, const Mat& AQ, int requested,
                  const std::string& which, Mat& Y, Mat& AY, Mat& Rs) {
  Mat P(m, m);
  for (int j = 0; j < m; ++j)
    for (int i = 0; i < m; ++i) P(i, j) = 0.5 * (projected(i, j) + projected(j, i));
  Vec w;
  Mat V;
  if (syev(P, w, V) != 0) throw std::runtime_error("projected eigensolve failed.");
  auto idx = sort_indices(w, which);
  Ritz out;
  out.values.resize(requested);
  out.coeffs = Mat(m, requested);
  for (int j = 0; j < requested; ++j) {
    out.values[j] = w[idx[j]];
    std::copy(V.col(idx[j]), V.col(idx[j]) + m, out.coeffs.col(j));
  }
  Y = AC(Q, m, out.coeffs);
  AY = AC(AQ, m, out.coeffs);
  Rs = AY;
  for (int j = 0; j < requested; ++j) {
    double* r = Rs.col(j);
    const double* y = Y.col(j);
    for (int t = 0; t < Rs.r; ++t) r[t] -= out.values[j] * y[t];
  }
  Vec rn = column_norms(Rs, requested), yn = column_norms(Y, requested);
  out.relres.resize(requested);
  for (int j = 0; j < requested; ++j) {
    const double den = (1.0 + std::fabs(out.values[j])) * yn[j];
    out.relres[j] = den > 0.0 ? rn[j] / den : rn[j];
  }
  return out;
}

// ---------------------------------------------------------------- Lanczos bounds

// 20-step Lanczos with two-pass full reorthogonalization; a = min(theta - |beta*y_last|),
// b = max(theta + |beta*y_last|). Reference: src/ChebyshevSolver.cpp:26-49,
// src/core/KrylovSubspace.cpp:112-178,232-259, src/core/LanczosWorkspace.cpp:25-39,
// src/core/RitzExtraction.cpp:181-216,252-270,435-441.
void lanczos_bounds(const Csr& A, double sign, unsigned seed, double& lower, double& upper) {
  const int n = A.n;
  const int steps = std::min(20, n);
  Mat r = random_block(n, 1, seed);
  Mat basis(n, steps), av(n, 1);
  Mat T(steps, steps);
  double beta_next = 0.0;
  const double eps = std::numeric_limits<double>::epsilon();
  auto two_pass = [&](int active, const Mat& src, Mat& res) -> double {
    res = src;
    if (active == 0) return 0.0;
    Mat h = AtB(basis, active, src, 1);
    Mat vh = AC(basis, active, h);
    for (int t = 0; t < n; ++t) res.a[t] -= vh.a[t];
    Mat s = AtB(basis, active, res, 1);
    Mat vs = AC(basis, active, s);
    for (int t = 0; t < n; ++t) res.a[t] -= vs.a[t];
    return h.a[active - 1] + s.a[active - 1];
  };
  for (int it = 0; it < steps; ++it) {
    double beta = column_norms(r, 1)[0];
    bool restarted = false;
    if (!std::isfinite(beta)) throw std::runtime_error("non-finite residual norm.");
    if (!(beta > eps)) {
      Mat cand = random_block(n, 1, seed + static_cast<unsigned>(it + 1));
      two_pass(it, cand, r);
      restarted = true;
      beta = column_norms(r, 1)[0];
      if (beta <= eps) throw std::runtime_error("Lanczos breakdown.");
    }
    for (int t = 0; t < n; ++t) r.a[t] /= beta;
    std::copy(r.a.begin(), r.a.end(), basis.col(it));
    const double recorded_beta = restarted ? 0.0 : beta;
    spmm(A, basis, it, 1, av, sign, 0.0);
    const double alpha = two_pass(it + 1, av, r);
    beta_next = column_norms(r, 1)[0];
    if (it != 0) {
      T(it, it - 1) = recorded_beta;
      T(it - 1, it) = recorded_beta;
    }
    T(it, it) = alpha;
  }
  Vec w;
  Mat V;
  if (syev(T, w, V) != 0) throw std::runtime_error("Lanczos snapshot eigensolve failed.");
  lower = std::numeric_limits<double>::infinity();
  upper = -std::numeric_limits<double>::infinity();
  for (int j = 0; j < steps; ++j) {
    const double rad = std::fabs(beta_next * V(steps - 1, j));
    lower = std::min(lower, w[j] - rad);
    upper = std::max(upper, w[j] + rad);
  }
  if (!std::isfinite(lower) || !std::isfinite(upper) || lower > upper)
    throw std::runtime_error("ChebyshevSolver: invalid spectral bound estimates.");
}

// Filter scalars, reference src/ChebyshevSolver.cpp:232-236.
void cheb_scalars(int degree, double lower, double upper, double& c, double& d) {
  const double xlb = std::cosh(std::acosh(1.0e3) / degree);
  const double cutoff = (2.0 * lower + upper * (xlb - 1.0)) / (1.0 + xlb);
  c = 2.0 / (upper - cutoff);
  d = -(upper + cutoff) / (upper - cutoff);
}

// Reference: src/ChebyshevSolver.cpp:223-258. Operates on sign*A.
void cheb_filter(const Csr& A, double sign, Mat& Xb, int degree, double lower, double upper) {
  if (upper == lower) return;
  double c, d;
  cheb_scalars(degree, lower, upper, c, d);
  if (!std::isfinite(c) || !std::isfinite(d)) throw std::runtime_error("invalid filter scaling.");
  const int k = Xb.c;
  Mat prev = Xb;
  Mat cur(Xb.r, k), nxt(Xb.r, k);
  spmm(A, prev, 0, k, cur, sign * c, 0.0);
  for (size_t t = 0; t < cur.a.size(); ++t) cur.a[t] += d * prev.a[t];
  for (int deg = 2; deg <= degree; ++deg) {
    nxt = prev;
    spmm(A, cur, 0, k, nxt, 2.0 * c * sign, -1.0);
    for (size_t t = 0; t < nxt.a.size(); ++t) nxt.a[t] += 2.0 * d * cur.a[t];
    std::swap(prev, cur);
    std::swap(cur, nxt);
  }
  Xb = std::move(cur);
}

// Scaled Chebyshev filter on [cutoff, upper] normalised at a0 (Zhou & Saad,
// "A Chebyshev-Davidson algorithm for large symmetric eigenproblems", SIAM J.
// Matrix Anal. Appl. 29 (2007), Algorithm 4.1). The adaptive-bound extension
// of the reference's filter: the damped interval starts at the largest Ritz
// value of the current subspace instead of the fixed-amplification cutoff of
// src/ChebyshevSolver.cpp:232-236. Operates on sign*A.
void cheb_filter_scaled(const Csr& A, double sign, Mat& Xb, int degree, double a0, double cutoff, double upper) {
  const double e = 0.5 * (upper - cutoff), c = 0.5 * (upper + cutoff);
  double sigma = e / (a0 - c);
  const double s1 = sigma;
  const int k = Xb.c;
  Mat X = Xb, Y(Xb.r, k), Yn(Xb.r, k);
  spmm(A, X, 0, k, Y, sign * s1 / e, 0.0);
  for (size_t t = 0; t < Y.a.size(); ++t) Y.a[t] -= (s1 / e) * c * X.a[t];
  for (int deg = 2; deg <= degree; ++deg) {
    const double sn = 1.0 / (2.0 / s1 - sigma);
    Yn = X;
    spmm(A, Y, 0, k, Yn, sign * 2.0 * sn / e, -sigma * sn);
    for (size_t t = 0; t < Yn.a.size(); ++t) Yn.a[t] -= (2.0 * sn / e) * c * Y.a[t];
    std::swap(X, Y);
    std::swap(Y, Yn);
    sigma = sn;
  }
  Xb = std::move(Y);
}

// ---------------------------------------------------------------- solvers

struct Result {
  Vec evals, relres;
  Mat evecs;
  long iterations = 0;
  int status = 0;
};

// Reference: src/ChebyshevSolver.cpp:67-221. Extension: block > nev sets the
// subspace capacity (the reference uses nev + min(nev, n - nev)).
Result chebyshev(const Csr& A, int nev, int block, int degree, int maxiter, double tol,
                 const std::string& which, unsigned seed, bool ritz_bounds = false) {
  const int n = A.n;
  if (n <= 0 || nev <= 0 || nev > n) throw std::invalid_argument("ChebyshevSolver: nev must be in [1,n].");
  if (which != "sr" && which != "lr") throw std::invalid_argument("ChebyshevSolver: only SR and LR are supported.");
  const double sign = which == "lr" ? -1.0 : 1.0;
  const int capacity = block > nev ? std::min(block, n) : nev + std::min(nev, n - nev);
  Mat vectors = random_block(n, capacity, seed);
  Mat Rq;
  if (cholesky_qr(vectors, Rq) != 0) throw std::runtime_error("ChebyshevSolver: initial Cholesky QR failed.");
  int locked = 0, subspace_cols = ritz_bounds ? capacity : nev;
  double lower, upper;
  lanczos_bounds(A, sign, seed, lower, upper);
  double cutoff = std::numeric_limits<double>::quiet_NaN();  // adaptive bound (ritz_bounds)
  Ritz ritz;
  Mat ritz_vectors;
  Result res;
  for (int iter = 0; iter < maxiter; ++iter) {
    const int active = subspace_cols - locked;
    Mat filtered(n, active);
    copy_cols(filtered, 0, vectors, locked, active);
    if (ritz_bounds && std::isfinite(cutoff) && cutoff > lower && cutoff < upper)
      cheb_filter_scaled(A, sign, filtered, degree, lower, cutoff, upper);
    else
      cheb_filter(A, sign, filtered, degree, lower, upper);
    if (locked > 0) {
      Mat out, H;
      block_gram_schmidt(vectors, locked, filtered, out, H);
      filtered = std::move(out);
    }
    copy_cols(vectors, locked, filtered, 0, active);
    Mat basis = left_cols(vectors, subspace_cols);
    if (cholesky_qr(basis, Rq) != 0) throw std::runtime_error("ChebyshevSolver: Cholesky QR failed during filtering.");
    Mat AB(n, subspace_cols);
    spmm(A, basis, 0, subspace_cols, AB, sign, 0.0);
    Mat P = AtB(basis, subspace_cols, AB, subspace_cols);
    Mat AY, Rs;
    ritz = extract_ritz(P, subspace_cols, basis, AB, subspace_cols, "sr", ritz_vectors, AY, Rs);
    for (int j = 0; j < subspace_cols; ++j)
      if (!std::isfinite(ritz.values[j]) || !std::isfinite(ritz.relres[j]))
        throw std::runtime_error("ChebyshevSolver: non-finite Ritz values or residuals.");
    copy_cols(vectors, 0, ritz_vectors, 0, subspace_cols);
    cutoff = ritz.values[subspace_cols - 1];
    locked = 0;
    while (locked < nev && ritz.relres[locked] <= tol) ++locked;
    subspace_cols = std::max(subspace_cols, nev + std::min(locked, capacity - nev));
    res.iterations = iter + 1;
    double mx = 0.0;
    for (int j = 0; j < nev; ++j) mx = std::max(mx, ritz.relres[j]);
    if (mx <= tol) break;
  }
  res.evals.assign(ritz.values.begin(), ritz.values.begin() + nev);
  for (double& v : res.evals) v *= sign;
  res.relres.assign(ritz.relres.begin(), ritz.relres.begin() + nev);
  res.evecs = left_cols(ritz_vectors, nev);
  return res;
}

// Reference: src/LOBPCGSolver.cpp:25-173. Extension: block > nev widens X (the
// reference fixes the block to nev); convergence and output use the first nev.
Result lobpcg(const Csr& A, int nev, int block, int maxiter, double tol, const std::string& which,
              unsigned seed) {
  const int n = A.n;
  if (n <= 0 || nev <= 0 || nev > n || maxiter <= 0)
    throw std::invalid_argument("LOBPCGSolver: require 0 < nev <= n and maxiter > 0.");
  if (which != "sr" && which != "lr") throw std::invalid_argument("LOBPCGSolver: only SR and LR are supported.");
  const double sign = which == "lr" ? -1.0 : 1.0;
  const int bw = std::min(std::max(block, nev), n);
  Mat X = random_block(n, bw, seed);
  Mat R;
  householder_qr(X, R);
  const int capacity = static_cast<int>(std::min<long>(n, 3L * bw));
  Mat AX(n, bw);
  spmm(A, X, 0, bw, AX, sign, 0.0);
  Mat P0 = AtB(X, bw, AX, bw);
  Mat Xn, AXn, Res;
  Ritz ritz = extract_ritz(P0, bw, X, AX, bw, "sr", Xn, AXn, Res);
  X = Xn;
  AX = AXn;
  Mat P;
  bool haveP = false;
  Result out;
  auto conv_max = [&](const Ritz& rz) {
    double m = 0;
    for (int j = 0; j < nev; ++j) m = std::max(m, rz.relres[j]);
    return m;
  };
  for (int iter = 0; iter < maxiter; ++iter) {
    for (int j = 0; j < bw; ++j)
      if (!std::isfinite(ritz.values[j]) || !std::isfinite(ritz.relres[j]))
        throw std::runtime_error("LOBPCGSolver: non-finite Ritz values or residuals.");
    if (conv_max(ritz) <= tol) break;
    std::vector<int> active;
    for (int j = 0; j < bw; ++j)
      if (ritz.relres[j] > tol) active.push_back(j);
    const int na = static_cast<int>(active.size());
    Mat Wsrc(n, na);
    for (int j = 0; j < na; ++j) copy_cols(Wsrc, j, Res, active[j], 1);
    householder_qr(X, R);
    Mat basis(n, capacity);
    copy_cols(basis, 0, X, 0, bw);
    int sc = bw;
    Mat W;
    if (orthonormalize_block(basis, sc, Wsrc, W) > 0) {
      const int add = std::min(W.c, capacity - sc);
      copy_cols(basis, sc, W, 0, add);
      sc += add;
    }
    if (haveP && sc < capacity) {
      Mat Psrc(n, na);
      for (int j = 0; j < na; ++j) copy_cols(Psrc, j, P, active[j], 1);
      Mat Po;
      if (orthonormalize_block(basis, sc, Psrc, Po) > 0) {
        const int add = std::min(Po.c, capacity - sc);
        copy_cols(basis, sc, Po, 0, add);
        sc += add;
      }
    }
    Mat B = left_cols(basis, sc);
    Mat AB(n, sc);
    spmm(A, B, 0, sc, AB, sign, 0.0);
    Mat Pm = AtB(B, sc, AB, sc);
    ritz = extract_ritz(Pm, sc, B, AB, bw, "sr", X, AX, Res);
    for (int j = 0; j < bw; ++j)
      if (!std::isfinite(ritz.values[j]) || !std::isfinite(ritz.relres[j]))
        throw std::runtime_error("LOBPCGSolver: non-finite Ritz values or residuals.");
    Mat C = ritz.coeffs;
    for (int j = 0; j < C.c; ++j)
      for (int i = 0; i < bw; ++i) C(i, j) = 0.0;
    P = AC(B, sc, C);
    haveP = true;
    out.iterations = iter + 1;
    const bool exhausted = sc == bw;
    if (conv_max(ritz) <= tol || exhausted) break;
  }
  out.evals.assign(ritz.values.begin(), ritz.values.begin() + nev);
  for (double& v : out.evals) v *= sign;
  out.relres.assign(ritz.relres.begin(), ritz.relres.begin() + nev);
  out.evecs = left_cols(X, nev);
  return out;
}

// Reference: src/DavidsonSolver.cpp:27-264.
Result davidson(const Csr& A, int nev, int maxblks, int maxiter, double tol, const std::string& which,
                unsigned seed) {
  const int n = A.n;
  if (n <= 0) throw std::invalid_argument("DavidsonSolver: n must be > 0.");
  if (nev <= 0 || nev > n) throw std::invalid_argument("DavidsonSolver: nev must be in [1,n].");
  if (maxiter <= 0) throw std::invalid_argument("DavidsonSolver: maxiter must be > 0.");
  if (maxblks < 1) throw std::invalid_argument("DavidsonSolver: maxblks must be at least 1.");
  Mat V0 = random_block(n, nev, seed);
  Mat R;
  if (cholesky_qr(V0, R) != 0) throw std::runtime_error("DavidsonSolver: initial Cholesky QR failed.");
  const int maxcols = std::min(n, maxblks * nev);
  Mat basis(n, maxcols), abasis(n, maxcols);
  Mat AV0(n, nev);
  spmm(A, V0, 0, nev, AV0, 1.0, 0.0);
  copy_cols(basis, 0, V0, 0, nev);
  copy_cols(abasis, 0, AV0, 0, nev);
  Mat projected(maxcols, maxcols);
  {
    Mat p = AtB(basis, nev, abasis, nev);
    for (int j = 0; j < nev; ++j)
      for (int i = 0; i < nev; ++i) projected(i, j) = p(i, j);
  }
  auto top = [&](int m) {
    Mat p(m, m);
    for (int j = 0; j < m; ++j)
      for (int i = 0; i < m; ++i) p(i, j) = projected(i, j);
    return p;
  };
  Mat Y, AY, Res;
  Ritz ritz = extract_ritz(top(nev), nev, basis, abasis, nev, which, Y, AY, Res);
  const double locktol = std::min(1.0e-5, tol);
  Result out;
  for (int iter = 0; iter < maxiter; ++iter) {
    householder_qr(Y, R);
    spmm(A, Y, 0, nev, AY, 1.0, 0.0);
    std::fill(basis.a.begin(), basis.a.end(), 0.0);
    std::fill(abasis.a.begin(), abasis.a.end(), 0.0);
    copy_cols(basis, 0, Y, 0, nev);
    copy_cols(abasis, 0, AY, 0, nev);
    std::fill(projected.a.begin(), projected.a.end(), 0.0);
    {
      Mat p = AtB(basis, nev, abasis, nev);
      for (int j = 0; j < nev; ++j)
        for (int i = 0; i < nev; ++i) projected(i, j) = p(i, j);
    }
    int act = nev;
    ritz = extract_ritz(top(nev), nev, basis, abasis, nev, which, Y, AY, Res);
    for (int blk = 1; blk < maxblks; ++blk) {
      std::vector<int> ai;
      for (int j = 0; j < nev; ++j)
        if (ritz.relres[j] > locktol) ai.push_back(j);
      const int cc = static_cast<int>(ai.size());
      if (cc == 0 || act + cc > maxcols) break;
      Mat src(n, cc);
      for (int j = 0; j < cc; ++j) copy_cols(src, j, Res, ai[j], 1);
      Mat corr, H;
      block_gram_schmidt(basis, act, src, corr, H);
      householder_qr(corr, R);
      Mat acorr(n, cc);
      spmm(A, corr, 0, cc, acorr, 1.0, 0.0);
      Mat cross = AtB(basis, act, acorr, cc);
      Mat dg = AtB(corr, cc, acorr, cc);
      for (int j = 0; j < cc; ++j)
        for (int i = 0; i < act; ++i) {
          projected(i, act + j) = cross(i, j);
          projected(act + j, i) = cross(i, j);
        }
      for (int j = 0; j < cc; ++j)
        for (int i = 0; i < cc; ++i) projected(act + i, act + j) = 0.5 * (dg(i, j) + dg(j, i));
      copy_cols(basis, act, corr, 0, cc);
      copy_cols(abasis, act, acorr, 0, cc);
      act += cc;
      ritz = extract_ritz(top(act), act, basis, abasis, nev, which, Y, AY, Res);
    }
    out.iterations = iter + 1;
    double mx = 0;
    for (double r : ritz.relres) mx = std::max(mx, r);
    if (!ritz.relres.empty() && mx < tol) break;
  }
  out.evals = ritz.values;
  out.relres = ritz.relres;
  out.evecs = Y;
  return out;
}

// Final residuals rho_i = ||A x - lambda x|| / ((1+|lambda|) ||x||).
// Reference: src/fasteig_solver.cpp:89-129.
// ---------------------------------------------------------------- block Krylov-Schur

Vec residual_norms(const Csr& A, const Vec& evals, const Mat& X);

// Two-pass projection of one vector against the first `active` columns of V.
// Reference: src/core/MultiVector.cpp:119-151 (orthogonalize_active_columns).
void orth_vector(const Mat& V, int active, Vec& v) {
  const int n = V.r;
  for (int pass = 0; pass < 2; ++pass) {
    Vec h(active, 0.0);
    for (int j = 0; j < active; ++j) {
      const double* c = V.col(j);
      double s = 0.0;
      for (int i = 0; i < n; ++i) s += c[i] * v[i];
      h[j] = s;
    }
    for (int j = 0; j < active; ++j) {
      const double* c = V.col(j);
      for (int i = 0; i < n; ++i) v[i] -= h[j] * c[i];
    }
  }
}

double norm2(const Vec& v) {
  double s = 0.0;
  for (double x : v) s += x * x;
  return std::sqrt(s);
}

// Rank-aware block QR of `source` against basis[:, :active]: Cholesky QR when
// it is accurate (R diagonal above 64 eps * reference_norm, Q^T Q = I and
// basis^T Q = 0 to 1e-10), else column-by-column with pivoting over the source
// columns and seeded random completion of missing directions; R = Q^T source
// with roundoff-sized columns zeroed. Returns true when directions were
// replenished. Reference: src/core/Orthogonalization.cpp:609-715.
bool complete_block_qr(const Mat& basis, int active, const Mat& source, Mat& Q, Mat& R, unsigned seed,
                       double reference_norm) {
  const int n = source.r, b = source.c;
  const int capacity = std::min(b, n - active);
  const double threshold = 64.0 * std::numeric_limits<double>::epsilon() * reference_norm;
  if (capacity == b) {
    Mat Qc = source, Rc;
    if (cholesky_qr(Qc, Rc) == 0) {
      bool ok = true;
      double dmin = std::numeric_limits<double>::infinity();
      for (int j = 0; j < b; ++j) {
        dmin = std::min(dmin, std::fabs(Rc(j, j)));
        for (int i = 0; i < b; ++i) ok = ok && std::isfinite(Rc(i, j));
      }
      Mat G = AtB(Qc, b, Qc, b);
      double ge = 0.0;
      for (int j = 0; j < b; ++j)
        for (int i = 0; i < b; ++i) {
          const double d = G(i, j) - (i == j ? 1.0 : 0.0);
          ge += d * d;
        }
      double ce = 0.0;
      if (active > 0) {
        Mat Cr = AtB(basis, active, Qc, b);
        for (double x : Cr.a) ce += x * x;
      }
      if (ok && dmin > threshold && std::isfinite(ge) && std::isfinite(ce) && std::sqrt(ge) < 1e-10 &&
          std::sqrt(ce) < 1e-10) {
        Q = std::move(Qc);
        R = std::move(Rc);
        return false;
      }
    }
  }
  bool replenished = false;
  Q = Mat(n, b);
  auto project = [&](int j, Vec& v) {
    orth_vector(basis, active, v);
    orth_vector(Q, j, v);
  };
  for (int j = 0; j < capacity; ++j) {
    double length = 0.0, original_scale = 0.0;
    Vec best(n, 0.0), v(n);
    for (int col = 0; col < b; ++col) {
      std::copy(source.col(col), source.col(col) + n, v.begin());
      original_scale = std::max(original_scale, norm2(v));
      project(j, v);
      const double cn = norm2(v);
      if (cn > length) {
        length = cn;
        best = v;
      }
    }
    const double cutoff = std::max(threshold, 1e-12 * original_scale);
    if (length > cutoff) {
      v = best;
      length = norm2(v);
    } else {
      replenished = true;
      length = 0.0;
      for (unsigned attempt = 0; attempt < 4; ++attempt) {
        Mat rv = random_block(n, 1, seed + 17u * static_cast<unsigned>(j + 1) + attempt);
        std::copy(rv.col(0), rv.col(0) + n, v.begin());
        const double rn = norm2(v);
        project(j, v);
        length = norm2(v);
        if (length > 1e-12 * rn) break;
        length = 0.0;
      }
      if (length == 0.0) throw std::runtime_error("complete_block_qr: unable to complete the basis.");
    }
    for (int i = 0; i < n; ++i) Q(i, j) = v[i] / length;
  }
  R = AtB(Q, b, source, b);
  for (int j = 0; j < b; ++j) {
    double cn = 0.0;
    for (int i = 0; i < b; ++i) cn += R(i, j) * R(i, j);
    if (std::sqrt(cn) <= threshold)
      for (int i = 0; i < b; ++i) R(i, j) = 0.0;
  }
  return replenished;
}

// Block Krylov-Schur with thick restart (standard problems, which = sr/lr/sm/
// lm/be). Reference: src/BlockKrylovSchurSolver.cpp:26-214 with
// src/core/KrylovSubspace.cpp:261-433 (basis expansion and restart),
// src/core/LanczosWorkspace.cpp:42-73,109-135 (projected matrix),
// src/core/RitzExtraction.cpp:272-380 (block Ritz estimates),
// src/core/RitzExtraction.cpp:34-63 (final residuals).
Result block_krylov_schur(const Csr& A, int nev, int bsz, int maxdim, int maxiter, double tol,
                          const std::string& which, unsigned seed) {
  const int n = A.n;
  if (n <= 0) throw std::invalid_argument("BlockKrylovSchurSolver: n must be > 0.");
  if (nev <= 0 || nev > n) throw std::invalid_argument("BlockKrylovSchurSolver: nev must be in [1,n].");
  if (maxiter <= 0) throw std::invalid_argument("BlockKrylovSchurSolver: maxiter must be > 0.");
  if (bsz <= 0 || bsz > n) throw std::invalid_argument("BlockKrylovSchurSolver: blksize must be in [1,n].");
  const int requested_dim = maxdim > 0 ? maxdim : std::max(2 * nev, 20);
  const int block_count = requested_dim >= n ? 1 + (n - 1) / bsz : requested_dim / bsz;
  const int cycle_dim = std::min(n, block_count * bsz);
  const int wanted_blocks = 1 + (nev - 1) / bsz;
  if (cycle_dim < n && block_count <= wanted_blocks)
    throw std::invalid_argument("BlockKrylovSchurSolver: maxdim must accommodate ceil(nev/blksize) retained blocks "
                                "plus at least one expansion block (or the full matrix dimension).");
  const int pdim = (block_count + 1) * bsz;
  const int retained_blocks = cycle_dim == n ? 0 : std::max(wanted_blocks, block_count / 2);
  const int retained_dim = retained_blocks * bsz;
  const int snap_count = std::max(nev, retained_dim);

  Mat Qb(n, pdim);   // basis
  Mat T(pdim, pdim); // projected matrix
  Mat beta(bsz, bsz);
  Mat init = random_block(n, bsz, seed);
  {
    Mat Q0, R0;
    complete_block_qr(Qb, 0, init, Q0, R0, seed, 0.0);
    copy_cols(Qb, 0, Q0, 0, bsz);
  }
  auto expand = [&](int first) {
    for (int blk = first; blk < block_count; ++blk) {
      Mat Z(n, bsz);
      spmm(A, Qb, blk * bsz, bsz, Z, 1.0, 0.0);
      const int active = std::min(n, (blk + 1) * bsz);
      Mat Rz, H;
      block_gram_schmidt(Qb, active, Z, Rz, H);
      const int width = 1 + (active - 1) % bsz;
      Mat D(bsz, bsz);
      for (int j = 0; j < width; ++j)
        for (int i = 0; i < width; ++i) D(i, j) = H(active - width + i, j);
      double scale = 0.0;
      for (double x : column_norms(Z, bsz)) scale += x * x;
      scale = std::sqrt(scale);
      Mat Qn, B;
      const int nf = std::min((blk + 1) * bsz, n);
      complete_block_qr(Qb, nf, Rz, Qn, B, seed + static_cast<unsigned>((blk + 1) * bsz), scale);
      copy_cols(Qb, (blk + 1) * bsz, Qn, 0, bsz);
      const int row = blk * bsz, nrow = row + bsz;
      for (int j = 0; j < bsz; ++j)
        for (int i = 0; i < bsz; ++i) {
          T(row + i, row + j) = D(i, j);
          T(nrow + i, row + j) = B(i, j);
          T(row + j, nrow + i) = B(i, j);
        }
      beta = B;
    }
  };
  expand(0);
  Result res;
  Vec lam_keep;
  Mat C_keep;
  for (int iter = 1; iter <= maxiter; ++iter) {
    // block Ritz snapshot of T[:m, :m]
    const int m = cycle_dim;
    Mat P(m, m);
    for (int j = 0; j < m; ++j)
      for (int i = 0; i < m; ++i) P(i, j) = T(i, j);
    Vec w;
    Mat V;
    if (syev(P, w, V) != 0) throw std::runtime_error("BlockKrylovSchurSolver: projected eigensolve failed.");
    const auto idx = sort_indices(w, which);
    const int cnt = std::min(snap_count, m);
    const int tail = 1 + (m - 1) % bsz;
    Vec lam(cnt), rr(cnt);
    Mat Csel(m, cnt);
    for (int j = 0; j < cnt; ++j) {
      const int c = idx[j];
      lam[j] = w[c];
      Vec ylast(bsz, 0.0);
      for (int t = 0; t < tail; ++t) ylast[t] = V(m - tail + t, c);
      double a = 0.0;
      for (int i = 0; i < bsz; ++i) {
        double s = 0.0;
        for (int t = 0; t < bsz; ++t) s += beta(i, t) * ylast[t];
        a += s * s;
      }
      rr[j] = std::sqrt(a) / (1.0 + std::fabs(w[c]));
      std::copy(V.col(c), V.col(c) + m, Csel.col(j));
    }
    res.iterations = iter;
    double mx = 0.0;
    for (int j = 0; j < std::min(nev, cnt); ++j) mx = std::max(mx, rr[j]);
    bool converged = mx < tol;
    const bool at_limit = iter == maxiter || cycle_dim == n;
    if (converged || at_limit) {
      Mat Cn(m, nev);
      copy_cols(Cn, 0, Csel, 0, nev);
      Mat X = AC(Qb, m, Cn);
      Vec ev(lam.begin(), lam.begin() + nev);
      Vec r = residual_norms(A, ev, X);
      converged = true;
      for (double x : r) converged = converged && x <= tol;
      res.evals = ev;
      res.relres = r;
      res.evecs = X;
      res.status = converged ? 0 : 1;
      break;
    }
    // thick restart: Q <- [Q C_keep, Q_residual], T <- [diag(theta), F^T; F, *]
    Mat coupling(bsz, cnt);
    for (int j = 0; j < cnt; ++j)
      for (int i = 0; i < bsz; ++i) {
        double s = 0.0;
        for (int t = 0; t < bsz; ++t) s += beta(i, t) * Csel(m - bsz + t, j);
        coupling(i, j) = s;
      }
    Mat U = AC(Qb, m, Csel);
    Mat Qres(n, bsz);
    copy_cols(Qres, 0, Qb, block_count * bsz, bsz);
    Qb = Mat(n, pdim);
    copy_cols(Qb, 0, U, 0, cnt);
    copy_cols(Qb, cnt, Qres, 0, bsz);
    T = Mat(pdim, pdim);
    for (int j = 0; j < cnt; ++j) T(j, j) = lam[j];
    for (int j = 0; j < cnt; ++j)
      for (int i = 0; i < bsz; ++i) {
        T(cnt + i, j) = coupling(i, j);
        T(j, cnt + i) = coupling(i, j);
      }
    expand(retained_blocks);
  }
  return res;
}

// Block Lanczos without restart (standard problems). Reference:
// src/BlockLanczosSolver.cpp:24-172: Q_0 from the rank-aware QR of the start
// block; block iter expands Q_{iter-1} (two-pass block Gram-Schmidt, block QR
// of the residual into Q_iter), then checks the block Ritz estimates of the
// nev wanted pairs once iter * bsz >= nev; at most maxiter blocks (and the
// whole space).
Result block_lanczos(const Csr& A, int nev, int bsz, int maxiter, double tol, const std::string& which,
                     unsigned seed) {
  const int n = A.n;
  if (n <= 0) throw std::invalid_argument("BlockLanczosSolver: n must be > 0.");
  if (nev <= 0) throw std::invalid_argument("BlockLanczosSolver: nev must be > 0.");
  if (maxiter <= 0) throw std::invalid_argument("BlockLanczosSolver: maxiter must be > 0.");
  if (bsz <= 0) throw std::invalid_argument("BlockLanczosSolver: blksize must be > 0.");
  if (bsz > n || nev > n) throw std::invalid_argument("BlockLanczosSolver: blksize and nev must be <= n.");
  if (static_cast<long>(maxiter) * bsz < nev) throw std::invalid_argument("BlockLanczosSolver: need maxiter*blksize >= nev.");
  if (which != "sr" && which != "lr" && which != "sm" && which != "lm" && which != "be")
    throw std::invalid_argument("BlockLanczosSolver: invalid which option (use SR/LR/SM/LM/BE).");
  const int blocks = std::min(maxiter, 1 + (n - 1) / bsz);
  const int pdim = (blocks + 1) * bsz;
  Mat Qb(n, pdim), T(pdim, pdim), beta(bsz, bsz);
  {
    Mat Q0, R0;
    complete_block_qr(Qb, 0, random_block(n, bsz, seed), Q0, R0, seed, 0.0);
    copy_cols(Qb, 0, Q0, 0, bsz);
  }
  auto snapshot = [&](int m, Vec& lam, Vec& rr, Mat& Csel) {
    Mat P(m, m);
    for (int j = 0; j < m; ++j)
      for (int i = 0; i < m; ++i) P(i, j) = T(i, j);
    Vec w;
    Mat V;
    if (syev(P, w, V) != 0) throw std::runtime_error("BlockLanczosSolver: projected eigensolve failed.");
    const auto idx = sort_indices(w, which);
    const int cnt = std::min(nev, m), tail = 1 + (m - 1) % bsz;
    lam.assign(cnt, 0.0);
    rr.assign(cnt, 0.0);
    Csel = Mat(m, cnt);
    for (int j = 0; j < cnt; ++j) {
      const int c = idx[j];
      lam[j] = w[c];
      double a = 0.0;
      for (int i = 0; i < bsz; ++i) {
        double s = 0.0;
        for (int t = 0; t < tail; ++t) s += beta(i, t) * V(m - tail + t, c);
        a += s * s;
      }
      rr[j] = std::sqrt(a) / (1.0 + std::fabs(w[c]));
      std::copy(V.col(c), V.col(c) + m, Csel.col(j));
    }
  };
  Vec lam, rr;
  Mat Csel;
  int it_stop = blocks;
  for (int iter = 1; iter <= blocks; ++iter) {
    const int blk = iter - 1;
    Mat Z(n, bsz);
    spmm(A, Qb, blk * bsz, bsz, Z, 1.0, 0.0);
    const int active = std::min(n, iter * bsz);
    Mat Rz, H;
    block_gram_schmidt(Qb, active, Z, Rz, H);
    const int width = 1 + (active - 1) % bsz;
    Mat D(bsz, bsz);
    for (int j = 0; j < width; ++j)
      for (int i = 0; i < width; ++i) D(i, j) = H(active - width + i, j);
    double scale = 0.0;
    for (double x : column_norms(Z, bsz)) scale += x * x;
    scale = std::sqrt(scale);
    Mat Qn, B;
    complete_block_qr(Qb, active, Rz, Qn, B, seed + static_cast<unsigned>(iter * bsz), scale);
    copy_cols(Qb, iter * bsz, Qn, 0, bsz);
    const int row = blk * bsz, nrow = row + bsz;
    for (int j = 0; j < bsz; ++j)
      for (int i = 0; i < bsz; ++i) {
        T(row + i, row + j) = D(i, j);
        T(nrow + i, row + j) = B(i, j);
        T(row + j, nrow + i) = B(i, j);
      }
    beta = B;
    snapshot(active, lam, rr, Csel);
    double mx = 0.0;
    for (double x : rr) mx = std::max(mx, x);
    if (static_cast<long>(iter) * bsz >= nev && mx < tol) {
      it_stop = iter;
      break;
    }
  }
  const int m = std::min(n, it_stop * bsz);
  snapshot(m, lam, rr, Csel);
  Result res;
  res.evals = lam;
  res.evecs = AC(Qb, m, Csel);
  res.relres = residual_norms(A, res.evals, res.evecs);
  res.iterations = it_stop;
  res.status = 0;
  return res;
}

// Scalar Krylov basis with full two-pass reorthogonalization. Reference:
// src/core/KrylovSubspace.cpp:124-260 (normalize_residual_into_basis with the
// random restart of normalize_or_restart / restart_residual,
// apply_operator_to_basis_column, orthogonalize_operator_result, residual_norm,
// rotate_basis) and src/core/MultiVector.cpp:153-187 (two-pass projection
// returning the coefficient of the newest column).
struct ScalarKrylov {
  const Csr& A;
  int n;
  unsigned seed;
  Mat basis, r, av;
  ScalarKrylov(const Csr& a, int cols, const Mat& start, unsigned s)
      : A(a), n(a.n), seed(s), basis(a.n, cols), r(start), av(a.n, 1) {}
  double project(int active, const Mat& src, Mat& res) {
    res = src;
    if (active == 0) return 0.0;
    Mat h = AtB(basis, active, src, 1);
    Mat vh = AC(basis, active, h);
    for (int t = 0; t < n; ++t) res.a[t] -= vh.a[t];
    Mat s2 = AtB(basis, active, res, 1);
    Mat vs = AC(basis, active, s2);
    for (int t = 0; t < n; ++t) res.a[t] -= vs.a[t];
    return h.a[active - 1] + s2.a[active - 1];
  }
  double norm() const {
    const double b = column_norms(r, 1)[0];
    if (!std::isfinite(b)) throw std::runtime_error("core::KrylovSubspace: non-finite residual norm encountered.");
    return b;
  }
  // the recorded coupling: 0 after a random restart (invariant subspace)
  double normalize_into(int col) {
    const double eps = std::numeric_limits<double>::epsilon();
    double beta = norm();
    bool restarted = false;
    if (!(beta > eps)) {
      Mat cand = random_block(n, 1, seed + static_cast<unsigned>(col + 1));
      project(col, cand, r);
      restarted = true;
      beta = norm();
      if (beta <= eps) throw std::runtime_error("core::KrylovSubspace: breakdown, unable to generate new basis vector.");
    }
    for (int t = 0; t < n; ++t) r.a[t] /= beta;
    std::copy(r.a.begin(), r.a.end(), basis.col(col));
    return restarted ? 0.0 : beta;
  }
  // z = A q_col, orthogonalized against q_0..q_col into r; returns alpha
  double step(int col) {
    spmm(A, basis, col, 1, av, 1.0, 0.0);
    return project(col + 1, av, r);
  }
  void rotate(int active, const Mat& C) {
    Mat U = AC(basis, active, C);
    copy_cols(basis, 0, U, 0, C.c);
  }
};

// Projected tridiagonal (arrowhead after a restart). Reference:
// src/core/LanczosWorkspace.cpp:25-40 (record_step), :97-108
// (prepare_krylov_schur_restart), src/core/RitzExtraction.cpp:195-230,252-268
// (snapshot: |beta_next * y_last| / (1 + |theta|)), :470-492 (restart matrix).
struct Projection {
  Mat T;
  int coupling_row = -1;
  double beta_next = 0.0;
  explicit Projection(int m) : T(m, m) {}
  void record(int it, double beta, double alpha, double bn) {
    if (it != 0 && it != coupling_row) {
      T(it, it - 1) = beta;
      T(it - 1, it) = beta;
    }
    T(it, it) = alpha;
    beta_next = bn;
    if (it == coupling_row) coupling_row = -1;
  }
  void snapshot(int m, int count, const std::string& which, Vec& lam, Vec& rel, Mat& Csel) const {
    Mat P(m, m);
    for (int j = 0; j < m; ++j)
      for (int i = 0; i < m; ++i) P(i, j) = T(i, j);
    Vec w;
    Mat V;
    if (syev(P, w, V) != 0) throw std::runtime_error("projected eigensolve failed.");
    const auto idx = sort_indices(w, which);
    const int cnt = std::min(count, m);
    lam.assign(cnt, 0.0);
    rel.assign(cnt, 0.0);
    Csel = Mat(m, cnt);
    for (int j = 0; j < cnt; ++j) {
      const int c = idx[j];
      lam[j] = w[c];
      rel[j] = std::fabs(beta_next * V(m - 1, c)) / (1.0 + std::fabs(w[c]));
      std::copy(V.col(c), V.col(c) + m, Csel.col(j));
    }
  }
  void restart(int retained, const Vec& lam, const Mat& Csel) {
    const int m = T.r;
    T = Mat(m, m);
    for (int j = 0; j < retained; ++j) {
      T(j, j) = lam[j];
      const double coupling = beta_next * Csel(Csel.r - 1, j);
      T(retained, j) = coupling;
      T(j, retained) = coupling;
    }
    coupling_row = retained;
  }
};

void check_lanczos_which(const std::string& which, const char* who) {
  if (which != "sr" && which != "lr" && which != "sm" && which != "lm" && which != "be")
    throw std::invalid_argument(std::string(who) + ": invalid which option (use SR/LR/SM/LM/BE).");
}

// Scalar Lanczos with full reorthogonalization, no restart (standard problems).
// Reference: src/LanczosSolver.cpp:28-164; convergence checks every 5 steps
// (print_every = 0) once m >= nev, and at maxiter.
Result lanczos(const Csr& A, int nev, int maxiter, double tol, const std::string& which, unsigned seed) {
  const int n = A.n;
  if (maxiter <= 0) throw std::invalid_argument("LanczosSolver: maxiter must be > 0.");
  if (nev <= 0) throw std::invalid_argument("LanczosSolver: nev must be > 0.");
  if (maxiter < nev) throw std::invalid_argument("LanczosSolver: maxiter must be >= nev.");
  if (maxiter > n) throw std::invalid_argument("LanczosSolver: maxiter must be <= n.");
  check_lanczos_which(which, "LanczosSolver");
  ScalarKrylov K(A, maxiter, random_block(n, 1, seed), seed);
  Projection P(maxiter);
  Vec lam, rel;
  Mat Csel;
  int it_stop = maxiter;
  bool converged = false;
  for (int it = 0; it < maxiter; ++it) {
    const double beta = K.normalize_into(it);
    const double alpha = K.step(it);
    P.record(it, beta, alpha, K.norm());
    const int m = it + 1;
    if (!(m >= nev && (m % 5 == 0 || m == maxiter))) continue;
    P.snapshot(m, nev, which, lam, rel, Csel);
    double mx = 0.0;
    for (double x : rel) mx = std::max(mx, x);
    if (!rel.empty() && mx < tol) {
      converged = true;
      it_stop = m;
      break;
    }
  }
  if (!converged) P.snapshot(maxiter, nev, which, lam, rel, Csel);
  Result res;
  res.evals = lam;
  res.evecs = AC(K.basis, it_stop, Csel);
  res.relres = residual_norms(A, res.evals, res.evecs);
  res.iterations = it_stop;
  res.status = 0;
  return res;
}

// Scalar Krylov-Schur with thick restart (standard problems).
// Reference: src/KrylovSchurSolver.cpp:28-167.
Result krylov_schur(const Csr& A, int nev, int maxdim, int maxiter, double tol, const std::string& which,
                    unsigned seed) {
  const int n = A.n;
  if (n <= 0) throw std::invalid_argument("KrylovSchurSolver: n must be > 0.");
  if (nev <= 0 || nev > n) throw std::invalid_argument("KrylovSchurSolver: nev must be in [1,n].");
  if (maxiter <= 0) throw std::invalid_argument("KrylovSchurSolver: maxiter must be > 0.");
  check_lanczos_which(which, "KrylovSchurSolver");
  const int m = std::min(std::max({2 * nev, 20, maxdim}), n);
  ScalarKrylov K(A, m, random_block(n, 1, seed), seed);
  Projection P(m);
  auto expand = [&](int begin) {
    for (int col = begin; col < m; ++col) {
      const double beta = K.normalize_into(col);
      const double alpha = K.step(col);
      P.record(col, beta, alpha, K.norm());
    }
  };
  expand(0);
  Result res;
  for (int iter = 1; iter <= maxiter; ++iter) {
    Vec lam, rel;
    Mat Csel;
    P.snapshot(m, m, which, lam, rel, Csel);
    int nconv = 0;
    for (int j = 0; j < nev; ++j)
      if (rel[j] * (1.0 + std::fabs(lam[j])) < tol * std::fabs(lam[j])) ++nconv;
    if (nconv >= nev || m == n || iter == maxiter) {
      Mat Ch(m, nev);
      copy_cols(Ch, 0, Csel, 0, nev);
      res.evals.assign(lam.begin(), lam.begin() + nev);
      res.evecs = AC(K.basis, m, Ch);
      res.relres = residual_norms(A, res.evals, res.evecs);
      res.iterations = iter;
      res.status = 0;
      return res;
    }
    const int spare = (m - nev) / 2;
    int retained = nev + std::min(nconv, spare);
    if (retained == 1 && m > 3) retained = m / 2;
    Mat Ck(m, retained);
    copy_cols(Ck, 0, Csel, 0, retained);
    K.rotate(m, Ck);
    P.restart(retained, lam, Csel);
    expand(retained);
  }
  throw std::logic_error("KrylovSchurSolver: unreachable solve termination.");
}

Vec residual_norms(const Csr& A, const Vec& evals, const Mat& X) {
  const int k = static_cast<int>(evals.size());
  Mat AX(X.r, k);
  spmm(A, X, 0, k, AX, 1.0, 0.0);
  Vec out(k);
  for (int j = 0; j < k; ++j) {
    double num = 0, xn = 0;
    for (int t = 0; t < X.r; ++t) {
      const double r = AX(t, j) - evals[j] * X(t, j);
      num += r * r;
      xn += X(t, j) * X(t, j);
    }
    const double den = (1.0 + std::fabs(evals[j])) * std::sqrt(xn);
    out[j] = den > 0 ? std::sqrt(num) / den : std::sqrt(num);
  }
  return out;
}

}  // namespace orc

// ================================================================ C ABI

namespace {
thread_local std::string g_err;
orc::Mat wrap(const double* p, int r, int c) {
  orc::Mat m(r, c);
  std::copy(p, p + static_cast<size_t>(r) * c, m.a.begin());
  return m;
}
void unwrap(const orc::Mat& m, double* p) { std::copy(m.a.begin(), m.a.end(), p); }
orc::Csr csr(int n, const int* rp, const int* ci, const double* v) {
  orc::Csr a;
  a.n = n;
  a.rowptr = rp;
  a.colind = ci;
  a.vals = v;
  return a;
}
template <class F>
int guarded(F f) {
  try {
    return f();
  } catch (const std::invalid_argument& e) {
    g_err = e.what();
    return -2;
  } catch (const std::exception& e) {
    g_err = e.what();
    return -3;
  }
}
int finish(const orc::Result& r, int nev, int n, double* ev, double* evec, double* res, long* iters) {
  for (int j = 0; j < nev; ++j) {
    ev[j] = r.evals[j];
    res[j] = r.relres[j];
  }
  if (evec) std::copy(r.evecs.a.begin(), r.evecs.a.begin() + static_cast<size_t>(n) * nev, evec);
  if (iters) *iters = r.iterations;
  return r.status;
}
}  // namespace

extern "C" {

const char* orc_last_error() { return g_err.c_str(); }

void orc_set_threads(int t) {
#ifdef _OPENMP
  if (t > 0) omp_set_num_threads(t);
#else
  (void)t;
#endif
}

int orc_max_threads() {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

void orc_random_block(int n, int k, unsigned seed, double* out) { unwrap(orc::random_block(n, k, seed), out); }

// G = X^T Y (ka x kb) and Y = X C (n x p): the dense products of the solvers.
void orc_gram(int n, int ka, const double* X, int kb, const double* Y, double* G) {
  unwrap(orc::AtB(wrap(X, n, ka), ka, wrap(Y, n, kb), kb), G);
}
void orc_right_multiply(int n, int k, const double* X, int p, const double* Cm, double* Y) {
  unwrap(orc::AC(wrap(X, n, k), k, wrap(Cm, k, p)), Y);
}

void orc_spmm(int n, const int* rp, const int* ci, const double* v, int k, const double* X, double* Y,
              double alpha, double beta) {
  orc::Mat Xm = wrap(X, n, k), Ym = wrap(Y, n, k);
  orc::spmm(csr(n, rp, ci, v), Xm, 0, k, Ym, alpha, beta);
  unwrap(Ym, Y);
}

int orc_syev(int m, const double* A, double* w, double* V) {
  orc::Mat Am = wrap(A, m, m), Vm;
  orc::Vec wv;
  int info = orc::syev(Am, wv, Vm);
  if (info == 0) {
    std::copy(wv.begin(), wv.end(), w);
    unwrap(Vm, V);
  }
  return info;
}

int orc_cholqr(int n, int k, double* X, double* R) {
  orc::Mat Xm = wrap(X, n, k), Rm;
  int info = orc::cholesky_qr(Xm, Rm);
  if (info == 0) {
    unwrap(Xm, X);
    unwrap(Rm, R);
  }
  return info;
}

int orc_householder_qr(int n, int k, double* X, double* R) {
  orc::Mat Xm = wrap(X, n, k), Rm;
  int info = orc::householder_qr(Xm, Rm);
  unwrap(Xm, X);
  unwrap(Rm, R);
  return info;
}

void orc_jacobi_svd(int r, int c, const double* R, double* s, double* U) {
  orc::Mat Rm = wrap(R, r, c), Um;
  orc::Vec sv;
  orc::jacobi_svd(Rm, sv, Um);
  std::copy(sv.begin(), sv.end(), s);
  unwrap(Um, U);
}

int orc_orthonormalize_block(int n, int active, const double* basis, int b, const double* src,
                             double* out) {
  return guarded([&] {
    orc::Mat B = wrap(basis, n, std::max(active, 1)), S = wrap(src, n, b), O;
    if (active == 0) B = orc::Mat(n, 0);
    int r = orc::orthonormalize_block(B, active, S, O);
    if (r > 0) unwrap(O, out);
    return r;
  });
}

int orc_lanczos_bounds(int n, const int* rp, const int* ci, const double* v, double sign, unsigned seed,
                       double* lower, double* upper) {
  return guarded([&] {
    orc::lanczos_bounds(csr(n, rp, ci, v), sign, seed, *lower, *upper);
    return 0;
  });
}

void orc_cheb_scalars(int degree, double lower, double upper, double* c, double* d) {
  orc::cheb_scalars(degree, lower, upper, *c, *d);
}

int orc_cheb_filter(int n, const int* rp, const int* ci, const double* v, int k, double* X, int degree,
                    double lower, double upper) {
  return guarded([&] {
    orc::Mat Xm = wrap(X, n, k);
    orc::cheb_filter(csr(n, rp, ci, v), 1.0, Xm, degree, lower, upper);
    unwrap(Xm, X);
    return 0;
  });
}

int orc_cheb_filter_scaled(int n, const int* rp, const int* ci, const double* v, int k, double* X, int degree,
                           double a0, double cutoff, double upper) {
  return guarded([&] {
    orc::Mat Xm = wrap(X, n, k);
    orc::cheb_filter_scaled(csr(n, rp, ci, v), 1.0, Xm, degree, a0, cutoff, upper);
    unwrap(Xm, X);
    return 0;
  });
}

int orc_chebyshev(int n, const int* rp, const int* ci, const double* v, int nev, int block, int degree,
                  int ritz_bounds, int maxiter, double tol, const char* which, unsigned seed, double* ev,
                  double* evec, double* res, long* iters) {
  return guarded([&] {
    auto r = orc::chebyshev(csr(n, rp, ci, v), nev, block, degree, maxiter, tol, which, seed, ritz_bounds != 0);
    return finish(r, nev, n, ev, evec, res, iters);
  });
}

int orc_lobpcg(int n, const int* rp, const int* ci, const double* v, int nev, int block, int maxiter,
               double tol, const char* which, unsigned seed, double* ev, double* evec, double* res,
               long* iters) {
  return guarded([&] {
    auto r = orc::lobpcg(csr(n, rp, ci, v), nev, block, maxiter, tol, which, seed);
    return finish(r, nev, n, ev, evec, res, iters);
  });
}

int orc_davidson(int n, const int* rp, const int* ci, const double* v, int nev, int maxblks, int maxiter,
                 double tol, const char* which, unsigned seed, double* ev, double* evec, double* res,
                 long* iters) {
  return guarded([&] {
    auto r = orc::davidson(csr(n, rp, ci, v), nev, maxblks, maxiter, tol, which, seed);
    return finish(r, nev, n, ev, evec, res, iters);
  });
}

int orc_block_krylov_schur(int n, const int* rp, const int* ci, const double* v, int nev, int block, int maxdim,
                           int maxiter, double tol, const char* which, unsigned seed, double* ev, double* evec,
                           double* res, long* iters) {
  return guarded([&] {
    auto r = orc::block_krylov_schur(csr(n, rp, ci, v), nev, block, maxdim, maxiter, tol, which, seed);
    return finish(r, nev, n, ev, evec, res, iters);
  });
}

int orc_lanczos(int n, const int* rp, const int* ci, const double* v, int nev, int maxiter, double tol,
                const char* which, unsigned seed, double* ev, double* evec, double* res, long* iters) {
  return guarded([&] {
    auto r = orc::lanczos(csr(n, rp, ci, v), nev, maxiter, tol, which, seed);
    return finish(r, nev, n, ev, evec, res, iters);
  });
}

int orc_block_lanczos(int n, const int* rp, const int* ci, const double* v, int nev, int block, int maxiter,
                      double tol, const char* which, unsigned seed, double* ev, double* evec, double* res,
                      long* iters) {
  return guarded([&] {
    auto r = orc::block_lanczos(csr(n, rp, ci, v), nev, block, maxiter, tol, which, seed);
    return finish(r, nev, n, ev, evec, res, iters);
  });
}

int orc_krylov_schur(int n, const int* rp, const int* ci, const double* v, int nev, int maxdim, int maxiter,
                     double tol, const char* which, unsigned seed, double* ev, double* evec, double* res,
                     long* iters) {
  return guarded([&] {
    auto r = orc::krylov_schur(csr(n, rp, ci, v), nev, maxdim, maxiter, tol, which, seed);
    return finish(r, nev, n, ev, evec, res, iters);
  });
}

void orc_residual_norms(int n, const int* rp, const int* ci, const double* v, int k, const double* evals,
                        const double* X, double* out) {
  orc::Vec ev(evals, evals + k);
  auto r = orc::residual_norms(csr(n, rp, ci, v), ev, wrap(X, n, k));
  std::copy(r.begin(), r.end(), out);
}

}  // extern "C"
