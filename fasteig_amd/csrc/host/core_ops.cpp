// Device-resident orthogonalization, Rayleigh-Ritz extraction and the Lanczos
// bound estimate (see core_ops.hpp for the reference anchors).
#include "core_ops.hpp"

#include <algorithm>
#include <cmath>
#include <limits>
#include <numeric>
#include <stdexcept>

namespace fasteig {
namespace hip {

Workspace::Workspace(HipExecutor& e, int max_m, int gram) : ex(e), gram_mode(gram) { ensure(std::max(max_m, 1)); }

void Workspace::ensure(int m) {
  if (m <= cap) return;
  cap = std::max(m, cap + cap / 2);
  G = Small(ex, cap, cap);
  R = Small(ex, cap, cap);
  V = Small(ex, cap, cap);
  C = Small(ex, cap, cap);
  W = Small(ex, cap, 1);
  theta = Small(ex, cap, 1);
  rn = Small(ex, cap, 1);
  yn = Small(ex, cap, 1);
}

static void scale_columns(HipExecutor& ex, const BlockRef& X, const double* s_dev) {
  check(fe_axpby_cols(ex.ctx(), X.prec, X.rows, X.cols, X.ptr, X.ld, X.ptr, X.ld, 0.0, 0.0, s_dev),
        "fe_axpby_cols(scale)");
}

int cholesky_qr(Workspace& ws, Block& X, Block& tmp) {
  HipExecutor& ex = ws.ex;
  const int k = X.ncols();
  if (k <= 0 || X.rows() < k) throw std::invalid_argument("thin_qr_in_place: row count must be >= column count.");
  ws.ensure(k);
  gram(ex, ws.gram_mode, X.ref(), X.ref(), ws.G.data(), ws.cap, nullptr, ws.need);
  const int info = fe_potrf(ex.ctx(), k, ws.G.data(), ws.cap);
  check(info, "fe_potrf");
  if (info != 0) return -1;
  check(fe_chol_rinv(ex.ctx(), k, ws.G.data(), ws.cap, ws.R.data(), ws.cap), "fe_chol_rinv");
  if (tmp.ncols() == k && tmp.rows() == X.rows() && tmp.prec() == X.prec()) {
    trmm_upper(ex, X.ref(), ws.R.data(), ws.cap, tmp.ref(), ws.gram_mode, ws.need);
    X.swap(tmp);
  } else {
    const BlockRef t = tmp.cols(0, k);
    if (tmp.ncols() < k || tmp.rows() != X.rows() || tmp.prec() != X.prec())
      throw std::invalid_argument("cholesky_qr: scratch block too small");
    trmm_upper(ex, X.ref(), ws.R.data(), ws.cap, t, ws.gram_mode, ws.need);
    copy(ex, t, X.ref());
  }
  return 0;
}

int orthonormal_qr(Workspace& ws, Block& X, Block& tmp) {
  HipExecutor& ex = ws.ex;
  const int k = X.ncols();
  int shifted = 0;
  if (cholesky_qr(ws, X, tmp) != 0) {
    // shifted Cholesky QR: G + s I = R^T R with s = 11 (n k + k (k+1)) u ||X||_2^2
    shifted = 1;
    ws.ensure(k);
    gram(ex, ws.gram_mode, X.ref(), X.ref(), ws.G.data(), ws.cap, nullptr, ws.need);
    MatrixXd Gh = ws.G.to_host();
    double tr = 0.0;
    for (int i = 0; i < k; ++i) tr += Gh(i, i);
    const double u = X.prec() == FE_F64 ? std::numeric_limits<double>::epsilon() * 0.5
                                        : std::numeric_limits<float>::epsilon() * 0.5;
    const double s = 11.0 * (static_cast<double>(X.rows()) * k + static_cast<double>(k) * (k + 1)) * u * tr;
    for (int i = 0; i < k; ++i) Gh(i, i) += s;
    ws.G.from_host(Gh);
    const int info = fe_potrf(ex.ctx(), k, ws.G.data(), ws.cap);
    check(info, "fe_potrf(shifted)");
    if (info != 0) throw std::runtime_error("orthonormal_qr: shifted Cholesky QR failed (non-finite block?)");
    check(fe_chol_rinv(ex.ctx(), k, ws.G.data(), ws.cap, ws.R.data(), ws.cap), "fe_chol_rinv");
    const BlockRef t = tmp.cols(0, k);
    trmm_upper(ex, X.ref(), ws.R.data(), ws.cap, t, ws.gram_mode, ws.need);
    copy(ex, t, X.ref());
    if (cholesky_qr(ws, X, tmp) != 0) throw std::runtime_error("orthonormal_qr: Cholesky QR failed after shift");
  }
  if (cholesky_qr(ws, X, tmp) != 0) throw std::runtime_error("orthonormal_qr: second Cholesky QR pass failed");
  return shifted;
}

void block_gram_schmidt(Workspace& ws, const BlockRef& basis, const BlockRef& S, const BlockRef& out) {
  HipExecutor& ex = ws.ex;
  if (out.ptr != S.ptr) copy(ex, S, out);
  const int a = basis.cols, b = S.cols;
  if (a == 0 || b == 0) return;
  ws.ensure(std::max(a, b));
  gram(ex, ws.gram_mode, basis, out, ws.G.data(), ws.cap, nullptr, ws.need);
  gemm(ex, basis, ws.G.data(), ws.cap, out, -1.0, 1.0);
  gram(ex, ws.gram_mode, basis, out, ws.R.data(), ws.cap, nullptr, ws.need);
  gemm(ex, basis, ws.R.data(), ws.cap, out, -1.0, 1.0);
}

namespace {
// Directions of `cur` (already projected off the basis) whose singular values
// exceed `cut` (absolute), as an orthonormal block, at most `capacity`
// columns, largest singular value first: Gram G = cur^T cur, G = V diag(w) V^T,
// Q = cur V_keep diag(w_keep)^(-1/2). The Gram eigenvalues carry an absolute
// error of a few eps * w_max, so only w > resolve * w_max are trusted here;
// *resolved receives the singular value below which this level cannot judge.
Block svqb_level(Workspace& ws, const BlockRef& cur, double cut, int capacity, double resolve, double* resolved) {
  HipExecutor& ex = ws.ex;
  const int c = cur.cols;
  ws.ensure(c);
  gram(ex, ws.gram_mode, cur, cur, ws.G.data(), ws.cap, nullptr, ws.need);
  check(fe_syev(ex.ctx(), c, ws.G.data(), ws.cap, ws.W.data(), ws.V.data(), ws.cap), "fe_syev");
  std::vector<double> w(c);
  check(fe_memcpy_d2h(ex.ctx(), w.data(), ws.W.data(), sizeof(double) * c), "d2h");
  const double wmax = std::max(w.back(), 0.0);
  *resolved = std::sqrt(resolve * wmax);
  std::vector<int> good;
  for (int i = c - 1; i >= 0; --i)  // descending singular values, like the SVD ordering
    if (w[i] > resolve * wmax && w[i] > cut * cut && static_cast<int>(good.size()) < capacity) good.push_back(i);
  if (good.empty()) return Block();
  const int r = static_cast<int>(good.size());
  MatrixXd Vh = ws.V.to_host();
  MatrixXd Cm(ws.cap, r);
  for (int j = 0; j < r; ++j) {
    const double f = 1.0 / std::sqrt(w[good[j]]);
    for (int i = 0; i < c; ++i) Cm(i, j) = Vh(i, good[j]) * f;
  }
  check(fe_memcpy_h2d(ex.ctx(), ws.C.data(), Cm.data(), sizeof(double) * Cm.size()), "h2d");
  Block q(ex, cur.rows, r, cur.prec);
  gemm(ex, cur, ws.C.data(), ws.cap, q.ref(), 1.0, 0.0, ws.gram_mode, ws.need);
  return q;
}
}  // namespace

Block orthonormalize_block(Workspace& ws, const BlockRef& basis, const BlockRef& src, double rank_tol) {
  HipExecutor& ex = ws.ex;
  if (src.cols <= 0 || src.rows < src.cols) throw std::invalid_argument("orthonormalize_block: invalid source block.");
  if (!std::isfinite(rank_tol) || rank_tol <= 0.0 || rank_tol >= 1.0)
    throw std::invalid_argument("orthonormalize_block: rank tolerance must be in (0,1).");
  if (basis.cols == basis.rows) return Block();
  // Gram eigenvalues trusted per level: sqrt of it is the singular-value ratio
  // one level resolves (1e-6 in FP64, 3e-3 in FP32)
  const double resolve = src.prec == FE_F64 ? 1e-12 : 1e-5;
  Block cur(ex, src.rows, src.cols, src.prec);
  copy(ex, src, cur.ref());
  for (int pass = 0; pass < 2; ++pass) {
    const std::vector<double> n0 = column_norms(ex, cur.ref());
    double scale = 0.0;
    for (double v : n0) {
      if (!std::isfinite(v)) throw std::runtime_error("orthonormalize_block: non-finite column norms.");
      scale = std::max(scale, v);
    }
    if (scale == 0.0) return Block();
    const double cut = rank_tol * scale;  // the reference keeps sigma(R) > rank_tol * max column norm
    block_gram_schmidt(ws, basis, cur.ref(), cur.ref());
    const int capacity = static_cast<int>(std::min<int64_t>(cur.ncols(), src.rows - basis.cols));
    if (capacity <= 0) return Block();
    // level 1: the directions the Gram of the projected block resolves
    double resolved = 0.0;
    Block q1 = svqb_level(ws, cur.ref(), cut, capacity, resolve, &resolved);
    Block next;
    const int r1 = q1.empty() ? 0 : q1.ncols();
    if (r1 < capacity && resolved > cut) {
      // level 2: singular values between cut and what level 1 resolved. Remove
      // the level-1 directions (and the basis again) from the projected block;
      // the remainder holds only the small singular values, which its own Gram
      // resolves relative to its much smaller scale.
      Block rest(ex, src.rows, cur.ncols(), src.prec);
      copy(ex, cur.ref(), rest.ref());
      if (r1 > 0) block_gram_schmidt(ws, q1.ref(), rest.ref(), rest.ref());
      block_gram_schmidt(ws, basis, rest.ref(), rest.ref());
      double resolved2 = 0.0;
      Block q2 = svqb_level(ws, rest.ref(), cut, capacity - r1, resolve, &resolved2);
      if (!q2.empty() && r1 > 0) {
        next = Block(ex, src.rows, r1 + q2.ncols(), src.prec);
        copy(ex, q1.ref(), next.cols(0, r1));
        copy(ex, q2.ref(), next.cols(r1, q2.ncols()));
      } else if (!q2.empty()) {
        next = std::move(q2);
      } else {
        next = std::move(q1);
      }
    } else {
      next = std::move(q1);
    }
    if (next.empty()) return Block();
    cur = std::move(next);
  }
  return cur;
}

std::vector<int> select_indices(const std::vector<double>& v, const std::string& which) {
  const int m = static_cast<int>(v.size());
  std::vector<int> idx(m);
  std::iota(idx.begin(), idx.end(), 0);
  if (which == "sr") std::stable_sort(idx.begin(), idx.end(), [&](int a, int b) { return v[a] < v[b]; });
  else if (which == "lr") std::stable_sort(idx.begin(), idx.end(), [&](int a, int b) { return v[a] > v[b]; });
  else if (which == "sm")
    std::stable_sort(idx.begin(), idx.end(), [&](int a, int b) { return std::fabs(v[a]) < std::fabs(v[b]); });
  else if (which == "lm")
    std::stable_sort(idx.begin(), idx.end(), [&](int a, int b) { return std::fabs(v[a]) > std::fabs(v[b]); });
  else if (which == "be") {
    int pos = 0;
    for (int i = 0; i < m / 2; ++i) {
      idx[pos++] = m - 1 - i;
      idx[pos++] = i;
    }
    if (m % 2 == 1) idx[pos] = m / 2;
  } else {
    throw std::invalid_argument("invalid which option (use SR/LR/SM/LM/BE).");
  }
  return idx;
}

void rotate_rest(Workspace& ws, int m, const BlockRef& Q, int from, const BlockRef& Y) {
  if (from >= Y.cols) return;
  gemm(ws.ex, Q.range(0, m), ws.C.data() + static_cast<int64_t>(from) * ws.cap, ws.cap, Y.range(from, Y.cols - from),
       1.0, 0.0, ws.gram_mode, ws.need);
}

RitzPairs extract_ritz(Workspace& ws, int m, const BlockRef& Q, const BlockRef& AQ, int requested,
                       const std::string& which, const BlockRef& Y, const BlockRef& AY, const BlockRef* R,
                       int res_count, bool keep_ay, int rotate_count, const double* rinv) {
  HipExecutor& ex = ws.ex;
  if (m <= 0 || requested <= 0 || requested > m)
    throw std::invalid_argument("extract_projected_ritz_pairs: invalid active/requested Ritz dimensions.");
  const int nr = res_count < 0 ? requested : res_count;
  if (nr <= 0 || nr > requested || Q.cols < m || AQ.cols < m || Y.cols != requested || AY.cols != nr)
    throw std::invalid_argument("extract_projected_ritz_pairs: multivector dimension mismatch.");
  ws.ensure(m);
  check(fe_sym_small(ex.ctx(), m, ws.G.data(), ws.cap), "fe_sym_small");
  if (rinv) {  // rinv^T G rinv (ws.V holds G rinv in between)
    check(fe_gemm_small(ex.ctx(), 0, 0, m, m, m, 1.0, ws.G.data(), ws.cap, rinv, ws.cap, 0.0, ws.V.data(), ws.cap),
          "fe_gemm_small");
    check(fe_gemm_small(ex.ctx(), 1, 0, m, m, m, 1.0, rinv, ws.cap, ws.V.data(), ws.cap, 0.0, ws.G.data(), ws.cap),
          "fe_gemm_small");
    check(fe_sym_small(ex.ctx(), m, ws.G.data(), ws.cap), "fe_sym_small");
  }
  const int info = fe_syev(ex.ctx(), m, ws.G.data(), ws.cap, ws.W.data(), ws.V.data(), ws.cap);
  check(info, "fe_syev");
  if (info != 0) throw std::runtime_error("extract_projected_ritz_pairs: projected eigensolve failed.");
  std::vector<double> w(m);
  check(fe_memcpy_d2h(ex.ctx(), w.data(), ws.W.data(), sizeof(double) * m), "d2h(ritz values)");
  std::vector<int> idx = select_indices(w, which);
  idx.resize(requested);
  const double* vecs = ws.V.data();
  if (rinv) {  // rinv times every eigenvector (ws.G is free after the eigensolve), then the selection
    check(fe_gemm_small(ex.ctx(), 0, 0, m, m, m, 1.0, rinv, ws.cap, ws.V.data(), ws.cap, 0.0, ws.G.data(), ws.cap),
          "fe_gemm_small");
    vecs = ws.G.data();
  }
  check(fe_gather_cols(ex.ctx(), m, requested, vecs, ws.cap, idx.data(), ws.C.data(), ws.cap), "fe_gather_cols");
  RitzPairs out;
  out.selected = idx;
  out.values.resize(requested);
  for (int j = 0; j < requested; ++j) out.values[j] = w[idx[j]];
  check(fe_memcpy_h2d(ex.ctx(), ws.theta.data(), out.values.data(), sizeof(double) * nr), "h2d(theta)");
  // Y = Q C (the first rotate_count columns now; rotate_rest does the others),
  // then AY = (AQ) C for C's first nr columns with R = AY - Y Theta and the
  // column norms in the same product's epilogue
  const int now = rotate_count < 0 ? requested : std::max(nr, std::min(rotate_count, requested));
  gemm(ex, Q.range(0, m), ws.C.data(), ws.cap, Y.range(0, now), 1.0, 0.0, ws.gram_mode, ws.need);
  const BlockRef Yr = Y.range(0, nr);
  gemm_resid(ex, AQ.range(0, m), ws.C.data(), ws.cap, AY, keep_ay, Yr, ws.theta.data(), R, ws.rn.data(),
             ws.yn.data(), ws.gram_mode, ws.need);
  std::vector<double> rn(nr), yn(nr);
  check(fe_memcpy_d2h(ex.ctx(), rn.data(), ws.rn.data(), sizeof(double) * nr), "d2h");
  check(fe_memcpy_d2h(ex.ctx(), yn.data(), ws.yn.data(), sizeof(double) * nr), "d2h");
  out.relres.resize(nr);
  for (int j = 0; j < nr; ++j) {
    const double den = (1.0 + std::fabs(out.values[j])) * yn[j];
    out.relres[j] = den > 0.0 ? rn[j] / den : rn[j];
  }
  return out;
}

void lanczos_bounds(HipExecutor& ex, const HipCsrMatrix& A, double sign, unsigned seed, double& lower,
                    double& upper) {
  const int64_t n = A.rows();
  const int steps = static_cast<int>(std::min<int64_t>(20, n));
  Workspace ws(ex, std::max(steps, 2), FE_GRAM_NATIVE);
  Block r(ex, n, 1, FE_F64), av(ex, n, 1, FE_F64), basis(ex, n, steps, FE_F64), xb(ex, n, 1, FE_F64);
  const BlockRef xv{xb.data(), n, 1, 1, FE_F64};  // unit-stride vector
  copy(ex, ex.random_block(n, 1, seed).ref(), r.ref());
  MatrixXd T(steps, steps);
  const double eps = std::numeric_limits<double>::epsilon();
  double beta_next = 0.0;
  // two-pass classical Gram-Schmidt of src against basis[:, :active] into res;
  // returns the coefficient of the last basis column (the Lanczos alpha)
  auto two_pass = [&](int active, const BlockRef& src, const BlockRef& res) -> double {
    if (src.ptr != res.ptr) copy(ex, src, res);
    if (active == 0) return 0.0;
    const BlockRef Vb = basis.cols(0, active);
    gram(ex, FE_GRAM_NATIVE, Vb, res, ws.G.data(), ws.cap);
    gemm(ex, Vb, ws.G.data(), ws.cap, res, -1.0, 1.0);
    gram(ex, FE_GRAM_NATIVE, Vb, res, ws.R.data(), ws.cap);
    gemm(ex, Vb, ws.R.data(), ws.cap, res, -1.0, 1.0);
    double h = 0.0, s = 0.0;
    check(fe_memcpy_d2h(ex.ctx(), &h, ws.G.data() + (active - 1), sizeof(double)), "d2h");
    check(fe_memcpy_d2h(ex.ctx(), &s, ws.R.data() + (active - 1), sizeof(double)), "d2h");
    return h + s;
  };
  for (int it = 0; it < steps; ++it) {
    double beta = column_norms(ex, r.ref())[0];
    if (!std::isfinite(beta)) throw std::runtime_error("core::KrylovSubspace: non-finite residual norm encountered.");
    bool restarted = false;
    if (!(beta > eps)) {
      const Block& cand = ex.random_block(n, 1, seed + static_cast<unsigned>(it + 1));
      two_pass(it, cand.ref(), r.ref());
      restarted = true;
      beta = column_norms(ex, r.ref())[0];
      if (beta <= eps) throw std::runtime_error("core::KrylovSubspace: breakdown, unable to generate new basis vector.");
    }
    check(fe_axpby_cols(ex.ctx(), FE_F64, n, 1, r.data(), r.ld(), r.data(), r.ld(), 1.0 / beta, 0.0, nullptr),
          "scale");
    copy(ex, r.ref(), basis.cols(it, 1));
    // the product reads the contiguous copy: its gathers then touch neighbouring
    // entries instead of one entry per basis row
    copy(ex, r.ref(), xv);
    spmm(ex, A, xv, av.ref(), sign);
    const double alpha = two_pass(it + 1, av.ref(), r.ref());
    beta_next = column_norms(ex, r.ref())[0];
    if (!std::isfinite(beta_next))
      throw std::runtime_error("core::KrylovSubspace: non-finite next residual norm encountered.");
    if (it != 0) {
      const double b = restarted ? 0.0 : beta;
      T(it, it - 1) = b;
      T(it - 1, it) = b;
    }
    T(it, it) = alpha;
  }
  MatrixXd Tp(ws.cap, ws.cap);
  for (int j = 0; j < steps; ++j)
    for (int i = 0; i < steps; ++i) Tp(i, j) = T(i, j);
  ws.G.from_host(Tp);
  const int info = fe_syev(ex.ctx(), steps, ws.G.data(), ws.cap, ws.W.data(), ws.V.data(), ws.cap);
  check(info, "fe_syev(lanczos)");
  if (info != 0) throw std::runtime_error("ChebyshevSolver: Lanczos snapshot eigensolve failed.");
  std::vector<double> w(steps);
  check(fe_memcpy_d2h(ex.ctx(), w.data(), ws.W.data(), sizeof(double) * steps), "d2h");
  MatrixXd Vh = ws.V.to_host();
  lower = std::numeric_limits<double>::infinity();
  upper = -std::numeric_limits<double>::infinity();
  for (int j = 0; j < steps; ++j) {
    const double rad = std::fabs(beta_next * Vh(steps - 1, j));
    lower = std::min(lower, w[j] - rad);
    upper = std::max(upper, w[j] + rad);
  }
  if (!std::isfinite(lower) || !std::isfinite(upper) || lower > upper)
    throw std::runtime_error("ChebyshevSolver: invalid spectral bound estimates.");
}

}  // namespace hip
}  // namespace fasteig
