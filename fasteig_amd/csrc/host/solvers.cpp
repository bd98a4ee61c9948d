// Chebyshev, LOBPCG and Davidson on the HIP backend. The control flow of each
// solver follows the reference source line by line (cited per step); every
// n-sized operation runs on the device through libfe_hip.so.
#include <algorithm>
#include <numeric>
#include <cmath>
#include <iostream>
#include <limits>
#include <stdexcept>

#include "core_ops.hpp"
#include "fasteig_problem.hpp"
#include "fasteig_solver.hpp"

namespace fasteig {

using hip::Block;
using hip::BlockRef;
using hip::HipCsrMatrix;
using hip::HipExecutor;
using hip::Workspace;

namespace {
constexpr int kMaxProjected = 2048;  // largest projected matrix fe_syev takes
std::string lower_copy(std::string s) {
  std::transform(s.begin(), s.end(), s.begin(), [](unsigned char c) { return static_cast<char>(std::tolower(c)); });
  return s;
}
// s starts with the first n characters of prefix (src/fasteig_solver.cpp:21-23)
bool has_prefix(const std::string& s, const std::string& prefix, size_t n) {
  return s.size() >= n && s.compare(0, n, prefix, 0, n) == 0;
}
double max_head(const std::vector<double>& v, int count) {
  double m = 0.0;
  for (int j = 0; j < count && j < static_cast<int>(v.size()); ++j) m = std::max(m, v[j]);
  return m;
}
bool all_finite(const std::vector<double>& v) {
  for (double x : v)
    if (!std::isfinite(x)) return false;
  return true;
}
// Per-iteration progress (print_every > 0, verbosity >= 2): the largest
// relative residual of the wanted pairs and the block precision; verbosity >= 3
// adds every wanted value and residual.
void progress(const fasteig_options& o, long long iter, const std::vector<double>& lam, const std::vector<double>& rr,
              int count, double sign, int prec = FE_F64) {
  if (o.print_every() <= 0 || o.verbosity() < 2 || iter % o.print_every() != 0) return;
  double mx = 0.0;
  for (int j = 0; j < count && j < static_cast<int>(rr.size()); ++j) mx = std::max(mx, rr[j]);
  std::cout << "  iter " << iter << " [" << (prec == FE_F32 ? "fp32" : "fp64") << "] max relres " << mx;
  if (o.verbosity() >= 3)
    for (int j = 0; j < count && j < static_cast<int>(lam.size()); ++j)
      std::cout << " " << sign * lam[j] << "(" << rr[j] << ")";
  std::cout << std::endl;
}
// Start block of `cols` columns in precision prec: the cached seeded block,
// with the first columns of options.v0 when provided (reference:
// src/ChebyshevSolver.cpp:95-103, src/LOBPCGSolver.cpp:48-56).
void start_block(HipExecutor& ex, const fasteig_options& o, int64_t n, int cols, int v0_max, unsigned seed,
                 const Block& dst, const char* who) {
  hip::copy(ex, ex.random_block(n, cols, seed).ref(), dst.ref());
  if (o.v0().size() > 0) {
    if (o.v0().rows() != n) throw std::invalid_argument(std::string(who) + ": v0 row count does not match the problem.");
    const int pc = std::min(v0_max, o.v0().cols());
    if (pc > 0) {
      MatrixXd head(static_cast<int>(n), pc);
      std::copy(o.v0().data(), o.v0().data() + n * pc, head.data());
      Block tmp(ex, n, pc, FE_F64);
      hip::upload(ex, head, tmp.ref());
      hip::copy(ex, tmp.ref(), dst.cols(0, pc));
    }
  }
}
}  // namespace

fasteig_solver::fasteig_solver(::fasteig_problem& problem, const fasteig_options& options)
    : problem_(problem), eigvals_(problem.eigenvalues()), resnrms_(problem.residual_norms()), options_(options) {}

ChebyshevSolver::ChebyshevSolver(::fasteig_problem& p, const fasteig_options& o) : fasteig_solver(p, o) {}
LOBPCGSolver::LOBPCGSolver(::fasteig_problem& p, const fasteig_options& o) : fasteig_solver(p, o) {}
DavidsonSolver::DavidsonSolver(::fasteig_problem& p, const fasteig_options& o) : fasteig_solver(p, o) {}

// ============================================================== Chebyshev
// Reference: src/ChebyshevSolver.cpp:67-221 (solve) and :223-258 (filter).
FASTEIG_INT ChebyshevSolver::solve() {
  options_.validate();
  const int64_t n = problem_.n();
  const int nev = options_.nev();
  const int maxiter = options_.maxiter();
  const int degree = options_.degree();
  const double tol = options_.tol();
  const std::string which = lower_copy(options_.which());
  if (n <= 0 || nev <= 0 || nev > n) throw std::invalid_argument("ChebyshevSolver: nev must be in [1,n].");
  if (which != "sr" && which != "lr") throw std::invalid_argument("ChebyshevSolver: only SR and LR are supported.");

  auto exp = problem_.executor(options_);
  HipExecutor& ex = *exp;
  HipCsrMatrix& A = problem_.device_matrix(options_);
  const double sign = which == "lr" ? -1.0 : 1.0;
  const unsigned seed = options_.seed().value_or(1u);
  const int capacity = options_.blksize() > nev ? static_cast<int>(std::min<int64_t>(options_.blksize(), n))
                                                : nev + static_cast<int>(std::min<int64_t>(nev, n - nev));
  const bool mixed = options_.mixed();
  int prec = mixed ? FE_F32 : FE_F64;
  const std::string& gopt = options_.gram();
  const int gram_mode = gopt == "bf16x3" ? FE_GRAM_BF16X3 : gopt == "bf16x6" ? FE_GRAM_BF16X6
                        : gopt == "auto" ? hip::kProductAuto : FE_GRAM_NATIVE;
  Workspace ws(ex, capacity, mixed ? gram_mode : FE_GRAM_NATIVE);

  // persistent blocks: V (subspace), F/P (filter recurrence), Q/AQ (basis), AY
  Block V(ex, n, capacity, prec), F(ex, n, capacity, prec), P(ex, n, capacity, prec);
  Block Q(ex, n, capacity, prec), AQ(ex, n, capacity, prec), AY(ex, n, capacity, prec);
  start_block(ex, options_, n, capacity, nev, seed, V, "ChebyshevSolver");
  // the random start block is well conditioned (kappa near 1 for n >> m) and
  // only spans the first filter's input: in the mixed path its Cholesky QR
  // takes the products of a first cycle (about 16 bits)
  ws.need = mixed ? 1e-4 : 0.0;
  if (hip::cholesky_qr(ws, V, Q) != 0) throw std::runtime_error("ChebyshevSolver: initial Cholesky QR failed.");
  ws.need = 0.0;

  double lower = 0.0, upper = 0.0;
  hip::lanczos_bounds(ex, A, sign, seed, lower, upper);

  const bool ritz_bounds = options_.cheb_bounds() == "ritz";
  int locked = 0;
  int subspace_cols = ritz_bounds ? capacity : nev;
  hip::RitzPairs ritz;
  long long it_done = 0;
  double prev_max = std::numeric_limits<double>::infinity();
  double last_mx = std::numeric_limits<double>::quiet_NaN();  // previous cycle's max residual
  double best_rate = 0.0;  // fastest residual decay seen, decades per filter step
  int cycle_degree = degree;  // filter degree of the coming cycle
  bool refine_wanted = false;  // FP64 cycles filter the wanted columns only
  bool closing = false;        // the coming cycle is the closing FP64 sweep of the mixed path
  std::vector<double> prev_relres;  // the previous FP32 cycle's residuals of the wanted pairs
  std::vector<int> col_deg;    // per-column degrees of the closing sweep (empty: cycle_degree for all)
  double cutoff = std::numeric_limits<double>::quiet_NaN();  // adaptive bound (ritz_bounds)
  hip::Small inv_norms(ex, capacity, 1);

  for (int iter = 0; iter < maxiter; ++iter) {
    // the FP64 cycles of the mixed path refine the wanted columns only when
    // the FP32 phase brought the whole block near its floor; the guard
    // columns then keep their FP32-phase vectors and stay in the Rayleigh-Ritz
    // subspace
    const int active = (mixed && prec == FE_F64 && refine_wanted) ? std::max(0, nev - locked) : subspace_cols - locked;
    // accuracy the FP32 products of this cycle need (automatic product mode):
    // while the previous cycle's residual is at least 1e-2 (or unknown) the
    // cycle cannot resolve errors of 1e-4 and compensated BF16x3 products
    // (about 16 bits) serve; below it FP32-class products (1e-6)
    ws.need = prec == FE_F32 ? ((std::isfinite(last_mx) && last_mx < 1e-2) ? 1e-6 : 1e-4) : 0.0;
    // 1. filter the active columns (ChebyshevSolver.cpp:151-155, 223-258)
    bool equil = false;
    if (active > 0 && upper != lower) {
      BlockRef prev = F.cols(0, active), cur = P.cols(0, active);
      // per-column degrees (the closing sweep): the columns are gathered in
      // order of decreasing degree, so step k filters the prefix of the
      // columns whose degree is at least k; all degrees have the parity of the
      // largest, so every column's last iterate ends in the same buffer
      const bool percol = static_cast<int>(col_deg.size()) == active && locked == 0;
      std::vector<int> perm, back;
      if (percol) {
        perm.resize(active);
        std::iota(perm.begin(), perm.end(), 0);
        std::stable_sort(perm.begin(), perm.end(), [&](int a, int b) { return col_deg[a] > col_deg[b]; });
        back.resize(active);
        for (int i = 0; i < active; ++i) back[perm[i]] = i;
      }
      auto width = [&](int k) {  // columns filtered at step k
        if (!percol) return active;
        int w = 0;
        while (w < active && col_deg[perm[w]] >= k) ++w;
        return w;
      };
      // the filter reads its start block in place when the active columns
      // start on a 16-byte boundary (the LDS-tiled product needs aligned
      // rows); otherwise from an aligned copy (or the gathered one)
      const BlockRef start = V.cols(locked, active);
      const bool in_place = !percol && (locked * static_cast<int64_t>(V.ref().esize())) % 16 == 0;
      if (percol) hip::gather_cols(ex, start, perm, prev);
      else if (!in_place) hip::copy(ex, start, prev);
      const BlockRef y0 = in_place ? start : prev;
      if (ritz_bounds && std::isfinite(cutoff) && cutoff > lower && cutoff < upper) {
        // scaled filter on [cutoff, upper] normalised at the lower bound
        // (Zhou & Saad 2007, Alg. 4.1): sigma_1 = e/(a0 - c),
        // Y_1 = (sigma_1/e)(A - cI) X, Y_{k+1} = (2 s_{k+1}/e)(A - cI) Y_k - s_k s_{k+1} Y_{k-1}
        const double e = 0.5 * (upper - cutoff), c = 0.5 * (upper + cutoff);
        double sigma = e / (lower - c);
        const double s1 = sigma;
        hip::spmm(ex, A, y0, cur, sign * s1 / e, -c * s1 / e);
        for (int k = 2; k <= cycle_degree; ++k) {
          const double sn = 1.0 / (2.0 / s1 - sigma);
          const int w = width(k);
          const BlockRef yw = y0.range(0, w), pw = prev.range(0, w);
          hip::spmm(ex, A, cur.range(0, w), pw, sign * 2.0 * sn / e, -c * 2.0 * sn / e, k == 2 ? &yw : &pw,
                    -sigma * sn);
          std::swap(prev, cur);
          sigma = sn;
        }
      } else {
        const double xlb = std::cosh(std::acosh(1.0e3) / degree);
        const double cut = (2.0 * lower + upper * (xlb - 1.0)) / (1.0 + xlb);
        const double c = 2.0 / (upper - cut);
        const double d = -(upper + cut) / (upper - cut);
        if (!std::isfinite(c) || !std::isfinite(d)) throw std::runtime_error("ChebyshevSolver: invalid filter scaling.");
        hip::spmm(ex, A, y0, cur, sign * c, d);                         // Y1 = (cA + dI) Y0
        for (int k = 2; k <= degree; ++k) {                            // Y_k = 2(cA+dI)Y_{k-1} - Y_{k-2}
          hip::spmm(ex, A, cur, prev, 2.0 * sign * c, 2.0 * d, k == 2 ? &y0 : &prev, -1.0);
          std::swap(prev, cur);
        }
      }
      // 2. deflate against the locked Ritz vectors (ChebyshevSolver.cpp:159-166)
      if (locked > 0) hip::block_gram_schmidt(ws, V.cols(0, locked), cur, cur);
      // FP32 blocks and the closing sweep: the first CholQR pass factors the
      // block with its columns scaled to unit norm, so that its Gram is not
      // limited by their norm spread. An FP32 block is scaled in place: the
      // filter leaves the damped columns so small that their BF16 parts and
      // FP32 part products leave the normal range. An FP64 block keeps its
      // columns; the scaling then acts on the m x m Gram and factor only.
      if (prec == FE_F32) {
        const std::vector<double> nr = hip::column_norms(ex, cur);
        std::vector<double> inv(active);
        for (int j = 0; j < active; ++j) inv[j] = nr[j] > 0.0 ? 1.0 / nr[j] : 1.0;
        hip::check(fe_memcpy_h2d(ex.ctx(), inv_norms.data(), inv.data(), sizeof(double) * active), "h2d");
        hip::check(fe_axpby_cols(ex.ctx(), prec, n, active, cur.ptr, cur.ld, cur.ptr, cur.ld, 0.0, 0.0,
                                 inv_norms.data()),
                   "scale");
      } else {
        equil = closing || cycle_degree > degree;
      }
      if (percol) {
        hip::gather_cols(ex, cur, back, V.cols(locked, active));  // back to the Ritz order
      } else if (locked == 0 && active == V.ncols() && (cur.ptr == F.data() || cur.ptr == P.data())) {
        V.swap(cur.ptr == F.data() ? F : P);  // the filtered block becomes V, the old V scratch
      } else {
        hip::copy(ex, cur, V.cols(locked, active));
      }
    }
    // CholQR of V[:, :subspace_cols] into Q (ChebyshevSolver.cpp:167-171); the
    // FP32 phase runs it twice (CholQR2) for orthogonality at FP32 accuracy
    // (and CholQR2 after the long closing sweep); a failed first factorization
    // is retried with a shift (shifted CholQR, Fukaya et al. 2020) and two
    // more passes
    int passes = (prec == FE_F32 || closing || cycle_degree > degree) ? 2 : 1;
    const double cycle_need = ws.need;
    // The last pass's R^{-1} goes into the Rayleigh-Ritz below instead of a
    // tall product Q R^{-1}: the projected matrix becomes R^{-T} (Q^T A Q)
    // R^{-1} and the coefficients R^{-1} C (small FP64 products), the same
    // Ritz pairs of the same span. A first pass folds only when it covers all
    // of V, which then trades places with Q (the basis moves to Q, V receives
    // the Ritz vectors).
    bool folded = false;
    for (int pass = 0; pass < passes; ++pass) {
      ws.ensure(subspace_cols);
      // the first pass factors the filtered block, whose condition number
      // kappa sets the accuracy it needs (a Gram error e fails the Cholesky
      // once kappa^2 e ~ 1, and X R^{-1} moves the span by about e kappa):
      // FP32-class products; later passes work on a nearly orthonormal block
      // and take the cycle's accuracy
      ws.need = pass == 0 ? std::min(cycle_need, 1e-6) : cycle_need;
      // pass 0 takes V into Q; a later pass takes Q into AY (scratch here)
      // and swaps the two
      const BlockRef Vb = pass == 0 ? V.cols(0, subspace_cols) : Q.cols(0, subspace_cols);
      const BlockRef Qout = pass == 0 ? Q.cols(0, subspace_cols) : AY.cols(0, subspace_cols);
      hip::gram(ex, ws.gram_mode, Vb, Vb, ws.G.data(), ws.cap, nullptr, ws.need);
      // Gram of V D (D = diag(1 / ||v_j||), kept in inv_norms)
      const bool scaled = pass == 0 && equil;
      if (scaled) hip::check(fe_gram_equilibrate(ex.ctx(), subspace_cols, ws.G.data(), ws.cap, inv_norms.data()),
                             "fe_gram_equilibrate");
      if (pass == 0 && passes == 2) {
        // The equilibrated block's Gram G = I + E: kappa(V)^2 <= (1 + |E|)/(1 - |E|).
        // One pass leaves Q^T Q - I at about the Gram's error times kappa^2:
        // with FP32-class products, for |E|_F <= 1/3 (kappa^2 <= 2) what the
        // second pass would reach with the same products; with FP64 products,
        // for |E|_F <= 0.9 (kappa^2 <= 19) about 1e-14. Then the second pass is
        // skipped (the closing sweep, whose filtered columns stay nearly
        // orthogonal: |E|_F = 0.19 on C3).
        double dev = 0.0;
        hip::check(fe_identity_deviation(ex.ctx(), subspace_cols, ws.G.data(), ws.cap, &dev), "fe_identity_deviation");
        if (dev <= (prec == FE_F64 ? 0.9 : 1.0 / 3.0)) passes = 1;
        if (options_.print_every() > 0 && options_.verbosity() >= 2)
          std::cout << "  CholQR: |G - I|_F = " << dev << ", " << passes << " pass(es)" << std::endl;
      }
      int info = fe_potrf(ex.ctx(), subspace_cols, ws.G.data(), ws.cap);
      hip::check(info, "fe_potrf");
      if (info != 0 && passes < 4) {
        hip::gram(ex, ws.gram_mode, Vb, Vb, ws.G.data(), ws.cap, nullptr, ws.need);
        if (scaled) hip::check(fe_gram_equilibrate(ex.ctx(), subspace_cols, ws.G.data(), ws.cap, inv_norms.data()),
                               "fe_gram_equilibrate");
        MatrixXd Gh = ws.G.to_host();
        double tr = 0.0;
        for (int i = 0; i < subspace_cols; ++i) tr += Gh(i, i);
        const double u = prec == FE_F64 ? 0.5 * std::numeric_limits<double>::epsilon()
                                        : 0.5 * std::numeric_limits<float>::epsilon();
        const double sh = 11.0 * (static_cast<double>(n) * subspace_cols +
                                  static_cast<double>(subspace_cols) * (subspace_cols + 1)) * u * tr;
        for (int i = 0; i < subspace_cols; ++i) Gh(i, i) += sh;
        ws.G.from_host(Gh);
        info = fe_potrf(ex.ctx(), subspace_cols, ws.G.data(), ws.cap);
        hip::check(info, "fe_potrf");
        passes = std::min(4, pass + 3);
      }
      if (info != 0)
        throw std::runtime_error("ChebyshevSolver: Cholesky QR failed during filtering (cycle " +
                                 std::to_string(iter + 1) + ", pass " + std::to_string(pass + 1) + ", " +
                                 (prec == FE_F32 ? "fp32" : "fp64") + ", degree " + std::to_string(cycle_degree) +
                                 ", column " + std::to_string(info) + ").");
      hip::check(fe_chol_rinv(ex.ctx(), subspace_cols, ws.G.data(), ws.cap, ws.R.data(), ws.cap), "fe_chol_rinv");
      // V D R^{-1} = V (D R^{-1}): the factor of the unscaled block
      if (scaled)
        hip::check(fe_scale_rows_small(ex.ctx(), subspace_cols, subspace_cols, inv_norms.data(), ws.R.data(), ws.cap),
                   "fe_scale_rows_small");
      if (pass + 1 == passes && (pass > 0 || subspace_cols == V.ncols())) {
        if (pass == 0) V.swap(Q);
        folded = true;
        break;
      }
      hip::trmm_upper(ex, Vb, ws.R.data(), ws.cap, Qout, ws.gram_mode, ws.need);
      if (pass > 0) Q.swap(AY);
    }
    ws.need = cycle_need;
    // 3. Rayleigh-Ritz (ChebyshevSolver.cpp:175-183); Ritz vectors go straight into V
    hip::spmm(ex, A, Q.cols(0, subspace_cols), AQ.cols(0, subspace_cols), sign);
    // Q^T (A Q) is symmetric: its upper tiles only
    hip::gram(ex, ws.gram_mode, Q.cols(0, subspace_cols), AQ.cols(0, subspace_cols), ws.G.data(), ws.cap, nullptr,
              ws.need, true);
    // residuals of the wanted pairs only: AQ C for their nev columns, fused
    // with the residual norms (AY itself is not kept); the guard columns of V
    // are rotated once the solve is known to go on (a converged last cycle
    // returns only the wanted vectors)
    const int m_rr = subspace_cols;
    ritz = hip::extract_ritz(ws, m_rr, Q.ref(), AQ.ref(), m_rr, "sr", V.cols(0, m_rr), AY.cols(0, nev), nullptr, nev,
                             false, nev, folded ? ws.R.data() : nullptr);
    if (!all_finite(ritz.values) || !all_finite(ritz.relres))
      throw std::runtime_error("ChebyshevSolver: non-finite Ritz values or residuals.");
    it_done = iter + 1;
    cutoff = ritz.values[subspace_cols - 1];
    const double mx = max_head(ritz.relres, nev);
    progress(options_, it_done, ritz.values, ritz.relres, nev, sign, prec);

    if (prec == FE_F32) {
      hip::rotate_rest(ws, m_rr, Q.ref(), nev, V.cols(0, m_rr));
      // Mixed precision: FP32 cycles until the FP32 residuals stall or reach the
      // FP32 floor; then the subspace is promoted to FP64 and the remaining
      // cycles (filter, CholQR, Rayleigh-Ritz) run in FP64.
      // A slow decay far above the FP32 floor (relative residuals of 1e-6 to
      // 1e-5) is the filter's convergence rate, not a precision limit: above
      // 1e-4 the FP32 cycles continue while the residual still falls by 1%
      // per cycle; below it a cycle must halve the residual.
      const bool floor_reached = mx <= std::max(tol, 2e-5);
      const bool stalled = iter >= 2 && mx > (mx <= 1e-4 ? 0.5 : 0.99) * prev_max;
      prev_max = std::min(prev_max, mx);
      if (std::isfinite(last_mx) && last_mx > mx) best_rate = std::max(best_rate, std::log10(last_mx / mx) / degree);
      if (floor_reached || stalled || iter + 1 == maxiter) {
        // The closing FP64 sweep: its filter degree is sized from the fastest
        // residual decay per filter step of the FP32 cycles (a stall reflects
        // the FP32 floor, not the filter), to bring the residual from mx to
        // a quarter of the tolerance in one cycle (between degree and 2 x degree;
        // further FP64 cycles follow at the base degree if needed).
        if (ritz_bounds && best_rate > 0.0 && mx > tol) {
          const double need = std::log10(mx / (0.25 * tol)) / best_rate;
          cycle_degree = std::clamp(static_cast<int>(std::ceil(need)), degree, 2 * degree);
        }
        prec = FE_F64;
        refine_wanted = mx <= 1e-4;
        closing = true;
        // Per-column degrees: column j needs log10(r_j / (tol / 4)) / rate_j
        // steps, rate_j its own decay over the last FP32 cycle but at least the
        // fastest decay of the largest residual (so no column gets more steps
        // than the uniform degree above), at least 2; the columns far below the
        // cutoff converge many times faster than the ones next to it.
        col_deg.clear();
        if (ritz_bounds && best_rate > 0.0 && mx > tol && refine_wanted &&
            static_cast<int>(prev_relres.size()) == nev) {
          col_deg.assign(nev, 2);
          int D = 2;
          for (int j = 0; j < nev; ++j) {
            const double r = ritz.relres[j], rp = prev_relres[j];
            double rate = best_rate;
            if (r > 0.0 && rp > r) rate = std::max(rate, std::log10(rp / r) / degree);
            const int d = r > 0.25 * tol ? static_cast<int>(std::ceil(std::log10(r / (0.25 * tol)) / rate)) : 2;
            col_deg[j] = std::clamp(d, 2, cycle_degree);
            D = std::max(D, col_deg[j]);
          }
          for (int& d : col_deg) d += (D - d) & 1;  // the parity of the largest
          cycle_degree = D;
          if (options_.print_every() > 0 && options_.verbosity() >= 2) {
            double mean = 0.0;
            for (int d : col_deg) mean += d;
            std::cout << "  closing sweep: column degrees mean " << mean / nev << ", max " << D << std::endl;
          }
        }
        Block V64(ex, n, capacity, FE_F64);
        hip::copy(ex, V.ref(), V64.ref());
        V = std::move(V64);
        F = Block(ex, n, capacity, FE_F64);
        P = Block(ex, n, capacity, FE_F64);
        Q = Block(ex, n, capacity, FE_F64);
        AQ = Block(ex, n, capacity, FE_F64);
        AY = Block(ex, n, capacity, FE_F64);
        ws.gram_mode = FE_GRAM_NATIVE;
        locked = 0;
        if (iter + 1 == maxiter) --iter;  // always finish with an FP64 cycle
      }
      last_mx = mx;
      prev_relres.assign(ritz.relres.begin(), ritz.relres.begin() + nev);
      continue;  // no locking on FP32 residuals
    }
    cycle_degree = degree;
    closing = false;
    col_deg.clear();

    // 4. ordered soft locking (ChebyshevSolver.cpp:188-192)
    locked = 0;
    while (locked < nev && ritz.relres[locked] <= tol) ++locked;
    subspace_cols = std::max(subspace_cols, nev + std::min(locked, capacity - nev));
    if (mx <= tol) break;  // (ChebyshevSolver.cpp:197-207)
    if (iter + 1 < maxiter) hip::rotate_rest(ws, m_rr, Q.ref(), nev, V.cols(0, m_rr));
  }

  eigvals_.resize(nev);
  resnrms_.resize(nev);
  for (int j = 0; j < nev; ++j) {
    eigvals_(j) = sign * ritz.values[j];
    resnrms_(j) = ritz.relres[j];
  }
  Block X(ex, n, nev, FE_F64);
  hip::copy(ex, V.cols(0, nev), X.ref());
  problem_.set_device_eigenvectors(std::move(X));
  set_iterations(it_done);
  return status_;
}

// ============================================================== LOBPCG
// Reference: src/LOBPCGSolver.cpp:25-173.
FASTEIG_INT LOBPCGSolver::solve() {
  options_.validate();
  const int64_t n = problem_.n();
  const int nev = options_.nev();
  const int maxiter = options_.maxiter();
  const double tol = options_.tol();
  const std::string which = lower_copy(options_.which());
  if (n <= 0 || nev <= 0 || nev > n || maxiter <= 0)
    throw std::invalid_argument("LOBPCGSolver: require 0 < nev <= n and maxiter > 0.");
  if (which != "sr" && which != "lr") throw std::invalid_argument("LOBPCGSolver: only SR and LR are supported.");
  auto exp = problem_.executor(options_);
  HipExecutor& ex = *exp;
  HipCsrMatrix& A = problem_.device_matrix(options_);
  const double sign = which == "lr" ? -1.0 : 1.0;
  const unsigned seed = options_.seed().value_or(1u);
  const int bw = static_cast<int>(std::min<int64_t>(std::max(options_.blksize(), nev), n));
  const int capacity = static_cast<int>(std::min<int64_t>(n, 3LL * bw));
  const int prec = FE_F64;
  Workspace ws(ex, capacity, FE_GRAM_NATIVE);

  Block X(ex, n, bw, prec), AX(ex, n, bw, prec), Res(ex, n, bw, prec), tmp(ex, n, bw, prec);
  Block basis(ex, n, capacity, prec), abasis(ex, n, capacity, prec), P(ex, n, bw, prec);
  start_block(ex, options_, n, bw, nev, seed, X, "LOBPCGSolver");
  hip::orthonormal_qr(ws, X, tmp);
  // initial Rayleigh-Ritz (LOBPCGSolver.cpp:72-78)
  hip::copy(ex, X.ref(), basis.cols(0, bw));
  hip::spmm(ex, A, basis.cols(0, bw), abasis.cols(0, bw), sign);
  hip::gram(ex, FE_GRAM_NATIVE, basis.cols(0, bw), abasis.cols(0, bw), ws.G.data(), ws.cap, nullptr, 0.0, true);
  const BlockRef res0 = Res.ref();
  hip::RitzPairs ritz = hip::extract_ritz(ws, bw, basis.ref(), abasis.ref(), bw, "sr", X.ref(), AX.ref(), &res0);
  bool haveP = false;
  long long it_done = 0;
  for (int iter = 0; iter < maxiter; ++iter) {
    if (!all_finite(ritz.values) || !all_finite(ritz.relres))
      throw std::runtime_error("LOBPCGSolver: non-finite Ritz values or residuals.");
    if (max_head(ritz.relres, nev) <= tol) break;
    // soft locking: residual directions of unconverged pairs (LOBPCGSolver.cpp:95-103)
    std::vector<int> act;
    for (int j = 0; j < bw; ++j)
      if (ritz.relres[j] > tol) act.push_back(j);
    const int na = static_cast<int>(act.size());
    Block Wsrc(ex, n, na, prec);
    hip::gather_cols(ex, Res.ref(), act, Wsrc.ref());
    // basis [X, orth(W), orth(P)] (LOBPCGSolver.cpp:105-129)
    hip::orthonormal_qr(ws, X, tmp);
    hip::copy(ex, X.ref(), basis.cols(0, bw));
    int sc = bw;
    Block W = hip::orthonormalize_block(ws, basis.cols(0, sc), Wsrc.ref());
    if (!W.empty()) {
      const int add = std::min(W.ncols(), capacity - sc);
      hip::copy(ex, W.cols(0, add), basis.cols(sc, add));
      sc += add;
    }
    if (haveP && sc < capacity) {
      Block Psrc(ex, n, na, prec);
      hip::gather_cols(ex, P.ref(), act, Psrc.ref());
      Block Po = hip::orthonormalize_block(ws, basis.cols(0, sc), Psrc.ref());
      if (!Po.empty()) {
        const int add = std::min(Po.ncols(), capacity - sc);
        hip::copy(ex, Po.cols(0, add), basis.cols(sc, add));
        sc += add;
      }
    }
    // Rayleigh-Ritz in S (LOBPCGSolver.cpp:131-141)
    hip::spmm(ex, A, basis.cols(0, sc), abasis.cols(0, sc), sign);
    ws.ensure(sc);
    hip::gram(ex, FE_GRAM_NATIVE, basis.cols(0, sc), abasis.cols(0, sc), ws.G.data(), ws.cap, nullptr, 0.0, true);
    const BlockRef rr = Res.ref();
    ritz = hip::extract_ritz(ws, sc, basis.ref(), abasis.ref(), bw, "sr", X.ref(), AX.ref(), &rr);
    if (!all_finite(ritz.values) || !all_finite(ritz.relres))
      throw std::runtime_error("LOBPCGSolver: non-finite Ritz values or residuals.");
    // P = S C with the X rows of C zeroed (LOBPCGSolver.cpp:143-150)
    hip::check(fe_zero_rows_small(ex.ctx(), bw, bw, ws.C.data(), ws.cap), "fe_zero_rows_small");
    hip::gemm(ex, basis.cols(0, sc), ws.C.data(), ws.cap, P.ref(), 1.0, 0.0);
    haveP = true;
    it_done = iter + 1;
    progress(options_, it_done, ritz.values, ritz.relres, nev, sign);
    const bool converged = max_head(ritz.relres, nev) <= tol;
    const bool exhausted = sc == bw;
    if (converged || exhausted) break;
  }
  eigvals_.resize(nev);
  resnrms_.resize(nev);
  for (int j = 0; j < nev; ++j) {
    eigvals_(j) = sign * ritz.values[j];
    resnrms_(j) = ritz.relres[j];
  }
  Block Xo(ex, n, nev, FE_F64);
  hip::copy(ex, X.cols(0, nev), Xo.ref());
  problem_.set_device_eigenvectors(std::move(Xo));
  set_iterations(it_done);
  return status_;
}

// ============================================================== Davidson
// Reference: src/DavidsonSolver.cpp:27-264.
FASTEIG_INT DavidsonSolver::solve() {
  const int64_t n = problem_.n();
  const int nev = options_.nev();
  const int maxiter = options_.maxiter();
  const int maxblks = options_.maxblks();
  const double tol = options_.tol();
  const std::string which = lower_copy(options_.which());
  if (n <= 0) throw std::invalid_argument("DavidsonSolver: n must be > 0.");
  if (nev <= 0 || nev > n) throw std::invalid_argument("DavidsonSolver: nev must be in [1,n].");
  if (maxiter <= 0) throw std::invalid_argument("DavidsonSolver: maxiter must be > 0.");
  if (maxblks < 1) throw std::invalid_argument("DavidsonSolver: maxblks must be at least 1.");
  if (which == "sai") throw std::invalid_argument("DavidsonSolver: SAI is not supported yet; use SR/LR/SM/LM/BE.");
  if (which != "sr" && which != "lr" && which != "sm" && which != "lm" && which != "be")
    throw std::invalid_argument("DavidsonSolver: invalid which option (use SR/LR/SM/LM/BE).");
  auto exp = problem_.executor(options_);
  HipExecutor& ex = *exp;
  HipCsrMatrix& A = problem_.device_matrix(options_);
  const unsigned seed = options_.seed().value_or(1u);
  const int prec = FE_F64;
  const int maxcols = static_cast<int>(std::min<int64_t>(n, static_cast<int64_t>(maxblks) * nev));
  Workspace ws(ex, maxcols, FE_GRAM_NATIVE);
  hip::Small proj(ex, maxcols, maxcols);  // bordered projected matrix, column-major, ld maxcols

  Block V0(ex, n, nev, prec), tmp(ex, n, nev, prec);
  start_block(ex, options_, n, nev, nev, seed, V0, "DavidsonSolver");
  if (hip::cholesky_qr(ws, V0, tmp) != 0) throw std::runtime_error("DavidsonSolver: initial Cholesky QR failed.");
  Block basis(ex, n, maxcols, prec), abasis(ex, n, maxcols, prec);
  Block Y(ex, n, nev, prec), AY(ex, n, nev, prec), Res(ex, n, nev, prec);
  const BlockRef resref = Res.ref();

  auto load_projected = [&](int m) {  // proj[:m, :m] -> ws.G (ld cap)
    ws.ensure(m);
    hip::check(fe_copy2d(ex.ctx(), m, m, proj.data(), FE_F64, maxcols, ws.G.data(), FE_F64, ws.cap), "copy(proj)");
  };
  // initial projection (DavidsonSolver.cpp:91-102)
  hip::copy(ex, V0.ref(), basis.cols(0, nev));
  hip::spmm(ex, A, basis.cols(0, nev), abasis.cols(0, nev), 1.0);
  hip::gram(ex, FE_GRAM_NATIVE, basis.cols(0, nev), abasis.cols(0, nev), proj.data(), maxcols);
  load_projected(nev);
  hip::RitzPairs ritz = hip::extract_ritz(ws, nev, basis.ref(), abasis.ref(), nev, which, Y.ref(), AY.ref(), &resref);
  const double locktol = std::min(1.0e-5, tol);
  long long it_done = 0;
  for (int iter = 0; iter < maxiter; ++iter) {
    // thick restart (DavidsonSolver.cpp:131-155)
    hip::orthonormal_qr(ws, Y, tmp);
    hip::spmm(ex, A, Y.ref(), AY.ref(), 1.0);
    hip::copy(ex, Y.ref(), basis.cols(0, nev));
    hip::copy(ex, AY.ref(), abasis.cols(0, nev));
    hip::gram(ex, FE_GRAM_NATIVE, basis.cols(0, nev), abasis.cols(0, nev), proj.data(), maxcols);
    int act = nev;
    load_projected(nev);
    ritz = hip::extract_ritz(ws, nev, basis.ref(), abasis.ref(), nev, which, Y.ref(), AY.ref(), &resref);
    for (int blk = 1; blk < maxblks; ++blk) {
      std::vector<int> ai;
      for (int j = 0; j < nev; ++j)
        if (ritz.relres[j] > locktol) ai.push_back(j);
      const int cc = static_cast<int>(ai.size());
      if (cc == 0 || act + cc > maxcols) break;
      // correction block from raw residuals (DavidsonSolver.cpp:172-198)
      Block corr(ex, n, cc, prec), ctmp(ex, n, cc, prec), acorr(ex, n, cc, prec);
      hip::gather_cols(ex, Res.ref(), ai, corr.ref());
      hip::block_gram_schmidt(ws, basis.cols(0, act), corr.ref(), corr.ref());
      hip::orthonormal_qr(ws, corr, ctmp);
      hip::spmm(ex, A, corr.ref(), acorr.ref(), 1.0);
      // bordered projection (DavidsonSolver.cpp:204-220): [V, corr]^T A corr and its transpose
      double* col_act = proj.data() + static_cast<int64_t>(act) * maxcols;
      hip::copy(ex, corr.ref(), basis.cols(act, cc));
      hip::copy(ex, acorr.ref(), abasis.cols(act, cc));
      hip::gram(ex, FE_GRAM_NATIVE, basis.cols(0, act + cc), acorr.ref(), col_act, maxcols);
      hip::gram(ex, FE_GRAM_NATIVE, acorr.ref(), basis.cols(0, act), proj.data() + act, maxcols);
      act += cc;
      load_projected(act);
      ritz = hip::extract_ritz(ws, act, basis.ref(), abasis.ref(), nev, which, Y.ref(), AY.ref(), &resref);
    }
    it_done = iter + 1;
    progress(options_, it_done, ritz.values, ritz.relres, nev, 1.0);
    double mx = 0.0;
    for (double r : ritz.relres) mx = std::max(mx, r);
    if (!ritz.relres.empty() && mx < tol) break;  // strict (DavidsonSolver.cpp:236-238)
  }
  eigvals_.resize(nev);
  resnrms_.resize(nev);
  for (int j = 0; j < nev; ++j) {
    eigvals_(j) = ritz.values[j];
    resnrms_(j) = ritz.relres[j];
  }
  Block Xo(ex, n, nev, FE_F64);
  hip::copy(ex, Y.ref(), Xo.ref());
  problem_.set_device_eigenvectors(std::move(Xo));
  set_iterations(it_done);
  return status_;
}

// ============================================================== Block Krylov-Schur
// Reference: src/BlockKrylovSchurSolver.cpp:26-214, src/core/KrylovSubspace.cpp:261-433,
// src/core/Orthogonalization.cpp:609-715 (complete_block_qr),
// src/core/LanczosWorkspace.cpp:42-73,109-135, src/core/RitzExtraction.cpp:34-63,272-380.
BlockKrylovSchurSolver::BlockKrylovSchurSolver(::fasteig_problem& problem, const fasteig_options& options)
    : BlockKrylovBase(problem, options) {}
BlockLanczosSolver::BlockLanczosSolver(::fasteig_problem& problem, const fasteig_options& options)
    : BlockKrylovBase(problem, options) {}
FASTEIG_INT BlockKrylovSchurSolver::solve() { return run(true); }
FASTEIG_INT BlockLanczosSolver::solve() { return run(false); }

// Block Krylov-Schur (restarted) and block Lanczos (not restarted) share the
// basis expansion, the rank-aware block QR and the block Ritz estimates.
FASTEIG_INT BlockKrylovBase::run(bool restarted) {
  options_.validate();
  const int64_t n = problem_.n();
  const int nev = options_.nev();
  const int maxiter = options_.maxiter();
  const double tol = options_.tol();
  const std::string which = lower_copy(options_.which());
  const std::string who = restarted ? "BlockKrylovSchurSolver" : "BlockLanczosSolver";
  if (n <= 0) throw std::invalid_argument(who + ": n must be > 0.");
  const int b = static_cast<int>(options_.blksize());
  int block_count = 0, cycle_dim = 0, retained_blocks = 0, snap_count = nev;
  if (restarted) {
    if (nev <= 0 || nev > n) throw std::invalid_argument("BlockKrylovSchurSolver: nev must be in [1,n].");
    if (maxiter <= 0) throw std::invalid_argument("BlockKrylovSchurSolver: maxiter must be > 0.");
    if (which != "sr" && which != "lr" && which != "sm" && which != "lm" && which != "be")
      throw std::invalid_argument("BlockKrylovSchurSolver: which must be SR, LR, SM, LM or BE on the HIP backend.");
    if (b <= 0 || b > n) throw std::invalid_argument("BlockKrylovSchurSolver: blksize must be in [1,n].");
    const int64_t requested_dim = options_.maxdim() > 0 ? options_.maxdim() : std::max<int64_t>(2 * nev, 20);
    block_count = static_cast<int>(requested_dim >= n ? 1 + (n - 1) / b : requested_dim / b);
    cycle_dim = static_cast<int>(std::min<int64_t>(n, static_cast<int64_t>(block_count) * b));
    const int wanted_blocks = 1 + (nev - 1) / b;
    if (cycle_dim < n && block_count <= wanted_blocks)
      throw std::invalid_argument("BlockKrylovSchurSolver: maxdim must accommodate ceil(nev/blksize) retained blocks "
                                  "plus at least one expansion block (or the full matrix dimension).");
    retained_blocks = cycle_dim == n ? 0 : std::max(wanted_blocks, block_count / 2);
    snap_count = std::max(nev, retained_blocks * b);
  } else {
    // src/BlockLanczosSolver.cpp:33-49
    if (nev <= 0) throw std::invalid_argument("BlockLanczosSolver: nev must be > 0.");
    if (maxiter <= 0) throw std::invalid_argument("BlockLanczosSolver: maxiter must be > 0.");
    if (b <= 0) throw std::invalid_argument("BlockLanczosSolver: blksize must be > 0.");
    if (b > n || nev > n) throw std::invalid_argument("BlockLanczosSolver: blksize and nev must be <= n.");
    if (static_cast<int64_t>(maxiter) * b < nev)
      throw std::invalid_argument("BlockLanczosSolver: need maxiter*blksize >= nev.");
    if (which == "sai")
      throw std::invalid_argument("BlockLanczosSolver: SAI needs a sparse direct solver and is not part of the HIP "
                                  "backend.");
    if (which != "sr" && which != "lr" && which != "sm" && which != "lm" && which != "be")
      throw std::invalid_argument("BlockLanczosSolver: invalid which option (use SR/LR/SM/LM/BE).");
    // blocks that can be held: the whole space, maxiter blocks, or the
    // largest projected matrix the device eigensolver takes (2048)
    const int64_t span_blocks = 1 + (n - 1) / b;
    block_count = static_cast<int>(std::min<int64_t>({static_cast<int64_t>(maxiter), span_blocks,
                                                      static_cast<int64_t>(kMaxProjected / b)}));
    if (block_count < 1) throw std::invalid_argument("BlockLanczosSolver: blksize above 2048 is not supported.");
    cycle_dim = static_cast<int>(std::min<int64_t>(n, static_cast<int64_t>(block_count) * b));
  }
  const int pdim = (block_count + 1) * b;
  const unsigned seed = options_.seed().value_or(1u);

  auto exp = problem_.executor(options_);
  HipExecutor& ex = *exp;
  HipCsrMatrix& A = problem_.device_matrix(options_);
  Workspace ws(ex, pdim, FE_GRAM_NATIVE);
  Block Qb(ex, n, pdim, FE_F64), Z(ex, n, b, FE_F64), Rz(ex, n, b, FE_F64), Qn(ex, n, b, FE_F64);
  Block v(ex, n, 1, FE_F64), Sw(ex, n, b, FE_F64);
  hip::zero(ex, Qb.ref());
  MatrixXd T(pdim, pdim), beta(b, b);
  const double eps = std::numeric_limits<double>::epsilon();

  auto host_small = [&](const hip::Small& S, int rows, int cols) {
    if (rows > S.rows() || cols > S.cols()) throw std::logic_error("BlockKrylovSchurSolver: small matrix too small.");
    MatrixXd full = S.to_host();
    MatrixXd out(rows, cols);
    for (int j = 0; j < cols; ++j)
      for (int i = 0; i < rows; ++i) out(i, j) = full(i, j);
    return out;
  };
  auto frob = [](const MatrixXd& M) {
    double s = 0.0;
    for (int j = 0; j < M.cols(); ++j)
      for (int i = 0; i < M.rows(); ++i) s += M(i, j) * M(i, j);
    return std::sqrt(s);
  };
  auto vnorm = [&](const BlockRef& x) { return hip::column_norms(ex, x)[0]; };

  // rank-aware block QR of src against Qb[:, :active] into Qn (reference
  // complete_block_qr): R = Qn^T src with columns at or below the threshold
  // zeroed; returns whether random directions had to complete the basis.
  // Fast path: Cholesky QR twice on the column-equilibrated block, accepted
  // when every diagonal of R clears the threshold and Qn is orthonormal and
  // orthogonal to the basis to 1e-10. Otherwise Gram-Schmidt with column
  // pivoting (the reference's greedy choice of the longest projected column)
  // on a working copy deflated one accepted direction at a time; a direction
  // at or below max(threshold, 1e-12 * longest source column) is replaced by
  // a seeded random vector projected against the basis.
  auto complete_qr = [&](int active, const BlockRef& src, MatrixXd& R, unsigned qseed, double refnorm) {
    const int capacity = static_cast<int>(std::min<int64_t>(b, n - active));
    const double threshold = 64.0 * eps * refnorm;
    const std::vector<double> cn0 = hip::column_norms(ex, src);
    double original = 0.0;
    for (double c : cn0) original = std::max(original, c);
    auto finish_r = [&]() {
      hip::gram(ex, FE_GRAM_NATIVE, Qn.ref(), src, ws.G.data(), ws.cap);
      R = host_small(ws.G, b, b);
    };
    auto rinv_into_r = [&](const std::vector<double>* rowscale) {
      if (fe_potrf(ex.ctx(), b, ws.G.data(), ws.cap) != 0) return false;
      hip::check(fe_chol_rinv(ex.ctx(), b, ws.G.data(), ws.cap, ws.R.data(), ws.cap), "fe_chol_rinv");
      if (rowscale) {
        MatrixXd Ri = ws.R.to_host();
        for (int j = 0; j < b; ++j)
          for (int i = 0; i < b; ++i) Ri(i, j) /= (*rowscale)[i];
        ws.R.from_host(Ri);
      }
      return true;
    };
    bool finite = std::isfinite(original);
    for (double c : cn0) finite = finite && c > 0.0;
    if (capacity == b && finite) {
      hip::gram(ex, FE_GRAM_NATIVE, src, src, ws.G.data(), ws.cap);
      MatrixXd G = ws.G.to_host();
      for (int j = 0; j < b; ++j)
        for (int i = 0; i < b; ++i) G(i, j) /= cn0[i] * cn0[j];
      ws.G.from_host(G);
      bool ok = rinv_into_r(&cn0);
      if (ok) {
        hip::trmm_upper(ex, src, ws.R.data(), ws.cap, Sw.ref());
        hip::gram(ex, FE_GRAM_NATIVE, Sw.ref(), Sw.ref(), ws.G.data(), ws.cap);
        ok = rinv_into_r(nullptr);
      }
      if (ok) {
        hip::trmm_upper(ex, Sw.ref(), ws.R.data(), ws.cap, Qn.ref());
        hip::gram(ex, FE_GRAM_NATIVE, Qn.ref(), Qn.ref(), ws.C.data(), ws.cap);
        MatrixXd Gq = host_small(ws.C, b, b);
        for (int i = 0; i < b; ++i) Gq(i, i) -= 1.0;
        double cross = 0.0;
        if (active > 0) {
          hip::gram(ex, FE_GRAM_NATIVE, Qb.cols(0, active), Qn.ref(), ws.V.data(), ws.cap);
          cross = frob(host_small(ws.V, active, b));
        }
        finish_r();
        double dmin = std::numeric_limits<double>::infinity();
        bool rfin = true;
        for (int j = 0; j < b; ++j) {
          dmin = std::min(dmin, std::fabs(R(j, j)));
          for (int i = 0; i < b; ++i) rfin = rfin && std::isfinite(R(i, j));
        }
        const double ge = frob(Gq);
        if (rfin && dmin > threshold && std::isfinite(ge) && std::isfinite(cross) && ge < 1e-10 && cross < 1e-10) {
          for (int j = 0; j < b; ++j)
            for (int i = j + 1; i < b; ++i) R(i, j) = 0.0;
          return false;
        }
      }
    }
    bool replenished = false;
    hip::zero(ex, Qn.ref());
    hip::copy(ex, src, Sw.ref());
    if (active > 0) hip::block_gram_schmidt(ws, Qb.cols(0, active), Sw.ref(), Sw.ref());
    auto project = [&](int j) {
      if (active > 0) hip::block_gram_schmidt(ws, Qb.cols(0, active), v.ref(), v.ref());
      if (j > 0) hip::block_gram_schmidt(ws, Qn.cols(0, j), v.ref(), v.ref());
    };
    const double cutoff = std::max(threshold, 1e-12 * original);
    for (int j = 0; j < capacity; ++j) {
      const std::vector<double> cn = hip::column_norms(ex, Sw.ref());
      int best = 0;
      for (int c = 1; c < b; ++c)
        if (cn[c] > cn[best]) best = c;
      double length = 0.0;
      if (cn[best] > cutoff) {
        hip::copy(ex, Sw.ref().range(best, 1), v.ref());
        project(j);
        length = vnorm(v.ref());
        if (!(length > cutoff)) length = 0.0;
      }
      if (length == 0.0) {
        replenished = true;
        for (unsigned attempt = 0; attempt < 4; ++attempt) {
          hip::copy(ex, ex.random_block(n, 1, qseed + 17u * static_cast<unsigned>(j + 1) + attempt).ref(), v.ref());
          const double rn = vnorm(v.ref());
          project(j);
          length = vnorm(v.ref());
          if (length > 1e-12 * rn) break;
          length = 0.0;
        }
        if (length == 0.0) throw std::runtime_error("BlockKrylovSchurSolver: unable to complete the basis.");
      }
      const BlockRef qj = Qn.cols(j, 1);
      hip::check(fe_axpby_cols(ex.ctx(), FE_F64, n, 1, v.data(), v.ld(), qj.ptr, Qn.ld(), 1.0 / length, 0.0,
                               nullptr),
                 "scale");
      // deflate the working copy by the accepted direction
      hip::gram(ex, FE_GRAM_NATIVE, qj, Sw.ref(), ws.G.data(), ws.cap);
      hip::gemm(ex, qj, ws.G.data(), ws.cap, Sw.ref(), -1.0, 1.0);
    }
    finish_r();
    for (int j = 0; j < b; ++j) {
      double cn = 0.0;
      for (int i = 0; i < b; ++i) cn += R(i, j) * R(i, j);
      if (std::sqrt(cn) <= threshold)
        for (int i = 0; i < b; ++i) R(i, j) = 0.0;
    }
    return replenished;
  };

  // start block (Executor::create_random_multivector or v0) and its QR
  if (options_.v0().size() > 0) {
    if (options_.v0().rows() != n || options_.v0().cols() < b)
      throw std::invalid_argument("BlockKrylovSchurSolver: v0 must have n rows and at least blksize columns.");
    MatrixXd head(static_cast<int>(n), b);
    std::copy(options_.v0().data(), options_.v0().data() + n * b, head.data());
    hip::upload(ex, head, Rz.ref());
  } else {
    hip::copy(ex, ex.random_block(n, b, seed).ref(), Rz.ref());
  }
  {
    MatrixXd R0;
    complete_qr(0, Rz.ref(), R0, seed, 0.0);
    hip::copy(ex, Qn.ref(), Qb.cols(0, b));
  }

  auto expand_block = [&](int blk) {
    hip::spmm(ex, A, Qb.cols(blk * b, b), Z.ref(), 1.0);
    const int active = static_cast<int>(std::min<int64_t>(n, static_cast<int64_t>(blk + 1) * b));
    double scale = 0.0;
    for (double c : hip::column_norms(ex, Z.ref())) scale += c * c;
    scale = std::sqrt(scale);
    // two-pass block Gram-Schmidt keeping the coefficients H = H1 + H2
    hip::copy(ex, Z.ref(), Rz.ref());
    const BlockRef Vb = Qb.cols(0, active);
    hip::gram(ex, FE_GRAM_NATIVE, Vb, Rz.ref(), ws.G.data(), ws.cap);
    hip::gemm(ex, Vb, ws.G.data(), ws.cap, Rz.ref(), -1.0, 1.0);
    hip::gram(ex, FE_GRAM_NATIVE, Vb, Rz.ref(), ws.R.data(), ws.cap);
    hip::gemm(ex, Vb, ws.R.data(), ws.cap, Rz.ref(), -1.0, 1.0);
    const MatrixXd H1 = host_small(ws.G, active, b), H2 = host_small(ws.R, active, b);
    const int width = 1 + (active - 1) % b;
    MatrixXd D(b, b);
    for (int j = 0; j < width; ++j)
      for (int i = 0; i < width; ++i) D(i, j) = H1(active - width + i, j) + H2(active - width + i, j);
    MatrixXd B;
    const int nf = static_cast<int>(std::min<int64_t>(static_cast<int64_t>(blk + 1) * b, n));
    complete_qr(nf, Rz.ref(), B, seed + static_cast<unsigned>((blk + 1) * b), scale);
    hip::copy(ex, Qn.ref(), Qb.cols((blk + 1) * b, b));
    const int row = blk * b, nrow = row + b;
    for (int j = 0; j < b; ++j)
      for (int i = 0; i < b; ++i) {
        T(row + i, row + j) = D(i, j);
        T(nrow + i, row + j) = B(i, j);
        T(row + j, nrow + i) = B(i, j);
      }
    beta = B;
  };
  auto expand = [&](int first) {
    for (int blk = first; blk < block_count; ++blk) expand_block(blk);
  };
  Block X(ex, n, nev, FE_F64), AX(ex, n, nev, FE_F64), Rr(ex, n, nev, FE_F64);
  // block Ritz snapshot of T[:m, :m]: the `count` leading pairs in `which`
  // order with estimates ||beta y_last|| / (1 + |theta|)
  // (src/core/RitzExtraction.cpp:272-380)
  auto snapshot = [&](int m, int count, std::vector<double>& lam, std::vector<double>& rr, MatrixXd& Csel) {
    MatrixXd Tp(ws.cap, ws.cap);
    for (int j = 0; j < m; ++j)
      for (int i = 0; i < m; ++i) Tp(i, j) = 0.5 * (T(i, j) + T(j, i));
    ws.G.from_host(Tp);
    const int info = fe_syev(ex.ctx(), m, ws.G.data(), ws.cap, ws.W.data(), ws.V.data(), ws.cap);
    hip::check(info, "fe_syev(block Krylov)");
    if (info != 0) throw std::runtime_error(who + ": projected eigensolve failed.");
    std::vector<double> w(m);
    hip::check(fe_memcpy_d2h(ex.ctx(), w.data(), ws.W.data(), sizeof(double) * m), "d2h");
    const MatrixXd Vh = host_small(ws.V, m, m);
    const std::vector<int> idx = hip::select_indices(w, which);
    const int cnt = std::min(count, m);
    const int tail = 1 + (m - 1) % b;
    lam.assign(cnt, 0.0);
    rr.assign(cnt, 0.0);
    Csel.resize(m, cnt);
    for (int j = 0; j < cnt; ++j) {
      const int c = idx[j];
      lam[j] = w[c];
      double a2 = 0.0;
      for (int i = 0; i < b; ++i) {
        double sacc = 0.0;
        for (int t = 0; t < tail; ++t) sacc += beta(i, t) * Vh(m - tail + t, c);
        a2 += sacc * sacc;
      }
      rr[j] = std::sqrt(a2) / (1.0 + std::fabs(w[c]));
      for (int i = 0; i < m; ++i) Csel(i, j) = Vh(i, c);
    }
  };
  // Ritz vectors X = Q[:, :m] C[:, :nev] and true residuals into rr
  // (src/core/RitzExtraction.cpp:34-63); true when all meet tol
  auto finish = [&](int m, const MatrixXd& Csel, const std::vector<double>& lam, std::vector<double>& rr) {
    hip::Small Cs(ex, m, nev), th(ex, nev, 1), rn(ex, nev, 1), yn(ex, nev, 1);
    MatrixXd Ch(m, nev);
    for (int j = 0; j < nev; ++j)
      for (int i = 0; i < m; ++i) Ch(i, j) = Csel(i, j);
    Cs.from_host(Ch);
    hip::gemm(ex, Qb.cols(0, m), Cs.data(), m, X.ref(), 1.0, 0.0);
    hip::spmm(ex, A, X.ref(), AX.ref(), 1.0);
    hip::check(fe_memcpy_h2d(ex.ctx(), th.data(), lam.data(), sizeof(double) * nev), "h2d");
    hip::check(fe_ritz_residuals(ex.ctx(), FE_F64, n, nev, X.data(), X.ld(), AX.data(), AX.ld(), th.data(),
                                 Rr.data(), Rr.ld(), rn.data(), yn.data()),
               "fe_ritz_residuals");
    std::vector<double> rh(nev), yh(nev);
    hip::check(fe_memcpy_d2h(ex.ctx(), rh.data(), rn.data(), sizeof(double) * nev), "d2h");
    hip::check(fe_memcpy_d2h(ex.ctx(), yh.data(), yn.data(), sizeof(double) * nev), "d2h");
    bool ok = true;
    rr.assign(nev, 0.0);
    for (int j = 0; j < nev; ++j) {
      const double den = (1.0 + std::fabs(lam[j])) * yh[j];
      if (!(den > 0.0) || !std::isfinite(rh[j])) throw std::runtime_error(who + ": invalid vector or non-finite residual.");
      rr[j] = rh[j] / den;
      ok = ok && rr[j] <= tol;
    }
    return ok;
  };
  auto publish = [&](const std::vector<double>& lam, const std::vector<double>& rr, long long iterations,
                     int status) {
    eigvals_.resize(nev);
    resnrms_.resize(nev);
    for (int j = 0; j < nev; ++j) {
      eigvals_(j) = lam[j];
      resnrms_(j) = rr[j];
    }
    problem_.set_device_eigenvectors(std::move(X));
    status_ = status;
    set_iterations(iterations);
    return status_;
  };

  if (!restarted) {
    // block Lanczos: one block per iteration, checked once iter * b >= nev
    // (src/BlockLanczosSolver.cpp:95-170)
    const int64_t span_blocks = 1 + (n - 1) / b;
    const bool device_limited = block_count < maxiter && block_count < span_blocks;
    std::vector<double> lam, rr;
    MatrixXd Csel;
    int it_stop = block_count;
    bool converged = false;
    for (int iter = 1; iter <= block_count; ++iter) {
      expand_block(iter - 1);
      const int m = static_cast<int>(std::min<int64_t>(static_cast<int64_t>(iter) * b, n));
      snapshot(m, nev, lam, rr, Csel);
      progress(options_, iter, lam, rr, static_cast<int>(lam.size()), 1.0);
      if (static_cast<int64_t>(iter) * b >= nev && max_head(rr, static_cast<int>(rr.size())) < tol) {
        converged = true;
        it_stop = iter;
        break;
      }
    }
    if (!converged && device_limited)
      throw std::runtime_error("BlockLanczosSolver: projected dimension above 2048 is not supported on the HIP "
                               "backend.");
    const int m = static_cast<int>(std::min<int64_t>(static_cast<int64_t>(it_stop) * b, n));
    snapshot(m, nev, lam, rr, Csel);
    if (static_cast<int>(lam.size()) < nev) throw std::runtime_error("BlockLanczosSolver: subspace smaller than nev.");
    finish(m, Csel, lam, rr);
    return publish(lam, rr, it_stop, 0);
  }

  expand(0);
  const int m = cycle_dim;
  std::vector<double> lam, rr;
  MatrixXd Csel;
  long long it_done = 0;
  bool converged = false;
  Block U(ex, n, std::min(snap_count, m), FE_F64);
  for (int iter = 1; iter <= maxiter; ++iter) {
    snapshot(m, snap_count, lam, rr, Csel);
    const int cnt = static_cast<int>(lam.size());
    it_done = iter;
    progress(options_, it_done, lam, rr, nev, 1.0);
    converged = max_head(rr, nev) < tol;
    const bool at_limit = iter == maxiter || cycle_dim == n;
    if (converged || at_limit) {
      converged = finish(m, Csel, lam, rr);
      break;
    }
    // thick restart: Q <- [Q C_keep, Q_residual]; T <- [diag(theta) F^T; F *]
    MatrixXd coupling(b, cnt);
    for (int j = 0; j < cnt; ++j)
      for (int i = 0; i < b; ++i) {
        double sacc = 0.0;
        for (int t = 0; t < b; ++t) sacc += beta(i, t) * Csel(m - b + t, j);
        coupling(i, j) = sacc;
      }
    {
      hip::Small Cs(ex, m, cnt);
      Cs.from_host(Csel);
      hip::gemm(ex, Qb.cols(0, m), Cs.data(), m, U.cols(0, cnt), 1.0, 0.0);
      hip::copy(ex, Qb.cols(block_count * b, b), Rz.ref());
      hip::zero(ex, Qb.ref());
      hip::copy(ex, U.cols(0, cnt), Qb.cols(0, cnt));
      hip::copy(ex, Rz.ref(), Qb.cols(cnt, b));
    }
    T.setZero();
    for (int j = 0; j < cnt; ++j) T(j, j) = lam[j];
    for (int j = 0; j < cnt; ++j)
      for (int i = 0; i < b; ++i) {
        T(cnt + i, j) = coupling(i, j);
        T(j, cnt + i) = coupling(i, j);
      }
    expand(retained_blocks);
  }

  return publish(lam, rr, it_done, converged ? 0 : 1);
}

// ============================================================== scalar Krylov
namespace {

// Scalar Krylov basis on the device with full two-pass reorthogonalization.
// Reference: src/core/KrylovSubspace.cpp:124-260 (normalize_residual_into_basis
// with the random restart of normalize_or_restart / restart_residual,
// apply_operator_to_basis_column, orthogonalize_operator_result, residual_norm,
// rotate_basis) and src/core/MultiVector.cpp:153-187 (the two-pass projection
// returns the coefficient of the newest column).
class DeviceKrylov {
 public:
  DeviceKrylov(HipExecutor& ex, HipCsrMatrix& A, const fasteig_options& o, int cols, unsigned seed)
      : ex_(ex), A_(A), n_(A.rows()), seed_(seed), ws_(ex, cols, FE_GRAM_NATIVE), basis_(ex, A.rows(), cols, FE_F64),
        r_(ex, A.rows(), 1, FE_F64), av_(ex, A.rows(), 1, FE_F64), xb_(ex, A.rows(), 1, FE_F64),
        xv_{xb_.data(), A.rows(), 1, 1, FE_F64} {
    if (o.v0().size() > 0) {
      if (o.v0().rows() != n_) throw std::invalid_argument("v0 row count does not match the problem.");
      MatrixXd head(static_cast<int>(n_), 1);
      std::copy(o.v0().data(), o.v0().data() + n_, head.data());
      hip::upload(ex_, head, r_.ref());
    } else {
      hip::copy(ex_, ex_.random_block(n_, 1, seed_).ref(), r_.ref());
    }
  }
  Workspace& ws() { return ws_; }
  const Block& basis() const { return basis_; }
  double norm() {
    const double b = hip::column_norms(ex_, r_.ref())[0];
    if (!std::isfinite(b)) throw std::runtime_error("core::KrylovSubspace: non-finite residual norm encountered.");
    return b;
  }
  // q_col = r / ||r|| (a seeded random vector projected against q_0..q_col-1
  // when ||r|| <= eps); returns the coupling to record (0 after such a restart)
  double normalize_into(int col) {
    const double eps = std::numeric_limits<double>::epsilon();
    double beta = norm();
    bool restarted = false;
    if (!(beta > eps)) {
      project(col, ex_.random_block(n_, 1, seed_ + static_cast<unsigned>(col + 1)).ref(), r_.ref());
      restarted = true;
      beta = norm();
      if (beta <= eps) throw std::runtime_error("core::KrylovSubspace: breakdown, unable to generate new basis vector.");
    }
    hip::check(fe_axpby_cols(ex_.ctx(), FE_F64, n_, 1, r_.data(), r_.ld(), r_.data(), r_.ld(), 1.0 / beta, 0.0,
                             nullptr),
               "scale");
    hip::copy(ex_, r_.ref(), basis_.cols(col, 1));
    hip::copy(ex_, r_.ref(), xv_);  // unit-stride copy for the product
    return restarted ? 0.0 : beta;
  }
  // r = A q_col orthogonalized against q_0..q_col; returns alpha
  double step(int col) {
    hip::spmm(ex_, A_, xv_, av_.ref(), 1.0);
    return project(col + 1, av_.ref(), r_.ref());
  }
  // basis[:, :C.cols()] = basis[:, :active] C
  void rotate(int active, const MatrixXd& C) {
    hip::Small Cs(ex_, active, C.cols());
    Cs.from_host(C);
    Block U(ex_, n_, C.cols(), FE_F64);
    hip::gemm(ex_, basis_.cols(0, active), Cs.data(), active, U.ref(), 1.0, 0.0);
    hip::copy(ex_, U.ref(), basis_.cols(0, C.cols()));
  }
  // Ritz vectors X = basis[:, :active] C and their residual norms
  // ||A x - theta x|| / ((1 + |theta|) ||x||) (src/core/RitzExtraction.cpp:34-63)
  std::vector<double> ritz(int active, const MatrixXd& C, const std::vector<double>& theta, Block& X) {
    const int k = C.cols();
    hip::Small Cs(ex_, active, k), th(ex_, k, 1), rn(ex_, k, 1), yn(ex_, k, 1);
    Cs.from_host(C);
    X = Block(ex_, n_, k, FE_F64);
    Block AX(ex_, n_, k, FE_F64);
    hip::gemm(ex_, basis_.cols(0, active), Cs.data(), active, X.ref(), 1.0, 0.0);
    hip::spmm(ex_, A_, X.ref(), AX.ref(), 1.0);
    hip::check(fe_memcpy_h2d(ex_.ctx(), th.data(), theta.data(), sizeof(double) * k), "h2d");
    hip::check(fe_ritz_residuals(ex_.ctx(), FE_F64, n_, k, X.data(), X.ld(), AX.data(), AX.ld(), th.data(), nullptr,
                                 0, rn.data(), yn.data()),
               "fe_ritz_residuals");
    std::vector<double> rh(k), yh(k), out(k);
    hip::check(fe_memcpy_d2h(ex_.ctx(), rh.data(), rn.data(), sizeof(double) * k), "d2h");
    hip::check(fe_memcpy_d2h(ex_.ctx(), yh.data(), yn.data(), sizeof(double) * k), "d2h");
    for (int j = 0; j < k; ++j) {
      const double den = (1.0 + std::fabs(theta[j])) * yh[j];
      if (!(den > 0.0) || !std::isfinite(den) || !std::isfinite(rh[j]))
        throw std::runtime_error("evaluate_ritz_pairs: invalid vector or non-finite residual.");
      out[j] = rh[j] / den;
    }
    return out;
  }

 private:
  double project(int active, const BlockRef& src, const BlockRef& res) {
    if (src.ptr != res.ptr) hip::copy(ex_, src, res);
    if (active == 0) return 0.0;
    const BlockRef Vb = basis_.cols(0, active);
    hip::gram(ex_, FE_GRAM_NATIVE, Vb, res, ws_.G.data(), ws_.cap);
    hip::gemm(ex_, Vb, ws_.G.data(), ws_.cap, res, -1.0, 1.0);
    hip::gram(ex_, FE_GRAM_NATIVE, Vb, res, ws_.R.data(), ws_.cap);
    hip::gemm(ex_, Vb, ws_.R.data(), ws_.cap, res, -1.0, 1.0);
    double h = 0.0, s = 0.0;
    hip::check(fe_memcpy_d2h(ex_.ctx(), &h, ws_.G.data() + (active - 1), sizeof(double)), "d2h");
    hip::check(fe_memcpy_d2h(ex_.ctx(), &s, ws_.R.data() + (active - 1), sizeof(double)), "d2h");
    return h + s;
  }

  HipExecutor& ex_;
  HipCsrMatrix& A_;
  int64_t n_;
  unsigned seed_;
  Workspace ws_;
  Block basis_, r_, av_, xb_;
  BlockRef xv_;
};

// Projected tridiagonal (an arrowhead after a restart), kept on the host and
// diagonalized on the device. Reference: src/core/LanczosWorkspace.cpp:25-40
// (record_step), :97-108 (prepare_krylov_schur_restart),
// src/core/RitzExtraction.cpp:195-230,252-268 (snapshot estimates
// |beta_next y_last| / (1 + |theta|)), :470-492 (restart matrix).
struct Projection {
  MatrixXd T;
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
  // the `count` leading Ritz pairs of T[:m, :m] in `which` order
  void snapshot(HipExecutor& ex, Workspace& ws, int m, int count, const std::string& which, bool vectors,
                std::vector<double>& lam, std::vector<double>& rel, MatrixXd& Csel, const char* who) {
    if (m > kMaxProjected)
      throw std::runtime_error(std::string(who) + ": projected dimension above 2048 is not supported on the HIP backend.");
    MatrixXd Tp(ws.cap, ws.cap);
    for (int j = 0; j < m; ++j)
      for (int i = 0; i < m; ++i) Tp(i, j) = T(i, j);
    ws.G.from_host(Tp);
    const int info = fe_syev(ex.ctx(), m, ws.G.data(), ws.cap, ws.W.data(), ws.V.data(), ws.cap);
    hip::check(info, "fe_syev");
    if (info != 0) throw std::runtime_error(std::string(who) + ": projected eigensolve failed.");
    std::vector<double> w(m);
    hip::check(fe_memcpy_d2h(ex.ctx(), w.data(), ws.W.data(), sizeof(double) * m), "d2h");
    const MatrixXd V = ws.V.to_host();
    const std::vector<int> idx = hip::select_indices(w, which);
    const int cnt = std::min(count, m);
    lam.assign(cnt, 0.0);
    rel.assign(cnt, 0.0);
    if (vectors) Csel = MatrixXd(m, cnt);
    for (int j = 0; j < cnt; ++j) {
      const int c = idx[j];
      lam[j] = w[c];
      rel[j] = std::fabs(beta_next * V(m - 1, c)) / (1.0 + std::fabs(w[c]));
      if (vectors)
        for (int i = 0; i < m; ++i) Csel(i, j) = V(i, c);
    }
  }
  void restart(int retained, const std::vector<double>& lam, const MatrixXd& Csel) {
    const int m = T.rows();
    T = MatrixXd(m, m);
    for (int j = 0; j < retained; ++j) {
      T(j, j) = lam[j];
      const double coupling = beta_next * Csel(Csel.rows() - 1, j);
      T(retained, j) = coupling;
      T(j, retained) = coupling;
    }
    coupling_row = retained;
  }
};

void check_lanczos_which(const std::string& which, const char* who) {
  if (which == "sai")
    throw std::invalid_argument(std::string(who) +
                                ": SAI needs a sparse direct solver and is not part of the HIP backend.");
  if (which != "sr" && which != "lr" && which != "sm" && which != "lm" && which != "be")
    throw std::invalid_argument(std::string(who) + ": invalid which option (use SR/LR/SM/LM/BE).");
}
}  // namespace

LanczosSolver::LanczosSolver(::fasteig_problem& p, const fasteig_options& o) : fasteig_solver(p, o) {}
KrylovSchurSolver::KrylovSchurSolver(::fasteig_problem& p, const fasteig_options& o) : fasteig_solver(p, o) {}

// Reference: src/LanczosSolver.cpp:28-164.
FASTEIG_INT LanczosSolver::solve() {
  options_.validate();
  const int64_t n = problem_.n();
  const int nev = options_.nev();
  const int maxiter = options_.maxiter();
  const double tol = options_.tol();
  const std::string which = lower_copy(options_.which());
  if (maxiter <= 0) throw std::invalid_argument("LanczosSolver: maxiter must be > 0.");
  if (nev <= 0) throw std::invalid_argument("LanczosSolver: nev must be > 0.");
  if (maxiter < nev) throw std::invalid_argument("LanczosSolver: maxiter must be >= nev.");
  if (maxiter > n) throw std::invalid_argument("LanczosSolver: maxiter must be <= n.");
  check_lanczos_which(which, "LanczosSolver");
  auto exp = problem_.executor(options_);
  HipExecutor& ex = *exp;
  HipCsrMatrix& A = problem_.device_matrix(options_);
  const unsigned seed = options_.seed().value_or(1u);
  const int pe = static_cast<int>(options_.print_every());
  const int check_every = pe > 0 ? pe : 5;
  const int cap = std::min(maxiter, kMaxProjected);  // basis columns kept
  DeviceKrylov K(ex, A, options_, cap, seed);
  Projection P(cap);
  std::vector<double> lam, rel;
  MatrixXd Csel;
  int it_stop = maxiter;
  bool converged = false;
  for (int it = 0; it < maxiter; ++it) {
    if (it >= kMaxProjected)
      throw std::runtime_error("LanczosSolver: projected dimension above 2048 is not supported on the HIP backend.");
    const double beta = K.normalize_into(it);
    const double alpha = K.step(it);
    P.record(it, beta, alpha, K.norm());
    const int m = it + 1;
    const bool do_progress = pe > 0 && m % pe == 0;
    const bool do_check = m >= nev && (m % check_every == 0 || m == maxiter);
    if (!do_progress && !do_check) continue;
    P.snapshot(ex, K.ws(), m, nev, which, false, lam, rel, Csel, "LanczosSolver");
    progress(options_, m, lam, rel, static_cast<int>(lam.size()), 1.0);
    if (do_check && !rel.empty() && max_head(rel, static_cast<int>(rel.size())) < tol) {
      converged = true;
      it_stop = m;
      break;
    }
  }
  P.snapshot(ex, K.ws(), it_stop, nev, which, true, lam, rel, Csel, "LanczosSolver");
  Block X;
  const std::vector<double> rr = K.ritz(it_stop, Csel, lam, X);
  const int cnt = static_cast<int>(lam.size());
  eigvals_.resize(cnt);
  resnrms_.resize(cnt);
  for (int j = 0; j < cnt; ++j) {
    eigvals_(j) = lam[j];
    resnrms_(j) = rr[j];
  }
  (void)converged;
  problem_.set_device_eigenvectors(std::move(X));
  status_ = 0;
  set_iterations(it_stop);
  return status_;
}

// Reference: src/KrylovSchurSolver.cpp:28-167.
FASTEIG_INT KrylovSchurSolver::solve() {
  options_.validate();
  const int64_t n = problem_.n();
  const int nev = options_.nev();
  const int maxiter = options_.maxiter();
  const double tol = options_.tol();
  const std::string which = lower_copy(options_.which());
  if (n <= 0) throw std::invalid_argument("KrylovSchurSolver: n must be > 0.");
  if (nev <= 0 || nev > n) throw std::invalid_argument("KrylovSchurSolver: nev must be in [1,n].");
  if (maxiter <= 0) throw std::invalid_argument("KrylovSchurSolver: maxiter must be > 0.");
  check_lanczos_which(which, "KrylovSchurSolver");
  const int64_t want = std::max<int64_t>({2LL * nev, 20LL, static_cast<int64_t>(options_.maxdim())});
  const int m = static_cast<int>(std::min<int64_t>(want, n));
  if (m > kMaxProjected)
    throw std::invalid_argument("KrylovSchurSolver: subspace dimension above 2048 is not supported on the HIP backend.");
  auto exp = problem_.executor(options_);
  HipExecutor& ex = *exp;
  HipCsrMatrix& A = problem_.device_matrix(options_);
  const unsigned seed = options_.seed().value_or(1u);
  DeviceKrylov K(ex, A, options_, m, seed);
  Projection P(m);
  auto expand = [&](int begin) {
    for (int col = begin; col < m; ++col) {
      const double beta = K.normalize_into(col);
      const double alpha = K.step(col);
      P.record(col, beta, alpha, K.norm());
    }
  };
  expand(0);
  for (int iter = 1; iter <= maxiter; ++iter) {
    std::vector<double> lam, rel;
    MatrixXd Csel;
    P.snapshot(ex, K.ws(), m, m, which, true, lam, rel, Csel, "KrylovSchurSolver");
    int nconv = 0;
    for (int j = 0; j < nev; ++j)
      if (rel[j] * (1.0 + std::fabs(lam[j])) < tol * std::fabs(lam[j])) ++nconv;
    {
      // progress reports absolute estimates, like krylov_schur_progress_snapshot
      std::vector<double> ab(nev);
      for (int j = 0; j < nev; ++j) ab[j] = rel[j] * (1.0 + std::fabs(lam[j]));
      progress(options_, iter, lam, ab, nev, 1.0);
    }
    if (nconv >= nev || m == n || iter == maxiter) {
      MatrixXd Ch(m, nev);
      for (int j = 0; j < nev; ++j)
        for (int i = 0; i < m; ++i) Ch(i, j) = Csel(i, j);
      std::vector<double> head(lam.begin(), lam.begin() + nev);
      Block X;
      const std::vector<double> rr = K.ritz(m, Ch, head, X);
      eigvals_.resize(nev);
      resnrms_.resize(nev);
      for (int j = 0; j < nev; ++j) {
        eigvals_(j) = head[j];
        resnrms_(j) = rr[j];
      }
      problem_.set_device_eigenvectors(std::move(X));
      status_ = 0;
      set_iterations(iter);
      return status_;
    }
    const int spare = (m - nev) / 2;
    int retained = nev + std::min(nconv, spare);
    if (retained == 1 && m > 3) retained = m / 2;
    MatrixXd Ck(m, retained);
    for (int j = 0; j < retained; ++j)
      for (int i = 0; i < m; ++i) Ck(i, j) = Csel(i, j);
    K.rotate(m, Ck);
    P.restart(retained, lam, Csel);
    expand(retained);
  }
  throw std::logic_error("KrylovSchurSolver: unreachable solve termination.");
}

std::unique_ptr<fasteig_solver> make_fasteig_solver(const std::string& solver, ::fasteig_problem& problem,
                                                    const fasteig_options& options) {
  if (has_prefix(solver, "Lanczos", 3)) return std::make_unique<LanczosSolver>(problem, options);
  if (has_prefix(solver, "BlockLanczos", 8)) return std::make_unique<BlockLanczosSolver>(problem, options);
  if (has_prefix(solver, "BlockKrylovSchur", 11)) return std::make_unique<BlockKrylovSchurSolver>(problem, options);
  if (has_prefix(solver, "KrylovSchur", 7)) return std::make_unique<KrylovSchurSolver>(problem, options);
  if (has_prefix(solver, "Davidson", 3)) return std::make_unique<DavidsonSolver>(problem, options);
  if (has_prefix(solver, "LOBPCG", 3)) return std::make_unique<LOBPCGSolver>(problem, options);
  if (has_prefix(solver, "Chebyshev", 3)) return std::make_unique<ChebyshevSolver>(problem, options);
  if (has_prefix(solver, "auto", 3)) return nullptr;
  throw std::invalid_argument("solver '" + solver +
                              "' is not available on the HIP backend (Chebyshev, LOBPCG, Davidson, Lanczos, "
                              "BlockLanczos, KrylovSchur, BlockKrylovSchur)");
}

}  // namespace fasteig
