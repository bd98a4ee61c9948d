// fasteig_problem on the HIP backend (reference: src/fasteig_problem.cpp:81-437,
// 554-622 for the symmetric CSR expansion, :340-437 for solve()).
#include "fasteig_problem.hpp"

#include <algorithm>
#include <cmath>
#include <iomanip>
#include <iostream>
#include <limits>
#include <numeric>
#include <stdexcept>

using fasteig::hip::Block;
using fasteig::hip::HipExecutor;

fasteig_problem::fasteig_problem(FASTEIG_INT n, CsrMatrixData A, std::string solver)
    : n_(n), solver_(std::move(solver)), csrA_(std::move(A)) {
  if (n <= 0) throw std::invalid_argument("fasteig_problem: n must be positive.");
  if (static_cast<int>(csrA_.rowptr.size()) != n + 1 || csrA_.colind.size() != csrA_.values.size())
    throw std::invalid_argument("fasteig_problem: inconsistent CSR arrays.");
  if (csrA_.rowptr[0] != 0 || csrA_.rowptr[n] != static_cast<int>(csrA_.values.size()))
    throw std::invalid_argument("fasteig_problem: rowptr does not span the nonzeros.");
  for (int i = 0; i < n; ++i)
    if (csrA_.rowptr[i] > csrA_.rowptr[i + 1]) throw std::invalid_argument("fasteig_problem: rowptr not monotone.");
  for (int c : csrA_.colind)
    if (c < 0 || c >= n) throw std::invalid_argument("fasteig_problem: column index out of range.");
}

// Triplets -> full CSR with summed duplicates and sorted columns; a stored
// symmetric triangle is mirrored (reference: build_csr_from_sparse_matrix with
// expand_symmetric, src/fasteig_problem.cpp:554-622).
void fasteig_problem::build_from_triplets(FASTEIG_INT n, const std::vector<int>& r, const std::vector<int>& c,
                                          const std::vector<double>& v, bool sym, CsrMatrixData& out) {
  std::vector<int64_t> cnt(static_cast<size_t>(n) + 1, 0);
  for (size_t t = 0; t < v.size(); ++t) {
    if (r[t] < 0 || r[t] >= n || c[t] < 0 || c[t] >= n) throw std::invalid_argument("matrix index out of range.");
    cnt[r[t] + 1]++;
    if (sym && r[t] != c[t]) cnt[c[t] + 1]++;
  }
  for (int i = 0; i < n; ++i) cnt[i + 1] += cnt[i];
  const int64_t total = cnt[n];
  if (total > std::numeric_limits<int>::max()) throw std::invalid_argument("matrix has more than 2^31 nonzeros.");
  std::vector<int> ci(total);
  std::vector<double> vv(total);
  std::vector<int64_t> pos(cnt.begin(), cnt.end() - 1);
  for (size_t t = 0; t < v.size(); ++t) {
    ci[pos[r[t]]] = c[t];
    vv[pos[r[t]]++] = v[t];
    if (sym && r[t] != c[t]) {
      ci[pos[c[t]]] = r[t];
      vv[pos[c[t]]++] = v[t];
    }
  }
  out.rowptr.assign(static_cast<size_t>(n) + 1, 0);
  out.colind.clear();
  out.values.clear();
  out.colind.reserve(total);
  out.values.reserve(total);
  std::vector<int> order;
  for (int i = 0; i < n; ++i) {
    const int64_t b = cnt[i], e = cnt[i + 1];
    order.resize(static_cast<size_t>(e - b));
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return ci[b + x] < ci[b + y]; });
    for (size_t q = 0; q < order.size(); ++q) {
      const int col = ci[b + order[q]];
      const double val = vv[b + order[q]];
      if (!out.colind.empty() && static_cast<int>(out.colind.size()) > out.rowptr[i] && out.colind.back() == col)
        out.values.back() += val;
      else {
        out.colind.push_back(col);
        out.values.push_back(val);
      }
    }
    out.rowptr[i + 1] = static_cast<int>(out.colind.size());
  }
}

fasteig_problem::fasteig_problem(const fasteig_params& p) : n_(p.n), solver_(p.solver) {
  if (p.n <= 0) throw std::invalid_argument("fasteig_problem: deck has no matrix.");
  build_from_triplets(p.n, p.arowind, p.acolind, p.anzvals, p.mtx_symmetric, csrA_);
  if (!p.bnzvals.empty()) {
    // B must be diagonal for the HIP backend: congruence M^{-1/2} K M^{-1/2}
    std::vector<double> mass(static_cast<size_t>(p.n), 0.0);
    for (size_t t = 0; t < p.bnzvals.size(); ++t) {
      if (p.browind[t] != p.bcolind[t]) {
        if (p.bnzvals[t] != 0.0)
          throw std::invalid_argument(
              "fasteig_problem: the HIP backend supports generalized problems with a diagonal B only.");
        continue;
      }
      mass[p.browind[t]] += p.bnzvals[t];
    }
    set_mass_diagonal(std::move(mass));
  }
}

fasteig_problem::~fasteig_problem() {
  dev_evecs_ = Block();
  dA_.reset();
}

void fasteig_problem::set_mass_diagonal(std::vector<double> mass) {
  if (static_cast<int>(mass.size()) != n_) throw std::invalid_argument("set_mass_diagonal: size mismatch.");
  for (double m : mass)
    if (!(m > 0.0) || !std::isfinite(m)) throw std::invalid_argument("set_mass_diagonal: B must be positive definite.");
  mass_ = std::move(mass);
  dA_.reset();
  norm_est_.reset();
}

std::shared_ptr<HipExecutor> fasteig_problem::executor(const fasteig_options& opts) {
  if (!executor_) executor_ = std::make_shared<HipExecutor>(opts.device());
  return executor_;
}

void fasteig_problem::set_executor(std::shared_ptr<HipExecutor> exec) {
  if (!exec) throw std::invalid_argument("fasteig_problem::set_executor: null executor.");
  dev_evecs_ = Block();
  dA_.reset();
  executor_ = std::move(exec);
}

fasteig::hip::HipCsrMatrix& fasteig_problem::device_matrix(const fasteig_options& opts) {
  auto ex = executor(opts);
  if (!dA_) {
    dA_ = std::make_unique<fasteig::hip::HipCsrMatrix>(*ex, n_, csrA_.rowptr, csrA_.colind, csrA_.values);
    if (!mass_.empty()) {
      std::vector<double> d(mass_.size());
      for (size_t i = 0; i < d.size(); ++i) d[i] = 1.0 / std::sqrt(mass_[i]);
      fasteig::hip::Small dd(*ex, n_, 1);
      fasteig::hip::check(fe_memcpy_h2d(ex->ctx(), dd.data(), d.data(), sizeof(double) * d.size()), "h2d");
      fasteig::hip::check(fe_csr_scale_sym(ex->ctx(), dA_->handle(), dd.data()), "fe_csr_scale_sym");
      ex->synchronize();
    }
  }
  const std::string& f = opts.format();
  if (f == "sell") {
    dA_->use_sell(32, 256);
  } else if (f == "tiled") {
    if (!dA_->use_tiled())
      throw std::invalid_argument("format=tiled: the matrix has too little column locality for the tiled copy");
  } else if (f == "blocked" || f == "auto") {
    // auto: the LDS-tiled copy, else the row-block copy, else CSR
    if ((f == "blocked" || !dA_->use_tiled()) && !dA_->use_blocked()) {
      if (f == "blocked")
        throw std::invalid_argument("format=blocked: the matrix has too little column locality for the row-block copy");
      dA_->use_csr();
    }
  } else {
    dA_->use_csr();
  }
  return *dA_;
}

void fasteig_problem::set_device_eigenvectors(Block&& X) {
  dev_evecs_ = std::move(X);
  eigvecs_valid_ = false;
}

double fasteig_problem::norm_estimate(const fasteig_options& opts) {
  if (!norm_est_) {
    auto ex = executor(opts);
    double lo = 0.0, hi = 0.0;
    fasteig::hip::lanczos_bounds(*ex, device_matrix(opts), 1.0, 1u, lo, hi);
    norm_est_ = std::max(std::fabs(lo), std::fabs(hi));
  }
  return *norm_est_;
}

const MatrixXd& fasteig_problem::eigenvectors() {
  if (!eigvecs_valid_) {
    if (dev_evecs_.empty()) {
      eigvecs_.resize(n_, eigvals_.size());
    } else {
      eigvecs_ = fasteig::hip::download(*executor_, dev_evecs_.ref());
      if (!mass_.empty())  // u = M^{-1/2} y
        for (int j = 0; j < eigvecs_.cols(); ++j)
          for (int i = 0; i < n_; ++i) eigvecs_(i, j) /= std::sqrt(mass_[i]);
    }
    eigvecs_valid_ = true;
  }
  return eigvecs_;
}

// Final FP64 residuals on the device (the reference recomputes them on the
// host after the timed solve, src/fasteig_problem.cpp:419-422 and
// src/fasteig_solver.cpp:89-129). For a diagonal B the residual of the pencil
// is ||K u - lambda M u|| / ((1 + |lambda|) ||M u||) = ||M^{1/2} r|| /
// ((1 + |lambda|) ||M^{1/2} y||), r = A y - lambda y.
void fasteig_problem::finalize_residuals(const fasteig_options& opts) {
  auto ex = executor(opts);
  const int k = eigvals_.size();
  if (dev_evecs_.empty() || k == 0) return;
  auto& A = device_matrix(opts);
  const bool sell = A.sell();
  if (sell) A.use_csr();
  Block AX(*ex, n_, k, FE_F64), R(*ex, n_, k, FE_F64), Y(*ex, n_, k, FE_F64);
  fasteig::hip::copy(*ex, dev_evecs_.ref(), Y.ref());
  fasteig::hip::spmm(*ex, A, Y.ref(), AX.ref(), 1.0);
  fasteig::hip::Small th(*ex, k, 1), rn(*ex, k, 1), yn(*ex, k, 1);
  fasteig::hip::check(fe_memcpy_h2d(ex->ctx(), th.data(), eigvals_.data(), sizeof(double) * k), "h2d");
  fasteig::hip::check(fe_ritz_residuals(ex->ctx(), FE_F64, n_, k, Y.data(), Y.ld(), AX.data(), AX.ld(), th.data(),
                                        R.data(), R.ld(), rn.data(), yn.data()),
                      "fe_ritz_residuals");
  if (!mass_.empty()) {
    std::vector<double> s(mass_.size());
    for (size_t i = 0; i < s.size(); ++i) s[i] = std::sqrt(mass_[i]);
    fasteig::hip::Small sd(*ex, n_, 1);
    fasteig::hip::check(fe_memcpy_h2d(ex->ctx(), sd.data(), s.data(), sizeof(double) * s.size()), "h2d");
    fasteig::hip::check(fe_scale_rows(ex->ctx(), FE_F64, n_, k, R.data(), R.ld(), sd.data()), "scale");
    fasteig::hip::check(fe_scale_rows(ex->ctx(), FE_F64, n_, k, Y.data(), Y.ld(), sd.data()), "scale");
    fasteig::hip::check(fe_colnorms(ex->ctx(), FE_F64, n_, k, R.data(), R.ld(), rn.data()), "norms");
    fasteig::hip::check(fe_colnorms(ex->ctx(), FE_F64, n_, k, Y.data(), Y.ld(), yn.data()), "norms");
  }
  std::vector<double> r(k), y(k);
  fasteig::hip::check(fe_memcpy_d2h(ex->ctx(), r.data(), rn.data(), sizeof(double) * k), "d2h");
  fasteig::hip::check(fe_memcpy_d2h(ex->ctx(), y.data(), yn.data(), sizeof(double) * k), "d2h");
  if (sell) A.use_sell();
  const double anorm = norm_estimate(opts);
  scaled_resnrms_.resize(k);
  for (int j = 0; j < k; ++j) {
    const double lam = eigvals_(j);
    const double den = (1.0 + std::fabs(lam)) * y[j];
    resnrms_(j) = den > 0.0 ? r[j] / den : r[j];
    const double den2 = anorm * std::fabs(lam) * y[j];
    scaled_resnrms_(j) = den2 > 0.0 ? r[j] / den2 : r[j];
  }
}

FASTEIG_INT fasteig_problem::solve() {
  fasteig_options o(n_);
  o.set("solver", solver_);
  return solve(o);
}

FASTEIG_INT fasteig_problem::solve(const fasteig_options& opts) {
  opts.validate();
  if (opts.nev() > n_) throw std::invalid_argument("fasteig_problem::solve: nev must be in [1,n].");
  last_iterations_.reset();
  const std::string solver_name = opts.solver().empty() ? solver_ : opts.solver();
  const int nev = opts.nev();
  eigvals_.resize(nev);
  resnrms_.resize(nev);
  eigvecs_valid_ = false;
  dev_evecs_ = Block();
  const auto t0 = std::chrono::steady_clock::now();
  auto ex = executor(opts);
  device_matrix(opts);
  auto solver = fasteig::make_fasteig_solver(solver_name, *this, opts);
  if (!solver) {
    status_ = 0;
    return status_;
  }
  solver_ = solver_name;
  status_ = solver->solve();
  ex->synchronize();
  const auto t1 = std::chrono::steady_clock::now();
  last_runtime_ = std::chrono::duration<double>(t1 - t0).count();
  last_iterations_ = solver->iterations();
  finalize_residuals(opts);
  if (log_) write_results_text(*log_);
  return status_;
}

// RAW RESULTS and DIAGNOSTICS blocks in the reference's text layout
// (src/fasteig_logger.cpp:440-472), which its test harness parses.
void fasteig_problem::write_results_text(std::ostream& os) const {
  os << "\nRAW RESULTS (FULL PRECISION)\n============================\n\nIndex Eigenvalue ResidualNorm\n";
  const auto f = os.flags();
  const auto pr = os.precision();
  os << std::setprecision(std::numeric_limits<double>::max_digits10) << std::scientific;
  for (int i = 0; i < eigvals_.size(); ++i)
    os << (i + 1) << " " << eigvals_(i) << " " << (i < resnrms_.size() ? resnrms_(i) : 0.0) << "\n";
  os << "\nDIAGNOSTICS\n===========\n\n";
  os << "  Status               : " << status_ << "\n";
  if (last_iterations_) os << "  Iterations           : " << *last_iterations_ << "\n";
  os << "  Runtime (sec)        : " << std::fixed << std::setprecision(6) << last_runtime_ << "\n\n";
  os.flags(f);
  os.precision(pr);
}
