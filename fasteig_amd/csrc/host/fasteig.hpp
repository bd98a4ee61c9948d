// Core host types of the fasteig API, kept name-compatible with the reference
// (include/fasteig.hpp:1-207) without Eigen: VectorXd / MatrixXd are small
// column-major containers, and fasteig_params carries the parsed input deck.
#pragma once

#include <algorithm>
#include <cmath>
#include <cstddef>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#ifndef FASTEIG_DOUBLE
#define FASTEIG_DOUBLE double
#endif
#ifndef FASTEIG_INT
#define FASTEIG_INT int
#endif

// Dense column vector (reference alias: Eigen::VectorXd).
class VectorXd {
 public:
  VectorXd() = default;
  explicit VectorXd(int n, double v = 0.0) : v_(static_cast<size_t>(n), v) {}
  int size() const noexcept { return static_cast<int>(v_.size()); }
  void resize(int n) { v_.assign(static_cast<size_t>(n), 0.0); }
  double& operator()(int i) { return v_[static_cast<size_t>(i)]; }
  double operator()(int i) const { return v_[static_cast<size_t>(i)]; }
  double& operator[](int i) { return v_[static_cast<size_t>(i)]; }
  double operator[](int i) const { return v_[static_cast<size_t>(i)]; }
  double* data() noexcept { return v_.data(); }
  const double* data() const noexcept { return v_.data(); }
  double maxCoeff() const {
    if (v_.empty()) throw std::logic_error("VectorXd::maxCoeff on empty vector");
    return *std::max_element(v_.begin(), v_.end());
  }
  double minCoeff() const {
    if (v_.empty()) throw std::logic_error("VectorXd::minCoeff on empty vector");
    return *std::min_element(v_.begin(), v_.end());
  }
  bool allFinite() const {
    for (double x : v_)
      if (!std::isfinite(x)) return false;
    return true;
  }
  std::vector<double>& vec() noexcept { return v_; }
  const std::vector<double>& vec() const noexcept { return v_; }

 private:
  std::vector<double> v_;
};

// Dense column-major matrix (reference alias: Eigen::MatrixXd).
class MatrixXd {
 public:
  MatrixXd() = default;
  MatrixXd(int r, int c) { resize(r, c); }
  int rows() const noexcept { return r_; }
  int cols() const noexcept { return c_; }
  int64_t size() const noexcept { return static_cast<int64_t>(r_) * c_; }
  void resize(int r, int c) {
    r_ = r;
    c_ = c;
    a_.assign(static_cast<size_t>(r) * static_cast<size_t>(c), 0.0);
  }
  void setZero() { std::fill(a_.begin(), a_.end(), 0.0); }
  double& operator()(int i, int j) { return a_[static_cast<size_t>(j) * r_ + i]; }
  double operator()(int i, int j) const { return a_[static_cast<size_t>(j) * r_ + i]; }
  double* data() noexcept { return a_.data(); }
  const double* data() const noexcept { return a_.data(); }
  double* col(int j) noexcept { return a_.data() + static_cast<size_t>(j) * r_; }
  const double* col(int j) const noexcept { return a_.data() + static_cast<size_t>(j) * r_; }

 private:
  int r_ = 0, c_ = 0;
  std::vector<double> a_;
};

enum class MatrixFormat { MTX, RSA, FEI, HDF, UNKNOWN };

// Parsed input deck (reference: include/fasteig.hpp:116-207). Matrices are kept
// as 0-based triplets of the stored entries; for symmetric Matrix Market files
// only one triangle is stored (mtx_symmetric) and is expanded when the problem
// builds its full CSR matrix.
class fasteig_params {
 public:
  std::string amatfname;
  std::string bmatfname;
  std::string logprefix = "logfile";
  std::string output_format = "text";
  MatrixFormat matrixformat = MatrixFormat::UNKNOWN;

  FASTEIG_INT n = 0;
  int64_t annz = 0;
  int64_t bnnz = 0;
  std::vector<FASTEIG_INT> arowind, acolind;
  std::vector<FASTEIG_DOUBLE> anzvals;
  std::vector<FASTEIG_INT> browind, bcolind;
  std::vector<FASTEIG_DOUBLE> bnzvals;
  bool mtx_symmetric = false;
  bool b_mtx_symmetric = false;

  FASTEIG_INT nev = 100;
  std::string which = "LM";
  std::string solver = "Lanczos";
  FASTEIG_INT maxiter = 10;
  FASTEIG_INT maxdim = 20;
  FASTEIG_INT degree = 10;
  FASTEIG_INT blksize = 5;
  FASTEIG_INT maxblks = 0;
  FASTEIG_DOUBLE shift = 0.0;
  FASTEIG_DOUBLE tol = 1.0e-6;
  FASTEIG_INT seed = -1;
  FASTEIG_INT verbosity = -1;
  FASTEIG_INT print_every = 0;
  // MI355X backend extensions (deck keys `precision`, `gram`, `format`,
  // `cheb_bounds`; empty cheb_bounds keeps the options' default)
  std::string precision = "fp64";
  std::string gram = "native";
  std::string format = "auto";
  std::string cheb_bounds;
};
