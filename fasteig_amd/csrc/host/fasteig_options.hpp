// Solver configuration (reference: include/fasteig_options.hpp:22-134,
// src/fasteig_options.cpp). Same typed accessors, string key/value interface,
// defaults and validation; the HIP backend adds four keys:
//   precision  "fp64" (plain FP64 on the GPU) | "mixed" (FP32 filter and Gram,
//              FP64 Cholesky, Rayleigh-Ritz, residuals and closing sweeps)
//   gram       "native" (FP64/FP32 MFMA) | "bf16x3" (split-BF16 Gram, FP32 blocks)
//   format     "csr" | "sell" (SELL-32-sigma sparse product) | "blocked"
//              (4-row blocks sharing one column list: one X gather per column of
//              a block instead of one per nonzero) | "tiled" (16-row tiles whose
//              X rows are staged in LDS, loading only the rows each tile adds) |
//              "auto" (tiled, else blocked, when the matrix has the column
//              locality, else csr)
//   device     HIP device index
//   cheb_bounds "amplification" (the reference's fixed cutoff) | "ritz" (the
//              damped interval starts at the largest Ritz value of the current
//              full-capacity subspace; scaled recurrence)
#pragma once

#include <optional>
#include <string>
#include <string_view>
#include <unordered_map>

#include "fasteig.hpp"

class fasteig_options {
 public:
  fasteig_options() = default;
  explicit fasteig_options(FASTEIG_INT n) : n_(n) {}

  bool set(std::string_view key, std::string_view value);
  std::string get(std::string_view key) const;
  bool apply(const std::unordered_map<std::string, std::string>& kv, bool strict = false);
  // "key=value;key=value" convenience form used by the C ABI
  void apply_string(const std::string& kv);

  FASTEIG_INT n() const noexcept { return n_; }
  FASTEIG_INT nev() const noexcept { return nev_; }
  FASTEIG_INT maxiter() const noexcept { return maxiter_; }
  FASTEIG_INT maxdim() const noexcept { return maxdim_; }
  FASTEIG_INT maxblks() const noexcept { return maxblks_; }
  FASTEIG_INT degree() const noexcept { return degree_; }
  FASTEIG_INT blksize() const noexcept { return blksize_; }
  const std::string& which() const noexcept { return which_; }
  const std::string& solver() const noexcept { return solver_; }
  FASTEIG_DOUBLE shift() const noexcept { return shift_; }
  FASTEIG_DOUBLE tol() const noexcept { return tol_; }
  std::optional<unsigned int> seed() const noexcept { return seed_; }
  int verbosity() const noexcept { return verbosity_; }
  bool verbose() const noexcept { return verbosity_ > 0; }
  FASTEIG_INT print_every() const noexcept { return print_every_; }
  std::string logfile() const noexcept { return logprefix_; }
  std::string output_prefix() const noexcept { return logprefix_; }
  std::string output_format() const noexcept { return output_format_; }
  bool gpu_backend() const noexcept { return gpu_backend_; }
  void set_gpu_backend(bool enabled) noexcept { gpu_backend_ = enabled; }

  // HIP backend extensions
  const std::string& precision() const noexcept { return precision_; }
  const std::string& gram() const noexcept { return gram_; }
  const std::string& format() const noexcept { return format_; }
  const std::string& cheb_bounds() const noexcept { return cheb_bounds_; }
  int device() const noexcept { return device_; }
  bool mixed() const noexcept { return precision_ == "mixed"; }

  std::string output_path(std::string_view solver_name) const {
    if (logprefix_.empty()) return {};
    return logprefix_ + "_" + std::string(solver_name) + ".log";
  }

  const MatrixXd& v0() const noexcept { return v0_; }
  MatrixXd& v0() noexcept { return v0_; }

  void configure_from(const fasteig_params& p);
  void validate() const;

 private:
  std::string logprefix_ = "logfile";
  std::string output_format_ = "text";
  FASTEIG_INT n_ = 0;
  FASTEIG_INT nev_ = 10;
  FASTEIG_INT maxiter_ = 1000;
  FASTEIG_INT maxdim_ = 0;
  FASTEIG_INT maxblks_ = 10;
  FASTEIG_INT degree_ = 10;
  FASTEIG_INT blksize_ = 1;
  std::string which_ = "LM";
  std::string solver_ = "Lanczos";
  FASTEIG_DOUBLE shift_ = 0.0;
  FASTEIG_DOUBLE tol_ = 1.0e-8;
  std::optional<unsigned int> seed_;
  int verbosity_ = 1;
  FASTEIG_INT print_every_ = 0;
  bool gpu_backend_ = true;
  std::string precision_ = "fp64";
  std::string gram_ = "auto";  // product mode of FP32 blocks (mixed precision)
  std::string format_ = "auto";
  std::string cheb_bounds_ = "amplification";
  int device_ = 0;
  MatrixXd v0_;

  static FASTEIG_INT parse_int(std::string_view key, std::string_view v);
  static FASTEIG_DOUBLE parse_double(std::string_view key, std::string_view v);
  static bool parse_bool(std::string_view key, std::string_view v);
};
