// The problem object (reference: include/fasteig_problem.hpp:24-186,
// src/fasteig_problem.cpp). It owns the HIP executor, the device copy of the
// full CSR matrix, and the solve results. Eigenvectors stay resident on the
// device after a solve; eigenvectors() copies them to the host on first use.
//
// Generalized problems K u = lambda M u with a diagonal M (the input decks'
// bmatrix when it is diagonal) are solved through the congruence
// A = M^{-1/2} K M^{-1/2}, u = M^{-1/2} y, built on the device.
#pragma once

#include <chrono>
#include <iosfwd>
#include <memory>
#include <optional>
#include <string>
#include <vector>

#include "core_ops.hpp"
#include "fasteig.hpp"
#include "fasteig_options.hpp"
#include "fasteig_solver.hpp"
#include "hip_backend.hpp"

class fasteig_problem {
 public:
  struct CsrMatrixData {
    std::vector<int> rowptr;
    std::vector<int> colind;
    std::vector<double> values;
  };

  // Full (both triangles) CSR matrix with int32 indices.
  fasteig_problem(FASTEIG_INT n, CsrMatrixData A, std::string solver = "Chebyshev");
  // From a parsed input deck (triplets; a symmetric Matrix Market triangle is expanded).
  explicit fasteig_problem(const fasteig_params& p);
  ~fasteig_problem();

  // Generalized problem with diagonal M (n positive entries).
  void set_mass_diagonal(std::vector<double> mass);

  FASTEIG_INT solve();
  FASTEIG_INT solve(const fasteig_options& opts);

  VectorXd& eigenvalues() noexcept { return eigvals_; }
  const VectorXd& eigenvalues() const noexcept { return eigvals_; }
  VectorXd& residual_norms() noexcept { return resnrms_; }
  const VectorXd& residual_norms() const noexcept { return resnrms_; }
  // ||A x - lambda x|| / (||A|| |lambda| ||x||): the norm used by the MI355X
  // acceptance tests (||A|| from the Lanczos spectral bound estimate).
  const VectorXd& scaled_residual_norms() const noexcept { return scaled_resnrms_; }
  const MatrixXd& eigenvectors();
  FASTEIG_INT status() const noexcept { return status_; }
  FASTEIG_INT n() const noexcept { return n_; }
  int64_t nnz() const noexcept { return static_cast<int64_t>(csrA_.values.size()); }
  const std::string& solver() const noexcept { return solver_; }
  bool is_generalized() const noexcept { return !mass_.empty(); }
  const CsrMatrixData* csr_matrix() const noexcept { return &csrA_; }
  std::optional<long long> last_iterations() const noexcept { return last_iterations_; }
  double last_runtime() const noexcept { return last_runtime_; }

  std::shared_ptr<fasteig::hip::HipExecutor> executor(const fasteig_options& opts);
  void set_executor(std::shared_ptr<fasteig::hip::HipExecutor> exec);

  // Device access used by the solvers.
  fasteig::hip::HipCsrMatrix& device_matrix(const fasteig_options& opts);
  // Final Ritz vectors (FP64, n x nev) produced by a solver; kept resident.
  void set_device_eigenvectors(fasteig::hip::Block&& X);
  const fasteig::hip::Block& device_eigenvectors() const { return dev_evecs_; }
  // ||A||_2 estimate (Lanczos bounds of A, cached).
  double norm_estimate(const fasteig_options& opts);

  // Logging (reference: src/fasteig_logger.cpp:423-473 RAW RESULTS / DIAGNOSTICS).
  void set_log_stream(std::ostream* os) noexcept { log_ = os; }
  void write_results_text(std::ostream& os) const;

  // the executor of the last solve (null before the first)
  std::shared_ptr<fasteig::hip::HipExecutor> current_executor() const { return executor_; }

 private:
  void build_from_triplets(FASTEIG_INT n, const std::vector<int>& r, const std::vector<int>& c,
                           const std::vector<double>& v, bool symmetric_triangle, CsrMatrixData& out);
  void finalize_residuals(const fasteig_options& opts);

  FASTEIG_INT n_ = 0;
  std::string solver_ = "Chebyshev";
  CsrMatrixData csrA_;
  std::vector<double> mass_;
  std::shared_ptr<fasteig::hip::HipExecutor> executor_;
  std::unique_ptr<fasteig::hip::HipCsrMatrix> dA_;
  bool dA_scaled_ = false;
  fasteig::hip::Block dev_evecs_;
  std::optional<double> norm_est_;

  VectorXd eigvals_, resnrms_, scaled_resnrms_;
  MatrixXd eigvecs_;
  bool eigvecs_valid_ = false;
  FASTEIG_INT status_ = 0;
  std::optional<long long> last_iterations_;
  double last_runtime_ = 0.0;
  std::ostream* log_ = nullptr;
};
