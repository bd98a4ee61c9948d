// Continuous-batching solve server on one GPU (no reference counterpart: the
// reference solves one problem per fasteig_problem::solve call,
// src/fasteig_problem.cpp:340-437).
//
// Requests (CSR matrix, nev, tol) queue on the host and are admitted in
// arrival order while their HBM footprint fits the budget and a slot is
// free; every admitted problem keeps its matrix and Ritz blocks resident.
// Each step() runs one cycle for all active problems together: a Chebyshev
// filter of `degree` grouped SpMM launches (FP32 blocks, or FP64 once a
// problem switches), two CholQR passes and a Rayleigh-Ritz step with FP64
// Gram accumulation, FP64 small factorizations (equilibrated Cholesky,
// Jacobi eigensolver) and FP64 residual norms. One table upload and one
// graph launch (or the same launches eagerly) per cycle; then converged
// problems retire and new ones join on the next step. A problem's arithmetic
// never depends on the rest of the batch, so its result is bitwise the same
// alone and in any mix.
//
// Per problem (m = subspace columns = min(64, round_up(2 nev, 4)) <= n):
// CheFSI with adaptive bounds (the damped interval runs from the largest
// Ritz value of the subspace to a Gershgorin upper bound; first cycle:
// CholQR + Rayleigh-Ritz of the seeded random block, no filter). FP32 cycles
// until the largest relative residual of the wanted pairs is below
// switch_tol or stops improving, then FP64 cycles; converged when every
// wanted ||A y - theta y|| / ((1 + |theta|) ||y||) <= tol.
#pragma once

#include <chrono>
#include <cstdint>
#include <deque>
#include <memory>
#include <vector>

#include "hip_backend.hpp"

namespace fasteig {
namespace hip {

struct ServeRequest {
  int64_t id = 0;
  int n = 0;
  std::vector<int> rowptr, colind;
  std::vector<double> values;
  int nev = 0;
  double tol = 1e-8;
};

struct ServeResult {
  int64_t id = 0;
  int status = 0;       // 0 converged, 1 maxiter reached, 2 rejected (too large / invalid)
  int cycles = 0;
  int fp64_cycles = 0;
  int n = 0, nev = 0;
  std::vector<double> eigenvalues, residuals;
  std::vector<double> eigenvectors;  // n x nev column-major (if kept)
  double queued_s = 0.0, admitted_s = 0.0, finished_s = 0.0;  // server clock
};

struct ServeOptions {
  size_t hbm_budget = 0;  // bytes for resident problems; 0: 80% of free device memory
  int max_slots = 128;    // problems in flight
  int degree = 16;        // Chebyshev filter degree per cycle
  int maxiter = 300;      // cycles per problem
  double switch_tol = 2e-5;
  bool use_graph = true;
  bool keep_vectors = true;
  uint64_t seed = 1;
};

struct ServeStats {
  int64_t cycles = 0, admitted = 0, retired = 0;
  double device_s = 0.0;  // host wall time spent in cycles
  int64_t peak_active = 0;
  size_t peak_bytes = 0;
};

// Device bytes that one problem holds while resident (the admission unit).
size_t serve_footprint(int n, int64_t nnz, int nev);

class BatchServer {
 public:
  BatchServer(int device, const ServeOptions& opts);
  ~BatchServer();
  BatchServer(const BatchServer&) = delete;
  BatchServer& operator=(const BatchServer&) = delete;

  void submit(ServeRequest&& r);
  // One cycle: admit, run all active problems, retire. Returns the number of
  // problems still active or queued.
  int step();
  int active() const { return static_cast<int>(slots_used_); }
  int queued() const { return static_cast<int>(queue_.size()); }
  std::vector<ServeResult> take_results();
  const ServeStats& stats() const { return stats_; }
  double now() const;

 private:
  struct Slot;
  struct Tables;
  void admit();
  void run_cycle();
  void retire(Slot& s, int status);
  void build_tables(std::vector<Slot*>& act);
  void capture_graph();

  ServeOptions o_;
  std::unique_ptr<HipExecutor> ex_;
  std::deque<std::pair<ServeRequest, double>> queue_;
  std::vector<std::unique_ptr<Slot>> slots_;  // max_slots entries, null when free
  size_t slots_used_ = 0;
  size_t used_bytes_ = 0;
  std::vector<ServeResult> done_;
  ServeStats stats_;
  std::chrono::steady_clock::time_point t0_;
  // device tables: one buffer, fixed offsets per launch of the cycle
  void* tab_dev_ = nullptr;
  void* tab_host_ = nullptr;  // pinned staging
  size_t tab_bytes_ = 0;
  std::vector<size_t> tab_off_;
  double* res_dev_ = nullptr;  // per slot: theta[64], rnorm[64], ynorm[64], info[...]
  std::vector<double> res_host_;
  void* graph_ = nullptr;
  void* graph_exec_ = nullptr;
};

}  // namespace hip
}  // namespace fasteig
