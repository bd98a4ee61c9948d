// Solver base class and factory (reference: include/fasteig_solver.hpp:84-169,
// src/fasteig_solver.cpp:131-168). The block solvers named by the hot path run
// on the HIP backend; the Krylov, SPPC and RMM-DIIS families of the reference
// are not part of this library and the factory rejects them with a message.
#pragma once

#include <memory>
#include <optional>
#include <string>

#include "fasteig.hpp"
#include "fasteig_options.hpp"

class fasteig_problem;

namespace fasteig {

class fasteig_solver {
 public:
  virtual ~fasteig_solver() = default;
  virtual FASTEIG_INT solve() = 0;
  FASTEIG_INT status() const noexcept { return status_; }
  std::optional<long long> iterations() const noexcept {
    return iterations_ >= 0 ? std::optional<long long>(iterations_) : std::nullopt;
  }

 protected:
  fasteig_solver(::fasteig_problem& problem, const fasteig_options& options);
  void set_iterations(long long it) noexcept { iterations_ = it; }

  ::fasteig_problem& problem_;
  VectorXd& eigvals_;
  VectorXd& resnrms_;
  fasteig_options options_;
  FASTEIG_INT status_ = 0;
  long long iterations_ = -1;
};

// Chebyshev-filtered subspace iteration (src/ChebyshevSolver.cpp:67-258).
// Extensions: blksize > nev sets the subspace capacity; precision=mixed runs
// the filter, CholQR Gram and Rayleigh-Ritz blocks in FP32 until the FP32
// residuals stop improving, then finishes with FP64 sweeps.
class ChebyshevSolver final : public fasteig_solver {
 public:
  ChebyshevSolver(::fasteig_problem& problem, const fasteig_options& options);
  FASTEIG_INT solve() override;
};

// Block Krylov-Schur with thick restart (src/BlockKrylovSchurSolver.cpp:26-214):
// block Lanczos expansion with full two-pass block Gram-Schmidt and a
// rank-aware block QR, block Ritz estimates from the coupling block, restart
// keeping max(ceil(nev/blksize), blocks/2) blocks of Ritz vectors. Standard
// problems, which = SR, LR, SM, LM, BE (shift-invert needs a sparse direct
// solver and is not part of the HIP backend).
class BlockKrylovBase : public fasteig_solver {
 protected:
  using fasteig_solver::fasteig_solver;
  // restarted: block Krylov-Schur; otherwise block Lanczos
  FASTEIG_INT run(bool restarted);
};
class BlockKrylovSchurSolver final : public BlockKrylovBase {
 public:
  BlockKrylovSchurSolver(::fasteig_problem& problem, const fasteig_options& options);
  FASTEIG_INT solve() override;
};

// Block Lanczos without restart (src/BlockLanczosSolver.cpp:24-172): the same
// block expansion and block QR as block Krylov-Schur, a convergence check
// after every block once iter * blksize >= nev, at most maxiter blocks (the
// projected dimension at most 2048 here).
class BlockLanczosSolver final : public BlockKrylovBase {
 public:
  BlockLanczosSolver(::fasteig_problem& problem, const fasteig_options& options);
  FASTEIG_INT solve() override;
};

// Scalar Lanczos with full two-pass reorthogonalization and no restart
// (src/LanczosSolver.cpp:28-164): convergence checked every print_every steps
// (5 when 0) once m >= nev. Standard problems, which = SR, LR, SM, LM, BE; the
// projected dimension is limited to 2048 on the HIP backend.
class LanczosSolver final : public fasteig_solver {
 public:
  LanczosSolver(::fasteig_problem& problem, const fasteig_options& options);
  FASTEIG_INT solve() override;
};

// Scalar Krylov-Schur with thick restart (src/KrylovSchurSolver.cpp:28-167):
// subspace max(2 nev, 20, maxdim) (at most 2048 here), restart keeping
// nev + min(nconv, (m - nev) / 2) Ritz vectors.
class KrylovSchurSolver final : public fasteig_solver {
 public:
  KrylovSchurSolver(::fasteig_problem& problem, const fasteig_options& options);
  FASTEIG_INT solve() override;
};

// LOBPCG without preconditioner (src/LOBPCGSolver.cpp:25-173). Extension:
// blksize > nev widens the block; convergence and output use the first nev.
class LOBPCGSolver final : public fasteig_solver {
 public:
  LOBPCGSolver(::fasteig_problem& problem, const fasteig_options& options);
  FASTEIG_INT solve() override;
};

// Block Davidson with thick restart (src/DavidsonSolver.cpp:27-264).
class DavidsonSolver final : public fasteig_solver {
 public:
  DavidsonSolver(::fasteig_problem& problem, const fasteig_options& options);
  FASTEIG_INT solve() override;
};

// Prefix dispatch like the reference (src/fasteig_solver.cpp:131-168).
std::unique_ptr<fasteig_solver> make_fasteig_solver(const std::string& solver, ::fasteig_problem& problem,
                                                    const fasteig_options& options);

}  // namespace fasteig
