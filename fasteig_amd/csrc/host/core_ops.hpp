// Device-resident core numerics of the block eigensolvers, the HIP counterparts
// of the reference's src/core/{Orthogonalization,RitzExtraction,KrylovSubspace}.cpp.
// Small matrices never leave the device except the Ritz values and residual
// norms that the solvers' control flow needs.
#pragma once

#include <string>
#include <vector>

#include "hip_backend.hpp"

namespace fasteig {
namespace hip {

// Scratch blocks a solve keeps for its whole lifetime (no allocation inside
// the iteration loop, unlike the CUDA backend's per-call cudaMalloc,
// src/gpu/CudaMultiVector.cpp:347,389).
struct Workspace {
  HipExecutor& ex;
  // product mode of FP32 blocks (fe_gram_mode or kProductAuto) and, for
  // kProductAuto, the relative product error the current step tolerates
  // (0: native products only)
  int gram_mode = FE_GRAM_NATIVE;
  double need = 0.0;
  Small G, R, W, V, C, theta, rn, yn;  // sized for the largest projected problem
  int cap = 0;
  Workspace(HipExecutor& e, int max_m, int gram);
  void ensure(int m);
};

// Cholesky QR in place: G = X^T X = R^T R, X <- X R^{-1} (via tmp, same shape).
// Returns 0, or -1 with X unchanged when the Gram matrix is not positive
// definite. Reference: src/core/Orthogonalization.cpp:49-57,185-246,290-297.
int cholesky_qr(Workspace& ws, Block& X, Block& tmp);

// Orthonormal basis of span(X) in place, the role Householder QR plays in the
// reference (thin_qr_in_place(..., Householder), src/core/Orthogonalization.cpp:351-365,
// 137-183): Cholesky QR twice; if the first factorization fails, a shifted
// Cholesky QR (shift 11 (n k + k (k+1)) u ||X||^2, Fukaya et al. 2020) runs
// first (shifted CholQR3). Returns 1 when the shifted path ran, else 0.
int orthonormal_qr(Workspace& ws, Block& X, Block& tmp);

// Two-pass classical block Gram-Schmidt of S against basis[:, :active] into
// out (out may be S itself). Reference: src/core/Orthogonalization.cpp:482-527.
void block_gram_schmidt(Workspace& ws, const BlockRef& basis, const BlockRef& S, const BlockRef& out);

// Rank-revealing orthonormalization of src against basis[:, :active]; returns
// an empty Block when nothing survives. Two passes of [block Gram-Schmidt,
// keep the directions with singular value above rank_tol * max column norm,
// orthonormalized]. Reference: src/core/Orthogonalization.cpp:717-775 (there:
// Householder QR + Jacobi SVD of R). Here the singular values come from Gram
// eigen-decompositions (SVQB, Stathopoulos & Wu 2002) on MFMA: a first level
// resolves those above 1e-6 of the largest (FP64); when it keeps fewer than
// the capacity, a second level on the remainder (projected off the first
// level's directions) resolves the rest down to the cutoff.
Block orthonormalize_block(Workspace& ws, const BlockRef& basis, const BlockRef& src, double rank_tol = 1e-12);

struct RitzPairs {
  std::vector<double> values;       // selected Ritz values
  std::vector<double> relres;       // ||r|| / ((1 + |theta|) ||y||)
  std::vector<int> selected;        // indices into the ascending eigenvalues
};

// Rayleigh-Ritz on the device. P (m x m, ld ws.cap) holds the projected matrix
// and is overwritten. Y = Q C, AY = AQ C, R = AY - Y Theta (R may be null: norms
// only). Reference: src/core/RitzExtraction.cpp:65-141, selection :162-179 and
// include/fasteig_solver.hpp:59-82. The coefficient matrix stays in ws.C.
// res_count >= 0: AY, R and the residual norms only for the first res_count
// selected pairs (relres then holds res_count entries); Y for all requested.
// The residual and its norms come from the epilogue of the AY product
// (fe_gemm_nn_resid); keep_ay false leaves AY unwritten (its block is then
// only the tuning target of that product).
// rotate_count >= 0: only the first max(res_count, rotate_count) columns of
// Y are rotated here; rotate_rest(ws, m, Q, from, Y) rotates Y's columns from
// `from` on later (with the coefficients kept in ws.C and Q unchanged), so a
// solve that stops after this extraction skips the rotation of its guard
// columns.
// rinv (upper triangular m x m, ld ws.cap, zero below the diagonal): Q is not
// orthonormal but Q rinv is (the last Cholesky QR pass left unapplied); the
// projected matrix Q^T A Q becomes rinv^T (Q^T A Q) rinv and the coefficients
// rinv C, so Y = Q (rinv C) and AY = AQ (rinv C) are the Ritz vectors of
// span(Q) without the tall product Q rinv.
RitzPairs extract_ritz(Workspace& ws, int m, const BlockRef& Q, const BlockRef& AQ, int requested,
                       const std::string& which, const BlockRef& Y, const BlockRef& AY, const BlockRef* R,
                       int res_count = -1, bool keep_ay = true, int rotate_count = -1,
                       const double* rinv = nullptr);
void rotate_rest(Workspace& ws, int m, const BlockRef& Q, int from, const BlockRef& Y);

// Stable index sort by `which` (include/fasteig_solver.hpp:59-82; "be" from
// src/core/RitzExtraction.cpp:162-179).
std::vector<int> select_indices(const std::vector<double>& v, const std::string& which);

// 20-step Lanczos bound estimate of sign*A with two-pass reorthogonalization:
// lower = min(theta - |beta y_last|), upper = max(theta + |beta y_last|).
// Reference: src/ChebyshevSolver.cpp:26-49, src/core/KrylovSubspace.cpp:112-259.
void lanczos_bounds(HipExecutor& ex, const HipCsrMatrix& A, double sign, unsigned seed, double& lower,
                    double& upper);

}  // namespace hip
}  // namespace fasteig
