/*
 * fasteig_hip.h -- C ABI of the solver layer (libfasteig_hip.so), for FFI
 * callers (ctypes, cgo, JNI) that drive whole solves.
 *
 * Each entry point wraps the C++ API that replaces the reference's
 *   fasteig_problem (include/fasteig_problem.hpp:24-186)
 *   fasteig_options (include/fasteig_options.hpp:22-134; key/value `set`, :32)
 *   fasteig_problem::solve(opts) (src/fasteig_problem.cpp:340-437)
 * and returns 0 on success, -2 for invalid arguments (std::invalid_argument),
 * -3 for numerical or device failures (std::runtime_error); feh_last_error()
 * gives the message. Options are passed as "key=value;key=value" with the
 * reference's keys (nev, which, solver, maxiter, degree, blksize, maxblks, tol,
 * seed, verbosity, print_every) plus precision, gram, format and device.
 */
#ifndef FASTEIG_HIP_H
#define FASTEIG_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct feh_problem_s* feh_problem;

const char* feh_last_error(void);

/* Full (both triangles) CSR matrix, int32 indices, host arrays (copied). */
int feh_problem_create(int n, int64_t nnz, const int* rowptr, const int* colind, const double* vals,
                       feh_problem* out);
/* Problem from an input deck (reference parseinput, src/parseinput.cpp:282-530);
 * the deck's options are applied first, then `options` overrides them. */
int feh_problem_create_from_deck(const char* deck_path, feh_problem* out);
/* Generalized K u = lambda M u with diagonal M (n host values). */
int feh_problem_set_mass(feh_problem p, const double* mass);
/* Share the device context of `src` (one HIP context per device). */
int feh_problem_share_executor(feh_problem dst, feh_problem src);
int feh_problem_destroy(feh_problem p);

int feh_problem_n(feh_problem p);
/* Run a solve; returns the solver status (>= 0) or a negative error code. */
int feh_problem_solve(feh_problem p, const char* options);
/* Results of the last solve: nev values each (NULL to skip). */
int feh_problem_nev(feh_problem p);
int feh_problem_results(feh_problem p, double* eigenvalues, double* residual_norms, double* scaled_residual_norms,
                        int64_t* iterations, double* runtime_sec);
/* Eigenvectors, n x nev column-major host array. */
int feh_problem_eigenvectors(feh_problem p, double* out);
/* RAW RESULTS / DIAGNOSTICS text of the last solve into buf (truncated). */
int feh_problem_results_text(feh_problem p, char* buf, int len);
/* The product-mode measurements of the problem's executor (options
 * gram=auto; no counterpart in the reference, whose products are cuBLAS
 * calls): one line per tuned product shape, "kind rows m p" then the time
 * (ms) of the native, BF16x3 and BF16x6 modes and their errors relative to the
 * FP64 product of the same FP32 data on a row sample. Returns the text
 * length; truncated to len - 1. */
int feh_problem_product_tuning(feh_problem p, char* buf, int len);

/* ---- core numerics of the block solvers on the problem's device (the
 * reference's L4 free functions, which the solvers call between SpMMs).
 * Host blocks are n x k column-major FP64 (copied in and out); the work runs
 * in FP64 on the device of the problem's executor. ----------------------- */
/* 20-step Lanczos bound estimate of sign * A (the problem's matrix), two-pass
 * reorthogonalized, seeded start vector: lower = min(theta - |beta y_last|),
 * upper = max(theta + |beta y_last|). Reference: estimate_bounds_with_executor,
 * src/ChebyshevSolver.cpp:26-49; KrylovSubspace::build, src/core/KrylovSubspace.cpp:112-131. */
int feh_core_lanczos_bounds(feh_problem p, double sign, unsigned seed, double* lower, double* upper);
/* thin_qr_in_place (src/core/Orthogonalization.cpp:403-480) on X (n x k, in
 * place; R is not returned). method 0: Cholesky QR (QrMethod::Cholesky) --
 * returns 1 and leaves X unchanged when the Gram matrix is not positive
 * definite (the reference's -1 status, :49-57,216-225). method 1: the
 * orthonormal-basis role of QrMethod::Householder (:351-365) -- Cholesky QR
 * twice, with a shifted Cholesky QR first when the plain one fails (shifted
 * CholQR3); returns 2 when the shifted path ran, else 0. */
int feh_core_thin_qr(feh_problem p, int64_t n, int k, double* X, int method);
/* orthonormalize_block (src/core/Orthogonalization.cpp:717-775): orthonormal
 * directions of src (n x b) against basis[:, 0:active] (n x active, orthonormal)
 * whose singular values exceed rank_tol * max column norm (two passes).
 * out holds n x b doubles; *rank receives the kept column count (0: none). */
int feh_core_orthonormalize_block(feh_problem p, int64_t n, int active, const double* basis, int b,
                                  const double* src, double rank_tol, double* out, int* rank);

/* ---- continuous-batching solve server (no reference counterpart: the
 * reference runs one solve per fasteig_problem::solve call) ------------------
 * Requests (CSR matrix, nev, tol; smallest eigenpairs) queue on the host and
 * are admitted in arrival order while their device footprint fits the HBM
 * budget and a slot is free. Each step runs one cycle for every active
 * problem with grouped launches (Chebyshev filter in FP32, FP64 after the
 * precision switch; CholQR; Rayleigh-Ritz with FP64 Gram, Jacobi eigensolver
 * and residuals), then retires converged problems. A problem's arithmetic is
 * independent of the batch, so its results are bitwise the same alone and in
 * any mix. Options "key=value;...": max_slots (128), degree (16), maxiter
 * (300), hbm_budget_mb (75% of device memory), switch_tol (2e-5), graph (1:
 * replay one captured graph per cycle), keep_vectors (1), seed (1). nev may be
 * 1..32 and the subspace (2 nev rounded up to 4 columns) must be below n;
 * other requests come back with status 2. */
typedef struct feh_server_s* feh_server;

/* Device bytes one request holds while resident (admission unit); host-only. */
size_t feh_serve_footprint(int n, int64_t nnz, int nev);
int feh_server_create(int device, const char* options, feh_server* out);
int feh_server_destroy(feh_server s);
/* Queue a request (arrays are copied). */
int feh_server_submit(feh_server s, int64_t id, int n, int64_t nnz, const int* rowptr, const int* colind,
                      const double* vals, int nev, double tol);
/* One cycle; returns the number of requests still active or queued. */
int feh_server_step(feh_server s);
int feh_server_counts(feh_server s, int* active, int* queued, int* ready);
/* Front finished result: returns 1 and its metadata (status 0 converged,
 * 1 maxiter, 2 rejected; times = queued, admitted, finished on the server
 * clock in seconds), or 0 when none is ready. */
int feh_server_peek(feh_server s, int64_t* id, int* status, int* n, int* nev, int* cycles, int* fp64_cycles,
                    double* times);
/* Copy the front result (nev values, nev residuals, n x nev column-major
 * vectors; NULL to skip) and drop it; returns 1, or 0 when none is ready. */
int feh_server_pop(feh_server s, double* eigenvalues, double* residuals, double* eigenvectors);
int feh_server_stats(feh_server s, int64_t* cycles, int64_t* admitted, int64_t* retired, int64_t* peak_active,
                     double* cycle_seconds, size_t* peak_bytes);

#ifdef __cplusplus
}
#endif
#endif /* FASTEIG_HIP_H */
