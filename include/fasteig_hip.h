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

#ifdef __cplusplus
}
#endif
#endif /* FASTEIG_HIP_H */
