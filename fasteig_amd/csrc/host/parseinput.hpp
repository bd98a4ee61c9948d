// Input deck and Matrix Market reading (reference: include/parseinput.hpp:36,
// src/parseinput.cpp:171-530, src/mmio.cpp).
//
// Deck: whitespace-separated "key value" pairs, '#'/'%' comment lines, keys
// matched case-insensitively by prefix (the key must start with the keyword).
// Keys: amatrix, bmatrix, logfile, nev, solver, which, maxiter, maxdim,
// blksize, verbosity, print_every, output_prefix, output_format, verbose, tol,
// shift, seed, and -- not parsed by the reference, which silently ignores them
// -- degree, maxblks, precision, gram, format, cheb_bounds.
// Matrix files by extension: .mtx, .rsa (text CSC), .fei (binary CSC).
// Matrix Market: "coordinate real general|symmetric" files; for symmetric
// files the dominant stored triangle is kept (as the reference does, :249-279)
// and expanded to full storage by fasteig_problem.
#pragma once

#include <string>
#include <vector>

#include "fasteig.hpp"

extern "C" int parseinput(char* filename, fasteig_params* params);

namespace fasteig {
// Returns n; fills 0-based triplets. Throws std::runtime_error on format errors.
int read_matrix_market(const std::string& path, std::vector<int>& row, std::vector<int>& col,
                       std::vector<double>& val, bool& symmetric);
// RSA text and FEI binary compressed-column files (full storage), the file
// format from the extension, and a reader dispatching on it.
int read_rsa(const std::string& path, std::vector<int>& row, std::vector<int>& col, std::vector<double>& val);
int read_fei(const std::string& path, std::vector<int>& row, std::vector<int>& col, std::vector<double>& val);
MatrixFormat detect_format(const std::string& path);
int read_matrix(const std::string& path, MatrixFormat fmt, std::vector<int>& row, std::vector<int>& col,
                std::vector<double>& val, bool& symmetric);
}  // namespace fasteig
