// Input deck parser and Matrix Market reader (see parseinput.hpp).
#include "parseinput.hpp"

#include <algorithm>
#include <cctype>
#include <cstdio>
#include <fstream>
#include <iostream>
#include <limits>
#include <sstream>
#include <stdexcept>

namespace fasteig {

namespace {
std::string lower(std::string s) {
  std::transform(s.begin(), s.end(), s.begin(), [](unsigned char c) { return static_cast<char>(std::tolower(c)); });
  return s;
}
bool starts_with_key(const std::string& key, const char* prefix) {
  const std::string k = lower(key), p = lower(prefix);
  return k.size() >= p.size() && k.compare(0, p.size(), p) == 0;
}
std::string resolve(const std::string& deck, const std::string& path) {
  std::ifstream t(path);
  if (t.good() || path.empty() || path[0] == '/') return path;
  const size_t slash = deck.find_last_of('/');
  if (slash == std::string::npos) return path;
  return deck.substr(0, slash + 1) + path;
}
}  // namespace

namespace {
// Compressed-column arrays (0-based colptr of n+1, rowind, values) to
// triplets, with the checks of the reference's build_sparse_matrix_from_RSA
// (src/fasteig_problem.cpp:509-550). Full storage: nothing is mirrored.
int csc_to_triplets(const std::string& path, long long nrows, long long ncols, long long nnz,
                    const std::vector<long long>& colptr, const std::vector<long long>& rowind,
                    std::vector<double>& values, std::vector<int>& row, std::vector<int>& col, std::vector<double>& val) {
  if (nrows <= 0 || ncols <= 0 || nnz < 0 || nrows != ncols || nrows > std::numeric_limits<int>::max())
    throw std::runtime_error("invalid matrix sizes in " + path);
  if (static_cast<long long>(colptr.size()) != ncols + 1 || colptr.front() != 0 || colptr.back() != nnz)
    throw std::runtime_error("colptr must start at 0 and end at nnz: " + path);
  row.clear();
  col.clear();
  row.reserve(static_cast<size_t>(nnz));
  col.reserve(static_cast<size_t>(nnz));
  for (long long j = 0; j < ncols; ++j) {
    const long long b = colptr[static_cast<size_t>(j)], e = colptr[static_cast<size_t>(j) + 1];
    if (b < 0 || e < b || e > nnz) throw std::runtime_error("invalid colptr range in " + path);
    for (long long k = b; k < e; ++k) {
      const long long i = rowind[static_cast<size_t>(k)];
      if (i < 0 || i >= nrows) throw std::runtime_error("row index out of range in " + path);
      row.push_back(static_cast<int>(i));
      col.push_back(static_cast<int>(j));
    }
  }
  val = std::move(values);
  return static_cast<int>(nrows);
}
}  // namespace

// RSA text file (reference readRSA, src/parseinput.cpp:27-67): five header
// lines, the third holding "type nrows ncols nnz ...", then the 1-based column
// pointers, row indices and values, whitespace separated.
int read_rsa(const std::string& path, std::vector<int>& row, std::vector<int>& col, std::vector<double>& val) {
  std::ifstream f(path);
  if (!f) throw std::runtime_error("cannot open RSA file: " + path);
  std::string lines[5];
  for (auto& l : lines) std::getline(f, l);
  std::istringstream h(lines[2]);
  std::string type;
  long long nrows = 0, ncols = 0, nnz = 0;
  h >> type >> nrows >> ncols >> nnz;
  if (!h || nrows <= 0 || ncols <= 0 || nnz < 0) throw std::runtime_error("invalid RSA header in " + path);
  std::vector<long long> colptr(static_cast<size_t>(ncols) + 1), rowind(static_cast<size_t>(nnz));
  std::vector<double> values(static_cast<size_t>(nnz));
  for (auto& x : colptr) {
    f >> x;
    --x;
  }
  for (auto& x : rowind) {
    f >> x;
    --x;
  }
  for (auto& x : values) f >> x;
  if (!f) throw std::runtime_error("truncated RSA file: " + path);
  return csc_to_triplets(path, nrows, ncols, nnz, colptr, rowind, values, row, col, val);
}

// FEI binary file (reference readFEI, src/parseinput.cpp:69-111, written by
// examples/MTX2FEI.cpp): int32 nrows, ncols, nnz; int32 colptr[ncols + 1]
// (0-based); int32 rowind[nnz]; float64 values[nnz]; native byte order.
int read_fei(const std::string& path, std::vector<int>& row, std::vector<int>& col, std::vector<double>& val) {
  std::ifstream f(path, std::ios::binary);
  if (!f) throw std::runtime_error("cannot open FEI file: " + path);
  int32_t hdr[3] = {0, 0, 0};
  f.read(reinterpret_cast<char*>(hdr), sizeof hdr);
  if (!f) throw std::runtime_error("Failed reading matrix header: " + path);
  if (hdr[0] <= 0 || hdr[1] <= 0 || hdr[2] < 0) throw std::runtime_error("Invalid matrix sizes in binary file: " + path);
  std::vector<int32_t> cp(static_cast<size_t>(hdr[1]) + 1), ri(static_cast<size_t>(hdr[2]));
  std::vector<double> values(static_cast<size_t>(hdr[2]));
  f.read(reinterpret_cast<char*>(cp.data()), static_cast<std::streamsize>(sizeof(int32_t) * cp.size()));
  f.read(reinterpret_cast<char*>(ri.data()), static_cast<std::streamsize>(sizeof(int32_t) * ri.size()));
  f.read(reinterpret_cast<char*>(values.data()), static_cast<std::streamsize>(sizeof(double) * values.size()));
  if (!f) throw std::runtime_error("Failed reading CSC arrays: " + path);
  std::vector<long long> colptr(cp.begin(), cp.end()), rowind(ri.begin(), ri.end());
  return csc_to_triplets(path, hdr[0], hdr[1], hdr[2], colptr, rowind, values, row, col, val);
}

int read_matrix_market(const std::string& path, std::vector<int>& row, std::vector<int>& col,
                       std::vector<double>& val, bool& symmetric) {
  FILE* f = std::fopen(path.c_str(), "r");
  if (!f) throw std::runtime_error("cannot open MatrixMarket file: " + path);
  char line[1024];
  if (!std::fgets(line, sizeof line, f)) {
    std::fclose(f);
    throw std::runtime_error("empty MatrixMarket file: " + path);
  }
  std::istringstream banner(line);
  std::string tag, object, format, field, sym;
  banner >> tag >> object >> format >> field >> sym;
  tag = lower(tag);
  object = lower(object);
  format = lower(format);
  field = lower(field);
  sym = lower(sym);
  if (tag != "%%matrixmarket" || object != "matrix" || format != "coordinate" ||
      (field != "real" && field != "integer" && field != "pattern")) {
    std::fclose(f);
    throw std::runtime_error("unsupported Matrix Market banner in: " + path);
  }
  if (sym != "general" && sym != "symmetric") {
    std::fclose(f);
    throw std::runtime_error("Matrix Market symmetry '" + sym + "' is not supported: " + path);
  }
  symmetric = sym == "symmetric";
  long long M = 0, N = 0, nz = 0;
  while (std::fgets(line, sizeof line, f)) {
    if (line[0] == '%') continue;
    if (std::sscanf(line, "%lld %lld %lld", &M, &N, &nz) == 3) break;
  }
  if (M != N || M <= 0) {
    std::fclose(f);
    throw std::runtime_error("MatrixMarket file is not square: " + path);
  }
  row.clear();
  col.clear();
  val.clear();
  row.reserve(nz);
  col.reserve(nz);
  val.reserve(nz);
  long long upper = 0, lowr = 0;
  for (long long i = 0; i < nz; ++i) {
    long long r = 0, c = 0;
    double x = 1.0;
    const int got = field == "pattern" ? std::fscanf(f, "%lld %lld", &r, &c) : std::fscanf(f, "%lld %lld %lg", &r, &c, &x);
    if (got < 2) {
      std::fclose(f);
      throw std::runtime_error("unexpected EOF while reading: " + path);
    }
    row.push_back(static_cast<int>(r - 1));
    col.push_back(static_cast<int>(c - 1));
    val.push_back(x);
    if (r < c) ++upper;
    else if (r > c) ++lowr;
  }
  std::fclose(f);
  if (symmetric) {  // keep the dominant stored triangle (src/parseinput.cpp:249-279)
    const bool up = upper >= lowr;
    size_t w = 0;
    for (size_t k = 0; k < row.size(); ++k) {
      const bool keep = row[k] == col[k] || (up ? row[k] < col[k] : row[k] > col[k]);
      if (!keep) continue;
      row[w] = row[k];
      col[w] = col[k];
      val[w] = val[k];
      ++w;
    }
    row.resize(w);
    col.resize(w);
    val.resize(w);
  }
  return static_cast<int>(N);
}

// Matrix file format from the extension (reference detect_format,
// src/parseinput.cpp:114-137); .hdf is recognised there but not read.
MatrixFormat detect_format(const std::string& path) {
  const size_t dot = path.find_last_of('.');
  const size_t slash = path.find_last_of('/');
  if (dot == std::string::npos || (slash != std::string::npos && dot < slash)) return MatrixFormat::UNKNOWN;
  const std::string ext = lower(path.substr(dot));
  if (ext == ".mtx") return MatrixFormat::MTX;
  if (ext == ".rsa") return MatrixFormat::RSA;
  if (ext == ".fei") return MatrixFormat::FEI;
  if (ext == ".hdf") return MatrixFormat::HDF;
  return MatrixFormat::UNKNOWN;
}

int read_matrix(const std::string& path, MatrixFormat fmt, std::vector<int>& row, std::vector<int>& col,
                std::vector<double>& val, bool& symmetric) {
  symmetric = false;
  switch (fmt) {
    case MatrixFormat::MTX:
      return read_matrix_market(path, row, col, val, symmetric);
    case MatrixFormat::RSA:
      return read_rsa(path, row, col, val);
    case MatrixFormat::FEI:
      return read_fei(path, row, col, val);
    default:
      throw std::runtime_error("unsupported matrix file format (use .mtx, .rsa or .fei): " + path);
  }
}

}  // namespace fasteig

extern "C" int parseinput(char* filename, fasteig_params* params) {
  if (!filename || !params) return -1;
  const std::string deck(filename);
  std::ifstream in(deck);
  if (!in) {
    std::cerr << "failed to open " << deck << "\n";
    return -5;
  }
  fasteig_params& p = *params;
  try {
    std::string key, value;
    while (in >> key) {
      if (!key.empty() && (key[0] == '#' || key[0] == '%')) {
        in.ignore(std::numeric_limits<std::streamsize>::max(), '\n');
        continue;
      }
      if (!(in >> value)) break;
      using fasteig::starts_with_key;
      if (starts_with_key(key, "amatrix")) {
        p.amatfname = fasteig::resolve(deck, value);
        p.matrixformat = fasteig::detect_format(p.amatfname);
        p.n = fasteig::read_matrix(p.amatfname, p.matrixformat, p.arowind, p.acolind, p.anzvals, p.mtx_symmetric);
        p.annz = static_cast<int64_t>(p.anzvals.size());
      } else if (starts_with_key(key, "bmatrix")) {
        p.bmatfname = fasteig::resolve(deck, value);
        const MatrixFormat bf = fasteig::detect_format(p.bmatfname);
        if (p.matrixformat != MatrixFormat::UNKNOWN && bf != p.matrixformat)
          throw std::runtime_error("bmatrix format differs from amatrix format");
        const int nb = fasteig::read_matrix(p.bmatfname, bf, p.browind, p.bcolind, p.bnzvals, p.b_mtx_symmetric);
        if (p.n > 0 && nb != p.n) throw std::runtime_error("bmatrix dimension differs from amatrix");
        p.bnnz = static_cast<int64_t>(p.bnzvals.size());
      } else if (starts_with_key(key, "logfile") || starts_with_key(key, "output_prefix")) {
        p.logprefix = value;
      } else if (starts_with_key(key, "nev")) {
        p.nev = std::stoi(value);
      } else if (starts_with_key(key, "solver")) {
        p.solver = value;
      } else if (starts_with_key(key, "which")) {
        p.which = value;
      } else if (starts_with_key(key, "maxiter")) {
        p.maxiter = std::stoi(value);
      } else if (starts_with_key(key, "maxdim")) {
        p.maxdim = std::stoi(value);
      } else if (starts_with_key(key, "maxblks")) {
        p.maxblks = std::stoi(value);
      } else if (starts_with_key(key, "blksize")) {
        p.blksize = std::stoi(value);
      } else if (starts_with_key(key, "degree")) {
        p.degree = std::stoi(value);
      } else if (starts_with_key(key, "verbosity")) {
        p.verbosity = std::stoi(value);
      } else if (starts_with_key(key, "print_every")) {
        p.print_every = std::stoi(value);
      } else if (starts_with_key(key, "output_format")) {
        p.output_format = value;
      } else if (starts_with_key(key, "verbose")) {
        p.verbosity = 1;
      } else if (starts_with_key(key, "tol")) {
        p.tol = std::stod(value);
      } else if (starts_with_key(key, "shift")) {
        p.shift = std::stod(value);
      } else if (starts_with_key(key, "seed")) {
        p.seed = std::stoi(value);
      } else if (starts_with_key(key, "precision")) {
        p.precision = value;
      } else if (starts_with_key(key, "gram")) {
        p.gram = value;
      } else if (starts_with_key(key, "format")) {
        p.format = value;
      } else if (starts_with_key(key, "cheb_bounds")) {
        p.cheb_bounds = value;
      }
      // unknown keys are ignored, as in the reference
    }
  } catch (const std::exception& e) {
    std::cerr << "parseinput: " << e.what() << "\n";
    return -1;
  }
  return 0;
}
