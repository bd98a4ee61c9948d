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
        p.matrixformat = MatrixFormat::MTX;
        p.n = fasteig::read_matrix_market(p.amatfname, p.arowind, p.acolind, p.anzvals, p.mtx_symmetric);
        p.annz = static_cast<int64_t>(p.anzvals.size());
      } else if (starts_with_key(key, "bmatrix")) {
        p.bmatfname = fasteig::resolve(deck, value);
        const int nb = fasteig::read_matrix_market(p.bmatfname, p.browind, p.bcolind, p.bnzvals, p.b_mtx_symmetric);
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
