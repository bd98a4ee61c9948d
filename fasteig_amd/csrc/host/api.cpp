// C ABI of the solver layer (include/fasteig_hip.h).
#include "fasteig_hip.h"

#include <cstring>
#include <memory>
#include <sstream>
#include <stdexcept>
#include <string>

#include "fasteig_problem.hpp"
#include "parseinput.hpp"

struct feh_problem_s {
  std::unique_ptr<fasteig_problem> prob;
  fasteig_options deck_options;
  bool from_deck = false;
};

namespace {
thread_local std::string g_err;

template <class F>
int guarded(F&& f) {
  try {
    return f();
  } catch (const std::invalid_argument& e) {
    g_err = e.what();
    return -2;
  } catch (const std::exception& e) {
    g_err = e.what();
    return -3;
  }
}
}  // namespace

extern "C" {

const char* feh_last_error(void) { return g_err.c_str(); }

int feh_problem_create(int n, int64_t nnz, const int* rowptr, const int* colind, const double* vals,
                       feh_problem* out) {
  return guarded([&] {
    if (!out || !rowptr || (nnz > 0 && (!colind || !vals))) throw std::invalid_argument("null argument");
    fasteig_problem::CsrMatrixData d;
    d.rowptr.assign(rowptr, rowptr + n + 1);
    d.colind.assign(colind, colind + nnz);
    d.values.assign(vals, vals + nnz);
    auto h = std::make_unique<feh_problem_s>();
    h->prob = std::make_unique<fasteig_problem>(n, std::move(d));
    *out = h.release();
    return 0;
  });
}

int feh_problem_create_from_deck(const char* path, feh_problem* out) {
  return guarded([&] {
    if (!out || !path) throw std::invalid_argument("null argument");
    fasteig_params p;
    if (parseinput(const_cast<char*>(path), &p) != 0) throw std::invalid_argument(std::string("cannot parse ") + path);
    auto h = std::make_unique<feh_problem_s>();
    h->prob = std::make_unique<fasteig_problem>(p);
    h->deck_options.configure_from(p);
    h->from_deck = true;
    *out = h.release();
    return 0;
  });
}

int feh_problem_set_mass(feh_problem p, const double* mass) {
  return guarded([&] {
    if (!p || !mass) throw std::invalid_argument("null argument");
    p->prob->set_mass_diagonal(std::vector<double>(mass, mass + p->prob->n()));
    return 0;
  });
}

int feh_problem_share_executor(feh_problem dst, feh_problem src) {
  return guarded([&] {
    if (!dst || !src) throw std::invalid_argument("null argument");
    fasteig_options o;
    dst->prob->set_executor(src->prob->executor(o));
    return 0;
  });
}

int feh_problem_destroy(feh_problem p) {
  delete p;
  return 0;
}

int feh_problem_n(feh_problem p) { return p ? p->prob->n() : -2; }

int feh_problem_solve(feh_problem p, const char* options) {
  return guarded([&] {
    if (!p) throw std::invalid_argument("null problem");
    fasteig_options o = p->from_deck ? p->deck_options : fasteig_options(p->prob->n());
    if (options) o.apply_string(options);
    return p->prob->solve(o);
  });
}

int feh_problem_nev(feh_problem p) { return p ? p->prob->eigenvalues().size() : -2; }

int feh_problem_results(feh_problem p, double* ev, double* res, double* sres, int64_t* iters, double* rt) {
  return guarded([&] {
    if (!p) throw std::invalid_argument("null problem");
    const int k = p->prob->eigenvalues().size();
    for (int j = 0; j < k; ++j) {
      if (ev) ev[j] = p->prob->eigenvalues()(j);
      if (res) res[j] = p->prob->residual_norms()(j);
      if (sres) sres[j] = j < p->prob->scaled_residual_norms().size() ? p->prob->scaled_residual_norms()(j) : 0.0;
    }
    if (iters) *iters = p->prob->last_iterations().value_or(-1);
    if (rt) *rt = p->prob->last_runtime();
    return 0;
  });
}

int feh_problem_eigenvectors(feh_problem p, double* out) {
  return guarded([&] {
    if (!p || !out) throw std::invalid_argument("null argument");
    const MatrixXd& V = p->prob->eigenvectors();
    std::memcpy(out, V.data(), sizeof(double) * static_cast<size_t>(V.size()));
    return 0;
  });
}

int feh_problem_results_text(feh_problem p, char* buf, int len) {
  return guarded([&] {
    if (!p || !buf || len <= 0) throw std::invalid_argument("null argument");
    std::ostringstream os;
    p->prob->write_results_text(os);
    const std::string s = os.str();
    std::strncpy(buf, s.c_str(), static_cast<size_t>(len - 1));
    buf[len - 1] = '\0';
    return static_cast<int>(s.size());
  });
}

}  // extern "C"
