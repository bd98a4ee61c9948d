// C ABI of the solver layer (include/fasteig_hip.h).
#include "fasteig_hip.h"

#include <cstring>
#include <deque>
#include <memory>
#include <sstream>
#include <stdexcept>
#include <string>

#include "core_ops.hpp"
#include "fasteig_problem.hpp"
#include "parseinput.hpp"
#include "serving.hpp"

struct feh_problem_s {
  std::unique_ptr<fasteig_problem> prob;
  fasteig_options deck_options;
  bool from_deck = false;
};

struct feh_server_s {
  std::unique_ptr<fasteig::hip::BatchServer> srv;
  std::deque<fasteig::hip::ServeResult> ready;
};

namespace {
thread_local std::string g_err;

fasteig::hip::ServeOptions parse_serve_options(const char* text) {
  fasteig::hip::ServeOptions o;
  std::string s = text ? text : "";
  size_t pos = 0;
  while (pos < s.size()) {
    size_t end = s.find(';', pos);
    if (end == std::string::npos) end = s.size();
    const std::string kv = s.substr(pos, end - pos);
    pos = end + 1;
    if (kv.find_first_not_of(" \t") == std::string::npos) continue;
    const size_t eq = kv.find('=');
    if (eq == std::string::npos) throw std::invalid_argument("serve option without '=': " + kv);
    std::string k = kv.substr(0, eq), v = kv.substr(eq + 1);
    auto trim = [](std::string& x) {
      x.erase(0, x.find_first_not_of(" \t"));
      x.erase(x.find_last_not_of(" \t") + 1);
    };
    trim(k);
    trim(v);
    try {
      if (k == "max_slots") o.max_slots = std::stoi(v);
      else if (k == "degree") o.degree = std::stoi(v);
      else if (k == "maxiter") o.maxiter = std::stoi(v);
      else if (k == "hbm_budget_mb") o.hbm_budget = static_cast<size_t>(std::stod(v) * 1048576.0);
      else if (k == "switch_tol") o.switch_tol = std::stod(v);
      else if (k == "graph") o.use_graph = std::stoi(v) != 0;
      else if (k == "keep_vectors") o.keep_vectors = std::stoi(v) != 0;
      else if (k == "seed") o.seed = std::stoull(v);
      else throw std::invalid_argument("unknown serve option: " + k);
    } catch (const std::invalid_argument&) {
      throw;
    } catch (const std::exception&) {
      throw std::invalid_argument("bad value for serve option " + k + ": " + v);
    }
  }
  return o;
}

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

int feh_core_lanczos_bounds(feh_problem p, double sign, unsigned seed, double* lower, double* upper) {
  return guarded([&] {
    if (!p || !lower || !upper) throw std::invalid_argument("null argument");
    if (sign != 1.0 && sign != -1.0) throw std::invalid_argument("feh_core_lanczos_bounds: sign must be +1 or -1");
    fasteig_options o(p->prob->n());
    auto ex = p->prob->executor(o);
    fasteig::hip::lanczos_bounds(*ex, p->prob->device_matrix(o), sign, seed, *lower, *upper);
    return 0;
  });
}

int feh_core_thin_qr(feh_problem p, int64_t n, int k, double* X, int method) {
  return guarded([&] {
    if (!p || !X) throw std::invalid_argument("null argument");
    if (k <= 0 || n < k) throw std::invalid_argument("thin_qr_in_place: row count must be >= column count.");
    if (method != 0 && method != 1) throw std::invalid_argument("feh_core_thin_qr: method must be 0 or 1");
    fasteig_options o(p->prob->n());
    auto ex = p->prob->executor(o);
    namespace H = fasteig::hip;
    MatrixXd host(static_cast<int>(n), k);
    std::copy(X, X + n * k, host.data());
    H::Block B(*ex, n, k, FE_F64), tmp(*ex, n, k, FE_F64);
    H::upload(*ex, host, B.ref());
    H::Workspace ws(*ex, k, FE_GRAM_NATIVE);
    int rc = 0;
    if (method == 0) {
      if (H::cholesky_qr(ws, B, tmp) != 0) return 1;
    } else {
      rc = H::orthonormal_qr(ws, B, tmp) ? 2 : 0;
    }
    const MatrixXd out = H::download(*ex, B.ref());
    std::copy(out.data(), out.data() + n * k, X);
    return rc;
  });
}

int feh_core_orthonormalize_block(feh_problem p, int64_t n, int active, const double* basis, int b,
                                  const double* src, double rank_tol, double* out, int* rank) {
  return guarded([&] {
    if (!p || !src || !out || !rank || (active > 0 && !basis)) throw std::invalid_argument("null argument");
    if (b <= 0 || n < b || active < 0 || active > n)
      throw std::invalid_argument("orthonormalize_block: invalid basis or source dimensions.");
    fasteig_options o(p->prob->n());
    auto ex = p->prob->executor(o);
    namespace H = fasteig::hip;
    MatrixXd hs(static_cast<int>(n), b);
    std::copy(src, src + n * b, hs.data());
    H::Block S(*ex, n, b, FE_F64), V(*ex, n, std::max(active, 1), FE_F64);
    H::upload(*ex, hs, S.ref());
    if (active > 0) {
      MatrixXd hb(static_cast<int>(n), active);
      std::copy(basis, basis + n * active, hb.data());
      H::upload(*ex, hb, V.ref());
    }
    H::Workspace ws(*ex, std::max(b, active), FE_GRAM_NATIVE);
    H::Block Q = H::orthonormalize_block(ws, V.cols(0, active), S.ref(), rank_tol);
    *rank = Q.empty() ? 0 : Q.ncols();
    if (*rank > 0) {
      const MatrixXd h = H::download(*ex, Q.ref());
      std::copy(h.data(), h.data() + n * (*rank), out);
    }
    return 0;
  });
}

size_t feh_serve_footprint(int n, int64_t nnz, int nev) {
  return n > 0 && nev > 0 ? fasteig::hip::serve_footprint(n, nnz, nev) : 0;
}

int feh_server_create(int device, const char* options, feh_server* out) {
  return guarded([&] {
    if (!out) throw std::invalid_argument("null argument");
    const auto o = parse_serve_options(options);
    auto h = std::make_unique<feh_server_s>();
    h->srv = std::make_unique<fasteig::hip::BatchServer>(device, o);
    *out = h.release();
    return 0;
  });
}

int feh_server_destroy(feh_server s) {
  delete s;
  return 0;
}

int feh_server_submit(feh_server s, int64_t id, int n, int64_t nnz, const int* rowptr, const int* colind,
                      const double* vals, int nev, double tol) {
  return guarded([&] {
    if (!s || !rowptr || (nnz > 0 && (!colind || !vals)) || n <= 0 || nnz < 0)
      throw std::invalid_argument("feh_server_submit: bad argument");
    fasteig::hip::ServeRequest r;
    r.id = id;
    r.n = n;
    r.rowptr.assign(rowptr, rowptr + n + 1);
    r.colind.assign(colind, colind + nnz);
    r.values.assign(vals, vals + nnz);
    r.nev = nev;
    r.tol = tol;
    s->srv->submit(std::move(r));
    return 0;
  });
}

int feh_server_step(feh_server s) {
  return guarded([&] {
    if (!s) throw std::invalid_argument("null server");
    const int left = s->srv->step();
    for (auto& r : s->srv->take_results()) s->ready.push_back(std::move(r));
    return left;
  });
}

int feh_server_counts(feh_server s, int* active, int* queued, int* ready) {
  return guarded([&] {
    if (!s) throw std::invalid_argument("null server");
    if (active) *active = s->srv->active();
    if (queued) *queued = s->srv->queued();
    if (ready) *ready = static_cast<int>(s->ready.size());
    return 0;
  });
}

int feh_server_peek(feh_server s, int64_t* id, int* status, int* n, int* nev, int* cycles, int* fp64_cycles,
                    double* times) {
  return guarded([&] {
    if (!s) throw std::invalid_argument("null server");
    if (s->ready.empty()) return 0;
    const auto& r = s->ready.front();
    if (id) *id = r.id;
    if (status) *status = r.status;
    if (n) *n = r.n;
    if (nev) *nev = r.nev;
    if (cycles) *cycles = r.cycles;
    if (fp64_cycles) *fp64_cycles = r.fp64_cycles;
    if (times) {
      times[0] = r.queued_s;
      times[1] = r.admitted_s;
      times[2] = r.finished_s;
    }
    return 1;
  });
}

int feh_server_pop(feh_server s, double* eigenvalues, double* residuals, double* eigenvectors) {
  return guarded([&] {
    if (!s) throw std::invalid_argument("null server");
    if (s->ready.empty()) return 0;
    auto& r = s->ready.front();
    if (eigenvalues) std::copy(r.eigenvalues.begin(), r.eigenvalues.end(), eigenvalues);
    if (residuals) std::copy(r.residuals.begin(), r.residuals.end(), residuals);
    if (eigenvectors) std::copy(r.eigenvectors.begin(), r.eigenvectors.end(), eigenvectors);
    s->ready.pop_front();
    return 1;
  });
}

int feh_server_stats(feh_server s, int64_t* cycles, int64_t* admitted, int64_t* retired, int64_t* peak_active,
                     double* cycle_seconds, size_t* peak_bytes) {
  return guarded([&] {
    if (!s) throw std::invalid_argument("null server");
    const auto& st = s->srv->stats();
    if (cycles) *cycles = st.cycles;
    if (admitted) *admitted = st.admitted;
    if (retired) *retired = st.retired;
    if (peak_active) *peak_active = st.peak_active;
    if (cycle_seconds) *cycle_seconds = st.device_s;
    if (peak_bytes) *peak_bytes = st.peak_bytes;
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

int feh_problem_product_tuning(feh_problem p, char* buf, int len) {
  return guarded([&] {
    if (!p || !buf || len <= 0) throw std::invalid_argument("null argument");
    std::ostringstream os;
    static const char* kinds[] = {"gram_sym", "gram", "gram_weighted", "gemm", "trmm", "gram_rr"};
    if (auto ex = p->prob->current_executor()) {
      os.precision(6);
      for (const auto& [key, t] : ex->tunings()) {
        const int kind = std::get<0>(key);
        os << (kind >= 0 && kind < 6 ? kinds[kind] : "?") << ' ' << std::get<1>(key) << ' ' << std::get<2>(key) << ' '
           << std::get<3>(key);
        for (int i = 0; i < 3; ++i) os << ' ' << t.ms[i];
        for (int i = 0; i < 3; ++i) os << ' ' << t.err[i];
        os << '\n';
      }
    }
    const std::string s = os.str();
    std::strncpy(buf, s.c_str(), static_cast<size_t>(len - 1));
    buf[len - 1] = '\0';
    return static_cast<int>(s.size());
  });
}

}  // extern "C"
