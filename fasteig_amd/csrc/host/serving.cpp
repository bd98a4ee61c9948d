// Continuous-batching solve server (see serving.hpp).
#include "serving.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>
#include <stdexcept>

namespace fasteig {
namespace hip {

namespace {

constexpr int kRes = 256;  // doubles per slot in the results area

int subspace_cols(int /*n*/, int nev) { return std::min<int>(FE_GRP_MAXCOLS, (2 * nev + 3) / 4 * 4); }

int64_t chunks(int n) { return (static_cast<int64_t>(n) + FE_GRP_CHUNK - 1) / FE_GRP_CHUNK; }

size_t align256(size_t b) { return (b + 255) / 256 * 256; }

// table kinds of one cycle, in launch order after the filter steps
enum Tab { T_G1, T_C1, T_M1, T_G2, T_C2, T_M2, T_A, T_G3, T_E, T_M3, T_M4, T_R, T_COUNT };

size_t item_bytes(int t) {
  switch (t) {
    case T_G1: case T_G2: case T_G3: return sizeof(fe_grp_gram);
    case T_C1: case T_C2: case T_E: return sizeof(fe_grp_small);
    case T_M1: case T_M2: case T_M3: case T_M4: return sizeof(fe_grp_gemm);
    case T_A: return sizeof(fe_grp_spmm);
    default: return sizeof(fe_grp_resid);
  }
}

}  // namespace

size_t serve_footprint(int n, int64_t nnz, int nev) {
  const int m = subspace_cols(n, nev);
  const size_t csr = static_cast<size_t>(n + 1) * 4 + static_cast<size_t>(nnz) * (4 + 8 + 4);
  const size_t blocks = 4 * align256(static_cast<size_t>(n) * m * 8);
  const size_t small = 2 * align256(sizeof(double) * FE_GRP_MAXCOLS * FE_GRP_MAXCOLS) +
                       align256(sizeof(double) * chunks(n) * m * m) + 256;
  return csr + blocks + small + 4096;
}

struct BatchServer::Slot {
  ServeRequest req;
  int index = 0;
  double queued_s = 0.0, admitted_s = 0.0;
  fe_csr A = nullptr;
  const int* rp = nullptr;
  const int* ci = nullptr;
  const double* v64 = nullptr;
  const float* v32 = nullptr;
  void* mem = nullptr;
  size_t bytes = 0;
  int n = 0, m = 0, nev = 0, prec = FE_F32;
  int64_t ld = 0;
  void* B[4] = {};
  double* G = nullptr;
  double* C = nullptr;
  double* part = nullptr;
  int* info = nullptr;
  int x = 0;            // buffer holding the current basis
  bool init = true;     // first cycle: no filter
  int cycles = 0, fp64_cycles = 0;
  double upper = 0.0, cut = 0.0, lo = 0.0;
  double best = std::numeric_limits<double>::infinity();
  int stall = 0;
  // roles of this cycle
  int f = 0, g = 0, h = 0, s = 0;
  fe_ctx ctx = nullptr;  // owner of A and mem
  Slot() = default;
  Slot(const Slot&) = delete;
  Slot& operator=(const Slot&) = delete;
  ~Slot() {
    if (A) fe_csr_destroy(A);
    if (mem && ctx) fe_free(ctx, mem);
  }
};

namespace {
// A request's CSR is usable: sizes agree, rowptr starts at 0, never
// decreases and ends at nnz, and every column index lies in [0, n).
bool csr_valid(const ServeRequest& r) {
  const int64_t nnz = static_cast<int64_t>(r.values.size());
  if (r.n <= 0 || static_cast<int64_t>(r.rowptr.size()) != static_cast<int64_t>(r.n) + 1 ||
      static_cast<int64_t>(r.colind.size()) != nnz || r.rowptr[0] != 0 || r.rowptr[r.n] != nnz)
    return false;
  for (int i = 0; i < r.n; ++i)
    if (r.rowptr[i + 1] < r.rowptr[i]) return false;
  for (int c : r.colind)
    if (c < 0 || c >= r.n) return false;
  return true;
}
}  // namespace

BatchServer::BatchServer(int device, const ServeOptions& opts) : o_(opts) {
  if (o_.max_slots < 1 || o_.max_slots > 4096) throw std::invalid_argument("serve: max_slots must be 1..4096");
  if (o_.degree < 1 || o_.degree > 64) throw std::invalid_argument("serve: degree must be 1..64");
  ex_ = std::make_unique<HipExecutor>(device);
  t0_ = std::chrono::steady_clock::now();
  if (o_.hbm_budget == 0) {
    int cus = 0;
    size_t hbm = 0;
    check(fe_ctx_info(ex_->ctx(), &cus, &hbm, nullptr, 0), "fe_ctx_info");
    o_.hbm_budget = static_cast<size_t>(0.75 * static_cast<double>(hbm));
  }
  slots_.resize(o_.max_slots);
  // table layout: degree filter tables, then T_COUNT tables
  size_t off = 0;
  const int ntab = o_.degree + T_COUNT;
  for (int t = 0; t < ntab; ++t) {
    tab_off_.push_back(off);
    const size_t ib = t < o_.degree ? sizeof(fe_grp_spmm) : item_bytes(t - o_.degree);
    off += align256(sizeof(fe_grp_header) + ib * o_.max_slots);
  }
  tab_bytes_ = off;
  check(fe_alloc(ex_->ctx(), tab_bytes_, &tab_dev_), "fe_alloc(tables)");
  check(fe_host_alloc(tab_bytes_, &tab_host_), "fe_host_alloc(tables)");
  std::memset(tab_host_, 0, tab_bytes_);
  void* r = nullptr;
  check(fe_alloc(ex_->ctx(), sizeof(double) * kRes * o_.max_slots, &r), "fe_alloc(results)");
  res_dev_ = static_cast<double*>(r);
  res_host_.assign(static_cast<size_t>(kRes) * o_.max_slots, 0.0);
}

BatchServer::~BatchServer() {
  if (ex_) {
    fe_ctx_sync(ex_->ctx());
    slots_.clear();  // each slot frees its matrix and blocks
    if (graph_exec_) fe_graph_destroy(static_cast<fe_graph>(graph_exec_));
    if (tab_dev_) fe_free(ex_->ctx(), tab_dev_);
    if (res_dev_) fe_free(ex_->ctx(), res_dev_);
  }
  if (tab_host_) fe_host_free(tab_host_);
}

double BatchServer::now() const {
  return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0_).count();
}

void BatchServer::submit(ServeRequest&& r) { queue_.emplace_back(std::move(r), now()); }

std::vector<ServeResult> BatchServer::take_results() {
  std::vector<ServeResult> out;
  out.swap(done_);
  return out;
}

void BatchServer::admit() {
  while (!queue_.empty() && slots_used_ < slots_.size()) {
    ServeRequest& r = queue_.front().first;
    const double queued = queue_.front().second;
    const int64_t nnz = static_cast<int64_t>(r.values.size());
    const int m = subspace_cols(r.n, r.nev);
    const bool valid = r.nev >= 1 && 2 * r.nev <= FE_GRP_MAXCOLS && m < r.n && r.tol > 0.0 && csr_valid(r);
    const size_t fp = valid ? serve_footprint(r.n, nnz, r.nev) : 0;
    auto reject = [&](ServeRequest& q, double qs) {  // status 2: can never run here
      ServeResult res;
      res.id = q.id;
      res.status = 2;
      res.n = q.n;
      res.nev = q.nev;
      res.queued_s = qs;
      res.admitted_s = res.finished_s = now();
      done_.push_back(std::move(res));
    };
    if (!valid || fp > o_.hbm_budget) {
      reject(r, queued);
      queue_.pop_front();
      continue;
    }
    if (used_bytes_ + fp > o_.hbm_budget) break;  // wait for memory (arrival order)
    size_t idx = 0;
    while (slots_[idx]) ++idx;
    auto s = std::make_unique<Slot>();
    s->req = std::move(r);
    queue_.pop_front();
    s->index = static_cast<int>(idx);
    s->queued_s = queued;
    s->admitted_s = now();
    s->n = s->req.n;
    s->nev = s->req.nev;
    s->m = m;
    s->ld = m;
    s->bytes = fp;
    s->ctx = ex_->ctx();
    const size_t blk = align256(static_cast<size_t>(s->n) * m * 8);
    const size_t sm = align256(sizeof(double) * FE_GRP_MAXCOLS * FE_GRP_MAXCOLS);
    const size_t pb = align256(sizeof(double) * chunks(s->n) * m * m);
    if (fe_csr_create(ex_->ctx(), s->n, nnz, s->req.rowptr.data(), s->req.colind.data(), s->req.values.data(),
                      &s->A) != 0 ||
        fe_csr_arrays(s->A, &s->rp, &s->ci, &s->v64, &s->v32) != 0 ||
        fe_alloc(ex_->ctx(), 4 * blk + 2 * sm + pb + 256, &s->mem) != 0) {
      // the device could not take it (the slot's destructor frees what was made)
      reject(s->req, s->queued_s);
      continue;
    }
    char* base = static_cast<char*>(s->mem);
    for (int b = 0; b < 4; ++b) s->B[b] = base + b * blk;
    s->G = reinterpret_cast<double*>(base + 4 * blk);
    s->C = reinterpret_cast<double*>(base + 4 * blk + sm);
    s->part = reinterpret_cast<double*>(base + 4 * blk + 2 * sm);
    s->info = reinterpret_cast<int*>(base + 4 * blk + 2 * sm + pb);
    // Gershgorin upper bound of the spectrum
    double up = -std::numeric_limits<double>::infinity();
    for (int i = 0; i < s->n; ++i) {
      double d = 0.0, off = 0.0;
      for (int p = s->req.rowptr[i]; p < s->req.rowptr[i + 1]; ++p) {
        if (s->req.colind[p] == i) d += s->req.values[p];
        else off += std::fabs(s->req.values[p]);
      }
      up = std::max(up, d + off);
    }
    s->upper = up;
    s->prec = FE_F32;
    s->x = 0;
    // seeded start block (counter-based: identical for the same seed and id)
    check(fe_fill_random(ex_->ctx(), FE_F32, s->n, m, s->B[0], s->ld,
                         o_.seed * 0x9E3779B97F4A7C15ull + static_cast<uint64_t>(s->req.id)),
          "fe_fill_random");
    s->req.rowptr.clear();
    s->req.rowptr.shrink_to_fit();
    s->req.colind.clear();
    s->req.colind.shrink_to_fit();
    s->req.values.clear();
    s->req.values.shrink_to_fit();
    used_bytes_ += fp;
    ++slots_used_;
    ++stats_.admitted;
    stats_.peak_active = std::max<int64_t>(stats_.peak_active, static_cast<int64_t>(slots_used_));
    stats_.peak_bytes = std::max(stats_.peak_bytes, used_bytes_);
    slots_[idx] = std::move(s);
  }
}

void BatchServer::build_tables(std::vector<Slot*>& act) {
  char* hb = static_cast<char*>(tab_host_);
  const char* db = static_cast<const char*>(tab_dev_);
  (void)db;
  const int d = o_.degree;
  // filter steps
  std::vector<std::vector<fe_grp_spmm>> fs(d);
  for (Slot* s : act) {
    const int R[3] = {s->x, (s->x + 1) % 4, (s->x + 2) % 4};
    s->s = (s->x + 3) % 4;
    if (s->init) {
      s->f = R[0];
      s->g = R[1];
      s->h = R[2];
      continue;
    }
    const double e = (s->upper - s->cut) / 2.0, c = (s->upper + s->cut) / 2.0;
    const double sigma1 = e / (s->lo - c);
    double sigma = sigma1;
    const void* vals = s->prec == FE_F64 ? static_cast<const void*>(s->v64) : static_cast<const void*>(s->v32);
    for (int j = 1; j <= d; ++j) {
      fe_grp_spmm it{};
      it.n = s->n;
      it.k = s->m;
      it.prec = s->prec;
      it.rowptr = s->rp;
      it.colind = s->ci;
      it.vals = vals;
      it.X = s->B[R[(j - 1) % 3]];
      it.Y = s->B[R[j % 3]];
      it.ldx = it.ldy = it.ldz = s->ld;
      if (j == 1) {
        it.alpha = sigma1 / e;
        it.beta = -c * sigma1 / e;
        it.Z = nullptr;
        it.gamma = 0.0;
      } else {
        const double sn = 1.0 / (2.0 / sigma1 - sigma);
        it.alpha = 2.0 * sn / e;
        it.beta = -2.0 * sn * c / e;
        it.Z = s->B[R[(j - 2) % 3]];
        it.gamma = -sigma * sn;
        sigma = sn;
      }
      fs[j - 1].push_back(it);
    }
    s->f = R[d % 3];
    s->g = R[(d + 1) % 3];
    s->h = R[(d + 2) % 3];
  }
  for (int j = 0; j < d; ++j) {
    auto* hdr = reinterpret_cast<fe_grp_header*>(hb + tab_off_[j]);
    auto* items = reinterpret_cast<fe_grp_spmm*>(hb + tab_off_[j] + sizeof(fe_grp_header));
    std::copy(fs[j].begin(), fs[j].end(), items);
    check(fe_grp_spmm_plan(hdr, items, static_cast<int>(fs[j].size())), "fe_grp_spmm_plan");
  }
  auto tab = [&](int t) { return hb + tab_off_[d + t]; };
  auto items_of = [&](int t) { return tab(t) + sizeof(fe_grp_header); };
  const int na = static_cast<int>(act.size());
  auto* g1 = reinterpret_cast<fe_grp_gram*>(items_of(T_G1));
  auto* g2 = reinterpret_cast<fe_grp_gram*>(items_of(T_G2));
  auto* g3 = reinterpret_cast<fe_grp_gram*>(items_of(T_G3));
  auto* c1 = reinterpret_cast<fe_grp_small*>(items_of(T_C1));
  auto* c2 = reinterpret_cast<fe_grp_small*>(items_of(T_C2));
  auto* ee = reinterpret_cast<fe_grp_small*>(items_of(T_E));
  auto* m1 = reinterpret_cast<fe_grp_gemm*>(items_of(T_M1));
  auto* m2 = reinterpret_cast<fe_grp_gemm*>(items_of(T_M2));
  auto* m3 = reinterpret_cast<fe_grp_gemm*>(items_of(T_M3));
  auto* m4 = reinterpret_cast<fe_grp_gemm*>(items_of(T_M4));
  auto* aa = reinterpret_cast<fe_grp_spmm*>(items_of(T_A));
  auto* rr = reinterpret_cast<fe_grp_resid*>(items_of(T_R));
  for (int i = 0; i < na; ++i) {
    Slot* s = act[i];
    double* res = res_dev_ + static_cast<size_t>(s->index) * kRes;
    auto gram = [&](const void* X, const void* Y) {
      fe_grp_gram it{};
      it.n = s->n;
      it.m = it.p = s->m;
      it.prec = s->prec;
      it.X = X;
      it.Y = Y;
      it.ldx = it.ldy = s->ld;
      it.part = s->part;
      it.G = s->G;
      it.ldg = s->m;
      return it;
    };
    auto small = [&](double* w) {
      fe_grp_small it{};
      it.m = s->m;
      it.A = s->G;
      it.lda = s->m;
      it.C = s->C;
      it.ldc = s->m;
      it.w = w;
      it.info = s->info;
      return it;
    };
    auto gemm = [&](const void* X, void* Y) {
      fe_grp_gemm it{};
      it.n = s->n;
      it.m = it.p = s->m;
      it.prec = s->prec;
      it.X = X;
      it.ldx = s->ld;
      it.C = s->C;
      it.ldc = s->m;
      it.Y = Y;
      it.ldy = s->ld;
      return it;
    };
    void* F = s->B[s->f];
    void* S = s->B[s->s];
    void* W = s->B[s->g];
    void* Wn = s->B[s->h];
    g1[i] = gram(F, F);
    c1[i] = small(nullptr);
    m1[i] = gemm(F, S);
    g2[i] = gram(S, S);
    c2[i] = small(nullptr);
    m2[i] = gemm(S, F);
    fe_grp_spmm a{};
    a.n = s->n;
    a.k = s->m;
    a.prec = s->prec;
    a.rowptr = s->rp;
    a.colind = s->ci;
    a.vals = s->prec == FE_F64 ? static_cast<const void*>(s->v64) : static_cast<const void*>(s->v32);
    a.X = F;
    a.Y = W;
    a.ldx = a.ldy = a.ldz = s->ld;
    a.alpha = 1.0;
    aa[i] = a;
    g3[i] = gram(F, W);
    ee[i] = small(res);  // theta -> results area
    m3[i] = gemm(F, S);
    m4[i] = gemm(W, Wn);
    fe_grp_resid r{};
    r.n = s->n;
    r.p = s->nev;
    r.prec = s->prec;
    r.Y = S;
    r.W = Wn;
    r.ldy = r.ldw = s->ld;
    r.theta = res;
    r.part = s->part;
    r.rnorm = res + 64;
    r.ynorm = res + 128;
    rr[i] = r;
  }
  auto H = [&](int t) { return reinterpret_cast<fe_grp_header*>(tab(t)); };
  check(fe_grp_gram_plan(H(T_G1), g1, na), "gram plan");
  check(fe_grp_gram_plan(H(T_G2), g2, na), "gram plan");
  check(fe_grp_gram_plan(H(T_G3), g3, na), "gram plan");
  check(fe_grp_small_plan(H(T_C1), c1, na), "small plan");
  check(fe_grp_small_plan(H(T_C2), c2, na), "small plan");
  check(fe_grp_small_plan(H(T_E), ee, na), "small plan");
  check(fe_grp_gemm_plan(H(T_M1), m1, na), "gemm plan");
  check(fe_grp_gemm_plan(H(T_M2), m2, na), "gemm plan");
  check(fe_grp_gemm_plan(H(T_M3), m3, na), "gemm plan");
  check(fe_grp_gemm_plan(H(T_M4), m4, na), "gemm plan");
  check(fe_grp_spmm_plan(H(T_A), aa, na), "spmm plan");
  check(fe_grp_resid_plan(H(T_R), rr, na), "resid plan");
}

// the launches of one cycle, in a fixed order over fixed table addresses
static void launch_cycle(fe_ctx ctx, const char* db, const std::vector<size_t>& off, int d, int ms) {
  for (int j = 0; j < d; ++j) check(fe_grp_spmm_run(ctx, db + off[j]), "fe_grp_spmm_run");
  auto t = [&](int k) { return db + off[d + k]; };
  check(fe_grp_gram_run(ctx, t(T_G1), ms), "gram");
  check(fe_grp_cholinv_run(ctx, t(T_C1), ms), "cholinv");
  check(fe_grp_gemm_run(ctx, t(T_M1)), "gemm");
  check(fe_grp_gram_run(ctx, t(T_G2), ms), "gram");
  check(fe_grp_cholinv_run(ctx, t(T_C2), ms), "cholinv");
  check(fe_grp_gemm_run(ctx, t(T_M2)), "gemm");
  check(fe_grp_spmm_run(ctx, t(T_A)), "spmm");
  check(fe_grp_gram_run(ctx, t(T_G3), ms), "gram");
  check(fe_grp_eig_run(ctx, t(T_E), ms), "eig");
  check(fe_grp_gemm_run(ctx, t(T_M3)), "gemm");
  check(fe_grp_gemm_run(ctx, t(T_M4)), "gemm");
  check(fe_grp_resid_run(ctx, t(T_R), ms), "resid");
}

void BatchServer::capture_graph() {
  fe_ctx ctx = ex_->ctx();
  check(fe_graph_begin(ctx), "fe_graph_begin");
  try {
    launch_cycle(ctx, static_cast<const char*>(tab_dev_), tab_off_, o_.degree, o_.max_slots);
  } catch (...) {
    fe_graph g = nullptr;
    fe_graph_end(ctx, &g);
    fe_graph_destroy(g);
    throw;
  }
  fe_graph g = nullptr;
  check(fe_graph_end(ctx, &g), "fe_graph_end");
  graph_exec_ = g;
}

void BatchServer::retire(Slot& s, int status) {
  ServeResult r;
  r.id = s.req.id;
  r.status = status;
  r.cycles = s.cycles;
  r.fp64_cycles = s.fp64_cycles;
  r.n = s.n;
  r.nev = s.nev;
  const double* res = res_host_.data() + static_cast<size_t>(s.index) * kRes;
  r.eigenvalues.assign(res, res + s.nev);
  r.residuals.resize(s.nev);
  for (int j = 0; j < s.nev; ++j) r.residuals[j] = res[64 + j] / ((1.0 + std::fabs(res[j])) * res[128 + j]);
  if (o_.keep_vectors) {
    r.eigenvectors.resize(static_cast<size_t>(s.n) * s.nev);
    check(fe_download_colmajor(ex_->ctx(), s.n, s.nev, s.B[s.x], s.prec, s.ld, r.eigenvectors.data()),
          "fe_download_colmajor");
  }
  r.queued_s = s.queued_s;
  r.admitted_s = s.admitted_s;
  r.finished_s = now();
  done_.push_back(std::move(r));
  fe_csr_destroy(s.A);
  s.A = nullptr;
  fe_free(ex_->ctx(), s.mem);
  s.mem = nullptr;
  used_bytes_ -= s.bytes;
  --slots_used_;
  ++stats_.retired;
}

void BatchServer::run_cycle() {
  std::vector<Slot*> act;
  for (auto& s : slots_)
    if (s) act.push_back(s.get());
  if (act.empty()) return;
  const double t = now();
  fe_ctx ctx = ex_->ctx();
  build_tables(act);
  check(fe_memcpy_h2d_async(ctx, tab_dev_, tab_host_, tab_bytes_), "table upload");
  if (o_.use_graph) {
    if (!graph_exec_) capture_graph();
    check(fe_graph_launch(ctx, static_cast<fe_graph>(graph_exec_)), "fe_graph_launch");
  } else {
    launch_cycle(ctx, static_cast<const char*>(tab_dev_), tab_off_, o_.degree, o_.max_slots);
  }
  check(fe_memcpy_d2h_async(ctx, res_host_.data(), res_dev_, sizeof(double) * res_host_.size()), "results");
  ex_->synchronize();
  ++stats_.cycles;
  for (Slot* s : act) {
    s->cycles += 1;
    if (s->prec == FE_F64) s->fp64_cycles += 1;
    s->x = s->s;  // the rotated basis
    s->init = false;
    const double* res = res_host_.data() + static_cast<size_t>(s->index) * kRes;
    double mx = 0.0;
    for (int j = 0; j < s->nev; ++j) {
      const double rj = res[64 + j] / ((1.0 + std::fabs(res[j])) * res[128 + j]);
      mx = std::max(mx, std::isfinite(rj) ? rj : std::numeric_limits<double>::infinity());
    }
    if (mx <= s->req.tol) {
      retire(*s, 0);
      slots_[s->index].reset();
      continue;
    }
    if (s->cycles >= o_.maxiter) {
      retire(*s, 1);
      slots_[s->index].reset();
      continue;
    }
    // FP32 -> FP64 once the FP32 cycles have done their part
    if (s->prec == FE_F32) {
      if (mx < 0.5 * s->best) {
        s->best = mx;
        s->stall = 0;
      } else {
        ++s->stall;
      }
      if (mx <= o_.switch_tol || s->stall >= 2) {
        const int t64 = (s->x + 1) % 4;
        check(fe_copy2d(ctx, s->n, s->m, s->B[s->x], FE_F32, s->ld, s->B[t64], FE_F64, s->ld), "fe_copy2d");
        s->x = t64;
        s->prec = FE_F64;
      }
    }
    // filter interval for the next cycle
    s->lo = res[0];
    s->cut = res[s->m - 1];
    if (!(s->cut < s->upper)) s->cut = s->lo + 0.5 * (s->upper - s->lo);
    if (!(s->lo < s->cut)) s->lo = s->cut - 1e-3 * (s->upper - s->cut);
  }
  stats_.device_s += now() - t;
}

int BatchServer::step() {
  admit();
  run_cycle();
  return static_cast<int>(slots_used_ + queue_.size());
}

}  // namespace hip
}  // namespace fasteig
