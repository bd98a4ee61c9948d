// fasteig::hip host classes over the libfe_hip.so C ABI.
#include "hip_backend.hpp"

#include <cmath>
#include <functional>
#include <limits>

#include <random>
#include <stdexcept>

namespace fasteig {
namespace hip {

int check(int rc, const char* what) {
  if (rc >= 0) return rc;
  const std::string msg = std::string(what) + ": " + fe_last_error();
  if (rc == -2) throw std::invalid_argument(msg);
  throw std::runtime_error(msg);
}

static int64_t padded_ld(int cols) { return cols <= 0 ? 4 : (static_cast<int64_t>(cols) + 3) / 4 * 4; }

// ---------------------------------------------------------------- Block
Block::Block(HipExecutor& ex, int64_t rows, int cols, int prec)
    : ex_(&ex), rows_(rows), cols_(cols), ld_(padded_ld(cols)), prec_(prec) {
  if (rows < 0 || cols < 0) throw std::invalid_argument("Block: negative dimensions");
  const size_t bytes = static_cast<size_t>(rows) * static_cast<size_t>(ld_) * (prec == FE_F64 ? 8 : 4);
  check(fe_alloc(ex.ctx(), bytes, &ptr_), "fe_alloc(block)");
}

Block::~Block() {
  if (ptr_ && ex_) fe_free(ex_->ctx(), ptr_);
}

Block& Block::operator=(Block&& o) noexcept {
  if (this != &o) {
    if (ptr_ && ex_) fe_free(ex_->ctx(), ptr_);
    ex_ = o.ex_;
    ptr_ = o.ptr_;
    rows_ = o.rows_;
    cols_ = o.cols_;
    ld_ = o.ld_;
    prec_ = o.prec_;
    o.ptr_ = nullptr;
    o.ex_ = nullptr;
  }
  return *this;
}

void Block::swap(Block& o) noexcept {
  std::swap(ex_, o.ex_);
  std::swap(ptr_, o.ptr_);
  std::swap(rows_, o.rows_);
  std::swap(cols_, o.cols_);
  std::swap(ld_, o.ld_);
  std::swap(prec_, o.prec_);
}

// ---------------------------------------------------------------- Small
Small::Small(HipExecutor& ex, int rows, int cols) : ex_(&ex), r_(rows), c_(cols) {
  void* p = nullptr;
  check(fe_alloc(ex.ctx(), sizeof(double) * static_cast<size_t>(std::max(rows, 1)) * std::max(cols, 1), &p),
        "fe_alloc(small)");
  ptr_ = static_cast<double*>(p);
}

Small::~Small() {
  if (ptr_ && ex_) fe_free(ex_->ctx(), ptr_);
}

Small& Small::operator=(Small&& o) noexcept {
  if (this != &o) {
    if (ptr_ && ex_) fe_free(ex_->ctx(), ptr_);
    ex_ = o.ex_;
    ptr_ = o.ptr_;
    r_ = o.r_;
    c_ = o.c_;
    o.ptr_ = nullptr;
    o.ex_ = nullptr;
  }
  return *this;
}

MatrixXd Small::to_host() const {
  MatrixXd M(r_, c_);
  check(fe_memcpy_d2h(ex_->ctx(), M.data(), ptr_, sizeof(double) * M.size()), "Small::to_host");
  return M;
}

void Small::from_host(const MatrixXd& M) {
  if (M.rows() != r_ || M.cols() != c_) throw std::invalid_argument("Small::from_host: shape mismatch");
  check(fe_memcpy_h2d(ex_->ctx(), ptr_, M.data(), sizeof(double) * M.size()), "Small::from_host");
}

// ---------------------------------------------------------------- CSR
HipCsrMatrix::HipCsrMatrix(HipExecutor& ex, int n, const std::vector<int>& rowptr, const std::vector<int>& colind,
                           const std::vector<double>& vals)
    : ex_(ex), n_(n), nnz_(static_cast<int64_t>(vals.size())) {
  if (static_cast<int>(rowptr.size()) != n + 1 || colind.size() != vals.size())
    throw std::invalid_argument("HipCsrMatrix: inconsistent CSR arrays");
  check(fe_csr_create(ex.ctx(), n, nnz_, rowptr.data(), colind.data(), vals.data(), &h_), "fe_csr_create");
}

HipCsrMatrix::~HipCsrMatrix() {
  if (h_) fe_csr_destroy(h_);
}

void HipCsrMatrix::use_sell(int C, int sigma) {
  if (!sell_built_) {
    check(fe_csr_build_sell(ex_.ctx(), h_, C, sigma), "fe_csr_build_sell");
    sell_built_ = true;
  }
  check(fe_csr_use(h_, FE_FMT_SELL), "fe_csr_use");
  sell_ = true;
  blocked_ = false;
  tiled_ = false;
}

bool HipCsrMatrix::use_blocked() {
  if (blocked_state_ == 0) {
    double reuse = 0.0;
    blocked_state_ = check(fe_csr_build_blocked(ex_.ctx(), h_, 4, &reuse), "fe_csr_build_blocked") == 0 ? 1 : -1;
    // grid stencils: strip-by-strip block order for L2 reuse across planes
    if (blocked_state_ > 0) check(fe_csr_block_schedule(ex_.ctx(), h_, -1, 0), "fe_csr_block_schedule");
  }
  if (blocked_state_ < 0) return false;
  check(fe_csr_use(h_, FE_FMT_BLOCKED), "fe_csr_use");
  blocked_ = true;
  sell_ = false;
  tiled_ = false;
  return true;
}

bool HipCsrMatrix::use_tiled() {
  if (tiled_state_ == 0) {
    double loads = 0.0;
    tiled_state_ = check(fe_csr_build_tiled(ex_.ctx(), h_, &loads), "fe_csr_build_tiled") == 0 ? 1 : -1;
  }
  if (tiled_state_ < 0) return false;
  check(fe_csr_use(h_, FE_FMT_TILED), "fe_csr_use");
  tiled_ = true;
  blocked_ = false;
  sell_ = false;
  return true;
}

void HipCsrMatrix::use_csr() {
  check(fe_csr_use(h_, FE_FMT_CSR), "fe_csr_use");
  sell_ = false;
  blocked_ = false;
  tiled_ = false;
}

// ---------------------------------------------------------------- executor
HipExecutor::HipExecutor(int device) : device_(device) { check(fe_ctx_create(device, &ctx_), "fe_ctx_create"); }

HipExecutor::~HipExecutor() {
  random_cache_.clear();
  if (ctx_) fe_ctx_destroy(ctx_);
}

void HipExecutor::synchronize() const { check(fe_ctx_sync(ctx_), "fe_ctx_sync"); }

const Block& HipExecutor::random_block(int64_t rows, int cols, unsigned seed) {
  auto key = std::make_tuple(rows, cols, seed);
  auto it = random_cache_.find(key);
  if (it != random_cache_.end()) return it->second;
  std::mt19937 rng(seed);
  std::uniform_real_distribution<double> dist(-1.0, 1.0);
  MatrixXd host(static_cast<int>(rows), cols);
  for (int j = 0; j < cols; ++j)
    for (int64_t i = 0; i < rows; ++i) host(static_cast<int>(i), j) = dist(rng);
  Block b(*this, rows, cols, FE_F64);
  upload(*this, host, b.ref());
  return random_cache_.emplace(key, std::move(b)).first->second;
}

void HipExecutor::clear_cache() { random_cache_.clear(); }

// ---------------------------------------------------------------- primitives
void spmm(HipExecutor& ex, const HipCsrMatrix& A, const BlockRef& X, const BlockRef& Y, double alpha, double beta,
          const BlockRef* Z, double gamma) {
  if (X.prec != Y.prec || (Z && Z->prec != X.prec)) throw std::invalid_argument("spmm: precision mismatch");
  if (X.cols != Y.cols || X.rows != A.rows() || Y.rows != A.rows())
    throw std::invalid_argument("spmm: dimension mismatch");
  check(fe_spmm(ex.ctx(), A.handle(), X.prec, X.cols, X.ptr, X.ld, Y.ptr, Y.ld, alpha, beta, Z ? Z->ptr : nullptr,
                Z ? Z->ld : 0, gamma),
        "fe_spmm");
}

namespace {

constexpr int kModes[3] = {FE_GRAM_NATIVE, FE_GRAM_BF16X3, FE_GRAM_BF16X6};
enum ProductKind { kGramSym = 0, kGram = 1, kGramW = 2, kGemm = 3, kTrmm = 4, kGramSymProd = 5 };
constexpr int kTuneEvent0 = 14, kTuneEvent1 = 15;  // context event slots used by the tuning
constexpr int64_t kTuneRowsGram = 65536;            // row sample of a Gram's error
constexpr int64_t kTuneRowsTall = 8192;             // row sample of a tall product's error

// Time of one call (after one untimed call that absorbs first-use costs).
float time_call(HipExecutor& ex, const std::function<void()>& f) {
  f();
  check(fe_event_record(ex.ctx(), kTuneEvent0), "fe_event_record");
  f();
  check(fe_event_record(ex.ctx(), kTuneEvent1), "fe_event_record");
  float ms = 0.0f;
  check(fe_event_elapsed_ms(ex.ctx(), kTuneEvent0, kTuneEvent1, &ms), "fe_event_elapsed_ms");
  return ms;
}

template <typename T>
double max_rel_diff(const std::vector<T>& a, const std::vector<double>& ref) {
  double num = 0.0, den = 0.0;
  for (size_t i = 0; i < a.size(); ++i) {
    const double d = std::fabs(static_cast<double>(a[i]) - ref[i]);
    num = std::isfinite(d) ? std::max(num, d) : std::numeric_limits<double>::infinity();
    den = std::max(den, std::fabs(ref[i]));
  }
  return den > 0.0 ? num / den : num;
}

// The fastest mode whose measured error (against the FP64 product of the
// same FP32 data) is within need, or within twice the native product's own
// error: a mode as accurate as native FP32 qualifies for any need that
// native FP32 products are used for.
int choose(const HipExecutor::ProductTuning& t, double need) {
  int best = 0;
  for (int i = 1; i < 3; ++i)
    if ((t.err[i] <= need || t.err[i] <= 2.0 * t.err[0]) && t.ms[i] < t.ms[best]) best = i;
  return kModes[best];
}

// FP64 copy of the first rows of an FP32 block (the tuning reference)
Block fp64_rows(HipExecutor& ex, const BlockRef& X, int64_t rows) {
  Block b(ex, rows, X.cols, FE_F64);
  BlockRef src = X;
  src.rows = rows;
  check(fe_copy2d(ex.ctx(), rows, X.cols, src.ptr, src.prec, src.ld, b.data(), FE_F64, b.ld()), "fe_copy2d");
  return b;
}

}  // namespace

void gram(HipExecutor& ex, int mode, const BlockRef& X, const BlockRef& Y, double* G, int64_t ldg, const double* w,
          double need, bool symmetric) {
  if (X.prec != Y.prec || X.rows != Y.rows) throw std::invalid_argument("gram: operand mismatch");
  if (symmetric && X.cols != Y.cols) throw std::invalid_argument("gram: a symmetric product needs m == p");
  const bool same = X.ptr == Y.ptr && X.cols == Y.cols && X.ld == Y.ld;
  const int flag = (symmetric && !same) ? FE_GRAM_SYMMETRIC : 0;
  auto run = [&](int md, const BlockRef& x, const BlockRef& y) {
    check(fe_gram(ex.ctx(), md | flag, x.prec, x.rows, x.cols, y.cols, x.ptr, x.ld, y.ptr, y.ld, w, G, ldg),
          "fe_gram");
  };
  if (X.prec == FE_F64 || mode == FE_GRAM_NATIVE || (mode == kProductAuto && !(need > 0.0)))
    return run(FE_GRAM_NATIVE, X, Y);
  if (mode != kProductAuto) return run(mode, X, Y);
  const int kind = w ? kGramW : same ? kGramSym : flag ? kGramSymProd : kGram;
  const HipExecutor::TuningKey key{kind, X.rows, X.cols, Y.cols};
  const HipExecutor::ProductTuning* t = ex.tuning(key);
  if (!t) {
    HipExecutor::ProductTuning nt;
    const int64_t ns = std::min<int64_t>(X.rows, kTuneRowsGram);
    BlockRef xs = X, ys = Y;
    xs.rows = ys.rows = ns;
    auto download = [&] {
      std::vector<double> g(static_cast<size_t>(X.cols) * Y.cols);
      for (int c = 0; c < Y.cols; ++c)
        check(fe_memcpy_d2h(ex.ctx(), g.data() + static_cast<size_t>(c) * X.cols, G + static_cast<size_t>(c) * ldg,
                            sizeof(double) * X.cols),
              "d2h(gram)");
      return g;
    };
    std::vector<double> ref;
    {
      Block x64 = fp64_rows(ex, X, ns);
      Block y64 = same ? Block() : fp64_rows(ex, Y, ns);
      run(FE_GRAM_NATIVE, x64.ref(), same ? x64.ref() : y64.ref());
      ref = download();
    }
    for (int i = 0; i < 3; ++i) {
      run(kModes[i], xs, ys);
      nt.err[i] = max_rel_diff(download(), ref);
      nt.ms[i] = time_call(ex, [&] { run(kModes[i], X, Y); });
    }
    t = &ex.set_tuning(key, nt);
  }
  run(choose(*t, need), X, Y);
}

// the tall products Y = X C: tuning as for the Gram, the error from a row
// sample against the FP64 product (run(mode, x, y) with an FP64 x and y for the
// reference)
static int tall_mode(HipExecutor& ex, int kind, int mode, double need, const BlockRef& X, const BlockRef& Y,
                     const std::function<void(int, const BlockRef&, const BlockRef&)>& run) {
  if (X.prec == FE_F64 || mode == FE_GRAM_NATIVE || (mode == kProductAuto && !(need > 0.0))) return FE_GRAM_NATIVE;
  if (mode != kProductAuto) return mode;
  const HipExecutor::TuningKey key{kind, X.rows, X.cols, Y.cols};
  const HipExecutor::ProductTuning* t = ex.tuning(key);
  if (!t) {
    HipExecutor::ProductTuning nt;
    const int64_t ns = std::min<int64_t>(X.rows, kTuneRowsTall);
    std::vector<double> ref(static_cast<size_t>(ns) * Y.cols);
    {
      Block x64 = fp64_rows(ex, X, ns), y64(ex, ns, Y.cols, FE_F64);
      run(FE_GRAM_NATIVE, x64.ref(), y64.ref());
      check(fe_memcpy_d2h(ex.ctx(), ref.data(), y64.data(), sizeof(double) * ref.size()), "d2h(sample)");
    }
    Block sample(ex, ns, Y.cols, Y.prec);
    BlockRef xs = X;
    xs.rows = ns;
    std::vector<float> got(ref.size());
    for (int i = 0; i < 3; ++i) {
      run(kModes[i], xs, sample.ref());
      check(fe_memcpy_d2h(ex.ctx(), got.data(), sample.data(), sizeof(float) * got.size()), "d2h(sample)");
      nt.err[i] = max_rel_diff(got, ref);
      nt.ms[i] = time_call(ex, [&] { run(kModes[i], X, Y); });
    }
    t = &ex.set_tuning(key, nt);
  }
  return choose(*t, need);
}

static void tall_product(HipExecutor& ex, int kind, int mode, double need, const BlockRef& X, const BlockRef& Y,
                         const std::function<void(int, const BlockRef&, const BlockRef&)>& run) {
  run(tall_mode(ex, kind, mode, need, X, Y, run), X, Y);
}

void gemm_resid(HipExecutor& ex, const BlockRef& X, const double* C, int64_t ldc, const BlockRef& AY, bool keep_ay,
                const BlockRef& Y, const double* theta, const BlockRef* R, double* rn, double* yn, int mode,
                double need) {
  if (X.prec != AY.prec || X.rows != AY.rows || Y.prec != X.prec || Y.rows != X.rows || Y.cols < AY.cols ||
      (R && (R->prec != X.prec || R->rows != X.rows || R->cols < AY.cols)))
    throw std::invalid_argument("gemm_resid: operand mismatch");
  const int md = tall_mode(ex, kGemm, mode, need, X, AY, [&](int m, const BlockRef& x, const BlockRef& y) {
    check(fe_gemm_nn_mode(ex.ctx(), m, x.prec, x.rows, x.cols, y.cols, x.ptr, x.ld, C, ldc, y.ptr, y.ld, 1.0, 0.0),
          "fe_gemm_nn");
  });
  check(fe_gemm_nn_resid(ex.ctx(), md, X.prec, X.rows, X.cols, AY.cols, X.ptr, X.ld, C, ldc,
                         keep_ay ? AY.ptr : nullptr, AY.ld, Y.ptr, Y.ld, theta, R ? R->ptr : nullptr, R ? R->ld : 0,
                         rn, yn),
        "fe_gemm_nn_resid");
}

void gemm(HipExecutor& ex, const BlockRef& X, const double* C, int64_t ldc, const BlockRef& Y, double alpha,
          double beta, int mode, double need) {
  if (X.prec != Y.prec || X.rows != Y.rows) throw std::invalid_argument("gemm: operand mismatch");
  if (beta != 0.0) mode = FE_GRAM_NATIVE;  // a tuning run would overwrite the Y it reads
  tall_product(ex, kGemm, mode, need, X, Y, [&](int md, const BlockRef& x, const BlockRef& y) {
    check(fe_gemm_nn_mode(ex.ctx(), md, x.prec, x.rows, x.cols, y.cols, x.ptr, x.ld, C, ldc, y.ptr, y.ld, alpha,
                          beta),
          "fe_gemm_nn");
  });
}

void trmm_upper(HipExecutor& ex, const BlockRef& X, const double* C, int64_t ldc, const BlockRef& Y, int mode,
                double need) {
  if (X.prec != Y.prec || X.rows != Y.rows || X.cols != Y.cols) throw std::invalid_argument("trmm_upper: operand mismatch");
  tall_product(ex, kTrmm, mode, need, X, Y, [&](int md, const BlockRef& x, const BlockRef& y) {
    check(fe_trmm_upper_mode(ex.ctx(), md, x.prec, x.rows, x.cols, x.ptr, x.ld, C, ldc, y.ptr, y.ld),
          "fe_trmm_upper");
  });
}

void copy(HipExecutor& ex, const BlockRef& src, const BlockRef& dst) {
  if (src.rows != dst.rows || src.cols != dst.cols) throw std::invalid_argument("copy: shape mismatch");
  check(fe_copy2d(ex.ctx(), src.rows, src.cols, src.ptr, src.prec, src.ld, dst.ptr, dst.prec, dst.ld), "fe_copy2d");
}

void gather_cols(HipExecutor& ex, const BlockRef& src, const std::vector<int>& idx, const BlockRef& dst) {
  if (src.rows != dst.rows || static_cast<int>(idx.size()) != dst.cols || src.prec != dst.prec)
    throw std::invalid_argument("gather_cols: shape mismatch");
  for (int c : idx)
    if (c < 0 || c >= src.cols) throw std::invalid_argument("gather_cols: column out of range");
  check(fe_block_gather_cols(ex.ctx(), src.prec, src.rows, dst.cols, src.ptr, src.ld, idx.data(), dst.ptr, dst.ld),
        "fe_block_gather_cols");
}

void zero(HipExecutor& ex, const BlockRef& X) {
  check(fe_zero2d(ex.ctx(), X.ptr, X.rows, X.cols, X.ld, static_cast<int>(X.esize())), "fe_zero2d");
}

std::vector<double> column_norms(HipExecutor& ex, const BlockRef& X) {
  std::vector<double> out(X.cols);
  if (X.cols == 0) return out;
  Small d(ex, X.cols, 1);
  check(fe_colnorms(ex.ctx(), X.prec, X.rows, X.cols, X.ptr, X.ld, d.data()), "fe_colnorms");
  check(fe_memcpy_d2h(ex.ctx(), out.data(), d.data(), sizeof(double) * X.cols), "column_norms");
  return out;
}

void upload(HipExecutor& ex, const MatrixXd& host, const BlockRef& dst) {
  if (host.rows() != dst.rows || host.cols() != dst.cols) throw std::invalid_argument("upload: shape mismatch");
  check(fe_upload_colmajor(ex.ctx(), dst.rows, dst.cols, host.data(), dst.ptr, dst.prec, dst.ld), "fe_upload");
}

MatrixXd download(HipExecutor& ex, const BlockRef& src) {
  MatrixXd host(static_cast<int>(src.rows), src.cols);
  check(fe_download_colmajor(ex.ctx(), src.rows, src.cols, src.ptr, src.prec, src.ld, host.data()), "fe_download");
  return host;
}

}  // namespace hip
}  // namespace fasteig
