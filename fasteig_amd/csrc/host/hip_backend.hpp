// fasteig::hip -- the MI355X backend namespace that replaces the reference's
// CUDA backend (include/gpu/*, src/gpu/*). It plays the roles of
//   core::Executor      (include/core/Executor.hpp:22-47)      -> HipExecutor
//   core::MultiVector   (include/core/MultiVector.hpp:12-57)   -> Block / BlockRef
//   core::CsrMatrix     (include/core/CsrMatrix.hpp:15-44)     -> HipCsrMatrix
// and calls the device only through the C ABI of libfe_hip.so (fe_hip.h).
//
// Device layout: blocks are row-major n x cols (ld >= cols, padded to a
// multiple of 4 elements); small matrices are FP64 column-major. Work is
// ordered on the executor's single stream; calls that return host values
// synchronize it. Failures of the C ABI become std::invalid_argument (-2) or
// std::runtime_error (otherwise), like the reference's CUDA checks
// (include/gpu/CudaCheck.hpp:18-23).
#pragma once

#include <map>
#include <memory>
#include <string>
#include <tuple>
#include <vector>

#include "fasteig.hpp"
#include "fe_hip.h"

namespace fasteig {
namespace hip {

int check(int rc, const char* what);  // throws on rc < 0, else returns rc

class HipExecutor;

// Non-owning view of a row-major device block (possibly a column range).
struct BlockRef {
  void* ptr = nullptr;
  int64_t rows = 0;
  int cols = 0;
  int64_t ld = 0;
  int prec = FE_F64;
  size_t esize() const { return prec == FE_F64 ? 8 : 4; }
  BlockRef range(int first, int count) const {
    BlockRef r = *this;
    r.ptr = static_cast<char*>(ptr) + static_cast<size_t>(first) * esize();
    r.cols = count;
    return r;
  }
};

// Owning row-major device block (replaces CudaMultiVector storage).
class Block {
 public:
  Block() = default;
  Block(HipExecutor& ex, int64_t rows, int cols, int prec);
  ~Block();
  Block(Block&& o) noexcept { *this = std::move(o); }
  Block& operator=(Block&& o) noexcept;
  Block(const Block&) = delete;
  Block& operator=(const Block&) = delete;

  BlockRef ref() const { return BlockRef{ptr_, rows_, cols_, ld_, prec_}; }
  BlockRef cols(int first, int count) const { return ref().range(first, count); }
  void* data() const { return ptr_; }
  int64_t rows() const { return rows_; }
  int ncols() const { return cols_; }
  int64_t ld() const { return ld_; }
  int prec() const { return prec_; }
  bool empty() const { return ptr_ == nullptr; }
  void swap(Block& o) noexcept;

 private:
  HipExecutor* ex_ = nullptr;
  void* ptr_ = nullptr;
  int64_t rows_ = 0;
  int cols_ = 0;
  int64_t ld_ = 0;
  int prec_ = FE_F64;
};

// Device FP64 column-major small matrix (Gram, Cholesky factor, Ritz coefficients).
class Small {
 public:
  Small() = default;
  Small(HipExecutor& ex, int rows, int cols);
  ~Small();
  Small(Small&& o) noexcept { *this = std::move(o); }
  Small& operator=(Small&& o) noexcept;
  Small(const Small&) = delete;
  Small& operator=(const Small&) = delete;
  double* data() const { return ptr_; }
  int rows() const { return r_; }
  int cols() const { return c_; }
  int64_t ld() const { return r_; }
  MatrixXd to_host() const;
  void from_host(const MatrixXd& M);

 private:
  HipExecutor* ex_ = nullptr;
  double* ptr_ = nullptr;
  int r_ = 0, c_ = 0;
};

// Device copy of a full CSR matrix (replaces CudaCsrMatrix, src/gpu/CudaCsrMatrix.cpp:31-42).
class HipCsrMatrix {
 public:
  HipCsrMatrix(HipExecutor& ex, int n, const std::vector<int>& rowptr, const std::vector<int>& colind,
               const std::vector<double>& vals);
  ~HipCsrMatrix();
  HipCsrMatrix(const HipCsrMatrix&) = delete;
  HipCsrMatrix& operator=(const HipCsrMatrix&) = delete;
  fe_csr handle() const { return h_; }
  int rows() const { return n_; }
  int64_t nonzeros() const { return nnz_; }
  void use_sell(int C = 32, int sigma = 256);
  // Row-block copy (4-row blocks sharing a column list, fe_csr_build_blocked); false
  // when the matrix has too little column locality for it (CSR stays active).
  bool use_blocked();
  // Tiled copy (16-row tiles with their X rows staged in LDS,
  // fe_csr_build_tiled); false when it would not save loads (CSR stays active).
  bool use_tiled();
  void use_csr();
  bool sell() const { return sell_; }
  bool blocked() const { return blocked_; }
  bool tiled() const { return tiled_; }

 private:
  HipExecutor& ex_;
  fe_csr h_ = nullptr;
  int n_ = 0;
  int64_t nnz_ = 0;
  bool sell_built_ = false, sell_ = false;
  int blocked_state_ = 0;  // 0 not tried, 1 built, -1 rejected
  bool blocked_ = false;
  int tiled_state_ = 0;    // 0 not tried, 1 built, -1 rejected
  bool tiled_ = false;
};

class HipExecutor {
 public:
  explicit HipExecutor(int device = 0);
  ~HipExecutor();
  HipExecutor(const HipExecutor&) = delete;
  HipExecutor& operator=(const HipExecutor&) = delete;
  fe_ctx ctx() const { return ctx_; }
  const char* name() const noexcept { return "HIP"; }
  bool is_gpu() const noexcept { return true; }
  void synchronize() const;
  int device() const { return device_; }

  // Seeded start blocks (reference: mt19937 + uniform(-1,1), column-major,
  // src/core/Executor.cpp:20-62), generated on the host once per
  // (rows, cols, seed) and kept resident; later solves copy the cached block.
  const Block& random_block(int64_t rows, int cols, unsigned seed);
  // Drop cached start blocks (frees HBM).
  void clear_cache();

  // Measured time (ms) and relative error against the FP64 product of the
  // three product modes (native, BF16x3, BF16x6) for one product shape.
  struct ProductTuning {
    float ms[3] = {0.0f, 0.0f, 0.0f};
    double err[3] = {0.0, 0.0, 0.0};
  };
  using TuningKey = std::tuple<int, int64_t, int, int>;  // kind, rows, m, p
  const ProductTuning* tuning(const TuningKey& key) const {
    auto it = tuning_.find(key);
    return it == tuning_.end() ? nullptr : &it->second;
  }
  const ProductTuning& set_tuning(const TuningKey& key, const ProductTuning& t) { return tuning_[key] = t; }
  const std::map<TuningKey, ProductTuning>& tunings() const { return tuning_; }

 private:
  fe_ctx ctx_ = nullptr;
  int device_ = 0;
  std::map<std::tuple<int64_t, int, unsigned>, Block> random_cache_;
  std::map<TuningKey, ProductTuning> tuning_;
};

// ---------------------------------------------------------------- primitives
// Y = alpha A X + beta X + gamma Z (Z may alias Y)
void spmm(HipExecutor& ex, const HipCsrMatrix& A, const BlockRef& X, const BlockRef& Y, double alpha,
          double beta = 0.0, const BlockRef* Z = nullptr, double gamma = 0.0);
// Product modes of FP32 blocks (fe_gram_mode), plus kProductAuto: the mode
// is chosen per product shape by measurement. On the first product of a
// shape the three modes run on the actual operands, each timed at full size
// (HIP events on the context stream) and its error measured on a row sample
// (65,536 rows of a Gram, 8,192 of a tall product) against the FP64 product
// of the same FP32 data. A call then takes the fastest mode whose error is at
// most `need` or at most twice the native FP32 product's (need <= 0: native
// only). FP64 blocks always use the native product.
constexpr int kProductAuto = -1;
// G (m x p, ldg) = X^T Y; symmetric: the caller asserts X^T Y is symmetric
// (the Rayleigh-Ritz matrix Q^T (A Q)), so only its upper tiles are computed
void gram(HipExecutor& ex, int mode, const BlockRef& X, const BlockRef& Y, double* G, int64_t ldg,
          const double* w = nullptr, double need = 0.0, bool symmetric = false);
// Y = X C with C upper triangular (m x m, FP64 column-major; nothing below
// its diagonal is read): the X R^{-1} of Cholesky QR
void trmm_upper(HipExecutor& ex, const BlockRef& X, const double* C, int64_t ldc, const BlockRef& Y,
                int mode = FE_GRAM_NATIVE, double need = 0.0);
// Y = alpha X C + beta Y, C FP64 column-major (m x p); the compensated modes
// and their tuning apply when beta == 0 (a tuning run overwrites Y)
void gemm(HipExecutor& ex, const BlockRef& X, const double* C, int64_t ldc, const BlockRef& Y, double alpha,
          double beta, int mode = FE_GRAM_NATIVE, double need = 0.0);
// Rotation with the fused Ritz residual: AY = X C (stored if keep_ay),
// R = AY - Y diag(theta) (if R), rn[j] = ||R[:, j]||, yn[j] = ||Y[:, j]||
// (theta, rn, yn device FP64, AY.cols of them); the product mode is chosen
// as for gemm (the tuning times the plain product into AY)
void gemm_resid(HipExecutor& ex, const BlockRef& X, const double* C, int64_t ldc, const BlockRef& AY, bool keep_ay,
                const BlockRef& Y, const double* theta, const BlockRef* R, double* rn, double* yn,
                int mode = FE_GRAM_NATIVE, double need = 0.0);
void copy(HipExecutor& ex, const BlockRef& src, const BlockRef& dst);
// dst[:, j] = src[:, idx[j]] in one pass (same precision)
void gather_cols(HipExecutor& ex, const BlockRef& src, const std::vector<int>& idx, const BlockRef& dst);
void zero(HipExecutor& ex, const BlockRef& X);
std::vector<double> column_norms(HipExecutor& ex, const BlockRef& X);
void upload(HipExecutor& ex, const MatrixXd& host, const BlockRef& dst);
MatrixXd download(HipExecutor& ex, const BlockRef& src);

}  // namespace hip
}  // namespace fasteig
